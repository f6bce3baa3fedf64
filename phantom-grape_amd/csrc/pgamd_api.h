// Visibility macro shared by the exported C entry points of every library.
#pragma once
#define PGAMD_API extern "C" __attribute__((visibility("default")))
