# Builds the reference Phantom-GRAPE libraries from their own C sources, where
# they lie under the reference tree, into oracle/_ref/ (git-ignored).
# These shared objects are test infrastructure only: the CPU baseline in
# bench.py and the parity tests load them as the checker, never as product.
#
#   make -f oracle/ref.mk REF=/root/reference
#
# Flags follow the reference's own setting.mk / Makefiles:
#   PG5/newton/setting.mk:3  CFLAGS = -O3 -ffast-math -funroll-loops, avx2 -> -mavx2 -mfma
#   PG5/table/libpg5/Makefile:31-32  compiled with $(ARCH) only (no CFLAGS)
#   PG6/setting.mk:3,15,33  -O3 -ffast-math -funroll-loops -fopenmp -DNORMAL
# Objects are compiled with the reference's flags but linked without
# -ffast-math: linking with it would add crtfastmath.o, whose constructor sets
# flush-to-zero for the whole host process that loads the library.
# Each library is built twice for PG5 (serial and ENABLE_OPENMP) because the
# reference selects g5_calculate_force_on_x's threading at compile time
# (PG5/newton/libpg5/phantom_g5.c:302-383).

REF ?= /root/reference
OUT ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
CC  ?= gcc

PG5N  := $(REF)/PG5/newton/libpg5
PG5T  := $(REF)/PG5/table/libpg5
PG5Q  := $(REF)/PG5/newton_quad/quad
PG6   := $(REF)/PG6/libpg6

NEWTON_FLAGS := -O3 -ffast-math -funroll-loops -mavx2 -mfma -fPIC
TABLE_FLAGS  := -mavx2 -mfma -fPIC
PG6_FLAGS    := -O3 -ffast-math -funroll-loops -fopenmp -DNORMAL -mavx2 -mfma -fPIC

# compile each source to its own object in a scratch dir, then link a .so
# $(call so,output,flags,includes,sources,linkflags)
define so
	rm -rf $(1).objs && mkdir -p $(1).objs
	cd $(1).objs && for f in $(4); do $(CC) $(2) $(3) -c $$f -o $$(basename $$f .c).o || exit 1; done
	$(CC) -shared $(5) $(1).objs/*.o -o $(1) -lm
	rm -rf $(1).objs
endef

ALL := $(OUT)/libref_pg5.so $(OUT)/libref_pg5_omp.so $(OUT)/libref_pg5sym.so $(OUT)/libref_pg5quad.so \
       $(OUT)/libref_pg5table.so $(OUT)/libref_pg5table_omp.so $(OUT)/libref_pg6.so \
       $(OUT)/direct_ref $(OUT)/direct_pgamd $(OUT)/banana_ref $(OUT)/banana_pgamd \
       $(OUT)/force_prof_ref $(OUT)/force_prof_pgamd

# The reference's own sample programs, compiled from their sources against
# their own headers, linked twice: with the reference library (*_ref) and with
# this repository's HIP library (*_pgamd) -- the drop-in check.  They are
# checkers like the libraries above and run only from tests/.
PGLIB   := $(abspath $(OUT)/../../phantom-grape_amd/lib)
DIRECT  := $(REF)/PG5/newton/direct
BANANA  := $(REF)/PG6/banana
DIRECT_SRC := $(DIRECT)/direct.c $(DIRECT)/cpu.c
BANANA_SRC := $(BANANA)/banana01.c $(BANANA)/io.c $(BANANA)/calc_energy.c $(BANANA)/hermite.c \
              $(BANANA)/timestep.c $(BANANA)/timeprof.c
FPROF   := $(REF)/PG5/table/force_prof
FPROF_SRC := $(FPROF)/force_prof.c $(FPROF)/cpu.c

# $(call prog,output,flags,includes,sources,linkflags)
define prog
	rm -rf $(1).objs && mkdir -p $(1).objs
	cd $(1).objs && for f in $(4); do $(CC) $(2) $(3) -c $$f -o $$(basename $$f .c).o || exit 1; done
	$(CC) $(1).objs/*.o -o $(1) $(5) -lm
	rm -rf $(1).objs
endef

all: $(ALL)

$(OUT):
	mkdir -p $(OUT)

$(OUT)/libref_pg5.so: $(PG5N)/gravity_avx2.c $(PG5N)/phantom_g5.c $(PG5N)/rsqrt.c $(PG5N)/pg5_fortran.c | $(OUT)
	$(call so,$(abspath $@),$(NEWTON_FLAGS),-I$(PG5N),$^,)

$(OUT)/libref_pg5_omp.so: $(PG5N)/gravity_avx2.c $(PG5N)/phantom_g5.c $(PG5N)/rsqrt.c $(PG5N)/pg5_fortran.c | $(OUT)
	$(call so,$(abspath $@),$(NEWTON_FLAGS) -fopenmp -DENABLE_OPENMP,-I$(PG5N),$^,-fopenmp)

# enable_symmetric=yes (PG5/newton/setting.mk:6,17-19)
$(OUT)/libref_pg5sym.so: $(PG5N)/gravity_avx2.c $(PG5N)/phantom_g5.c $(PG5N)/rsqrt.c $(PG5N)/pg5_fortran.c | $(OUT)
	$(call so,$(abspath $@),$(NEWTON_FLAGS) -DSYMMETRIC,-I$(PG5N),$^,)

# quad.a = libpg5 kernel + quad kernel + g5c API (which #includes phantom_g5.c)
$(OUT)/libref_pg5quad.so: $(PG5N)/gravity_avx2.c $(PG5Q)/quad_avx2.c $(PG5Q)/quad_g5c.c $(PG5N)/rsqrt.c | $(OUT)
	$(call so,$(abspath $@),$(NEWTON_FLAGS) -Wno-unused-result,-I$(PG5N),$^,)

$(OUT)/libref_pg5table.so: $(PG5T)/gravity_kernel_avx2.c $(PG5T)/phantom_g5.c $(PG5T)/pg5_table.c | $(OUT)
	$(call so,$(abspath $@),$(TABLE_FLAGS),-I$(PG5T),$^,)

$(OUT)/libref_pg5table_omp.so: $(PG5T)/gravity_kernel_avx2.c $(PG5T)/phantom_g5.c $(PG5T)/pg5_table.c | $(OUT)
	$(call so,$(abspath $@),$(TABLE_FLAGS) -fopenmp -DENABLE_OPENMP,-I$(PG5T),$^,-fopenmp)

$(OUT)/libref_pg6.so: $(PG6)/phantom_g6.c $(PG6)/gravity_avx2.c $(PG6)/timeprof.c | $(OUT)
	$(call so,$(abspath $@),$(PG6_FLAGS),-I$(PG6),$^,-fopenmp)

# direct.c as PG5/newton/direct/Makefile builds it (setting.mk: no OpenMP)
$(OUT)/direct_ref: $(DIRECT_SRC) $(OUT)/libref_pg5.so | $(OUT)
	$(call prog,$(abspath $@),-O3 -ffast-math -funroll-loops -Wno-unused-result,-I$(DIRECT),$(DIRECT_SRC),-L$(abspath $(OUT)) -lref_pg5 -Wl$(COMMA)-rpath$(COMMA)'$$ORIGIN')

$(OUT)/direct_pgamd: $(DIRECT_SRC) $(PGLIB)/libpg5.so | $(OUT)
	$(call prog,$(abspath $@),-O3 -ffast-math -funroll-loops -Wno-unused-result,-I$(DIRECT),$(DIRECT_SRC),-L$(PGLIB) -lpg5 -Wl$(COMMA)-rpath$(COMMA)'$$ORIGIN/../../phantom-grape_amd/lib')

# banana01 as PG6/banana/Makefile builds it (PG6/setting.mk: OpenMP, NORMAL)
$(OUT)/banana_ref: $(BANANA_SRC) $(OUT)/libref_pg6.so | $(OUT)
	$(call prog,$(abspath $@),-O3 -ffast-math -funroll-loops -fopenmp -DNORMAL -Wno-unused-result,-I$(PG6) -I$(BANANA),$(BANANA_SRC),-fopenmp -L$(abspath $(OUT)) -lref_pg6 -Wl$(COMMA)-rpath$(COMMA)'$$ORIGIN')

$(OUT)/banana_pgamd: $(BANANA_SRC) $(PGLIB)/libpg6.so | $(OUT)
	$(call prog,$(abspath $@),-O3 -ffast-math -funroll-loops -fopenmp -DNORMAL -Wno-unused-result,-I$(PG6) -I$(BANANA),$(BANANA_SRC),-fopenmp -L$(PGLIB) -lpg6 -Wl$(COMMA)-rpath$(COMMA)'$$ORIGIN/../../phantom-grape_amd/lib')

# force_prof as PG5/table/force_prof/Makefile builds it (PG5/table/setting.mk:
# enable_openmp=no, so the serial table library)
$(OUT)/force_prof_ref: $(FPROF_SRC) $(OUT)/libref_pg5table.so | $(OUT)
	$(call prog,$(abspath $@),-O3 -ffast-math -funroll-loops -std=c99 -Wno-unused-result,-I$(FPROF),$(FPROF_SRC),-L$(abspath $(OUT)) -lref_pg5table -Wl$(COMMA)-rpath$(COMMA)'$$ORIGIN')

$(OUT)/force_prof_pgamd: $(FPROF_SRC) $(PGLIB)/libpg5_table.so | $(OUT)
	$(call prog,$(abspath $@),-O3 -ffast-math -funroll-loops -std=c99 -Wno-unused-result,-I$(FPROF),$(FPROF_SRC),-L$(PGLIB) -lpg5_table -Wl$(COMMA)-rpath$(COMMA)'$$ORIGIN/../../phantom-grape_amd/lib')

COMMA := ,

.PHONY: all
