"""The libraries' error state under concurrent misuse (no GPU needed: a
devid outside the board range is reported before any device work).

Eight threads report errors through g5_set_nMC with distinct bad devids
while another reads pgamd_last_error: every message read is one whole
message, never a mix of two."""
import ctypes as C
import re
import threading

import pgamd


def test_error_messages_stay_whole_under_concurrent_reports():
    L = pgamd.load("pg5")
    L.pgamd_set_abort_on_error.argtypes = [C.c_int]
    L.pgamd_last_error.argtypes = [C.c_char_p, C.c_int]
    L.pgamd_last_error.restype = C.c_int
    L.g5_set_nMC.argtypes = [C.c_int, C.c_int]
    L.g5_set_nMC.restype = None
    L.pgamd_set_abort_on_error(0)
    pat = re.compile(r"g5_set_n: devid -(\d+) outside \[0, \d+\)")
    stop = threading.Event()
    seen, bad = [], []

    def writer(k):
        for _ in range(3000):
            L.g5_set_nMC(-(1000 + k), 1)

    def reader():
        buf = C.create_string_buffer(512)
        while not stop.is_set():
            code = L.pgamd_last_error(buf, 512)
            if code == 0:
                continue
            msg = buf.value.decode()
            m = pat.fullmatch(msg)
            if code != 4 or not m or not 1000 <= int(m.group(1)) < 1008:
                bad.append((code, msg))
            seen.append(msg)

    r = threading.Thread(target=reader)
    r.start()
    ws = [threading.Thread(target=writer, args=(k,)) for k in range(8)]
    try:
        for w in ws:
            w.start()
        for w in ws:
            w.join()
    finally:
        stop.set()
        r.join()
        # the state is cleared by a read
        L.pgamd_last_error(None, 0)
        last = L.pgamd_last_error(None, 0)
        L.pgamd_set_abort_on_error(1)
    assert not bad, bad[:5]
    assert seen
    assert last == 0
