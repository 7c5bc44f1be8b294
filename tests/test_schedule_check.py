"""The race check of owner-tile schedules (tmrx_check_tile_schedule, run by
every tile build before upload) on hand-made schedules: a valid one passes,
and each kind of mistake -- two records of one colour writing a slot, records
out of colour order, a write beyond the accumulators -- is reported."""
import ctypes as C

import numpy as np

from tmr_amd._lib import lib

REC = np.dtype([("lc", np.uint16, 8), ("elem", np.int32), ("mask", np.uint8),
                ("color", np.uint8), ("cls", np.uint8), ("pad", np.uint8)])
assert REC.itemsize == 24


def tile_pos(s):
    return s ^ (((s >> 3) ^ (s >> 6) ^ (s >> 9)) & 7)


def rec(color, slots_written, other=()):
    """A record writing the given accumulator slots (mask bits set) and
    reading others."""
    r = np.zeros(1, REC)[0]
    lc = list(slots_written) + list(other)
    lc += [lc[-1]] * (8 - len(lc))
    r["lc"] = [tile_pos(s) for s in lc]
    r["mask"] = (1 << len(slots_written)) - 1
    r["color"] = color
    return r


def check(tiles, nacc):
    recs = np.array([r for t in tiles for r in t], dtype=REC)
    ptr = np.cumsum([0] + [len(t) for t in tiles]).astype(np.intc)
    na = np.asarray(nacc, dtype=np.intc)
    msg = C.create_string_buffer(256)
    st = lib().tmrx_check_tile_schedule(len(tiles), ptr.ctypes.data, recs.ctypes.data,
                                        na.ctypes.data, msg, 256)
    return st, msg.value.decode()


def test_valid_schedule_passes():
    t0 = [rec(0, [0, 1], [9]), rec(0, [2, 3]), rec(1, [0, 2], [12]), rec(2, [1, 3])]
    t1 = [rec(0, [20, 21]), rec(1, [20, 22])]
    assert check([t0, t1], [8, 24]) == (0, "")


def test_same_colour_conflict_is_reported():
    t0 = [rec(0, [0, 1]), rec(0, [1, 2])]
    st, msg = check([t0], [8])
    assert st == -1 and "twice" in msg and "slot 1" in msg


def test_colour_order_is_enforced():
    t0 = [rec(1, [0]), rec(0, [1])]
    st, msg = check([t0], [8])
    assert st == -1 and "not sorted" in msg


def test_write_beyond_accumulators_is_reported():
    t0 = [rec(0, [0, 9])]
    st, msg = check([t0], [8])
    assert st == -1 and "beyond the accumulators" in msg


def test_conflict_in_a_later_tile():
    ok = [rec(0, [0, 1])]
    bad = [rec(3, [40, 41]), rec(3, [41, 42])]
    st, msg = check([ok, bad], [8, 48])
    assert st == -1 and msg.startswith("tile 1")
