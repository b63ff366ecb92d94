"""Brick launch choices: the library's element order (columns of element rows walked
layer by layer), the element-kernel instances (threads per workgroup, waves per
SIMD), the LDS residency cap, the XCD mapping and the nodal slot tile, and the
autotuner that picks among them.

Bars: nodes, element state and FSKY slots bitwise identical under every choice (each
slot is written by exactly one element and summed in slot order); internal and
kinetic energy bitwise (summed in the engine's element order); the cumulative
hourglass energy, whose per-workgroup partials follow the storage order, within
1e-12.  The per-group drop-in (engine-numbered groups) gives the whole-deck launch's
slots bitwise under a permuted order.
"""
import ctypes as C

import numpy as np
import pytest

from openradioss_amd import deck as dk
from openradioss_amd import engine as en

pytestmark = pytest.mark.gpu


def run(d, prec, ncyc, **launch):
    s = en.BrickSolver(d, prec)
    if launch:
        s.set_launch(**launch)
    s.run_cycles(ncyc)
    out = dict(nodes=s.nodes(), state=s.state(), fsky=s.fsky(), en=s.energies(),
               launch=s.launch(), time=s.time())
    s.close()
    return out


def same(a, b, what):
    for k in ("x", "v", "d"):
        assert np.array_equal(a["nodes"][k], b["nodes"][k]), (what, k)
    for k in a["state"]:
        assert np.array_equal(a["state"][k], b["state"][k]), (what, k)
    assert np.array_equal(a["fsky"], b["fsky"]), what
    assert a["en"][0] == b["en"][0] and a["en"][1] == b["en"][1], what
    assert a["en"][2] == pytest.approx(b["en"][2], rel=1e-12, abs=1e-30), what
    assert a["time"].tt == b["time"].tt, what


@pytest.mark.parametrize("prec", [en.MIXED, en.FP64])
def test_element_order_and_instances_give_identical_results(prec):
    d = dk.brick_block(23, 17, 12, v_top=-5.0e4)
    ref = run(d, prec, 120, order_rows=0)
    assert ref["launch"]["order_rows"] == 0
    shapes = [(256, 2), (256, 3), (128, 2), (64, 2)] if prec == en.MIXED else \
        [(64, 1), (128, 1), (256, 1), (64, 2)]
    for rows in (1, 5, 16, 33):
        got = run(d, prec, 120, order_rows=rows)
        assert got["launch"]["order_rows"] == rows
        same(got, ref, f"rows {rows}")
    for b, w in shapes:
        got = run(d, prec, 120, block=b, waves=w)
        assert (got["launch"]["block"], got["launch"]["waves"]) == (b, w)
        same(got, ref, f"instance {b}/{w}")
    for tile in ((1024, 2048, 4096) if prec == en.MIXED else (1024, 2048)):
        got = run(d, prec, 120, node_tile=tile)
        assert got["launch"]["node_tile"] == tile
        same(got, ref, f"tile {tile}")
    got = run(d, prec, 120, lds_kb=80, xcd_remap=0)
    same(got, ref, "lds 80, no xcd remap")


def test_default_order_is_columns_and_groups_follow_the_engine_numbering():
    """The default context stores bricks in column order; SFORC3-style groups are
    numbered in the engine's order and must give the slots of the whole launch."""
    d = dk.brick_block(19, 14, 9, v_top=-5.0e4)
    a = en.BrickSolver(d, en.FP64)
    assert a.launch()["order_rows"] == 10
    a.set_launch(order_rows=3)            # 14 rows in columns of 3
    t = a.run_cycles(30)
    b = en.BrickSolver(d, en.FP64)
    b.set_launch(order_rows=0)
    b.run_cycles(30)
    na = a.nodes()
    # one more element pass: whole deck in library order (a) vs engine groups of 128 (b)
    a.sforc3(0, a.numels, t.dt2)
    for nft in range(0, b.numels, 128):
        b.sforc3(nft, min(128, b.numels - nft), t.dt2)
    assert np.array_equal(a.fsky(), b.fsky())
    assert np.array_equal(na["x"], b.nodes()["x"])
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    a.close()
    b.close()


def test_set_state_round_trips_through_the_order():
    d = dk.brick_block(9, 8, 7, v_top=-5.0e4)
    s = en.BrickSolver(d, en.MIXED)
    s.set_launch(order_rows=3)
    rng = np.random.default_rng(1)
    sig = rng.uniform(-50.0, 50.0, (s.numels, 6)).astype(np.float32).astype(np.float64)
    pla = rng.uniform(0.0, 0.01, s.numels).astype(np.float32).astype(np.float64)
    s.set_state(sig=sig, pla=pla)
    st = s.state()
    assert np.array_equal(st["sig"], sig) and np.array_equal(st["pla"], pla)
    s.set_launch(order_rows=7)
    st = s.state()
    assert np.array_equal(st["sig"], sig) and np.array_equal(st["pla"], pla)
    s.close()


@pytest.mark.parametrize("prec", [en.MIXED, en.FP64])
def test_autotune_picks_instance_order_and_tile_and_keeps_results(prec):
    d = dk.brick_block(30, 24, 20, v_top=-5.0e4)
    ref = run(d, prec, 60)
    s = en.BrickSolver(d, prec)
    s.run_cycles(20)
    best, ms = s.autotune(cycles=6)
    got_launch = s.launch()
    print(f"autotuned {best}, {ms:.4f} ms per cycle", flush=True)
    assert ms > 0.0 and best == got_launch
    assert got_launch["block"] > 0 and got_launch["waves"] > 0 and got_launch["node_tile"] > 0
    assert s.time().ncycle == 20
    s.run_cycles(40)
    got = dict(nodes=s.nodes(), state=s.state(), fsky=s.fsky(), en=s.energies(),
               time=s.time())
    s.close()
    same(got, ref, "autotuned")


def test_node_shared_by_more_than_256_corners_is_refused():
    """Corner ranks are 8-bit: a node in more than 256 brick corners cannot be
    addressed and set_bricks says so (arrays are checked on the host only)."""
    lib = en.load_library()
    h = C.c_void_p()
    assert lib.orad_create(C.byref(h), 0, en.FP64) == en.ORAD_OK
    try:
        nb = 300                     # 300 bricks all sharing node 1 as corner 1
        nn = 1 + 7 * nb
        x = np.random.default_rng(0).uniform(0, 1, (nn, 3))
        v, ms, icodt = np.zeros((nn, 3)), np.ones(nn), np.zeros(nn, np.int32)
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        assert lib.orad_set_nodes(h, nn, P(x), P(v), P(ms), P(icodt)) == en.ORAD_OK
        ixs = np.zeros((nb, 11), np.int32)
        ixs[:, 0] = 1
        ixs[:, 1] = 1
        ixs[:, 2:9] = 2 + np.arange(7 * nb).reshape(nb, 7)
        ixs[:, 9] = 1
        ixs[:, 10] = np.arange(1, nb + 1)
        vol0 = np.ones(nb)
        assert lib.orad_set_bricks(h, nb, P(ixs), P(vol0)) == en.ORAD_EINVAL
        assert b"256" in lib.orad_last_error(h)
    finally:
        lib.orad_destroy(h)
