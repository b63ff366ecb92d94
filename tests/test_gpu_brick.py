"""GPU parity of the HIP brick path against the oracle, through the C ABI.

Tolerances:
  * connectivity, PARITH/ON CSR (ADSKY/IADS): bit-exact;
  * FP64 element form vs the FP64 oracle: 1e-12 relative per call (the only
    differences are last-ulp ones of the device pow/cbrt); after 1000 cycles
    1e-8 relative on nodes and stresses;
  * mixed form (FP32 element math, FP64 nodes) vs the FP64 form: nodal forces
    within 1e-4 relative (L2 over all nodes) after 1000 cycles;
  * run-to-run: bitwise identical (PARITH/ON assembly, fixed-order reductions).
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from openradioss_amd import deck as dk
from openradioss_amd import engine as en

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def perturbed_deck(nx, ny, nz, seed, vscale=2e3):
    d = dk.brick_block(nx, ny, nz, seed=seed)
    rng = np.random.Generator(np.random.Philox(seed + 1))
    d["v"] = d["v"] + rng.uniform(-vscale, vscale, d["v"].shape)
    ne = d["ixs"].shape[0]
    sig = rng.uniform(-300.0, 300.0, (ne, 6))
    pla = rng.uniform(0.0, 0.1, ne)
    eint = rng.uniform(0.0, 1.0, ne)
    return d, sig, pla, eint


def oracle_group(d, sig, pla, eint, dt1, iads):
    """Oracle element outputs for every element, placed into FSKY slot order."""
    ne = d["ixs"].shape[0]
    fs = np.zeros((8 * ne, 4))
    st_all = np.zeros((ne, po.NST))
    dts = []
    for e in range(ne):
        st = np.zeros(po.NST)
        st[0:6] = sig[e]
        st[po.ST_EINT] = eint[e]
        st[po.ST_RHO] = d["mat"]["rho0"]
        st[po.ST_VOL0] = d["vol0"][e]
        st[po.ST_PLA] = pla[e]
        st[po.ST_OFFG] = 1.0
        nd = d["ixs"][e]
        o = po.brick_element(d["x"][nd].T, d["v"][nd].T, st, dt1, d["mat"], d["prop"])
        for k in range(8):
            fs[iads[e, k] - 1, 0:3] = o["f"][:, k]
            fs[iads[e, k] - 1, 3] = o["sti"]
        st_all[e] = st
        dts.append(o["dtx"])
    return fs, st_all, min(dts)


@pytest.fixture(scope="module")
def deck20():
    return dk.brick_block(20, 20, 20)


def test_pon_csr_bitexact_vs_oracle():
    d = dk.brick_block(6, 5, 4, seed=3)
    rng = np.random.default_rng(3)
    perm = rng.permutation(d["ixs"].shape[0])
    d["ixs"] = d["ixs"][perm]
    d["uid"] = rng.permutation(np.arange(100, 100 + perm.size)).astype(np.int32)
    d["vol0"] = d["vol0"][perm]
    s = en.BrickSolver(d, en.FP64)
    adsky, iads = s.pon()
    a0, i0 = po.build_pon(d["x"].shape[0], d["ixs"], d["uid"])
    np.testing.assert_array_equal(adsky, a0 + 1)
    np.testing.assert_array_equal(iads, i0 + 1)


@pytest.mark.parametrize("prec,tol", [(en.FP64, 1e-12), (en.MIXED, 2e-4)])
def test_sforc3_groups_vs_oracle(prec, tol):
    d, sig, pla, eint = perturbed_deck(8, 7, 5, seed=11)
    s = en.BrickSolver(d, prec)
    s.set_state(sig=sig, pla=pla, eint=eint)
    dt1 = 1.25e-7
    dt2t = 1e30
    ne = s.numels
    for nft in range(0, ne, 128):                     # NVSIZ-sized groups, FORINT order
        dt2t = s.sforc3(nft, min(128, ne - nft), dt1, dt2t)
    fs = s.fsky()
    _, iads = s.pon()
    fo, sto, dto = oracle_group(d, sig, pla, eint, dt1, iads)
    assert rel(fs[:, 0:3], fo[:, 0:3]) < tol
    assert rel(fs[:, 3], fo[:, 3]) < tol
    st = s.state()
    assert rel(st["sig"], sto[:, 0:6]) < tol
    assert rel(st["pla"], sto[:, po.ST_PLA]) < max(tol, 1e-6)
    assert rel(st["eint"], sto[:, po.ST_EINT]) < max(tol, 1e-5)
    assert rel(st["rho"], sto[:, po.ST_RHO]) < tol
    assert abs(dt2t - dto) / dto < tol
    a, stifn = s.asspar4()
    ao = np.zeros_like(a)
    adsky, _ = s.pon()
    for n in range(s.numnod):
        ao[n] = fo[adsky[n] - 1:adsky[n + 1] - 1, 0:3].sum(0)
    assert rel(a, ao) < 10 * tol


def run_both(d, ncyc, prec=en.FP64):
    s = en.BrickSolver(d, prec)
    t = s.run_cycles(ncyc)
    m = po.Model(d)
    m.cycle(ncyc)
    return s, t, m


@pytest.mark.parametrize("vtop,ncyc", [(-1.0e4, 1000), (-2.0e5, 300)])
def test_cycles_fp64_vs_oracle(vtop, ncyc):
    d = dk.brick_block(20, 20, 20, v_top=vtop) if vtop != -1.0e4 else dk.brick_block(20, 20, 20)
    s, t, m = run_both(d, ncyc)
    assert t.ncycle == ncyc and t.mstop == 0
    assert abs(t.tt - m.tt) / m.tt < 1e-12
    assert abs(t.dt2 - m.dt2) / m.dt2 < 1e-12
    nd = s.nodes()
    assert rel(nd["x"] - d["x"], m.x - d["x"]) < 1e-8
    assert rel(nd["v"], m.v) < 1e-8
    assert rel(nd["d"], m.d) < 1e-8
    st = s.state()
    assert rel(st["sig"], m.st[:, 0:6]) < 1e-8
    assert rel(st["rho"], m.st[:, po.ST_RHO]) < 1e-12
    if vtop == -2.0e5:
        assert m.st[:, po.ST_PLA].max() > 1e-3          # plasticity exercised
        assert rel(st["pla"], m.st[:, po.ST_PLA]) < 1e-7
    ei, ek, eh = s.energies()
    eio, eko, eho = m.energies()
    assert abs(ei - eio) / eio < 1e-8
    assert abs(ek - eko) / eko < 1e-8
    assert abs(eh - eho) / max(eho, 1e-30) < 1e-7


def test_mixed_vs_fp64_nodal_force_error_after_1000_cycles(deck20):
    s64 = en.BrickSolver(deck20, en.FP64)
    s32 = en.BrickSolver(deck20, en.MIXED)
    s64.run_cycles(1000)
    s32.run_cycles(1000)
    a64, _ = s64.asspar4()
    a32, _ = s32.asspar4()
    err = rel_l2(a32, a64)
    assert err < 1e-4, err
    x64, x32 = s64.nodes()["d"], s32.nodes()["d"]
    assert rel_l2(x32, x64) < 1e-4


@pytest.mark.parametrize("prec", [en.FP64, en.MIXED])
def test_bitwise_reproducible_runs(prec):
    d = dk.brick_block(24, 20, 16, v_top=-5.0e4)
    outs = []
    for _ in range(2):
        s = en.BrickSolver(d, prec)
        s.run_cycles(150)
        outs.append((s.nodes(), s.state(), s.fsky(), s.energies()))
        s.close()
    (n0, s0, f0, e0), (n1, s1, f1, e1) = outs
    for k in ("x", "v", "d"):
        assert np.array_equal(n0[k], n1[k])
    for k in ("sig", "pla", "eint", "rho", "qvis"):
        assert np.array_equal(s0[k], s1[k])
    assert np.array_equal(f0, f1)
    assert e0 == e1


def test_graph_replay_matches_eager_launches():
    d = dk.brick_block(12, 10, 8, v_top=-5.0e4)
    a = en.BrickSolver(d, en.FP64)
    b = en.BrickSolver(d, en.FP64)
    a.run_cycles(64)                 # graph replays
    b.profile_cycles(64)             # eager launches with events
    na, nb = a.nodes(), b.nodes()
    assert np.array_equal(na["x"], nb["x"]) and np.array_equal(na["v"], nb["v"])
    assert np.array_equal(a.state()["sig"], b.state()["sig"])
    assert a.time().ncycle == b.time().ncycle == 64


def test_large_deck_spot_check_vs_oracle():
    # 100^3 bricks: after 20 cycles, one more element pass is compared with the
    # oracle on a sample of 512 elements (size-independent check of the full deck).
    d = dk.brick_block(100, 100, 100, v_top=-5.0e4)
    s = en.BrickSolver(d, en.FP64)
    t = s.run_cycles(20)
    nd = s.nodes()
    st = s.state()
    rng = np.random.default_rng(0)
    sample = rng.choice(s.numels, 512, replace=False)
    s.sforc3(0, s.numels, t.dt2)
    fs = s.fsky()
    st1 = s.state()
    _, iads = s.pon()
    for e in sample:
        row = np.zeros(po.NST)
        row[0:6] = st["sig"][e]
        row[po.ST_EINT] = st["eint"][e]
        row[po.ST_RHO] = st["rho"][e]
        row[po.ST_VOL0] = d["vol0"][e]
        row[po.ST_PLA] = st["pla"][e]
        row[po.ST_QVIS] = st["qvis"][e]
        row[po.ST_OFFG] = 1.0
        row[po.ST_EPSD] = st["epsd"][e]
        n8 = d["ixs"][e]
        o = po.brick_element(nd["x"][n8].T, nd["v"][n8].T, row, t.dt2, d["mat"], d["prop"])
        got = fs[iads[e] - 1, 0:3].T
        scale = max(np.abs(o["f"]).max(), 1e-6)
        assert np.abs(got - o["f"]).max() / scale < 1e-11
        assert np.allclose(st1["sig"][e], row[0:6], rtol=1e-11, atol=1e-9)


def test_error_paths_return_codes():
    d = dk.brick_block(3, 3, 3)
    s = en.BrickSolver(d, en.FP64)
    with pytest.raises(en.OradError) as ei:
        s.sforc3(20, 100, 1e-7)
    assert ei.value.code == en.ORAD_EINVAL
    # NFT + NEL would overflow a 32-bit int: refused, not wrapped into range
    with pytest.raises(en.OradError) as ei:
        s.sforc3(2**31 - 10, 100, 1e-7)
    assert ei.value.code == en.ORAD_EINVAL
    with pytest.raises(en.OradError) as ei:
        s.sforc3(5, 2**31 - 1, 1e-7)
    assert ei.value.code == en.ORAD_EINVAL
    bad = dict(d)
    bad["ixs"] = d["ixs"].copy()
    bad["ixs"][0, 0] = d["x"].shape[0] + 5
    with pytest.raises(en.OradError) as ei:
        en.BrickSolver(bad, en.FP64)
    assert ei.value.code == en.ORAD_EINVAL


def test_slot_index_range_is_checked():
    """8 * NUMELS slots must fit the 32-bit IADS indices; one brick more is refused
    before any array is read (the tiny arrays passed are never touched), as is an
    empty mesh."""
    lib = en.load_library()
    h = C.c_void_p()
    assert lib.orad_create(C.byref(h), 0, en.FP64) == en.ORAD_OK
    try:
        x = np.zeros((8, 3))
        x[:, 0] = [0, 1, 1, 0, 0, 1, 1, 0]
        x[:, 1] = [0, 0, 1, 1, 0, 0, 1, 1]
        x[:, 2] = [0, 0, 0, 0, 1, 1, 1, 1]
        v, ms, icodt = np.zeros((8, 3)), np.ones(8), np.zeros(8, np.int32)
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        assert lib.orad_set_nodes(h, 8, P(x), P(v), P(ms), P(icodt)) == en.ORAD_OK
        ixs = np.zeros(11, np.int32)
        vol0 = np.ones(1)
        assert lib.orad_set_bricks(h, 0, P(ixs), P(vol0)) == en.ORAD_EINVAL   # empty mesh
        assert lib.orad_set_bricks(h, 268435456, P(ixs), P(vol0)) == en.ORAD_EINVAL
        assert b"32-bit" in lib.orad_last_error(h)
    finally:
        lib.orad_destroy(h)


def test_set_nodes_validates_before_allocating():
    """A first orad_set_nodes with a bad ICODT fails before any device array exists;
    the retry with good codes then sets the nodes (and nothing is allocated twice)."""
    lib = en.load_library()
    h = C.c_void_p()
    assert lib.orad_create(C.byref(h), 0, en.FP64) == en.ORAD_OK
    try:
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        x, v, ms = np.zeros((4, 3)), np.zeros((4, 3)), np.ones(4)
        bad = np.array([0, 9, 0, 0], np.int32)
        assert lib.orad_set_nodes(h, 4, P(x), P(v), P(ms), P(bad)) == en.ORAD_EINVAL
        assert b"ICODT" in lib.orad_last_error(h)
        good = np.array([0, 7, 1, 0], np.int32)
        assert lib.orad_set_nodes(h, 4, P(x), P(v), P(ms), P(good)) == en.ORAD_OK
        x[:, 0] = 1.0
        assert lib.orad_set_nodes(h, 4, P(x), P(v), P(ms), P(good)) == en.ORAD_OK
        got = np.zeros((4, 3))
        assert lib.orad_get_nodes(h, P(got), None, None, None, None) == en.ORAD_OK
        assert np.array_equal(got, x)
    finally:
        lib.orad_destroy(h)


def test_empty_group_is_a_no_op():
    """NEL = 0 (an empty group) touches nothing: DT2T, FSKY and the state stay as
    they were, for bricks and both shell formulations."""
    d = dk.brick_block(4, 3, 2, v_top=-5.0e4)
    s = en.BrickSolver(d, en.FP64)
    s.sforc3(0, s.numels, 1.0e-7)
    f0, st0 = s.fsky(), s.state()
    assert s.sforc3(7, 0, 1.0e-7, dt2t=0.5) == 0.5
    assert np.array_equal(s.fsky(), f0)
    assert all(np.array_equal(s.state()[k], st0[k]) for k in st0)
    s.close()
    for ishell in (1, 24):
        ds = dk.shell_plate(n=6, patch=2)
        ds["prop"] = dict(ds["prop"], ishell=ishell)
        sh = en.ShellSolver(ds, en.FP64)
        sh.cforc3(0, sh.numelc, 1.0e-7)
        g0, r0 = sh.fsky(), sh.state()
        assert sh.cforc3(5, 0, 1.0e-7, dt2t=0.25) == 0.25
        assert np.array_equal(sh.fsky(), g0) and np.array_equal(sh.state(), r0)
        sh.close()


def test_negative_volume_sets_mstop():
    d = dk.brick_block(3, 3, 3)
    ixs = d["ixs"].copy()
    ixs[13] = ixs[13][[4, 5, 6, 7, 0, 1, 2, 3]]       # turn one brick inside out
    d["ixs"] = ixs
    s = en.BrickSolver(d, en.FP64)
    with pytest.raises(en.OradError) as ei:
        s.run_cycles(2)
    assert ei.value.code == en.ORAD_ENEGVOL
