"""GPU parity of the /INTER/TYPE7 contact pass through the C ABI.

Tolerances:
  * one contact pass vs the reference routines' outputs (tests/golden/int7_cases.npz,
    I7COR3/I7DST3/I7FOR3/I7ASS2 compiled from source): every skyline entry and its
    node, the contact time step, the penetrating-pair count and the deactivated nodes
    bit for bit; the elastic and friction energies within 1e-13 (their LOG is the
    device's, the reference's is the host libm's);
  * whole FP64 cycles with contact vs the FP64 oracle (a shell plate dropped on a
    clamped plate, a brick block dropped on a fixed block): 1e-8 relative on nodes and
    energies after the impact and rebound, time within 1e-12;
  * mixed form vs FP64 form on the shell drop: displacement relative L2 < 1e-4;
  * a tube crushed with self-contact: elastic, bit for bit the FP64 oracle's run (the
    element arithmetic is the same on both sides); plastic, within 1e-5 (the law's
    exp/pow differ in the last bit and the contact damping switches on at a finite value);
  * run-to-run and restart: bitwise.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from openradioss_amd import deck as dk
from openradioss_amd import engine as en

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "int7_cases.npz")


def load():
    z = np.load(GOLD)
    names = sorted({k.split("/")[0] for k in z.files})
    return {n: {k.split("/")[1]: z[k] for k in z.files if k.startswith(n + "/")} for n in names}


CASES = load()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def golden_deck(c):
    """Shells on the golden segments (so the context has a mesh) and the secondary
    nodes as free nodes; the case's positions, velocities and masses."""
    gap, fric, visc, viscf, _ = c["params"]
    x = c["x"]
    nn = x.shape[0]
    irect = c["irect"]
    ne = irect.shape[0]
    return dict(x=x, v=c["v"], vr=np.zeros_like(x), ms=c["ms"], **{"in": c["ms"] * 1e-3},
                icodt=np.zeros(nn, np.int32), icodr=np.zeros(nn, np.int32),
                ixc=irect.astype(np.int32), uid=np.arange(1, ne + 1, dtype=np.int32),
                thk=np.ones(ne), mat=dk.shell_mat_params(), prop=dk.shell_prop_params(),
                contact=dict(nsv=c["nsv"], irect=irect, stfm=c["stfm"], stfns=c["stfns"],
                             gap=gap, fric=fric, visc=visc, viscf=viscf))


@pytest.mark.parametrize("name", sorted(CASES))
def test_contact_pass_bitwise_vs_reference_routines(name):
    c = CASES[name]
    s = en.ShellSolver(golden_deck(c), en.FP64)
    try:
        dt2t = s.contact7(float(c["params"][4]))
        f, isky = s.contact7_sky()
        (ec, ev, ed), counts, stfns = s.contact7_state()
    finally:
        s.close()
    assert counts["npen"] == int(c["full_pen"].sum()) and counts["overflow"] == 0
    assert counts["ncand"] >= counts["npen"]
    assert np.array_equal(isky, c["act_isky"])
    assert np.array_equal(f, c["act_fsky"]), np.abs(f - c["act_fsky"]).max()
    assert dt2t == c["act_scal"][0]
    assert ec == pytest.approx(c["act_scal"][1], rel=1e-13, abs=1e-300)
    assert ev == pytest.approx(c["act_scal"][2], rel=1e-13, abs=1e-300)
    assert ed == 0.0
    assert np.array_equal(stfns, c["full_stfns"])


def contact_energies(m):
    k = m.contact.it
    return k.econt, k.econtv


@pytest.mark.parametrize("family", ["shell", "brick"])
def test_contact_cycles_fp64_vs_oracle(family):
    if family == "shell":
        d = dk.shell_plates_contact(n=12, m=6, fric=0.2, viscf=0.5, tilt=0.02)
        m, s = po.ShellModel(d, nthreads=4), en.ShellSolver(d, en.FP64)
    else:
        d = dk.brick_blocks_contact(fric=0.2, viscf=0.5)
        m, s = po.Model(d, nthreads=4), en.BrickSolver(d, en.FP64)
    peak = 0
    try:
        for _ in range(30):
            m.cycle(20)
            t = s.run_cycles(20)
            peak = max(peak, m.contact.it.npen)
            _, cnt, _ = s.contact7_state()
            assert cnt["npen"] == m.contact.it.npen
        nd = s.nodes()
        (ec, ev, _), _, _ = s.contact7_state()
        ei, ek, eh = s.energies()
    finally:
        s.close()
    assert peak > 0                                          # the bodies met
    assert t.tt == pytest.approx(m.m.tt, rel=1e-12) and t.dt2 == pytest.approx(m.m.dt2, rel=1e-12)
    assert rel(nd["x"], m.x) < 1e-12
    assert rel(nd["d"], m.d) < 1e-8 and rel(nd["v"], m.v) < 1e-8
    oc, ov = contact_energies(m)
    assert oc > 0.0 and ec == pytest.approx(oc, rel=1e-8)
    assert ev == pytest.approx(ov, rel=1e-8, abs=1e-12 * oc)
    oi, ok, oh = m.energies()
    assert ei == pytest.approx(oi, rel=1e-8) and ek == pytest.approx(ok, rel=1e-8)
    # the upper body came back up
    assert nd["v"][d["contact"]["nsv"], 2].mean() > 0.0


def test_contact_mixed_vs_fp64_shell_drop():
    d = dk.shell_plates_contact(n=16, m=8)
    out = {}
    for prec in (en.FP64, en.MIXED):
        s = en.ShellSolver(d, prec)
        s.run_cycles(600)
        out[prec] = (s.nodes()["d"], s.contact7_state()[0][0])
        s.close()
    e = np.linalg.norm(out[en.MIXED][0] - out[en.FP64][0]) / np.linalg.norm(out[en.FP64][0])
    print(f"contact shell drop mixed vs fp64 after 600 cycles: displacement {e:.2e}")
    assert e < 1e-4
    assert out[en.MIXED][1] == pytest.approx(out[en.FP64][1], rel=1e-3)


def test_contact_run_to_run_and_restart_bitwise(tmp_path):
    d = dk.shell_plates_contact(n=12, m=6, fric=0.3, viscf=1.0, tilt=0.05)
    res = []
    for _ in range(2):
        s = en.ShellSolver(d, en.MIXED)
        s.run_cycles(150)
        s.write_restart(str(tmp_path / "c.rst"))
        s.run_cycles(150)
        res.append((s.nodes()["d"], s.state(), s.contact7_state()[0]))
        s.close()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]
    r = en.ShellSolver(d, en.MIXED)
    r.read_restart(str(tmp_path / "c.rst"))
    r.run_cycles(150)
    got = (r.nodes()["d"], r.state(), r.contact7_state()[0])
    r.close()
    assert np.array_equal(got[0], res[0][0]) and np.array_equal(got[1], res[0][1])
    assert got[2] == res[0][2]


def test_contact_group_path_adds_the_entries_once():
    d = dk.brick_blocks_contact()
    s = en.BrickSolver(d, en.FP64)
    try:
        for _ in range(40):                   # into the impact
            s.run_cycles(2)
            if s.contact7_state()[1]["npen"] > 4:
                break
        t = s.time()
        s.sforc3(0, s.numels, t.dt2)
        a0, _ = s.asspar4()
        s.contact7(t.dt2)
        f, isky = s.contact7_sky()
        a1, _ = s.asspar4()
        a2, _ = s.asspar4()
    finally:
        s.close()
    assert len(isky) > 0
    exp = a0.copy()
    for k in range(len(isky)):
        exp[isky[k]] += f[k, :3]
    assert np.allclose(a1, exp, rtol=1e-12, atol=1e-9 * np.abs(f[:, :3]).max())
    assert np.array_equal(a2, a0)            # the entries are added once


def test_contact_moved_nodes_are_searched_again():
    """The last search's candidates stand while every node stays within GAP/4 of its place
    then; a larger move (the engine's new X through orad_update_xv) searches again, so the
    pass finds the pairs an every-cycle search finds (the oracle's on the same X)."""
    d = dk.shell_plates_contact(n=12, m=6)
    c = d["contact"]
    nsv = np.asarray(c["nsv"])
    low = np.unique(np.asarray(c["irect"]))
    k = po.Contact7(dict(nsv=nsv, irect=c["irect"], stfm=c["stfm"], stfns=c["stfns"],
                         gap=c["gap"], fric=c.get("fric", 0.0), visc=c.get("visc", 0.05),
                         viscf=c.get("viscf", 0.0)))
    x0 = np.asarray(d["x"], float).copy()
    v = np.zeros_like(x0)
    s = en.ShellSolver(d, en.FP64)
    try:
        got = []
        for dz in (0.0, x0[nsv, 2].min() - x0[low, 2].max() - 0.5 * c["gap"], 0.1 * c["gap"]):
            x0[nsv, 2] -= dz
            s.update_xv(x0, v, np.zeros_like(x0))
            s.contact7(1.0e-7)
            cn, ce = k.candidates(x0)
            exp = k.forces(x0, v, d["ms"], cn, ce, 1.0e-7)["npen"]
            got.append((s.contact7_state()[1]["npen"], exp))
    finally:
        s.close()
    assert got[0][0] == got[0][1]
    assert got[1][1] > 0 and got[1][0] == got[1][1]      # moved by more than GAP/4
    assert got[2][1] > 0 and got[2][0] == got[2][1]      # moved by less: the kept candidates


def test_contact_capacity_overflow_is_reported():
    d = dk.shell_plates_contact(n=12, m=6)
    d["contact"]["cand_cap"] = 8
    s = en.ShellSolver(d, en.FP64)
    try:
        with pytest.raises(en.OradError, match="capacity"):
            s.run_cycles(200)
    finally:
        s.close()


def _tube_run(d, chunks, every):
    m, s = po.ShellModel(d, nthreads=8), en.ShellSolver(d, en.FP64)
    worst, peak = 0.0, 0
    try:
        for _ in range(chunks):
            m.cycle(every)
            s.run_cycles(every)
            _, cnt, _ = s.contact7_state()
            assert cnt["npen"] == m.contact.it.npen
            peak = max(peak, m.contact.it.npen)
            worst = max(worst, rel(s.nodes()["d"], m.d))
        nd = s.nodes()
        (ec, ev, _), _, _ = s.contact7_state()
        t = s.time()
    finally:
        s.close()
    return m, nd, ec, ev, t, worst, peak


def test_tube_crush_self_contact_elastic_bitwise():
    """A square tube crushed at 200 m/s with a self-contact interface (every node against
    every shell of the tube), elastic so that the element arithmetic of the device and the
    oracle is the same: the whole run, through up to 160 penetrating pairs, is bit for bit
    the FP64 oracle's."""
    d = dk.tube_self_contact(v_imp=-2.0e5)
    d["mat"] = dict(d["mat"], ca=1.0e7)          # the yield stress is never reached
    m, nd, ec, ev, t, worst, peak = _tube_run(d, 12, 50)
    assert peak > 50
    assert worst == 0.0
    assert np.array_equal(nd["x"], m.x) and np.array_equal(nd["v"], m.v)
    assert np.array_equal(nd["vr"], m.vr)
    assert t.tt == m.m.tt and t.dt2 == m.m.dt2
    # the energies' LOG is the device's
    assert ec == pytest.approx(m.contact.it.econt, rel=1e-13)
    assert ev == pytest.approx(m.contact.it.econtv, rel=1e-13, abs=1e-300)


def test_tube_crush_self_contact_plastic_vs_oracle():
    """The same crush at 50 m/s with law-2 plasticity: the folds meet from about cycle 500.
    The device's exp/pow in the law differ from the host's in the last bit, and TYPE7's
    damping switches on at a finite value when a node enters the gap, so a pair at the
    threshold can start one cycle apart: the runs stay within 1e-5 of each other (the same
    penetrating pairs every 100 cycles), where without contact they agree to 1e-13."""
    d = dk.tube_self_contact()
    m, nd, ec, ev, t, worst, peak = _tube_run(d, 15, 100)
    print(f"tube self-contact, plastic: worst displacement rel err {worst:.2e}, peak pairs {peak}")
    assert peak > 0 and m.contact.it.econt > 0.0
    assert worst < 1e-5
    assert rel(nd["v"], m.v) < 1e-4
    assert t.tt == pytest.approx(m.m.tt, rel=1e-6)      # the contact time step differs too
    assert ec == pytest.approx(m.contact.it.econt, rel=1e-4)


def _tube_alone(d, m):
    """Tube m of a batch as its own deck, with its own self-contact interface."""
    n0, n1 = d["node_off"][m], d["node_off"][m + 1]
    e0, e1 = d["elem_off"][m], d["elem_off"][m + 1]
    out = {k: d[k][n0:n1] for k in ("x", "v", "vr", "ms", "in", "icodt", "icodr", "fv_dir",
                                    "fv_val")}
    out["ixc"] = d["ixc"][e0:e1] - n0
    out["uid"] = d["uid"][e0:e1]
    out["thk"] = d["thk"][e0:e1]
    out["mat"] = d["mats"][m]
    out["prop"] = d["prop"]
    c = d["contact"]
    out["contact"] = dict(c, nsv=np.arange(n1 - n0, dtype=np.int32), irect=out["ixc"].copy(),
                          stfm=c["stfm"][e0:e1], stfns=c["stfns"][n0:n1])
    return out


@pytest.mark.parametrize("prec", [en.FP64, en.MIXED])
def test_batched_tubes_with_self_contact_equal_each_tube_alone(prec):
    """Five tubes crushed at 50 m/s in one batched context, each with self-contact (the
    tubes overlap in space; pairs never cross models): every tube's nodes, state and time
    record are bit for bit those of the tube run alone with its own interface."""
    d = dk.shell_tubes(ntube=5, side=(16.0, 24.0), length=(30.0, 50.0), h=(1.8, 2.2),
                       thk=(1.2, 1.2), ca=(200.0, 400.0), v_imp=-5.0e4, seed=11)
    d = dk.tubes_self_contact(d, gap=0.72)
    ncyc = 900
    s = en.ShellSolver(d, prec)
    try:
        s.run_cycles(ncyc)
        nd, st = s.nodes(), s.state()
        times = [s.model_time(m) for m in range(s.nmodel)]
        (ec, _, _), cnt, _ = s.contact7_state()
    finally:
        s.close()
    assert ec > 0.0
    for m in range(5):
        dm = _tube_alone(d, m)
        r = en.ShellSolver(dm, prec)
        try:
            r.run_cycles(ncyc)
            ndm, stm, tm = r.nodes(), r.state(), r.time()
        finally:
            r.close()
        n0, n1 = d["node_off"][m], d["node_off"][m + 1]
        e0, e1 = d["elem_off"][m], d["elem_off"][m + 1]
        for k in ("x", "v", "d", "vr"):
            assert np.array_equal(nd[k][n0:n1], ndm[k]), (m, k)
        assert np.array_equal(st[e0:e1], stm), m
        assert (times[m].tt, times[m].dt2, times[m].ncycle) == (tm.tt, tm.dt2, tm.ncycle), m
