"""The /INTER/TYPE7 contact oracle against the reference's own routines, plus closed
forms and invariants.

tests/golden/int7_cases.npz holds inputs and outputs of the reference routines I7COR3,
I7DST3, I7FOR3 and I7ASS2 (compiled from source, driven in I7MAINF order by
oracle/ref/harness_int7.F; written by tests/golden/make_golden_int7.py).  The oracle
restates them in C and must reproduce them bit for bit: every PARITH/ON skyline entry
(force and stiffness) and its node, which candidates penetrate, the contact time step,
the elastic and friction contact energies, and the deactivation of a node at full
penetration.
"""
import os

import numpy as np
import pytest

import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "int7_cases.npz")


def load():
    z = np.load(GOLD)
    names = sorted({k.split("/")[0] for k in z.files})
    return {n: {k.split("/")[1]: z[k] for k in z.files if k.startswith(n + "/")} for n in names}


CASES = load()


def contact(c, stfns=None):
    gap, fric, visc, viscf, _ = c["params"]
    return po.Contact7(dict(nsv=c["nsv"], irect=c["irect"], stfm=c["stfm"],
                            stfns=c["stfns"] if stfns is None else stfns, gap=gap, fric=fric,
                            visc=visc, viscf=viscf))


@pytest.mark.parametrize("name", sorted(CASES))
def test_int7_oracle_bitwise_vs_reference_routines(name):
    c = CASES[name]
    dt1 = float(c["params"][4])
    k = contact(c)
    cn, ce = k.candidates(c["x"])
    assert np.array_equal(cn, c["cand_n"]) and np.array_equal(ce, c["cand_e"])
    o = k.forces(c["x"], c["v"], c["ms"], cn, ce, dt1)
    # the library's batching: NVSIZ penetrating pairs per batch, in candidate order
    assert np.array_equal(o["isky"], c["act_isky"])
    assert np.array_equal(o["fsky"], c["act_fsky"]), np.abs(o["fsky"] - c["act_fsky"]).max()
    assert o["npen"] == int(c["full_pen"].sum())
    assert o["dt2t"] == c["act_scal"][0]
    assert o["econt"] == c["act_scal"][1] and o["econtv"] == c["act_scal"][2]
    assert np.array_equal(o["stfns"], c["full_stfns"])
    # the penetration filter itself: the kept candidates all penetrate, the others none
    pen = c["full_pen"]
    assert contact(c).forces(c["x"], c["v"], c["ms"], cn[pen], ce[pen], dt1)["npen"] == pen.sum()
    r = contact(c).forces(c["x"], c["v"], c["ms"], cn[~pen], ce[~pen], dt1)
    assert r["npen"] == 0 and len(r["isky"]) == 0 and r["dt2t"] == 1.0e30


def test_one_batch_lists_batch_as_the_reference_does():
    """With at most NVSIZ candidates the reference's own batching (over candidates,
    compressed to the penetrating ones) is the library's (over penetrating pairs)."""
    for name, c in CASES.items():
        if len(c["cand_n"]) <= 128:
            assert np.array_equal(c["full_fsky"], c["act_fsky"]), name
            assert np.array_equal(c["full_scal"], c["act_scal"]), name
    assert len(CASES["multi_batch"]["cand_n"]) > 2 * 128


def test_cases_exercise_the_branches():
    assert CASES["on_surface"]["full_stfns"].min() < 0.0        # a node deactivated
    assert np.all(CASES["frictionless"]["act_scal"][2] == 0.0)   # no friction energy
    assert CASES["friction"]["act_scal"][2] > 0.0
    tri = CASES["triangles"]["irect"]
    assert np.all(tri[:, 2] == tri[:, 3])


def test_candidates_are_the_box_test():
    c = CASES["multi_batch"]
    k = contact(c)
    tz = 0.37
    cn, ce = k.candidates(c["x"], tz)
    x = c["x"]
    exp = []
    for i, n in enumerate(c["nsv"]):
        for l, r in enumerate(c["irect"]):
            if n in r:
                continue
            lo, hi = x[r].min(axis=0) - tz, x[r].max(axis=0) + tz
            if np.all(x[n] > lo) and np.all(x[n] < hi):
                exp.append((i, l))
    assert list(zip(cn.tolist(), ce.tolist())) == exp


def single(h, w=0.0, vt=(0.0, 0.0), gap=0.2, stf=2.0e4, fric=0.0, visc=0.0, viscf=0.0,
           dt1=1.0e-6, xy=(0.5, 0.5), ms=1.0e-3):
    """One secondary node at height h over the unit square segment [0,1]^2, z = 0,
    moving with (vt, -w); the segment at rest."""
    x = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [xy[0], xy[1], h]], float)
    v = np.zeros_like(x)
    v[4] = [vt[0], vt[1], -w]
    k = po.Contact7(dict(nsv=[4], irect=[[0, 1, 2, 3]], stfm=[stf], stfns=[1.0], gap=gap,
                         fric=fric, visc=visc, viscf=viscf))
    return k.forces(x, v, np.full(5, ms), [0], [0], dt1)


def test_penalty_closed_form_over_the_centre():
    """Over the segment centre every triangle projects onto X0: the normal is z, the
    shape functions are 1/4 each, PENE = GAP - h, the force is the nonlinear penalty
    1/2 K GAP/(GAP-PENE) PENE (i7for3.F:941-991) and the elastic energy
    1/2 K GAP^2 (f - 1 - ln f), f = (GAP-PENE)/GAP; the time step 1/2 (GAP-PENE)/w."""
    gap, stf, h, w = 0.2, 2.0e4, 0.05, 3.0
    o = single(h, w=w, gap=gap, stf=stf)
    pene = gap - h
    fn = 0.5 * stf * gap / (gap - pene) * pene
    assert o["npen"] == 1 and len(o["isky"]) == 5
    assert list(o["isky"]) == [0, 1, 2, 3, 4]
    f = o["fsky"]
    assert f[4, 0] == 0.0 and f[4, 1] == 0.0
    assert f[4, 2] == pytest.approx(fn, rel=1e-14)
    assert np.allclose(f[:4, 2], -fn / 4, rtol=1e-14)
    assert np.allclose(f[:4, :2], 0.0, atol=1e-12 * fn)
    k_sec = 0.5 * stf * gap / (gap - pene) * gap / (gap - pene)
    assert f[4, 3] == pytest.approx(k_sec, rel=1e-14)
    assert np.allclose(f[:4, 3], k_sec / 4, rtol=1e-14)
    r = (gap - pene) / gap
    assert o["econt"] == pytest.approx(0.5 * stf * gap ** 2 * (r - 1 - np.log(r)), rel=1e-12)
    assert o["dt2t"] == pytest.approx(0.5 * (gap - pene) / w, rel=1e-14)


def test_damping_closed_form_over_the_centre():
    """IVIS2=0 damping adds VISC sqrt(2 K' m (GAP/(GAP-PENE))) w + 0.4^2 2 m w/(GAP-PENE)
    to the normal force of an approaching node, K' = the nonlinear stiffness
    (i7for3.F:1086-1121)."""
    gap, stf, h, w, ms, visc = 0.2, 2.0e4, 0.05, 3.0, 1.0e-3, 0.05
    o0 = single(h, w=w, gap=gap, stf=stf, ms=ms)
    o1 = single(h, w=w, gap=gap, stf=stf, ms=ms, visc=visc)
    pene = gap - h
    k1 = 0.5 * stf * gap / (gap - pene)
    vis = np.sqrt(2 * k1 * ms / ((gap - pene) / gap))
    ff = visc * vis + 0.4 ** 2 * 2 * ms * w / (gap - pene)
    assert o1["fsky"][4, 2] - o0["fsky"][4, 2] == pytest.approx(ff * w, rel=1e-10)


def test_coulomb_friction_caps_at_mu_fn():
    """Fast sliding: the tangential force is mu |FN| against the sliding velocity; slow
    sliding: the viscous force |vt| sqrt(VISf 2 K' m), K' the nonlinear stiffness
    (i7for3.F:2444-2480)."""
    gap, stf, h, mu, ms = 0.2, 2.0e4, 0.05, 0.3, 1.0e-3
    pene = gap - h
    k1 = 0.5 * stf * gap / (gap - pene)
    o0 = single(h, vt=(1000.0, 0.0), gap=gap, stf=stf, ms=ms)
    fn = o0["fsky"][4, 2]
    o = single(h, vt=(1000.0, 0.0), gap=gap, stf=stf, ms=ms, fric=mu, viscf=1.0)
    f = o["fsky"][4]
    assert f[2] == pytest.approx(fn, rel=1e-14)
    assert f[0] == pytest.approx(-mu * fn, rel=1e-12) and f[1] == 0.0
    assert o["econtv"] == pytest.approx(mu * fn * 1000.0 * 1.0e-6, rel=1e-12)
    # the segment nodes carry the reaction: the entries sum to zero
    assert np.abs(o["fsky"][:, :3].sum(axis=0)).max() < 1e-12 * fn
    o = single(h, vt=(0.0, 50.0), gap=gap, stf=stf, ms=ms, fric=mu, viscf=1.0)
    assert o["fsky"][4, 1] == pytest.approx(-50.0 * np.sqrt(2 * k1 * ms), rel=1e-12)
    assert o["fsky"][4, 0] == 0.0


def test_outside_the_gap_no_contact():
    for h in (0.2000001, 0.25, -0.21):
        o = single(h, w=3.0)
        assert o["npen"] == 0 and len(o["isky"]) == 0
    # below the segment the force points down (the gap is on both sides)
    o = single(-0.05, w=-3.0)
    assert o["fsky"][4, 2] < 0.0


@pytest.mark.parametrize("name", ["friction", "multi_batch", "triangles", "first_cycle"])
def test_every_pair_is_in_equilibrium(name):
    """The shape functions of the projection sum to 1 (after the hourglass correction), so
    each pair's main-node forces balance its secondary-node force."""
    c = CASES[name]
    f = c["act_fsky"][:, :3]
    assert np.abs(f.sum(axis=0)).max() < 1e-12 * np.abs(f).max() * len(f)
