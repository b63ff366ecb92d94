"""Adjoint of the brick cycle (orad_adjoint_forward / orad_adjoint_reverse on a brick
context) against central finite differences of forward runs, and the hand-written
reverse pass of the brick chain against the dual-number VJP of the same chain.

Setup: a 10 x 8 x 6 block whose top nodes start downward at 50 m/s, so the bricks yield
within the run; the time step is pinned by DTMX at 0.8 x the initial element time step,
so the time-step sequence does not depend on the parameters.  Objective J = final z
displacement of a top node + sum over the bricks of w_e * PLA_e (w_e = V0_e * 1e3).

Bars:
  * FP64 forward, 60 cycles, checkpoint every 7: adjoint vs central FD with h = 1e-6
    relative, relative error < 1e-5 on A, B, n, an initial velocity and a nodal mass;
  * reverse pass vs dual numbers (launch field adj_vjp = 1) on the same forward run,
    FP64 and mixed forward: every output within 1e-10 of the largest of its kind;
  * FP32-element forward with the FP64 adjoint vs the all-FP64 gradient: A, B, n within
    1e-3 (printed);
  * C != 0 is refused.
"""
import numpy as np
import pytest

from openradioss_amd import deck as dk
from openradioss_amd import engine as en

pytestmark = pytest.mark.gpu

NCYC, EVERY = 60, 7


def block():
    d = dk.brick_block(10, 8, 6, v_top=-5.0e4)
    s = en.BrickSolver(d, en.FP64)
    dtmx = 0.8 * s.sforc3(0, s.numels, 0.0)
    s.close()
    top = np.flatnonzero(d["x"][:, 2] > d["x"][:, 2].max() - 1e-6)
    node = int(top[len(top) // 2])
    w = d["vol0"] * 1.0e3
    return d, dtmx, node, w


def objective(d, dtmx, node, w, prec=en.FP64):
    s = en.BrickSolver(d, prec)
    s.set_time(dtmx=dtmx)
    t = s.run_cycles(NCYC)
    J = s.nodes()["d"][node, 2] + float(np.dot(w, s.state()["pla"]))
    s.close()
    return J, t.tt


def adjoint(d, dtmx, node, w, prec=en.FP64, vjp=0):
    s = en.BrickSolver(d, prec)
    s.set_time(dtmx=dtmx)
    s.set_launch(adj_vjp=vjp)
    s.adjoint_forward(NCYC, EVERY)
    lam_d = np.zeros((d["x"].shape[0], 3))
    lam_d[node, 2] = 1.0
    lam_state = np.zeros((d["ixs"].shape[0], 7))
    lam_state[:, 6] = w
    g = s.adjoint_reverse(lam_d=lam_d, lam_state=lam_state)
    pla = s.state()["pla"]
    s.close()
    return g, pla


def test_brick_adjoint_vs_fd():
    d, dtmx, node, w = block()
    g, pla = adjoint(d, dtmx, node, w)
    assert pla.max() > 1e-4                      # the top layers yield
    J0, tt = objective(d, dtmx, node, w)
    assert tt == pytest.approx(NCYC * dtmx, rel=1e-12)
    vnode = node - 3
    mnode = node - 1
    cases = [("A", g["g_jc"][0]), ("B", g["g_jc"][1]), ("n", g["g_jc"][2]),
             ("v0", g["lam_v0"][vnode, 2]), ("ms", g["g_ms"][mnode])]
    for name, ga in cases:
        vals = []
        for sgn in (1.0, -1.0):
            dd = dict(d, mat=dict(d["mat"]), v=d["v"].copy(), ms=d["ms"].copy())
            if name in ("A", "B", "n"):
                key = {"A": "ca", "B": "cb", "n": "cn"}[name]
                step = 1e-6 * abs(d["mat"][key])
                dd["mat"][key] = d["mat"][key] + sgn * step
            elif name == "v0":
                step = 1e-6 * max(abs(d["v"][vnode, 2]), 1.0e3)
                dd["v"][vnode, 2] += sgn * step
            else:
                step = 1e-6 * d["ms"][mnode]
                dd["ms"][mnode] += sgn * step
            vals.append(objective(dd, dtmx, node, w)[0])
        gf = (vals[0] - vals[1]) / (2.0 * step)
        print(f"dJ/d{name}: adjoint {ga:.9e} fd {gf:.9e}", flush=True)
        assert abs(ga - gf) <= 1e-5 * abs(gf) + 1e-12, (name, ga, gf)


@pytest.mark.parametrize("prec", [en.FP64, en.MIXED])
def test_brick_reverse_pass_matches_duals(prec):
    d, dtmx, node, w = block()
    rev, _ = adjoint(d, dtmx, node, w, prec, 0)
    dual, _ = adjoint(d, dtmx, node, w, prec, 1)
    for k in rev:
        scale = np.abs(dual[k]).max()
        err = np.abs(rev[k] - dual[k]).max() / max(scale, 1e-300)
        print(f"prec {prec} {k}: max |reverse - duals| / max|duals| = {err:.2e}", flush=True)
        assert err < 1e-10, (k, err)


def test_brick_mixed_forward_close_to_fp64():
    d, dtmx, node, w = block()
    g64, _ = adjoint(d, dtmx, node, w, en.FP64)
    g32, _ = adjoint(d, dtmx, node, w, en.MIXED)
    e = np.abs(g32["g_jc"] - g64["g_jc"]) / np.abs(g64["g_jc"])
    ev = np.linalg.norm(g32["lam_v0"] - g64["lam_v0"]) / np.linalg.norm(g64["lam_v0"])
    print(f"FP32-forward brick gradient vs FP64: A, B, n {e}, initial velocity L2 {ev:.2e}")
    assert e.max() < 1e-3


def test_brick_adjoint_refuses_rate_dependent_law():
    d, dtmx, node, w = block()
    d["mat"] = dict(d["mat"], cc=0.05, israte=1)
    s = en.BrickSolver(d, en.FP64)
    with pytest.raises(en.OradError) as ei:
        s.adjoint_forward(4, 2)
    assert ei.value.code == en.ORAD_EINVAL
    s.close()
