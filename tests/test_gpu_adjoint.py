"""Adjoint of the shell cycle through the C ABI, wrapped as a torch.autograd.Function,
checked against central finite differences of forward runs.

Setup (all tests): the time step is pinned by DTMX at 0.8 x the initial element
time step, so the time-step sequence does not depend on the parameters; masses and
rotational inertias are torch functions of the thicknesses (CINMAS), so their
sensitivities reach the thickness gradient.

Objective J = final z displacement of the centre node + total plastic work, the
plastic work of a point being V_pt * (A*pla + B*pla^(n+1)/(n+1)) with
V_pt = wf_pt * thickness * initial area.

Bars:
  * short runs (checkpoint boundaries inside the run): adjoint vs central FD with
    h = 1e-6 relative, relative error < 1e-5 on A, B, n, thicknesses and an
    initial velocity.  FD converges to the adjoint as h falls (1e-3 .. 1e-7); at
    h = 1e-4 a difference quotient can straddle the cycle at which a point starts
    to yield (the radial return is not differentiable there), which after 60
    cycles of the 16 x 16 plate leaves up to 1% in the quotient of A;
  * 128 x 128 plate, 2000 cycles, checkpoint every 100 (the adjoint study, FP32
    forward as stated and FP64 forward): relative error < 1e-3 against FD of FP64
    forward runs with h = 1e-6 on 20 sampled parameters (A, B, n and 17
    thicknesses).  The quotients with h = 1e-4 are printed beside them: over 2000
    cycles so many points start to yield inside a 1e-4 step that they are not
    derivative estimates (measured: 1% to 5x away from both the adjoint and the
    h = 1e-6 quotients).  FD of the FP32 forward itself is not a derivative
    estimate either: its rounding (about 1e-7 relative per cycle) is the size of
    the perturbation;
  * QEPH (Ishell=24): the same short runs at the same bar; over 2000 cycles a
    twisted elastic plate (test_adjoint_qeph_plate_128_vs_fd says why);
  * FP32-element forward with the FP64 adjoint vs the all-FP64 gradient: printed;
  * the hand-written reverse passes of the Belytschko-Tsay and QEPH elements (the
    default) vs the forward-mode dual-number VJP (adj_vjp=1) on the same forward
    run: every output within 1e-10 of the largest of its kind (16 x 16 with
    plasticity and random adjoints of the whole final state, both forms; BT with
    triangles, QEPH flat and twisted; the 128 x 128 adjoint study for both).
"""
import numpy as np
import pytest
import torch

import pyoracle as po
from openradioss_amd import deck as dk
from openradioss_amd import engine as en
from openradioss_amd.autograd import ShellRun, cinmas, displacement_plus_plastic_work, force_weights

pytestmark = pytest.mark.gpu


masses_torch = cinmas


def objective(deck, theta, prec, ncyc, every, dtmx, centre, wf):
    """theta = (A, B, n, thk) tensors; returns J (a 0-d tensor)."""
    A, B, n, thk = theta
    return displacement_plus_plastic_work(deck, A, B, n, thk, prec, ncyc, every, dtmx, centre)


def forward_only(deck, theta, ncyc, dtmx, centre, wf):
    """J from a plain FP64 forward run (no adjoint bookkeeping)."""
    A, B, n, thk = [float(v) if np.ndim(v) == 0 else np.asarray(v) for v in theta]
    x = torch.from_numpy(deck["x"])
    ixc = torch.from_numpy(deck["ixc"].astype(np.int64))
    ms, inr, area0 = masses_torch(x, ixc, torch.from_numpy(thk), deck["mat"]["rho0"])
    d = dict(deck, thk=thk, ms=ms.numpy(), **{"in": inr.numpy()})
    d["mat"] = dict(deck["mat"], ca=A, cb=B, cn=n)
    s = en.ShellSolver(d, en.FP64)
    s.set_time(dtmx=dtmx)
    t = s.run_cycles(ncyc)
    nd, st = s.nodes(), s.state()
    s.close()
    pla = st[:, po.SH_PLA0:po.SH_PLA0 + 5]
    w = np.where(pla > 0, A * pla + B * np.where(pla > 0, pla, 1.0) ** (n + 1.0) / (n + 1.0), 0.0)
    vpt = wf[None, :] * (thk * area0.numpy())[:, None]
    return nd["d"][centre, 2] + (vpt * w).sum(), t.tt


def pinned_dt(deck):
    s = en.ShellSolver(deck, en.FP64)
    dt = s.cforc3(0, s.numelc, 0.0)
    s.close()
    return 0.8 * dt


def setup_plate(n, patch, thk_noise=0.1, seed=20261203, ishell=1, twist=0.0, ca_scale=1.0):
    """The clamped plate; `twist` c lifts it to the hyperbolic paraboloid
    z = c (x - n/2)(y - n/2), which gives every element the same warp c/4."""
    deck = dk.shell_plate(n=n, patch=patch, thk_noise=thk_noise, seed=seed)
    deck["prop"] = dict(deck["prop"], ishell=ishell)
    deck["mat"] = dict(deck["mat"], ca=deck["mat"]["ca"] * ca_scale)
    if twist:
        x = deck["x"].copy()
        x[:, 2] = twist * (x[:, 0] - 0.5 * n) * (x[:, 1] - 0.5 * n)
        deck["x"] = x
    ni = n + 1
    centre = (n // 2) + ni * (n // 2)
    _, wf, _ = po.shell_tables(deck["prop"]["npt"])
    return deck, centre, wf, pinned_dt(deck)


def gradients(deck, prec, ncyc, every, dtmx, centre, wf):
    m = deck["mat"]
    A = torch.tensor(m["ca"], dtype=torch.float64, requires_grad=True)
    B = torch.tensor(m["cb"], dtype=torch.float64, requires_grad=True)
    n = torch.tensor(m["cn"], dtype=torch.float64, requires_grad=True)
    thk = torch.tensor(deck["thk"], dtype=torch.float64, requires_grad=True)
    J = objective(deck, (A, B, n, thk), prec, ncyc, every, dtmx, centre, wf)
    J.backward()
    return float(J.detach()), np.array([A.grad.item(), B.grad.item(), n.grad.item()]), thk.grad.numpy()


def fd(deck, which, idx, h, ncyc, dtmx, centre, wf):
    m = deck["mat"]
    base = [m["ca"], m["cb"], m["cn"], deck["thk"].copy()]
    vals = []
    for sgn in (1.0, -1.0):
        th = [base[0], base[1], base[2], base[3].copy()]
        if which < 3:
            step = h * abs(base[which])
            th[which] = base[which] + sgn * step
        else:
            step = h * abs(base[3][idx])
            th[3][idx] = base[3][idx] + sgn * step
        J, tt = forward_only(deck, th, ncyc, dtmx, centre, wf)
        vals.append(J)
    return (vals[0] - vals[1]) / (2.0 * step), tt


@pytest.mark.parametrize("ishell", [1, 24])
def test_adjoint_short_runs_vs_fd(ishell):
    deck, centre, wf, dtmx = setup_plate(n=16, patch=6, ishell=ishell)
    ncyc, every = 60, 7                         # segment boundaries inside the run
    J, gjc, gthk = gradients(deck, en.FP64, ncyc, every, dtmx, centre, wf)
    J0, tt = forward_only(deck, [deck["mat"]["ca"], deck["mat"]["cb"], deck["mat"]["cn"],
                                 deck["thk"]], ncyc, dtmx, centre, wf)
    assert J == pytest.approx(J0, rel=1e-13)
    assert tt == pytest.approx(ncyc * dtmx, rel=1e-12)       # the time step stayed pinned
    rng = np.random.default_rng(1)
    elems = [int(i) for i in rng.choice(deck["ixc"].shape[0], 6, replace=False)]
    for which, idx, g in [(0, 0, gjc[0]), (1, 0, gjc[1]), (2, 0, gjc[2])] + \
                         [(3, e, gthk[e]) for e in elems]:
        gf, _ = fd(deck, which, idx, 1e-6, ncyc, dtmx, centre, wf)
        assert abs(g - gf) <= 1e-5 * max(abs(gf), 1e-3 * np.abs(gthk).max()), (which, idx, g, gf)


@pytest.mark.parametrize("ishell", [1, 24])
def test_adjoint_initial_velocity(ishell):
    deck, centre, wf, dtmx = setup_plate(n=12, patch=4, ishell=ishell)
    s = en.ShellSolver(deck, en.FP64)
    s.set_time(dtmx=dtmx)
    s.adjoint_forward(40, 9)
    lam_d = np.zeros((deck["x"].shape[0], 3))
    lam_d[centre, 2] = 1.0
    g = s.adjoint_reverse(lam_d=lam_d)
    s.close()
    node = centre + 1
    h = 1e-6 * abs(deck["v"][node, 2])
    vals = []
    for sgn in (1.0, -1.0):
        d = dict(deck, v=deck["v"].copy())
        d["v"][node, 2] += sgn * h
        s = en.ShellSolver(d, en.FP64)
        s.set_time(dtmx=dtmx)
        s.run_cycles(40)
        vals.append(s.nodes()["d"][centre, 2])
        s.close()
    gf = (vals[0] - vals[1]) / (2 * h)
    print(f"d(centre z)/d(v0 z): adjoint {g['lam_v0'][node, 2]:.9e} fd {gf:.9e}")
    assert g["lam_v0"][node, 2] == pytest.approx(gf, rel=1e-5)


def sample_params(deck, k_patch=9, k_any=8, seed=20261203):
    rng = np.random.default_rng(seed)
    ne = deck["ixc"].shape[0]
    n = int(round(np.sqrt(ne)))
    patch_elems = np.flatnonzero(np.abs(deck["x"][deck["ixc"][:, 0], 0] - 0.5 * n) < 18.0)
    elems = list(rng.choice(patch_elems, k_patch, replace=False)) + \
        list(rng.choice(ne, k_any, replace=False))
    return [(0, 0), (1, 0), (2, 0)] + [(3, int(e)) for e in elems]


def adjoint_vs_fd(deck, centre, wf, dtmx, params, precs, hs, ncyc=2000, every=100):
    """Adjoint gradients under each forward precision against central FD of FP64 forward
    runs; returns {(prec, h): [relative errors]} and prints one line per parameter."""
    grads = {}
    for prec in precs:
        _, gjc, gthk = gradients(deck, prec, ncyc, every, dtmx, centre, wf)
        grads[prec] = (gjc, gthk)
    pick = lambda prec, which, idx: (grads[prec][0][which] if which < 3 else grads[prec][1][idx])
    errs = {(prec, h): [] for prec in precs for h in hs}
    name = {en.MIXED: "FP32-fwd", en.FP64: "FP64-fwd"}
    for which, idx in params:
        line = f"param {which}:{idx} adjoint " + " ".join(
            f"{name[p]} {pick(p, which, idx):.6e}" for p in precs)
        for h in hs:
            gf, tt = fd(deck, which, idx, h, ncyc, dtmx, centre, wf)
            assert tt == pytest.approx(ncyc * dtmx, rel=1e-12)
            for prec in precs:
                g = pick(prec, which, idx)
                errs[(prec, h)].append(abs(g - gf) / max(abs(gf), 1e-30))
            line += f" | fd(h={h:g}) {gf:.6e} rel " + "/".join(
                f"{errs[(p, h)][-1]:.2e}" for p in precs)
        print(line, flush=True)
    for (prec, h), e in errs.items():
        print(f"{name[prec]}, h={h:g}: max relative error over {len(e)} parameters "
              f"{max(e):.2e}, median {np.median(e):.2e}", flush=True)
    return errs, grads


def test_adjoint_plate_128_vs_fd():
    """The adjoint study: 128 x 128 BT plate, 2000 cycles, checkpoint every 100,
    J = centre z displacement + plastic work, 20 sampled parameters (A, B, n, 17
    thicknesses) against central FD of FP64 forward runs at h = 1e-4 (printed) and
    h = 1e-6.  FP64-forward adjoint: every parameter within 1e-3.  FP32-element
    forward with the FP64 adjoint: A, B, n within 1e-2; its thickness gradients are
    printed -- they are the exact derivative of the FP32 run (the FP64 chain takes the
    FP32 run's plastic/elastic branch at every point), but over 2000 cycles the FP32
    and FP64 runs differ in which of the ~200 yielding points yield when, and a
    single thickness's gradient follows that (test_adjoint_plate_128_fp32_elastic_vs_fd
    shows the FP32 forward within the bar once no point yields)."""
    deck, centre, wf, dtmx = setup_plate(n=128, patch=32)
    errs, _ = adjoint_vs_fd(deck, centre, wf, dtmx, sample_params(deck), (en.MIXED, en.FP64),
                            (1e-4, 1e-6))
    assert max(errs[(en.FP64, 1e-6)]) < 1e-3
    assert max(errs[(en.MIXED, 1e-6)][:3]) < 1e-2


def test_adjoint_plate_128_fp32_elastic_vs_fd():
    """The same study with a yield stress 1000x the plate's (no point yields): the
    FP32-forward adjoint within 1e-3 of FD (FP64 forward, h = 1e-6) on 17 thicknesses."""
    deck, centre, wf, dtmx = setup_plate(n=128, patch=32, ca_scale=1.0e3)
    params = sample_params(deck)[3:]
    errs, grads = adjoint_vs_fd(deck, centre, wf, dtmx, params, (en.MIXED, en.FP64), (1e-6,))
    assert np.all(grads[en.MIXED][0] == 0.0)
    assert max(errs[(en.MIXED, 1e-6)]) < 1e-3
    assert max(errs[(en.FP64, 1e-6)]) < 1e-3


@pytest.mark.parametrize("ishell", [1, 24])
def test_adjoint_mixed_forward_close_to_fp64(ishell):
    deck, centre, wf, dtmx = setup_plate(n=32, patch=8, ishell=ishell)
    _, g64, t64 = gradients(deck, en.FP64, 300, 50, dtmx, centre, wf)
    _, g32, t32 = gradients(deck, en.MIXED, 300, 50, dtmx, centre, wf)
    ej = np.abs(g32 - g64) / np.abs(g64)
    et = np.linalg.norm(t32 - t64) / np.linalg.norm(t64)
    print(f"FP32-forward gradient vs FP64: A,B,n {ej}, thickness L2 {et:.2e}")
    assert ej.max() < 1e-2 and et < 1e-2


def test_adjoint_qeph_plate_128_vs_fd():
    """QEPH, 128 x 128, 2000 cycles, checkpoint every 100, on a plate twisted to
    z = 0.005 (x - 64)(y - 64) with a yield stress 1000x the plate's, so no point
    yields: relative error < 1e-3 against FD with h = 1e-5 and 1e-6 on 17
    thicknesses (measured on MI355X: 5.8e-6 and 8.4e-5 at most).

    A flat or a yielding QEPH plate has no such check over 2000 cycles, because
    the CZFORC3 cycle is discontinuous in its inputs there, and FD quotients of
    it are not derivative estimates (measured: off by 1 to 1e4x at h = 1e-5 ..
    1e-7 on both):
      * CZCORP5 switches an element between the flat and the warped projection
        when Z1^2 crosses LM * 1e-8 (czcorp5.F:84-88), and the two differ by a
        finite force at the switch; a flat plate starts at Z1 = 0 and its bending
        moves thousands of elements across it; the twist keeps every element at
        |Z1| >= 7.6e-4, 5x above the threshold, for the whole run;
      * CZFINTN1 cuts the hourglass increment by EH = max(.., 1 - ZCFAC) once
        a point has yielded or the hourglass stress passes the yield surface
        (czfintn.F:303-345), ZCFAC being the mean and the minimum of the points'
        ET/E (1 elastic, H/(H+E) plastic, m2cplr.F:99,227); it jumps whenever a
        point starts or stops yielding.
    The 60-cycle runs above (plastic, flat) stay clear of both and pass at 1e-5."""
    deck, centre, wf, dtmx = setup_plate(n=128, patch=32, ishell=24, twist=0.005,
                                         ca_scale=1.0e3)
    ncyc, every = 2000, 100
    J, gjc, gthk = gradients(deck, en.FP64, ncyc, every, dtmx, centre, wf)
    assert np.all(gjc == 0.0)                   # nothing yields
    rng = np.random.default_rng(20261203)
    ne = deck["ixc"].shape[0]
    patch_elems = np.flatnonzero(np.abs(deck["x"][deck["ixc"][:, 0], 0] - 64.0) < 18.0)
    elems = list(rng.choice(patch_elems, 9, replace=False)) + list(rng.choice(ne, 8, replace=False))
    errs = {1e-5: [], 1e-6: []}
    for e in elems:
        line = f"thickness {int(e)} adjoint {gthk[int(e)]:.6e}"
        for h in errs:
            gf, tt = fd(deck, 3, int(e), h, ncyc, dtmx, centre, wf)
            errs[h].append(abs(gthk[int(e)] - gf) / max(abs(gf), 1e-30))
            line += f" | fd(h={h:g}) {gf:.6e} rel {errs[h][-1]:.2e}"
            assert tt == pytest.approx(ncyc * dtmx, rel=1e-12)
        print(line, flush=True)
    for h, e in errs.items():
        print(f"h={h:g}: max relative error over {len(e)} thicknesses {max(e):.2e}")
        assert max(e) < 1e-3


def reverse_vs_duals(deck, prec, ncyc, every, dtmx, seeds):
    """adjoint_reverse outputs under the reverse pass (adj_vjp=0) and the duals (1)."""
    outs = []
    for vjp in (0, 1):
        s = en.ShellSolver(deck, prec)
        s.set_time(dtmx=dtmx)
        s.set_launch(adj_vjp=vjp)
        assert s.launch()["adj_vjp"] == vjp
        s.adjoint_forward(ncyc, every)
        outs.append(s.adjoint_reverse(**seeds))
        s.close()
    return outs


def assert_same_gradients(a, b, bar, what):
    for k in a:
        scale = np.abs(b[k]).max()
        err = np.abs(a[k] - b[k]).max() / max(scale, 1e-300)
        print(f"{what} {k}: max |reverse - duals| / max|duals| = {err:.2e} (scale {scale:.3e})",
              flush=True)
        assert err < bar, (what, k, err)


@pytest.mark.parametrize("prec", [en.FP64, en.MIXED])
def test_reverse_pass_matches_dual_vjp(prec):
    deck, centre, wf, dtmx = setup_plate(n=16, patch=6, ca_scale=0.5)
    ixc = deck["ixc"].copy()
    for e in (3, 100, 201):
        ixc[e, 3] = ixc[e, 2]                  # three triangles (corner 4 = corner 3)
    deck["ixc"] = ixc
    nn, ne = deck["x"].shape[0], ixc.shape[0]
    rng = np.random.default_rng(11)
    seeds = dict(lam_x=rng.standard_normal((nn, 3)), lam_v=1e-4 * rng.standard_normal((nn, 3)),
                 lam_vr=1e-4 * rng.standard_normal((nn, 3)), lam_d=rng.standard_normal((nn, 3)),
                 lam_state=1e-3 * rng.standard_normal((ne, 33)))
    rev, dual = reverse_vs_duals(deck, prec, 60, 7, dtmx, seeds)
    assert np.all(rev["g_jc"] != 0.0)          # points yielded
    assert_same_gradients(rev, dual, 1e-10, f"16x16 prec {prec}")


@pytest.mark.parametrize("prec", [en.FP64, en.MIXED])
@pytest.mark.parametrize("twist", [0.0, 0.01])
def test_reverse_pass_matches_dual_vjp_qeph(prec, twist):
    """QEPH: the hand-written reverse pass of CZFORC3 (the default) vs the dual-number
    VJP on the same forward run, random adjoints of the whole final state (SIG, PLA,
    HOURG(1:12)).  Flat (the flat projection) and twisted (CZCORP5's warped
    projection, every element at the same warp); points yield, so the hourglass
    limiter's ZCFAC and SIGY terms carry derivatives."""
    deck, centre, wf, dtmx = setup_plate(n=16, patch=6, ishell=24, twist=twist, ca_scale=0.5)
    nn, ne = deck["x"].shape[0], deck["ixc"].shape[0]
    rng = np.random.default_rng(12)
    seeds = dict(lam_x=rng.standard_normal((nn, 3)), lam_v=1e-4 * rng.standard_normal((nn, 3)),
                 lam_vr=1e-4 * rng.standard_normal((nn, 3)), lam_d=rng.standard_normal((nn, 3)),
                 lam_state=1e-3 * rng.standard_normal((ne, 42)))
    rev, dual = reverse_vs_duals(deck, prec, 60, 7, dtmx, seeds)
    assert np.all(rev["g_jc"] != 0.0)          # points yielded
    assert_same_gradients(rev, dual, 1e-10, f"QEPH 16x16 twist {twist} prec {prec}")


@pytest.mark.parametrize("ishell", [1, 24])
def test_reverse_pass_matches_dual_vjp_plate_128(ishell):
    deck, centre, wf, dtmx = setup_plate(n=128, patch=32, ishell=ishell)
    lam_d = np.zeros((deck["x"].shape[0], 3))
    lam_d[centre, 2] = 1.0
    rev, dual = reverse_vs_duals(deck, en.FP64, 2000, 100, dtmx, dict(lam_d=lam_d))
    assert_same_gradients(rev, dual, 1e-10, f"128x128 2000 cycles ishell {ishell}")


def test_adjoint_refuses_rate_dependent_law_and_second_backward():
    """C != 0 makes the rate filter EPSDP and the thickness THK part of the forces,
    which the adjoint state does not carry: the forward run is refused.  A second
    backward through the same run raises a clear error (the checkpoints are released
    after the first)."""
    deck, centre, wf, dtmx = setup_plate(8, 2)
    rate = dict(deck)
    rate["mat"] = dict(deck["mat"], cc=0.05, epdr=1.0)
    s = en.ShellSolver(rate, en.FP64)
    with pytest.raises(en.OradError) as ei:
        s.adjoint_forward(4, 2)
    assert ei.value.code == en.ORAD_EINVAL and "C != 0" in str(ei.value)
    s.close()
    m = deck["mat"]
    A = torch.tensor(m["ca"], dtype=torch.float64, requires_grad=True)
    B = torch.tensor(m["cb"], dtype=torch.float64)
    n = torch.tensor(m["cn"], dtype=torch.float64)
    thk = torch.tensor(deck["thk"], dtype=torch.float64)
    x = torch.from_numpy(deck["x"])
    ixc = torch.from_numpy(deck["ixc"].astype(np.int64))
    ms, inr, _ = masses_torch(x, ixc, thk, m["rho0"])
    d, pla = ShellRun.apply(A, B, n, thk, ms, inr, deck, en.FP64, 6, 3, dtmx)
    J = d[centre, 2] + pla.sum()
    J.backward(retain_graph=True)
    assert np.isfinite(A.grad.item())
    with pytest.raises(RuntimeError, match="already"):
        J.backward()
