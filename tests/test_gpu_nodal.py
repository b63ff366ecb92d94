"""The device nodal kernel (DT2 selection, ACCELE, BCS10, FIXVEL, VELOCITY, DEPLA)
against the oracle's nodal update, which tests/test_oracle_nodal.py pins bit for bit
to the reference's own routines.  Cycle by cycle, the oracle is fed the forces the
device itself assembled that cycle (orad_asspar4 re-sums the slots the cycle's
element launch wrote) and the device's new DT2 as the element minimum (the DT2
selection is idempotent on its own result), so the comparison isolates the nodal
update: it must be bitwise.  The runs cover the first cycle (DT1 = 0), the 1.1
growth limit, DTMX imposing a drop and being lifted again, fixed DOFs, FIXVEL nodes
(bricks and shells) and the rotational DOF of shells.
"""
import numpy as np
import pytest

import pyoracle as po
from openradioss_amd import deck as dk
from openradioss_amd import engine as en

pytestmark = pytest.mark.gpu


def tdict(t):
    return dict(tt=t.tt, dt1=t.dt1, dt2=t.dt2, dt12=t.dt12, dt2old=t.dt2old, dtmx=t.dtmx,
                wfext=0.0, ncycle=t.ncycle)


def check_cycles(s, deck, rot, ncyc, dtmx_at):
    """dtmx_at: {cycle: factor}; before that cycle DTMX becomes factor x the current DT2
    (a factor above 1e10 lifts it)."""
    for c in range(ncyc):
        if c in dtmx_at:
            f = dtmx_at[c]
            s.set_time(dtmx=1.0e20 if f > 1e10 else f * s.time().dt2)
        before, t0 = s.nodes(), tdict(s.time())
        t1 = s.run_cycles(1)
        after = s.nodes()
        if rot:
            a, ar, _, _ = s.asspar4()
        else:
            a, _ = s.asspar4()
            ar = None
        nodes = dict(x=before["x"], v=before["v"], d=before["d"], a=a, ms=deck["ms"],
                     icodt=deck["icodt"])
        if deck.get("fv_dir") is not None:
            nodes.update(fv_dir=deck["fv_dir"], fv_val=deck["fv_val"])
        if rot:
            nodes.update(vr=before["vr"], ar=ar, icodr=deck["icodr"], **{"in": deck["in"]})
        out, t = po.nodal_update(nodes, t0, t1.dt2, rot=rot)
        print(f"cycle {c}: dt1 {t['dt1']:.6e} dt2 {t['dt2']:.6e} dt12 {t['dt12']:.6e}",
              flush=True)
        for k in ("x", "v", "d") + (("vr",) if rot else ()):
            assert np.array_equal(after[k], out[k]), (c, k, np.abs(after[k] - out[k]).max())
        for k in ("tt", "dt1", "dt2", "dt12", "dt2old"):
            assert getattr(t1, k) == t[k], (c, k, getattr(t1, k), t[k])
        assert t1.ncycle == t["ncycle"]


@pytest.mark.parametrize("prec", [en.FP64, en.MIXED])
def test_brick_node_kernel_bitwise_vs_oracle_nodal_update(prec):
    d = dk.brick_block(5, 4, 6, v_top=-5.0e4)
    nn = d["x"].shape[0]
    rng = np.random.default_rng(7)
    fv_dir = np.zeros(nn, np.int32)
    fv_val = np.zeros(nn)
    top = np.flatnonzero(d["x"][:, 2] == d["x"][:, 2].max())
    fv_dir[top[::3]] = rng.integers(1, 4, top[::3].size)
    fv_val[top[::3]] = rng.uniform(-2.0e4, 2.0e4, top[::3].size)
    d["fv_dir"], d["fv_val"] = fv_dir, fv_val
    s = en.BrickSolver(d, prec)
    try:
        # cycle 0: DT1 = 0; 1: growth limit; 4: DTMX below the element step; 6: lifted
        check_cycles(s, d, False, 9, {4: 0.8, 6: 1.0e20})
        t = s.time()
        assert t.ncycle == 9
    finally:
        s.close()


@pytest.mark.parametrize("ishell", [1, 24])
def test_shell_node_kernel_bitwise_vs_oracle_nodal_update(ishell):
    d = dk.shell_plate(n=10, patch=4)
    d["prop"] = dict(d["prop"], ishell=ishell)
    nn = d["x"].shape[0]
    fv_dir = np.zeros(nn, np.int32)
    fv_val = np.zeros(nn)
    fv_dir[nn // 2: nn // 2 + 6] = [1, 2, 3, 3, 2, 1]
    fv_val[nn // 2: nn // 2 + 6] = np.linspace(-8.0e3, 8.0e3, 6)
    d["fv_dir"], d["fv_val"] = fv_dir, fv_val
    s = en.ShellSolver(d, en.FP64)
    try:
        check_cycles(s, d, True, 8, {3: 0.7, 5: 1.0e20})
    finally:
        s.close()
