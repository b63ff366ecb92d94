"""ECRIT's energy balance on the device (orad_energy_balance) against the oracle's,
which tests/test_oracle_nodal.py pins with closed forms (rigid translation, imposed
velocity jump) on top of the bitwise-pinned nodal update and FIXVEL work.

Bars: FP64 form vs oracle, kinetic / internal / external work within 1e-9 relative
and ERR within 1e-9 absolute (the device sums per workgroup and per node where the
oracle sums in node order); the imposed-velocity jump closes to 1e-12; the mixed form's
ERR within 1e-4 of the FP64 form's.  A tube crushed at 10 m/s states its ERR.
"""
import numpy as np
import pytest

import pyoracle as po
from openradioss_amd import deck as dk
from openradioss_amd import engine as en

pytestmark = pytest.mark.gpu

KEYS = ("encin", "enrot", "enint", "wfext", "entot")


def close(g, o, what):
    for k in KEYS:
        scale = max(abs(o["entot"]), 1e-300)
        assert abs(g[k] - o[k]) <= 1e-9 * scale, (what, k, g[k], o[k])
    assert abs(g["err"] - o["err"]) < 1e-9, (what, g["err"], o["err"])
    assert g["ncycle"] == o["ncycle"]


def brick_deck():
    d = dk.brick_block(10, 8, 12, v_top=-5.0e3)
    nn = d["x"].shape[0]
    top = np.flatnonzero(d["x"][:, 2] > d["x"][:, 2].max() - 0.5)
    d["fv_dir"] = np.zeros(nn, np.int32)
    d["fv_val"] = np.zeros(nn)
    d["fv_dir"][top] = 3
    d["fv_val"][top] = -1.0e4               # driven down harder than it started
    return d


def test_brick_balance_vs_oracle():
    d = brick_deck()
    s = en.BrickSolver(d, en.FP64)
    m = po.Model(d)
    for n in (1, 49, 150):
        s.run_cycles(n)
        m.cycle(n)
        g, o = s.balance(), m.balance()
        print(f"cycle {g['ncycle']}: ERR {g['err']:.3e} (oracle {o['err']:.3e}) WFEXT "
              f"{g['wfext']:.4e} ENCIN {g['encin']:.4e} ENINT {g['enint']:.4e}", flush=True)
        close(g, o, n)
    assert g["wfext"] > 0.0 and g["entot0"] == pytest.approx(m.entot0, rel=1e-12)
    s.close()


def test_shell_balance_vs_oracle():
    d = dk.shell_plate(n=16, patch=6)
    nn = d["x"].shape[0]
    d["fv_dir"] = np.zeros(nn, np.int32)
    d["fv_val"] = np.zeros(nn)
    mid = nn // 2
    d["fv_dir"][mid - 3: mid + 3] = 3
    d["fv_val"][mid - 3: mid + 3] = 8.0e3
    for ishell in (1, 24):
        d["prop"] = dict(d["prop"], ishell=ishell)
        s = en.ShellSolver(d, en.FP64)
        m = po.ShellModel(d)
        for n in (1, 99, 200):
            s.run_cycles(n)
            m.cycle(n)
            g, o = s.balance(), m.balance()
            close(g, o, (ishell, n))
        print(f"Ishell={ishell}: ERR {g['err']:.3e} ENROT {g['enrot']:.3e} WFEXT {g['wfext']:.3e}",
              flush=True)
        assert g["enrot"] > 0.0
        s.close()


def test_imposed_velocity_jump_closes():
    d = dk.brick_block(3, 2, 2, v_top=0.0)
    nn = d["x"].shape[0]
    d["icodt"][:] = 0
    d["v"][:, 2] = -2.0e3
    d["fv_dir"] = np.full(nn, 3, np.int32)
    d["fv_val"] = np.full(nn, -1.0e4)
    for prec in (en.FP64, en.MIXED):
        s = en.BrickSolver(d, prec)
        s.run_cycles(7)
        b = s.balance()
        M = d["ms"].sum()
        assert b["wfext"] == pytest.approx(0.5 * M * (1.0e8 - 4.0e6), rel=1e-12)
        assert abs(b["err"]) < 1e-12, b
        s.close()


def test_mixed_balance_close_to_fp64():
    d = brick_deck()
    errs = {}
    for prec in (en.FP64, en.MIXED):
        s = en.BrickSolver(d, prec)
        s.run_cycles(300)
        errs[prec] = s.balance()
        s.close()
    assert abs(errs[en.MIXED]["err"] - errs[en.FP64]["err"]) < 1e-4
    assert errs[en.MIXED]["wfext"] == pytest.approx(errs[en.FP64]["wfext"], rel=1e-4)


def test_tube_crush_balance_is_stated_and_batches_are_refused():
    t = dk.shell_tubes(ntube=2, seed=5)
    n1, e1 = t["node_off"][1], t["elem_off"][1]
    keys = ("x", "v", "vr", "ms", "in", "icodt", "icodr", "fv_dir", "fv_val")
    one = {k: t[k][:n1] for k in keys}
    one.update(ixc=t["ixc"][:e1], uid=t["uid"][:e1], thk=t["thk"][:e1], mat=t["mats"][0],
               prop=t["prop"])
    s = en.ShellSolver(one, en.FP64)
    out = []
    for _ in range(6):
        s.run_cycles(500)
        b = s.balance()
        out.append(b["err"])
    print(f"tube crush, {e1} shells, 3000 cycles: ERR " +
          " ".join(f"{e:.2e}" for e in out) + f"; WFEXT {b['wfext']:.3e} "
          f"EINT {b['enint']:.3e} EHOUR {b['ehour']:.3e}", flush=True)
    assert all(np.isfinite(out)) and b["wfext"] > 0.0 and b["enint"] > 0.0
    assert abs(out[-1]) < 0.1
    s.close()
    sb = en.ShellSolver(t, en.FP64)
    sb.run_cycles(2)
    with pytest.raises(en.OradError) as ei:
        sb.balance()
    assert ei.value.code == en.ORAD_ESTATE
    sb.close()
