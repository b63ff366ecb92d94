"""GPU tests of batched independent shell models (one element launch over uneven
models, per-model material and time step) and of the launch configuration.

Bars:
  * FP64 batch vs each model run alone: nodes, element state and time bitwise
    equal; internal and kinetic energy bitwise; hourglass energy within 1e-12
    (its sum is grouped per 256-element chunk in the batch);
  * one model of the batch vs the FP64 oracle: 1e-8 relative after 200 cycles;
  * mixed batch vs each model run alone in the mixed form: 1e-6 relative;
  * every launch configuration (grid / persistent tiles, LDS caps, XCD mapping)
    and the autotuner: bitwise identical results, hourglass energy included.
"""
import ctypes

import numpy as np
import pytest

import pyoracle as po
from openradioss_amd import deck as dk
from openradioss_amd import engine as en

pytestmark = pytest.mark.gpu

NODE_KEYS = ("x", "v", "vr", "ms", "in", "icodt", "icodr", "fv_dir", "fv_val")


def small_tubes(n=5, seed=7):
    return dk.shell_tubes(ntube=n, side=(8.0, 16.0), length=(12.0, 30.0), h=(2.0, 4.0),
                          thk=(0.8, 2.5), ca=(200.0, 400.0), seed=seed)


def model_deck(d, m):
    n0, n1 = d["node_off"][m], d["node_off"][m + 1]
    e0, e1 = d["elem_off"][m], d["elem_off"][m + 1]
    out = {k: d[k][n0:n1] for k in NODE_KEYS}
    out["ixc"] = d["ixc"][e0:e1] - n0
    out["uid"] = d["uid"][e0:e1]
    out["thk"] = d["thk"][e0:e1]
    out["mat"] = d["mats"][m]
    out["prop"] = d["prop"]
    return out


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def run_batch(d, prec, ncyc, launch=None):
    s = en.ShellSolver(d, prec)
    if launch:
        s.set_launch(**launch)
    s.run_cycles(ncyc)
    out = dict(nodes=s.nodes(), state=s.state(),
               times=[s.model_time(m) for m in range(s.nmodel)],
               en=[s.model_energies(m) for m in range(s.nmodel)], total=s.energies())
    s.close()
    return out


def test_batch_fp64_equals_models_run_alone():
    d = small_tubes()
    b = run_batch(d, en.FP64, 300)
    for m in range(len(d["mats"])):
        n0, n1 = d["node_off"][m], d["node_off"][m + 1]
        e0, e1 = d["elem_off"][m], d["elem_off"][m + 1]
        s = en.ShellSolver(model_deck(d, m), en.FP64)
        s.run_cycles(300)
        nd, st, t, e = s.nodes(), s.state(), s.time(), s.energies()
        s.close()
        for k in ("x", "v", "vr", "d"):
            assert np.array_equal(b["nodes"][k][n0:n1], nd[k]), (m, k)
        assert np.array_equal(b["state"][e0:e1], st), m
        tb = b["times"][m]
        assert (tb.tt, tb.dt2, tb.ncycle) == (t.tt, t.dt2, t.ncycle), m
        assert b["en"][m][0] == e[0] and b["en"][m][1] == e[1], m
        assert b["en"][m][2] == pytest.approx(e[2], rel=1e-12, abs=1e-30), m
    # the models step with different time steps
    assert len({t.dt2 for t in b["times"]}) == len(b["times"])
    assert b["total"][0] == pytest.approx(sum(x[0] for x in b["en"]), rel=1e-13)


def test_batch_model_vs_oracle():
    d = small_tubes(n=3, seed=11)
    m = 1
    md = model_deck(d, m)
    o = po.ShellModel(md, nthreads=4)
    o.cycle(200)
    b = run_batch(d, en.FP64, 200)
    n0, n1 = d["node_off"][m], d["node_off"][m + 1]
    e0, e1 = d["elem_off"][m], d["elem_off"][m + 1]
    assert b["times"][m].tt == pytest.approx(o.tt, rel=1e-12)
    for k, ref in (("x", o.x), ("v", o.v), ("vr", o.vr), ("d", o.d)):
        assert rel(b["nodes"][k][n0:n1], ref) < 1e-8, k
    assert rel(b["state"][e0:e1, po.SH_SIG0:po.SH_SIG0 + 25], o.st[:, po.SH_SIG0:po.SH_SIG0 + 25]) < 1e-8
    top = md["fv_dir"] == 3
    assert np.allclose(b["nodes"]["v"][n0:n1][top, 2], -1.0e4, rtol=0, atol=1e-6)
    assert o.st[:, po.SH_PLA0:po.SH_PLA0 + 5].max() > 0.0     # the crush yields


def test_batch_mixed_close_to_models_run_alone():
    d = small_tubes(n=4, seed=3)
    b = run_batch(d, en.MIXED, 200)
    same = True
    for m in range(4):
        n0, n1 = d["node_off"][m], d["node_off"][m + 1]
        s = en.ShellSolver(model_deck(d, m), en.MIXED)
        s.run_cycles(200)
        nd = s.nodes()
        s.close()
        for k in ("x", "v", "vr", "d"):
            same = same and np.array_equal(b["nodes"][k][n0:n1], nd[k])
            assert rel(b["nodes"][k][n0:n1], nd[k]) < 1e-6, (m, k)
    print(f"mixed batch bitwise equal to the models run alone: {same}")


@pytest.mark.parametrize("prec", [en.FP64, en.MIXED])
def test_launch_configurations_bitwise_identical(prec):
    d = small_tubes(n=6, seed=5)
    ref = run_batch(d, prec, 120)
    for launch in (dict(persistent=1, lds_kb=0, xcd_remap=0),
                   dict(persistent=1, lds_kb=80, xcd_remap=0),
                   dict(persistent=0, lds_kb=40, xcd_remap=0),
                   dict(persistent=0, lds_kb=0, xcd_remap=1)):
        got = run_batch(d, prec, 120, launch)
        for k in ("x", "v", "vr", "d"):
            assert np.array_equal(got["nodes"][k], ref["nodes"][k]), (launch, k)
        assert np.array_equal(got["state"], ref["state"]), launch
        assert got["en"] == ref["en"], launch


def test_autotune_restores_state_and_keeps_results():
    d = small_tubes(n=6, seed=9)
    ref = run_batch(d, en.MIXED, 100)
    s = en.ShellSolver(d, en.MIXED)
    s.run_cycles(40)
    best, ms = s.autotune(cycles=10)
    assert ms > 0.0 and best == s.launch()
    assert s.model_time(0).ncycle == 40
    s.run_cycles(60)
    nd = s.nodes()
    s.close()
    for k in ("x", "v", "vr", "d"):
        assert np.array_equal(nd[k], ref["nodes"][k]), k


def test_autotune_bricks():
    d = dk.brick_block(12, 12, 12)
    a = en.BrickSolver(d, en.MIXED)
    a.run_cycles(20)
    best, ms = a.autotune(cycles=6)
    assert ms > 0.0 and best["persistent"] == 0
    a.run_cycles(20)
    b = en.BrickSolver(d, en.MIXED)
    b.run_cycles(40)
    assert np.array_equal(a.nodes()["x"], b.nodes()["x"])
    a.close()
    b.close()


def test_batch_graph_replay_matches_eager():
    d = small_tubes(n=4, seed=13)
    a = en.ShellSolver(d, en.FP64)
    b = en.ShellSolver(d, en.FP64)
    a.set_launch(persistent=1, lds_kb=0, xcd_remap=0)
    b.set_launch(persistent=1, lds_kb=0, xcd_remap=0)
    a.run_cycles(64)
    b.profile_cycles(64)
    assert np.array_equal(a.nodes()["x"], b.nodes()["x"])
    assert np.array_equal(a.state(), b.state())
    assert a.model_energies(2) == b.model_energies(2)
    a.close()
    b.close()


def test_batch_error_paths():
    d = small_tubes(n=3)
    s = en.ShellSolver({k: v for k, v in d.items() if k not in ("node_off", "elem_off")},
                       en.FP64)
    lib, h = s.lib, s.h
    i32 = lambda a: np.ascontiguousarray(a, np.int32)
    bad = i32([0, 10, d["x"].shape[0] - 1])            # does not span every node
    eo = i32([0, d["elem_off"][1], d["elem_off"][3]])
    assert lib.orad_set_shell_models(h, 2, en._p(bad), en._p(eo), None, None) == en.ORAD_EINVAL
    no = i32([0, 10, d["x"].shape[0]])                  # shells straddle the split
    assert lib.orad_set_shell_models(h, 2, en._p(no), en._p(eo), None, None) == en.ORAD_EINVAL
    l = en.Launch(1, 0, 0, 0)                           # persistent needs models
    assert lib.orad_set_launch(h, ctypes.byref(l)) == en.ORAD_EINVAL
    s.set_models(d["node_off"], d["elem_off"], d["mats"])
    t = en.TimeState()
    assert lib.orad_get_model_time(h, 3, ctypes.byref(t)) == en.ORAD_EINVAL
    dt2t = ctypes.c_double(1.0)
    e1 = int(d["elem_off"][1])
    assert lib.orad_cforc3(h, e1 - 2, 4, 1e-7, ctypes.byref(dt2t)) == en.ORAD_EINVAL
    assert lib.orad_cforc3(h, e1, 4, 1e-7, ctypes.byref(dt2t)) == en.ORAD_OK
    s.close()
