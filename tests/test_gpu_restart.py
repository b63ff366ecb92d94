"""Restart files through the C ABI (orad_write_restart / orad_read_restart).

A run of N + M cycles and a run of N cycles written to a file, read back into a new
context built from the same model and continued for M cycles must agree bit for bit:
nodes (X, V, D, A, STIFN and VR, AR for shells), every element state value, the time
record, the energies and ECRIT's balance (ENTOT0 and WFEXT included).  Covers
bricks and the three shell formulations (Belytschko-Tsay, QEPH, QBAT) in both forms,
batched shell models, rewinding a context to its own file, and the refusals: a file
of another model, another precision or another element order, a corrupt or truncated
file (nothing on the device changes then).
"""
import numpy as np
import pytest

from openradioss_amd import deck as dk
from openradioss_amd import engine as en

pytestmark = pytest.mark.gpu

TIME_KEYS = ("tt", "dt1", "dt2", "dt12", "dt2old", "dtmx", "ncycle", "mstop")


def snapshot(s, balance=True):
    t = s.time()
    out = dict(nodes=s.nodes(), state=s.state(), time=tuple(getattr(t, k) for k in TIME_KEYS),
               energies=s.energies())
    if balance:
        out["balance"] = s.balance()
    return out


def assert_same(a, b):
    for k in a["nodes"]:
        assert np.array_equal(a["nodes"][k], b["nodes"][k]), k
    if isinstance(a["state"], dict):
        for k in a["state"]:
            assert np.array_equal(a["state"][k], b["state"][k]), k
    else:
        assert np.array_equal(a["state"], b["state"])
    assert a["time"] == b["time"]
    assert a["energies"] == b["energies"]
    if "balance" in a:
        assert a["balance"] == b["balance"]


def resume_matches(make, deck, prec, n, m, tmp_path, balance=True):
    path = str(tmp_path / "run.rst")
    s = make(deck, prec)
    s.run_cycles(n)
    s.write_restart(path)
    s.run_cycles(m)
    ref = snapshot(s, balance)
    s.close()
    r = make(deck, prec)
    r.read_restart(path)
    assert r.time().ncycle == n
    r.run_cycles(m)
    got = snapshot(r, balance)
    r.close()
    assert ref["time"][6] == n + m
    assert_same(got, ref)


@pytest.mark.parametrize("prec", [en.FP64, en.MIXED])
def test_brick_resume_bitwise(prec, tmp_path):
    d = dk.brick_block(10, 8, 6, v_top=-2.0e4)
    resume_matches(en.BrickSolver, d, prec, 37, 50, tmp_path)


@pytest.mark.parametrize("ishell,prec", [(1, en.FP64), (1, en.MIXED), (24, en.FP64),
                                         (24, en.MIXED), (12, en.FP64), (12, en.MIXED)])
def test_shell_resume_bitwise(ishell, prec, tmp_path):
    d = dk.shell_plate(n=24, patch=6, v0=1.5e4, ipla=1, ishell=ishell)
    resume_matches(en.ShellSolver, d, prec, 61, 80, tmp_path)


def test_batched_shell_models_resume_bitwise(tmp_path):
    d = dk.shell_tubes(ntube=4, side=(8.0, 16.0), length=(12.0, 30.0), h=(2.0, 4.0),
                       thk=(0.8, 2.5), ca=(200.0, 400.0), seed=11)
    path = str(tmp_path / "batch.rst")
    s = en.ShellSolver(d, en.MIXED)
    s.run_cycles(40)
    s.write_restart(path)
    s.run_cycles(55)
    ref = (s.nodes(), s.state(), [tuple(getattr(s.model_time(k), f) for f in TIME_KEYS)
                                  for k in range(s.nmodel)])
    s.close()
    r = en.ShellSolver(d, en.MIXED)
    r.read_restart(path)
    r.run_cycles(55)
    got = (r.nodes(), r.state(), [tuple(getattr(r.model_time(k), f) for f in TIME_KEYS)
                                  for k in range(r.nmodel)])
    r.close()
    for k in ref[0]:
        assert np.array_equal(got[0][k], ref[0][k]), k
    assert np.array_equal(got[1], ref[1])
    assert got[2] == ref[2]
    assert len({t[0] for t in ref[2]}) > 1          # the models keep their own times


def test_rewind_a_context_to_its_own_file(tmp_path):
    d = dk.shell_plate(n=16, patch=4, v0=1.5e4, ishell=24)
    path = str(tmp_path / "rw.rst")
    s = en.ShellSolver(d, en.FP64)
    s.run_cycles(30)
    s.write_restart(path)
    s.run_cycles(45)
    first = snapshot(s)
    s.read_restart(path)
    assert s.time().ncycle == 30
    s.run_cycles(45)
    again = snapshot(s)
    s.close()
    assert_same(again, first)


def test_refusals_leave_the_context_unchanged(tmp_path):
    d = dk.shell_plate(n=12, patch=4, v0=1.0e4, ishell=1)
    path = str(tmp_path / "a.rst")
    s = en.ShellSolver(d, en.FP64)
    s.run_cycles(20)
    s.write_restart(path)
    raw = open(path, "rb").read()
    s.close()

    other = {
        "another mesh": (en.ShellSolver, dk.shell_plate(n=14, patch=4, v0=1.0e4), en.FP64),
        "another precision": (en.ShellSolver, d, en.MIXED),
        "another formulation": (en.ShellSolver, dk.shell_plate(n=12, patch=4, v0=1.0e4,
                                                               ishell=24), en.FP64),
        "another family": (en.BrickSolver, dk.brick_block(4, 4, 4), en.FP64),
    }
    for why, (make, deck, prec) in other.items():
        t = make(deck, prec)
        t.run_cycles(3)
        before = snapshot(t)
        with pytest.raises(en.OradError, match="different model|another mesh"):
            t.read_restart(path)
        assert_same(snapshot(t), before)
        t.close()

    t = en.ShellSolver(d, en.FP64)
    t.run_cycles(5)
    before = snapshot(t)
    bad = bytearray(raw)
    bad[len(bad) - 200] ^= 0x10         # inside the last section's bytes
    (tmp_path / "bad.rst").write_bytes(bytes(bad))
    with pytest.raises(en.OradError, match="checksum"):
        t.read_restart(str(tmp_path / "bad.rst"))
    (tmp_path / "short.rst").write_bytes(raw[:len(raw) - 100])
    with pytest.raises(en.OradError, match="truncated|checksum"):
        t.read_restart(str(tmp_path / "short.rst"))
    (tmp_path / "junk.rst").write_bytes(b"not a restart file at all")
    with pytest.raises(en.OradError, match="not a restart file"):
        t.read_restart(str(tmp_path / "junk.rst"))
    with pytest.raises(en.OradError, match="cannot open"):
        t.read_restart(str(tmp_path / "missing.rst"))
    assert_same(snapshot(t), before)
    # the intact file still loads into the same context
    t.read_restart(path)
    assert t.time().ncycle == 20
    t.close()


def test_brick_element_order_is_part_of_the_model(tmp_path):
    d = dk.brick_block(6, 24, 4)        # 24 element rows: columns of 10 rows reorder
    path = str(tmp_path / "b.rst")
    s = en.BrickSolver(d, en.FP64)
    s.run_cycles(10)
    s.write_restart(path)
    s.close()
    t = en.BrickSolver(d, en.FP64)
    t.set_launch(order_rows=0)          # engine order: another state layout
    with pytest.raises(en.OradError, match="element order"):
        t.read_restart(path)
    t.set_launch(order_rows=10)
    t.read_restart(path)
    assert t.time().ncycle == 10
    t.close()
