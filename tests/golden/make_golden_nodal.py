"""Writes tests/golden/nodal_cycles.npz: inputs and outputs of the reference's own
nodal-update routines (ACCELE, BCS10/BCS1V, FIXVEL, VELOCITY, DEPLA, built from
source into oracle/_ref and driven in RESOL order by oracle/ref/harness_nodal.F,
with RESOL's own time-step lines) over several consecutive cycles.  Run in the
build container, where the reference tree exists:

    python tests/golden/make_golden_nodal.py

Every case starts at the first cycle (DT1 = 0, DT2OLD = DTINI/1.1) unless stated,
and the element time-step minimum DT2T is chosen per cycle so that the growth
limit 1.1*DT2OLD, a drop, DTMX and a steady step each decide DT2 at least once.
Nodes carry every ICODT code 0..7 (and ICODR 0..7 with rotations), a zero mass
(and a zero inertia), and FIXVEL entries in x, y and z, one of them on a node that
is also fixed in that direction.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

NN = 16
ARRAYS = ("x", "v", "d", "a", "vr", "ar", "fv_mem")


def make_case(seed, rot, dt2t_seq, dtmx_seq=None, start=None):
    rng = np.random.default_rng(seed)
    nodes = dict(x=rng.uniform(-50.0, 50.0, (NN, 3)), v=rng.uniform(-1.0e4, 1.0e4, (NN, 3)),
                 d=np.zeros((NN, 3)), ms=rng.uniform(1.0e-9, 5.0e-9, NN),
                 icodt=np.array([0, 1, 2, 3, 4, 5, 6, 7] + [0] * 8, np.int32),
                 fv_dir=np.array([0] * 8 + [1, 2, 3, 0, 3, 0, 1, 0], np.int32),
                 fv_val=rng.uniform(-2.0e4, 2.0e4, NN))
    nodes["icodt"][14] = 4          # fixed in x and imposed in x: FIXVEL wins
    nodes["fv_dir"][14] = 1
    nodes["ms"][11] = 0.0           # ACCELE's MS <= 0 branch
    if rot:
        nodes["vr"] = rng.uniform(-20.0, 20.0, (NN, 3))
        nodes["in"] = rng.uniform(1.0e-10, 4.0e-10, NN)
        nodes["icodr"] = np.array([7, 6, 5, 4, 3, 2, 1, 0] + [0, 1, 2, 4] + [0] * 4, np.int32)
        nodes["in"][10] = 0.0
    time = dict(tt=0.0, dt1=0.0, dt2=0.0, dt12=0.0, dt2old=1.0e6 / 1.1, dtmx=1.0e20,
                wfext=0.0, ncycle=0)
    if start is not None:
        time.update(start)
    init = {k: np.array(v) for k, v in nodes.items()}
    init_time = dict(time)
    n = len(dt2t_seq)
    forces = rng.uniform(-2.0e-3, 2.0e-3, (n, NN, 3))
    moments = rng.uniform(-1.0e-4, 1.0e-4, (n, NN, 3))
    out = {k: np.zeros((n, NN, 3)) for k in ARRAYS}
    out["fv_mem"] = np.zeros((n, NN))
    tout = np.zeros((n, 8))
    cur = dict(nodes)
    for c in range(n):
        if dtmx_seq is not None:
            time["dtmx"] = dtmx_seq[c]
        cur["a"] = forces[c].copy()
        cur["ar"] = moments[c].copy() if rot else np.zeros((NN, 3))
        cur, time = po.ref_nodal_cycle(cur, time, dt2t_seq[c], rot=rot)
        for k in ARRAYS:
            out[k][c] = cur[k]
        tout[c] = [time[k] for k in po.TIME_KEYS] + [time["wfext"], time["ncycle"]]
    rec = {f"in_{k}": v for k, v in init.items()}
    rec.update({f"out_{k}": v for k, v in out.items()})
    rec["time0"] = np.array([init_time[k] for k in po.TIME_KEYS] +
                            [init_time["wfext"], init_time["ncycle"]])
    rec["time"] = tout
    rec["forces"] = forces
    rec["moments"] = moments
    rec["dt2t"] = np.asarray(dt2t_seq, float)
    rec["dtmx"] = np.asarray(dtmx_seq if dtmx_seq is not None else [time["dtmx"]] * n, float)
    rec["rot"] = np.array(int(rot))
    return rec


def main():
    if not po.ref_available():
        raise SystemExit("oracle/_ref/libref.so missing: run oracle/build_ref.py first")
    # DT2T per cycle: 1st cycle takes it whole; then the 1.1 growth limit; a drop;
    # steady; the growth limit again
    seq = [1.0e-7, 1.5e-7, 0.8e-7, 0.8e-7, 2.0e-7, 2.0e-7]
    cases = {
        "trans": make_case(1, False, seq),
        "rot": make_case(2, True, seq),
        # DTMX decides from the third cycle, then is lifted
        "dtmx": make_case(3, True, [1.2e-7] * 6,
                          dtmx_seq=[1e20, 1e20, 0.9e-7, 0.9e-7, 1e20, 1e20]),
        # a run picked up mid-way (TT, DT1, DT2 set), with the external work memory
        "restart": make_case(4, False, [1.3e-7, 1.25e-7, 1.4e-7, 1.1e-7],
                             start=dict(tt=3.1e-5, dt1=1.2e-7, dt2=1.21e-7, dt12=1.205e-7,
                                        dt2old=1.21e-7, wfext=0.25, ncycle=250)),
    }
    flat = {}
    for name, rec in cases.items():
        for k, v in rec.items():
            flat[f"{name}:{k}"] = v
    path = os.path.join(HERE, "nodal_cycles.npz")
    np.savez_compressed(path, **flat)
    print(f"wrote {path}: {len(cases)} cases")


if __name__ == "__main__":
    main()
