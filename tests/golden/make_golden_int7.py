"""Writes tests/golden/int7_cases.npz: inputs and outputs of the reference's own
/INTER/TYPE7 routines (I7COR3, I7DST3, I7FOR3, I7ASS2, built from source into
oracle/_ref and driven in I7MAINF order by oracle/ref/harness_int7.F).  Run in the
build container, where the reference tree exists:

    python tests/golden/make_golden_int7.py

Each case is a small surface of quadrangle (and triangle) segments with secondary
nodes scattered around it; the candidate list is every (node, segment) pair of
I7TRI's box test, in (node, segment) order.  Recorded per case: the reference's
output on the full candidate list (skyline entries, contact time step, energies and,
per candidate, whether I7DST3 kept it) and on the list of the kept candidates only
(the batching the library uses: NVSIZ penetrating pairs per batch).  Cases cover
Coulomb friction with friction damping, frictionless contact, no damping at all, the
first cycle (DT1 = 0), warped quadrangles, triangles, a node on a segment (zero
distance: full penetration, the node's deactivation and the batch's time-step
reset) and lists of several batches.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402


def grid(nx, ny, h, warp, rng, tri=()):
    g = np.array([[i * h, j * h, 0.0] for j in range(ny + 1) for i in range(nx + 1)], float)
    g[:, 2] += rng.uniform(-warp, warp, len(g))
    w = nx + 1
    irect = np.array([[i + w * j, i + 1 + w * j, i + 1 + w * (j + 1), i + w * (j + 1)]
                      for j in range(ny) for i in range(nx)], np.int32)
    for t in tri:
        irect[t, 3] = irect[t, 2]
    return g, irect


def case(rng, nx=3, ny=3, h=1.0, warp=0.05, ns=40, gap=0.2, fric=0.3, visc=0.05, viscf=1.0,
         dt1=1e-6, tri=(4,), vscale=3.0, extra=None):
    g, irect = grid(nx, ny, h, warp, rng, tri)
    xs = np.column_stack([rng.uniform(-0.3 * h, (nx + 0.3) * h, ns),
                          rng.uniform(-0.3 * h, (ny + 0.3) * h, ns),
                          rng.uniform(-1.25 * gap, 1.25 * gap, ns)])
    if extra is not None:
        xs = np.vstack([xs, extra])
    x = np.vstack([g, xs])
    nn = len(x)
    v = rng.normal(size=(nn, 3)) * vscale
    ms = rng.uniform(0.5, 2.0, nn) * 1e-3
    nsv = np.arange(len(g), nn, dtype=np.int32)
    stfm = rng.uniform(1e4, 3e4, len(irect))
    stfns = np.ones(len(nsv))
    c = dict(nsv=nsv, irect=irect, stfm=stfm, stfns=stfns, gap=gap, fric=fric, visc=visc,
             viscf=viscf)
    cn, ce = po.Contact7(c).candidates(x)
    args = (x, v, ms, nsv, irect, stfm, stfns)
    full = po.ref_int7(*args, cn, ce, gap, fric, visc, viscf, dt1, dt1)
    kept = full["pen"]
    act = po.ref_int7(*args, cn[kept], ce[kept], gap, fric, visc, viscf, dt1, dt1)
    assert act["npen"] == kept.sum()
    return dict(x=x, v=v, ms=ms, nsv=nsv, irect=irect, stfm=stfm, stfns=stfns, cand_n=cn,
                cand_e=ce, params=np.array([gap, fric, visc, viscf, dt1]),
                full_fsky=full["fsky"], full_isky=full["isky"], full_pen=full["pen"],
                full_scal=np.array([full["dt2t"], full["econt"], full["econtv"],
                                    full["econtd"]]),
                full_stfns=full["stfns"], act_fsky=act["fsky"], act_isky=act["isky"],
                act_scal=np.array([act["dt2t"], act["econt"], act["econtv"], act["econtd"]]))


def main():
    if not po.ref_available():
        raise SystemExit("oracle/_ref/libref.so not built (python oracle/build_ref.py)")
    rng = np.random.Generator(np.random.Philox(20261207))
    cases = {
        "friction": case(rng),
        "frictionless": case(rng, fric=0.0, viscf=0.0),
        "no_damping": case(rng, fric=0.0, visc=0.0, viscf=0.0),
        "friction_damping_only": case(rng, fric=0.2, visc=0.0, viscf=0.5),
        "first_cycle": case(rng, dt1=0.0),
        "flat": case(rng, warp=0.0, tri=()),
        "triangles": case(rng, tri=tuple(range(9))),
        "on_surface": case(rng, warp=0.0, tri=(), ns=20,
                           extra=np.array([[1.5, 1.5, 0.0], [0.25, 2.5, 0.0]])),
        "multi_batch": case(rng, nx=8, ny=8, ns=600, h=0.8),
        "slow_approach": case(rng, vscale=0.05, gap=0.3),
    }
    out = {}
    for name, c in cases.items():
        for k, a in c.items():
            out[f"{name}/{k}"] = a
        print(f"{name}: {len(c['cand_n'])} candidates, {int(c['full_pen'].sum())} penetrating, "
              f"{len(c['act_isky'])} skyline entries")
    np.savez_compressed(os.path.join(HERE, "int7_cases.npz"), **out)


if __name__ == "__main__":
    main()
