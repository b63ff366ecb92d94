"""Writes tests/golden/brick_groups.npz: inputs and outputs of the reference's own
brick kernels (SRCOOR3 ... SCUMU3P, built from source into oracle/_ref) on small
element groups.  Run in the build container, where the reference tree exists:

    python tests/golden/make_golden.py

Cases cover the branches the path can take: elastic and plastic steps, n < 1 and
n = 1 hardening, rate-dependent yield (C > 0), compression (bulk viscosity on),
the first cycle (DT1 = 0) and strongly distorted elements.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402
from openradioss_amd import deck as dk  # noqa: E402

CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float).T


def make_case(rng, nel, dist, vmag, smag, plamax, dt1, mat, prop, compress=False):
    xg = np.stack([CUBE * rng.uniform(0.5, 2.0) + rng.uniform(-dist, dist, (3, 8))
                   + rng.uniform(-100, 100, (3, 1)) for _ in range(nel)])
    vg = rng.uniform(-vmag, vmag, (nel, 3, 8)) + rng.uniform(-1e4, 1e4, (nel, 3, 1))
    if compress:
        vg[:, 2, 4:] -= 2e3          # top face moving down
    st = np.zeros((nel, po.NST))
    vol = np.array([po.brick_volume(x) for x in xg])
    st[:, po.ST_VOL0] = vol * rng.uniform(0.98, 1.02, nel)
    st[:, po.ST_RHO] = mat["rho0"] * rng.uniform(0.98, 1.02, nel)
    st[:, po.ST_OFFG] = 1.0
    st[:, 0:6] = rng.uniform(-smag, smag, (nel, 6))
    st[:, po.ST_PLA] = rng.uniform(0, plamax, nel)
    st[:, po.ST_EINT] = rng.uniform(0, 2.0, nel)
    st[:, po.ST_QVIS] = rng.uniform(0, 3.0, nel)
    st[:, po.ST_EPSD] = rng.uniform(0, 1e3, nel)
    out_st = st.copy()
    r = po.ref_brick_group(xg, vg, out_st, dt1, mat, prop)
    return dict(xg=xg, vg=vg, state=st, dt1=np.float64(dt1), fsky=r["fsky"], state_out=out_st,
                dt2t=np.float64(r["dt2t"]), ehou=r["ehou"], voln=r["voln"],
                mat=np.array([mat[k] for k in MKEYS]), imat=np.array([mat[k] for k in IKEYS]),
                prop=np.array([prop[k] for k in PKEYS]))


MKEYS = ("rho0", "shear", "bulk", "ca", "cb", "cn", "epmx", "sigmx", "cc", "epdr", "mexp")
IKEYS = ("icc", "vp", "israte")
PKEYS = ("qa", "qb", "hvis", "dtfac")


def main():
    if not po.ref_available():
        raise SystemExit("oracle/_ref/libref.so not built (python oracle/build_ref.py)")
    rng = np.random.Generator(np.random.Philox(dk.SEED_BASE + 100))
    steel = dk.law2_params()
    prop = dk.solid_params()
    cases = {
        "elastic": make_case(rng, 64, 0.05, 50.0, 50.0, 0.0, 1.2e-7, steel, prop),
        "plastic": make_case(rng, 64, 0.10, 3e3, 400.0, 0.2, 1.3e-7, steel, prop),
        "compress": make_case(rng, 64, 0.05, 500.0, 300.0, 0.05, 1.3e-7, steel, prop, True),
        "first_cycle": make_case(rng, 32, 0.05, 1e3, 0.0, 0.0, 0.0, steel, prop),
        "distorted": make_case(rng, 64, 0.30, 2e3, 300.0, 0.1, 1.0e-7, steel, prop),
        "hard_n1": make_case(rng, 64, 0.10, 3e3, 400.0, 0.1, 1.3e-7,
                             dk.law2_params(n=1.0), prop),
        "rate_c": make_case(rng, 64, 0.10, 3e3, 400.0, 0.1, 1.3e-7,
                            dk.law2_params(C=0.014, eps0=1.0), prop),
        "sigmax": make_case(rng, 64, 0.10, 3e3, 400.0, 0.3, 1.3e-7,
                            dk.law2_params(sigmx=500.0), prop),
    }
    flat = {}
    for name, c in cases.items():
        for k, v in c.items():
            flat[f"{name}/{k}"] = v
    path = os.path.join(HERE, "brick_groups.npz")
    np.savez_compressed(path, **flat)
    c = po.ref_constants()
    np.save(os.path.join(HERE, "ref_constants.npy"), c)
    print(f"wrote {path} ({os.path.getsize(path)} bytes), {len(cases)} cases")


if __name__ == "__main__":
    main()
