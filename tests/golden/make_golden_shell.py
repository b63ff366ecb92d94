"""Writes tests/golden/shell_groups.npz: inputs and outputs of the reference's own
shell routines (CCOOR3 ... CUPDT3P and SIGEPS02C/M2CPLR, built from source into
oracle/_ref and driven in CFORC3 order by oracle/ref/harness_shell.F) on small
element groups, plus tests/golden/shell_constants.npy and shell_tables.npz.
Run in the build container, where the reference tree exists:

    python tests/golden/make_golden_shell.py

Each case first advances a group through a few reference steps from rest so the
recorded step starts from a realistic state (stresses, plastic strain,
hourglass and generalised-stress history), then records one step.  Cases cover
the plasticity algorithms IPLA 0/1/2, 2..5 through-thickness points, the rate
forms VP 1/2/3, SIG_max capping, the first cycle (DT1 = 0), warped elements,
triangles (nodes 3 = 4) and a deleted element (OFF < 0).
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

SQUARE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float).T
MKEYS = ("rho0", "young", "nu", "shear", "a11", "a12", "ssp", "ca", "cb", "cn", "epmx",
         "ymax", "cc", "epdr", "mexp", "asrate_pm9")
IKEYS = ("icc", "vp", "israte", "ipla")
PKEYS = ("hm", "hf", "hr", "dm", "fsh", "dtfac", "dtmin")
PIKEYS = ("ish", "npt")


def make_case(rng, nel, warp, vmag, wmag, dt1, mat, prop, warm=4, tri=False, dead=False):
    size = rng.uniform(0.6, 2.0, (nel, 1, 1))
    xg = np.repeat(SQUARE[None], nel, 0) * size + rng.uniform(-warp, warp, (nel, 3, 4))
    xg += rng.uniform(-50, 50, (nel, 3, 1))
    thk = rng.uniform(0.7, 1.3, nel)
    itri = np.zeros(nel, np.int32)
    if tri:
        itri[::2] = 1
        xg[::2, :, 3] = xg[::2, :, 2]
    st = po.shell_state(nel, thk)
    if dead:
        st[: nel // 4, po.SH_OFFG] = -1.0
    def velocities():
        vg = rng.normal(0, vmag, (nel, 3, 4))
        vrg = rng.normal(0, wmag, (nel, 3, 4))
        vg[itri == 1, :, 3] = vg[itri == 1, :, 2]      # corner 4 is node 3
        vrg[itri == 1, :, 3] = vrg[itri == 1, :, 2]
        return vg, vrg

    for _ in range(warm):
        vg, vrg = velocities()
        po.ref_shell_group(xg, vg, vrg, thk, st, 1.5e-7, mat, prop, tri=itri)
        xg = xg + 1.5e-7 * vg
    vg, vrg = velocities()
    out = st.copy()
    r = po.ref_shell_group(xg, vg, vrg, thk, out, dt1, mat, prop, tri=itri)
    return dict(xg=xg, vg=vg, vrg=vrg, thk=thk, tri=itri, state=st, dt1=np.float64(dt1),
                fsky=r["fsky"], dt=r["dt"], ehou=r["ehou"], state_out=out,
                mat=np.array([mat[k] for k in MKEYS]), imat=np.array([mat[k] for k in IKEYS]),
                prop=np.array([prop[k] for k in PKEYS]),
                iprop=np.array([prop[k] for k in PIKEYS]))


def main():
    if not po.ref_available():
        raise SystemExit("oracle/_ref/libref.so not built (python oracle/build_ref.py)")
    rng = np.random.Generator(np.random.Philox(dk.SEED_BASE + 200))
    m0 = dk.shell_mat_params()
    p5 = dk.shell_prop_params(npt=5)
    cases = {
        "elastic": make_case(rng, 48, 0.02, 300.0, 300.0, 1.4e-7, m0, p5),
        "plastic_ipla0": make_case(rng, 64, 0.05, 4e3, 4e3, 1.4e-7, m0, p5),
        "plastic_ipla1": make_case(rng, 64, 0.05, 4e3, 4e3, 1.4e-7,
                                   dk.shell_mat_params(ipla=1), p5),
        "plastic_ipla2": make_case(rng, 64, 0.05, 4e3, 4e3, 1.4e-7,
                                   dk.shell_mat_params(ipla=2), p5),
        "npt2": make_case(rng, 48, 0.05, 4e3, 4e3, 1.4e-7, m0, dk.shell_prop_params(npt=2)),
        "npt3": make_case(rng, 48, 0.05, 4e3, 4e3, 1.4e-7, m0, dk.shell_prop_params(npt=3)),
        "npt4": make_case(rng, 48, 0.05, 4e3, 4e3, 1.4e-7, m0, dk.shell_prop_params(npt=4)),
        "rate_vp1": make_case(rng, 48, 0.05, 4e3, 4e3, 1.4e-7,
                              dk.shell_mat_params(cc=0.014, eps0=1.0, vp=1), p5),
        "rate_vp2": make_case(rng, 48, 0.05, 4e3, 4e3, 1.4e-7,
                              dk.shell_mat_params(cc=0.014, eps0=1.0, vp=2), p5),
        "rate_vp3": make_case(rng, 48, 0.05, 4e3, 4e3, 1.4e-7,
                              dk.shell_mat_params(cc=0.014, eps0=1.0, vp=3), p5),
        "sigmax": make_case(rng, 48, 0.05, 4e3, 4e3, 1.4e-7,
                            dk.shell_mat_params(sigmx=400.0), p5),
        "first_cycle": make_case(rng, 32, 0.03, 2e3, 2e3, 0.0, m0, p5, warm=0),
        "warped": make_case(rng, 64, 0.25, 3e3, 3e3, 1.2e-7, m0, p5),
        "triangles": make_case(rng, 32, 0.05, 3e3, 3e3, 1.4e-7, m0, p5, tri=True),
        "deleted": make_case(rng, 32, 0.05, 3e3, 3e3, 1.4e-7, m0, p5, dead=True),
        "damped": make_case(rng, 32, 0.05, 3e3, 3e3, 1.4e-7, m0,
                            dk.shell_prop_params(npt=5, dm=0.05)),
    }
    flat = {}
    for name, c in cases.items():
        for k, v in c.items():
            flat[f"{name}/{k}"] = v
    path = os.path.join(HERE, "shell_groups.npz")
    np.savez_compressed(path, **flat)
    np.save(os.path.join(HERE, "shell_constants.npy"), po.ref_shell_constants())
    tabs = {}
    for npt in range(1, 6):
        z, w, m = po.ref_shell_tables(npt)
        tabs[f"z{npt}"], tabs[f"w{npt}"], tabs[f"m{npt}"] = z, w, m
    np.savez(os.path.join(HERE, "shell_tables.npz"), **tabs)
    print(f"wrote {path} ({os.path.getsize(path)} bytes), {len(cases)} cases")


if __name__ == "__main__":
    main()
