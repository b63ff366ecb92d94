"""Writes tests/golden/qbat_groups.npz: inputs and outputs of the reference's own
QBAT shell routines (CBACOOR, CNCOEF3, CBADEFSH, CBADEF, CBASTRA3, CBAENER(S),
SIGEPS02C/M2CPLR, CBAVISC, CBAFORI, CBAFORCT, CBAPROJ, CNDT3 and CUPDTN3P, built
from source into oracle/_ref and driven in CBAFORC3 order by
oracle/ref/harness_qbat.F) on small element groups.  Run in the build container,
where the reference tree exists:

    python tests/golden/make_golden_qbat.py

Each case first advances a group through a few reference steps so the recorded
step starts from a realistic state (per Gauss point stresses, plastic strain and
resultants), then records one step.  Cases cover IPLA 0/1/2, 2..5 through-thickness
points, warped elements (per Gauss point frames and the rigid-body projection of
CBACOOR) and exactly flat ones (its flat branch), the first cycle (DT1 = 0), a
deleted element (OFF < 0), membrane damping, the default Dn (0 -> 1e-3) and a set
one, and the strain-rate term.
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
PKEYS = ("hm", "hf", "hr", "dm", "fsh", "dtfac", "dtmin", "dn", "cvis")
PIKEYS = ("ish", "npt", "ishell")


def bprop(**kw):
    kw.setdefault("dn", 0.0)
    return dk.shell_prop_params(ishell=12, **kw)


def rotations(rng, n):
    q, r = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    return q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]


def make_case(rng, nel, warp, vmag, wmag, dt1, mat, prop, warm=4, dead=False, flat=False):
    size = rng.uniform(0.6, 2.0, (nel, 1, 1))
    xg = np.repeat(SQUARE[None], nel, 0) * size + rng.uniform(-warp, warp, (nel, 3, 4))
    if flat:
        xg[:, 2, :] = 0.0
    else:
        xg = np.einsum("eij,ejk->eik", rotations(rng, nel), xg)
    xg += rng.uniform(-50, 50, (nel, 3, 1)) * (0.0 if flat else 1.0)
    thk = rng.uniform(0.7, 1.3, nel)
    st = po.qbat_state(nel, thk)
    if dead:
        st[: nel // 4, po.SH_OFFG] = -1.0

    def velocities():
        vg = rng.normal(0, vmag, (nel, 3, 4))
        vrg = rng.normal(0, wmag, (nel, 3, 4))
        if flat:
            vg[:, 2, :] = 0.0
            vrg[:, 0:2, :] = 0.0
        return vg, vrg

    for _ in range(warm):
        vg, vrg = velocities()
        po.ref_qbat_group(xg, vg, vrg, thk, st, 1.5e-7, mat, prop)
        xg = xg + 1.5e-7 * vg
    vg, vrg = velocities()
    out = st.copy()
    r = po.ref_qbat_group(xg, vg, vrg, thk, out, dt1, mat, prop)
    return dict(xg=xg, vg=vg, vrg=vrg, thk=thk, state=st, dt1=np.float64(dt1),
                fsky=r["fsky"], dtel=r["dt"], dt2t=np.float64(r["dt2t"]), evis=r["evis"],
                state_out=out,
                mat=np.array([mat[k] for k in MKEYS]), imat=np.array([mat[k] for k in IKEYS]),
                prop=np.array([prop[k] for k in PKEYS]),
                iprop=np.array([prop[k] for k in PIKEYS]))


def main():
    if not po.ref_available():
        raise SystemExit("oracle/_ref/libref.so not built (python oracle/build_ref.py)")
    rng = np.random.Generator(np.random.Philox(dk.SEED_BASE + 400))
    m0 = dk.shell_mat_params()
    p5 = bprop(npt=5)
    cases = {
        "elastic": make_case(rng, 48, 0.02, 300.0, 300.0, 1.4e-7, m0, p5),
        "plastic_ipla0": make_case(rng, 64, 0.05, 4e3, 4e3, 1.4e-7, m0, p5),
        "plastic_ipla1": make_case(rng, 64, 0.05, 4e3, 4e3, 1.4e-7,
                                   dk.shell_mat_params(ipla=1), p5),
        "plastic_ipla2": make_case(rng, 64, 0.05, 4e3, 4e3, 1.4e-7,
                                   dk.shell_mat_params(ipla=2), p5),
        "npt2": make_case(rng, 48, 0.05, 4e3, 4e3, 1.4e-7, m0, bprop(npt=2)),
        "npt3": make_case(rng, 48, 0.05, 4e3, 4e3, 1.4e-7, m0, bprop(npt=3)),
        "npt4": make_case(rng, 48, 0.05, 4e3, 4e3, 1.4e-7, m0, bprop(npt=4)),
        "flat": make_case(rng, 48, 0.08, 4e3, 4e3, 1.4e-7, m0, p5, flat=True),
        "warped": make_case(rng, 64, 0.25, 3e3, 3e3, 1.2e-7, m0, p5),
        "first_cycle": make_case(rng, 32, 0.03, 2e3, 2e3, 0.0, m0, p5, warm=0),
        "deleted": make_case(rng, 32, 0.05, 3e3, 3e3, 1.4e-7, m0, p5, dead=True),
        "damped": make_case(rng, 32, 0.05, 3e3, 3e3, 1.4e-7, m0, bprop(npt=5, dm=0.05)),
        "dn_set": make_case(rng, 32, 0.05, 3e3, 3e3, 1.4e-7, m0, bprop(npt=3, dn=0.02)),
        "rate_vp1": make_case(rng, 32, 0.05, 4e3, 4e3, 1.4e-7,
                              dk.shell_mat_params(cc=0.014, eps0=1.0, vp=1), p5),
    }
    flat = {}
    for name, c in cases.items():
        for k, v in c.items():
            flat[f"{name}/{k}"] = v
    path = os.path.join(HERE, "qbat_groups.npz")
    np.savez_compressed(path, **flat)
    print(f"wrote {path} ({os.path.getsize(path)} bytes), {len(cases)} cases")


if __name__ == "__main__":
    main()
