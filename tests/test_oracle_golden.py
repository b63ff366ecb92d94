"""The oracle against the reference's own brick kernels.

tests/golden/brick_groups.npz holds inputs and outputs of the reference
routines SRCOOR3, SDERI3, SDLEN3, SDEFO3, SRHO3, M2LAW, MQVISCB, SHVIS3, SFINT3,
SRROTA3 and SCUMU3P (compiled from source, driven in SFORC3 order by
oracle/ref/harness_brick.F; written by tests/golden/make_golden.py).
The oracle restates them in C and must reproduce them bit for bit.
"""
import os

import numpy as np
import pytest

import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "brick_groups.npz")
MKEYS = ("rho0", "shear", "bulk", "ca", "cb", "cn", "epmx", "sigmx", "cc", "epdr", "mexp")
IKEYS = ("icc", "vp", "israte")
PKEYS = ("qa", "qb", "hvis", "dtfac")


def load_cases():
    z = np.load(GOLD)
    names = sorted({k.split("/")[0] for k in z.files})
    out = {}
    for n in names:
        c = {k.split("/")[1]: z[k] for k in z.files if k.startswith(n + "/")}
        c["mat"] = dict(zip(MKEYS, c["mat"].tolist()))
        c["mat"].update({k: int(v) for k, v in zip(IKEYS, c["imat"].tolist())})
        c["prop"] = dict(zip(PKEYS, c["prop"].tolist()))
        out[n] = c
    return out


CASES = load_cases()


def run_oracle(c):
    st = c["state"].copy()
    nel = st.shape[0]
    f = np.zeros((nel, 8, 3))
    sti = np.zeros(nel)
    dts, eh, vol = [], np.zeros(nel), np.zeros(nel)
    for e in range(nel):
        row = st[e].copy()
        o = po.brick_element(c["xg"][e], c["vg"][e], row, float(c["dt1"]), c["mat"], c["prop"])
        st[e] = row
        f[e] = o["f"].T
        sti[e] = o["sti"]
        dts.append(o["dtx"])
        eh[e] = o["ehou"]
        vol[e] = o["voln"]
    return f, sti, st, min(dts), eh, vol


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_bitwise_vs_reference_kernels(name):
    c = CASES[name]
    f, sti, st, dt, eh, vol = run_oracle(c)
    # FSKY(1:3,slot) forces and FSKY(7,slot) = STI/4 per corner (SCUMU3P)
    np.testing.assert_array_equal(f, c["fsky"][:, :, 0:3])
    np.testing.assert_array_equal(np.repeat(sti[:, None], 8, 1), c["fsky"][:, :, 6])
    np.testing.assert_array_equal(st, c["state_out"])
    assert dt == float(c["dt2t"])
    np.testing.assert_array_equal(eh, c["ehou"])
    np.testing.assert_array_equal(vol, c["voln"])


def test_golden_cases_exercise_branches():
    """The fixtures really reach plasticity, viscosity and the rate branch."""
    pl = CASES["plastic"]
    assert np.any(pl["state_out"][:, po.ST_PLA] > pl["state"][:, po.ST_PLA])
    cp = CASES["compress"]
    assert np.any(cp["state_out"][:, po.ST_QVIS] > 0)
    rc = CASES["rate_c"]
    assert np.any(rc["state_out"][:, po.ST_PLA] > rc["state"][:, po.ST_PLA])
    fc = CASES["first_cycle"]       # DT1 = 0: stress unchanged apart from pressure
    assert np.all(np.isfinite(fc["fsky"]))


def test_reference_constants_match_oracle_literals():
    c = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_constants.npy"))
    expect = [float.fromhex("0x1.553f7ced91687p+0"), float.fromhex("0x1.5555555555555p-2"),
              float.fromhex("0x1.5555555555555p-1"), 2.0 ** -6,
              float.fromhex("0x1.203af9ee75616p-50"), float.fromhex("0x1.79ca10c924223p-67"),
              float.fromhex("0x1.199999999999ap+0"), 0.25, 1.0e20,
              float(np.float32(1.1)), 1.0e6]
    np.testing.assert_array_equal(c[:11], expect)


