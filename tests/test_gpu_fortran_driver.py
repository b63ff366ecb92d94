"""The Fortran binding (openradioss_amd/fortran/orad_bind.F90) drives the C ABI the
way the reference engine would: a RESOL-style loop calling the SFORC3 drop-in per
128-element group and the ASSPAR4 drop-in, with the nodal update in Fortran on
the host.  It must give the same nodes, bit for bit, as the device-resident
cycle, and agree with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
from openradioss_amd import build as ob
from openradioss_amd import deck as dk
from openradioss_amd import engine as en

pytestmark = pytest.mark.gpu


def run_driver(deckf, n, mode, out, prec="fp64", driver=None):
    args = [driver or ob.DRIVER, str(deckf), str(n), mode, str(out)] + \
        (["mixed"] if prec == "mixed" else [])
    p = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def test_fortran_group_loop_equals_device_cycle(tmp_path):
    d = dk.brick_block(9, 7, 6, v_top=-5.0e4)
    deckf = tmp_path / "deck.bin"
    en.write_deck(d, deckf)
    nn = d["x"].shape[0]
    run_driver(deckf, 60, "group", tmp_path / "g.bin")
    run_driver(deckf, 60, "device", tmp_path / "d.bin")
    xg, vg, dg = en.read_driver_output(tmp_path / "g.bin", nn)
    xd, vd, dd = en.read_driver_output(tmp_path / "d.bin", nn)
    assert np.array_equal(xg, xd)
    assert np.array_equal(vg, vd)
    assert np.array_equal(dg, dd)
    m = po.Model(d)
    m.cycle(60)
    assert np.abs(vg - m.v).max() / np.abs(m.v).max() < 1e-10
    assert np.abs(dg - m.d).max() / np.abs(m.d).max() < 1e-10


def test_fortran_driver_mixed_runs(tmp_path):
    d = dk.brick_block(6, 6, 6)
    deckf = tmp_path / "deck.bin"
    en.write_deck(d, deckf)
    run_driver(deckf, 40, "device", tmp_path / "m.bin", "mixed")
    x, v, dd = en.read_driver_output(tmp_path / "m.bin", d["x"].shape[0])
    m = po.Model(d)
    m.cycle(40)
    assert np.linalg.norm(dd - m.d) / np.linalg.norm(m.d) < 1e-4


@pytest.mark.parametrize("ishell", [1, 24, 12])
def test_fortran_shell_group_loop_equals_device_cycle(tmp_path, ishell):
    # CFORC3 / CZFORC3 / CBAFORC3 drop-in per 128-shell group, ASSPAR4 with rotations, the 6-DOF
    # nodal update on the host: bit for bit the device-resident cycle, and the oracle
    d = dk.shell_plate(n=14, patch=4, v0=1.5e4, ishell=ishell)
    deckf = tmp_path / "shell.bin"
    en.write_shell_deck(d, deckf)
    nn = d["x"].shape[0]
    run_driver(deckf, 80, "group", tmp_path / "g.bin", driver=ob.DRIVER_SHELL)
    run_driver(deckf, 80, "device", tmp_path / "d.bin", driver=ob.DRIVER_SHELL)
    g = en.read_shell_driver_output(tmp_path / "g.bin", nn)
    dv = en.read_shell_driver_output(tmp_path / "d.bin", nn)
    for a, b in zip(g, dv):
        assert np.array_equal(a, b)
    m = po.ShellModel(d)
    m.cycle(80)
    xg, vg, dg, vrg = g
    assert np.abs(vg - m.v).max() / np.abs(m.v).max() < 1e-10
    assert np.abs(dg - m.d).max() / np.abs(m.d).max() < 1e-10
    assert np.abs(vrg - m.vr).max() / np.abs(m.vr).max() < 1e-10
