"""orad_energies_noalloc, the entry point that allocates no device memory, against
orad_energies through the C ABI: the same kernels over the same data, so the three
energies must agree bit for bit.  Two calls in a row on one GPU share the library's
scratch array; the second call is on a smaller mesh, and a third family (shells,
rotational kinetic energy included) follows on the same array."""
import ctypes as C

import numpy as np
import pytest

from openradioss_amd import deck as dk
from openradioss_amd import engine as en

pytestmark = pytest.mark.gpu


def both(s):
    ref = s.energies()
    ei, ek, eh = C.c_double(), C.c_double(), C.c_double()
    rc = s.lib.orad_energies_noalloc(s.h, C.byref(ei), C.byref(ek), C.byref(eh))
    assert rc == en.ORAD_OK, s.lib.orad_last_error(s.h)
    return np.array(ref), np.array([ei.value, ek.value, eh.value])


@pytest.mark.parametrize("prec", [en.FP64, en.MIXED])
def test_noalloc_energies_equal_allocating_entry(prec):
    big = en.BrickSolver(dk.brick_block(24, 24, 24, v_top=-5.0e4), prec)
    big.run_cycles(30)
    small = en.BrickSolver(dk.brick_block(5, 4, 3, v_top=-5.0e4), prec)
    small.run_cycles(30)
    plate = dk.shell_plate(n=16, patch=4)
    sh = en.ShellSolver(plate, prec)
    sh.run_cycles(30)
    try:
        for s in (big, small, sh, big):
            ref, got = both(s)
            print(f"energies {ref} noalloc {got}", flush=True)
            assert ref[1] > 0.0
            assert np.array_equal(ref, got)
        # state moves on between calls: the scratch keeps the last call's partials,
        # the next call overwrites every one it reads
        small.run_cycles(10)
        ref, got = both(small)
        assert np.array_equal(ref, got)
    finally:
        for s in (big, small, sh):
            s.close()
