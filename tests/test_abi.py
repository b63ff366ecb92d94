"""The C-ABI library loads and exports every entry point include/orad.h declares.
No compute call is made here (no GPU in the build container)."""
import os
import re

import pytest

from openradioss_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "orad.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(orad_\w+)\s*\(", txt, re.M)))


def test_header_and_python_binding_agree():
    assert set(declared()) == set(engine.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = engine.load_library()
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        engine._lib = None
        try:
            engine.load_library(str(tmp_path / "nope.so"))
        finally:
            engine._lib = None


def test_ixs_layout_is_the_engines():
    from openradioss_amd import deck as dk
    d = dk.brick_block(2, 2, 2)
    ixs = engine.ixs_from_deck(d)
    assert ixs.shape == (8, 11)
    assert ixs[:, 1:9].min() == 1 and ixs[:, 1:9].max() == d["x"].shape[0]
    assert list(ixs[:, 10]) == list(range(1, 9))


def test_fortran_binding_declares_every_abi_entry():
    """Every C entry point the engine needs has an ISO_C_BINDING interface."""
    txt = open(os.path.join(ROOT, "openradioss_amd", "fortran", "orad_bind.F90")).read()
    bound = set(re.findall(r"bind\(c, name='(orad_\w+)'\)", txt))
    needed = set(declared()) - {"orad_profile_cycles", "orad_get_brick_state",
                                "orad_set_brick_state", "orad_get_fsky"}
    assert needed <= bound, needed - bound


def test_deck_file_roundtrip(tmp_path):
    from openradioss_amd import deck as dk
    import numpy as np
    d = dk.brick_block(2, 3, 2)
    engine.write_deck(d, tmp_path / "d.bin")
    raw = np.fromfile(tmp_path / "d.bin", np.int32, count=4)
    assert list(raw) == [engine.DECK_MAGIC, 1, d["x"].shape[0], d["ixs"].shape[0]]
