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
