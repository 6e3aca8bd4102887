"""The C ABI library loads and exports every symbol include/*.h declares (no GPU calls)."""

import re
from pathlib import Path

import pytest

from isogriddft_amd import _lib

HEADERS = sorted((Path(__file__).resolve().parents[1] / "include").glob("*.h"))


def _declared():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(ig_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_and_binding_agree():
    declared = _declared()
    assert declared, "no declarations found"
    assert declared == set(_lib.SIGNATURES), (declared ^ set(_lib.SIGNATURES))


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.fail(f"{_lib.LIB_PATH} not built (run python -m isogriddft_amd.build)")
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name


def test_host_only_entry_points():
    lib = _lib.load()
    assert lib.ig_reduction_work_size(10) >= 3
    assert lib.ig_gram_work_size(16, 16, 1000) >= 256
    assert lib.ig_gram_work_size(0, 16, 1000) == 0


def test_invalid_arguments_are_rejected_before_any_launch():
    # NULL device pointers are rejected on the host side with status -1 -> ValueError
    with pytest.raises(ValueError):
        _lib.call("ig_monitor_stencil_f64", None, None, 1, 1000, 10, 10, 10,
                  None, None, None, 1.0, 1, 0, 0, 0, 0, None)
    with pytest.raises(ValueError):
        _lib.call("ig_gram_f32", None, 10, 2, None, 10, 2, None, 10, None, 0, None, 0, None)
