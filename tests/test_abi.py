"""The C ABI library loads and exports every function include/exh.h declares."""
import ctypes
import os
import re

import pytest

from exciting_amd import lapw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "exh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(exh_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "exh_hamiltonandoverlapsetup" in names and "exh_match" in names
    assert set(names) == set(lapw.EXPORTS)


def test_library_exports_all_symbols():
    if not os.path.exists(lapw.LIB_PATH):
        pytest.fail("libexh.so not built (python -m exciting_amd.build)")
    L = ctypes.CDLL(lapw.LIB_PATH)
    for n in declared():
        assert hasattr(L, n), n


def test_fortran_module_binds_every_entry_point():
    """The ISO_C_BINDING module declares an interface for every C entry
    point, under the same name, and makes it public."""
    text = open(os.path.join(ROOT, "exciting_amd", "fortran", "exh_lapw.f90")).read()
    bound = set(re.findall(r'bind\(C, name="(exh_[a-z_0-9]+)"\)', text))
    assert bound == set(declared())
    public = re.search(r"^\s*public\s*::(.*?)\n\s*\n", text, flags=re.S | re.M).group(1)
    public = set(re.findall(r"exh_[a-z_0-9]+", public))
    # exh_error returns a C string: the module keeps that interface private
    # and exports exh_error_message, which copies it into a Fortran string
    wrapped = {"exh_error": "exh_error_message"}
    assert {wrapped.get(n, n) for n in declared()} <= public


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(lapw, "_lib", None)
    monkeypatch.setattr(lapw, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(lapw.ExhError, match="not built"):
        lapw.lib()


def test_batch_plan_table():
    """Every plan names one of the compiled tile shapes; plans enumerate
    shape x {tile grid, persistent} x {k-major, XCD order} exactly once."""
    n = lapw.plan_count()
    seen = set()
    for p in range(n):
        info = lapw.plan_info(p)
        assert info["tile_rows"] in (64, 128) and info["tile_cols"] in (64, 128)
        assert info["k_slab"] in (8, 16)
        assert info["persistent"] in (0, 1) and info["xcd_order"] in (0, 1)
        seen.add(tuple(info.values()))
    assert len(seen) == n
    with pytest.raises(lapw.ExhError):
        lapw.plan_info(n)
    with pytest.raises(lapw.ExhError):
        lapw.plan_info(-1)
