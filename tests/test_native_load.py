"""The C-ABI library loads and exports every symbol of include/fsgpu.h
(no compute call: this runs without a GPU)."""
import os
import re

import fractalshades_amd._native as native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "fsgpu.h")).read()
    return sorted(set(re.findall(r"^[\w ]+?\*?\s*(fsg_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "fsg_tile" in syms and "fsg_ctx_create" in syms
    assert set(syms) == set(native.EXPORTS)


def test_library_exports_all_declared():
    lib = native.load()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert lib.fsg_abi_version() == 6
    assert lib.fsg_work_bytes() >= 40


def test_desc_layout_matches_header():
    # offsets of a few fields of fsg_ref_desc, as laid out by the C compiler
    d = native.RefDesc
    assert d.Zn_path.offset == 40
    assert d.dx_mant.offset == 80
    assert d.lin_mat.offset == 120
    assert d.precision.offset == 208
    assert d.calc_orbit.offset == 212 and d.backshift.offset == 216
    assert d.deriv_mant.offset == 224 and d.deriv_exp.offset == 232
    assert native.ctypes.sizeof(d) == 240
    p = native.PostDesc
    assert p.skew.offset == 48 and p.proj.offset == 80 and p.mod.offset == 84
    assert p.w_scale.offset == 88 and p.mod_scale.offset == 96 and native.ctypes.sizeof(p) == 104


def test_argument_errors_without_gpu():
    """Argument checks that return before any device call: a null context
    or a negative grid width is FSG_E_INVALID with a message."""
    import ctypes
    lib = native.load()
    st = native.Stats()
    rc = lib.fsg_tile_grid(None, None, 0, 8, None, None, None, None, None, ctypes.byref(st))
    assert rc == native.FSG_E_INVALID
    assert b"null context" in lib.fsg_last_error()
    rc = lib.fsg_tile_grid_device(None, None, 0, 0, None, None, None, None, None, None)
    assert rc == native.FSG_E_INVALID
