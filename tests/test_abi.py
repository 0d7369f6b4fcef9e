"""The C-ABI library loads and exports every symbol include/kpm_amd.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kpm_amd.h")
LIB = os.path.join(ROOT, "kwant_amd", "libkpm_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kpm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("kpm_plan_create", "kpm_cheb_steps", "kpm_block_dots", "kpm_spmm",
                     "kpm_corr_gemm", "kpm_last_error"):
        assert required in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_all_declared_symbols():
    import torch  # noqa: F401  (the library shares torch's HIP runtime)
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    lib.kpm_abi_version.restype = ctypes.c_int
    assert lib.kpm_abi_version() >= 100


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_engine_binding_covers_header():
    from kwant_amd import _engine
    assert sorted(_engine.EXPORTS) == declared_functions()


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_argument_errors_without_gpu():
    """Argument validation happens before any device work."""
    import torch  # noqa: F401
    lib = ctypes.CDLL(LIB)
    lib.kpm_last_error.restype = ctypes.c_char_p
    out = ctypes.c_void_p()
    seg = (ctypes.c_int32 * 2)(0, 10)
    rc = lib.kpm_plan_create(10, 3, 1, 1, seg, -1, ctypes.byref(out))  # width 3
    assert rc != 0 and b"power of two" in lib.kpm_last_error()
    rc = lib.kpm_plan_create(10, 4, 7, 1, seg, -1, ctypes.byref(out))  # bad type
    assert rc == 1
    seg_bad = (ctypes.c_int32 * 2)(0, 9)
    rc = lib.kpm_plan_create(10, 4, 1, 1, seg_bad, -1, ctypes.byref(out))
    assert rc == 1 and b"cover" in lib.kpm_last_error()


def test_pow2_widths():
    from kwant_amd._engine import pow2_widths
    assert pow2_widths(10, 64) == [8, 2]
    assert pow2_widths(64, 64) == [64]
    assert pow2_widths(100, 64) == [64, 32, 4]
    assert pow2_widths(1, 64) == [1]
    assert sum(pow2_widths(1000, 16)) == 1000


def test_no_gpu_no_fallback():
    """Without a visible GPU the product path raises; it never computes on
    the host."""
    import numpy as np
    import scipy.sparse as sp
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from kwant_amd import kpm
    h = sp.diags([np.ones(9), np.ones(9)], [1, -1], format='csr', dtype=complex)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        kpm.SpectralDensity(h, num_moments=10, num_vectors=1, bounds=(-2.1, 2.1))


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_tiled_argument_errors_without_gpu():
    """The two-step entry points validate their arguments before any device
    work: shapes, NULL arrays, tile boundaries."""
    import torch  # noqa: F401
    from kwant_amd._engine import KpmCsr
    lib = ctypes.CDLL(LIB)
    lib.kpm_last_error.restype = ctypes.c_char_p
    count = ctypes.c_int32()
    assert lib.kpm_tiled_variant_count(64, 1, ctypes.byref(count)) == 0 and count.value >= 3
    assert lib.kpm_tiled_variant_count(3, 1, ctypes.byref(count)) == 0 and count.value == 0
    gw, waves, rows = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    for v in range(lib_count(lib, 64, 1)):
        assert lib.kpm_tiled_variant_info(64, 1, v, ctypes.byref(gw), ctypes.byref(waves),
                                          ctypes.byref(rows)) == 0
        # the largest tile is four workgroup iterations of rows
        lpr = gw.value // 2
        assert rows.value == 4 * waves.value * 64 // lpr
    assert lib.kpm_tiled_variant_info(64, 1, 99, ctypes.byref(gw), ctypes.byref(waves),
                                      ctypes.byref(rows)) == 1
    out = ctypes.c_void_p()
    tb = (ctypes.c_int32 * 3)(0, 5, 10)
    csr = KpmCsr(10, 10, 30, None, None, None, 1, 4)
    rc = lib.kpm_tiled_create(ctypes.byref(csr), 64, 2, tb, -1, ctypes.byref(out))
    assert rc == 1 and b"NULL" in lib.kpm_last_error()
    dummy = (ctypes.c_int32 * 1)(0)
    csr = KpmCsr(10, 10, 30, ctypes.addressof(dummy), ctypes.addressof(dummy),
                 ctypes.addressof(dummy), 1, 4)
    bad = (ctypes.c_int32 * 3)(0, 5, 9)
    rc = lib.kpm_tiled_create(ctypes.byref(csr), 64, 2, bad, -1, ctypes.byref(out))
    assert rc == 1 and b"cover" in lib.kpm_last_error()
    empty = (ctypes.c_int32 * 3)(0, 10, 10)
    rc = lib.kpm_tiled_create(ctypes.byref(csr), 64, 2, empty, -1, ctypes.byref(out))
    assert rc == 1 and b"at least one row" in lib.kpm_last_error()
    rc = lib.kpm_tiled_create(ctypes.byref(csr), 3, 2, tb, -1, ctypes.byref(out))
    assert rc == 3  # no tiled kernel for width 3
    rc = lib.kpm_cheb_steps2(None, None, None, None, 0, 4, None, 0, 0, 1, 1, None)
    assert rc == 1


def lib_count(lib, width, value_type):
    count = ctypes.c_int32()
    lib.kpm_tiled_variant_count(width, value_type, ctypes.byref(count))
    return count.value


def test_march_argument_errors_without_gpu():
    """The lattice-stencil work list validates its arguments before any device
    work; MarchPlan.shape_ok accepts the widths the library has kernels for."""
    import torch  # noqa: F401
    from kwant_amd._engine import MarchPlan
    lib = ctypes.CDLL(LIB)
    lib.kpm_last_error.restype = ctypes.c_char_p
    out = ctypes.c_void_p()
    sides = (ctypes.c_int32 * 2)(3, 5)
    create = lib.kpm_march_create
    assert create(2, sides, 16, 1, 5, 1, 1, ctypes.byref(out)) == 1          # waves
    assert create(2, sides, 16, 1, 4, 0, 1, ctypes.byref(out)) == 1          # depth
    assert create(0, sides, 16, 1, 4, 1, 1, ctypes.byref(out)) == 1          # no systems
    assert create(2, sides, 4, 1, 4, 1, 1, ctypes.byref(out)) == 3           # 2 lanes per row
    bad = (ctypes.c_int32 * 1)(0)
    assert create(1, bad, 16, 1, 4, 1, 1, ctypes.byref(out)) == 1            # side 0
    for width, vt in ((8, 1), (16, 1), (64, 1), (4, 2), (16, 2), (64, 2)):
        for shape in range(len(MarchPlan.SHAPES)):
            assert MarchPlan.shape_ok(width, vt, shape)
        assert not MarchPlan.shape_ok(width, vt, len(MarchPlan.SHAPES))
        for tile_y in (0, 1025):
            rc = create(2, sides, width, vt, 8, 1, tile_y, ctypes.byref(out))
            assert rc == 1 and b"tile_y" in lib.kpm_last_error(), (width, vt, tile_y)
    assert not MarchPlan.shape_ok(4, 1, 0) and not MarchPlan.shape_ok(256, 2, 0)
    rc = lib.kpm_cheb_steps_march(None, None, None, ctypes.c_double(1.0), None, 0, 4, None, 0,
                                  0, 1, 1, None)
    assert rc == 1
