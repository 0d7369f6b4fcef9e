"""ctypes binding of libkpm_amd.so and the device-side objects built on it.

Device memory and streams come from PyTorch-ROCm (tensors are the buffers,
``torch.cuda.current_stream()`` the stream); every computation runs in the
HIP kernels of ``kwant_amd/csrc``.  There is no CPU fallback: without the
built library or without a visible GPU every entry point raises.
"""
import ctypes
import os

import numpy as np
import scipy.sparse

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkpm_amd.so")

KPM_OK = 0
KPM_ERR_ARG = 1
KPM_C64 = 1
KPM_C128 = 2
DOTS_NONE = 0
DOTS_DOUBLING = 1
DOTS_PROJECT = 2
DOTS_LOCAL = 3
DOTS_LOCAL_COLS = 4
GEMM_FP64 = 0
GEMM_FP32 = 1
GEMM_FP32X3 = 2
GEMM_VECTOR_ALU = 16
MLP_PLAIN = 0
MLP_BIAS_GELU = 1
MLP_GELU_GRAD = 2
MLP_BF16X3 = 16
MLP_ACCUMULATE = 32

# Exported symbols, in the order include/kpm_amd.h declares them.
EXPORTS = (
    "kpm_last_error", "kpm_abi_version", "kpm_device_info", "kpm_variant_count",
    "kpm_variant_info", "kpm_plan_create", "kpm_plan_destroy", "kpm_plan_set_scale",
    "kpm_plan_set_dots_grid", "kpm_plan_set_step", "kpm_plan_get_step", "kpm_cheb_steps",
    "kpm_cheb_steps_static", "kpm_bsr_steps", "kpm_tiled_variant_count",
    "kpm_tiled_variant_info", "kpm_tiled_create", "kpm_tiled_destroy", "kpm_tiled_info",
    "kpm_cheb_steps2", "kpm_block_dots",
    "kpm_bond_records", "kpm_spmm", "kpm_corr_gemm", "kpm_corr_gemm_sliced_work", "kpm_corr_gemm_sliced",
    "kpm_corr_gemm_islice_work", "kpm_corr_gemm_islice", "kpm_islice_prep_bytes",
    "kpm_islice_prep_rows", "kpm_corr_gemm_islice_prep_work", "kpm_corr_gemm_islice_prep",
    "kpm_moment_shift", "kpm_cubic4_nnz", "kpm_gen_cubic4", "kpm_gen_cubic4_onsite",
    "kpm_march_create", "kpm_march_destroy", "kpm_cheb_steps_march",
    "kpm_mlp_gemm", "kpm_mlp_colsum", "kpm_mlp_mse",
    "kpm_mlp_ensemble", "kpm_mlp_adam",
)


class KpmCsr(ctypes.Structure):
    _fields_ = [
        ("n_rows", ctypes.c_int32),
        ("n_cols", ctypes.c_int32),
        ("nnz", ctypes.c_int64),
        ("indptr", ctypes.c_void_p),
        ("indices", ctypes.c_void_p),
        ("values", ctypes.c_void_p),
        ("value_type", ctypes.c_int32),
        ("row_nnz_hint", ctypes.c_int32),
    ]


def row_nnz_hint(row_lengths, quantile=0.99):
    """A row length that ``quantile`` of the rows stay within (the recursion
    kernel's prefetch depth follows it; results never depend on it).
    ``row_lengths``: numpy array or torch tensor (on any device)."""
    cap = 64
    if isinstance(row_lengths, np.ndarray):
        if row_lengths.size == 0:
            return 0
        counts = np.bincount(np.minimum(row_lengths, cap), minlength=cap + 1)
    else:
        import torch
        if row_lengths.numel() == 0:
            return 0
        counts = torch.bincount(torch.clamp(row_lengths, max=cap).to(torch.int64),
                                minlength=cap + 1).cpu().numpy()
    cum = np.cumsum(counts)
    return int(np.searchsorted(cum, quantile * cum[-1]))


class MlpGemmDesc(ctypes.Structure):
    _fields_ = [
        ("a", ctypes.c_void_p), ("a_rs", ctypes.c_int64), ("a_cs", ctypes.c_int64),
        ("a_bs", ctypes.c_int64),
        ("b", ctypes.c_void_p), ("b_rs", ctypes.c_int64), ("b_cs", ctypes.c_int64),
        ("b_bs", ctypes.c_int64),
        ("c", ctypes.c_void_p), ("c_rs", ctypes.c_int64), ("c_bs", ctypes.c_int64),
        ("c2", ctypes.c_void_p),
        ("bias", ctypes.c_void_p), ("bias_bs", ctypes.c_int64),
        ("aux", ctypes.c_void_p), ("aux_rs", ctypes.c_int64), ("aux_bs", ctypes.c_int64),
        ("m", ctypes.c_int32), ("n", ctypes.c_int32), ("k", ctypes.c_int32),
        ("batch", ctypes.c_int32), ("flags", ctypes.c_int32),
    ]


_lib = None


def load_library(path=LIB_PATH):
    """Load libkpm_amd.so (after torch, so that both share one HIP runtime)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            "libkpm_amd.so is not built (expected at %s); run "
            "`python -m kwant_amd.build`" % path)
    import torch  # noqa: F401  -- loads torch's HIP runtime first
    lib = ctypes.CDLL(path)
    i32, i64, vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
    f32 = ctypes.c_float
    pi32 = ctypes.POINTER(ctypes.c_int32)
    sig = {
        "kpm_last_error": (ctypes.c_char_p, []),
        "kpm_abi_version": (ctypes.c_int, []),
        "kpm_device_info": (ctypes.c_int, [i32, pi32, pi32]),
        "kpm_variant_count": (ctypes.c_int, [i32, i32, pi32]),
        "kpm_variant_info": (ctypes.c_int, [i32, i32, i32] + [pi32] * 7),
        "kpm_plan_create": (ctypes.c_int, [i32, i32, i32, i32, pi32, i32,
                                           ctypes.POINTER(vp)]),
        "kpm_plan_destroy": (ctypes.c_int, [vp]),
        "kpm_plan_set_scale": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double)]),
        "kpm_plan_set_dots_grid": (ctypes.c_int, [vp, vp, vp]),
        "kpm_plan_set_step": (ctypes.c_int, [vp, i32, vp]),
        "kpm_plan_get_step": (ctypes.c_int, [vp, pi32, vp]),
        "kpm_cheb_steps": (ctypes.c_int, [vp, ctypes.POINTER(KpmCsr), vp, i64, i32, vp,
                                          vp, i32, i32, i32, i32, vp]),
        "kpm_cheb_steps_static": (ctypes.c_int, [ctypes.POINTER(KpmCsr), ctypes.c_double,
                                                 ctypes.c_double, i32, i32, vp, i64, i32, vp,
                                                 vp, i32, i32, vp, i32, i32, i32, vp]),
        "kpm_bsr_steps": (ctypes.c_int, [vp, ctypes.POINTER(KpmBsr), vp, i64, i32, vp, i32,
                                         i32, i32, i32, vp]),
        "kpm_tiled_variant_count": (ctypes.c_int, [i32, i32, pi32]),
        "kpm_tiled_variant_info": (ctypes.c_int, [i32, i32, i32, pi32, pi32, pi32]),
        "kpm_tiled_create": (ctypes.c_int, [ctypes.POINTER(KpmCsr), i32, i32, pi32, i32,
                                            ctypes.POINTER(vp)]),
        "kpm_tiled_destroy": (ctypes.c_int, [vp]),
        "kpm_tiled_info": (ctypes.c_int, [vp, pi32, pi32, pi32] +
                           [ctypes.POINTER(ctypes.c_int64)] * 3 + [pi32, pi32]),
        "kpm_cheb_steps2": (ctypes.c_int, [vp, vp, ctypes.POINTER(KpmCsr), vp, i64, i32, vp, i32,
                                           i32, i32, i32, vp]),
        "kpm_block_dots": (ctypes.c_int, [vp, vp, vp, vp, i32, vp]),
        "kpm_bond_records": (ctypes.c_int, [ctypes.POINTER(KpmCsr), vp, vp, vp, vp, i32, i32,
                                            vp, vp]),
        "kpm_spmm": (ctypes.c_int, [ctypes.POINTER(KpmCsr), vp, i64, vp, i64, i32, i32, vp]),
        "kpm_corr_gemm": (ctypes.c_int, [vp, i64, vp, i64, vp, i64, i32, i32, i64, i32,
                                         i32, vp, i64, vp]),
        "kpm_corr_gemm_sliced_work": (ctypes.c_int, [i32, i32, i64, i32,
                                                     ctypes.POINTER(ctypes.c_int64)]),
        "kpm_corr_gemm_sliced": (ctypes.c_int, [vp, i64, vp, i64, vp, i64, i32, i32, i64, i32,
                                                vp, i64, vp]),
        "kpm_corr_gemm_islice_work": (ctypes.c_int, [i32, i32, i64, i32,
                                                     ctypes.POINTER(ctypes.c_int64)]),
        "kpm_corr_gemm_islice": (ctypes.c_int, [vp, i64, vp, i64, vp, i64, i32, i32, i64, i32,
                                                vp, i64, vp]),
        "kpm_islice_prep_bytes": (ctypes.c_int, [i32, i64, i32, ctypes.POINTER(ctypes.c_int64)]),
        "kpm_islice_prep_rows": (ctypes.c_int, [vp, i64, i32, i32, i32, i64, i32, vp, i64, vp]),
        "kpm_corr_gemm_islice_prep_work": (ctypes.c_int, [i32, i32, i64, i32,
                                                          ctypes.POINTER(ctypes.c_int64)]),
        "kpm_corr_gemm_islice_prep": (ctypes.c_int, [vp, i64, i32, vp, i64, i32, i64, i32, vp,
                                                     i64, vp, i64, vp]),
        "kpm_moment_shift": (ctypes.c_int, [vp, i64, i64, i32, i64, i32, vp, vp,
                                            ctypes.c_double, vp, vp, i64, i64, vp]),
        "kpm_cubic4_nnz": (ctypes.c_int, [i32, ctypes.POINTER(ctypes.c_int64)]),
        "kpm_gen_cubic4": (ctypes.c_int, [i32, ctypes.c_uint64, ctypes.c_double,
                                          ctypes.c_double, i64, i64, vp, vp, vp, i32, vp, vp]),
        "kpm_gen_cubic4_onsite": (ctypes.c_int, [i32, ctypes.c_uint64, ctypes.c_double,
                                                 ctypes.c_double, i64, vp, i32, vp, vp]),
        "kpm_march_create": (ctypes.c_int, [i32, pi32, i32, i32, i32, i32, i32,
                                            ctypes.POINTER(vp)]),
        "kpm_march_destroy": (ctypes.c_int, [vp]),
        "kpm_cheb_steps_march": (ctypes.c_int, [vp, vp, vp, ctypes.c_double, vp, i64, i32, vp,
                                                i32, i32, i32, i32, vp]),
        "kpm_mlp_gemm": (ctypes.c_int, [ctypes.POINTER(MlpGemmDesc), vp]),
        "kpm_mlp_colsum": (ctypes.c_int, [vp, i64, i64, i32, i32, i32, vp, i64, vp]),
        "kpm_mlp_mse": (ctypes.c_int, [vp, vp, vp, i64, f32, vp, i32, vp]),
        "kpm_mlp_ensemble": (ctypes.c_int, [vp, i64, i64, i32, i64, i32, vp, i64, vp, vp]),
        "kpm_mlp_adam": (ctypes.c_int, [vp, vp, vp, vp, i64, f32, f32, f32, f32, i32, vp, vp,
                                        f32, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what):
    if rc == KPM_OK:
        return
    msg = load_library().kpm_last_error().decode(errors="replace")
    if rc == KPM_ERR_ARG:
        raise ValueError("%s: %s" % (what, msg))
    raise RuntimeError("%s failed (status %d): %s" % (what, rc, msg))


def torch_device():
    """The GPU the engine runs on (torch's current device)."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError(
            "kwant_amd runs its KPM kernels on an AMD GPU (ROCm) and none is "
            "visible to PyTorch; there is no CPU fallback")
    load_library()
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def complex_dtype(precision):
    import torch
    if precision in ("c128", "complex128", np.complex128, torch.complex128):
        return torch.complex128, KPM_C128
    if precision in ("c64", "complex64", np.complex64, torch.complex64):
        return torch.complex64, KPM_C64
    raise ValueError("precision must be 'complex128' or 'complex64'")


def variant_count(width, value_type):
    """Number of recursion kernel launch shapes for (width, value type)."""
    n = ctypes.c_int32()
    check(load_library().kpm_variant_count(width, value_type, ctypes.byref(n)),
          "kpm_variant_count")
    return n.value


def variant_info(width, value_type, variant):
    """(lanes per row, complex values per lane, waves, residency rounds,
    row slots per wave iteration, interleave chunk, spare workgroups per CU)."""
    vals = [ctypes.c_int32() for _ in range(7)]
    check(load_library().kpm_variant_info(width, value_type, variant,
                                          *[ctypes.byref(v) for v in vals]),
          "kpm_variant_info")
    return tuple(v.value for v in vals)


def device_info():
    lib = load_library()
    cu, arch = ctypes.c_int32(), ctypes.c_int32()
    check(lib.kpm_device_info(0, ctypes.byref(cu), ctypes.byref(arch)), "kpm_device_info")
    return cu.value, arch.value


class DeviceCSR:
    """A CSR matrix resident on the GPU (int32 indices, complex values).

    Keeps kwant's CSR exactly as given (entry order within rows included),
    so the SpMM sums each row in the order scipy's csr_matvec does.  With
    ``order`` (kwant_amd.ordering) the rows and columns are stored permuted,
    P H P^T with device row i = caller row order[i]; vectors cross the
    boundary through rows_in / rows_out, and every row sum stays the same
    sequence of operations."""

    def __init__(self, matrix, precision="c128", order=None):
        import torch
        from .ordering import permute_csr
        dev = torch_device()
        csr = scipy.sparse.csr_matrix(matrix)
        self.order = self.inverse = None
        if order is not None:
            csr = permute_csr(csr, order)
            self.order = np.asarray(order, dtype=np.int64)
            self.inverse = np.empty_like(self.order)
            self.inverse[self.order] = np.arange(len(self.order))
        if csr.nnz >= 2 ** 31 or csr.shape[0] >= 2 ** 31:
            raise ValueError("matrix too large for int32 CSR indices")
        self.shape = csr.shape
        self.nnz = int(csr.nnz)
        self.precision = precision
        cdt, self.value_type = complex_dtype(precision)
        np_cdt = np.complex128 if self.value_type == KPM_C128 else np.complex64
        self.indptr = torch.from_numpy(np.ascontiguousarray(csr.indptr, dtype=np.int32)).to(dev)
        self.indices = torch.from_numpy(
            np.ascontiguousarray(csr.indices, dtype=np.int32)).to(dev)
        vals = np.ascontiguousarray(csr.data, dtype=np.complex128).astype(np_cdt)
        if vals.size == 0:
            vals = np.zeros(1, dtype=np_cdt)
        self.values = torch.from_numpy(vals).to(dev)
        assert self.values.dtype == cdt
        self.desc = KpmCsr(int(csr.shape[0]), int(csr.shape[1]), self.nnz,
                           self.indptr.data_ptr(), self.indices.data_ptr(),
                           self.values.data_ptr(), self.value_type,
                           row_nnz_hint(np.diff(csr.indptr)))

    def row_block(self, r0, r1):
        """Rows r0 .. r1-1 of this device matrix as an (r1 - r0) x n_cols
        DeviceCSR that shares its column indices and values (an SpMM with it
        writes only those rows, wherever its output points)."""
        if not 0 <= r0 < r1 <= self.shape[0]:
            raise ValueError("row block outside the matrix")
        k0, k1 = int(self.indptr[r0].item()), int(self.indptr[r1].item())
        sub = DeviceCSR.__new__(DeviceCSR)
        sub.order = sub.inverse = None
        sub.shape = (r1 - r0, self.shape[1])
        sub.nnz = k1 - k0
        sub.precision, sub.value_type = self.precision, self.value_type
        sub.indptr = (self.indptr[r0:r1 + 1] - k0).contiguous()
        sub.indices = self.indices[k0:max(k1, k0 + 1)] if self.indices.numel() else self.indices
        sub.values = self.values[k0:max(k1, k0 + 1)]
        sub._parent = self  # the shared arrays stay alive
        sub.desc = KpmCsr(r1 - r0, int(self.shape[1]), sub.nnz, sub.indptr.data_ptr(),
                          sub.indices.data_ptr(), sub.values.data_ptr(), self.value_type,
                          self.desc.row_nnz_hint)
        return sub

    def rows_in(self, host):
        """Caller-order rows (axis 0) of a host array, in device order."""
        return host if self.order is None else np.asarray(host)[self.order]

    def rows_out(self, host):
        """Device-order rows (axis 0) of a host array, in caller order."""
        return host if self.order is None else np.asarray(host)[self.inverse]

    @classmethod
    def from_device_arrays(cls, shape, indptr, indices, values, precision):
        """Wrap CSR arrays that already live on the GPU (generated there)."""
        self = cls.__new__(cls)
        self.order = self.inverse = None
        cdt, self.value_type = complex_dtype(precision)
        if values.dtype != cdt:
            raise ValueError("values dtype does not match precision")
        self.shape = tuple(shape)
        self.nnz = int(indices.numel())
        self.precision = precision
        self.indptr, self.indices, self.values = indptr, indices, values
        self.desc = KpmCsr(int(shape[0]), int(shape[1]), self.nnz, indptr.data_ptr(),
                           indices.data_ptr(), values.data_ptr(), self.value_type,
                           row_nnz_hint(indptr[1:] - indptr[:-1]))
        return self

    def bond_records(self, bra, ket, v, w, width, out, per_column=False):
        """Bond records of the entries of this matrix (see kpm_bond_records):
        out[k] (or out[k, c] with per_column) = i (conj(h) conj(bra_s) v_r -
        conj(w_r) h ket_s) summed over the block's columns."""
        n = self.shape[0]
        for t in (bra, ket, v, w):
            if tuple(t.shape) != (n, width) or not t.is_contiguous():
                raise ValueError("bond operands must be contiguous (N, width) blocks")
        want = (self.nnz, width) if per_column else (self.nnz,)
        import torch
        if out.dtype != torch.complex128 or tuple(out.shape) != want or not out.is_contiguous():
            raise ValueError("bond records must be a contiguous complex128 %s tensor" % (want,))
        if self.nnz == 0:
            return
        check(load_library().kpm_bond_records(
            ctypes.byref(self.desc), bra.data_ptr(), ket.data_ptr(), v.data_ptr(), w.data_ptr(),
            width, int(bool(per_column)), out.data_ptr(), stream_handle()), "kpm_bond_records")

    def spmm(self, x, y, width, n_slots=1, x_stride=0, y_stride=0):
        """y[slot] = self @ x[slot] for row-major blocks of ``width`` columns."""
        lib = load_library()
        check(lib.kpm_spmm(ctypes.byref(self.desc), x.data_ptr(), x_stride, y.data_ptr(),
                           y_stride, width, n_slots, stream_handle()), "kpm_spmm")


class KpmBsr(ctypes.Structure):
    _fields_ = [
        ("n_block_rows", ctypes.c_int32),
        ("n_block_cols", ctypes.c_int32),
        ("nnzb", ctypes.c_int64),
        ("indptr", ctypes.c_void_p),
        ("indices", ctypes.c_void_p),
        ("values", ctypes.c_void_p),
        ("value_type", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
    ]


class DeviceBSR:
    """The same matrix in 4 x 4 block-CSR on the GPU (kpm_bsr_steps): the
    site blocks of a 4-orbital system, the four rows of a site gathered as one
    block.  Built from the column-sorted CSR, so the moments equal the CSR
    path's bit for bit.  Rows stay in the caller's order."""

    format = "bsr"
    BLOCK = 4

    def __init__(self, matrix, precision="c128"):
        import torch
        dev = torch_device()
        csr = scipy.sparse.csr_matrix(matrix)
        n = csr.shape[0]
        if csr.shape[1] != n or n % self.BLOCK:
            raise ValueError("block-CSR needs a square matrix whose size is a multiple of 4")
        csr = csr.copy()
        csr.sort_indices()
        bsr = scipy.sparse.bsr_matrix(csr, blocksize=(self.BLOCK, self.BLOCK))
        bsr.sort_indices()
        if bsr.indices.size * 16 >= 2 ** 31:
            raise ValueError("matrix too large for int32 block indices")
        self.order = self.inverse = None
        self.shape = csr.shape
        self.nnz = int(csr.nnz)                 # entries of the CSR it holds
        self.nnzb = int(bsr.indices.size)       # stored 4 x 4 blocks
        self.precision = precision
        cdt, self.value_type = complex_dtype(precision)
        np_cdt = np.complex128 if self.value_type == KPM_C128 else np.complex64
        self.indptr = torch.from_numpy(np.ascontiguousarray(bsr.indptr, dtype=np.int32)).to(dev)
        self.indices = torch.from_numpy(
            np.ascontiguousarray(bsr.indices, dtype=np.int32)).to(dev)
        vals = np.ascontiguousarray(bsr.data, dtype=np.complex128).astype(np_cdt).reshape(-1)
        if vals.size == 0:
            vals = np.zeros(1, dtype=np_cdt)
        self.values = torch.from_numpy(vals).to(dev)
        self.desc = KpmBsr(n // self.BLOCK, n // self.BLOCK, self.nnzb, self.indptr.data_ptr(),
                           self.indices.data_ptr(), self.values.data_ptr(), self.value_type,
                           self.BLOCK)

    def rows_in(self, host):
        return host

    def rows_out(self, host):
        return host

    @staticmethod
    def variant_ok(width, value_type, variant):
        """Block-CSR steps run 4-wave shapes with at most 16 lanes per row."""
        lpr, _, waves, *_ = variant_info(width, value_type, variant)
        return lpr <= 16 and waves == 4


class Plan:
    """A recursion plan (kernel shape, row partition into segments, workspace)."""

    def __init__(self, n_rows, width, value_type, seg_rows=None, variant=-1):
        lib = load_library()
        if seg_rows is None:
            seg_rows = [0, n_rows]
        seg = np.ascontiguousarray(seg_rows, dtype=np.int32)
        self.n_rows, self.width, self.value_type = n_rows, width, value_type
        self.n_segments = len(seg) - 1
        handle = ctypes.c_void_p()
        check(lib.kpm_plan_create(n_rows, width, value_type, self.n_segments,
                                  seg.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                  variant, ctypes.byref(handle)), "kpm_plan_create")
        self.handle = handle

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            try:
                _lib.kpm_plan_destroy(h)
            except Exception:
                pass
            self.handle = None

    def set_scale(self, ab):
        """(a, b) per segment, flattened: 2 * n_segments values."""
        ab = np.ascontiguousarray(np.asarray(ab, dtype=np.float64).reshape(-1))
        if ab.size != 2 * self.n_segments:
            raise ValueError("need (a, b) for each of the plan's %d segments, got %d values"
                             % (self.n_segments, ab.size))
        check(load_library().kpm_plan_set_scale(
            self.handle, ab.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
            "kpm_plan_set_scale")

    def set_dots_grid(self, a0):
        """Fix the grid of the exact Im <a_{n+1}|a_n> sums from the block's
        start vectors (an (N, width) device block); raises ValueError when
        they hold NaN or Inf."""
        check(load_library().kpm_plan_set_dots_grid(self.handle, a0.data_ptr(),
                                                    stream_handle()),
              "kpm_plan_set_dots_grid")

    def set_step(self, n):
        check(load_library().kpm_plan_set_step(self.handle, n, stream_handle()),
              "kpm_plan_set_step")

    def get_step(self):
        n = ctypes.c_int32()
        check(load_library().kpm_plan_get_step(self.handle, ctypes.byref(n),
                                               stream_handle()), "kpm_plan_get_step")
        return n.value

    def steps(self, mat, ring, slot_stride, slots, n_steps, mode, dots=None,
              out_base=0, w=None):
        out_cap = 0 if dots is None else int(dots.shape[0])
        if isinstance(mat, DeviceBSR):
            if w is not None:
                raise ValueError("block-CSR steps have no projection modes")
            check(load_library().kpm_bsr_steps(
                self.handle, ctypes.byref(mat.desc), ring.data_ptr(), slot_stride, slots,
                None if dots is None else dots.data_ptr(), out_base, out_cap, n_steps, mode,
                stream_handle()), "kpm_bsr_steps")
            return
        check(load_library().kpm_cheb_steps(
            self.handle, ctypes.byref(mat.desc), ring.data_ptr(), slot_stride, slots,
            None if w is None else w.data_ptr(),
            None if dots is None else dots.data_ptr(), out_base, out_cap, n_steps, mode,
            stream_handle()), "kpm_cheb_steps")

    def block_dots(self, x, out, mode, w=None):
        check(load_library().kpm_block_dots(
            self.handle, x.data_ptr(), None if w is None else w.data_ptr(), out.data_ptr(),
            mode, stream_handle()), "kpm_block_dots")


def tiled_variant_count(width, value_type):
    """Number of two-step (tiled) launch shapes for (width, value type)."""
    n = ctypes.c_int32()
    check(load_library().kpm_tiled_variant_count(width, value_type, ctypes.byref(n)),
          "kpm_tiled_variant_count")
    return n.value


def tiled_variant_info(width, value_type, variant):
    """(columns per workgroup, waves per workgroup, suggested tile rows)."""
    vals = [ctypes.c_int32() for _ in range(3)]
    check(load_library().kpm_tiled_variant_info(width, value_type, variant,
                                                *[ctypes.byref(v) for v in vals]),
          "kpm_tiled_variant_info")
    return tuple(v.value for v in vals)


class TiledPlan:
    """Tiles of consecutive device rows of ``mat`` (tile t owns rows
    tile_begin[t] .. tile_begin[t+1]-1) for the two-step launches
    (kpm_cheb_steps2): the library computes each tile's halo and the LDS slot
    of every entry from the matrix pattern, and packs each row's first
    entries."""

    def __init__(self, mat, width, tile_begin, variant=-1):
        lib = load_library()
        tb = np.ascontiguousarray(tile_begin, dtype=np.int32)
        pi32 = ctypes.POINTER(ctypes.c_int32)
        handle = ctypes.c_void_p()
        check(lib.kpm_tiled_create(ctypes.byref(mat.desc), int(width), len(tb) - 1,
                                   tb.ctypes.data_as(pi32), int(variant),
                                   ctypes.byref(handle)), "kpm_tiled_create")
        self.mat = mat   # the packed entries are a copy of this matrix's
        self.handle = handle
        self.width, self.n_rows, self.value_type = int(width), int(mat.shape[0]), mat.value_type
        self.tile_begin = tb

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            try:
                _lib.kpm_tiled_destroy(h)
            except Exception:
                pass
            self.handle = None

    def info(self):
        """Tiles, largest tile, largest halo, all halo rows and their entries,
        the gathered rows of a_n, workgroups, LDS bytes."""
        vals = [ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64(),
                ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()]
        check(load_library().kpm_tiled_info(self.handle, *[ctypes.byref(v) for v in vals]),
              "kpm_tiled_info")
        keys = ("tiles", "max_tile_rows", "max_halo_rows", "halo_rows", "halo_entries",
                "gather_rows", "workgroups", "lds_bytes")
        return dict(zip(keys, (v.value for v in vals)))

    def kernel_bytes(self, nnz, s_v, s_m):
        """Bytes one two-step launch reads and writes if nothing is shared
        between tiles: a_{n-1} of every tile's rows and halo, a_n of the rows
        each tile gathers, a_{n+1} and a_{n+2} written, the CSR of the tile
        and halo rows (phase A) and of the tile rows with their local indices
        (phase B), the halo ids."""
        i = self.info()
        n, w = self.n_rows, self.width
        rows_a = n + i["halo_rows"]
        vec = w * s_v * (rows_a + i["gather_rows"] + 2 * n)
        csr = 4 * rows_a + (nnz + i["halo_entries"]) * (4 + s_m) + 4 * n + nnz * (4 + s_m)
        return vec + csr + 4 * i["halo_rows"]

    def steps(self, plan, mat, ring, slot_stride, slots, n_pairs, mode, dots=None, out_base=0):
        out_cap = 0 if dots is None else int(dots.shape[0])
        check(load_library().kpm_cheb_steps2(
            plan.handle, self.handle, ctypes.byref(mat.desc), ring.data_ptr(), slot_stride,
            slots, None if dots is None else dots.data_ptr(), out_base, out_cap, n_pairs, mode,
            stream_handle()), "kpm_cheb_steps2")


def cheb_steps_static(mat, ab, ring, slots, first_step, n_steps, mode, dots=None,
                      out_base=0, w=None, grid_exp=None, variant=0):
    """kpm_cheb_steps_static: ``n_steps`` steps from step ``first_step`` of
    the (slots, N, width) ring, with no plan and no device allocation (the
    library's static workspace).  ``grid_exp``: (width,) int32 device tensor,
    written when first_step == 0, read otherwise (doubling mode)."""
    width = int(ring.shape[-1])
    out_cap = 0 if dots is None else int(dots.shape[0])
    a, b = (float(x) for x in ab)
    check(load_library().kpm_cheb_steps_static(
        ctypes.byref(mat.desc), a, b, width, variant, ring.data_ptr(),
        mat.shape[0] * width, slots, None if w is None else w.data_ptr(),
        None if grid_exp is None else grid_exp.data_ptr(), first_step, n_steps,
        None if dots is None else dots.data_ptr(), out_base, out_cap, mode,
        stream_handle()), "kpm_cheb_steps_static")


def pow2_widths(count, max_width):
    """Split ``count`` vectors into power-of-two block widths <= max_width."""
    out = []
    while count > 0:
        w = max_width
        while w > count:
            w //= 2
        out.append(w)
        count -= w
    return out


class BlockRecursion:
    """Chebyshev recursion of one N x W block of vectors on the GPU.

    The ring holds a_{n-1}, a_n in two slots (in-place update), or every
    a_n when ``store_all`` (Correlator expansions)."""

    def __init__(self, mat, ab, vectors, width, mode, w=None, store_slots=None,
                 variant=-1, plan=None, wrap=False, slot_pad=0, tiles=None):
        """``wrap``: a ring of ``store_slots`` slots that the caller drains
        as it goes -- a_k lives in slot k mod slots, and the steps may run
        past the ring's end (the caller uses each a_k before slots steps
        overwrite it).  ``slot_pad``: elements of padding after each slot, so
        that slots read together as the rows of a GEMM are not a power of two
        apart (the ring is then a strided (slots, N, W) view).  ``tiles``: a
        TiledPlan of ``mat`` for this width; the steps then run two per launch
        (step 0 and an odd last step one per launch) on a ring of four
        slots."""
        import torch
        self.wrap = bool(wrap)
        self.tiles = tiles
        if tiles is not None:
            if mode not in (DOTS_DOUBLING, DOTS_NONE):
                raise ValueError("two-step launches take the doubling or no-dots mode")
            if tiles.width != width or tiles.n_rows != mat.shape[0]:
                raise ValueError("the tiling was built for another width or matrix")
            if store_slots is None:  # a rolling ring: a_k in slot k mod 4
                store_slots = 4
                self.wrap = True
        dev = torch_device()
        n = mat.shape[0]
        self.mat, self.width, self.mode, self.n_rows = mat, width, mode, n
        self.variant = variant
        self.cdt, self.vt = complex_dtype(mat.precision)
        self.plan = plan if plan is not None else Plan(n, width, self.vt, variant=variant)
        ab = np.asarray(ab, dtype=np.float64).reshape(-1)
        if ab.size == 2 and self.plan.n_segments > 1:
            ab = np.tile(ab, self.plan.n_segments)   # one (a, b) for every segment
        self.plan.set_scale(ab)
        self.slots = 2 if store_slots is None else int(store_slots)
        if self.slots < 2:
            raise ValueError("a ring needs at least two slots")
        self.pitch = n * width + int(slot_pad)
        if slot_pad:
            flat = torch.zeros((self.slots, self.pitch), dtype=self.cdt, device=dev)
            self.ring = flat[:, :n * width].view(self.slots, n, width)
        else:
            self.ring = torch.zeros((self.slots, n, width), dtype=self.cdt, device=dev)
        self.ring[0, :, :vectors.shape[1]] = vectors
        if mode == DOTS_DOUBLING:
            self.plan.set_dots_grid(self.ring[0])
        self.w = w
        self.n = 0  # steps done: ring holds a_0 .. a_n
        self.plan.set_step(0)

    @property
    def slot_stride(self):
        return self.pitch

    def step(self, n_steps, want_dots=True):
        """Advance ``n_steps``; returns the (n_steps, W, 4) FP64 dots on host."""
        import torch
        if n_steps <= 0:
            return np.zeros((0, self.width, 4))
        if self.n + n_steps >= self.slots and self.slots > 2 and not self.wrap:
            raise ValueError("ring too small for the requested moments")
        mode = self.mode if want_dots else DOTS_NONE
        dots = None
        if mode != DOTS_NONE:
            dots = torch.empty((n_steps, 1, self.width, 4), dtype=torch.float64,
                               device=self.ring.device)
        if self.tiles is None:
            self.plan.steps(self.mat, self.ring, self.slot_stride, self.slots, n_steps, mode,
                            dots=dots, out_base=self.n, w=self.w)
        else:
            base, left = self.n, n_steps
            if self.n == 0:  # a_1 has no a_{-1}: one step on its own
                self.plan.steps(self.mat, self.ring, self.slot_stride, self.slots, 1, mode,
                                dots=dots, out_base=base, w=self.w)
                left -= 1
            if left >= 2:
                self.tiles.steps(self.plan, self.mat, self.ring, self.slot_stride, self.slots,
                                 left // 2, mode, dots=dots, out_base=base)
            if left % 2:
                self.plan.steps(self.mat, self.ring, self.slot_stride, self.slots, 1, mode,
                                dots=dots, out_base=base, w=self.w)
        self.n += n_steps
        if dots is None:
            return None
        return dots[:, 0].cpu().numpy()

    def _local_shape(self):
        if self.mode not in (DOTS_LOCAL, DOTS_LOCAL_COLS):
            raise ValueError("not a site-resolved recursion")
        return (self.n_rows, self.width) if self.mode == DOTS_LOCAL_COLS else (self.n_rows,)

    def local_steps(self, n_steps):
        """Advance ``n_steps`` in a site-resolved mode; returns the device
        records <w|a_{n+1}> per row (device row order) of every step:
        (n_steps, N) complex128 for DOTS_LOCAL (summed over the block's
        columns), (n_steps, N, W) for DOTS_LOCAL_COLS."""
        import torch
        shape = self._local_shape()
        if self.n + n_steps >= self.slots and self.slots > 2:
            raise ValueError("ring too small for the requested moments")
        out = torch.empty((n_steps,) + shape, dtype=torch.complex128, device=self.ring.device)
        if n_steps > 0:
            self.plan.steps(self.mat, self.ring, self.slot_stride, self.slots, n_steps,
                            self.mode, dots=out, out_base=self.n, w=self.w)
            self.n += n_steps
        return out

    def local_now(self):
        """The site-resolved record of the current a_n alone (no step)."""
        import torch
        out = torch.empty(self._local_shape(), dtype=torch.complex128, device=self.ring.device)
        self.plan.block_dots(self.slot(self.n), out, self.mode, w=self.w)
        return out

    def slot(self, k):
        return self.ring[k % self.slots]

    def norms(self):
        """(W, 4) dots of the current a_n alone (odd-moment tail)."""
        import torch
        out = torch.empty((1, self.width, 4), dtype=torch.float64, device=self.ring.device)
        self.plan.block_dots(self.slot(self.n), out, self.mode, w=self.w)
        return out[0].cpu().numpy()

    def last_two(self, col):
        """(a_{n-1}, a_n) of column ``col`` as host complex128 arrays."""
        return tuple(self.mat.rows_out(self.slot(k)[:, col].cpu().numpy().astype(np.complex128))
                     for k in (self.n - 1, self.n))


def corr_gemm(a, b, mu, precision, k_chunk=256, work=None):
    """mu += conj(a) @ b.T on the GPU (a: m x k, b: n x k complex tensors with
    unit stride along k; mu: m x n complex128)."""
    import torch
    m, k = a.shape
    n = b.shape[0]
    if b.shape[1] != k or tuple(mu.shape) != (m, n):
        raise ValueError("corr_gemm shape mismatch")
    if mu.dtype != torch.complex128:
        raise ValueError("mu must be complex128")
    want = torch.complex64 if (precision & ~GEMM_VECTOR_ALU) == GEMM_FP32 else torch.complex128
    if a.dtype != want or b.dtype != want:
        raise ValueError("operand dtype does not match the GEMM precision")
    if a.stride(1) != 1 or b.stride(1) != 1 or mu.stride(1) != 1:
        raise ValueError("operands must be contiguous along k")
    if work is None:
        tiles = -(-m // 64) * -(-n // 64)
        slices = max(1, min(-(-2048 // tiles), -(-k // 4096)))
        work = torch.empty(2 * m * n * slices, dtype=torch.float64, device=mu.device)
    check(load_library().kpm_corr_gemm(
        a.data_ptr(), a.stride(0), b.data_ptr(), b.stride(0), mu.data_ptr(), mu.stride(0),
        m, n, k, precision, k_chunk, work.data_ptr(), work.numel(), stream_handle()),
        "kpm_corr_gemm")


def corr_gemm_sliced(a, b, mu, slices=4, work=None, kind="bf16"):
    """mu += conj(a) @ b.T from complex128 a (m x k) and b (n x k) with
    ``slices`` exact integer slices per component: on the BF16 matrix cores
    (kind 'bf16': slices of 8 bits, 1..4) or the INT8 matrix cores (kind
    'int8': slices of 7 bits, 1..6)."""
    import torch
    m, k = a.shape
    n = b.shape[0]
    if b.shape[1] != k or tuple(mu.shape) != (m, n):
        raise ValueError("corr_gemm_sliced shape mismatch")
    if a.dtype != torch.complex128 or b.dtype != torch.complex128 or mu.dtype != torch.complex128:
        raise ValueError("corr_gemm_sliced takes complex128 operands and output")
    if a.stride(1) != 1 or b.stride(1) != 1 or mu.stride(1) != 1:
        raise ValueError("operands must be contiguous along k")
    if kind not in ("bf16", "int8"):
        raise ValueError("kind must be 'bf16' or 'int8'")
    lib = load_library()
    work_fn, gemm_fn, name = ((lib.kpm_corr_gemm_sliced_work, lib.kpm_corr_gemm_sliced,
                               "kpm_corr_gemm_sliced") if kind == "bf16" else
                              (lib.kpm_corr_gemm_islice_work, lib.kpm_corr_gemm_islice,
                               "kpm_corr_gemm_islice"))
    need = ctypes.c_int64()
    check(work_fn(m, n, k, int(slices), ctypes.byref(need)), name + "_work")
    if work is None or work.numel() * work.element_size() < need.value:
        work = torch.empty(need.value, dtype=torch.uint8, device=mu.device)
    check(gemm_fn(a.data_ptr(), a.stride(0), b.data_ptr(), b.stride(0),
                  mu.data_ptr(), mu.stride(0), m, n, k, int(slices),
                  work.data_ptr(), work.numel() * work.element_size(),
                  stream_handle()), name)
    return work


class IsliceOperand:
    """An n_rows x k complex128 operand cut once into ``slices`` int8 slices
    per component (kpm_islice_prep_rows), for many products
    mu += conj(a) @ b.T on the INT8 matrix cores (``gemm``)."""

    def __init__(self, n_rows, k, slices, device=None):
        import torch
        self.n_rows, self.k, self.slices = int(n_rows), int(k), int(slices)
        need = ctypes.c_int64()
        check(load_library().kpm_islice_prep_bytes(self.n_rows, self.k, self.slices,
                                                   ctypes.byref(need)), "kpm_islice_prep_bytes")
        self.data = torch.empty(need.value, dtype=torch.uint8,
                                device=device if device is not None else torch_device())
        self.work = None

    def fill_rows(self, rows, row0):
        """Slice ``rows`` (r x k complex128, unit stride along k) into rows
        row0 .. row0 + r - 1."""
        import torch
        if rows.dtype != torch.complex128 or rows.stride(1) != 1 or rows.shape[1] != self.k:
            raise ValueError("rows must be r x k complex128 with unit stride along k")
        check(load_library().kpm_islice_prep_rows(
            rows.data_ptr(), rows.stride(0), int(row0), rows.shape[0], self.n_rows, self.k,
            self.slices, self.data.data_ptr(), self.data.numel(), stream_handle()),
            "kpm_islice_prep_rows")

    def gemm(self, a, mu):
        """mu (m x n_rows) += conj(a) @ b.T, a: m x k complex128."""
        import torch
        m, k = a.shape
        if k != self.k or tuple(mu.shape) != (m, self.n_rows):
            raise ValueError("IsliceOperand.gemm shape mismatch")
        if a.dtype != torch.complex128 or mu.dtype != torch.complex128:
            raise ValueError("IsliceOperand.gemm takes complex128 operands and output")
        if a.stride(1) != 1 or mu.stride(1) != 1:
            raise ValueError("operands must be contiguous along k")
        lib = load_library()
        need = ctypes.c_int64()
        check(lib.kpm_corr_gemm_islice_prep_work(m, self.n_rows, k, self.slices,
                                                 ctypes.byref(need)),
              "kpm_corr_gemm_islice_prep_work")
        if self.work is None or self.work.numel() < need.value:
            self.work = torch.empty(need.value, dtype=torch.uint8, device=mu.device)
        check(lib.kpm_corr_gemm_islice_prep(
            a.data_ptr(), a.stride(0), m, self.data.data_ptr(), self.data.numel(), self.n_rows,
            k, self.slices, mu.data_ptr(), mu.stride(0), self.work.data_ptr(),
            self.work.numel(), stream_handle()), "kpm_corr_gemm_islice_prep")


def moment_shift(nu, alpha, beta, out, in_scale=None, out_scale=None):
    """Moments of H + mu from those of H (see kpm_moment_shift): nu, out:
    (Z, P, M) float32 with unit stride along M; beta: (P,) float64 device
    tensor with |alpha| + |beta| <= 1 (checked by the caller); scales: (M,)
    float64 device tensors or None."""
    import torch
    if nu.dtype != torch.float32 or out.dtype != torch.float32 or nu.dim() != 3:
        raise ValueError("nu and out must be (Z, P, M) float32")
    z, p, m = nu.shape
    if tuple(out.shape) != (z, p, m) or nu.stride(2) != 1 or out.stride(2) != 1:
        raise ValueError("moment_shift layout")
    if beta.dtype != torch.float64 or beta.numel() != p or not beta.is_contiguous():
        raise ValueError("beta must be a contiguous (P,) float64 tensor")
    for sc in (in_scale, out_scale):
        if sc is not None and (sc.dtype != torch.float64 or sc.numel() != m):
            raise ValueError("scales must be (M,) float64")
    check(load_library().kpm_moment_shift(
        nu.data_ptr(), nu.stride(0), nu.stride(1), z, p, m,
        None if in_scale is None else in_scale.data_ptr(),
        None if out_scale is None else out_scale.data_ptr(), float(alpha), beta.data_ptr(),
        out.data_ptr(), out.stride(0), out.stride(1), stream_handle()), "kpm_moment_shift")


def cubic4_nnz(L):
    """Entries of the CSR of one cubic 4-orbital system of side L."""
    out = ctypes.c_int64()
    check(load_library().kpm_cubic4_nnz(int(L), ctypes.byref(out)), "kpm_cubic4_nnz")
    return out.value


def gen_cubic4(L, seed, sigma, t, row0, nz0, indptr, indices, values, radius_bits):
    """Write one cubic 4-orbital system into its diagonal block of a
    block-diagonal CSR on the device (see kpm_gen_cubic4)."""
    import torch
    if indptr.dtype != torch.int32 or indices.dtype != torch.int32:
        raise ValueError("indptr and indices must be int32")
    if radius_bits.dtype != torch.int32 or radius_bits.numel() != 1:
        raise ValueError("radius_bits must be one int32 element")
    _, vt = complex_dtype(values.dtype)
    n = 4 * L ** 3
    if row0 + n + 1 > indptr.numel() or nz0 + cubic4_nnz(L) > indices.numel() \
            or nz0 + cubic4_nnz(L) > values.numel():
        raise ValueError("cubic4 block outside the CSR buffers")
    check(load_library().kpm_gen_cubic4(int(L), int(seed) & (2 ** 64 - 1), float(sigma),
                                        float(t), int(row0), int(nz0), indptr.data_ptr(),
                                        indices.data_ptr(), values.data_ptr(), vt,
                                        radius_bits.data_ptr(), stream_handle()),
          "kpm_gen_cubic4")


def gen_cubic4_onsite(L, seed, sigma, t, row0, onsite, radius_bits):
    """The on-site blocks of one cubic 4-orbital system, rows row0 ..
    row0 + 4 L^3 of the (N, 4) complex device array ``onsite`` (see
    kpm_gen_cubic4_onsite)."""
    import torch
    if radius_bits.dtype != torch.int32 or radius_bits.numel() != 1:
        raise ValueError("radius_bits must be one int32 element")
    if onsite.dim() != 2 or onsite.shape[1] != 4 or not onsite.is_contiguous():
        raise ValueError("onsite must be a contiguous (N, 4) complex tensor")
    _, vt = complex_dtype(onsite.dtype)
    if row0 + 4 * L ** 3 > onsite.shape[0]:
        raise ValueError("cubic4 block outside the on-site buffer")
    check(load_library().kpm_gen_cubic4_onsite(int(L), int(seed) & (2 ** 64 - 1), float(sigma),
                                               float(t), int(row0), onsite.data_ptr(), vt,
                                               radius_bits.data_ptr(), stream_handle()),
          "kpm_gen_cubic4_onsite")


class MarchPlan:
    """Work list of the lattice-stencil steps for cubic 4-orbital systems of
    the given sides (kpm_march_create)."""

    # (waves per workgroup, depth, tile_y) of each shape: depth 1 or 2 planes
    # of prefetch, 3 or 4 the same with the plane exchanged through LDS;
    # tile_y rows along y of the bands whose sites (z-major) workgroups take
    SHAPES = ((4, 1, 1), (8, 1, 1), (4, 2, 1), (8, 2, 1), (4, 3, 1), (8, 3, 1), (4, 4, 1),
              (8, 4, 1), (4, 3, 2), (8, 3, 2), (4, 3, 4), (8, 3, 4), (4, 1, 2), (4, 1, 4),
              (8, 3, 8), (4, 4, 2))

    @staticmethod
    def shape_ok(width, value_type, shape):
        """Whether the library has a stencil kernel for this width (at least
        4 lanes per row; every shape takes every such width)."""
        lpr = width // 2 if value_type == KPM_C64 else width
        return 0 <= int(shape) < len(MarchPlan.SHAPES) and 4 <= lpr <= 64

    def __init__(self, sides, width, value_type, shape=0):
        lib = load_library()
        sd = np.ascontiguousarray(sides, dtype=np.int32)
        waves, depth, tile_y = self.SHAPES[int(shape)]
        handle = ctypes.c_void_p()
        check(lib.kpm_march_create(len(sd), sd.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                   int(width), int(value_type), waves, depth, tile_y,
                                   ctypes.byref(handle)), "kpm_march_create")
        self.handle = handle
        self.width, self.value_type = int(width), int(value_type)
        self.shape, self.waves, self.depth, self.tile_y = int(shape), waves, depth, tile_y

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            try:
                _lib.kpm_march_destroy(h)
            except Exception:
                pass
            self.handle = None

    def steps(self, plan, onsite, t, ring, slot_stride, slots, n_steps, mode, dots=None,
              out_base=0):
        out_cap = 0 if dots is None else int(dots.shape[0])
        check(load_library().kpm_cheb_steps_march(
            plan.handle, self.handle, onsite.data_ptr(), float(t), ring.data_ptr(), slot_stride,
            slots, None if dots is None else dots.data_ptr(), out_base, out_cap, n_steps, mode,
            stream_handle()), "kpm_cheb_steps_march")


def _f32_view(t, name, dims=3):
    import torch
    if t.dtype != torch.float32 or t.dim() != dims:
        raise ValueError("%s must be a %d-d float32 tensor" % (name, dims))
    return t


def mlp_gemm(a, b, c, bias=None, flags=MLP_PLAIN, c2=None, aux=None):
    """c[z] = a[z] @ b[z].T (+ bias[z]) with the epilogue/modifiers of
    ``flags`` on the matrix cores (BF16 operands, FP32 accumulation).

    a: (Z, M, K), b: (Z, N, K), c / c2 / aux: (Z, M, N) float32 views with
    any strides (stride 0 along Z broadcasts one operand to every batch);
    c, c2 and aux need unit stride along N; bias: (Z, N) with unit stride."""
    _f32_view(a, "a"), _f32_view(b, "b"), _f32_view(c, "c")
    z, m, k = a.shape
    n = b.shape[1]
    if b.shape[0] != z or b.shape[2] != k or tuple(c.shape) != (z, m, n):
        raise ValueError("mlp_gemm shape mismatch")
    if c.stride(2) != 1:
        raise ValueError("c needs unit stride along N")
    d = MlpGemmDesc()
    d.a, d.a_bs, d.a_rs, d.a_cs = a.data_ptr(), *a.stride()
    d.b, d.b_bs, d.b_rs, d.b_cs = b.data_ptr(), *b.stride()
    d.c, d.c_bs, d.c_rs = c.data_ptr(), c.stride(0), c.stride(1)
    if c2 is not None:
        if c2.shape != c.shape or c2.stride() != c.stride():
            raise ValueError("c2 must have the shape and strides of c")
        d.c2 = c2.data_ptr()
    if bias is not None:
        if tuple(bias.shape) != (z, n) or bias.stride(1) != 1:
            raise ValueError("bias must be (Z, N) with unit stride")
        d.bias, d.bias_bs = bias.data_ptr(), bias.stride(0)
    if aux is not None:
        if aux.shape != c.shape or aux.stride(2) != 1:
            raise ValueError("aux must match c with unit stride along N")
        d.aux, d.aux_bs, d.aux_rs = aux.data_ptr(), aux.stride(0), aux.stride(1)
    d.m, d.n, d.k, d.batch, d.flags = m, n, k, z, flags
    check(load_library().kpm_mlp_gemm(ctypes.byref(d), stream_handle()), "kpm_mlp_gemm")


def mlp_colsum(a, out):
    """out[z, n] = sum_m a[z, m, n] (a: (Z, M, N), unit stride along N)."""
    _f32_view(a, "a")
    z, m, n = a.shape
    if a.stride(2) != 1 or tuple(out.shape) != (z, n) or out.stride(1) != 1:
        raise ValueError("mlp_colsum layout")
    check(load_library().kpm_mlp_colsum(a.data_ptr(), a.stride(1), a.stride(0), m, n, z,
                                        out.data_ptr(), out.stride(0), stream_handle()),
          "kpm_mlp_colsum")


def mlp_mse(y, t, dy, scale, partial):
    """dy = scale (y - t) elementwise over contiguous tensors; partial (FP64,
    one per workgroup) receives sums of (y - t)^2."""
    for x in (y, t, dy):
        if not x.is_contiguous() or x.numel() != y.numel():
            raise ValueError("mlp_mse needs contiguous tensors of one size")
    check(load_library().kpm_mlp_mse(y.data_ptr(), t.data_ptr(), dy.data_ptr(), y.numel(),
                                     float(scale), partial.data_ptr(), partial.numel(),
                                     stream_handle()), "kpm_mlp_mse")


def mlp_ensemble(y, mean=None, dev=None):
    """Ensemble mean and largest member L1 deviation per row of y (Z, P, K)
    (unit stride along K); mean: (P, K) or None, dev: (P,) or None."""
    _f32_view(y, "y")
    z, rows, width = y.shape
    if y.stride(2) != 1:
        raise ValueError("y needs unit stride along its last axis")
    if mean is not None and (tuple(mean.shape) != (rows, width) or mean.stride(1) != 1):
        raise ValueError("mean must be (P, K) with unit stride")
    if dev is not None and (dev.numel() != rows or not dev.is_contiguous()):
        raise ValueError("dev must be a contiguous (P,) tensor")
    check(load_library().kpm_mlp_ensemble(
        y.data_ptr(), y.stride(0), y.stride(1), z, rows, width,
        None if mean is None else mean.data_ptr(), 0 if mean is None else mean.stride(0),
        None if dev is None else dev.data_ptr(), stream_handle()), "kpm_mlp_ensemble")


def mlp_adam(p, g, m, v, lr, b1, b2, eps, t, sched=None, ema=None, ema_decay=0.0):
    """One bias-corrected Adam update of the flat parameter buffer p; with
    ``sched`` (a float32 device tensor holding lr, 1/(1-b1^t), 1/(1-b2^t),
    ema decay) the kernel reads those when it runs and ignores lr, t and
    ema_decay.  With ``ema`` the new parameters are folded into it."""
    for x in (p, g, m, v) + (() if ema is None else (ema,)):
        if not x.is_contiguous() or x.numel() != p.numel():
            raise ValueError("mlp_adam needs contiguous buffers of one size")
    if sched is not None and (sched.numel() < 4 or not sched.is_contiguous()):
        raise ValueError("sched needs four contiguous floats")
    check(load_library().kpm_mlp_adam(p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(),
                                      p.numel(), float(lr), float(b1), float(b2), float(eps),
                                      int(t), None if sched is None else sched.data_ptr(),
                                      None if ema is None else ema.data_ptr(),
                                      float(ema_decay), stream_handle()), "kpm_mlp_adam")
