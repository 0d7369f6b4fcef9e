"""The batched SpMM (kpm_spmm): slots taken 8 per workgroup give bitwise the
single-slot kernel's products and scipy's to rounding, on ragged rows, padded
slot strides and ragged slot counts; a row block of the device matrix
(DeviceCSR.row_block) writes bitwise those rows of the full product."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def ragged(n, m, seed, max_row=13):
    rng = np.random.RandomState(seed)
    rows, cols = [], []
    for r in range(n):
        k = rng.randint(0, max_row + 1)
        rows += [r] * k
        cols += list(rng.randint(0, m, k))
    vals = rng.standard_normal(len(rows)) + 1j * rng.standard_normal(len(rows))
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, m))


@pytest.mark.parametrize("prec,width", [("complex64", 1), ("complex64", 8), ("complex64", 16),
                                        ("complex128", 1), ("complex128", 8)])
def test_spmm_slots_bitwise_single_slot(prec, width):
    import torch
    from kwant_amd import _engine as eng
    n = 700
    h = ragged(n, n, 3)
    dm = eng.DeviceCSR(h, prec)
    cdt, _ = eng.complex_dtype(prec)
    pad = 24
    for n_slots in (3, 8, 11, 17):
        gen = torch.Generator(device="cuda")
        gen.manual_seed(n_slots)
        xs = torch.randn((n_slots, n * width + pad), generator=gen, device="cuda",
                         dtype=torch.complex128).to(cdt)
        ys = torch.zeros((n_slots, n * width + pad), dtype=cdt, device="cuda")
        dm.spmm(xs, ys, width, n_slots=n_slots, x_stride=xs.stride(0), y_stride=ys.stride(0))
        one = torch.zeros_like(ys)
        for s in range(n_slots):
            dm.spmm(xs[s], one[s], width, n_slots=1)
        assert torch.equal(ys, one), n_slots
        x = xs[:, :n * width].cpu().numpy().astype(np.complex128).reshape(n_slots, n, width)
        want = np.stack([h @ x[s] for s in range(n_slots)])
        got = ys[:, :n * width].cpu().numpy().reshape(n_slots, n, width)
        tol = 1e-12 if prec == "complex128" else 1e-5
        assert np.max(np.abs(got - want)) < tol * np.max(np.abs(want))
        assert torch.all(ys[:, n * width:] == 0)   # nothing written past the rows


def test_row_block_writes_those_rows():
    import torch
    from kwant_amd import _engine as eng
    n, width, n_slots = 500, 8, 9
    h = ragged(n, n, 5)
    dm = eng.DeviceCSR(h, "complex128")
    xs = torch.randn((n_slots, n, width), dtype=torch.complex128, device="cuda")
    full = torch.zeros_like(xs)
    dm.spmm(xs, full, width, n_slots=n_slots, x_stride=n * width, y_stride=n * width)
    for r0, r1 in ((0, 123), (123, 124), (124, 500)):
        sub = dm.row_block(r0, r1)
        assert sub.shape == (r1 - r0, n)
        out = torch.zeros((n_slots, (r1 - r0) * width + 5), dtype=torch.complex128, device="cuda")
        sub.spmm(xs, out, width, n_slots=n_slots, x_stride=n * width, y_stride=out.stride(0))
        got = out[:, :(r1 - r0) * width].reshape(n_slots, r1 - r0, width)
        assert torch.equal(got, full[:, r0:r1])
    with pytest.raises(ValueError):
        dm.row_block(10, 10)
