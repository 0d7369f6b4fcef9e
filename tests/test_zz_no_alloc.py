"""The allocation-free entry point kpm_cheb_steps_static against the plan
path kpm_cheb_steps: the same dot records and ring contents, bit for bit, on
two calls in a row (the second with a smaller system), across a resumed
expansion, and a clear refusal when the static workspace is too small."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import brute

pytestmark = pytest.mark.gpu


def _system(n, seed):
    """Sparse Hermitian matrix with ragged rows (a band plus random far
    entries) and its Gershgorin radius."""
    rng = np.random.RandomState(seed)
    rows = np.concatenate([np.arange(n - 1), rng.randint(0, n, 3 * n)])
    cols = np.concatenate([np.arange(1, n), rng.randint(0, n, 3 * n)])
    vals = rng.normal(size=len(rows)) + 1j * rng.normal(size=len(rows))
    h = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    h = (h + h.getH() + sp.diags(rng.normal(size=n))).tocsr()
    h.sum_duplicates()
    h.sort_indices()
    return h, float(np.abs(h).sum(axis=1).max())


def _compare(n, width, precision, steps, seed, variant):
    import torch
    from kwant_amd import _engine as eng
    h, r = _system(n, seed)
    ab = (1.05 * r, 0.1)
    mat = eng.DeviceCSR(h, precision)
    cdt, _ = eng.complex_dtype(precision)
    dev = eng.torch_device()
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    ph = torch.rand((n, width), generator=gen, device=dev, dtype=torch.float64)
    block = torch.polar(torch.ones_like(ph), 2 * np.pi * ph).to(cdt)
    # the plan path
    rec = eng.BlockRecursion(mat, ab, block, width, eng.DOTS_DOUBLING, variant=variant)
    want = rec.step(steps)
    # the static path, in two calls (a fresh start, then a resume)
    ring = torch.zeros((2, n, width), dtype=cdt, device=dev)
    ring[0] = block
    grid = torch.full((width,), -12345, dtype=torch.int32, device=dev)
    dots = torch.empty((steps, 1, width, 4), dtype=torch.float64, device=dev)
    half = steps // 2
    eng.cheb_steps_static(mat, ab, ring, 2, 0, half, eng.DOTS_DOUBLING, dots=dots,
                          grid_exp=grid, variant=variant)
    assert int(grid.min()) != -12345          # written from a_0
    eng.cheb_steps_static(mat, ab, ring, 2, half, steps - half, eng.DOTS_DOUBLING, dots=dots,
                          grid_exp=grid, variant=variant)
    got = dots[:, 0].cpu().numpy()
    assert np.array_equal(got, want), (n, width, precision)
    assert torch.equal(ring, rec.ring)
    # recursion only (no dots) agrees too
    ring2 = torch.zeros_like(ring)
    ring2[0] = block
    eng.cheb_steps_static(mat, ab, ring2, 2, 0, steps, eng.DOTS_NONE, variant=variant)
    assert torch.equal(ring2, rec.ring)


def test_static_steps_match_the_plan_path_on_two_calls():
    _compare(5000, 16, "complex128", 30, 1, 0)
    _compare(1200, 8, "complex64", 12, 2, 5)     # smaller, another shape


def test_static_steps_refuse_an_oversized_workspace():
    import torch
    from kwant_amd import _engine as eng
    n, width = 400000, 64
    h = sp.diags(np.linspace(-1, 1, n)).tocsr().astype(complex)
    mat = eng.DeviceCSR(h, "complex64")
    ring = torch.zeros((2, n, width), dtype=torch.complex64, device=eng.torch_device())
    dots = torch.empty((4, 1, width, 4), dtype=torch.float64, device=ring.device)
    grid = torch.zeros(width, dtype=torch.int32, device=ring.device)
    # variant 3: 32 residency rounds -> ~50,000 workgroups of partials
    with pytest.raises(RuntimeError, match="exceeds the"):
        eng.cheb_steps_static(mat, (1.1, 0.0), ring, 2, 0, 4, eng.DOTS_DOUBLING, dots=dots,
                              grid_exp=grid, variant=3)
