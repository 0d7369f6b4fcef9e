"""Size-independent checks at the full benchmark sizes, where the oracle is
too slow to run: the 2M-site graphene block with 64 vectors and the
256-system cubic ensemble.

* mu_0 = 1 for every normalised vector and |mu_n| <= 1 (the rescaled
  spectrum lies in (-1, 1), so |<a_0|T_n(H~)|a_0>| <= <a_0|a_0>);
* complex64 moments agree with complex128 moments of the same vectors to
  1e-5 (the parity tolerance of the complex64 path);
* resuming an expansion (add_moments) is bitwise the one-shot expansion;
* the ensemble's graph replays equal its eager launches bitwise, and every
  system's moments obey the same bounds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graphene_block(width, precision, moments_first, moments_more=0):
    import torch
    from kwant_amd import _engine as eng
    from kwant_amd.lattices import honeycomb, honeycomb_bounds
    h, _ = honeycomb(1000, disorder=0.5, seed=2)
    lmin, lmax = honeycomb_bounds(1.0, 0.5)
    a = (lmax - lmin) / (2 - 0.05)
    cdt, _ = eng.complex_dtype(precision)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    ph = torch.rand((h.shape[0], width), generator=gen, device="cuda", dtype=torch.float64)
    block = torch.polar(torch.ones_like(ph), 2 * np.pi * ph).to(cdt)
    del ph
    mat = eng.DeviceCSR(h, precision)
    rec = eng.BlockRecursion(mat, (a, 0.0), block, width, eng.DOTS_DOUBLING)
    d1 = rec.step(moments_first // 2)
    d2 = rec.step(moments_more // 2) if moments_more else None
    return h.shape[0], d1, d2


def _moments(dots):
    """Doubling moments (steps, W, 4) -> (W, 2 * steps) normalised by mu_0."""
    mu0 = dots[0, :, 0]
    mu1 = dots[0, :, 2]
    m = np.empty((dots.shape[1], 2 * dots.shape[0]))
    m[:, 0], m[:, 1] = mu0, mu1
    m[:, 2::2] = (2 * dots[1:, :, 0] - mu0[None, :]).T
    m[:, 3::2] = (2 * dots[1:, :, 2] - mu1[None, :]).T
    return m / mu0[:, None]


def test_graphene_full_size_invariants_and_precision():
    n, d128, _ = _graphene_block(64, "complex128", 256)
    _, d64, _ = _graphene_block(64, "complex64", 256)
    m128, m64 = _moments(d128), _moments(d64)
    assert np.allclose(d128[0, :, 0], n, rtol=1e-12)        # <a_0|a_0> = N (unit phases)
    assert np.all(np.abs(m128) <= 1 + 1e-12)
    assert np.max(np.abs(m64 - m128)) < 1e-5


def test_graphene_full_size_resume_bitwise():
    _, once, _ = _graphene_block(64, "complex64", 128)
    _, first, more = _graphene_block(64, "complex64", 64, 64)
    assert np.array_equal(once, np.concatenate([first, more]))


def test_ensemble_full_size_replay_and_bounds():
    import torch
    from kwant_amd.ensemble import CubicEnsemble, ensemble_sizes
    ens = CubicEnsemble(ensemble_sizes(256, 20, 60, seed=0), seed=0, num_moments=64,
                        num_vectors=16)
    ens.run(graph=False)
    eager = ens.dots.cpu().numpy().copy()
    ens.dots.zero_()
    ens.run(graph=True)
    assert np.array_equal(ens.dots.cpu().numpy(), eager)
    m = ens.moments(damped=False)
    assert m.shape == (256, 64)
    assert np.allclose(m[:, 0], 1.0, rtol=0, atol=1e-12)
    assert np.all(np.abs(m) <= 1 + 1e-5)
    del ens
    torch.cuda.empty_cache()
