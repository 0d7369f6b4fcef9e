"""The 4 x 4 block-CSR recursion (kpm_bsr_steps) against the CSR path and
the oracle: the four-orbital workloads (the cubic ensemble's systems and the
four-band Dirac model) give moments bitwise equal to the column-sorted CSR
path's for every launch shape it runs, and within 1e-11 of the oracle."""
import numpy as np
import pytest

from oracle import kpm_oracle as ko

pytestmark = pytest.mark.gpu


def _systems():
    from kwant_amd.lattices import cubic_4orb, dirac_4band
    h1 = cubic_4orb(5, seed=3)
    h1 = h1[0] if isinstance(h1, tuple) else h1
    h2, _ = dirac_4band(6, flux=1 / 36, disorder=1.0, seed=2)
    return [("cubic4", h1), ("dirac4", h2)]


def _bounds(h):
    r = float(np.abs(h).sum(axis=1).max())
    return (-r, r)


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_bsr_moments_equal_csr_bitwise(precision):
    import torch
    from kwant_amd import _engine as eng
    from kwant_amd import autotune
    for name, h in _systems():
        h = h.tocsr()
        h.sort_indices()
        r = _bounds(h)[1]
        csr = eng.DeviceCSR(h, precision)
        bsr = eng.DeviceBSR(h, precision)
        cdt, vt = eng.complex_dtype(precision)
        for width in (1, 4, 16):
            gen = torch.Generator(device="cuda")
            gen.manual_seed(width)
            ph = torch.rand((h.shape[0], width), generator=gen, device="cuda",
                            dtype=torch.float64)
            block = torch.polar(torch.ones_like(ph), 2 * np.pi * ph).to(cdt)
            base = eng.BlockRecursion(csr, (1.05 * r, 0.05), block, width,
                                      eng.DOTS_DOUBLING).step(30)
            for v in autotune.candidates(bsr, width):
                got = eng.BlockRecursion(bsr, (1.05 * r, 0.05), block, width,
                                         eng.DOTS_DOUBLING, variant=v).step(30)
                assert np.array_equal(got, base), (name, precision, width, v)


def test_bsr_spectral_density_matches_oracle():
    from kwant_amd import kpm
    for name, h in _systems():
        kw = dict(num_moments=90, num_vectors=5, bounds=_bounds(h), rng=4)
        gpu = kpm.SpectralDensity(h, matrix_format='bsr', **kw)
        csr = kpm.SpectralDensity(h, **kw)
        ora = ko.SpectralDensity(h, **kw)
        g, o = np.array(gpu._moments_list), np.array(ora._moments_list)
        assert np.max(np.abs(g - o)) < 1e-11 * h.shape[0], name
        assert np.array_equal(g, np.array(csr._moments_list)), name
        gpu.add_moments(20)
        ora.add_moments(20)
        assert np.max(np.abs(np.array(gpu._moments_list) - np.array(ora._moments_list))) \
            < 1e-11 * h.shape[0]


def test_bsr_rejects_what_it_cannot_run():
    import scipy.sparse as sp
    from kwant_amd import kpm
    h = sp.random(30, 30, density=0.2, random_state=1) + 0j
    h = h + h.getH()
    with pytest.raises(ValueError, match="multiple of 4"):
        kpm.SpectralDensity(h, matrix_format='bsr', bounds=(-9, 9), num_moments=10)
    with pytest.raises(ValueError, match="operator=None"):
        kpm.SpectralDensity(np.eye(8) + 0j, matrix_format='bsr', operator=np.eye(8),
                            bounds=(-2, 2))
