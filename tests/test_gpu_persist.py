"""Save / load of an expansion (kwant_amd.persist): after a round trip
through the file, add_moments and add_vectors give bitwise the moments of
the uninterrupted run, for every moment path (doubling, matrix operator,
site-resolved Density, Current bond records, user callable), both
precisions and a reordered matrix."""
import numpy as np
import pytest

from fake_system import FakeSystem

pytestmark = pytest.mark.gpu


def _bounds(h):
    r = float(np.max(np.asarray(abs(h).sum(axis=1)).ravel()))
    return (-r, r)


SYST = FakeSystem([(30, 2), (20, 1), (10, 3)], seed=3)


def _operators():
    from kwant_amd.operator import Current, Density
    op = (SYST.h + SYST.h.T.conj()) * 0.3
    return {
        'dos': None,
        'matrix': op,
        'density': Density(SYST, 0.5, where=[3, 40, 40, 55]),
        'density_per_vector': Density(SYST, 0.5, where=[3, 40]),
        'current': Current(SYST, 1, where=list(SYST.graph)[::9]),
        'callable': (lambda bra, ket: np.vdot(bra, op @ ket)),
    }


def _make(name, precision='complex128', **kw):
    from kwant_amd import kpm
    ops = _operators()
    mean = name != 'density_per_vector'
    return kpm.SpectralDensity(SYST, operator=ops[name], num_moments=30, num_vectors=10,
                               bounds=_bounds(SYST.h), rng=5, precision=precision,
                               mean=mean, **kw), ops[name]


@pytest.mark.parametrize("name", ['dos', 'matrix', 'density', 'density_per_vector', 'current',
                                  'callable'])
def test_round_trip_resumes_bitwise(name, tmp_path):
    from kwant_amd import kpm
    a, _ = _make(name)
    a.add_moments(15)
    a.add_vectors(3)
    b, op = _make(name)
    path = tmp_path / ("%s.npz" % name)
    b.save(path)
    c = kpm.SpectralDensity.load(path, SYST, operator=op)
    assert np.array_equal(c._moments(), b._moments())
    c.add_moments(15)
    c.add_vectors(3)
    assert np.array_equal(c._moments(), a._moments())
    assert np.array_equal(c.densities, a.densities)


@pytest.mark.parametrize("precision,reorder", [('complex64', False), ('complex128', True)])
def test_round_trip_precision_and_order(precision, reorder, tmp_path):
    from kwant_amd import kpm
    a, _ = _make('dos', precision=precision, reorder=reorder)
    a.add_moments(11)
    b, _ = _make('dos', precision=precision, reorder=reorder)
    b.save(tmp_path / "x.npz")
    c = kpm.SpectralDensity.load(tmp_path / "x.npz", SYST)
    assert (c._dmat.order is not None) == reorder
    c.add_moments(11)
    assert np.array_equal(c._moments(), a._moments())


def test_load_checks_the_hamiltonian(tmp_path):
    import scipy.sparse as sp
    from kwant_amd import kpm
    b, _ = _make('dos')
    b.save(tmp_path / "x.npz")
    with pytest.raises(ValueError):
        kpm.SpectralDensity.load(tmp_path / "x.npz", sp.identity(7, format='csr'))


def test_ensemble_round_trip(tmp_path):
    """An ensemble's moments survive save/load bitwise; the loaded ensemble
    regenerates the same systems (running it again gives the same dots)."""
    import torch
    from kwant_amd.ensemble import CubicEnsemble
    ens = CubicEnsemble([6, 9, 7], seed=4, num_moments=32, num_vectors=4)
    ens.run(graph=False)
    ens.save(tmp_path / "ens.npz")
    back = CubicEnsemble.load(tmp_path / "ens.npz")
    assert np.array_equal(back.moments(), ens.moments())
    assert torch.equal(back.values, ens.values) and torch.equal(back.indices, ens.indices)
    saved = back.dots.clone()
    back.run(graph=False)
    assert torch.equal(back.dots, saved)
    fresh = CubicEnsemble([6, 9, 7], seed=4, num_moments=32, num_vectors=4)
    with pytest.raises(ValueError):
        fresh.save(tmp_path / "not_run.npz")


def test_correlator_round_trip(tmp_path):
    """A saved Correlator evaluates sigma(mu, T) bitwise as the original;
    it refuses to add moments or vectors without the Hamiltonian."""
    from kwant_amd import kpm
    h = SYST.h
    x = np.asarray([p[0] for p in np.random.default_rng(1).normal(size=(h.shape[0], 2))])
    vel = 1j * (h.multiply(x[None, :]) - h.multiply(x[:, None]))
    cor = kpm.Correlator(h, operator1=vel, operator2=vel, num_moments=40, num_vectors=2,
                         rng=3, bounds=_bounds(h))
    cor.save(tmp_path / "cor.npz")
    back = kpm.Correlator.load(tmp_path / "cor.npz")
    for mu, temp in ((0.0, 0.0), (0.3, 0.05)):
        assert back(mu, temp) == cor(mu, temp)
    assert np.array_equal(back.moments_matrix, cor.moments_matrix)
    with pytest.raises(ValueError):
        back.add_moments(10)
    with pytest.raises(ValueError):
        back.add_vectors(1)


def test_round_trip_keeps_the_kernel(tmp_path):
    """jackson and lorentz kernels come back by name; a custom kernel must be
    passed again."""
    from kwant_amd import kpm
    from oracle import brute
    h = brute.gaussian_unitary(16, 3)
    kw = dict(num_moments=40, num_vectors=3, bounds=(-12.0, 12.0), rng=1)
    sd = kpm.SpectralDensity(h, kernel=kpm.lorentz_kernel, **kw)
    p = tmp_path / "lorentz.npz"
    sd.save(p)
    back = kpm.SpectralDensity.load(p, h)
    assert back.kernel is kpm.lorentz_kernel
    assert np.array_equal(back.densities, sd.densities)

    def halved(m):
        return kpm.jackson_kernel(m) / 2
    sd = kpm.SpectralDensity(h, kernel=halved, **kw)
    q = tmp_path / "custom.npz"
    sd.save(q)
    with pytest.raises(ValueError, match="custom kernel"):
        kpm.SpectralDensity.load(q, h)
    back = kpm.SpectralDensity.load(q, h, kernel=halved)
    assert np.array_equal(back.densities, sd.densities)
