"""Vectorised Hamiltonian assembly (kwant_amd.assembly.LatticeSystem)
against the oracle's restatement of kwant's one-site-at-a-time assembly
(kwant/_system.pyx:94-153), and the oracle against closed forms."""
import numpy as np
import pytest
import scipy.sparse

from kwant_amd.assembly import LatticeSystem, honeycomb_system
from oracle import assembly_oracle as ao


def _same(a, b):
    a, b = scipy.sparse.csr_matrix(a), scipy.sparse.csr_matrix(b)
    assert a.shape == b.shape
    assert np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
    assert np.max(np.abs(a.data - b.data), initial=0) <= 1e-15 * max(1, np.max(np.abs(b.data)))


def _chain(n, t):
    syst = LatticeSystem([(1.0,)], [(0.0,)], [(0, n)])
    syst.add_hoppings((1,), 0, 0, -t)
    return syst


def test_oracle_chain_closed_form():
    n, t = 30, 0.8
    h = ao.csr(_chain(n, t)).toarray()
    want = -t * (np.eye(n, k=1) + np.eye(n, k=-1))
    assert np.array_equal(h, want)
    ev = np.sort(np.linalg.eigvalsh(h))
    k = np.arange(1, n + 1)
    assert np.allclose(ev, np.sort(-2 * t * np.cos(k * np.pi / (n + 1))), atol=1e-12)


def test_oracle_square_spectrum():
    L, t = 9, 1.3
    syst = LatticeSystem([(1, 0), (0, 1)], [(0, 0)], [(0, L), (0, L)])
    syst.add_hoppings((1, 0), 0, 0, -t)
    syst.add_hoppings((0, 1), 0, 0, -t)
    ev = np.sort(np.linalg.eigvalsh(ao.csr(syst).toarray()))
    c = np.cos(np.arange(1, L + 1) * np.pi / (L + 1))
    want = np.sort((-2 * t * (c[:, None] + c[None, :])).ravel())
    assert np.allclose(ev, want, atol=1e-12)


def test_oracle_graphene_entry_count():
    cells = 12
    syst = honeycomb_system(cells, onsite=0.1)
    h = ao.csr(syst)
    n = 2 * cells * cells
    # bond A(i,j)-B(i,j): c^2; A(i,j)-B(i+1,j-1): (c-1)^2; A(i,j)-B(i,j-1): c(c-1)
    hops = cells ** 2 + (cells - 1) ** 2 + cells * (cells - 1)
    assert h.nnz == n + 2 * hops


def _disordered_honeycomb(cells):
    syst = honeycomb_system(cells)

    def onsite(pos, w, seed):
        # a deterministic position hash: the same value whichever way it is evaluated
        x = np.sin(pos[:, 0] * 12.9898 + pos[:, 1] * 78.233 + seed) * 43758.5453
        return w * (x - np.floor(x) - 0.5)

    def hop(to, fr, t, phi):
        return -t * np.exp(1j * phi * (to[:, 0] - fr[:, 0]) * (to[:, 1] + fr[:, 1]))

    syst.set_onsite(onsite)
    syst._hops = [(d, s, hop) for d, s, _ in syst._hops]
    return syst


def test_vectorised_matches_oracle():
    syst = _disordered_honeycomb(10)
    params = dict(w=0.7, seed=3.0, t=1.1, phi=0.05)
    _same(syst.hamiltonian_submatrix(params=params), ao.csr(syst, params))
    h = syst.hamiltonian_submatrix(params=params)
    assert abs(h - h.conj().T).max() < 1e-15


def test_two_orbitals_shape_and_zero_entries():
    sz = np.diag([1.0, -1.0])
    sx = np.array([[0, 1], [1, 0]], complex)

    def hop(to, fr, t):
        return -t * np.broadcast_to(np.eye(2) + 0.3j * sx, (len(to), 2, 2))

    syst = LatticeSystem([(1, 0), (0, 1)], [(0, 0)], [(-6, 7), (-6, 7)],
                         shape=lambda p: np.hypot(p[:, 0], p[:, 1]) < 5.5, norbs=2)
    syst.set_onsite(lambda pos, m: m * np.broadcast_to(sz, (len(pos), 2, 2)))
    syst.add_hoppings((1, 0), 0, 0, hop)
    syst.add_hoppings((0, 1), 0, 0, np.diag([0.5, 0.0]))      # a zero entry in every block
    params = dict(m=0.2, t=1.0)
    _same(syst.hamiltonian_submatrix(params=params), ao.csr(syst, params))
    assert syst.site_ranges[-1][2] == 2 * len(syst.pos_array)


def test_errors():
    syst = _chain(5, 1.0)
    syst.add_hoppings((1,), 0, 0, 2.0)
    with pytest.raises(ValueError):
        syst.hamiltonian_submatrix()                          # hopping defined twice
    bad = LatticeSystem([(1.0,)], [(0.0,)], [(0, 4)])
    bad.set_onsite(lambda pos, w: w * np.ones(len(pos)))
    with pytest.raises(TypeError):
        bad.hamiltonian_submatrix()                           # missing parameter


def test_operators_accept_lattice_systems():
    from kwant_amd.operator import Current, Density
    from oracle.operator_oracle import current, density
    syst = _disordered_honeycomb(6)
    params = dict(w=0.7, seed=3.0, t=1.1, phi=0.05)
    n = len(syst.pos_array)
    rng = np.random.default_rng(2)
    bra, ket = rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))
    rho = Density(syst, where=[0, 5, 9])
    assert np.allclose(rho(bra, ket), density(syst.site_ranges, 1, [0, 5, 9])(bra, ket))
    hops = list(syst.graph)[:20]
    j = Current(syst, where=hops).bind(params=params)
    ref = current(syst.site_ranges, 1, hops,
                  [syst.hamiltonian(a, b, params=params) for a, b in hops])
    assert np.allclose(j(bra, ket), ref(bra, ket), rtol=1e-13, atol=1e-13)


def test_velocity_from_lattice_system():
    """conductivity's velocity operator (kwant/kpm.py:1042-1085) reads
    hamiltonian_submatrix(return_norb=True) and pos(i) from a system."""
    from kwant_amd.correlator import _velocity
    syst = _disordered_honeycomb(5)
    params = dict(w=0.7, seed=3.0, t=1.1, phi=0.05)
    h = syst.hamiltonian_submatrix(params=params)
    for direction in ('x', 'y'):
        from_syst = _velocity(syst, params, direction, None)
        from_matrix = _velocity(h, params, direction, syst.pos_array)
        assert abs(from_syst - from_matrix).max() == 0
    assert np.array_equal(syst.pos(3), syst.pos_array[3])
