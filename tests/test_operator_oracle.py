"""The Density oracle pinned by closed forms and brute force, and the host
side of kwant_amd.operator (Density, DensityLayout) checked against it.

CPU only: the GPU evaluation of the same densities is in test_gpu_density.py.
"""
import numpy as np
import pytest
import scipy.sparse

from fake_system import FakeSystem
from kwant_amd.operator import Density, density_layout
from oracle.operator_oracle import density as oracle_density


def _vec(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def _herm(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return a + a.conj().T


# -- oracle pins ---------------------------------------------------------------

def test_oracle_identity_closed_forms():
    ranges = [(0, 1, 0), (40, 0, 40)]
    bra, ket = _vec(40, 1), _vec(40, 2)
    per_site = oracle_density(ranges, 1, range(40))(bra, ket)
    assert np.allclose(per_site, np.conj(bra) * ket, rtol=0, atol=1e-14)
    total = oracle_density(ranges, 1, range(40), sum=True)(bra, ket)
    assert abs(total - np.vdot(bra, ket)) < 1e-12
    # real for bra is ket, complex otherwise
    assert np.isrealobj(oracle_density(ranges, 1, range(40))(bra, bra))
    assert np.iscomplexobj(oracle_density(ranges, 1, range(40))(bra, bra.copy()))


def test_oracle_matrix_onsite_brute_force():
    ranges = [(0, 3, 0), (10, 0, 30)]
    m = _herm(3, 3)
    where = [4, 0, 9, 4]
    bra, ket = _vec(30, 4), _vec(30, 5)
    got = oracle_density(ranges, m, where)(bra, ket)
    for w, site in enumerate(where):
        q = np.zeros((30, 30), dtype=complex)          # dense Q_w
        q[3 * site:3 * site + 3, 3 * site:3 * site + 3] = m
        assert abs(got[w] - np.vdot(bra, q @ ket)) < 1e-12


def test_oracle_scalar_onsite_on_smaller_sites():
    # sites of 1 and 2 orbitals next to 3-orbital sites: the scalar becomes
    # 2 * identity(3), read with the site's norbs as the row stride, so a
    # 2-orbital site sees the flattened entries [2, 0, 0, 0]
    ranges = [(0, 3, 0), (1, 2, 3), (2, 1, 5), (3, 0, 6)]
    bra = np.arange(1, 7) + 0j
    ket = np.arange(1, 7) * 1j
    got = oracle_density(ranges, 2.0, [0, 1, 2])(bra, ket)
    assert np.allclose(got, [2 * (1 * 1j + 2 * 2j + 3 * 3j), 2 * 4 * 4j, 2 * 6 * 6j])


# -- the host Density against the oracle -------------------------------------

RUNS = [(30, 2), (20, 1), (10, 3)]     # 60 sites, 100 orbitals


@pytest.mark.parametrize("sum_", [False, True])
def test_density_call_matches_oracle(sum_):
    syst = FakeSystem(RUNS)
    n = syst.site_ranges[-1][2]
    bra, ket = _vec(n, 6), _vec(n, 7)
    where = [5, 31, 31, 55, 0]
    for onsite in (1, 0.5):
        op = Density(syst, onsite, where=where, sum=sum_)
        ref = oracle_density(syst.site_ranges, onsite, where, sum=sum_)
        assert np.allclose(op(bra, ket), ref(bra, ket), rtol=1e-13, atol=1e-13)
        assert np.allclose(op(bra), ref(bra), rtol=1e-13, atol=1e-13)
        assert np.isrealobj(op(bra)) and np.iscomplexobj(op(bra, bra.copy()))


def test_density_callable_onsite_with_params():
    syst = FakeSystem([(20, 2)])
    mats = [_herm(2, s) for s in range(20)]

    def onsite(site, scale):
        return scale * mats[site]

    op = Density(syst, onsite, where=lambda s: s % 3 == 0)
    bound = op.bind(params=dict(scale=0.25))
    where = list(range(0, 20, 3))
    ref = oracle_density(syst.site_ranges, [0.25 * mats[s] for s in where], where)
    bra, ket = _vec(40, 8), _vec(40, 9)
    assert np.allclose(bound(bra, ket), ref(bra, ket), rtol=1e-13, atol=1e-13)
    assert np.allclose(op(bra, ket, params=dict(scale=0.25)), ref(bra, ket), atol=1e-13)
    with pytest.raises(ValueError):
        bound(bra, ket, params=dict(scale=1.0))
    with pytest.raises(TypeError):
        op(bra, ket, params=dict())


def test_density_act_and_tocoo():
    syst = FakeSystem([(15, 2), (5, 1)])
    m = _herm(2, 10)
    op = Density(syst, 1, where=[0, 3, 16])
    ket = _vec(35, 11)
    assert np.allclose(op.act(ket), op.tocoo().tocsr() @ ket)
    op2 = Density(FakeSystem([(15, 2)]), m)
    ket2 = _vec(30, 12)
    assert np.allclose(op2.act(ket2), op2.tocoo().tocsr() @ ket2)


def test_density_errors():
    syst = FakeSystem([(10, 2)])
    with pytest.raises(ValueError):
        Density(syst, np.array([[0, 1], [0, 0]]))            # not Hermitian
    Density(syst, np.array([[0, 1], [0, 0]]), check_hermiticity=False)
    with pytest.raises(ValueError):
        Density(syst, np.eye(3))                               # wrong norbs
    with pytest.raises(ValueError):
        Density(FakeSystem([(5, 2), (5, 1)]), np.eye(2))       # mixed norbs
    with pytest.raises(ValueError):
        Density(syst, where=[10])                              # not a site
    with pytest.raises(ValueError):
        Density(syst)(np.zeros(5))                             # wrong length
    bad = FakeSystem([(4, 1)])
    bad.site_ranges = None
    with pytest.raises(ValueError):
        Density(bad)


def test_layout_decomposition_matches_oracle():
    """The GPU path's algebra on the host: w = M^H bra, then per row
    conj(w) ket, summed over each where entry's orbitals."""
    syst = FakeSystem(RUNS)
    n = syst.site_ranges[-1][2]
    where = [5, 31, 31, 55, 0, 59]
    bra, ket = _vec(n, 13), _vec(n, 14)
    mats = {}

    def onsite(site):
        norbs = 2 if site < 30 else (1 if site < 50 else 3)
        return mats.setdefault(site, _herm(norbs, site))

    for onsite_arg, ref in (
            (0.7, oracle_density(syst.site_ranges, 0.7, where)),
            (onsite, oracle_density(syst.site_ranges, [onsite(s) for s in where], where))):
        lay = density_layout(Density(syst, onsite_arg, where=where).bind())
        w = lay.matrix.conj().T @ bra
        rows = (np.conj(w) * ket)[lay.orbs]
        got = lay.site_sums(rows[None])[0]
        assert np.allclose(got, ref(bra, ket), rtol=1e-13, atol=1e-13)
        total = density_layout(Density(syst, onsite_arg, where=where, sum=True).bind())
        assert abs(np.vdot(bra, total.matrix @ ket) - np.sum(ref(bra, ket))) < 1e-11


def test_plain_matrix_system():
    h = scipy.sparse.random(50, 50, density=0.1, random_state=1)
    op = Density(h, where=[3, 7])
    bra, ket = _vec(50, 15), _vec(50, 16)
    assert np.allclose(op(bra, ket), np.conj(bra[[3, 7]]) * ket[[3, 7]])
