"""The Density oracle pinned by closed forms and brute force, and the host
side of kwant_amd.operator (Density, Current, Source and their layouts) checked
against it.

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
        assert abs(np.vdot(bra, total.act_matrix @ ket) - np.sum(ref(bra, ket))) < 1e-11


def test_plain_matrix_system():
    h = scipy.sparse.random(50, 50, density=0.1, random_state=1)
    op = Density(h, where=[3, 7])
    bra, ket = _vec(50, 15), _vec(50, 16)
    assert np.allclose(op(bra, ket), np.conj(bra[[3, 7]]) * ket[[3, 7]])


# -- Current and Source --------------------------------------------------------

from kwant_amd.operator import Current, Source, operator_layout  # noqa: E402
from oracle.operator_oracle import current as oracle_current  # noqa: E402
from oracle.operator_oracle import source as oracle_source  # noqa: E402


def _blocks_of(syst, pairs):
    return [syst.hamiltonian(a, b) for a, b in pairs]


def test_oracle_chain_current_closed_form():
    n, t = 12, 0.7 - 0.2j
    h = scipy.sparse.diags([np.full(n - 1, t), np.full(n - 1, np.conj(t))], [1, -1]).tocsr()
    ranges = [(0, 1, 0), (n, 0, n)]
    psi = _vec(n, 20)
    hops = [(a, a + 1) for a in range(n - 1)]
    got = oracle_current(ranges, 1, hops, [h[a, b] for a, b in hops])(psi)
    want = [2 * np.imag(t * np.conj(psi[a]) * psi[a + 1]) for a, _ in hops]
    assert np.isrealobj(got)
    assert np.allclose(got, want, rtol=0, atol=1e-13)


def test_oracle_current_brute_force_and_continuity():
    syst = FakeSystem([(12, 2)], seed=21, density=0.3)
    hops = list(syst.graph)
    m = _herm(2, 22)
    bra, ket = _vec(24, 23), _vec(24, 24)
    got = oracle_current(syst.site_ranges, m, hops, _blocks_of(syst, hops))(bra, ket)
    for w, (a, b) in enumerate(hops):
        j = np.zeros((24, 24), dtype=complex)         # dense J_w
        hab = syst.h[2 * a:2 * a + 2, 2 * b:2 * b + 2].toarray()
        j[2 * b:2 * b + 2, 2 * a:2 * a + 2] = 1j * hab.conj().T @ m
        j[2 * a:2 * a + 2, 2 * b:2 * b + 2] = -1j * m @ hab
        assert abs(got[w] - np.vdot(bra, j @ ket)) < 1e-12
    # stationary continuity: no net current out of any site in an eigenstate
    psi = np.linalg.eigh(syst.h.toarray())[1][:, 3]
    cur = oracle_current(syst.site_ranges, 1, hops, _blocks_of(syst, hops))(psi)
    out = np.zeros(12)
    for (a, b), c in zip(hops, cur):
        out[a] += c
    assert np.max(np.abs(out)) < 1e-12


def test_oracle_source_closed_form():
    sx = np.array([[0, 1], [1, 0]], complex)
    sy = np.array([[0, -1j], [1j, 0]])
    sz = np.diag([1.0, -1.0]).astype(complex)
    ranges = [(0, 2, 0), (1, 0, 2)]
    bra, ket = _vec(2, 25), _vec(2, 26)
    got = oracle_source(ranges, sz, [0], [sx])(bra, ket)
    assert abs(got[0] - np.vdot(bra, 2 * sy @ ket)) < 1e-13
    assert abs(oracle_source(ranges, 1, [0], [sx])(bra, ket)[0]) < 1e-14


@pytest.mark.parametrize("sum_", [False, True])
def test_current_and_source_match_oracle(sum_):
    syst = FakeSystem(RUNS, seed=27, density=0.05)
    n = syst.site_ranges[-1][2]
    bra, ket = _vec(n, 28), _vec(n, 29)
    hops = list(syst.graph)[::7]
    hops = hops + hops[:2]                          # repeated hoppings
    for onsite in (1, 0.5):
        cur = Current(syst, onsite, where=hops, sum=sum_)
        ref = oracle_current(syst.site_ranges, onsite, hops, _blocks_of(syst, hops), sum=sum_)
        assert np.allclose(cur(bra, ket), ref(bra, ket), rtol=1e-12, atol=1e-12)
        assert np.allclose(cur(bra), ref(bra), rtol=1e-12, atol=1e-12)
    sites = [3, 40, 40, 55]
    mats = {s: _herm(syst.hamiltonian(s, s).shape[0], s) for s in set(sites)}
    src = Source(syst, lambda s: mats[s], where=sites, sum=sum_)
    ref = oracle_source(syst.site_ranges, [mats[s] for s in sites], sites,
                        _blocks_of(syst, [(s, s) for s in sites]), sum=sum_)
    assert np.allclose(src(bra, ket), ref(bra, ket), rtol=1e-12, atol=1e-12)


def test_current_defaults_and_errors():
    syst = FakeSystem([(10, 1)], seed=30, density=0.3)
    cur = Current(syst)
    assert cur.where.shape == (len(list(syst.graph)), 2)
    with pytest.raises(ValueError):
        Current(syst, where=[(0, 0)])               # not a hopping
    plain = Current(syst.h)                           # a matrix: its off-diagonal entries
    assert plain.where.shape[0] == syst.h.nnz - np.count_nonzero(syst.h.diagonal())
    ket = _vec(10, 31)
    assert np.allclose(cur.act(ket), cur.tocoo().tocsr() @ ket)


def test_bond_layout_decomposition_matches_oracle():
    """The bond kernel's algebra on the host: per entry (r, s, h)
    i (conj(h) conj(bra_s) v_r - conj(w_r) h ket_s), v = M ket, w = M^H bra,
    summed over each hopping's entries."""
    syst = FakeSystem(RUNS, seed=32, density=0.05)
    n = syst.site_ranges[-1][2]
    bra, ket = _vec(n, 33), _vec(n, 34)
    hops = list(syst.graph)[::5]
    hops = hops + hops[:1]
    for onsite in (1, 0.3):
        lay = operator_layout(Current(syst, onsite, where=hops).bind())
        ref = oracle_current(syst.site_ranges, onsite, hops, _blocks_of(syst, hops))
        m = scipy.sparse.identity(n, format='csr') if lay.onsite is None else lay.onsite
        v, w = m @ ket, m.conj().T @ bra
        r, c, h = lay.rows, lay.cols, lay.values
        rec = 1j * (np.conj(h) * np.conj(bra[c]) * v[r] - np.conj(w[r]) * h * ket[c])
        got = lay.hop_sums(rec[None])[0]
        assert np.allclose(got, ref(bra, ket), rtol=1e-12, atol=1e-12)
        total = operator_layout(Current(syst, onsite, where=hops, sum=True).bind())
        assert abs(np.vdot(bra, total.act_matrix @ ket) - np.sum(ref(bra, ket))) < 1e-11


# -- kwant's own operator objects, recognised through their public attributes --

class _BlockSparse:
    """The public fields of kwant.operator.BlockSparseMatrix: blocks stored
    flattened in ``data`` from ``data_offsets``."""

    def __init__(self, blocks):
        flat = [np.asarray(b, dtype=complex).ravel() for b in blocks]
        self.data = np.concatenate(flat) if flat else np.zeros(0, complex)
        self.data_offsets = np.concatenate([[0], np.cumsum([len(f) for f in flat])])[:-1]


def _kwant_like(kind, ours, bound_onsite=None, bound_ham=None):
    """An object with the class name and public attributes of kwant's own
    operator (operator.pxd: where, _site_ranges, onsite, sum,
    check_hermiticity, _bound_onsite, _bound_hamiltonian)."""
    obj = type(kind, (), {})()
    obj.where = np.asarray(ours.where, dtype=np.int32)
    obj._site_ranges = np.asarray(ours._site_ranges, dtype=np.int32)
    obj.onsite = ours.onsite
    obj.sum = ours.sum
    obj.check_hermiticity = ours.check_hermiticity
    obj._bound_onsite = bound_onsite
    obj._bound_hamiltonian = bound_ham
    return obj


def _same_layout(a, b):
    assert type(a) is type(b)
    assert abs(a.act_matrix - b.act_matrix).max() == 0
    assert abs(a.matrix - b.matrix).max() == 0
    for name in ('orbs', 'starts', 'rows', 'cols', 'values', 'where_index'):
        if hasattr(a, name):
            assert np.array_equal(getattr(a, name), getattr(b, name)), name


def test_kwant_objects_give_the_same_layouts():
    syst = FakeSystem(RUNS, seed=40, density=0.05)
    sites = [3, 40, 40, 55]
    mats = {s: _herm(syst.hamiltonian(s, s).shape[0], s + 1) for s in set(sites)}
    # Density with a constant onsite, and with a callable one (bound blocks)
    ours = Density(syst, 0.5, where=sites)
    _same_layout(operator_layout(_kwant_like('Density', ours)), operator_layout(ours))
    ours = Density(syst, lambda s: mats[s], where=sites).bind()
    kw = _kwant_like('Density', ours, bound_onsite=_BlockSparse(ours._bound_onsite))
    kw.onsite = lambda s: mats[s]
    _same_layout(operator_layout(kw), operator_layout(ours))
    # Source: the onsite Hamiltonian blocks come from _bound_hamiltonian
    ours = Source(syst, lambda s: mats[s], where=sites).bind()
    kw = _kwant_like('Source', ours, bound_onsite=_BlockSparse(ours._bound_onsite),
                     bound_ham=_BlockSparse(ours._bound_hamiltonian))
    kw.onsite = lambda s: mats[s]
    _same_layout(operator_layout(kw), operator_layout(ours))
    # Current over some hoppings, with a constant onsite
    hops = list(syst.graph)[::11]
    ours = Current(syst, 1, where=hops).bind()
    kw = _kwant_like('Current', ours, bound_ham=_BlockSparse(ours._bound_hamiltonian))
    _same_layout(operator_layout(kw), operator_layout(ours))
    # an unbound kwant Current (no Hamiltonian blocks yet) is not taken
    assert operator_layout(_kwant_like('Current', ours)) is None
