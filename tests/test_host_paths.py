"""Host-side set-up paths (CPU): the velocity operator, the random-phase
vectors and the orbital lists of kwant systems give bitwise what the oracle's
restatement of kwant gives, on inputs that reach every branch."""
import numpy as np
import scipy.sparse as sp

from oracle import kpm_oracle as ko


def _hermitian(n, seed, density=0.05):
    rng = np.random.default_rng(seed)
    a = sp.random(n, n, density=density, random_state=seed, dtype=complex,
                  data_rvs=lambda k: rng.normal(size=k) + 1j * rng.normal(size=k))
    return ((a + a.conj().T) / 2 + sp.diags(rng.uniform(-1, 1, n))).tocsr()


def _same_csr(a, b):
    a, b = sp.csr_matrix(a), sp.csr_matrix(b)
    return (a.shape == b.shape and np.array_equal(a.indptr, b.indptr)
            and np.array_equal(a.indices, b.indices) and np.array_equal(a.data, b.data))


def test_velocity_equals_oracle_bitwise():
    """The CSR route (canonical input) and the COO route (duplicates, COO
    input) both give the oracle's velocity operator bit for bit, in 2 and 3
    dimensions; a direction past the positions' dimension raises what the
    oracle raises (kwant's check lets 'z' of 2-d positions through to the
    indexing, which raises IndexError)."""
    from kwant_amd.correlator import _velocity
    h = _hermitian(300, 1)
    assert h.has_canonical_format
    rng = np.random.default_rng(2)
    for dim in (2, 3):
        pos = rng.uniform(-5, 5, (300, dim))
        for d in "xyz"[:dim]:
            assert _same_csr(_velocity(h, None, d, pos), ko.velocity(h, d, pos)), (dim, d)
    # duplicates: a COO with a repeated entry, and a CSR that is not canonical
    coo = h.tocoo()
    dup = sp.coo_matrix((np.r_[coo.data, 0.5 + 0.25j], (np.r_[coo.row, 3], np.r_[coo.col, 7])),
                        shape=h.shape)
    pos = rng.uniform(-5, 5, (300, 2))
    assert _same_csr(_velocity(dup, None, 'x', pos), ko.velocity(dup, 'x', pos))
    raw = sp.csr_matrix((dup.data, (dup.row, dup.col)), shape=h.shape)
    raw.has_canonical_format = False
    assert _same_csr(_velocity(raw, None, 'y', pos), ko.velocity(raw, 'y', pos))
    for f in (lambda: _velocity(h, None, 'z', pos), lambda: ko.velocity(h, 'z', pos)):
        try:
            f()
        except IndexError:
            pass
        else:
            raise AssertionError("direction z of 2-d positions must raise")


def test_random_vectors_equal_oracle_bitwise():
    """Every orbital (a slice) and a subset given as indices draw the oracle's
    vectors bit for bit from the same seed."""
    from kwant_amd.kpm import RandomVectors
    h = _hermitian(200, 3)
    for where in (None, [5, 0, 199, 17], np.arange(0, 200, 3)):
        ours, ref = RandomVectors(h, where=where, rng=11), ko.random_vectors(h, where, rng=11)
        for _ in range(3):
            assert np.array_equal(next(ours), next(ref)), where


def test_system_orbitals_in_site_order():
    """Orbitals of the sites of ``where`` (sites with 2, 1 and 3 orbitals, in
    the order given), and every orbital for ``where=None``."""
    from fake_system import FakeSystem
    from kwant_amd.kpm import _orbitals_of
    syst = FakeSystem([(3, 2), (2, 1), (2, 3)])
    tot, orbs = _orbitals_of(syst, [5, 0, 3])
    assert tot == 14
    assert list(orbs) == [8, 9, 10, 0, 1, 6]
    tot, orbs = _orbitals_of(syst, None)
    assert list(orbs) == list(range(14))
    tot, orbs = _orbitals_of(syst, [])
    assert len(orbs) == 0
