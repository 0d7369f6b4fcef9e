"""Row reordering (kwant_amd.ordering): the permuted matrix reproduces the
caller's products exactly, RCM shortens neighbour distances, and 'auto'
only reorders matrices whose order is poor."""
import numpy as np
import pytest
import scipy.sparse as sp

from kwant_amd import ordering
from kwant_amd.lattices import honeycomb, square


def test_permute_csr_is_exact_relabelling():
    h, _ = square(30, disorder=1.0, seed=3)
    h = h.tocsr()
    rng = np.random.RandomState(1)
    order = rng.permutation(h.shape[0])
    hp = ordering.permute_csr(h, order)
    x = rng.randn(h.shape[0]) + 1j * rng.randn(h.shape[0])
    # same row sums in the same order: bitwise equal products
    assert np.array_equal(hp @ x[order], (h @ x)[order])
    # every row keeps its entries in their original order
    for i in (0, 17, 899):
        r = order[i]
        inv = np.empty_like(order)
        inv[order] = np.arange(len(order))
        assert np.array_equal(hp.indices[hp.indptr[i]:hp.indptr[i + 1]],
                              inv[h.indices[h.indptr[r]:h.indptr[r + 1]]])
        assert np.array_equal(hp.data[hp.indptr[i]:hp.indptr[i + 1]],
                              h.data[h.indptr[r]:h.indptr[r + 1]])


def test_permute_csr_rejects_non_permutations():
    h = sp.identity(5, format="csr")
    with pytest.raises(ValueError):
        ordering.permute_csr(h, [0, 1, 2, 3, 3])
    with pytest.raises(ValueError):
        ordering.permute_csr(sp.csr_matrix((3, 4)), [0, 1, 2])


def test_rcm_restores_locality_and_auto_decides():
    h, _ = honeycomb(80, disorder=0.5, seed=2)
    n = h.shape[0]
    natural = ordering.mean_distance(h)
    assert ordering.choose_order(h, 'auto') is None          # lattice order is fine
    assert ordering.choose_order(h, False) is None
    shuffled = ordering.permute_csr(h, np.random.RandomState(0).permutation(n))
    assert ordering.mean_distance(shuffled) > 50 * natural
    order = ordering.choose_order(shuffled, 'auto')
    assert order is not None and sorted(order) == list(range(n))
    assert ordering.mean_distance(ordering.permute_csr(shuffled, order)) < 3 * natural
    explicit = ordering.choose_order(shuffled, np.arange(n)[::-1])
    assert np.array_equal(explicit, np.arange(n)[::-1])
    with pytest.raises(ValueError):
        ordering.choose_order(h, 'sometimes')
