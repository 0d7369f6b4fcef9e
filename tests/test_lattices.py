"""Hamiltonian generators: Hermiticity, the nnz counts of the benchmark
workloads, and equality with an element-by-element construction on tiny
sizes."""
import numpy as np
import scipy.sparse as sp

from kwant_amd import lattices as lat


def hermitian(m):
    return abs(m - m.conj().T).max() < 1e-14


def test_honeycomb_counts_and_structure():
    for l in (1, 2, 5, 17):
        h, pos = lat.honeycomb(l, disorder=0.5, seed=1)
        assert hermitian(h)
        bonds = 3 * l * l - 3 * l + 1
        assert h.nnz == 2 * bonds + 2 * l * l
        # every bond has length 1/sqrt(3)
        coo = sp.coo_matrix(h)
        off = coo.row != coo.col
        d = np.linalg.norm(pos[coo.row[off]] - pos[coo.col[off]], axis=1)
        assert np.allclose(d, 1 / np.sqrt(3))
    h, _ = lat.honeycomb(1000)
    assert h.nnz == 2 * (3 * 10 ** 6 - 3000 + 1)  # zero on-site dropped


def test_honeycomb_brute_force_neighbours():
    l = 4
    h, pos = lat.honeycomb(l)
    n = len(pos)
    dense = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j and abs(np.linalg.norm(pos[i] - pos[j]) - 1 / np.sqrt(3)) < 1e-9:
                dense[i, j] = 1
    assert np.array_equal(h.toarray().real, dense)


def test_square_lattice():
    h, pos = lat.square(5, 4, t=1.0, disorder=1.0, seed=3, flux=0.1)
    assert hermitian(h)
    n = 20
    assert h.shape == (n, n)
    off = h - sp.diags(h.diagonal())
    assert off.nnz == 2 * (4 * 4 + 5 * 3)
    # phase of the x-hop from (x, y) to (x+1, y) is exp(-2 pi i flux y)
    for i in range(n - 4):
        y = pos[i, 1]
        assert abs(h[i + 4, i] - np.exp(-2j * np.pi * 0.1 * y)) < 1e-14
    assert np.all(np.abs(h.diagonal()) <= 0.5)


def test_cubic_4orb_counts():
    L = 4
    h = lat.cubic_4orb(L, seed=1)
    assert hermitian(h)
    assert h.shape == (4 * L ** 3,) * 2
    assert h.nnz == 16 * L ** 3 + 24 * L * L * (L - 1)
    blk = h[:4, :4].toarray()
    assert np.allclose(blk, blk.conj().T)


def test_dirac_model_bulk_spectrum_symmetry():
    h, pos = lat.dirac_4band(6, flux=0.0)
    assert hermitian(h)
    assert h.shape == (144, 144)
    # the clean model has the chiral-like symmetry spectrum(H) = -spectrum(H)
    w = np.linalg.eigvalsh(h.toarray())
    assert np.allclose(np.sort(w), np.sort(-w), atol=1e-10)
    hf, _ = lat.dirac_4band(6, flux=1 / 36, disorder=1.0, seed=2)
    assert hermitian(hf)
    assert len(pos) == 144
