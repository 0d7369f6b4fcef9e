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


def _halo_rows(h, tile_begin):
    """Rows outside each tile that its entries reach, summed over tiles
    (the halo the two-step kernel recomputes)."""
    sizes = np.diff(tile_begin)
    tile_of = np.repeat(np.arange(len(sizes)), sizes)
    coo = sp.csr_matrix(h).tocoo()
    out = tile_of[coo.row] != tile_of[coo.col]
    pairs = np.unique(tile_of[coo.row[out]].astype(np.int64) * h.shape[0] + coo.col[out])
    return len(pairs)


def test_tile_order_groups_rows_into_compact_tiles():
    h, pos = square(64, disorder=1.0, seed=1)
    n = h.shape[0]
    tiling = ordering.tile_order(pos, 256)
    assert sorted(tiling.order) == list(range(n))
    sizes = np.diff(tiling.tile_begin)
    assert sizes.min() > 0 and sizes.max() <= 256
    # 16 x 16 boxes of a square lattice: 4 x 16 boundary neighbours per box,
    # against about 2 x 256 for tiles of 256 consecutive rows of the row-major
    # order (the rows above and below)
    hp = ordering.permute_csr(h, tiling.order)
    tiled = _halo_rows(hp, tiling.tile_begin)
    strips = _halo_rows(h, np.arange(0, n + 1, 256))
    assert tiled <= 0.3 * n
    assert tiled < 0.6 * strips
    # the rows of a tile keep their relative order
    for t in range(tiling.n_tiles):
        rows = tiling.order[tiling.tile_begin[t]:tiling.tile_begin[t + 1]]
        assert np.all(np.diff(rows) > 0)


def test_tile_order_graphene_and_three_dimensions():
    _, pos = honeycomb(50, disorder=0.5, seed=2)
    tiling = ordering.tile_order(pos, 128)
    assert np.diff(tiling.tile_begin).max() <= 128
    rng = np.random.RandomState(0)
    cloud = rng.uniform(0, 10, (5000, 3))
    tiling3 = ordering.tile_order(cloud, 100)
    assert sorted(tiling3.order) == list(range(5000))
    assert np.diff(tiling3.tile_begin).max() <= 100
    # orbitals of one site (equal positions) stay together when they fit a tile
    four = np.repeat(cloud[:500], 4, axis=0)
    t4 = ordering.tile_order(four, 64)
    site_of = t4.order // 4
    starts = t4.tile_begin[:-1]
    for s, e in zip(starts, t4.tile_begin[1:]):
        counts = np.bincount(site_of[s:e])
        assert np.all(counts[counts > 0] == 4)


def test_tiling_validation():
    with pytest.raises(ValueError):
        ordering.Tiling([0, 1, 2], [0, 2])               # does not reach the row count
    with pytest.raises(ValueError):
        ordering.Tiling([0, 1, 2], [0, 2, 2, 3])         # an empty tile
    with pytest.raises(ValueError):
        ordering.tile_order(np.zeros((0, 2)), 16)
    with pytest.raises(ValueError):
        ordering.tile_order(np.zeros((4, 2)), 0)
    t = ordering.Tiling([2, 0, 1], [0, 1, 3])
    assert t.n_tiles == 2


def test_tiles_argument_needs_positions(monkeypatch):
    """True needs site positions; 'auto' tiles only a kwant system with
    positions, at least TILES_AUTO_ROWS orbitals and the caller's consent
    (the plain DOS path with the default row order)."""
    from kwant_amd import kpm
    from kwant_amd.assembly import LatticeSystem
    h, _ = square(8, disorder=1.0, seed=1)
    with pytest.raises(ValueError, match="site positions"):
        kpm._tiling_of(h, True)
    with pytest.raises(ValueError):
        kpm._tiling_of(h, 'yes')
    assert kpm._tiling_of(h, 'auto', auto_ok=True) is None      # a matrix
    assert kpm._tiling_of(h, None) is None and kpm._tiling_of(h, False) is None
    syst = LatticeSystem([(1, 0), (0, 1)], [(0, 0)], [(0, 20), (0, 20)], norbs=2)
    assert kpm._tiling_of(syst, 'auto', auto_ok=True) is None   # 800 orbitals
    monkeypatch.setattr(kpm, "TILES_AUTO_ROWS", 800)
    assert kpm._tiling_of(syst, 'auto', auto_ok=False) is None
    tiling = kpm._tiling_of(syst, 'auto', auto_ok=True, tile_rows=64)
    assert tiling is not None and sorted(tiling.order) == list(range(800))
    # a site's two orbitals stay together and in order
    assert np.all(tiling.order[1::2] == tiling.order[0::2] + 1)


def test_graph_coordinates_tile_like_positions():
    """Tiles from hop distances alone (no positions) are about as compact as
    tiles from the site positions, on the square and honeycomb lattices; a
    graph of two components still tiles every row."""
    for h, pos in (square(48, disorder=1.0, seed=1), honeycomb(40, disorder=0.5, seed=2)):
        n = h.shape[0]
        coords = ordering.graph_coordinates(h)
        assert coords.shape == (n, 2)
        by_pos = ordering.tile_order(pos, 256)
        by_graph = ordering.tile_order(coords, 256)
        assert sorted(by_graph.order) == list(range(n))
        halo_pos = _halo_rows(ordering.permute_csr(h, by_pos.order), by_pos.tile_begin)
        halo_graph = _halo_rows(ordering.permute_csr(h, by_graph.order), by_graph.tile_begin)
        assert halo_graph < 1.3 * halo_pos, (halo_graph, halo_pos)
    a, _ = square(10, disorder=1.0, seed=3)
    b, _ = square(7, disorder=1.0, seed=4)
    two = sp.block_diag([a, b]).tocsr()
    coords = ordering.graph_coordinates(two)
    assert np.all(np.isfinite(coords))
    tiling = ordering.tile_order(coords, 32)
    assert sorted(tiling.order) == list(range(two.shape[0]))


def test_tiles_graph_for_a_matrix():
    """tiles='graph' tiles a plain matrix by its graph."""
    from kwant_amd import kpm
    h, _ = square(30, disorder=1.0, seed=1)
    tiling = kpm._tiling_of(h, 'graph', csr=sp.csr_matrix(h), tile_rows=128)
    assert sorted(tiling.order) == list(range(h.shape[0]))
    assert np.diff(tiling.tile_begin).max() <= 128
