"""Row orders with short neighbour distances for the recursion kernel.

The Chebyshev step gathers a_n at the columns of each row.  When neighbouring
sites are far apart in the row order (a system assembled in a random or
hashed site order), every gather misses the GPU's caches: on the 2M-site
graphene block a random site order costs 58 % more time per step than the
lattice order.  A symmetric permutation P H P^T with the vectors permuted
alike changes nothing in the moments -- each row keeps its entries in the
same order, so every row sum is the same sequence of operations, and the dot
products are order-independent compensated sums -- while reverse
Cuthill-McKee (scipy.sparse.csgraph) restores short distances (the same block
after a random shuffle and RCM: within 5 % of the lattice order).

Conventions: ``order`` maps device row i to caller row order[i]; ``inverse``
is its inverse (caller row j lives at device row inverse[j]).
"""
import numpy as np
import scipy.sparse
from scipy.sparse.csgraph import reverse_cuthill_mckee

# 'auto' reorders when the mean |i - j| over the entries exceeds this many
# times sqrt(N) (a 2D lattice in row-major order sits near 0.35 sqrt(N), a
# random order near N / 3) and RCM at least halves it.
AUTO_SPREAD = 16.0


def mean_distance(csr):
    """Mean |i - j| over the stored entries of a CSR matrix."""
    csr = scipy.sparse.csr_matrix(csr)
    if csr.nnz == 0:
        return 0.0
    rows = np.repeat(np.arange(csr.shape[0], dtype=np.int64), np.diff(csr.indptr))
    return float(np.mean(np.abs(csr.indices.astype(np.int64) - rows)))


def permute_csr(csr, order):
    """P H P^T for ``order`` (device row i = caller row order[i]): row i of
    the result is row order[i] with its entries in their original order and
    the columns relabelled, so y = (P H P^T)(P x) equals P (H x) exactly."""
    csr = scipy.sparse.csr_matrix(csr)
    n = csr.shape[0]
    if csr.shape[1] != n:
        raise ValueError("a symmetric permutation needs a square matrix")
    order = np.asarray(order, dtype=np.int64)
    if order.shape != (n,) or not np.array_equal(np.sort(order), np.arange(n)):
        raise ValueError("order must be a permutation of range(%d)" % n)
    inverse = np.empty(n, dtype=np.int64)
    inverse[order] = np.arange(n)
    lengths = np.diff(csr.indptr).astype(np.int64)[order]
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lengths, out=indptr[1:])
    src = np.repeat(csr.indptr[:-1].astype(np.int64)[order] - indptr[:-1], lengths) + \
        np.arange(indptr[-1], dtype=np.int64)
    out = scipy.sparse.csr_matrix((csr.data[src], inverse[csr.indices[src]], indptr),
                                  shape=csr.shape)
    out.has_sorted_indices = False  # entry order is the caller's, kept on purpose
    return out


def locality_order(csr):
    """Reverse Cuthill-McKee order of the symmetrised pattern of ``csr``."""
    csr = scipy.sparse.csr_matrix(csr)
    pattern = (abs(csr) + abs(csr).T).tocsr()
    return reverse_cuthill_mckee(pattern, symmetric_mode=True).astype(np.int64)


def choose_order(csr, reorder):
    """The device row order for ``reorder`` in (False, True, 'auto', or an
    explicit permutation); None means the caller's order."""
    if reorder is False or reorder is None:
        return None
    if reorder is True:
        return locality_order(csr)
    if isinstance(reorder, str):
        if reorder != 'auto':
            raise ValueError("reorder must be True, False, 'auto' or a permutation")
        n = csr.shape[0]
        spread = mean_distance(csr)
        if n < 4096 or spread <= AUTO_SPREAD * np.sqrt(n):
            return None
        order = locality_order(csr)
        if 2 * mean_distance(permute_csr(csr, order)) > spread:
            return None
        return order
    return np.asarray(reorder, dtype=np.int64)


class Tiling:
    """A device row order that groups the rows into compact tiles, and the
    tiles' row boundaries: device rows tile_begin[t] .. tile_begin[t+1]-1 form
    tile t.  ``order`` follows the module's convention (device row i = caller
    row order[i]).  With tiles, the recursion runs two Chebyshev steps per
    launch (each tile's halo recomputed from LDS, see kpm_cheb_steps2)."""

    def __init__(self, order, tile_begin):
        self.order = np.asarray(order, dtype=np.int64)
        self.tile_begin = np.asarray(tile_begin, dtype=np.int64)
        n = len(self.order)
        tb = self.tile_begin
        if tb.ndim != 1 or len(tb) < 2 or tb[0] != 0 or tb[-1] != n or np.any(np.diff(tb) <= 0):
            raise ValueError("tile_begin must rise strictly from 0 to the row count")

    @property
    def n_tiles(self):
        return len(self.tile_begin) - 1


def graph_coordinates(csr):
    """Pseudo-coordinates of the rows from the matrix's graph alone: hop
    distances to two far-apart rows.  Row a is the row farthest from row 0, b
    the row farthest from a, and c the row whose distance to the nearer of a
    and b is largest; the coordinates are (d_a, d_c).  On a lattice, hop
    distances to two such corners grow along two independent directions (the
    corners (0, 0) and (W, 0) of a rectangle see x + y and W - x + y), so boxes
    of these coordinates are compact patches of the lattice.  Rows that a
    cannot reach (other components) get one coordinate past the largest."""
    from scipy.sparse.csgraph import dijkstra
    csr = scipy.sparse.csr_matrix(csr)
    n = csr.shape[0]
    pattern = scipy.sparse.csr_matrix(
        (np.ones(len(csr.indices), dtype=np.int8), csr.indices, csr.indptr), shape=csr.shape)

    def dist(row):
        d = dijkstra(pattern, directed=False, indices=int(row), unweighted=True)
        finite = np.isfinite(d)
        d[~finite] = d[finite].max() + 1
        return d

    if n == 0:
        return np.zeros((0, 2))
    da = dist(np.argmax(dist(0)))
    db = dist(np.argmax(da))
    dc = dist(np.argmax(np.minimum(da, db)))
    return np.stack([da, dc], axis=1)


def tile_order(coords, tile_rows=512, iterations=4):
    """Tiling of rows with coordinates ``coords`` ((N,) or (N, d) array, one
    point per row -- the site position, repeated for each orbital of a site)
    into axis-aligned boxes of about ``tile_rows`` rows.

    The box edge starts from the mean density of the bounding box and is then
    scaled so that the 95th percentile of the occupied boxes' row counts meets
    ``tile_rows``; boxes that still hold more rows are cut into consecutive
    pieces.  Boxes are numbered with the first coordinate fastest, and within
    a box the rows keep their relative order.  Returns a Tiling."""
    x = np.asarray(coords, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    n, d = x.shape
    if n == 0:
        raise ValueError("no rows to tile")
    tile_rows = int(tile_rows)
    if tile_rows < 1:
        raise ValueError("tile_rows must be positive")
    lo = x.min(axis=0)
    span = np.maximum(x.max(axis=0) - lo, 1e-12)
    edge = (np.prod(span) * tile_rows / n) ** (1.0 / d)
    for _ in range(max(1, int(iterations))):
        box = np.floor((x - lo) / edge).astype(np.int64)
        dims = box.max(axis=0) + 1
        key = np.ravel_multi_index(tuple(box[:, k] for k in range(d - 1, -1, -1)),
                                   tuple(dims[::-1]))
        counts = np.bincount(key)
        occ = counts[counts > 0]
        p95 = float(np.percentile(occ, 95))
        if abs(p95 - tile_rows) <= 0.02 * tile_rows:
            break
        edge *= (tile_rows / p95) ** (1.0 / d)
    box = np.floor((x - lo) / edge).astype(np.int64)
    dims = box.max(axis=0) + 1
    key = np.ravel_multi_index(tuple(box[:, k] for k in range(d - 1, -1, -1)), tuple(dims[::-1]))
    order = np.argsort(key, kind="stable")
    sorted_key = key[order]
    starts = np.flatnonzero(np.r_[True, sorted_key[1:] != sorted_key[:-1]])
    ends = np.r_[starts[1:], n]
    # rows at the same point (the orbitals of one site) stay in one tile
    xs = x[order]
    site_start = np.flatnonzero(np.r_[True, np.any(xs[1:] != xs[:-1], axis=1)])
    bounds = [0]
    for s, e in zip(starts, ends):
        if e - s > tile_rows:
            # a box with more rows than a tile: pieces of about equal size, cut
            # between sites (a site larger than a tile is cut inside)
            target = -(-(e - s) // -(-(e - s) // tile_rows))
            inner = site_start[(site_start > s) & (site_start < e)]
            begin = prev = s
            for b in list(inner) + [e]:
                b = int(b)
                while b - begin > tile_rows and prev == begin:
                    begin += tile_rows
                    bounds.append(begin)
                    prev = begin
                if b - begin > target and prev > begin:
                    bounds.append(prev)
                    begin = prev
                prev = b
        bounds.append(e)
    return Tiling(order, np.asarray(bounds, dtype=np.int64))
