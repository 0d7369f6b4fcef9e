"""Tight-binding Hamiltonians of the benchmark workloads, written directly as
int32 CSR (the format kwant/_system.pyx:118-119 produces) instead of through
kwant's Builder.

Site order is cell-major with the orbitals of a site (or the sites of a
unit cell) contiguous, so the SpMM gathers of one matrix row touch
neighbouring rows of the vector block.  Exact zeros are dropped as
kwant/_system.pyx:129, 145 does; the Hermitian partner of every hopping is
stored explicitly.
"""
import numpy as np
import scipy.sparse as sp

SQRT3 = np.sqrt(3.0)


def _assemble(n, rows, cols, vals, diag=None):
    """CSR from directed hopping lists (both directions given) + diagonal."""
    if diag is not None:
        idx = np.arange(n)
        keep = diag != 0
        rows = np.concatenate([rows, idx[keep]])
        cols = np.concatenate([cols, idx[keep]])
        vals = np.concatenate([vals, diag[keep].astype(vals.dtype)])
    m = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    m.sum_duplicates()
    m.eliminate_zeros()
    m.sort_indices()
    m.indptr = m.indptr.astype(np.int32)
    m.indices = m.indices.astype(np.int32)
    return m


def square(lx, ly=None, t=1.0, disorder=0.0, seed=None, shift=0.0, flux=0.0):
    """Square lattice lx x ly, one orbital, hopping t, on-site
    U[-disorder/2, disorder/2] + shift, Peierls phase exp(-2 pi i flux y) on
    x-hops (Landau gauge).  Returns (csr, positions)."""
    ly = lx if ly is None else ly
    n = lx * ly
    x, y = np.divmod(np.arange(n), ly)
    rng = np.random.RandomState(seed)
    onsite = (rng.uniform(-0.5, 0.5, n) * disorder if disorder else np.zeros(n)) + shift
    rows, cols, vals = [], [], []
    hx = np.nonzero(x < lx - 1)[0]          # (x, y) -> (x+1, y)
    phase = np.exp(-2j * np.pi * flux * y[hx]) if flux else np.ones(len(hx))
    rows += [hx + ly, hx]
    cols += [hx, hx + ly]
    vals += [t * phase, np.conj(t * phase)]
    hy = np.nonzero(y < ly - 1)[0]          # (x, y) -> (x, y+1)
    rows += [hy + 1, hy]
    cols += [hy, hy + 1]
    vals += [np.full(len(hy), t, complex), np.full(len(hy), np.conj(t), complex)]
    csr = _assemble(n, np.concatenate(rows), np.concatenate(cols),
                    np.concatenate(vals).astype(complex), onsite.astype(complex))
    pos = np.stack([x, y], axis=1).astype(float)
    return csr, pos


class SquarePattern:
    """Sparsity pattern of :func:`square` (every diagonal entry kept) with the
    per-entry data that turns (disorder W, flux phi, shift mu) into values:

        value = exp(2 pi i phi peierls[e])     on hoppings (t = 1)
        value = W u[row] + mu                  on the diagonal

    with u = RandomState(seed).uniform(-0.5, 0.5, n), the draw of
    :func:`square`.  ``values(W, phi, mu)`` equals ``square(l, disorder=W,
    seed=seed, shift=mu, flux=phi)`` entry for entry wherever that matrix
    keeps the entry (it drops exact zeros)."""

    def __init__(self, l, seed=None):
        n = l * l
        x, y = np.divmod(np.arange(n), l)
        hx = np.nonzero(x < l - 1)[0]
        hy = np.nonzero(y < l - 1)[0]
        rows = np.concatenate([hx + l, hx, hy + 1, hy, np.arange(n)])
        cols = np.concatenate([hx, hx + l, hy, hy + 1, np.arange(n)])
        m = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
        m.sort_indices()
        self.l, self.n, self.nnz = l, n, m.nnz
        self.indptr = m.indptr.astype(np.int32)
        self.indices = m.indices.astype(np.int32)
        r = np.repeat(np.arange(n), np.diff(m.indptr))
        c = self.indices.astype(np.int64)
        self.is_diag = r == c
        # Landau gauge of square(): (x+1, y) <- (x, y) carries exp(-2 pi i phi y)
        self.peierls = np.where(np.abs(x[r] - x[c]) == 1, -(x[r] - x[c]) * y[r], 0).astype(float)
        self.u = np.random.RandomState(seed).uniform(-0.5, 0.5, n)
        self.u_entry = np.where(self.is_diag, self.u[r], 0.0)

    def values(self, disorder, flux, shift):
        hop = np.exp(2j * np.pi * flux * self.peierls)
        return np.where(self.is_diag, disorder * self.u_entry + shift, hop)

    def matrix(self, disorder, flux, shift):
        return sp.csr_matrix((self.values(disorder, flux, shift), self.indices, self.indptr),
                             shape=(self.n, self.n))


def honeycomb(l1, l2=None, t=1.0, disorder=0.0, seed=None):
    """Graphene parallelogram of l1 x l2 unit cells with kwant's honeycomb
    geometry (primitive vectors (1, 0), (1/2, sqrt3/2); basis (0, 0),
    (0, 1/sqrt3)), open boundaries, NN hopping t, on-site
    U[-disorder/2, disorder/2].  Site (i, j, s) has index 2 (i l2 + j) + s.
    A(i, j) bonds to B(i, j), B(i, j-1), B(i+1, j-1).  Returns (csr, pos)."""
    l2 = l1 if l2 is None else l2
    cells = l1 * l2
    n = 2 * cells
    i, j = np.divmod(np.arange(cells), l2)
    a_idx = 2 * np.arange(cells)
    rows, cols = [], []
    # B(i, j)
    rows.append(a_idx)
    cols.append(a_idx + 1)
    # B(i, j-1)
    m = j > 0
    rows.append(a_idx[m])
    cols.append(2 * (i[m] * l2 + j[m] - 1) + 1)
    # B(i+1, j-1)
    m = (j > 0) & (i < l1 - 1)
    rows.append(a_idx[m])
    cols.append(2 * ((i[m] + 1) * l2 + j[m] - 1) + 1)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.full(len(r), t, dtype=complex)
    rng = np.random.RandomState(seed)
    onsite = rng.uniform(-0.5, 0.5, n) * disorder if disorder else np.zeros(n)
    csr = _assemble(n, np.concatenate([r, c]), np.concatenate([c, r]),
                    np.concatenate([v, np.conj(v)]), onsite.astype(complex))
    px = i + 0.5 * j
    py = (SQRT3 / 2) * j
    pos = np.empty((n, 2))
    pos[0::2, 0], pos[0::2, 1] = px, py
    pos[1::2, 0], pos[1::2, 1] = px, py + 1 / SQRT3
    return csr, pos


def honeycomb_bounds(t=1.0, disorder=0.0):
    """Gershgorin bounds of honeycomb(): |onsite| + 3|t|."""
    r = 3 * abs(t) + disorder / 2
    return (-r, r)


def cubic_4orb(L, t=1.0, onsite_sigma=0.3, seed=None):
    """Cubic L^3 lattice with 4 orbitals per site: random Hermitian on-site
    block A = (X + X^H)/2, X_ij ~ N(0, s) + i N(0, s); hopping t * I_4 to the
    6 nearest neighbours, open boundaries.  Orbital index 4 * site + o with
    site = (x L + y) L + z.  Returns the CSR."""
    nsite = L ** 3
    n = 4 * nsite
    rng = np.random.RandomState(seed)
    x = rng.normal(0, onsite_sigma, (nsite, 4, 4)) + 1j * rng.normal(0, onsite_sigma,
                                                                     (nsite, 4, 4))
    blocks = (x + np.conj(np.swapaxes(x, 1, 2))) / 2
    base = 4 * np.arange(nsite)
    o_r, o_c = np.meshgrid(np.arange(4), np.arange(4), indexing='ij')
    rows = [(base[:, None, None] + o_r[None]).ravel()]
    cols = [(base[:, None, None] + o_c[None]).ravel()]
    vals = [blocks.ravel()]
    s = np.arange(nsite)
    sx, rem = np.divmod(s, L * L)
    sy, sz = np.divmod(rem, L)
    for coord, stride in ((sx, L * L), (sy, L), (sz, 1)):
        src = s[coord < L - 1]
        dst = src + stride
        for o in range(4):
            rows += [4 * src + o, 4 * dst + o]
            cols += [4 * dst + o, 4 * src + o]
            vals += [np.full(len(src), t, complex), np.full(len(src), np.conj(t), complex)]
    return _assemble(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def cubic_4orb_bounds(csr):
    """Gershgorin bounds of a Hermitian CSR: max row sum of |H_ij|."""
    r = np.asarray(abs(csr).sum(axis=1)).ravel().max()
    return (-r, r)


# 4x4 Dirac matrices: Gamma1 = s_x t_x, Gamma2 = s_y t_x, Gamma5 = s_0 t_z.
_S = [np.eye(2), np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]),
      np.array([[1, 0], [0, -1]])]
GAMMA1 = np.kron(_S[1], _S[1]).astype(complex)
GAMMA2 = np.kron(_S[2], _S[1]).astype(complex)
GAMMA5 = np.kron(_S[0], _S[3]).astype(complex)


def dirac_4band(L, A=1.0, B=-1.0, M=-2.0, flux=0.0, disorder=0.0, seed=None):
    """Four-band Dirac model on an L x L square lattice (open):
    on-site (M - 4B) G5 + W U[-1/2, 1/2] I4, hops T_x = B G5 - (iA/2) G1 and
    T_y = B G5 - (iA/2) G2 from r to r + e_x / r + e_y, the x-hop multiplied by
    exp(-2 pi i flux y).  Orbital index 4 (x L + y) + o.  Returns (csr, pos)."""
    nsite = L * L
    n = 4 * nsite
    sx, sy = np.divmod(np.arange(nsite), L)
    rng = np.random.RandomState(seed)
    eps = rng.uniform(-0.5, 0.5, nsite) * disorder if disorder else np.zeros(nsite)
    onsite = (M - 4 * B) * GAMMA5
    tx = B * GAMMA5 - 0.5j * A * GAMMA1
    ty = B * GAMMA5 - 0.5j * A * GAMMA2
    rows, cols, vals = [], [], []
    o_r, o_c = np.meshgrid(np.arange(4), np.arange(4), indexing='ij')
    o_r, o_c = o_r.ravel(), o_c.ravel()

    def put(src, dst, blocks):
        # H[dst, src] = blocks (k, 4, 4) and its Hermitian partner
        k = len(src)
        rr = (4 * dst[:, None] + o_r[None]).ravel()
        cc = (4 * src[:, None] + o_c[None]).ravel()
        vv = blocks.reshape(k, 16).ravel()
        rows.extend([rr, cc])
        cols.extend([cc, rr])
        vals.extend([vv, np.conj(vv)])

    s = np.arange(nsite)
    m = sx < L - 1
    src = s[m]
    phase = np.exp(-2j * np.pi * flux * sy[m]) if flux else np.ones(len(src))
    put(src, src + L, tx[None] * phase[:, None, None])
    m = sy < L - 1
    src = s[m]
    put(src, src + 1, np.broadcast_to(ty, (len(src), 4, 4)))
    diag_blocks = onsite[None] + eps[:, None, None] * np.eye(4)[None]
    rows.append((4 * s[:, None] + o_r[None]).ravel())
    cols.append((4 * s[:, None] + o_c[None]).ravel())
    vals.append(diag_blocks.reshape(nsite, 16).ravel())
    csr = _assemble(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))
    pos = np.repeat(np.stack([sx, sy], axis=1).astype(float), 4, axis=0)
    return csr, pos
