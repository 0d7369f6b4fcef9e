"""Vectorised Hamiltonian assembly for lattice systems.

kwant builds the sparse Hamiltonian of a finalized system one site and one
hopping at a time: ``make_sparse_full`` (kwant/_system.pyx:94-153) calls the
value function of every onsite and of every hopping (ts > fs) separately,
keeps the nonzero entries and the Hermitian conjugate of each hopping, and
returns a COO matrix.  For a 2M-site lattice that is millions of Python
calls before any KPM runs.  ``LatticeSystem`` evaluates each value function
once, over the arrays of all sites or of all hoppings of one kind, and
scatters the blocks into a CSR with numpy -- the same entries, in the
canonical (sorted) order that ``csr_matrix`` gives kwant's COO.

A LatticeSystem is a Bravais lattice with a basis (``prim_vecs``,
``basis``), the sites of the tag box ``tag_range`` that ``shape`` accepts,
``norbs`` orbitals per site, an onsite value and hopping kinds
(tag displacement, from-sublattice, to-sublattice, value).  Values are
constants (scalar or norbs x norbs) or functions of positions and keyword
parameters, vectorised over the first axis:
    onsite(pos, **params) -> (n,) or (n, norbs, norbs)
    hopping(pos_to, pos_from, **params) -> (m,) or (m, norbs, norbs)
It exposes what kwant_amd reads from a finalized kwant system --
``hamiltonian_submatrix``, ``site_ranges``, ``graph`` and
``hamiltonian(i, j, params=...)`` -- so SpectralDensity, Correlator,
conductivity and the Density / Current / Source operators take it directly.
"""
import inspect

import numpy as np
import scipy.sparse


def _call(fn, args, params):
    """fn(*args, **the params it names) (kwant passes parameters by name)."""
    if not callable(fn):
        return fn
    names = [p.name for p in inspect.signature(fn).parameters.values()
             if p.kind in (p.POSITIONAL_OR_KEYWORD, p.KEYWORD_ONLY)][len(args):]
    params = params or {}
    missing = [n for n in names if n not in params]
    if missing:
        raise TypeError("value function is missing parameters: " + ", ".join(missing))
    return fn(*args, **{n: params[n] for n in names})


def _blocks(value, count, norbs):
    """A value as a (count, norbs, norbs) complex array."""
    v = np.asarray(value, dtype=complex)
    if v.ndim == 0:
        v = v * np.eye(norbs)
    if v.ndim == 1:
        if norbs != 1 or v.shape[0] != count:
            raise ValueError("a value of shape %s does not fit %d sites of %d orbitals"
                             % (v.shape, count, norbs))
        v = v[:, None, None]
    if v.ndim == 2:
        v = np.broadcast_to(v, (count,) + v.shape)
    if v.shape != (count, norbs, norbs):
        raise ValueError("a value of shape %s does not fit %d sites of %d orbitals"
                         % (v.shape, count, norbs))
    return v


class _Graph:
    """Directed hopping graph (both directions of every hopping), edges
    ordered by (tail, head) as kwant's CGraph iterates them."""

    def __init__(self, num_nodes, tails, heads):
        self.num_nodes = num_nodes
        order = np.lexsort((heads, tails))
        self.tails = tails[order]
        self.heads = heads[order]

    def __iter__(self):
        return zip(self.tails.tolist(), self.heads.tolist())

    def has_edge(self, a, b):
        lo = np.searchsorted(self.tails, a, side='left')
        hi = np.searchsorted(self.tails, a, side='right')
        return bool(np.any(self.heads[lo:hi] == b))

    def out_neighbors(self, a):
        lo = np.searchsorted(self.tails, a, side='left')
        hi = np.searchsorted(self.tails, a, side='right')
        return self.heads[lo:hi]


class LatticeSystem:
    """A finite lattice system with vectorised Hamiltonian assembly (see
    the module docstring)."""

    def __init__(self, prim_vecs, basis, tag_range, shape=None, norbs=1):
        self.prim_vecs = np.atleast_2d(np.asarray(prim_vecs, dtype=float))
        self.basis = np.atleast_2d(np.asarray(basis, dtype=float))
        dim = self.prim_vecs.shape[0]
        ranges = [np.arange(lo, hi) for lo, hi in tag_range]
        if len(ranges) != dim:
            raise ValueError("tag_range needs one (lo, hi) per primitive vector")
        grids = np.meshgrid(*ranges, indexing='ij')
        tags = np.stack([g.ravel() for g in grids], axis=1).astype(np.int64)
        # sites: every tag, then every sublattice of it (tag-major order)
        nsub = self.basis.shape[0]
        tags = np.repeat(tags, nsub, axis=0)
        sub = np.tile(np.arange(nsub), len(tags) // nsub)
        pos = tags @ self.prim_vecs + self.basis[sub]
        keep = np.ones(len(pos), bool) if shape is None else \
            np.asarray(shape(pos), dtype=bool)
        self.tags, self.sublattice, self._pos = tags[keep], sub[keep], pos[keep]
        self.norbs = int(norbs)
        n = len(self._pos)
        self.site_ranges = [(0, self.norbs, 0), (n, 0, n * self.norbs)]
        # site index of every (tag, sublattice) of the box, -1 where none
        self._lo = np.array([lo for lo, _ in tag_range], dtype=np.int64)
        self._hi = np.array([hi for _, hi in tag_range], dtype=np.int64)
        self._grid = np.full(tuple(self._hi - self._lo) + (nsub,), -1, dtype=np.int64)
        self._grid[tuple((self.tags - self._lo).T) + (self.sublattice,)] = np.arange(n)
        self.onsite = 0
        self._hops = []
        self._graph = None
        self._lookup = None

    # -- definition ------------------------------------------------------------
    def set_onsite(self, value):
        self.onsite = value

    def add_hoppings(self, delta, from_sub, to_sub, value):
        """Hoppings from every site (t, from_sub) to (t + delta, to_sub)
        that is in the system; ``value(pos_to, pos_from, **params)`` is the
        block <to|H|from>, and its conjugate transpose is the reverse."""
        delta = np.asarray(delta, dtype=np.int64)
        src = np.nonzero(self.sublattice == from_sub)[0]
        target = self.tags[src] + delta
        inside = np.all((target >= self._lo) & (target < self._hi), axis=1)
        src, target = src[inside], target[inside]
        dst = self._grid[tuple((target - self._lo).T) + (np.full(len(src), to_sub),)]
        ok = dst >= 0
        src, dst = src[ok], dst[ok]
        if np.any(src == dst):
            raise ValueError("a hopping may not connect a site to itself")
        self._hops.append((dst, src, value))
        self._graph = None
        self._lookup = None

    # -- what kwant_amd reads from a finalized system ----------------------------
    @property
    def graph(self):
        if self._graph is None:
            if self._hops:
                to = np.concatenate([d for d, _, _ in self._hops])
                fr = np.concatenate([s for _, s, _ in self._hops])
            else:
                to = fr = np.zeros(0, dtype=np.int64)
            graph = _Graph(len(self._pos), np.concatenate([fr, to]), np.concatenate([to, fr]))
            if np.any((graph.tails[1:] == graph.tails[:-1]) &
                      (graph.heads[1:] == graph.heads[:-1])):
                raise ValueError("a hopping is defined twice")
            self._graph = graph
        return self._graph

    def hamiltonian(self, i, j, *args, params=None):
        """The block <i|H|j> (one site or one hopping, evaluated alone)."""
        if i == j:
            return _blocks(_call(self.onsite, (self._pos[[i]],), params), 1, self.norbs)[0]
        if self._lookup is None:
            self.graph  # noqa: B018  (checks for hoppings defined twice)
            self._lookup = {}
            for kind, (to, fr, _) in enumerate(self._hops):
                for a, b in zip(to.tolist(), fr.tolist()):
                    self._lookup[(a, b)] = kind
        kind = self._lookup.get((i, j))
        if kind is not None:
            value = self._hops[kind][2]
            return _blocks(_call(value, (self._pos[[i]], self._pos[[j]]), params),
                           1, self.norbs)[0]
        kind = self._lookup.get((j, i))
        if kind is not None:
            value = self._hops[kind][2]
            return _blocks(_call(value, (self._pos[[j]], self._pos[[i]]), params),
                           1, self.norbs)[0].conj().T
        raise ValueError("(%d, %d) is not a hopping of the system" % (i, j))

    def pos(self, i):
        """Real-space position of site i (kwant's System.pos)."""
        return self._pos[i]

    @property
    def pos_array(self):
        return self._pos

    def hamiltonian_submatrix(self, args=(), sparse=True, return_norb=False, params=None):
        """The Hamiltonian as a CSR matrix (dense with sparse=False): every
        value function is called once, over all sites or all hoppings of a
        kind.  return_norb adds the orbitals per site, twice, as kwant does."""
        if args:
            raise TypeError("pass parameters by name (params=...)")
        n, no = len(self._pos), self.norbs
        self.graph  # noqa: B018  (checks for hoppings defined twice)
        r0 = np.arange(no)
        rows, cols, vals = [], [], []
        onsite = _blocks(_call(self.onsite, (self._pos,), params), n, no)
        ii, jj = np.meshgrid(r0, r0, indexing='ij')
        base = (np.arange(n) * no)[:, None, None]
        rows.append((base + ii).ravel())
        cols.append((base + jj).ravel())
        vals.append(onsite.ravel())
        for to, fr, value in self._hops:
            h = _blocks(_call(value, (self._pos[to], self._pos[fr]), params), len(to), no)
            bt = (to * no)[:, None, None]
            bf = (fr * no)[:, None, None]
            rows += [(bt + ii).ravel(), (bf + jj).ravel()]
            cols += [(bf + jj).ravel(), (bt + ii).ravel()]
            vals += [h.ravel(), h.conj().ravel()]
        data = np.concatenate(vals)
        keep = data != 0
        coo = scipy.sparse.coo_matrix(
            (data[keep], (np.concatenate(rows)[keep], np.concatenate(cols)[keep])),
            shape=(n * no, n * no))
        out = coo.tocsr() if sparse else coo.toarray()
        if return_norb:
            norbs = np.full(n, no, dtype=np.int64)
            return out, norbs, norbs
        return out


def honeycomb_system(cells, t=1.0, onsite=0.0):
    """Graphene as a LatticeSystem: cells x cells unit cells of two sites,
    nearest-neighbour hopping -t (the lattice of kwant_amd.lattices.honeycomb)."""
    syst = LatticeSystem([(1, 0), (0.5, np.sqrt(3) / 2)], [(0, 0), (0, 1 / np.sqrt(3))],
                         [(0, cells), (0, cells)])
    syst.set_onsite(onsite)
    for delta in ((0, 0), (1, -1), (0, -1)):
        syst.add_hoppings(delta, 0, 1, -t)
    return syst

