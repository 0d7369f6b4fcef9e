"""Site-resolved densities for the KPM engine.

``Density`` is kwant.operator.Density (kwant/operator.pyx:723-817 with the
_LocalOperator plumbing of :418-721): the tensor Q_iab that equals the
onsite matrix M_i on the orbitals of site i and vanishes elsewhere, so that

    op(bra, ket)[w] = sum_ab conj(bra_a) M_{where[w], ab} ket_b

over the sites ``where`` (``sum=True``: the sum over them).  Called directly
it evaluates on the host with numpy, like kwant's operator.  Given to
SpectralDensity, the moments op(a_0, a_n) are formed on the GPU inside the
recursion kernel instead: with w = M^H a_0 (block-diagonal M) the kernel
writes sum_b conj(w_b) a_n,b for every matrix row, and the rows of each site
are summed afterwards (see ``density_layout``).

Systems are duck-typed kwant systems (``site_ranges``, ``graph.num_nodes``,
and ``sites``/``id_by_site`` for finalized builders).  As an extension, a
plain matrix or an integer N stands for N sites of one orbital each.
"""
import collections.abc
import inspect
import numbers

import numpy as np
import scipy.sparse


def _site_ranges(syst):
    """kwant's site_ranges rows (first site, norbs, first orbital), closed by
    (num_sites, 0, tot_norbs) (kwant/system.py:27-34)."""
    if hasattr(syst, 'site_ranges'):
        if syst.site_ranges is not None:
            return np.asarray(syst.site_ranges, dtype=np.int64).reshape(-1, 3)
        raise ValueError('Number of orbitals not defined.\n'
                         'Declare the number of orbitals using the '
                         '`norbs` keyword argument when constructing '
                         'the site families (lattices).')
    n = _matrix_size(syst)
    return np.array([[0, 1, 0], [n, 0, n]], dtype=np.int64)


def _matrix_size(syst):
    if isinstance(syst, numbers.Integral):
        return int(syst)
    shape = getattr(syst, 'shape', None)
    if shape is None or len(shape) != 2 or shape[0] != shape[1]:
        raise TypeError("'syst' is neither a kwant system, a square matrix nor an integer")
    return int(shape[0])


def _num_sites(syst, ranges):
    graph = getattr(syst, 'graph', None)
    if graph is not None:
        return int(graph.num_nodes)
    return int(ranges[-1][0])


def site_orbitals(ranges, site):
    """(first orbital, norbs) of ``site`` (kwant/operator.pyx:114-128)."""
    k = int(np.searchsorted(ranges[:, 0], site, side='right')) - 1
    first, norbs, offset = ranges[k]
    return int(offset + (site - first) * norbs), int(norbs)


def _normalize_site_where(syst, where, num_sites):
    """``where`` as an (n, 1) int array of site ids (kwant/operator.pyx:160-199)."""
    if where is None:
        where = list(range(num_sites))
    elif callable(where):
        ids = getattr(syst, 'id_by_site', None)
        sites = getattr(syst, 'sites', None)
        if ids is not None and sites is not None:
            where = [ids[s] for s in filter(where, sites)]
        else:
            where = [s for s in range(num_sites) if where(s)]
    else:
        where = list(where)
        if where and not isinstance(where[0], numbers.Integral):
            ids = getattr(syst, 'id_by_site', None)
            if ids is None:
                raise TypeError("'where' contains Sites, but the system is not "
                                "a finalized Builder.")
            where = [ids[s] for s in where]
    where = np.asarray(where, dtype=np.int64).reshape(-1, 1)
    if np.any((where < 0) | (where >= num_sites)):
        raise ValueError('`where` contains sites that are not in the system.')
    return where


def _is_herm_conj(a, b, atol=1e-300, rtol=1e-13):
    """kwant/operator.pyx:50-74, including its comparison of the squared
    difference with the unsquared tolerance."""
    a = np.asarray(a, dtype=complex)
    b = np.asarray(b, dtype=complex)
    max_a = np.sqrt(np.max(np.abs(a) ** 2)) if a.size else 0.0
    tol = rtol * max_a + atol
    diff = a - b.T.conj()
    return not np.any(diff.real ** 2 + diff.imag ** 2 > tol)


def _check_onsite(mat, norbs, check_hermiticity):
    """kwant/operator.pyx:85-94."""
    if mat.shape[0] != mat.shape[1]:
        raise ValueError('Onsite matrix is not square')
    if mat.shape[0] != norbs:
        raise ValueError('Onsite matrix dimensions do not match the declared '
                         'number of orbitals')
    if check_hermiticity and not _is_herm_conj(mat, mat):
        raise ValueError('Onsite matrix is not hermitian, use the option '
                         '`check_hermiticity=True` if this is intentional.')


def _as_matrix(value):
    mat = np.asarray(value, dtype=complex)
    if mat.ndim == 0:
        mat = mat.reshape(1, 1)
    if mat.ndim != 2:
        raise ValueError('Onsite matrix is not square')
    return mat


def _parameters(fn):
    """Parameter names of a value function (kwant/_common.py get_parameters)."""
    sig = inspect.signature(fn)
    return tuple(p.name for p in sig.parameters.values()
                 if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD, p.KEYWORD_ONLY))


def _normalize_onsite(syst, onsite, check_hermiticity, ranges):
    """A callable of (site id, *args), or a single matrix
    (kwant/operator.pyx:273-321)."""
    if callable(onsite):
        names = _parameters(onsite)[1:]
        sites = getattr(syst, 'sites', None)
        if sites is not None:
            return (lambda site_id, *args: onsite(sites[site_id], *args)), names
        return onsite, names
    if isinstance(onsite, collections.abc.Mapping):
        sites = getattr(syst, 'sites', None)
        if sites is None:
            raise TypeError('Provide `onsite` as a value or a function for '
                            'systems that are not finalized Builders.')
        for fam, value in onsite.items():
            _check_onsite(_as_matrix(value), fam.norbs, check_hermiticity)
        return (lambda site_id, *args: onsite[sites[site_id].family]), ()
    mat = _as_matrix(onsite)
    _check_onsite(mat, mat.shape[0], check_hermiticity)
    norbs = ranges[:-1, 1] if len(ranges) > 1 else ranges[:, 1]
    if mat.shape[0] == 1:
        # one value for every site: value * identity(max norbs), which
        # _operate then reads with each site's own norbs as the row stride
        mat = mat[0, 0] * np.eye(int(np.max(ranges[:, 1])), dtype=complex)
    elif len(set(norbs.tolist())) == 1:
        if mat.shape[0] != norbs[0]:
            raise ValueError('Single `onsite` matrix of shape ({0}, {0}) provided but there '
                             'are {1} orbitals per site in the system'
                             .format(mat.shape[0], norbs[0]))
    else:
        raise ValueError('Single `onsite` matrix provided, but there are different '
                         'numbers of orbitals on different sites')
    return mat, ()


def _normalize_hopping_where(syst, where, ranges):
    """``where`` as an (n, 2) int array of hoppings (kwant/operator.pyx:202-253):
    every hopping of the system graph when None.  Systems without a graph
    (plain matrices) take the off-diagonal entries of their Hamiltonian."""
    graph = getattr(syst, 'graph', None)
    if graph is not None and hasattr(graph, '__iter__'):
        edges = list(graph)
        has_edge = graph.has_edge if hasattr(graph, 'has_edge') else \
            (lambda a, b, _e=set(map(tuple, edges)): (a, b) in _e)
    else:
        edges = site_edges(_hamiltonian_csr(syst, None), ranges)
        eset = set(edges)
        has_edge = lambda a, b: (a, b) in eset  # noqa: E731
    if where is None:
        where = edges
    elif callable(where):
        sites = getattr(syst, 'sites', None)
        if sites is not None:
            where = [e for e in edges if where(sites[e[0]], sites[e[1]])]
        else:
            where = [e for e in edges if where(*e)]
    else:
        where = list(where)
        if where and not isinstance(where[0][0], numbers.Integral):
            ids = getattr(syst, 'id_by_site', None)
            if ids is None:
                raise TypeError("'where' contains Sites, but the system is not "
                                "a finalized Builder.")
            where = [(ids[a], ids[b]) for a, b in where]
        if any(not has_edge(int(a), int(b)) for a, b in where):
            raise ValueError('`where` contains hoppings that are not in the system.')
    return np.asarray(where, dtype=np.int64).reshape(-1, 2)


def orbital_sites(ranges):
    """The site of every orbital."""
    out = []
    for (first, norbs, _), (nxt, _, _) in zip(ranges[:-1], ranges[1:]):
        out.append(np.repeat(np.arange(first, nxt, dtype=np.int64), norbs))
    return np.concatenate(out) if out else np.zeros(0, dtype=np.int64)


def site_edges(hamiltonian, ranges):
    """The hoppings (a, b), a != b, of a Hamiltonian matrix at site level,
    ordered by (a, b): the graph of a system given only as a matrix."""
    h = scipy.sparse.csr_matrix(hamiltonian)
    site = orbital_sites(ranges)
    rows = site[np.repeat(np.arange(h.shape[0]), np.diff(h.indptr))]
    cols = site[h.indices]
    keep = rows != cols
    pairs = np.unique(np.stack([rows[keep], cols[keep]], axis=1), axis=0)
    return [tuple(p) for p in pairs.tolist()]


def _hamiltonian_csr(syst, params):
    if hasattr(syst, 'hamiltonian_submatrix'):
        return syst.hamiltonian_submatrix(params=params, sparse=True)
    return syst


class _LocalOperator:
    """Plumbing shared by Density, Current and Source
    (kwant/operator.pyx:418-721)."""

    def __init__(self, syst, onsite, where, *, check_hermiticity=True, sum=False):
        ranges = _site_ranges(syst)
        self.syst = syst
        self._site_ranges = ranges
        self._num_sites = _num_sites(syst, ranges)
        self.where = where(ranges) if callable(where) else where
        self.onsite, self._onsite_param_names = _normalize_onsite(
            syst, onsite, check_hermiticity, ranges)
        self.check_hermiticity = check_hermiticity
        self.sum = sum
        self._bound_onsite = None
        self._bound_hamiltonian = None

    @property
    def tot_norbs(self):
        return int(self._site_ranges[-1][2])

    def _orbs(self, site):
        return site_orbitals(self._site_ranges, int(site))

    # -- onsite and Hamiltonian blocks ---------------------------------------
    def _eval_onsites(self, args, params):
        """Onsite matrix at the first site of every ``where`` entry
        (kwant/operator.pyx:666-690)."""
        if params:
            missing = [p for p in self._onsite_param_names if p not in params]
            if missing:
                raise TypeError('Operator is missing required arguments: ' +
                                ', '.join(map('"{}"'.format, missing)))
            args = tuple(params[p] for p in self._onsite_param_names)
        out = []
        for site in self.where[:, 0]:
            _, norbs = self._orbs(site)
            mat = _as_matrix(self.onsite(int(site), *args))
            _check_onsite(mat, norbs, self.check_hermiticity)
            out.append(mat)
        return out

    def _onsite_blocks(self, args=(), params=None):
        """(first orbital, norbs, M) at the first site of every ``where``
        entry, M read as kwant's _operate loops read it: entry (i, j) of a
        site with norbs orbitals is element i * norbs + j of the flattened
        onsite (kwant/operator.pyx:791-797, 976-986, 1100-1114)."""
        if callable(self.onsite):
            mats = self._bound_onsite if self._bound_onsite is not None else \
                self._eval_onsites(args, params)
        else:
            mats = None
            flat = np.asarray(self.onsite, dtype=complex).reshape(-1)
        blocks = []
        for w, site in enumerate(self.where[:, 0]):
            start, norbs = self._orbs(site)
            m = np.asarray(mats[w], dtype=complex).reshape(-1) if mats is not None else flat
            blocks.append((start, norbs, m[:norbs * norbs].reshape(norbs, norbs)))
        return blocks

    def _eval_hamiltonian(self, args, params):
        """H_ab of every ``where`` entry (b = a for sites), from the system's
        ``hamiltonian(a, b, *args, params=...)`` when it has one, else from
        its Hamiltonian matrix (kwant/operator.pyx:692-705, 97-108)."""
        hop = self.where.shape[1] == 2
        fn = getattr(self.syst, 'hamiltonian', None)
        csr = None if callable(fn) else scipy.sparse.csr_matrix(
            _hamiltonian_csr(self.syst, params))
        out = []
        for row in self.where:
            a, b = int(row[0]), int(row[1] if hop else row[0])
            (a_s, a_n), (b_s, b_n) = self._orbs(a), self._orbs(b)
            if csr is None:
                mat = _as_matrix(fn(a, b, *args, params=params))
                partner = mat if a == b else _as_matrix(fn(b, a, *args, params=params))
            else:
                mat = csr[a_s:a_s + a_n, b_s:b_s + b_n].toarray()
                partner = csr[b_s:b_s + b_n, a_s:a_s + a_n].toarray()
            if mat.shape[0] != a_n and mat.shape[1] != b_n:
                raise ValueError('Hamiltonian matrix dimensions do not match the declared '
                                 'number of orbitals')
            if self.check_hermiticity and not _is_herm_conj(partner, mat):
                raise ValueError('Hamiltonian matrix is not hermitian, use the option '
                                 '`check_hermiticity=True` if this is intentional.')
            out.append(mat)
        return out

    def _hamiltonian_blocks(self, args, params):
        if self._bound_hamiltonian is not None:
            return self._bound_hamiltonian
        return self._eval_hamiltonian(args, params)

    def _check_args(self, args, params):
        if (self._bound_onsite is not None or self._bound_hamiltonian is not None) and \
                (args or params):
            raise ValueError("Extra arguments are already bound to this operator. You "
                             "should call this operator providing neither 'args' nor "
                             "'params'.")
        if args and params:
            raise TypeError("'args' and 'params' are mutually exclusive.")

    # -- kwant's operator interface -----------------------------------------
    def __call__(self, bra, ket=None, args=(), *, params=None):
        """The matrix elements of every ``where`` entry, or their sum
        (kwant/operator.pyx:470-560); real when check_hermiticity and bra
        and ket are one array."""
        self._check_args(args, params)
        if bra is None:
            raise TypeError('bra must be an array')
        bra = np.asarray(bra, dtype=complex)
        ket = bra if ket is None else np.asarray(ket, dtype=complex)
        # op(v, v) with a complex array v passes one object twice
        same = bra is ket
        if bra.shape != (self.tot_norbs,):
            raise ValueError('vector is incorrect shape')
        if ket.shape != (self.tot_norbs,):
            raise ValueError('ket vector is incorrect shape')
        result = self._mat_els(bra, ket, args, params)
        if self.check_hermiticity and same:
            result = result.real
        return np.sum(result) if self.sum else result

    def act(self, ket, args=(), *, params=None):
        """sum_i Q_i ket (kwant/operator.pyx:562-606)."""
        self._check_args(args, params)
        if ket is None:
            raise TypeError('ket must be an array')
        ket = np.asarray(ket, dtype=complex)
        if ket.shape != (self.tot_norbs,):
            raise ValueError('ket vector is incorrect shape')
        return self.tocoo(args, params=params).tocsr() @ ket

    def bind(self, args=(), *, params=None):
        """A copy with the onsite (and Hamiltonian) blocks evaluated
        (kwant/operator.pyx:608-637, 925-938)."""
        if args and params:
            raise TypeError("'args' and 'params' are mutually exclusive.")
        q = self.__class__.__new__(self.__class__)
        q.__dict__.update(self.__dict__)
        if callable(self.onsite):
            q._bound_onsite = self._eval_onsites(args, params)
        if self._needs_hamiltonian:
            q._bound_hamiltonian = self._eval_hamiltonian(args, params)
        return q

    def tocoo(self, args=(), *, params=None):
        """The operator's action sum_i Q_i as a tot_norbs x tot_norbs COO
        matrix (repeated ``where`` entries kept as separate entries)."""
        self._check_args(args, params)
        rows, cols, data = [], [], []
        for r0, c0, mat in self._act_blocks(args, params):
            r, c = np.meshgrid(np.arange(mat.shape[0]), np.arange(mat.shape[1]), indexing='ij')
            rows.append((r0 + r).ravel())
            cols.append((c0 + c).ravel())
            data.append(mat.ravel())
        n = self.tot_norbs
        if not data:
            return scipy.sparse.coo_matrix((n, n), dtype=complex)
        return scipy.sparse.coo_matrix(
            (np.concatenate(data), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


class _SiteOperator(_LocalOperator):
    """Operators that are one square block Q_a per ``where`` site."""

    def _site_blocks(self, args=(), params=None):
        raise NotImplementedError

    def _mat_els(self, bra, ket, args, params):
        result = np.zeros(self.where.shape[0], dtype=complex)
        for w, (start, norbs, q) in enumerate(self._site_blocks(args, params)):
            sl = slice(start, start + norbs)
            result[w] = np.vdot(bra[sl], q @ ket[sl])
        return result

    def _act_blocks(self, args, params):
        return [(s, s, q) for s, _, q in self._site_blocks(args, params)]


class Density(_SiteOperator):
    """Density operator Q_iab = M_i on the orbitals of site i
    (kwant/operator.pyx:723-817).

    Parameters as kwant.operator.Density: ``syst``, ``onsite`` (scalar,
    square matrix, family -> matrix mapping, or function of (site, *params)),
    ``where`` (site ids, Sites of a finalized builder, or a predicate),
    ``check_hermiticity`` and ``sum``."""

    _needs_hamiltonian = False

    def __init__(self, syst, onsite=1, where=None, *, check_hermiticity=True, sum=False):
        super().__init__(syst, onsite,
                         lambda ranges: _normalize_site_where(syst, where,
                                                              _num_sites(syst, ranges)),
                         check_hermiticity=check_hermiticity, sum=sum)

    def _site_blocks(self, args=(), params=None):
        return self._onsite_blocks(args, params)


class Source(_SiteOperator):
    """Source operator Q_iab = i (H_ii^H M_i - M_i H_ii) on the orbitals of
    site i (kwant/operator.pyx:1003-1126), H_ii the onsite Hamiltonian."""

    _needs_hamiltonian = True

    def __init__(self, syst, onsite=1, where=None, *, check_hermiticity=True, sum=False):
        super().__init__(syst, onsite,
                         lambda ranges: _normalize_site_where(syst, where,
                                                              _num_sites(syst, ranges)),
                         check_hermiticity=check_hermiticity, sum=sum)

    def _site_blocks(self, args=(), params=None):
        hams = self._hamiltonian_blocks(args, params)
        out = []
        for (start, norbs, m), h in zip(self._onsite_blocks(args, params), hams):
            h = np.asarray(h, dtype=complex).reshape(-1)[:norbs * norbs].reshape(norbs, norbs)
            out.append((start, norbs, 1j * (h.conj().T @ m - m @ h)))
        return out


class Current(_LocalOperator):
    """Current operator on the hoppings (a, b) of ``where``:
    i [(H_ab)^H M_a - M_a H_ab] between the orbitals of b and a
    (kwant/operator.pyx:871-1001), so that

        op(bra, ket)[w] = i (bra_b^H H_ab^H M_a ket_a - bra_a^H M_a H_ab ket_b).
    """

    _needs_hamiltonian = True

    def __init__(self, syst, onsite=1, where=None, *, check_hermiticity=True, sum=False):
        super().__init__(syst, onsite,
                         lambda ranges: _normalize_hopping_where(syst, where, ranges),
                         check_hermiticity=check_hermiticity, sum=sum)

    def _hop_blocks(self, args=(), params=None):
        """(a_start, a_norbs, b_start, b_norbs, M_a, H_ab) per ``where``
        entry; H_ab read flattened with stride b_norbs as in
        kwant/operator.pyx:976-986."""
        hams = self._hamiltonian_blocks(args, params)
        out = []
        for (a_s, a_n, m), (a, b), h in zip(self._onsite_blocks(args, params), self.where, hams):
            b_s, b_n = self._orbs(b)
            h = np.asarray(h, dtype=complex).reshape(-1)[:a_n * b_n].reshape(a_n, b_n)
            out.append((a_s, a_n, b_s, b_n, m, h))
        return out

    def _mat_els(self, bra, ket, args, params):
        result = np.zeros(self.where.shape[0], dtype=complex)
        for w, (a_s, a_n, b_s, b_n, m, h) in enumerate(self._hop_blocks(args, params)):
            ba, bb = bra[a_s:a_s + a_n], bra[b_s:b_s + b_n]
            ka, kb = ket[a_s:a_s + a_n], ket[b_s:b_s + b_n]
            result[w] = 1j * (np.vdot(h @ bb, m @ ka) - np.vdot(ba, m @ (h @ kb)))
        return result

    def _act_blocks(self, args, params):
        out = []
        for a_s, a_n, b_s, b_n, m, h in self._hop_blocks(args, params):
            out.append((b_s, a_s, 1j * (h.conj().T @ m)))
            out.append((a_s, b_s, -1j * (m @ h)))
        return out


# ---------------------------------------------------------------------------
# What the GPU path needs

def _block_diag(blocks, n, keep_repeats):
    """CSR of square blocks (start, norbs, Q) on the diagonal; repeated
    starts are summed when ``keep_repeats``, else the first one is used."""
    rows, cols, data = [], [], []
    seen = set()
    for start, norbs, q in blocks:
        if not keep_repeats and start in seen:
            continue
        seen.add(start)
        r, c = np.meshgrid(np.arange(norbs), np.arange(norbs), indexing='ij')
        rows.append((start + r).ravel())
        cols.append((start + c).ravel())
        data.append(np.asarray(q, dtype=complex).ravel())
    if not data:
        return scipy.sparse.csr_matrix((n, n), dtype=complex)
    return scipy.sparse.coo_matrix(
        (np.concatenate(data), (np.concatenate(rows), np.concatenate(cols))),
        shape=(n, n)).tocsr()


class SiteLayout:
    """A site operator (Density, Source) prepared for the recursion kernel.

    ``matrix``: the block-diagonal Q over the distinct ``where`` sites
    (tot_norbs x tot_norbs CSR; w = matrix^H a_0).  ``act_matrix``: sum_w Q_w
    with repeated sites counted each time (the operator's ``act``, and its
    sum=True form).  ``orbs``: the orbitals of the ``where`` entries,
    concatenated in ``where`` order (repeats included); ``starts``: where
    each entry begins in ``orbs`` (for numpy.add.reduceat).  ``sum``,
    ``real_mu0``: the sum flag, and whether op(a_0, a_0) is real
    (check_hermiticity with bra is ket, kwant/operator.pyx:555-557).
    """

    def __init__(self, blocks, tot_norbs, sum_, check_hermiticity):
        orbs, starts, pos = [], [], 0
        for start, norbs, _ in blocks:
            starts.append(pos)
            orbs.append(np.arange(start, start + norbs, dtype=np.int64))
            pos += norbs
        self.matrix = _block_diag(blocks, tot_norbs, keep_repeats=False)
        self.act_matrix = _block_diag(blocks, tot_norbs, keep_repeats=True)
        self.orbs = np.concatenate(orbs) if orbs else np.zeros(0, dtype=np.int64)
        self.starts = np.asarray(starts, dtype=np.int64)
        self.n_where = len(blocks)
        self.sum = bool(sum_)
        self.real_mu0 = bool(check_hermiticity)

    def site_sums(self, rows):
        """Per-``where`` sums over axis 1 of an array indexed by ``orbs``
        position: (k, len(orbs), ...) -> (k, n_where, ...)."""
        if self.n_where == 0:
            shape = list(rows.shape)
            shape[1] = 0
            return np.zeros(shape, dtype=rows.dtype)
        return np.add.reduceat(rows, self.starts, axis=1)



class BondLayout:
    """A Current prepared for the bond-record kernel.

    The distinct hoppings of ``where`` are expanded into entries (r, s, h):
    r runs over the orbitals of a, s over those of b, h = H_ab[r, s].
    ``rows``, ``cols``, ``values``: the entries, hopping by hopping;
    ``starts``: where each distinct hopping begins (for numpy.add.reduceat);
    ``where_index``: the distinct hopping of every ``where`` entry.
    ``onsite``: block-diagonal M over the hoppings' first sites, None when
    every M_a is the identity (then v = ket, w = bra).  ``act_matrix``: the
    operator's action (repeated hoppings counted each time).
    """

    def __init__(self, hops, tot_norbs, sum_, check_hermiticity):
        key_index, uniq = {}, []
        self.where_index = np.zeros(len(hops), dtype=np.int64)
        for w, hop in enumerate(hops):
            key = (hop[0], hop[2])
            if key not in key_index:
                key_index[key] = len(uniq)
                uniq.append(hop)
            self.where_index[w] = key_index[key]
        rows, cols, vals, starts, pos = [], [], [], [], 0
        for a_s, a_n, b_s, b_n, m, h in uniq:
            r, c = np.meshgrid(np.arange(a_n), np.arange(b_n), indexing='ij')
            rows.append((a_s + r).ravel())
            cols.append((b_s + c).ravel())
            vals.append(np.asarray(h, dtype=complex).ravel())
            starts.append(pos)
            pos += a_n * b_n
        cat = (lambda x, dt: np.concatenate(x) if x else np.zeros(0, dt))
        self.rows = cat(rows, np.int64)
        self.cols = cat(cols, np.int64)
        self.values = cat(vals, complex)
        self.starts = np.asarray(starts, dtype=np.int64)
        onsite = [(a_s, a_n, m) for a_s, a_n, _, _, m, _ in uniq]
        identity = all(np.array_equal(m, np.eye(a_n)) for _, a_n, m in onsite)
        self.onsite = None if identity else _block_diag(onsite, tot_norbs, keep_repeats=False)
        act = []
        for a_s, a_n, b_s, b_n, m, h in hops:
            act.append((b_s, a_s, 1j * (h.conj().T @ m)))
            act.append((a_s, b_s, -1j * (m @ h)))
        n = tot_norbs
        if act:
            self.act_matrix = scipy.sparse.coo_matrix(
                (np.concatenate([q.ravel() for _, _, q in act]),
                 (np.concatenate([(r0 + np.repeat(np.arange(q.shape[0]), q.shape[1]))
                                  for r0, _, q in act]),
                  np.concatenate([(c0 + np.tile(np.arange(q.shape[1]), q.shape[0]))
                                  for _, c0, q in act]))), shape=(n, n)).tocsr()
        else:
            self.act_matrix = scipy.sparse.csr_matrix((n, n), dtype=complex)
        self.matrix = self.act_matrix
        self.n_where = len(hops)
        self.sum = bool(sum_)
        self.real_mu0 = bool(check_hermiticity)

    def hop_sums(self, entries):
        """(k, n_entries, ...) entry records -> (k, n_where, ...) records of
        the ``where`` hoppings."""
        if self.n_where == 0:
            shape = list(entries.shape)
            shape[1] = 0
            return np.zeros(shape, dtype=entries.dtype)
        return np.add.reduceat(entries, self.starts, axis=1)[:, self.where_index]


def _kwant_blocks(bsm, w, shape):
    """Block w of a kwant BlockSparseMatrix (public data, data_offsets),
    flattened and cut to ``shape``."""
    beg = int(np.asarray(bsm.data_offsets)[w])
    size = shape[0] * shape[1]
    return np.asarray(bsm.data, dtype=complex)[beg:beg + size].reshape(shape)


def operator_layout(op):
    """The SiteLayout / BondLayout of a bound Density, Source or Current --
    this module's, or kwant's own (read through its public ``where``,
    ``_site_ranges``, ``onsite``, ``_bound_onsite``, ``_bound_hamiltonian``)
    -- or None for any other operator."""
    if isinstance(op, _SiteOperator):
        return SiteLayout(op._site_blocks(), op.tot_norbs, op.sum, op.check_hermiticity)
    if isinstance(op, Current):
        return BondLayout(op._hop_blocks(), op.tot_norbs, op.sum, op.check_hermiticity)
    kind = type(op).__name__
    if kind not in ('Density', 'Source', 'Current') or not all(
            hasattr(op, a) for a in ('where', '_site_ranges', 'onsite', 'sum')):
        return None
    ranges = np.asarray(op._site_ranges, dtype=np.int64)
    where = np.asarray(op.where, dtype=np.int64)
    where = where.reshape(len(where), -1)
    bound = getattr(op, '_bound_onsite', None)
    hams = getattr(op, '_bound_hamiltonian', None)
    if (callable(op.onsite) and bound is None) or (kind != 'Density' and hams is None):
        return None
    tot = int(ranges[-1][2])
    flat = None if callable(op.onsite) else np.asarray(op.onsite, dtype=complex).reshape(-1)
    onsites = []
    for w, site in enumerate(where[:, 0]):
        start, norbs = site_orbitals(ranges, int(site))
        m = flat[:norbs * norbs] if flat is not None else \
            np.asarray(bound.data, dtype=complex)[
                int(np.asarray(bound.data_offsets)[w]):][:norbs * norbs]
        onsites.append((start, norbs, m.reshape(norbs, norbs)))
    if kind == 'Density':
        return SiteLayout(onsites, tot, bool(op.sum), bool(op.check_hermiticity))
    if kind == 'Source':
        blocks = []
        for w, (start, norbs, m) in enumerate(onsites):
            h = _kwant_blocks(hams, w, (norbs, norbs))
            blocks.append((start, norbs, 1j * (h.conj().T @ m - m @ h)))
        return SiteLayout(blocks, tot, bool(op.sum), bool(op.check_hermiticity))
    hops = []
    for w, (a_s, a_n, m) in enumerate(onsites):
        b_s, b_n = site_orbitals(ranges, int(where[w, 1]))
        hops.append((a_s, a_n, b_s, b_n, m, _kwant_blocks(hams, w, (a_n, b_n))))
    return BondLayout(hops, tot, bool(op.sum), bool(op.check_hermiticity))


def density_layout(op):
    """operator_layout restricted to site operators (Density, Source)."""
    lay = operator_layout(op)
    return lay if isinstance(lay, SiteLayout) else None
