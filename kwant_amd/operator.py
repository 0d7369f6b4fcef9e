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


class Density:
    """Density operator Q_iab = M_i on the orbitals of site i
    (kwant/operator.pyx:723-817).

    Parameters as kwant.operator.Density: ``syst``, ``onsite`` (scalar,
    square matrix, family -> matrix mapping, or function of (site, *params)),
    ``where`` (site ids, Sites of a finalized builder, or a predicate),
    ``check_hermiticity`` and ``sum``."""

    def __init__(self, syst, onsite=1, where=None, *, check_hermiticity=True, sum=False):
        ranges = _site_ranges(syst)
        self.syst = syst
        self._site_ranges = ranges
        self._num_sites = _num_sites(syst, ranges)
        self.where = _normalize_site_where(syst, where, self._num_sites)
        self.onsite, self._onsite_param_names = _normalize_onsite(
            syst, onsite, check_hermiticity, ranges)
        self.check_hermiticity = check_hermiticity
        self.sum = sum
        self._bound_onsite = None

    @property
    def tot_norbs(self):
        return int(self._site_ranges[-1][2])

    # -- onsite matrices -----------------------------------------------------
    def _eval_onsites(self, args, params):
        """Onsite matrix of every ``where`` entry (kwant/operator.pyx:666-690)."""
        if params:
            missing = [p for p in self._onsite_param_names if p not in params]
            if missing:
                raise TypeError('Operator is missing required arguments: ' +
                                ', '.join(map('"{}"'.format, missing)))
            args = tuple(params[p] for p in self._onsite_param_names)
        out = []
        for site in self.where[:, 0]:
            _, norbs = site_orbitals(self._site_ranges, int(site))
            mat = _as_matrix(self.onsite(int(site), *args))
            _check_onsite(mat, norbs, self.check_hermiticity)
            out.append(mat)
        return out

    def _site_blocks(self, args=(), params=None):
        """(first orbital, norbs, M) of every ``where`` entry, each M read as
        kwant/operator.pyx:771-801 reads it: entry (i, j) of a site with
        norbs orbitals is element i * norbs + j of the flattened onsite."""
        if callable(self.onsite):
            mats = self._bound_onsite if self._bound_onsite is not None else \
                self._eval_onsites(args, params)
        else:
            mats = None
            flat = np.asarray(self.onsite, dtype=complex).reshape(-1)
        blocks = []
        for w, site in enumerate(self.where[:, 0]):
            start, norbs = site_orbitals(self._site_ranges, int(site))
            if mats is not None:
                mat = np.asarray(mats[w], dtype=complex).reshape(-1)[:norbs * norbs]
            else:
                mat = flat[:norbs * norbs]
            blocks.append((start, norbs, mat.reshape(norbs, norbs)))
        return blocks

    def _check_args(self, args, params):
        if self._bound_onsite is not None and (args or params):
            raise ValueError("Extra arguments are already bound to this operator. You "
                             "should call this operator providing neither 'args' nor "
                             "'params'.")
        if args and params:
            raise TypeError("'args' and 'params' are mutually exclusive.")

    # -- kwant's operator interface -----------------------------------------
    def __call__(self, bra, ket=None, args=(), *, params=None):
        """q_w = sum_ab conj(bra_a) M_ab ket_b per ``where`` site, or their
        sum (kwant/operator.pyx:470-560); real when check_hermiticity and
        bra and ket are one array."""
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
        result = np.zeros(self.where.shape[0], dtype=complex)
        for w, (start, norbs, mat) in enumerate(self._site_blocks(args, params)):
            sl = slice(start, start + norbs)
            result[w] = np.vdot(bra[sl], mat @ ket[sl])
        if self.check_hermiticity and same:
            result = result.real
        return np.sum(result) if self.sum else result

    def act(self, ket, args=(), *, params=None):
        """sum_ib Q_iab ket_b (kwant/operator.pyx:562-606)."""
        self._check_args(args, params)
        if ket is None:
            raise TypeError('ket must be an array')
        ket = np.asarray(ket, dtype=complex)
        if ket.shape != (self.tot_norbs,):
            raise ValueError('ket vector is incorrect shape')
        out = np.zeros(self.tot_norbs, dtype=complex)
        for start, norbs, mat in self._site_blocks(args, params):
            sl = slice(start, start + norbs)
            out[sl] += mat @ ket[sl]
        return out

    def bind(self, args=(), *, params=None):
        """A copy with the onsite function evaluated (kwant/operator.pyx:608-637)."""
        if args and params:
            raise TypeError("'args' and 'params' are mutually exclusive.")
        q = self.__class__.__new__(self.__class__)
        q.__dict__.update(self.__dict__)
        if callable(self.onsite):
            q._bound_onsite = self._eval_onsites(args, params)
        return q

    def tocoo(self, args=(), *, params=None):
        """The operator as a tot_norbs x tot_norbs COO matrix, duplicates of
        ``where`` entries kept as separate entries."""
        self._check_args(args, params)
        rows, cols, data = [], [], []
        for start, norbs, mat in self._site_blocks(args, params):
            r, c = np.meshgrid(np.arange(norbs), np.arange(norbs), indexing='ij')
            rows.append((start + r).ravel())
            cols.append((start + c).ravel())
            data.append(mat.ravel())
        n = self.tot_norbs
        if not data:
            return scipy.sparse.coo_matrix((n, n), dtype=complex)
        return scipy.sparse.coo_matrix(
            (np.concatenate(data), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


# ---------------------------------------------------------------------------
# What the GPU path needs

class DensityLayout:
    """A site-resolved density prepared for the recursion kernel.

    ``matrix``: the block-diagonal onsite matrix over the distinct ``where``
    sites (tot_norbs x tot_norbs CSR; w = matrix^H a_0).
    ``orbs``: the orbitals of the ``where`` entries, concatenated in
    ``where`` order (repeats included); ``starts``: where each entry begins
    in ``orbs`` (for numpy.add.reduceat).
    ``sum``, ``real_mu0``: the operator's sum flag, and whether op(a_0, a_0)
    is real (check_hermiticity with bra is ket, kwant/operator.pyx:555-557).
    """

    def __init__(self, blocks, tot_norbs, sum_, check_hermiticity):
        rows, cols, data, orbs, starts = [], [], [], [], []
        seen = set()
        pos = 0
        for start, norbs, mat in blocks:
            starts.append(pos)
            orbs.append(np.arange(start, start + norbs, dtype=np.int64))
            pos += norbs
            if sum_ or start not in seen:
                # sum=True keeps repeated sites (each adds its term, as the
                # reference's sum over where does); sum=False evaluates every
                # repeat from the same rows
                seen.add(start)
                r, c = np.meshgrid(np.arange(norbs), np.arange(norbs), indexing='ij')
                rows.append((start + r).ravel())
                cols.append((start + c).ravel())
                data.append(np.asarray(mat, dtype=complex).ravel())
        n = tot_norbs
        if data:
            coo = scipy.sparse.coo_matrix(
                (np.concatenate(data), (np.concatenate(rows), np.concatenate(cols))),
                shape=(n, n))
            self.matrix = coo.tocsr()   # duplicates summed
        else:
            self.matrix = scipy.sparse.csr_matrix((n, n), dtype=complex)
        self.orbs = np.concatenate(orbs) if orbs else np.zeros(0, dtype=np.int64)
        self.starts = np.asarray(starts, dtype=np.int64)
        self.n_where = len(blocks)
        self.sum = bool(sum_)
        self.real_mu0 = bool(check_hermiticity)

    def site_sums(self, rows):
        """Per-``where`` sums over the last-but-one axis of an array indexed
        by ``orbs`` position: (..., len(orbs), ...) -> (..., n_where, ...)."""
        if self.n_where == 0:
            shape = list(rows.shape)
            shape[1] = 0
            return np.zeros(shape, dtype=rows.dtype)
        return np.add.reduceat(rows, self.starts, axis=1)


def density_layout(op):
    """The DensityLayout of a bound Density -- this module's, or kwant's own
    (kwant.operator.Density: public ``where``, ``_site_ranges``, ``onsite``,
    ``_bound_onsite``) -- or None when ``op`` is some other operator."""
    if isinstance(op, Density):
        return DensityLayout(op._site_blocks(), op.tot_norbs, op.sum, op.check_hermiticity)
    if type(op).__name__ != 'Density' or not all(
            hasattr(op, a) for a in ('where', '_site_ranges', 'onsite', 'sum')):
        return None
    ranges = np.asarray(op._site_ranges, dtype=np.int64)
    where = np.asarray(op.where, dtype=np.int64).reshape(-1)
    bound = getattr(op, '_bound_onsite', None)
    if callable(op.onsite) and bound is None:
        return None
    blocks = []
    for w, site in enumerate(where):
        start, norbs = site_orbitals(ranges, int(site))
        if callable(op.onsite):
            beg = int(np.asarray(bound.data_offsets)[w])
            flat = np.asarray(bound.data)[beg:beg + norbs * norbs]
        else:
            flat = np.asarray(op.onsite, dtype=complex).reshape(-1)[:norbs * norbs]
        blocks.append((start, norbs, np.asarray(flat, dtype=complex).reshape(norbs, norbs)))
    return DensityLayout(blocks, int(ranges[-1][2]), bool(op.sum), bool(op.check_hermiticity))
