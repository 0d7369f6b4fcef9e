"""Drop-in ``kwant.kpm`` whose Chebyshev moment engine runs on an MI355X.

Same public surface as kwant/kpm.py:32-34 -- ``SpectralDensity``,
``Correlator``, ``conductivity``, ``RandomVectors``, ``LocalVectors``,
``jackson_kernel``, ``lorentz_kernel``, ``fermi_distribution`` -- with the
same arguments, attributes, resume methods (``add_moments``/``add_vectors``)
and exceptions.  What differs is where the work runs:

* the Hamiltonian is uploaded once as an int32 CSR (complex128 by default,
  complex64 with ``precision='complex64'``) and the rescaling (H - b)/a is
  applied inside the recursion kernel (kwant/kpm.py:1182-1183);
* random vectors are drawn on the host with the reference's exact
  ``RandomState`` consumption and processed in blocks of up to 64 vectors by
  one fused SpMM recursion kernel per Chebyshev step that also accumulates the
  FP64 dot products of the doubling identity (kwant/kpm.py:436-477);
* the Correlator expansions keep their Chebyshev vectors on the GPU and the
  moment matrix is a GEMM on the matrix cores (kwant/kpm.py:685-695);
* the moment-to-energy transforms (kernel, DCT-III, chebval, simpson) stay in
  FP64 on the host, exactly as kwant/kpm.py:378-388, 1197-1222 do.

Extra keyword arguments (all optional): ``precision`` ('complex128' |
'complex64'), ``max_block_width`` (power of two <= 64), ``variant`` (kernel
launch shape, see kwant_amd.autotune), ``reorder`` (True | False | 'auto' |
a permutation: the device row order, see kwant_amd.ordering; the moments do
not depend on it), ``matrix_format`` ('csr' | 'bsr': the recursion reads the
Hamiltonian as 4 x 4 site blocks -- for 4-orbital systems, operator=None; the
moments are bitwise the CSR ones), ``tiles`` (a kwant_amd.ordering.Tiling, or
True to tile a kwant system by its site positions: the device row order groups
the rows into compact tiles and the plain DOS path runs two Chebyshev steps per
launch; the moments are bitwise those without tiles.  The default 'auto' tiles
kwant systems with positions and at least 2^18 orbitals on the plain DOS path
with the default row order and random vectors, when orbitals x vectors x
num_moments / 2 reaches 2^32, and each block keeps whichever of the two-step and
one-step launches its timing picks; 'graph' tiles any matrix by hop distances
in its graph, about 1 s per 10^6 rows of host set-up), ``gemm_precision`` for the Correlator
('fp64' | 'fp32' | 'isliced4' | 'isliced5' | 'sliced3' | 'sliced4': the 'isliced'
modes are the split-precision GEMM on the INT8 matrix cores, 'sliced' on BF16).
"""
import math
import numbers
from collections.abc import Iterable

import numpy as np
import scipy.fftpack
import scipy.sparse
from numpy.polynomial.chebyshev import chebval
from scipy.integrate import simpson
from scipy.sparse import coo_matrix, csr_matrix
from scipy.sparse.linalg import LinearOperator

from . import _engine as eng
from . import autotune
from . import ordering
from .lanczos import lanczos_bounds
from .operator import BondLayout, SiteLayout, operator_layout

__all__ = ['SpectralDensity', 'Correlator', 'conductivity',
           'RandomVectors', 'LocalVectors', 'jackson_kernel', 'lorentz_kernel',
           'fermi_distribution']

SAMPLING = 2  # energies per moment (kwant/kpm.py:36)


# ---------------------------------------------------------------------------
# inputs: systems, operators, random numbers

def _is_system(obj):
    """Duck-typed kwant.system.System (kwant/kpm.py:177)."""
    return hasattr(obj, 'hamiltonian_submatrix') and hasattr(obj, 'site_ranges')


def _is_local_operator(obj):
    """Duck-typed kwant.operator._LocalOperator (bind + act)."""
    return hasattr(obj, 'bind') and hasattr(obj, 'act')


def ensure_rng(rng=None):
    """Seed contract of kwant/_common.py:115-131: None -> numpy's global
    RandomState, an integer -> a new RandomState(seed), else an object with
    the RandomState interface."""
    if rng is None:
        return np.random.mtrand._rand
    if isinstance(rng, numbers.Integral):
        return np.random.RandomState(rng)
    if all(hasattr(rng, a) for a in ('random_sample', 'randn', 'randint', 'choice')):
        return rng
    raise ValueError("Expecting a seed or an object that offers the "
                     "numpy.random.RandomState interface.")


def _as_csr(hamiltonian, params=None):
    if _is_system(hamiltonian):
        hamiltonian = hamiltonian.hamiltonian_submatrix(params=params, sparse=True)
    try:
        return csr_matrix(hamiltonian)
    except Exception:
        raise ValueError("'hamiltonian' is neither a matrix nor a Kwant system.")


# 'auto' tiles kwant systems with site positions from this many orbitals on,
# when the expansion asked for is at least this many orbital x vector x step
# products (the tiling and its timing cost about a second at 2 10^6 orbitals)
TILES_AUTO_ROWS = 1 << 18
TILES_AUTO_WORK = 1 << 32


def _has_positions(syst):
    return _is_system(syst) and (hasattr(syst, 'pos_array') or hasattr(syst, 'sites')
                                 or hasattr(syst, 'pos'))


def _tiling_of(hamiltonian, tiles, tile_rows=256, auto_ok=False, csr=None):
    """The Tiling for ``tiles``: None/False (no tiles), a Tiling, True (a
    kwant system tiled by its site positions, each site's orbitals kept
    together), 'graph' (any matrix, tiled by hop distances in the graph of
    ``csr``, its CSR form) or 'auto' (True for a kwant system with positions
    and at least TILES_AUTO_ROWS orbitals when ``auto_ok``, else no tiles)."""
    if isinstance(tiles, str) and tiles == 'graph':
        return ordering.tile_order(ordering.graph_coordinates(csr), tile_rows)
    if isinstance(tiles, str):
        if tiles != 'auto':
            raise ValueError("tiles must be None, False, True, 'auto', 'graph' or a Tiling")
        if not (auto_ok and _has_positions(hamiltonian)
                and int(np.asarray(hamiltonian.site_ranges)[-1][2]) >= TILES_AUTO_ROWS):
            return None
        tiles = True
    if tiles is None or tiles is False:
        return None
    if isinstance(tiles, ordering.Tiling):
        return tiles
    if tiles is True:
        syst = hamiltonian
        if not _has_positions(syst):
            raise ValueError("tiles=True needs a Kwant system with site positions; pass "
                             "a kwant_amd.ordering.Tiling for a matrix")
        ranges = np.asarray(syst.site_ranges, dtype=np.int64)
        if hasattr(syst, 'pos_array'):
            pos = np.asarray(syst.pos_array, dtype=float)
        elif hasattr(syst, 'sites'):
            pos = np.array([np.asarray(site.pos, dtype=float) for site in syst.sites])
        else:   # kwant's System.pos(i)
            pos = np.array([np.asarray(syst.pos(i), dtype=float)
                            for i in range(syst.graph.num_nodes)])
        norbs = np.zeros(len(pos), dtype=np.int64)
        for (first, n_orb, _), (nxt, _, _) in zip(ranges[:-1], ranges[1:]):
            norbs[first:nxt] = n_orb
        return ordering.tile_order(np.repeat(pos.reshape(len(pos), -1), norbs, axis=0),
                                   tile_rows)
    raise ValueError("tiles must be None, False, True, 'auto', 'graph' or a Tiling")


def _orbitals_of(syst, where):
    """(total orbitals, orbital indices) of ``where`` (kwant/kpm.py:1025-1039)."""
    if _is_system(syst):
        return _system_orbitals(syst, where)
    try:
        tot = csr_matrix(syst).shape[0]
    except TypeError:
        raise TypeError("'syst' is neither a matrix nor a Kwant system.")
    return tot, (range(tot) if where is None else np.asarray(where, int))


def _system_orbitals(syst, where):
    """Orbitals of the sites in ``where`` for a kwant system, from its
    ``site_ranges`` (kwant/system.py:27-34, kwant/operator.pyx:114-199)."""
    ranges = np.asarray(syst.site_ranges, dtype=np.int64)
    tot = int(ranges[-1][2])
    sites = list(range(syst.graph.num_nodes)) if where is None else None
    if sites is None:
        if callable(where):
            sites = [i for i, s in enumerate(syst.sites) if where(s)]
        else:
            ids = getattr(syst, 'id_by_site', None)
            sites = [ids[s] if ids is not None and not isinstance(s, numbers.Integral)
                     else int(s) for s in where]
    # each site's orbitals in site order, the sites in the given order
    sites = np.asarray(sites, dtype=np.int64).reshape(-1)
    k = np.searchsorted(ranges[:, 0], sites, side='right') - 1
    first, norbs, offset = ranges[k, 0], ranges[k, 1], ranges[k, 2]
    start = offset + (sites - first) * norbs
    within = np.arange(int(norbs.sum()), dtype=np.int64) - np.repeat(np.cumsum(norbs) - norbs, norbs)
    return tot, np.repeat(start, norbs) + within


def RandomVectors(syst, where=None, rng=None):
    """Random-phase vectors exp(2 pi i u), u ~ U[0, 1), on the orbitals of
    ``where`` and zero elsewhere; one ``rng.random_sample(len(where))`` draw
    per vector (kwant/kpm.py:932-951)."""
    rng = ensure_rng(rng)
    tot, orbs = _orbitals_of(syst, where)
    # every orbital in order: one slice assignment instead of a gather
    if len(orbs) == tot and np.array_equal(orbs, np.arange(tot)):
        orbs = slice(None)
        count = tot
    else:
        orbs = np.asarray(orbs, dtype=np.intp)
        count = len(orbs)
    while True:
        vec = np.zeros(tot, dtype=complex)
        vec[orbs] = np.exp(2j * np.pi * rng.random_sample(count))
        yield vec


class LocalVectors:
    """Unit (float64) vectors on each orbital of ``where``, in order
    (kwant/kpm.py:954-985)."""

    def __init__(self, syst, where=None, *args):
        self.tot_norbs, self.orbs = _orbitals_of(syst, where)
        self._idx = 0

    def __len__(self):
        return len(self.orbs)

    def __iter__(self):
        return self

    def __next__(self):
        if self._idx >= len(self):
            raise StopIteration('Too many vectors requested from this generator')
        vec = np.zeros(self.tot_norbs)
        vec[self.orbs[self._idx]] = 1
        self._idx += 1
        return vec


class _VectorFactory:
    """Indexed, lazily filled view of a vector iterable (kwant/kpm.py:850-929)."""

    def __init__(self, vectors=None, num_vectors=None, accumulate=True):
        assert isinstance(vectors, Iterable)
        try:
            size = len(vectors)
            if num_vectors is None:
                num_vectors = size
        except TypeError:
            size = np.inf
            if num_vectors is None:
                raise ValueError("'num_vectors' must be specified when "
                                 "'vectors' has no len() method.")
        self._max_vectors = size
        self._iterator = iter(vectors)
        self.accumulate = accumulate
        self.saved_vectors = []
        self.num_vectors = 0
        self.add_vectors(num_vectors=num_vectors)
        self._last_idx = -np.inf
        self._last_vector = None

    def __getitem__(self, index):
        if index < self._last_idx and not self.accumulate:
            raise ValueError("Cannot get previous values if 'accumulate' is False")
        if index >= self.num_vectors:
            raise IndexError('Requested more vectors than available')
        self._last_idx = index
        if self.accumulate:
            if self.saved_vectors[index] is None:
                self.saved_vectors[index] = next(self._iterator)
            return self.saved_vectors[index]
        self._last_vector = next(self._iterator)
        return self._last_vector

    def add_vectors(self, num_vectors=None):
        if num_vectors is None:
            raise ValueError("'num_vectors' must be specified.")
        if num_vectors <= 0 or num_vectors != int(num_vectors):
            raise ValueError("'num_vectors' must be a positive integer")
        if self.num_vectors + num_vectors > self._max_vectors:
            raise ValueError("'num_vectors' is larger than available vectors")
        self.num_vectors += num_vectors
        if self.accumulate:
            self.saved_vectors.extend([None] * num_vectors)


# ---------------------------------------------------------------------------
# kernels and transforms (host FP64; kwant/kpm.py:990-1009, 1106-1142, 1190-1222)

def jackson_kernel(moments):
    """Jackson damping g_m (Weisse et al., RMP 78, 275, Eq. 71) applied along
    axis 0 (kwant/kpm.py:1106-1123)."""
    count = len(moments)
    q = np.pi / (count + 1)
    m = np.arange(count)
    g = ((count - m + 1) * np.cos(q * m) + np.sin(q * m) / np.tan(q)) / (count + 1)
    return np.transpose(moments.transpose() * g)


def lorentz_kernel(moments, l=4):
    """Lorentz damping sinh(l (1 - m/M)) / sinh(l) (kwant/kpm.py:1126-1142)."""
    count = len(moments)
    g = np.sinh(l * (1 - np.arange(count) / count)) / np.sinh(l)
    return np.transpose(moments.transpose() * g)


def fermi_distribution(energy, mu, temperature):
    """Fermi function; a step at T = 0 (kwant/kpm.py:990-1009)."""
    if temperature < 0:
        raise ValueError("temperature must be non-negative")
    if temperature == 0:
        return np.array(np.less(energy - mu, 0), dtype=float)
    return 1 / (1 + np.exp((energy - mu) / temperature))


def _chebyshev_nodes(n_sampling):
    """x_j = cos(pi (1 - (2 j + 1) / (2 n))), ascending (kwant/kpm.py:1190-1194)."""
    raw, step = np.linspace(np.pi, 0, n_sampling, endpoint=False, retstep=True)
    return np.cos(raw + step / 2)


# Moment arrays with at least this many (moment x channel) entries are
# transformed on the GPU (channels: the trailing axes, e.g. sites x vectors of
# a site-resolved density with mean=False); smaller ones on the host.
GPU_TRANSFORM_MIN = 1 << 22


def _transform_on_gpu(moments):
    return moments.ndim > 1 and moments.size >= GPU_TRANSFORM_MIN


def _gemm_fp64(a, channels):
    """a (rows x M, host) @ channels (M x C, host) as one FP64 matrix-core
    GEMM on the GPU (kpm_corr_gemm: mu += conj(a) @ b^T, a real here)."""
    import torch
    dev = eng.torch_device()
    ta = torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).to(dev)
    tb = torch.from_numpy(np.ascontiguousarray(channels.T, dtype=np.complex128)).to(dev)
    mu = torch.zeros((a.shape[0], channels.shape[1]), dtype=torch.complex128, device=dev)
    eng.corr_gemm(ta, tb, mu, eng.GEMM_FP64)
    out = mu.cpu().numpy()
    return out if np.iscomplexobj(channels) else out.real.copy()


def _dct3_gemm(moments):
    """The unnormalised DCT-III of the zero-padded moments along axis 0,
    y_k = x_0 + 2 sum_{m>=1} x_m cos(pi m (2k + 1) / 2n) for n = SAMPLING * M
    nodes, as a GEMM on the GPU: the transform of kwant/kpm.py:1212-1216 for
    many channels at once."""
    count = moments.shape[0]
    n = SAMPLING * count
    k = np.arange(n, dtype=np.float64)[:, None]
    m = np.arange(count, dtype=np.float64)[None, :]
    basis = 2 * np.cos(np.pi * m * (2 * k + 1) / (2 * n))
    basis[:, 0] = 1.0
    flat = moments.reshape(count, -1)
    return _gemm_fp64(basis, flat).reshape((n,) + moments.shape[1:])


def _chebval_gemm(x, moments):
    """sum_m moments[m] T_m(x) for every channel (numpy.polynomial's chebval
    for coefficient arrays), as a GEMM on the GPU; x is 1-D."""
    count = moments.shape[0]
    t = np.empty((len(x), count))
    t[:, 0] = 1.0
    if count > 1:
        t[:, 1] = x
    for j in range(2, count):
        t[:, j] = 2 * x * t[:, j - 1] - t[:, j - 2]
    flat = moments.reshape(count, -1)
    return _gemm_fp64(t, flat).reshape((len(x),) + moments.shape[1:])


def _calc_fft_moments(moments):
    """Densities (and Gauss-Chebyshev weights) at the 2M Chebyshev nodes by an
    unnormalised DCT-III of the zero-padded moments (kwant/kpm.py:1197-1222).
    Large multi-channel arrays take the DCT as a GPU GEMM."""
    moments = np.asarray(moments)
    if _transform_on_gpu(moments):
        count = moments.shape[0]
        nodes = _chebyshev_nodes(SAMPLING * count)
        gammas = _dct3_gemm(moments)[::-1]
        rho = np.transpose(np.divide(gammas.transpose(), np.pi * np.sqrt(1 - nodes ** 2)))
        return rho, gammas
    count, *extra = moments.shape
    n_sampling = SAMPLING * count
    padded = np.zeros([n_sampling] + extra, dtype=moments.dtype)
    padded[:count] = moments
    nodes = _chebyshev_nodes(n_sampling)
    gammas = np.transpose(scipy.fftpack.dct(padded.transpose(), type=3))[::-1]
    rho = np.transpose(np.divide(gammas.transpose(), np.pi * np.sqrt(1 - nodes ** 2)))
    return rho, gammas


def _scale_from_bounds(lmin, lmax, eps):
    """(a, b) and the single-eigenvalue check of kwant/kpm.py:1164, 1174-1180."""
    tol = eps / 2
    a = np.abs(lmax - lmin) / (2. - eps)
    b = (lmax + lmin) / 2.
    if lmax - lmin <= abs(lmax + lmin) * tol / 2:
        raise ValueError('The Hamiltonian has a single eigenvalue, it is not possible '
                         'to obtain a spectral density.')
    return a, b


class _RescaledOperator(LinearOperator):
    """(H v - b v) / a evaluated on the GPU (the ``hamiltonian`` attribute of
    kwant/kpm.py:1185)."""

    def __init__(self, dmat, a, b):
        self._dmat, self._a, self._b = dmat, a, b
        super().__init__(dtype=np.complex128, shape=dmat.shape)

    def _matvec(self, v):
        import torch
        dev = eng.torch_device()
        cdt, _ = eng.complex_dtype(self._dmat.precision)
        v = self._dmat.rows_in(np.asarray(v).reshape(-1))
        x = torch.from_numpy(np.ascontiguousarray(v, dtype=np.complex128)).to(dev).to(cdt)
        y = torch.empty_like(x)
        self._dmat.spmm(x, y, 1)
        y = (y - self._b * x) / self._a
        return self._dmat.rows_out(y.cpu().numpy().astype(np.complex128))


def _rescale(hamiltonian, eps, v0, bounds, precision='complex128', order=None):
    """Upload H (rows in ``order``, see kwant_amd.ordering) and return (GPU
    rescaled operator, (a, b)) as kwant/kpm.py:1145-1187; bounds=None runs
    FP64 Lanczos on the GPU."""
    dmat = eng.DeviceCSR(hamiltonian, precision, order=order)
    if bounds:
        lmin, lmax = bounds
    else:
        dmat64 = dmat if dmat.precision in ('complex128', 'c128') else \
            eng.DeviceCSR(hamiltonian, 'complex128', order=order)
        lmin, lmax = lanczos_bounds(dmat64, dmat64.rows_in(v0), tol=eps / 2)
    a, b = _scale_from_bounds(lmin, lmax, eps)
    return _RescaledOperator(dmat, a, b), (a, b)


# ---------------------------------------------------------------------------
# GPU moment blocks

class _DosBlock:
    """W vectors expanded together: one BlockRecursion plus their moments."""

    def __init__(self, dmat, ab, vectors, width, mode, w=None, variant=-1, tiles=None):
        import torch
        dev = eng.torch_device()
        cdt, _ = eng.complex_dtype(dmat.precision)
        host = np.stack([np.asarray(v) for v in vectors], axis=1).astype(np.complex128)
        _check_finite_vectors(host)
        block = torch.from_numpy(np.ascontiguousarray(dmat.rows_in(host))).to(dev).to(cdt)
        self.count = len(vectors)
        self.rec = eng.BlockRecursion(dmat, ab, block, width, mode, w=w, variant=variant,
                                      tiles=tiles)
        self.moments = None  # (count, M) complex128


def _check_finite_vectors(host):
    """Start vectors with NaN or Inf would spread through every moment."""
    if not np.all(np.isfinite(host)):
        bad = sorted(set(np.nonzero(~np.isfinite(host))[1].tolist()))
        raise ValueError("vectors %s of the block contain NaN or Inf" % bad)


# |mu_n| <= mu_0 for every n while the rescaled spectrum lies in [-1, 1]
# (|T_n| <= 1 there); a spectrum outside makes the moments grow like
# cosh(n arccosh|x|), so a moment past mu_0 by this margin means the bounds
# miss the spectrum.
MOMENT_GROWTH_TOL = 1e-2


def _check_moment_growth(mom, bounds):
    """Raise if the doubling moments (count, M) of a block are not finite or
    grow past mu_0 (their first column)."""
    if not np.all(np.isfinite(mom)):
        raise ValueError("the Chebyshev moments are not finite (NaN or Inf in the "
                         "Hamiltonian or the vectors)")
    mu0 = np.abs(mom[:, :1])
    if np.any(np.abs(mom) > mu0 * (1 + MOMENT_GROWTH_TOL)):
        n = int(np.nonzero(np.any(np.abs(mom) > mu0 * (1 + MOMENT_GROWTH_TOL), axis=0))[0][0])
        raise ValueError("the Chebyshev moments grow past mu_0 (first at moment %d): the "
                         "bounds %s do not contain the spectrum of the Hamiltonian"
                         % (n, tuple(float(x) for x in bounds)))


def _doubling_fill(mom, dots, first_step, mu0, mu1):
    """Moments 2s, 2s+1 of steps s = first_step.. from the per-step dots
    (kwant/kpm.py:470-473)."""
    d0 = dots[:, :, 0]
    d1 = dots[:, :, 2] + 1j * dots[:, :, 3]
    for i in range(dots.shape[0]):
        s = first_step + i
        if s == 0:
            continue
        mom[:, 2 * s] = 2 * d0[i, :mom.shape[0]] - mu0
        mom[:, 2 * s + 1] = 2 * d1[i, :mom.shape[0]] - mu1


class SpectralDensity:
    """Spectral density of an operator by the kernel polynomial method,
    computed on the GPU.  Arguments and attributes as kwant/kpm.py:39-161."""

    def __init__(self, hamiltonian, params=None, operator=None,
                 num_vectors=10, num_moments=None, energy_resolution=None,
                 vector_factory=None, bounds=None, eps=0.05, rng=None,
                 kernel=None, mean=True, accumulate_vectors=True, *,
                 precision='complex128', max_block_width=64, variant=-1, reorder='auto',
                 matrix_format='csr', tiles='auto'):
        if num_moments and energy_resolution:
            raise TypeError("either 'num_moments' or 'energy_resolution' "
                            "must be provided.")
        self.eps = eps
        csr = _as_csr(hamiltonian, params)
        # 'auto' tiles only the plain DOS path with the default row order, and
        # only long expansions (the default 100 moments x 10 vectors is not)
        work = csr.shape[0] * (num_vectors or 0) * ((num_moments or 100) // 2)
        tiling = _tiling_of(hamiltonian, tiles, csr=csr, auto_ok=(
            operator is None and matrix_format == 'csr' and vector_factory is None
            and not energy_resolution and work >= TILES_AUTO_WORK
            and isinstance(reorder, str) and reorder == 'auto'))
        hamiltonian = csr

        self._set_operator(operator, params, hamiltonian)

        self.mean = mean
        rng = ensure_rng(rng)
        self._v0 = np.exp(2j * np.pi * rng.random_sample(hamiltonian.shape[0]))
        if eps <= 0:
            raise ValueError("'eps' must be positive")
        if max_block_width < 1 or max_block_width > 64 or \
                max_block_width & (max_block_width - 1):
            raise ValueError("max_block_width must be a power of two <= 64")
        self._precision = precision
        self._max_width = max_block_width
        self._variant = variant
        if matrix_format not in ('csr', 'bsr'):
            raise ValueError("matrix_format must be 'csr' or 'bsr'")
        if matrix_format == 'bsr' and self.operator is not None:
            raise ValueError("matrix_format='bsr' runs the doubling path (operator=None)")
        if tiling is not None and matrix_format == 'bsr':
            raise ValueError("tiles run the CSR recursion (matrix_format='csr')")
        if tiling is not None and len(tiling.order) != hamiltonian.shape[0]:
            raise ValueError("the tiling covers %d rows, the Hamiltonian has %d"
                             % (len(tiling.order), hamiltonian.shape[0]))
        self._tiling, self._tiled = tiling, {}
        if tiling is not None:
            order = tiling.order
        elif matrix_format == 'bsr':
            order = None
        else:
            order = ordering.choose_order(hamiltonian, reorder)
        self.hamiltonian, (self._a, self._b) = _rescale(
            hamiltonian, self.eps, self._v0, bounds, precision, order)
        self._dmat = self.hamiltonian._dmat
        if matrix_format == 'bsr':
            # the recursion reads 4 x 4 site blocks; the CSR copy serves the
            # rescaled matvec and the bounds
            self._dmat = eng.DeviceBSR(hamiltonian, precision)
        self.bounds = (self._b - self._a, self._b + self._a)

        if energy_resolution:
            num_moments = math.ceil((1.6 * self._a) / energy_resolution)
        elif num_moments is None:
            num_moments = 100
        if num_moments <= 0 or num_moments != int(num_moments):
            raise ValueError("'num_moments' must be a positive integer")

        self._rng, self._default_vectors = rng, vector_factory is None
        if vector_factory is None:
            self._vector_factory = _VectorFactory(
                RandomVectors(hamiltonian, rng=rng), num_vectors=num_vectors,
                accumulate=accumulate_vectors)
        else:
            if not isinstance(vector_factory, Iterable):
                raise TypeError('vector_factory must be iterable')
            try:
                len(vector_factory)
            except TypeError:
                if num_vectors is None:
                    raise ValueError('num_vectors must be provided if'
                                     'vector_factory has no length.')
            self._vector_factory = _VectorFactory(
                vector_factory, num_vectors=num_vectors, accumulate=accumulate_vectors)
        num_vectors = self._vector_factory.num_vectors

        self._blocks = []
        self._moments_list = []
        self.num_moments = num_moments
        self._update_moments_list(self.num_moments, num_vectors)

        self.kernel = kernel if kernel is not None else jackson_kernel
        self.densities, self._gammas = _calc_fft_moments(self._moments())

    def _set_operator(self, operator, params, hamiltonian):
        """Normalise ``operator`` (kwant/kpm.py:186-198)."""
        self._op_matrix = None
        self._layout = None
        if operator is None:
            self.operator = None
        elif _is_local_operator(operator):
            self.operator = operator.bind(params=params)
            # Density, Source and Current are evaluated on the GPU: sum=True
            # as the matrix operator sum_i Q_i, sum=False per site (inside the
            # recursion kernel) or per hopping (bond records)
            self._layout = operator_layout(self.operator)
            if self._layout is not None and \
                    self._layout.act_matrix.shape[0] != hamiltonian.shape[0]:
                raise ValueError('vector is incorrect shape')
            if self._layout is not None and self._layout.sum:
                self._op_matrix = self._layout.act_matrix
        elif callable(operator):
            self.operator = operator
        elif hasattr(operator, 'dot'):
            op = csr_matrix(operator)
            self._op_matrix = op
            self.operator = lambda bra, ket: np.vdot(bra, op.dot(ket))
        else:
            raise ValueError("Parameter 'operator' has no '.dot' attribute "
                             "and is not callable.")

    @property
    def energies(self):
        return self._a * _chebyshev_nodes(SAMPLING * self.num_moments) + self._b

    @property
    def num_vectors(self):
        return len(self._moments_list)

    @property
    def _last_two_alphas(self):
        """(alpha_{n-1}, alpha_n) of every vector, copied from the GPU."""
        out = []
        for blk in self._blocks:
            for c in range(blk.count):
                out.append(blk.rec.last_two(c) if blk.rec is not None else 0.)
        return out

    def __call__(self, energy=None):
        """Densities at the nodes, or chebval at ``energy`` (kwant/kpm.py:263-295)."""
        if energy is None:
            return self.energies, self.densities
        energy = np.asarray(energy)
        e = (energy - self._b) / self._a
        g_e = np.pi * np.sqrt(1 - e) * np.sqrt(1 + e)
        moments = self._moments()
        moments[1:] = 2 * moments[1:]
        if _transform_on_gpu(moments) and e.ndim <= 1:
            vals = _chebval_gemm(np.atleast_1d(e).astype(np.float64), moments)
            vals = np.moveaxis(vals, 0, -1)          # chebval's layout: channels, then x
            if e.ndim == 0:
                vals = vals[..., 0]
            return np.transpose(vals / g_e)
        return np.transpose(chebval(e, moments) / g_e)

    def integrate(self, distribution_function=None):
        """Gauss-Chebyshev integral a/(2M) sum_k gamma_k f(E_k)
        (kwant/kpm.py:297-316)."""
        factor = self._a / (2 * self.num_moments)
        if distribution_function is None:
            rho = self._gammas
        else:
            dist = distribution_function(self.energies)
            rho = np.transpose(self._gammas.transpose() * dist)
        return factor * np.sum(rho, axis=0)

    def add_moments(self, num_moments=None, *, energy_resolution=None):
        """Extend every expansion by ``num_moments`` (kwant/kpm.py:318-358)."""
        if not ((num_moments is None) ^ (energy_resolution is None)):
            raise TypeError("either 'num_moments' or 'energy_resolution' "
                            "must be provided.")
        if energy_resolution:
            if energy_resolution <= 0:
                raise ValueError("'energy_resolution' must be positive")
            if self._a * 1.6 / self.num_moments < energy_resolution:
                raise ValueError('Energy resolution is already smaller '
                                 'than the requested resolution')
            num_moments = math.ceil((1.6 * self._a) / energy_resolution)
        if (num_moments is None or num_moments <= 0
                or num_moments != int(num_moments)):
            raise ValueError("'num_moments' must be a positive integer")
        self._update_moments_list(self.num_moments + num_moments, self.num_vectors)
        self.num_moments += num_moments
        self.densities, self._gammas = _calc_fft_moments(self._moments())

    def add_vectors(self, num_vectors=None):
        """Expand ``num_vectors`` more vectors (kwant/kpm.py:360-376)."""
        self._vector_factory.add_vectors(num_vectors)
        num_vectors = self._vector_factory.num_vectors - self.num_vectors
        self._update_moments_list(self.num_moments, self.num_vectors + num_vectors)
        self.densities, self._gammas = _calc_fft_moments(self._moments())

    def save(self, path):
        """Write the expansion (moments, resume state, vectors drawn, the
        random generator's state) to ``path``; see kwant_amd.persist."""
        from . import persist
        persist.save(self, path)

    @classmethod
    def load(cls, path, hamiltonian, params=None, operator=None, kernel=None,
             vector_factory=None):
        """An expansion saved by ``save``, ready for add_moments /
        add_vectors; ``hamiltonian`` and ``operator`` are the ones it was
        made with."""
        from . import persist
        return persist.load(path, hamiltonian, params=params, operator=operator,
                            kernel=kernel, vector_factory=vector_factory)

    def _moments(self):
        """real_if_close, vectors to axis 1, mean, / a, kernel
        (kwant/kpm.py:378-388)."""
        if self._local_mean:
            # the site-resolved blocks hold their vectors' sum: the mean is
            # the total over blocks / R (real_if_close taken on the total)
            total = self._blocks[0].moments.copy()
            for blk in self._blocks[1:]:
                total += blk.moments
            moments = np.real_if_close(total) / self.num_vectors
            moments /= self._a
            return self.kernel(moments)
        moments = np.real_if_close(self._moments_list)
        moments = np.swapaxes(moments, 0, 1)
        if self.mean:
            moments = np.mean(moments, axis=1)
        moments /= self._a
        return self.kernel(moments)

    # -- the moment engine ---------------------------------------------------
    def _update_moments_list(self, n_moments, num_vectors):
        """Validation and bookkeeping of kwant/kpm.py:390-435; the moments of
        new vectors or new moments come from the GPU blocks."""
        if self.num_vectors == num_vectors:
            r_start, new_vectors = 0, 0
        elif self.num_vectors < num_vectors:
            r_start = self.num_vectors
            new_vectors = num_vectors - self.num_vectors
        else:
            raise ValueError('Cannot decrease number of vectors')
        if n_moments == self.num_moments:
            new_moments = 0
            if new_vectors == 0:
                return
        else:
            if not self._vector_factory.accumulate:
                raise ValueError("Cannot increase the number of moments if "
                                 "'accumulate_vectors' is 'False'.")
            new_moments = n_moments - self.num_moments
            if new_moments < 0:
                raise ValueError('Cannot decrease number of moments')
            if new_vectors != 0:
                raise ValueError("Only 'num_moments' *or* 'num_vectors' "
                                 "may be updated at a time.")

        if new_vectors:
            vectors = [self._vector_factory[r] for r in range(r_start, num_vectors)]
            if n_moments < 2:
                # kwant/kpm.py:439-448 writes moments 0 and 1 of every new vector
                raise IndexError('list assignment index out of range')
            start = 0
            for width in eng.pow2_widths(len(vectors), self._max_width):
                chunk = vectors[start:start + width]
                start += width
                blk = self._new_block(chunk, width)
                self._expand_new(blk, n_moments)
                self._moments_list.extend(self._entries(blk))
                if not self._vector_factory.accumulate:
                    blk.rec = None  # no resume without accumulated vectors
                self._blocks.append(blk)
        else:
            for blk in self._blocks:
                self._expand_more(blk, self.num_moments, n_moments)
            self._moments_list = [m for blk in self._blocks for m in self._entries(blk)]

    def _launch_shape(self, width, mode):
        """The requested kernel variant, or the autotuned one (variant=-1)."""
        if self._variant >= 0:
            return self._variant
        return autotune.pick_variant(self._dmat, (self._a, self._b), width, mode)

    @property
    def _local_mean(self):
        return self._layout is not None and not self._layout.sum and self.mean

    def _entries(self, blk):
        """The block's per-vector entries of _moments_list (placeholders
        where the block keeps only its vectors' sum)."""
        if self._local_mean:
            return [None] * blk.count
        return list(blk.moments)

    def _block_mode(self, width):
        """(dots mode, kernel variant) of a new block of ``width`` vectors."""
        if isinstance(self._layout, BondLayout) and not self._layout.sum:
            mode, tune = eng.DOTS_NONE, eng.DOTS_NONE
        elif isinstance(self._layout, SiteLayout) and not self._layout.sum:
            mode = eng.DOTS_LOCAL if self.mean else eng.DOTS_LOCAL_COLS
            tune = eng.DOTS_PROJECT
        elif self.operator is None:
            mode = tune = eng.DOTS_DOUBLING
        elif self._op_matrix is not None:
            mode = tune = eng.DOTS_PROJECT
        else:
            mode = tune = eng.DOTS_NONE
        return mode, self._launch_shape(width, tune)

    def _tiles_for(self, width, mode):
        """The two-step tiling of the device matrix for blocks of ``width``
        vectors (doubling mode only), or None (see autotune.pick_tiled)."""
        if self._tiling is None or mode != eng.DOTS_DOUBLING or \
                not isinstance(self._dmat, eng.DeviceCSR):
            return None
        if width not in self._tiled:
            self._tiled[width] = autotune.pick_tiled(self._dmat, (self._a, self._b), width,
                                                     self._tiling.tile_begin)
        return self._tiled[width]

    def _new_block(self, vectors, width):
        mode, variant = self._block_mode(width)
        blk = _DosBlock(self._dmat, (self._a, self._b), vectors, width, mode, variant=variant,
                        tiles=self._tiles_for(width, mode))
        self._attach(blk, width)
        return blk

    def _attach(self, blk, width):
        """The operator's per-block data, formed from alpha_0 (ring slot 0 of
        a block that has not stepped yet)."""
        import torch
        a0 = blk.rec.ring[0]
        if isinstance(self._layout, BondLayout) and not self._layout.sum:
            # a_0 is kept (the two-slot ring overwrites it); w = M^H a_0
            blk.bra = a0.clone()
            blk.w = blk.bra
            onsite = self._bond_device()[2]
            if onsite is not None:
                blk.w = torch.empty_like(blk.bra)
                onsite[1].spmm(blk.bra, blk.w, width)
        elif blk.rec.mode in (eng.DOTS_LOCAL, eng.DOTS_LOCAL_COLS, eng.DOTS_PROJECT):
            # w = O^H alpha_0, so <w|alpha_n> = vdot(alpha_0, O alpha_n); per row
            # for a site operator (kwant/operator.pyx:791-797)
            op = self._layout.matrix if blk.rec.mode != eng.DOTS_PROJECT else self._op_matrix
            op_h = eng.DeviceCSR(op.conj().T.tocsr(), self._dmat.precision,
                                 order=self._dmat.order)
            w = torch.empty_like(a0)
            op_h.spmm(a0, w, width)
            blk.rec.w = w
        elif blk.rec.mode == eng.DOTS_NONE:
            # a user callable gets alpha_0 on the host
            blk.alpha0 = self._dmat.rows_out(a0.cpu().numpy().astype(np.complex128))

    def _expand_new(self, blk, n_moments):
        rec = blk.rec
        if rec.mode == eng.DOTS_DOUBLING:
            mom = np.zeros((blk.count, n_moments), dtype=np.complex128)
            dots = rec.step(n_moments // 2)
            mu0 = dots[0, :blk.count, 0] + 0j
            mu1 = np.conj(dots[0, :blk.count, 2] + 1j * dots[0, :blk.count, 3])
            mom[:, 0], mom[:, 1] = mu0, mu1
            _doubling_fill(mom, dots, 0, mu0, mu1)
            if n_moments % 2:
                tail = rec.norms()
                mom[:, n_moments - 1] = 2 * tail[:blk.count, 0] - mu0
            _check_moment_growth(mom, self.bounds)
            blk.moments = mom
        elif rec.mode == eng.DOTS_PROJECT:
            mom = np.zeros((blk.count, n_moments), dtype=np.complex128)
            dots = rec.step(n_moments - 1)
            mom[:, 0] = dots[0, :blk.count, 0] + 1j * dots[0, :blk.count, 1]
            if self._layout is not None and self._layout.real_mu0:
                mom[:, 0] = mom[:, 0].real   # op(a_0, a_0) of a Hermitian density
            mom[:, 1:] = (dots[:, :blk.count, 2] + 1j * dots[:, :blk.count, 3]).T
            blk.moments = mom
        elif rec.mode in (eng.DOTS_LOCAL, eng.DOTS_LOCAL_COLS):
            blk.moments = self._local_moments(blk, 0, n_moments, None)
        elif hasattr(blk, 'bra'):
            blk.moments = self._bond_moments(blk, 0, n_moments, None)
        else:
            blk.moments = self._callable_moments(blk, 0, n_moments, None)

    def _expand_more(self, blk, m_old, m_new):
        rec = blk.rec
        if rec.mode == eng.DOTS_DOUBLING:
            mom = np.zeros((blk.count, m_new), dtype=np.complex128)
            mom[:, :m_old] = blk.moments
            mu0, mu1 = mom[:, 0].copy(), mom[:, 1].copy()
            first = m_old // 2
            dots = rec.step(m_new // 2 - first)
            _doubling_fill(mom, dots, first, mu0, mu1)
            if m_new % 2:
                tail = rec.norms()
                mom[:, m_new - 1] = 2 * tail[:blk.count, 0] - mu0
            _check_moment_growth(mom, self.bounds)
            blk.moments = mom
        elif rec.mode == eng.DOTS_PROJECT:
            mom = np.zeros((blk.count, m_new), dtype=np.complex128)
            mom[:, :m_old] = blk.moments
            dots = rec.step(m_new - m_old)
            mom[:, m_old:] = (dots[:, :blk.count, 2] + 1j * dots[:, :blk.count, 3]).T
            blk.moments = mom
        elif rec.mode in (eng.DOTS_LOCAL, eng.DOTS_LOCAL_COLS):
            blk.moments = self._local_moments(blk, m_old, m_new, blk.moments)
        elif hasattr(blk, 'bra'):
            blk.moments = self._bond_moments(blk, m_old, m_new, blk.moments)
        else:
            blk.moments = self._callable_moments(blk, m_old, m_new, blk.moments)

    # device records per step are kept under this many bytes at a time
    _LOCAL_CHUNK_BYTES = 1 << 30

    def _local_rows(self):
        """Device rows of the density's orbitals (None: all rows in order)."""
        import torch
        lay = self._layout
        if getattr(self, '_local_idx', None) is None:
            n = self._dmat.shape[0]
            rows = lay.orbs if self._dmat.inverse is None else self._dmat.inverse[lay.orbs]
            if len(rows) == n and np.array_equal(rows, np.arange(n)):
                self._local_idx = (None,)
            else:
                self._local_idx = (torch.from_numpy(rows).to(eng.torch_device()),)
        return self._local_idx[0]

    def _local_moments(self, blk, m_old, m_new, old):
        """Site-resolved Density moments op(alpha_0, alpha_n) of
        kwant/kpm.py:446-448, 484 for n in [m_old, m_new): the recursion
        kernel writes each row's <w|alpha_n>, the rows of each ``where`` site
        are summed on the host.  mean=True: (M, n_where), the block's vectors
        summed; mean=False: (count, M, n_where)."""
        rec, lay = blk.rec, self._layout
        idx = self._local_rows()
        cols = rec.mode == eng.DOTS_LOCAL_COLS

        def sites(records):
            if idx is not None:
                records = records.index_select(1, idx)
            host = records.cpu().numpy()
            if cols:
                host = host[:, :, :blk.count]
            return lay.site_sums(host)       # (k, n_where) or (k, n_where, count)

        parts = []
        n = m_old
        if n == 0:
            first = sites(rec.local_now()[None])
            if lay.real_mu0:
                first = first.real + 0j      # op(a_0, a_0) of a Hermitian density
            parts.append(first)
            n = 1
        per_step = rec.n_rows * 16 * (rec.width if cols else 1)
        chunk = max(1, self._LOCAL_CHUNK_BYTES // per_step)
        while n < m_new:
            k = min(chunk, m_new - n)
            parts.append(sites(rec.local_steps(k)))
            n += k
        new = np.concatenate(parts, axis=0) if parts else None
        if cols:
            new = np.moveaxis(new, 2, 0)     # (count, k, n_where)
            return new if old is None else np.concatenate([old, new], axis=1)
        return new if old is None else np.concatenate([old, new], axis=0)

    def _bond_device(self):
        """The Current's entries as a device CSR in device row order, the
        CSR position of every entry (entries are listed hopping by hopping),
        and the onsite matrix M with M^H on the device (None when M = 1)."""
        import torch
        if getattr(self, '_bond', None) is None:
            lay, n = self._layout, self._dmat.shape[0]
            inv = self._dmat.inverse
            r = lay.rows if inv is None else inv[lay.rows]
            c = lay.cols if inv is None else inv[lay.cols]
            perm = np.lexsort((c, r))
            indptr = np.zeros(n + 1, dtype=np.int64)
            np.cumsum(np.bincount(r, minlength=n), out=indptr[1:])
            csr = csr_matrix((lay.values[perm], c[perm], indptr), shape=(n, n))
            bond = eng.DeviceCSR(csr, self._dmat.precision)
            pos = np.empty_like(perm)
            pos[perm] = np.arange(len(perm))
            onsite = None
            if lay.onsite is not None:
                onsite = (eng.DeviceCSR(lay.onsite, self._dmat.precision, order=self._dmat.order),
                          eng.DeviceCSR(lay.onsite.conj().T.tocsr(), self._dmat.precision,
                                        order=self._dmat.order))
            self._bond = (bond, torch.from_numpy(pos).to(eng.torch_device()), onsite)
        return self._bond

    def _bond_moments(self, blk, m_old, m_new, old):
        """Current moments op(alpha_0, alpha_n) of kwant/kpm.py:446-448, 484
        for n in [m_old, m_new): after each recursion step the bond kernel
        writes every hopping entry's record (kpm_bond_records), the entries
        of each hopping are summed on the host.  mean=True: (M, n_where),
        the block's vectors summed; mean=False: (count, M, n_where)."""
        import torch
        rec, lay = blk.rec, self._layout
        bond, pos, onsite = self._bond_device()
        cols = not self.mean
        width = rec.width
        tail = (bond.nnz, width) if cols else (bond.nnz,)
        v = torch.empty_like(blk.bra) if onsite is not None else None
        dev = blk.bra.device

        def record(ket, out):
            if onsite is not None:
                onsite[0].spmm(ket, v, width)
            bond.bond_records(blk.bra, ket, v if onsite is not None else ket, blk.w, width,
                              out, per_column=cols)

        def hops(buf):
            host = buf.index_select(1, pos).cpu().numpy()
            if cols:
                host = host[:, :, :blk.count]
            return lay.hop_sums(host)      # (k, n_where) or (k, n_where, count)

        parts = []
        n = m_old
        if n == 0:
            buf = torch.empty((1,) + tail, dtype=torch.complex128, device=dev)
            record(rec.slot(0), buf[0])
            first = hops(buf)
            if lay.real_mu0:
                first = first.real + 0j    # op(a_0, a_0) of a Hermitian current
            parts.append(first)
            n = 1
        per_step = max(1, int(np.prod(tail)) * 16)
        chunk = max(1, self._LOCAL_CHUNK_BYTES // per_step)
        while n < m_new:
            k = min(chunk, m_new - n)
            buf = torch.empty((k,) + tail, dtype=torch.complex128, device=dev)
            for i in range(k):
                rec.step(1, want_dots=False)
                record(rec.slot(rec.n), buf[i])
            parts.append(hops(buf))
            n += k
        new = np.concatenate(parts, axis=0)
        if cols:
            new = np.moveaxis(new, 2, 0)   # (count, k, n_where)
            return new if old is None else np.concatenate([old, new], axis=1)
        return new if old is None else np.concatenate([old, new], axis=0)

    def _callable_moments(self, blk, m_old, m_new, old):
        """Operator path with a Python callable (kwant/kpm.py:478-484): the
        recursion runs on the GPU; each new alpha_n is handed to the callable
        on the host, one moment per step."""
        rec = blk.rec
        alpha0 = blk.alpha0     # kept on the host: the two-slot ring overwrites it
        moms = [list(old[c]) if old is not None else [] for c in range(blk.count)]
        if m_old == 0:
            alpha1 = None
            rec.step(1, want_dots=False)
            alpha1 = self._dmat.rows_out(rec.slot(1).cpu().numpy().astype(np.complex128))
            for c in range(blk.count):
                moms[c].append(self.operator(alpha0[:, c], alpha0[:, c]))
                moms[c].append(self.operator(alpha0[:, c], alpha1[:, c]))
            m_old = 2
        for _ in range(m_old, m_new):
            rec.step(1, want_dots=False)
            cur = self._dmat.rows_out(rec.slot(rec.n).cpu().numpy().astype(np.complex128))
            for c in range(blk.count):
                moms[c].append(self.operator(alpha0[:, c], cur[:, c]))
        return moms


# ---------------------------------------------------------------------------
# Correlator / conductivity

from .correlator import Correlator, conductivity  # noqa: E402
