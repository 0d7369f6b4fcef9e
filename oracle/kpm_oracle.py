"""CPU oracle: a NumPy/SciPy restatement of kwant's KPM moment engine.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this module, and only as
the checker (or the timed CPU baseline) -- the product package ``kwant_amd``
never imports it.

It follows ``kwant/kpm.py`` (reference tree) step by step for matrix
Hamiltonians: the same random-number consumption (the v0 draw first,
kwant/kpm.py:201-203, then one random-phase vector per draw,
kwant/kpm.py:932-951), the same per-vector loop of scipy ``csr @ v`` with the
rescaled matvec (kwant/kpm.py:1182-1183), the doubling identity
(kwant/kpm.py:436-477), ``real_if_close`` + mean + kernel
(kwant/kpm.py:378-388), the DCT-III transform (kwant/kpm.py:1197-1222), and
the Correlator / Kubo-Bastin chain (kwant/kpm.py:557-760, 763-847, 1042-1085).
Builder-made ``kwant.system.System`` inputs are out of its scope (kwant
cannot be imported here); every function takes matrices.

Pinning (what checks this oracle, so that it is not its own reference):
  * moments: dense eigendecomposition brute force (oracle/brute.py) --
    tests/test_oracle_pins.py (the reference ships no fixture data for this
    path, so nothing is committed under tests/golden/);
  * kernels, nodes, DCT, chebval, integrate: closed forms and direct cosine
    sums (tests/test_oracle_pins.py);
  * random vectors: numpy's MT19937 ``RandomState`` itself;
  * Correlator moments: brute force mu_mn = Tr-estimator via eigh on tiny H;
    the whole Kubo-Bastin evaluation (integral factor, weight, simpson,
    prefactor): with the exact trace, N * sigma_xy(mu, T) of a QWZ Chern
    insulator torus converges to the eigenstate Kubo formula
    (oracle/brute.py:kubo_hall_exact, whose normalisation is pinned by the
    Chern number), at T = 0 and T > 0, in the gap and in the bands; planted
    mistakes (conjugated e+, unhalved g_0, halved prefactor) fail it; plus
    sigma_xy = -sigma_yx;
  * ``_rescale`` with bounds=None: the invariant spec((H-b)/a) in (-1, 1).
"""
import math
from collections.abc import Iterable

import numpy as np
import scipy.fftpack
from numpy.polynomial.chebyshev import chebval
from scipy.integrate import simpson
from scipy.sparse import coo_matrix, csr_matrix
from scipy.sparse.linalg import LinearOperator, eigsh

SAMPLING = 2  # kwant/kpm.py:36


# --------------------------------------------------------------------------
# random numbers and vectors

def ensure_rng(rng=None):
    """Seed contract of kwant/_common.py:115-131."""
    if rng is None:
        return np.random.mtrand._rand
    if isinstance(rng, (int, np.integer)):
        return np.random.RandomState(rng)
    needed = ('random_sample', 'randn', 'randint', 'choice')
    if all(hasattr(rng, name) for name in needed):
        return rng
    raise ValueError("Expecting a seed or an object that offers the "
                     "numpy.random.RandomState interface.")


def _orbs_where(ham, where):
    """kwant/kpm.py:1025-1039 for matrix inputs."""
    try:
        n = csr_matrix(ham).shape[0]
    except TypeError:
        raise TypeError("'syst' is neither a matrix nor a Kwant system.")
    orbs = range(n) if where is None else np.asarray(where, int)
    return n, orbs


def random_vectors(ham, where=None, rng=None):
    """Random-phase vectors exp(2 pi i u) on ``where`` (kwant/kpm.py:932-951)."""
    rng = ensure_rng(rng)
    n, orbs = _orbs_where(ham, where)
    while True:
        vec = np.zeros(n, dtype=complex)
        vec[orbs] = np.exp(2j * np.pi * rng.random_sample(len(orbs)))
        yield vec


class LocalVectors:
    """Unit vectors (float64) on the orbitals of ``where``
    (kwant/kpm.py:954-985)."""

    def __init__(self, ham, where=None, *args):
        self.tot_norbs, self.orbs = _orbs_where(ham, where)
        self._next = 0

    def __len__(self):
        return len(self.orbs)

    def __iter__(self):
        return self

    def __next__(self):
        if self._next >= len(self):
            raise StopIteration('Too many vectors requested from this generator')
        vec = np.zeros(self.tot_norbs)
        vec[self.orbs[self._next]] = 1
        self._next += 1
        return vec


class VectorFactory:
    """Lazy indexed vector source (kwant/kpm.py:850-929)."""

    def __init__(self, vectors, num_vectors=None, accumulate=True):
        assert isinstance(vectors, Iterable)
        try:
            available = len(vectors)
            if num_vectors is None:
                num_vectors = available
        except TypeError:
            available = np.inf
            if num_vectors is None:
                raise ValueError("'num_vectors' must be specified when "
                                 "'vectors' has no len() method.")
        self._max_vectors = available
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
        if not self.accumulate:
            self._last_vector = next(self._iterator)
            return self._last_vector
        if self.saved_vectors[index] is None:
            self.saved_vectors[index] = next(self._iterator)
        return self.saved_vectors[index]

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


# --------------------------------------------------------------------------
# rescaling, kernels, transforms

def rescale(ham, eps, v0, bounds):
    """(H - b)/a as a LinearOperator and (a, b) -- kwant/kpm.py:1145-1187."""
    tol = eps / 2
    if bounds:
        lmin, lmax = bounds
    else:
        lmax = eigsh(ham, k=1, which='LA', return_eigenvectors=False,
                     tol=tol, v0=v0)[0]
        lmin = eigsh(ham, k=1, which='SA', return_eigenvectors=False,
                     tol=tol, v0=v0)[0]
    a = np.abs(lmax - lmin) / (2. - eps)
    b = (lmax + lmin) / 2.
    if lmax - lmin <= abs(lmax + lmin) * tol / 2:
        raise ValueError('The Hamiltonian has a single eigenvalue, it is not '
                         'possible to obtain a spectral density.')

    def matvec(v):
        return (ham.dot(v) - b * v) / a

    return LinearOperator(shape=ham.shape, matvec=matvec), (a, b)


def jackson_kernel(moments):
    """Eq. (71) of Weisse et al., RMP 78, 275 (kwant/kpm.py:1106-1123)."""
    m_count = len(moments)
    m = np.arange(m_count)
    q = np.pi / (m_count + 1)
    g = ((m_count - m + 1) * np.cos(q * m) + np.sin(q * m) / np.tan(q))
    g /= m_count + 1
    return np.transpose(moments.transpose() * g)


def lorentz_kernel(moments, l=4):
    """sinh(l (1 - m/M)) / sinh(l) (kwant/kpm.py:1126-1142)."""
    m_count = len(moments)
    m = np.arange(m_count)
    g = np.sinh(l * (1 - m / m_count)) / np.sinh(l)
    return np.transpose(moments.transpose() * g)


def fermi_distribution(energy, mu, temperature):
    """kwant/kpm.py:990-1009."""
    if temperature < 0:
        raise ValueError("temperature must be non-negative")
    if temperature == 0:
        return np.array(np.less(energy - mu, 0), dtype=float)
    return 1 / (1 + np.exp((energy - mu) / temperature))


def chebyshev_nodes(n_sampling):
    """cos of the midpoints of n equal steps from pi down to 0
    (kwant/kpm.py:1190-1194)."""
    raw, step = np.linspace(np.pi, 0, n_sampling, endpoint=False, retstep=True)
    return np.cos(raw + step / 2)


def fft_moments(moments):
    """Densities at the Chebyshev nodes by DCT-III (kwant/kpm.py:1197-1222)."""
    moments = np.asarray(moments)
    m_count, *extra = moments.shape
    n_sampling = SAMPLING * m_count
    padded = np.zeros([n_sampling] + extra, dtype=moments.dtype)
    padded[:m_count] = moments
    x = chebyshev_nodes(n_sampling)
    gammas = np.transpose(scipy.fftpack.dct(padded.transpose(), type=3))[::-1]
    rho = np.transpose(np.divide(gammas.transpose(), np.pi * np.sqrt(1 - x ** 2)))
    return rho, gammas


# --------------------------------------------------------------------------
# the moment loop (kwant/kpm.py:390-490)

def doubling_moments(matvec, alpha0, n_moments, state=None, old=None):
    """Moments of one vector by the doubling identity (kwant/kpm.py:436-477).

    Returns (moments, (alpha, alpha_next)).  ``state``/``old`` resume an
    expansion of len(old) moments (kwant/kpm.py:450-452)."""
    mom = [0.] * n_moments
    if state is None:
        alpha = alpha0
        alpha_next = matvec(alpha)
        mom[0] = np.vdot(alpha0, alpha0)
        mom[1] = np.vdot(alpha0, alpha_next)
        m_start = 2
    else:
        alpha, alpha_next = state
        mom[0:len(old)] = old
        m_start = len(old)
    for n in range(m_start // 2, n_moments // 2):
        alpha, alpha_next = alpha_next, 2 * matvec(alpha_next) - alpha
        mom[2 * n] = 2 * np.vdot(alpha, alpha) - mom[0]
        mom[2 * n + 1] = 2 * np.vdot(alpha_next, alpha) - mom[1]
    if n_moments % 2:
        mom[n_moments - 1] = 2 * np.vdot(alpha_next, alpha_next) - mom[0]
    return mom, (alpha, alpha_next)


def operator_moments(matvec, op, alpha0, n_moments, state=None, old=None):
    """One moment per matvec, mu_n = op(alpha0, alpha_n) (kwant/kpm.py:443-448,
    478-484)."""
    mom = [0.] * n_moments
    if state is None:
        alpha = alpha0
        alpha_next = matvec(alpha)
        mom[0] = op(alpha0, alpha0)
        mom[1] = op(alpha0, alpha_next)
        m_start = 2
    else:
        alpha, alpha_next = state
        mom[0:len(old)] = old
        m_start = len(old)
    for n in range(m_start, n_moments):
        alpha, alpha_next = alpha_next, 2 * matvec(alpha_next) - alpha
        mom[n] = op(alpha0, alpha_next)
    return mom, (alpha, alpha_next)


def _normalize_sd_operator(operator):
    """kwant/kpm.py:186-198 (matrix / callable operators)."""
    if operator is None:
        return None
    if callable(operator):
        return operator
    if hasattr(operator, 'dot'):
        op = csr_matrix(operator)
        return lambda bra, ket: np.vdot(bra, op.dot(ket))
    raise ValueError("Parameter 'operator' has no '.dot' attribute and is "
                     "not callable.")


class SpectralDensity:
    """kwant.kpm.SpectralDensity for matrix Hamiltonians (kwant/kpm.py:39-490)."""

    def __init__(self, hamiltonian, params=None, operator=None,
                 num_vectors=10, num_moments=None, energy_resolution=None,
                 vector_factory=None, bounds=None, eps=0.05, rng=None,
                 kernel=None, mean=True, accumulate_vectors=True):
        if num_moments and energy_resolution:
            raise TypeError("either 'num_moments' or 'energy_resolution' "
                            "must be provided.")
        self.eps = eps
        try:
            hamiltonian = csr_matrix(hamiltonian)
        except Exception:
            raise ValueError("'hamiltonian' is neither a matrix nor a Kwant system.")
        self.operator = _normalize_sd_operator(operator)
        self.mean = mean
        rng = ensure_rng(rng)
        self._v0 = np.exp(2j * np.pi * rng.random_sample(hamiltonian.shape[0]))
        if eps <= 0:
            raise ValueError("'eps' must be positive")
        self.hamiltonian, (self._a, self._b) = rescale(
            hamiltonian, eps=self.eps, v0=self._v0, bounds=bounds)
        self.bounds = (self._b - self._a, self._b + self._a)
        if energy_resolution:
            num_moments = math.ceil((1.6 * self._a) / energy_resolution)
        elif num_moments is None:
            num_moments = 100
        if num_moments <= 0 or num_moments != int(num_moments):
            raise ValueError("'num_moments' must be a positive integer")
        if vector_factory is None:
            self._vector_factory = VectorFactory(
                random_vectors(hamiltonian, rng=rng), num_vectors=num_vectors,
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
            self._vector_factory = VectorFactory(
                vector_factory, num_vectors=num_vectors,
                accumulate=accumulate_vectors)
        num_vectors = self._vector_factory.num_vectors
        self._last_two_alphas = []
        self._moments_list = []
        self.num_moments = num_moments
        self._update_moments_list(self.num_moments, num_vectors)
        self.kernel = kernel if kernel is not None else jackson_kernel
        self.densities, self._gammas = fft_moments(self._moments())

    @property
    def energies(self):
        return self._a * chebyshev_nodes(SAMPLING * self.num_moments) + self._b

    @property
    def num_vectors(self):
        return len(self._moments_list)

    def __call__(self, energy=None):
        """kwant/kpm.py:263-295."""
        if energy is None:
            return self.energies, self.densities
        energy = np.asarray(energy)
        e = (energy - self._b) / self._a
        g_e = np.pi * np.sqrt(1 - e) * np.sqrt(1 + e)
        moments = self._moments()
        moments[1:] = 2 * moments[1:]
        return np.transpose(chebval(e, moments) / g_e)

    def integrate(self, distribution_function=None):
        """Gauss-Chebyshev sum (kwant/kpm.py:297-316)."""
        factor = self._a / (2 * self.num_moments)
        if distribution_function is None:
            rho = self._gammas
        else:
            dist = distribution_function(self.energies)
            rho = np.transpose(self._gammas.transpose() * dist)
        return factor * np.sum(rho, axis=0)

    def add_moments(self, num_moments=None, *, energy_resolution=None):
        """kwant/kpm.py:318-358."""
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
        self.densities, self._gammas = fft_moments(self._moments())

    def add_vectors(self, num_vectors=None):
        """kwant/kpm.py:360-376."""
        self._vector_factory.add_vectors(num_vectors)
        num_vectors = self._vector_factory.num_vectors - self.num_vectors
        self._update_moments_list(self.num_moments, self.num_vectors + num_vectors)
        self.densities, self._gammas = fft_moments(self._moments())

    def _moments(self):
        """kwant/kpm.py:378-388."""
        moments = np.real_if_close(self._moments_list)
        moments = np.swapaxes(moments, 0, 1)
        if self.mean:
            moments = np.mean(moments, axis=1)
        moments /= self._a
        return self.kernel(moments)

    def _update_moments_list(self, n_moments, num_vectors):
        """kwant/kpm.py:390-490."""
        if self.num_vectors == num_vectors:
            r_start, new_vectors = 0, 0
        elif self.num_vectors < num_vectors:
            r_start = self.num_vectors
            new_vectors = num_vectors - self.num_vectors
        else:
            raise ValueError('Cannot decrease number of vectors')
        self._moments_list.extend([0.] * new_vectors)
        self._last_two_alphas.extend([0.] * new_vectors)
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
        matvec = self.hamiltonian.matvec
        for r in range(r_start, num_vectors):
            alpha0 = self._vector_factory[r]
            state = old = None
            if new_moments > 0:
                state = self._last_two_alphas[r]
                old = self._moments_list[r]
            if self.operator is None:
                mom, state = doubling_moments(matvec, alpha0, n_moments, state, old)
            else:
                mom, state = operator_moments(matvec, self.operator, alpha0,
                                              n_moments, state, old)
            if self._vector_factory.accumulate:
                self._last_two_alphas[r] = state
            self._moments_list[r] = mom


# --------------------------------------------------------------------------
# Correlator / conductivity (kwant/kpm.py:493-847, 1042-1103)

def _normalize_operator(op):
    """kwant/kpm.py:1088-1103 for matrices and callables."""
    if op is None:
        return lambda v: v
    if callable(op):
        return op
    if hasattr(op, 'dot'):
        return csr_matrix(op).dot
    raise TypeError("The operators must have a '.dot' attribute or must be "
                    "callable.")


def velocity(hamiltonian, op_type, positions):
    """V_ij = i (x_j - x_i)_d H_ij on the pattern of H (kwant/kpm.py:1042-1085)."""
    if isinstance(op_type, str):
        direction = dict(x=0, y=1, z=2)[op_type]
        op = coo_matrix(hamiltonian, copy=True)
        disp = positions[op.col] - positions[op.row]
        if direction > disp.shape[1]:
            raise ValueError("{} is not an allowed direction.".format(op_type))
        op.data = op.data * (1j * disp[:, direction])
        return op.tocsr()
    try:
        return csr_matrix(op_type)
    except Exception:
        raise ValueError("Velocity operator must be provided as a matrix, a "
                         "kwant operator, or a direction.")


def integral_factor(mu_kernel, e_scaled):
    """sum_mn mu~_mn Gamma_mn(e) for each rescaled energy, by the direct loop
    of kwant/kpm.py:745-760."""
    m_count = mu_kernel.shape[-1]
    m = np.arange(m_count)
    out = []
    for e in e_scaled:
        root = np.sqrt(1 - e ** 2)
        phase = np.exp(1j * np.arccos(e) * m)
        t_n = np.real(phase)
        e_plus = (e - 1j * root * m) * phase
        gamma = e_plus[None, :] * t_n[:, None]
        gamma = gamma + gamma.conj().T
        out.append(np.tensordot(mu_kernel, gamma.T))
    return np.array(out).T


class Correlator:
    """kwant.kpm.Correlator for matrix Hamiltonians and matrix/callable
    operators (kwant/kpm.py:493-760)."""

    def __init__(self, hamiltonian, operator1=None, operator2=None, **kwargs):
        self.mean = kwargs.get('mean', True)
        accumulate_vectors = kwargs.get('accumulate_vectors', False)
        kwargs['accumulate_vectors'] = True
        kwargs.pop('operator', None)
        self.operator1 = _normalize_operator(operator1)
        self.operator2 = _normalize_operator(operator2)

        def keep_ket(bra, ket):
            return ket

        self._spectrum_R = SpectralDensity(hamiltonian, operator=keep_ket, **kwargs)
        self._a = self._spectrum_R._a
        self._b = self._spectrum_R._b
        shrunk = self._a * (1 - self._spectrum_R.eps / 2)
        bounds = (self._b - shrunk, self._b + shrunk)
        self.num_vectors = self._spectrum_R.num_vectors
        self.num_moments = self._spectrum_R.num_moments
        self._update_psi()
        kwargs['accumulate_vectors'] = accumulate_vectors
        kwargs['num_vectors'] = self.num_vectors
        kwargs['num_moments'] = self.num_moments
        kwargs['energy_resolution'] = None
        kwargs['vector_factory'] = self._op_factory()
        kwargs['bounds'] = bounds
        self._spectrum_L = SpectralDensity(hamiltonian, operator=keep_ket, **kwargs)
        self._omega = np.array(self._spectrum_L._moments_list)
        self._calculate_moments_matrix()
        self._build_integral_factor()

    def __call__(self, mu=0, temperature=0):
        """kwant/kpm.py:608-630."""
        e = self.energies
        x = (e - self._b) / self._a
        f = fermi_distribution(e, mu, temperature)
        integrand = np.divide(f, (1 - x ** 2) ** 2)
        integrand = np.multiply(integrand, self._integral_factor)
        integral = simpson(integrand, x=x)
        return 2 * 4 ** 2 / ((2 * self._a) ** 2) * integral

    @property
    def energies(self):
        return self._spectrum_R.energies

    def add_moments(self, num_moments=None, *, energy_resolution=None):
        self._spectrum_R.add_moments(num_moments=num_moments,
                                     energy_resolution=energy_resolution)
        self.num_moments = self._spectrum_R.num_moments
        self._update_psi()
        self._spectrum_L.add_moments(num_moments=num_moments,
                                     energy_resolution=energy_resolution)
        self._omega = np.array(self._spectrum_L._moments_list)
        self._calculate_moments_matrix()
        self._build_integral_factor()

    def add_vectors(self, num_vectors=None):
        self._spectrum_R.add_vectors(num_vectors)
        self._update_psi()
        self._spectrum_L.add_vectors(num_vectors)
        self.num_vectors = self._spectrum_L.num_vectors
        self._omega = np.array(self._spectrum_L._moments_list)
        self._calculate_moments_matrix()
        self._build_integral_factor()

    def _calculate_moments_matrix(self):
        self.moments_matrix = self._omega.conjugate() @ self._psi
        if self.mean:
            self.moments_matrix = np.mean(self.moments_matrix, axis=0)

    def _op_factory(self):
        for vector in self._spectrum_R._vector_factory:
            yield self.operator1(vector)

    def _update_psi(self):
        spec = self._spectrum_R
        self._psi = np.array([[self.operator2(spec._moments_list[r][n])
                               for n in range(spec.num_moments)]
                              for r in range(spec.num_vectors)]).swapaxes(1, 2)

    def _build_integral_factor(self):
        g = self._spectrum_R.kernel(np.ones(self.num_moments))
        g[0] /= 2
        mu_kernel = np.outer(g, g) * self.moments_matrix
        x = (self.energies - self._b) / self._a
        self._integral_factor = integral_factor(mu_kernel, x)


def conductivity(hamiltonian, alpha='x', beta='x', positions=None, **kwargs):
    """kwant/kpm.py:763-847 for matrix Hamiltonians."""
    if positions is None:
        raise ValueError("If 'hamiltonian' is a matrix, positions must be provided")
    ham = csr_matrix(hamiltonian)
    va = velocity(ham, alpha, positions)
    vb = velocity(ham, beta, positions)
    return Correlator(hamiltonian, operator1=va, operator2=vb, **kwargs)


# --------------------------------------------------------------------------
# CPU-baseline helper: the reference's per-vector loop over a bounded sample

def time_doubling_steps(ham, a, b, alpha0, n_steps):
    """Run ``n_steps`` iterations of the kwant/kpm.py:464-473 loop for one
    vector (scipy csr @ v + numpy temporaries + two vdot per step) and return
    the last two moments; used to time the CPU baseline."""
    ham = csr_matrix(ham)

    def matvec(v):
        return (ham.dot(v) - b * v) / a

    alpha = alpha0
    alpha_next = matvec(alpha)
    mu0 = np.vdot(alpha0, alpha0)
    mu1 = np.vdot(alpha0, alpha_next)
    out = None
    for _ in range(n_steps):
        alpha, alpha_next = alpha_next, 2 * matvec(alpha_next) - alpha
        out = (2 * np.vdot(alpha, alpha) - mu0, 2 * np.vdot(alpha_next, alpha) - mu1)
    return out
