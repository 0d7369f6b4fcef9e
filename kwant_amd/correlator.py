"""Kubo-Bastin Correlator and conductivity with GPU expansions and a GPU GEMM.

Mirrors kwant/kpm.py:493-847.  For every block of random vectors r:

* the "right" expansion T_n(H~)|r> (kwant/kpm.py:574) runs through a GPU
  ring of OMEGA_CHUNK slots, and each run of new vectors becomes
  Psi_n = op2 T_n(H~)|r> by one batched SpMM (kwant/kpm.py:708-718): Psi
  (num_moments slots) is the only block kept whole;
* the "left" expansion T_m(H~) op1|r> (kwant/kpm.py:588-603) starts from
  op1|r> (SpMM) with the same (a, b) -- recomputed through the shrunken
  bounds exactly as kwant/kpm.py:578-579 does -- and streams through its own
  ring: every run of OMEGA_CHUNK rows of Omega is multiplied into
  mu[m0:m1, :] += sum_i conj(Omega_m[i]) Psi_n[i], a GPU GEMM over
  K = N x block width (kwant/kpm.py:685-695), FP64 by default
  (``gemm_precision='fp32'`` runs the complex64 path on the FP32 matrix cores
  with FP64 accumulation across k-chunks).  Memory per vector is
  (num_moments + OMEGA_CHUNK) slots instead of 2 num_moments, so wider
  blocks fit.

The integral factor (kwant/kpm.py:720-760) is evaluated in its GEMM form,
factor(x_k) = e+_k^T mu~ T_k + T_k^T mu~ conj(e+_k), and the final
simpson integral (kwant/kpm.py:608-630) is FP64 on the host.
"""
import numpy as np
from scipy.integrate import simpson
from scipy.sparse import coo_matrix, csr_matrix

from . import _engine as eng
from . import autotune
from . import ordering
from .operator import operator_layout
from .kpm import (SpectralDensity, _VectorFactory, RandomVectors, _as_csr, _check_finite_vectors,
                  _is_local_operator,
                  _is_system, _rescale, _scale_from_bounds, ensure_rng, fermi_distribution,
                  jackson_kernel, SAMPLING, _chebyshev_nodes)
from collections.abc import Iterable
import math

# The split-precision modes cut each FP64 component into exact integer
# slices: 'isliced4' / 'isliced5' (7-bit slices on the INT8 matrix cores,
# exact int32 sums; 5 slices are within ~1e-9 of FP64 and the fastest form)
# and 'sliced3' / 'sliced4' (8-bit slices on the BF16 matrix cores).  A hi/lo
# split into FP32 (KPM_GEMM_FP32X3 in the C ABI) is not offered: the FP32
# matrix cores round every product and their accumulator to 24 bits, so it is
# no more accurate than plain FP32.
_GEMM = {'fp64': eng.GEMM_FP64, 'fp32': eng.GEMM_FP32, 'sliced3': ('bf16', 3),
         'sliced4': ('bf16', 4), 'isliced4': ('int8', 4), 'isliced5': ('int8', 5)}


def _normalize_operator(op, params):
    """None -> identity, kwant operator -> op.act, callable, or a matrix
    (kwant/kpm.py:1088-1103).  Returns ('id'|'matrix'|'callable', payload)."""
    if op is None:
        return 'id', None
    if _is_local_operator(op):
        bound = op.bind(params=params)
        layout = operator_layout(bound)
        if layout is not None:
            # Density / Source / Current act as one sparse matrix on the GPU
            return 'matrix', layout.act_matrix
        return 'callable', bound.act
    if callable(op):
        return 'callable', op
    if hasattr(op, 'dot'):
        return 'matrix', csr_matrix(op)
    raise TypeError("The operators must have a '.dot' attribute or must be callable.")


class _Expansion:
    """Inputs of one SpectralDensity of kwant/kpm.py:163-241 without its
    moments: matrix on the GPU, (a, b), num_moments and the vector factory."""

    def __init__(self, hamiltonian, params=None, num_vectors=10, num_moments=None,
                 energy_resolution=None, vector_factory=None, bounds=None, eps=0.05,
                 rng=None, kernel=None, mean=True, accumulate_vectors=True,
                 precision='complex128', reorder='auto'):
        if num_moments and energy_resolution:
            raise TypeError("either 'num_moments' or 'energy_resolution' must be provided.")
        self.eps = eps
        ham = _as_csr(hamiltonian, params)
        self.csr = ham
        rng = ensure_rng(rng)
        self._v0 = np.exp(2j * np.pi * rng.random_sample(ham.shape[0]))
        if eps <= 0:
            raise ValueError("'eps' must be positive")
        order = ordering.choose_order(ham, reorder)
        self.hamiltonian, (self._a, self._b) = _rescale(ham, eps, self._v0, bounds, precision,
                                                        order)
        self.dmat = self.hamiltonian._dmat
        if energy_resolution:
            num_moments = math.ceil((1.6 * self._a) / energy_resolution)
        elif num_moments is None:
            num_moments = 100
        if num_moments <= 0 or num_moments != int(num_moments):
            raise ValueError("'num_moments' must be a positive integer")
        self.num_moments = num_moments
        if vector_factory is None:
            self._vector_factory = _VectorFactory(
                RandomVectors(ham, rng=rng), num_vectors=num_vectors,
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
        self.kernel = kernel if kernel is not None else jackson_kernel

    @property
    def energies(self):
        return self._a * _chebyshev_nodes(SAMPLING * self.num_moments) + self._b


class _SavedNodes:
    """The energies of a loaded Correlator (its expansions are not kept)."""

    def __init__(self, energies):
        self.energies = energies


class Correlator:
    """Response of the correlation between two operators (kwant/kpm.py:493-555).
    ``operator1`` acts to the right as O_a|r>."""

    def __init__(self, hamiltonian, operator1=None, operator2=None, **kwargs):
        params = kwargs.get('params')
        self.mean = kwargs.get('mean', True)
        self._accumulate = kwargs.get('accumulate_vectors', False)
        kwargs['accumulate_vectors'] = True
        kwargs.pop('operator', None)
        gp = kwargs.pop('gemm_precision', None)
        precision = kwargs.get('precision', 'complex128')
        if gp is None:
            gp = 'fp64' if precision in ('complex128', 'c128') else 'fp32'
        if gp not in _GEMM:
            raise ValueError("gemm_precision must be one of " + ", ".join(sorted(_GEMM)))
        if 'sliced' in gp and precision not in ('complex128', 'c128'):
            raise ValueError("the sliced GEMM takes complex128 expansions")
        self._gemm_precision = gp
        self._max_width = kwargs.pop('max_block_width', 16)
        self._variant = int(kwargs.pop('variant', -1))
        self._op1 = _normalize_operator(operator1, params)
        self._op2 = _normalize_operator(operator2, params)

        self._spectrum_R = _Expansion(hamiltonian, **kwargs)
        spec = self._spectrum_R
        self._a, self._b = spec._a, spec._b
        shrunk = self._a * (1 - spec.eps / 2)
        # the second expansion recomputes (a, b) from these bounds, as
        # kwant/kpm.py:578-579, 597 make it do
        self._aL, self._bL = _scale_from_bounds(self._b - shrunk, self._b + shrunk, spec.eps)
        self.num_moments = spec.num_moments
        self.num_vectors = 0
        self._dev_ops = {}
        self._mu_sum = None
        self._mu_list = []
        # kwant draws the first expansion's v0, then its random vectors
        # (inside its constructor), then the second expansion's v0
        # (kwant/kpm.py:573-603): with a shared generator (None or a
        # RandomState) the vectors must be drawn before that last v0
        self._add_vectors(spec._vector_factory.num_vectors)
        rng = ensure_rng(kwargs.get('rng'))
        rng.random_sample(spec.csr.shape[0])  # the v0 draw of the second expansion
        self._build_integral_factor()

    # -- public API (kwant/kpm.py:608-683) ---------------------------------
    def __call__(self, mu=0, temperature=0):
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

    def save(self, path):
        """Write what evaluating sigma needs -- (a, b), the moment count and
        the integral factor at the 2M nodes -- and the mean moment matrix
        mu_mn to ``path`` (.npz with a JSON header, nothing pickled)."""
        from . import persist
        meta = dict(a=float(self._a), b=float(self._b), num_moments=int(self.num_moments),
                    num_vectors=int(self.num_vectors), mean=bool(self.mean))
        arrays = dict(integral_factor=np.asarray(self._integral_factor),
                      energies=np.asarray(self.energies))
        if getattr(self, "moments_matrix", None) is not None:
            arrays["moments_matrix"] = np.asarray(self.moments_matrix)
        persist.save_arrays(path, persist.CORRELATOR_FORMAT, meta, arrays)

    @classmethod
    def load(cls, path):
        """A Correlator saved by ``save``, for evaluating sigma(mu, T) with
        ``__call__``; adding moments or vectors needs the Hamiltonian, so a
        loaded Correlator refuses them."""
        from . import persist
        meta, arrays = persist.load_arrays(path, persist.CORRELATOR_FORMAT)
        obj = cls.__new__(cls)
        obj._a, obj._b = meta["a"], meta["b"]
        obj.num_moments, obj.num_vectors = meta["num_moments"], meta["num_vectors"]
        obj.mean = meta["mean"]
        obj._integral_factor = arrays["integral_factor"]
        obj._spectrum_R = _SavedNodes(arrays["energies"])
        obj.moments_matrix = arrays.get("moments_matrix")
        obj._loaded = True
        return obj

    def _check_live(self):
        if getattr(self, "_loaded", False):
            raise ValueError("a loaded Correlator evaluates sigma only; build it again "
                             "from the Hamiltonian to add moments or vectors")

    def add_moments(self, num_moments=None, *, energy_resolution=None):
        """Extend both expansions (kwant/kpm.py:636-661); the moment matrix is
        recomputed from the saved vectors."""
        self._check_live()
        spec = self._spectrum_R
        if not ((num_moments is None) ^ (energy_resolution is None)):
            raise TypeError("either 'num_moments' or 'energy_resolution' must be provided.")
        if energy_resolution:
            if energy_resolution <= 0:
                raise ValueError("'energy_resolution' must be positive")
            if spec._a * 1.6 / spec.num_moments < energy_resolution:
                raise ValueError('Energy resolution is already smaller '
                                 'than the requested resolution')
            num_moments = math.ceil((1.6 * spec._a) / energy_resolution)
        if num_moments is None or num_moments <= 0 or num_moments != int(num_moments):
            raise ValueError("'num_moments' must be a positive integer")
        if not self._accumulate:
            raise ValueError("Cannot increase the number of moments if "
                             "'accumulate_vectors' is 'False'.")
        spec.num_moments += num_moments
        self.num_moments = spec.num_moments
        count = self.num_vectors
        self.num_vectors = 0
        self._mu_sum = None
        self._mu_list = []
        self._add_vectors(count)
        self._build_integral_factor()

    def add_vectors(self, num_vectors=None):
        """Expand more vectors (kwant/kpm.py:663-683)."""
        self._check_live()
        self._spectrum_R._vector_factory.add_vectors(num_vectors)
        self._add_vectors(num_vectors)
        self._build_integral_factor()

    # -- GPU work ------------------------------------------------------------
    def _device_op(self, key, kind, payload):
        if kind != 'matrix':
            return None
        if key not in self._dev_ops:
            dmat = self._spectrum_R.dmat
            self._dev_ops[key] = eng.DeviceCSR(payload, dmat.precision, order=dmat.order)
        return self._dev_ops[key]

    def _apply(self, key, op, src, dst, width, n_slots):
        """dst[slot] = op src[slot] on the GPU (matrix), on the host
        (callable), or a copy (identity)."""
        import torch
        kind, payload = op
        if kind == 'id':
            dst.copy_(src)
        elif kind == 'matrix':
            self._device_op(key, kind, payload).spmm(
                src, dst, width, n_slots=n_slots, x_stride=src.stride(0),
                y_stride=dst.stride(0))
        else:
            dmat = self._spectrum_R.dmat
            host = src.cpu().numpy().astype(np.complex128)
            out = np.empty_like(host)
            for s in range(n_slots):
                caller = dmat.rows_out(host[s])
                res = np.empty_like(caller)
                for c in range(width):
                    res[:, c] = payload(caller[:, c])
                out[s] = dmat.rows_in(res)
            dst.copy_(torch.from_numpy(out).to(dst.device).to(dst.dtype))

    # Chebyshev vectors per ring of a streamed expansion, and rows of Omega
    # per GEMM: Psi stays resident (num_moments slots), each expansion runs
    # through a ring of this many slots that is drained as it fills
    OMEGA_CHUNK = 256
    # elements of padding after every ring slot and Psi row: the GEMM reads
    # 64 of them at the same k, and rows a power of two apart (N x W
    # elements) all fall on one L2 set and channel
    SLOT_PAD = 32

    def _stream(self, ab, start, width, m_count, consume):
        """Expand ``start`` (an (N, width) device block) to m_count Chebyshev
        vectors through a wrap-around ring, handing each run of consecutive
        vectors to consume(m0, m1, slots) -- slots: the (m1 - m0, N, width)
        view of the ring holding T_m0 .. T_{m1-1} -- in order."""
        chunk = max(3, min(m_count, self.OMEGA_CHUNK))
        rec = eng.BlockRecursion(self._spectrum_R.dmat, ab, start, width, eng.DOTS_NONE,
                                 store_slots=chunk, wrap=True, plan=self._plan_for(width),
                                 slot_pad=self.SLOT_PAD)
        consumed, produced = 0, 1
        while consumed < m_count:
            target = min(m_count, consumed + chunk)
            if target > produced:
                with self._timed('expansion_s'):
                    rec.step(target - produced, want_dots=False)
                produced = target
            m0 = consumed
            while m0 < target:
                s0 = m0 % chunk
                m1 = min(target, m0 + chunk - s0)
                consume(m0, m1, rec.ring[s0:s0 + (m1 - m0)])
                m0 = m1
            consumed = target
        del rec

    def _plan_for(self, width):
        """The launch plan of the expansions at this block width: one per
        width, shared by every expansion (creating a plan synchronises the
        device), with the autotuned launch shape unless ``variant`` was
        given."""
        plans = self.__dict__.setdefault('_plans', {})
        plan = plans.get(width)
        if plan is None:
            dmat = self._spectrum_R.dmat
            _, vt = eng.complex_dtype(dmat.precision)
            variant = self._variant if self._variant >= 0 else autotune.pick_variant(
                dmat, (self._a, self._b), width, eng.DOTS_NONE)
            plan = plans[width] = eng.Plan(dmat.shape[0], width, vt, variant=variant)
        return plan

    def _timed(self, phase):
        """A context that adds the GPU time of the enclosed launches to
        self.timings[phase] (read after the next synchronisation)."""
        import contextlib
        import torch

        @contextlib.contextmanager
        def span():
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            yield
            e1.record()
            self._events.append((phase, e0, e1))
        return span()

    def _collect_timings(self):
        import torch
        torch.cuda.synchronize()
        for phase, e0, e1 in self._events:
            self.timings[phase] = self.timings.get(phase, 0.0) + e0.elapsed_time(e1) / 1e3
        self._events = []

    def _add_vectors(self, count):
        import torch
        spec = self._spectrum_R
        factory = spec._vector_factory
        dmat = spec.dmat
        dev = eng.torch_device()
        cdt, _ = eng.complex_dtype(dmat.precision)
        n, m_count = dmat.shape[0], self.num_moments
        r0 = self.num_vectors
        if not hasattr(self, 'timings'):
            self.timings, self._events = {}, []
        widths = [1] * count if not self.mean else eng.pow2_widths(
            count, self._width_cap(n, m_count, cdt))
        self.block_widths = getattr(self, 'block_widths', []) + list(widths)
        gp = _GEMM[self._gemm_precision]

        def draw(first, width):
            # the factory's vectors in index order (the reference's RNG order)
            vecs = [factory[r] for r in range(first, first + width)]
            host = np.stack([np.asarray(v) for v in vecs], axis=1).astype(np.complex128)
            _check_finite_vectors(host)
            return np.ascontiguousarray(dmat.rows_in(host))

        # Psi storage for the widest block, reused by every block (views with
        # padded rows); one launch plan per block width, shared by all
        # expansions.  With the INT8 GEMM, Psi is kept only as its int8
        # slices (prepared once per block, reused by every chunk of Omega):
        # each run of Psi rows passes through a buffer of OMEGA_CHUNK rows.
        pad = self.SLOT_PAD
        prepared = isinstance(gp, tuple) and gp[0] == 'int8'
        # the FP64 / FP32 GEMMs read Psi and Omega in row panels (see _panels)
        panelled = not isinstance(gp, tuple)
        psi_slots = min(m_count, self.OMEGA_CHUNK) if prepared else m_count
        psi_len = max([n * w + pad for w in widths] + [1])
        if panelled:
            psi_len = max([len(self._panels(n, w, cdt)) * (self._panels(n, w, cdt)[0][1] * w + pad)
                           for w in widths] + [psi_len])
        psi_store = torch.empty(psi_slots * psi_len, dtype=cdt, device=dev)
        a_pan = None
        if panelled and widths:
            a_pan = torch.empty(min(m_count, self.OMEGA_CHUNK) * max(
                self._panels(n, w, cdt)[0][1] * w + pad for w in widths), dtype=cdt, device=dev)
        start = r0
        host_next = draw(start, widths[0]) if widths else None
        for bi, width in enumerate(widths):
            host = host_next
            start += width
            block = torch.from_numpy(host).to(dev).to(cdt)
            # Psi_n = op2 T_n(H~)|r>, resident for the whole block
            k = n * width
            if panelled:
                ranges = self._panels(n, width, cdt)
                kp = ranges[0][1] * width + pad
                psi_pan = psi_store[:len(ranges) * psi_slots * kp].view(len(ranges), psi_slots, kp)
                psi = psi_rows = None
            else:
                psi_rows = psi_store[:psi_slots * (k + pad)].view(psi_slots, k + pad)[:, :k]
                psi = psi_rows.view(psi_slots, n, width)
            if prepared:
                prep = getattr(self, '_psi_prep', None)
                if prep is None or prep.k != k:
                    self._psi_prep = prep = None
                    prep = self._psi_prep = eng.IsliceOperand(m_count, k, gp[1])

            def to_psi(m0, m1, slots):
                if panelled:
                    with self._timed('operator_s'):
                        self._apply_panels('op2', self._op2, slots, psi_pan, m0, ranges, width)
                elif prepared:
                    with self._timed('operator_s'):
                        self._apply('op2', self._op2, slots, psi[:m1 - m0], width, m1 - m0)
                    with self._timed('gemm_s'):
                        prep.fill_rows(psi_rows[:m1 - m0], m0)
                else:
                    with self._timed('operator_s'):
                        self._apply('op2', self._op2, slots, psi[m0:m1], width, m1 - m0)

            self._stream((self._a, self._b), block, width, m_count, to_psi)
            # the next block's vectors are drawn on the host while the GPU
            # runs this block's expansions
            if bi + 1 < len(widths):
                host_next = draw(start, widths[bi + 1])
            # Omega_m = T_m(H~) op1|r>, streamed in chunks against Psi
            start_l = torch.empty_like(block)
            with self._timed('operator_s'):
                self._apply('op1', self._op1, block.unsqueeze(0), start_l.unsqueeze(0), width, 1)
            mu = torch.zeros((m_count, m_count), dtype=torch.complex128, device=dev)
            psi_k = psi_rows

            def gemm(m0, m1, slots):
                cnt = m1 - m0
                if panelled:
                    # each row panel of the ring slots is copied into a
                    # buffer of short rows and multiplied by Psi's same panel
                    with self._timed('gemm_s'):
                        for p, (r0, r1) in enumerate(ranges):
                            kk = (r1 - r0) * width
                            a = a_pan[:cnt * kp].view(cnt, kp)[:, :kk]
                            a.copy_(slots[:, r0:r1, :].reshape(cnt, kk))
                            eng.corr_gemm(a, psi_pan[p, :m_count, :kk], mu[m0:m1], gp)
                    return
                # the ring slots as GEMM rows: k elements each, padded pitch
                a = slots.as_strided((cnt, k), (slots.stride(0), 1))
                with self._timed('gemm_s'):
                    if prepared:
                        prep.gemm(a, mu[m0:m1])
                    elif isinstance(gp, tuple):
                        self._gemm_work = eng.corr_gemm_sliced(
                            a, psi_k, mu[m0:m1], slices=gp[1], kind=gp[0],
                            work=getattr(self, '_gemm_work', None))
                    else:
                        eng.corr_gemm(a, psi_k, mu[m0:m1], gp)

            self._stream((self._aL, self._bL), start_l, width, m_count, gemm)
            del psi, psi_k, psi_rows
            if self.mean:
                self._mu_sum = mu if self._mu_sum is None else self._mu_sum + mu
            else:
                self._mu_list.append(mu.cpu().numpy())
        self.num_vectors = r0 + count
        if self.mean:
            self.moments_matrix = self._mu_sum.cpu().numpy() / self.num_vectors
        else:
            self.moments_matrix = np.array(self._mu_list)
        self._collect_timings()

    # bytes of a GEMM row panel: the FP64 / FP32 GEMMs slow down ~2.3x when
    # their operand rows are 128 MB apart (8 complex128 vectors of 2^20
    # sites) rather than 32 MB, so Psi and Omega reach them in row panels
    PANEL_BYTES = 1 << 25

    def _panels(self, n, width, cdt):
        """Row ranges [(r0, r1), ...] that cut an (n, width) block into panels
        of at most about PANEL_BYTES, all but the last of equal length."""
        import torch
        size = n * width * torch.empty((), dtype=cdt).element_size()
        count = max(1, -(-size // self.PANEL_BYTES))
        rows = -(-n // count)
        return [(r0, min(n, r0 + rows)) for r0 in range(0, n, rows)]

    def _apply_panels(self, key, op, src, dst, m0, ranges, width):
        """dst[p, m0 + slot, :(r1 - r0) width] = rows r0 .. r1-1 of op src[slot]
        for each panel p = (r0, r1): the operator's row blocks on the GPU
        (matrix), a copy (identity), or the host path a slot at a time."""
        import torch
        kind, payload = op
        cnt = src.shape[0]
        if kind == 'matrix':
            full = self._device_op(key, kind, payload)
            blocks = self.__dict__.setdefault('_row_blocks', {})
            for p, (r0, r1) in enumerate(ranges):
                sub = blocks.get((key, r0, r1))
                if sub is None:
                    sub = blocks[(key, r0, r1)] = full.row_block(r0, r1)
                sub.spmm(src, dst[p, m0:m0 + cnt], width, n_slots=cnt, x_stride=src.stride(0),
                         y_stride=dst.stride(1))
            return
        tmp = src if kind == 'id' else torch.empty_like(src[:1])
        for s in range(cnt if kind != 'id' else 1):
            if kind != 'id':
                self._apply(key, op, src[s:s + 1], tmp, width, 1)
            for p, (r0, r1) in enumerate(ranges):
                kk = (r1 - r0) * width
                rows = tmp[:, r0:r1, :].reshape(tmp.shape[0], kk)
                if kind == 'id':
                    dst[p, m0:m0 + cnt, :kk].copy_(rows)
                else:
                    dst[p, m0 + s, :kk].copy_(rows[0])

    def _width_cap(self, n, m_count, cdt):
        """Largest power-of-two block width (<= max_block_width) whose live
        vectors -- Psi (m_count slots) and one streaming ring of OMEGA_CHUNK
        slots, plus two single blocks -- take at most 3/4 of the free device
        memory less the GEMMs' panel buffer (OMEGA_CHUNK row panels)."""
        import torch
        free, _ = torch.cuda.mem_get_info()
        # memory that torch's allocator holds but does not use is free for us
        free += torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
        slots = m_count + min(m_count, self.OMEGA_CHUNK) + 2
        per_width = slots * n * torch.empty((), dtype=cdt).element_size()
        width = self._max_width
        budget = free * 3 // 4 - min(m_count, self.OMEGA_CHUNK) * (self.PANEL_BYTES + 65536)
        while width > 1 and width * per_width > budget:
            width //= 2
        if per_width > budget:
            raise MemoryError("one vector's %d Chebyshev vectors of %d sites do not fit in "
                              "device memory" % (m_count, n))
        return width

    def _build_integral_factor(self):
        """sum_mn mu~_mn Gamma_mn(x_k) at the 2M nodes (kwant/kpm.py:720-760)
        in GEMM form."""
        m_count = self.num_moments
        g = self._spectrum_R.kernel(np.ones(m_count))
        g[0] /= 2
        mu_k = np.outer(g, g) * self.moments_matrix
        x = (self.energies - self._b) / self._a
        self._integral_factor = integral_factor_gemm(mu_k, x)


# moment matrices at least this large (M) form the integral factor's two
# products on the GPU matrix cores; smaller ones on the host
GPU_FACTOR_MIN = 256


def _matmul_gpu(a, b):
    """a @ b for host complex arrays, as one FP64 matrix-core GEMM
    (kpm_corr_gemm computes mu += conj(A) B^T)."""
    import torch
    dev = eng.torch_device()
    ta = torch.from_numpy(np.ascontiguousarray(np.conj(a), dtype=np.complex128)).to(dev)
    tb = torch.from_numpy(np.ascontiguousarray(np.transpose(b), dtype=np.complex128)).to(dev)
    out = torch.zeros((a.shape[0], b.shape[1]), dtype=torch.complex128, device=dev)
    eng.corr_gemm(ta, tb, out, eng.GEMM_FP64)
    return out.cpu().numpy()


def integral_factor_gemm(mu_kernel, x):
    """factor_k = sum_i e+_ki (mu T^T)_ik + sum_j conj(e+_kj) (T mu)_kj with
    T_kn = cos(n arccos x_k) and e+_kn = (x_k - i n sqrt(1-x_k^2)) e^{i n arccos x_k};
    equal to the per-energy loop of kwant/kpm.py:745-760.  The two products
    run on the GPU for M >= GPU_FACTOR_MIN."""
    m_count = mu_kernel.shape[-1]
    m = np.arange(m_count)
    theta = np.arccos(x)[:, None] * m[None, :]
    t = np.cos(theta)
    e_plus = (x[:, None] - 1j * np.sqrt(1 - x ** 2)[:, None] * m[None, :]) * np.exp(1j * theta)
    if mu_kernel.ndim == 2:
        if m_count >= GPU_FACTOR_MIN:
            y = _matmul_gpu(mu_kernel, t.T)      # (M, 2M)
            z = _matmul_gpu(t, mu_kernel)        # (2M, M)
        else:
            y = mu_kernel @ t.T                  # (M, 2M)
            z = t @ mu_kernel                    # (2M, M)
        return np.sum(e_plus * y.T, axis=1) + np.sum(np.conj(e_plus) * z, axis=1)
    y = mu_kernel @ t.T                          # (R, M, 2M)
    z = t[None] @ mu_kernel                      # (R, 2M, M)
    return (np.sum(e_plus[None] * np.swapaxes(y, 1, 2), axis=2)
            + np.sum(np.conj(e_plus)[None] * z, axis=2))


def _velocity(hamiltonian, params, op_type, positions):
    """Velocity operator i (x_j - x_i) H_ij on the pattern of H, or the given
    operator (kwant/kpm.py:1042-1085)."""
    directions = dict(x=0, y=1, z=2)
    if _is_local_operator(op_type):
        return op_type
    if isinstance(op_type, str):
        direction = directions[op_type]
        if _is_system(hamiltonian):
            op, norbs, _ = hamiltonian.hamiltonian_submatrix(
                params=params, sparse=True, return_norb=True)
            positions = np.vstack([[hamiltonian.pos(i)] * nb for i, nb in enumerate(norbs)])
        else:
            op = hamiltonian
        positions = np.asarray(positions)
        if direction > positions.shape[1]:
            raise ValueError("{} is not an allowed direction.".format(op_type))
        if isinstance(op, csr_matrix) and op.has_canonical_format:
            # the same entries as the COO route below, without its sort
            x = positions[:, direction]
            rows = np.repeat(np.arange(op.shape[0]), np.diff(op.indptr))
            data = op.data * (1j * (x[op.indices] - x[rows]))
            return csr_matrix((data, op.indices.copy(), op.indptr.copy()), shape=op.shape)
        op = coo_matrix(op, copy=True)
        disp = positions[op.col] - positions[op.row]
        op.data = op.data * (1j * disp[:, direction])
        return op.tocsr()
    try:
        return csr_matrix(op_type)
    except Exception:
        raise ValueError("Velocity operator must be provided as a matrix, "
                         "a kwant operator, or a direction.")


def conductivity(hamiltonian, alpha='x', beta='x', positions=None, **kwargs):
    """Kubo-Bastin conductivity tensor element (kwant/kpm.py:763-847)."""
    if positions is None and not _is_system(hamiltonian):
        raise ValueError("If 'hamiltonian' is a matrix, positions must be provided")
    params = kwargs.get('params')
    alpha = _velocity(hamiltonian, params, alpha, positions)
    beta = _velocity(hamiltonian, params, beta, positions)
    return Correlator(hamiltonian, operator1=alpha, operator2=beta, **kwargs)
