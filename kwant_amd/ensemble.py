"""Batched disorder ensembles: many systems of uneven size advanced by one
kernel launch per Chebyshev step.

The systems are the diagonal blocks of one CSR matrix, generated on the GPU
(``kpm_gen_cubic4``: no host assembly or upload) and run with one recursion
plan whose segments are the systems.  Each segment has its own rescaling
(a, b) from its own Gershgorin radius and its own FP64 dot products, so one
launch per step produces every system's moments, which kwant would compute
with one ``SpectralDensity`` per system (kwant/kpm.py:163-252, the recursion
of kwant/kpm.py:436-477 looping over systems and vectors).  The whole step
sequence of an expansion is captured once as a HIP graph and replayed.
"""
import time

import numpy as np

from . import _engine as eng
from .kpm import _scale_from_bounds, jackson_kernel


def ensemble_sizes(count, lo=20, hi=60, seed=0):
    """Side lengths L ~ U{lo..hi} of ``count`` systems (numpy MT19937)."""
    return np.random.RandomState(seed).randint(lo, hi + 1, size=count)


class CubicEnsemble:
    """Disorder ensemble of cubic L x L x L lattices with 4 orbitals per site
    (random Hermitian on-site blocks, entries sigma N(0, 1) + i sigma N(0, 1)
    symmetrised, hopping t I_4 to the 6 neighbours), one system per side
    length in ``sizes``, ``num_vectors`` random-phase vectors each."""

    def __init__(self, sizes, seed=0, sigma=0.3, t=1.0, num_moments=512, num_vectors=16,
                 precision="complex64", eps=0.05, variant=-1, vector_seed=1,
                 matrix_format="csr"):
        """``matrix_format``: 'csr' (the systems' CSR, the general recursion
        kernel) or 'stencil' (only the on-site blocks; the lattice-stencil
        kernel marches each system along x, bitwise the same moments).
        ``variant``: the CSR kernel's launch shape, or for 'stencil' the shape
        of the stencil kernel (``_engine.MarchPlan.SHAPES``: waves per
        workgroup, prefetch depth, tile of sites); -1 picks by timing."""
        import torch
        if matrix_format not in ("csr", "stencil"):
            raise ValueError("matrix_format must be 'csr' or 'stencil'")
        if num_moments < 2 or num_moments % 2:
            raise ValueError("num_moments must be even and >= 2")
        dev = eng.torch_device()
        self._config = dict(sizes=[int(x) for x in np.asarray(sizes).ravel()], seed=int(seed),
                            sigma=float(sigma), t=float(t), num_moments=int(num_moments),
                            num_vectors=int(num_vectors), precision=precision, eps=float(eps),
                            vector_seed=int(vector_seed), matrix_format=matrix_format)
        self.matrix_format = matrix_format
        self.t = float(t)
        self._ran = False
        self.sizes = np.asarray(sizes, dtype=np.int64)
        if self.sizes.ndim != 1 or len(self.sizes) == 0 or self.sizes.min() < 1:
            raise ValueError("sizes must be a non-empty list of positive side lengths")
        self.num_moments, self.num_vectors = num_moments, num_vectors
        self.rows = 4 * self.sizes ** 3
        self.nnzs = np.array([eng.cubic4_nnz(int(L)) for L in self.sizes], dtype=np.int64)
        self.row_off = np.concatenate([[0], np.cumsum(self.rows)])
        self.nz_off = np.concatenate([[0], np.cumsum(self.nnzs)])
        n, nnz = int(self.row_off[-1]), int(self.nz_off[-1])
        if n >= 2 ** 31 or nnz >= 2 ** 31:
            raise ValueError("ensemble too large for int32 CSR indices")
        self.n, self.nnz = n, nnz
        self.cdt, vt = eng.complex_dtype(precision)
        radius = torch.zeros(len(self.sizes), dtype=torch.int32, device=dev)
        self.seeds = [(int(seed) << 20) + s for s in range(len(self.sizes))]
        self.mat = self.onsite = self.march = None
        if matrix_format == "csr":
            self.indptr = torch.empty(n + 1, dtype=torch.int32, device=dev)
            self.indices = torch.empty(nnz, dtype=torch.int32, device=dev)
            self.values = torch.empty(nnz, dtype=self.cdt, device=dev)
            for s, L in enumerate(self.sizes):
                eng.gen_cubic4(int(L), self.seeds[s], sigma, t, int(self.row_off[s]),
                               int(self.nz_off[s]), self.indptr, self.indices, self.values,
                               radius[s:s + 1])
            self.mat = eng.DeviceCSR.from_device_arrays((n, n), self.indptr, self.indices,
                                                        self.values, precision)
        else:
            self.onsite = torch.empty((n, 4), dtype=self.cdt, device=dev)
            for s, L in enumerate(self.sizes):
                eng.gen_cubic4_onsite(int(L), self.seeds[s], sigma, t, int(self.row_off[s]),
                                      self.onsite, radius[s:s + 1])
        self.radius = radius.cpu().numpy().view(np.float32).astype(np.float64)
        self.scales = np.array([_scale_from_bounds(-r, r, eps) for r in self.radius])
        if matrix_format == "csr" and variant < 0:
            from . import autotune
            variant = autotune.pick_variant(self.mat, tuple(self.scales[0]), num_vectors,
                                            eng.DOTS_DOUBLING)
        plan_variant = variant if matrix_format == "csr" else 0
        self.plan = eng.Plan(n, num_vectors, vt, seg_rows=self.row_off.astype(np.int32),
                             variant=plan_variant)
        self.plan.set_scale(self.scales.ravel())
        self.ring = torch.empty((2, n, num_vectors), dtype=self.cdt, device=dev)
        self.vectors = torch.empty((n, num_vectors), dtype=self.cdt, device=dev)
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(vector_seed))
        chunk = 1 << 22
        for r0 in range(0, n, chunk):
            r1 = min(n, r0 + chunk)
            u = torch.rand((r1 - r0, num_vectors), generator=gen, device=dev,
                           dtype=torch.float64)
            self.vectors[r0:r1] = torch.polar(torch.ones_like(u), 2 * np.pi * u)
        self.plan.set_dots_grid(self.vectors)
        self.steps = num_moments // 2
        self.dots = torch.empty((self.steps, len(self.sizes), num_vectors, 4),
                                dtype=torch.float64, device=dev)
        self._graph = None
        if matrix_format == "stencil":
            if variant < 0:
                variant = self._pick_shape(vt)
            self.march = eng.MarchPlan(self.sizes, num_vectors, vt, shape=variant)
        self.variant = variant

    def _pick_shape(self, vt):
        """The stencil kernel's shape (``MarchPlan.SHAPES``): the fastest over
        a few steps of this ensemble, the three fastest timed again three
        times longer (below 2^20 vector elements, shape 0)."""
        if self.n * self.num_vectors < (1 << 20):
            return 0
        shapes = [w for w in range(len(eng.MarchPlan.SHAPES))
                  if eng.MarchPlan.shape_ok(self.num_vectors, vt, w)]
        times = {w: self._time_shape(vt, w, 6) for w in shapes}
        final = {w: self._time_shape(vt, w, 18)
                 for w in sorted(times, key=lambda w: times[w])[:3]}
        times.update(final)
        self.shape_ms = times
        return min(final, key=lambda w: final[w])

    def _time_shape(self, vt, shape, count):
        """Milliseconds per step of ``count`` steps (at most the expansion's)
        with the stencil kernel's ``shape``, after two untimed steps."""
        import torch
        self.march = eng.MarchPlan(self.sizes, self.num_vectors, vt, shape=shape)
        count = min(count, self.steps)
        for run in range(2):
            self.ring[0].copy_(self.vectors)
            self.plan.set_step(0)
            torch.cuda.synchronize()
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record()
            self._steps_on(min(2, count) if run == 0 else count)
            end.record()
            torch.cuda.synchronize()
        return start.elapsed_time(end) / count

    def _steps_on(self, count):
        """``count`` steps from the plan's counter (dots into self.dots)."""
        if self.march is not None:
            self.march.steps(self.plan, self.onsite, self.t, self.ring,
                             self.n * self.num_vectors, 2, count, eng.DOTS_DOUBLING,
                             dots=self.dots)
        else:
            self.plan.steps(self.mat, self.ring, self.n * self.num_vectors, 2, count,
                            eng.DOTS_DOUBLING, dots=self.dots)

    def bytes_per_step(self):
        """Algorithmic HBM bytes of one step with the CSR: a_n, a_{n-1} read,
        a_{n+1} written, the CSR read once (the same for both formats: the
        path's bytes per step)."""
        s_v = 8 if self.cdt.itemsize == 8 else 16
        return 3 * self.num_vectors * self.n * s_v + self.nnz * (s_v + 4) + 4 * (self.n + 1)

    def stencil_bytes_per_step(self):
        """The stencil kernel's own bytes per step: the three vector passes
        and the on-site blocks (4 complex per row)."""
        s_v = 8 if self.cdt.itemsize == 8 else 16
        return 3 * self.num_vectors * self.n * s_v + 4 * self.n * s_v

    def _launch(self):
        self.plan.set_step(0)
        self._steps_on(self.steps)

    def run(self, graph=True):
        """One expansion of every system; returns the seconds from the first
        launch to completion (the moments stay on the device, see moments()).
        With ``graph`` the step sequence is captured on the first call and
        replayed afterwards."""
        import torch
        self.ring[0].copy_(self.vectors)
        if graph and self._graph is None:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._launch()
            self._graph = g
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if graph:
            self._graph.replay()
        else:
            self._launch()
        torch.cuda.synchronize()
        self.plan.set_step(self.steps)
        self._ran = True
        return time.perf_counter() - t0

    def save(self, path):
        """Write the ensemble's definition and its per-step FP64 dots to
        ``path`` (.npz with a JSON header, nothing pickled).  The systems and
        vectors are not stored: ``load`` regenerates them from the seeds, and
        the generator is deterministic."""
        from . import persist
        if not self._ran:
            raise ValueError("run() the ensemble before saving its moments")
        persist.save_arrays(path, persist.ENSEMBLE_FORMAT,
                            dict(self._config, variant=int(self.variant)),
                            dict(dots=self.dots.cpu().numpy()))

    @classmethod
    def load(cls, path):
        """An ensemble saved by ``save``: systems regenerated on the GPU,
        moments restored (moments() works without running again)."""
        import torch
        from . import persist
        meta, arrays = persist.load_arrays(path, persist.ENSEMBLE_FORMAT)
        cfg = {k: meta[k] for k in ("sizes", "seed", "sigma", "t", "num_moments",
                                     "num_vectors", "precision", "eps", "vector_seed",
                                     "variant")}
        cfg["matrix_format"] = meta.get("matrix_format", "csr")
        ens = cls(**cfg)
        dots = torch.from_numpy(arrays["dots"])
        if tuple(dots.shape) != tuple(ens.dots.shape):
            raise ValueError("saved moments do not match the ensemble definition")
        ens.dots.copy_(dots.to(ens.dots.device))
        ens.plan.set_step(ens.steps)
        ens._ran = True
        return ens

    def moments(self, damped=True):
        """Per-system moments mu_n / mu_0 averaged over the vectors, (S, M);
        Jackson-damped (kwant/kpm.py:1106-1123) unless damped=False."""
        d = self.dots.mean(dim=2).cpu().numpy()   # (steps, S, 4)
        m = np.empty((d.shape[1], self.num_moments))
        mu0 = d[0, :, 0]
        m[:, 0] = mu0
        m[:, 1] = d[0, :, 2]
        m[:, 2::2] = (2 * d[1:, :, 0] - mu0[None, :]).T
        m[:, 3::2] = (2 * d[1:, :, 2] - d[0:1, :, 2]).T
        m /= mu0[:, None]
        if damped:
            m = jackson_kernel(m.T).T
        return m
