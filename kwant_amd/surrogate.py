"""Moment surrogate for parameter sweeps: a small MLP ensemble in place of the
KPM recursion, with plain KPM as the fallback and as the accuracy reference.

The family is the square-lattice Hamiltonian H(W, phi, mu) of
:class:`kwant_amd.lattices.SquarePattern` (L x L sites, hopping 1 with a
Peierls flux phi per plaquette in Landau gauge, Anderson disorder W u_i,
u_i ~ U[-1/2, 1/2] with a fixed draw, uniform on-site shift mu).  A sweep
asks for the Jackson-damped (kwant/kpm.py:1106-1123), normalised DOS moments
g_n mu_n / mu_0 that kwant's SpectralDensity averages over random vectors
(kwant/kpm.py:436-477), in one spectral window common to the whole family
(bounds (-7, 7), so that shift and disorder show in the rescaled moments).

The shift mu enters exactly: H(W, phi, mu) = H(W, phi, 0) + mu, so the
moments in the common window are a fixed linear function (kpm_moment_shift)
of the moments of H(W, phi, 0) in a tighter window that contains every
unshifted spectrum (bounds (-6, 6)).  The identity holds per random vector,
so it reproduces the recursion's own estimate, not only its mean.  The
network therefore learns the (W, phi) dependence only; mu is applied by the
exact shift layer after it.

* ``SquareFamily`` computes moments with the GPU recursion for many parameter
  points at once: ``batch`` systems are the diagonal blocks of one CSR matrix,
  one segment each, so one launch per Chebyshev step advances all of them.
* ``MomentMLP`` is an ensemble of MLPs (GELU, linear output) whose forward,
  backward and Adam update run on the kernels of libkpm_amd (BF16 matrix
  cores, FP32 accumulation); all members train in the same launches.
* ``MomentSurrogate`` trains the ensemble on tight-window moments of
  H(W, phi, 0) and answers sweeps over (W, phi, mu): where the L1 norm of
  the members' pointwise density standard deviation exceeds ``tol`` (the
  spread gate) the point is recomputed with plain KPM.

DOS L1 distance of two damped moment vectors d = g (mu - mu'):
    L1 = int |rho - rho'| dx = (1/K) sum_k |d_0 + 2 sum_n d_n cos(n theta_k)|,
theta_k = pi (k + 1/2) / K: Gauss-Chebyshev quadrature, whose weight
cancels the density's 1 / (pi sqrt(1 - x^2)); it is exact for int rho itself
and, with K = 2M nodes, resolves |rho - rho'| at the kernel's resolution.
With mu_0 = 1 the distance is relative to int rho.
"""
import time

import numpy as np

from . import _engine as eng
from .kpm import jackson_kernel, _scale_from_bounds
from .lattices import SquarePattern

PARAM_NAMES = ("disorder", "flux", "shift")
PARAM_LO = np.array([0.0, 0.0, -1.0])
PARAM_HI = np.array([4.0, 0.02, 1.0])
LAYERS = (2, 256, 256, 256, 256, 512)
WIDE_BOUNDS = (-7.0, 7.0)     # common window of the sweep's moments
TIGHT_BOUNDS = (-6.0, 6.0)    # contains every spectrum of H(W <= 4, phi, 0)


def jackson_factors(count):
    """g_n of the Jackson kernel for ``count`` moments."""
    return jackson_kernel(np.ones(count))


def normalize_params(params, lo=PARAM_LO, hi=PARAM_HI):
    """Map parameter points (P, 3) from [lo, hi] to [-1, 1]."""
    params = np.asarray(params, dtype=np.float64)
    return 2.0 * (params - lo) / (hi - lo) - 1.0


def dos_l1_matrix(num_moments, nodes=None):
    """(K, M) matrix C with C[k, n] = w_n cos(n theta_k) / K, w_0 = 1,
    w_n = 2: the L1 distance of damped moments d is sum_k |(C d)_k|."""
    k = 2 * num_moments if nodes is None else nodes
    theta = np.pi * (np.arange(k) + 0.5) / k
    c = np.cos(np.outer(theta, np.arange(num_moments)))
    c[:, 1:] *= 2.0
    return c / k


def dos_l1(d_a, d_b, nodes=None):
    """Host L1 distance between densities given by damped moments (..., M)."""
    d = np.asarray(d_a, dtype=np.float64) - np.asarray(d_b, dtype=np.float64)
    c = dos_l1_matrix(d.shape[-1], nodes)
    return np.abs(d @ c.T).sum(axis=-1)


class SquareFamily:
    """Damped DOS moments of H(W, phi, mu) from the GPU recursion, ``batch``
    parameter points per launch.

    The random vectors are those kwant's SpectralDensity(rng=vector_seed)
    draws: one random_sample(N) for its Lanczos start vector
    (kwant/kpm.py:203), then one exp(2 pi i random_sample(N)) per vector
    (kwant/kpm.py:932-951), the same vectors for every parameter point."""

    def __init__(self, l=512, num_moments=512, num_vectors=32, bounds=WIDE_BOUNDS, eps=0.05,
                 disorder_seed=0, vector_seed=1, batch=8, precision="complex64", variant=-1):
        import torch
        if num_moments < 2 or num_moments % 2:
            raise ValueError("num_moments must be even and >= 2")
        dev = eng.torch_device()
        self.pattern = pat = SquarePattern(l, seed=disorder_seed)
        self.l, self.n, self.batch = l, pat.n, batch
        self.num_moments, self.num_vectors = num_moments, num_vectors
        self.a, self.b = _scale_from_bounds(bounds[0], bounds[1], eps)
        self.bounds = bounds
        n, nnz = pat.n, pat.nnz
        offs = np.arange(batch, dtype=np.int64)
        indptr = np.concatenate([(pat.indptr[None, :-1] + nnz * offs[:, None]).ravel(),
                                 [batch * nnz]])
        indices = (pat.indices[None, :] + n * offs[:, None]).ravel()
        if batch * nnz >= 2 ** 31:
            raise ValueError("batch too large for int32 CSR indices")
        self.cdt, vt = eng.complex_dtype(precision)
        self._indptr = torch.from_numpy(indptr.astype(np.int32)).to(dev)
        self._indices = torch.from_numpy(indices.astype(np.int32)).to(dev)
        self._values = torch.zeros(batch * nnz, dtype=self.cdt, device=dev)
        self.mat = eng.DeviceCSR.from_device_arrays((batch * n, batch * n), self._indptr,
                                                    self._indices, self._values, precision)
        self._peierls = torch.from_numpy(pat.peierls).to(dev)
        self._u = torch.from_numpy(pat.u_entry).to(dev)
        self._diag = torch.from_numpy(pat.is_diag).to(dev)
        rng = np.random.RandomState(vector_seed)
        rng.random_sample(n)
        vecs = np.stack([np.exp(2j * np.pi * rng.random_sample(n))
                         for _ in range(num_vectors)], axis=1)
        self._vectors = torch.from_numpy(vecs).to(dev).to(self.cdt)
        if variant < 0:
            from . import autotune
            variant = autotune.pick_variant(self.mat, (self.a, self.b), num_vectors,
                                            eng.DOTS_DOUBLING)
        self.plan = eng.Plan(batch * n, num_vectors, vt, seg_rows=n * np.arange(batch + 1),
                             variant=variant)
        self.plan.set_scale([self.a, self.b] * batch)
        self.ring = torch.empty((2, batch * n, num_vectors), dtype=self.cdt, device=dev)
        # every batch's vectors are unit-modulus phases, so <a_0|a_0> = n for
        # the seeded and the fresh ones alike: one grid serves them all
        self.ring[0].view(batch, n, num_vectors).copy_(
            self._vectors[None].expand(batch, n, num_vectors))
        self.plan.set_dots_grid(self.ring[0])
        self.steps = num_moments // 2
        self.g = jackson_factors(num_moments)

    def _load(self, params):
        """Write the values of the ``batch`` systems into the block matrix."""
        import torch
        p = torch.from_numpy(np.asarray(params, dtype=np.float64)).to(self._u.device)
        w, phi, mu = p[:, 0:1], p[:, 1:2], p[:, 2:3]
        hop = torch.polar(torch.ones_like(self._peierls),
                          2 * np.pi * phi * self._peierls[None, :])
        diag = (w * self._u[None, :] + mu).to(torch.complex128)
        vals = torch.where(self._diag[None, :], diag, hop)
        self._values.view(self.batch, -1).copy_(vals)

    def dots(self, params, fresh_seed=None):
        """Launch the recursion for all points; returns the FP64 dot records
        (batches, steps, batch, R, 4) on the device (asynchronous).

        fresh_seed=None: every point uses the seeded kwant vectors; otherwise
        every point gets its own random-phase vectors exp(2 pi i u) from the
        device generator seeded with ``fresh_seed`` (independent stochastic
        errors across points, as training data wants)."""
        import torch
        params = np.asarray(params, dtype=np.float64).reshape(-1, 3)
        gen = None
        if fresh_seed is not None:
            gen = torch.Generator(device=self._u.device)
            gen.manual_seed(int(fresh_seed))
        nb = -(-len(params) // self.batch)
        pad = np.concatenate([params, np.repeat(params[-1:], nb * self.batch - len(params), 0)])
        out = torch.empty((nb, self.steps, self.batch, self.num_vectors, 4),
                          dtype=torch.float64, device=self._u.device)
        n, r = self.n, self.num_vectors
        for i in range(nb):
            self._load(pad[i * self.batch:(i + 1) * self.batch])
            if gen is None:
                self.ring[0].view(self.batch, n, r).copy_(
                    self._vectors[None].expand(self.batch, n, r))
            else:
                u = torch.rand((self.batch * n, r), generator=gen, device=self._u.device,
                               dtype=torch.float64)
                self.ring[0].copy_(torch.polar(torch.ones_like(u), 2 * np.pi * u))
            self.plan.set_step(0)
            self.plan.steps(self.mat, self.ring, self.batch * n * r, 2, self.steps,
                            eng.DOTS_DOUBLING, dots=out[i])
        return out

    def moments(self, params, fresh_seed=None):
        """Normalised moments mu_n / mu_0 averaged over vectors, (P, M)."""
        params = np.asarray(params, dtype=np.float64).reshape(-1, 3)
        # the doubling identity is linear in the dots, so average over the
        # vectors first (kwant/kpm.py:470-473 per vector, then the mean)
        d = self.dots(params, fresh_seed).mean(dim=3).cpu().numpy()
        nb, steps, batch, _ = d.shape
        d = d.transpose(0, 2, 1, 3).reshape(nb * batch, steps, 4)[:len(params)]
        out = np.empty((len(params), self.num_moments))
        mu0 = d[:, 0, 0]
        out[:, 0] = mu0
        out[:, 1] = d[:, 0, 2]                     # Re conj(<a1|a0>)
        out[:, 2::2] = 2 * d[:, 1:, 0] - mu0[:, None]
        out[:, 3::2] = 2 * d[:, 1:, 2] - d[:, 0:1, 2]
        return out / mu0[:, None]

    def damped_moments(self, params, fresh_seed=None):
        """Jackson-damped normalised moments g_n mu_n / mu_0, (P, M)."""
        return self.moments(params, fresh_seed) * self.g


class MomentMLP:
    """An ensemble of ``members`` MLPs ``layers[0] -> ... -> layers[-1]``
    (GELU between layers, linear output) on the GPU.

    Parameters live in one flat FP32 buffer, layer by layer: W_l as
    (members, out, in) then b_l as (members, out); gradients and the two
    Adam moments have the same layout, so one Adam launch updates everything.
    Every GEMM is one batched launch over the members.  ``split`` lists, per
    layer, whether its forward product uses the three-term BF16 split; by
    default every forward product does (outputs accurate to ~1e-5 relative,
    which moment regression needs), while the backward pass runs plain BF16
    (gradients tolerate the rounding)."""

    def __init__(self, layers=LAYERS, members=5, seed=0, split=None):
        import torch
        dev = eng.torch_device()
        self.layers, self.members = tuple(layers), members
        self.split = [True] * (len(layers) - 1) if split is None else list(split)
        if len(self.split) != len(layers) - 1:
            raise ValueError("split needs one flag per layer")
        self.shapes, off = [], 0
        for fan_in, fan_out in zip(layers[:-1], layers[1:]):
            self.shapes.append((off, off + members * fan_out * fan_in, fan_in, fan_out))
            off += members * fan_out * (fan_in + 1)
        self.size = off
        flat = np.zeros(off, dtype=np.float32)
        for z in range(members):
            rng = np.random.RandomState(1000 * seed + z)
            for li, (w0, b0, fi, fo) in enumerate(self.shapes):
                gain = np.sqrt(2.0 / fi) if li < len(self.shapes) - 1 else np.sqrt(1.0 / fi)
                w = (rng.standard_normal((fo, fi)) * gain).astype(np.float32)
                flat[w0:b0].reshape(members, fo, fi)[z] = w
        self.params = torch.from_numpy(flat).to(dev)
        self.grads = torch.zeros_like(self.params)
        self.adam_m = torch.zeros_like(self.params)
        self.adam_v = torch.zeros_like(self.params)
        self.t = 0
        self._ws = {}

    def weight(self, li, buf=None):
        w0, b0, fi, fo = self.shapes[li]
        return (self.params if buf is None else buf)[w0:b0].view(self.members, fo, fi)

    def bias(self, li, buf=None):
        w0, b0, fi, fo = self.shapes[li]
        return (self.params if buf is None else buf)[b0:b0 + self.members * fo].view(
            self.members, fo)

    def _workspace(self, batch):
        import torch
        ws = self._ws.get(batch)
        if ws is None:
            dev, z = self.params.device, self.members
            ws = {"z": [torch.empty((z, batch, fo), dtype=torch.float32, device=dev)
                        for (_, _, _, fo) in self.shapes],
                  "a": [torch.empty((z, batch, fo), dtype=torch.float32, device=dev)
                        for (_, _, _, fo) in self.shapes[:-1]],
                  "g": [torch.empty((z, batch, fo), dtype=torch.float32, device=dev)
                        for (_, _, _, fo) in self.shapes[:-1]],
                  "dy": torch.empty((z, batch, self.layers[-1]), dtype=torch.float32,
                                    device=dev),
                  "partial": torch.empty(256, dtype=torch.float64, device=dev)}
            self._ws[batch] = ws
        return ws

    def forward(self, x):
        """x: (members or 1, B, layers[0]) float32; returns (members, B, out)
        (a workspace view, valid until the next call with this B)."""
        z, b, _ = x.shape
        if z == 1 and self.members > 1:
            x = x.expand(self.members, b, x.shape[2])
        ws = self._workspace(b)
        act = x
        last = len(self.shapes) - 1
        for li in range(len(self.shapes)):
            flags = eng.MLP_BF16X3 if self.split[li] else 0
            if li < last:
                eng.mlp_gemm(act, self.weight(li), ws["z"][li], bias=self.bias(li),
                             flags=eng.MLP_BIAS_GELU | flags, c2=ws["a"][li])
                act = ws["a"][li]
            else:
                eng.mlp_gemm(act, self.weight(li), ws["z"][li], bias=self.bias(li),
                             flags=eng.MLP_PLAIN | flags)
        self._x = x
        return ws["z"][last]

    def backward(self, dy):
        """Gradients of the loss whose output gradient is dy (members, B,
        out) for the inputs of the last forward call, into ``self.grads``."""
        b = dy.shape[1]
        ws = self._workspace(b)
        g = dy
        for li in range(len(self.shapes) - 1, -1, -1):
            prev = self._x if li == 0 else ws["a"][li - 1]
            eng.mlp_gemm(g.transpose(1, 2), prev.transpose(1, 2), self.weight(li, self.grads))
            eng.mlp_colsum(g, self.bias(li, self.grads))
            if li > 0:
                eng.mlp_gemm(g, self.weight(li).transpose(1, 2), ws["g"][li - 1],
                             flags=eng.MLP_GELU_GRAD, aux=ws["z"][li - 1])
                g = ws["g"][li - 1]

    def loss_grad(self, y, target):
        """dy = d(mean squared error per member)/dy into the workspace;
        returns (dy, FP64 partial sums of squared errors)."""
        ws = self._workspace(y.shape[1])
        scale = 2.0 / (y.shape[1] * y.shape[2])
        eng.mlp_mse(y, target, ws["dy"], scale, ws["partial"])
        return ws["dy"], ws["partial"]

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8, sched=None, ema=None):
        """One Adam step; with ``sched`` (device lr, bias corrections, EMA
        decay) the step reads its schedule from device memory (graph
        replay); with ``ema`` the parameters' moving average is updated."""
        self.t += 1
        eng.mlp_adam(self.params, self.grads, self.adam_m, self.adam_v, lr, b1, b2, eps, self.t,
                     sched=sched, ema=ema)


class MomentSurrogate:
    """Ensemble surrogate of the damped moments of H(W, phi, mu) in the
    window of ``family`` (the plain-KPM family that also serves the
    fallback), learned from moments of H(W, phi, 0) in the tighter window of
    ``train_family`` and carried to any mu by the exact shift."""

    def __init__(self, family, train_family, members=5, layers=LAYERS, seed=0, tol=1e-3,
                 split=None):
        import torch
        m = family.num_moments
        layers = tuple(layers)
        # layers[0] == 2: (W, phi) -> tight-window moments, mu by the exact
        # shift; layers[0] == 3: the literal network (W, phi, mu) -> the
        # wide-window moments, trained on H(W, phi, mu) itself
        self.literal = layers[0] == 3
        if layers[0] not in (2, 3) or layers[-1] != m or train_family.num_moments != m:
            raise ValueError("layers must map (W, phi) or (W, phi, mu) to the family's moments")
        if self.literal:
            train_family = family
        self.family, self.train_family = family, train_family
        self.tol = tol
        self.alpha = train_family.a / family.a
        # H(W, phi, mu) in the wide window = alpha t + beta, t the tight variable
        self._beta0 = (train_family.b - family.b) / family.a
        # the largest |mu| whose shifted window stays inside the tight one
        self.max_shift = (1.0 - self.alpha) * family.a - abs(train_family.b - family.b)
        self.mlp = MomentMLP(layers, members, seed, split)
        self.active = members
        dev = self.mlp.params.device
        self._in_scale = torch.from_numpy(1.0 / train_family.g).to(dev)
        self._out_scale = torch.from_numpy(family.g).to(dev)
        c = dos_l1_matrix(m)
        # node densities c[k, n] d_n, row-scaled by K so that values are O(1)
        self._l1 = torch.from_numpy((c * c.shape[0]).astype(np.float32)).to(dev)
        self.history = []

    def _inputs(self, params):
        """Normalised (W, phi) -- or (W, phi, mu) for the literal network --
        of parameter rows (P, 2) or (P, 3)."""
        import torch
        d = 3 if self.literal else 2
        p = np.atleast_2d(np.asarray(params, dtype=np.float64))[:, :d]
        x = normalize_params(p, PARAM_LO[:d], PARAM_HI[:d]).astype(np.float32)
        return torch.from_numpy(x).to(self.mlp.params.device)

    def training_targets(self, params, fresh_seed=1):
        """Damped tight-window moments of H(W, phi, 0) for params (P, >= 2)
        (any shift column is ignored) -- for the literal network the
        wide-window moments of H(W, phi, mu) -- from the GPU recursion with
        fresh random vectors per point."""
        p = np.asarray(params, dtype=np.float64)
        if self.literal:
            return self.family.damped_moments(p[:, :3], fresh_seed=fresh_seed)
        p = np.concatenate([p[:, :2], np.zeros((len(p), 1))], axis=1)
        return self.train_family.damped_moments(p, fresh_seed=fresh_seed)

    def in_range(self, params):
        """Points the surrogate answers: (W, phi) inside the training box and,
        for the shift design, |mu| within the exact shift's range."""
        p = np.atleast_2d(np.asarray(params, dtype=np.float64)).reshape(-1, 3)
        ok = np.all((p >= PARAM_LO - 1e-12) & (p <= PARAM_HI + 1e-12), axis=1)
        if not self.literal:
            ok &= np.abs(p[:, 2]) <= self.max_shift + 1e-12
        return ok

    def fit(self, params, targets, epochs=400, batch=64, lr=3e-3, lr_min=3e-6, seed=0,
            log=None, graph=True, b1=0.9, b2=0.999, ema_decay=None, ema_start=0.5,
            warmup=0.1, health=10.0, collapse=0.5):
        """Train every member on (params, targets) with Adam and a cosine
        learning-rate schedule; members draw their own mini-batch order.

        An epoch is a fixed sequence of launches over fixed buffers (the
        epoch's shuffled copy of the data, one schedule row per iteration), so
        with ``graph`` it is captured once as a HIP graph and replayed every
        epoch; the host only reshuffles and writes the schedule between
        replays.  graph=False runs the same launches eagerly (bitwise the same
        result).  With ``ema_decay`` the members end with the exponential
        moving average of their parameters over the iterations after the
        fraction ``ema_start`` of training (Polyak averaging), which smooths
        the last optimiser noise out of the fit.  The learning rate ramps up
        linearly over the first ``warmup`` fraction of the iterations.  After
        training, a member whose training error exceeds ``health`` times the
        members' median is left out of the ensemble statistics (a member
        stuck in a poor minimum would otherwise set the spread everywhere),
        and so is a member whose error exceeds ``collapse`` times that of the
        best constant prediction (a collapsed network: when every member
        collapses to the same constant their spread is zero, so the spread
        gate alone would accept it).  With fewer than two members left every
        point falls back to plain KPM."""
        import torch
        dev = self.mlp.params.device
        x = self._inputs(params)
        t = torch.from_numpy(np.asarray(targets, dtype=np.float32)).to(dev)
        count, z = len(x), self.mlp.members
        per_epoch = max(1, count // batch)
        used = per_epoch * batch
        total = epochs * per_epoch
        # iteration s reads xs[s], ts[s]: contiguous (members, batch, features)
        xs = torch.empty((per_epoch, z, batch, x.shape[1]), dtype=torch.float32, device=dev)
        ts = torch.empty((per_epoch, z, batch, t.shape[1]), dtype=torch.float32, device=dev)
        sched = torch.empty((per_epoch, 4), dtype=torch.float32, device=dev)
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        ema = torch.zeros_like(self.mlp.params) if ema_decay else None

        def epoch_body():
            for s in range(per_epoch):
                y = self.mlp.forward(xs[s])
                dy, partial = self.mlp.loss_grad(y, ts[s])
                self.mlp.backward(dy)
                self.mlp.adam(0.0, b1, b2, sched=sched[s], ema=ema)
            return partial

        g = None
        if graph:
            # capture reads only the buffers above; nothing is allocated or cleared inside
            self.mlp._workspace(batch)
            g = torch.cuda.CUDAGraph()
            t_saved = self.mlp.t
            with torch.cuda.graph(g):
                last_partial = epoch_body()
            self.mlp.t = t_saved
        it = 0
        for ep in range(epochs):
            order = torch.argsort(torch.rand((z, count), generator=gen, device=dev), dim=1)
            order = order[:, :used].reshape(z, per_epoch, batch).transpose(0, 1)
            xs.copy_(x[order])
            ts.copy_(t[order])
            rows = np.empty((per_epoch, 4), dtype=np.float32)
            for s in range(per_epoch):
                frac = (it + s) / max(1, total - 1)
                step = self.mlp.t + s + 1
                decay = float(ema_decay) if ema_decay and frac >= ema_start else 0.0
                rate = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * frac))
                if warmup and frac < warmup:
                    rate *= 0.01 + 0.99 * frac / warmup
                rows[s] = (rate, 1.0 / (1.0 - b1 ** step), 1.0 / (1.0 - b2 ** step), decay)
            sched.copy_(torch.from_numpy(rows))
            if g is not None:
                g.replay()
                self.mlp.t += per_epoch
            else:
                last_partial = epoch_body()
            it += per_epoch
            if log is not None and (ep % max(1, epochs // 10) == 0 or ep == epochs - 1):
                mse = float(last_partial.sum()) / (batch * z * t.shape[1])
                self.history.append((ep, mse))
                log("epoch %d: mse %.3e (last mini-batch)" % (ep, mse))
        if ema is not None:
            self.mlp.params.copy_(ema)
        self._select_members(x, t, health, log, collapse)

    def _select_members(self, x, t, health, log=None, collapse=0.5):
        """Order the members by training error and keep those within
        ``health`` times the median and within ``collapse`` times the error
        of the best constant prediction (the targets' variance; not checked
        when the targets are constant)."""
        import torch
        count = min(len(x), 1024)
        y = self.mlp.forward(x[None, :count])
        err = ((y - t[None, :count]) ** 2).mean(dim=(1, 2)).double().cpu().numpy()
        tc = t[:count].double()
        base = float(((tc - tc.mean(dim=0)) ** 2).mean())
        med = float(np.median(err))
        order = np.argsort(err)
        healthy = err <= health * med
        if base > 1e-12:
            healthy &= err <= collapse * base
        keep = int(np.sum(healthy))
        self.constant_error = base
        self.member_errors = err
        self.active = keep
        if keep < self.mlp.members or np.any(order != np.arange(len(order))):
            idx = torch.from_numpy(order).to(self.mlp.params.device)
            for buf in (self.mlp.params, self.mlp.adam_m, self.mlp.adam_v, self.mlp.grads):
                for li in range(len(self.mlp.shapes)):
                    for view in (self.mlp.weight(li, buf), self.mlp.bias(li, buf)):
                        view.copy_(view[idx])
        if log is not None:
            log("members' training MSE %s (constant prediction %.2e); %d of %d kept"
                % (np.array2string(err, precision=2), base, keep, len(err)))

    def predict(self, params, chunk=4096, members=False):
        """Ensemble mean of the damped wide-window moments (P, M) and the
        spread (P,) on the device: the spread is the L1 norm (over energy)
        of the pointwise standard deviation of the members' densities.
        members=True also returns the per-member moments (Z, P, M)."""
        import torch
        p = np.asarray(params, dtype=np.float64).reshape(-1, 3)
        if not self.literal and np.any(np.abs(p[:, 2]) > self.max_shift + 1e-12):
            raise ValueError("shift outside the exact range |mu| <= %.4f" % self.max_shift)
        dev = self.mlp.params.device
        x = self._inputs(p)
        beta = torch.from_numpy(self._beta0 + p[:, 2] / self.family.a).to(dev)
        count, m, z = len(x), self.family.num_moments, self.mlp.members
        mean = torch.empty((count, m), dtype=torch.float32, device=dev)
        spread = torch.empty(count, dtype=torch.float32, device=dev)
        per = torch.empty((z, count, m), dtype=torch.float32, device=dev) if members else None
        k = self._l1.shape[0]
        step = min(chunk, count)
        shifted = torch.empty((z, step, m), dtype=torch.float32, device=dev)
        nodes = torch.empty((z, step, k), dtype=torch.float32, device=dev)
        for s in range(0, count, chunk):
            e = min(count, s + chunk)
            y = self.mlp.forward(x[None, s:e])
            if self.literal:
                sh = y                       # already the wide-window moments
            else:
                sh = shifted[:, :e - s]
                eng.moment_shift(y, self.alpha, beta[s:e], sh, in_scale=self._in_scale,
                                 out_scale=self._out_scale)
            if per is not None:
                per[:, s:e].copy_(sh)
            if self.active == 0:
                continue
            act = sh[:self.active]
            eng.mlp_ensemble(act, mean=mean[s:e])
            if self.active < 2:
                continue
            nd = nodes[:self.active, :e - s]
            eng.mlp_gemm(act, self._l1[None].expand(self.active, k, m), nd,
                         flags=eng.MLP_BF16X3)
            eng.mlp_ensemble(nd, dev=spread[s:e])
        if self.active == 0:
            mean.fill_(float("nan"))        # no healthy member: nothing to predict
        if self.active < 2:
            spread.fill_(float("inf"))      # no spread without two healthy members
        if members:
            return mean, spread, per
        return mean, spread

    def sweep(self, params):
        """Damped moments for every parameter point: the ensemble mean where
        the spread is within ``tol``, plain KPM (``family``) elsewhere.
        Points outside the surrogate's range (``in_range``) fall back too.
        Returns (moments (P, M) FP64 host array, fallback mask, seconds)."""
        import torch
        params = np.asarray(params, dtype=np.float64).reshape(-1, 3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        inside = self.in_range(params)
        out = np.zeros((len(params), self.family.num_moments))
        fall = ~inside
        if inside.any():
            mean, spread = self.predict(params[inside])
            out[inside] = mean.double().cpu().numpy()
            fall[inside] = spread.cpu().numpy() > self.tol
        if fall.any():
            out[fall] = self.family.damped_moments(params[fall])
        torch.cuda.synchronize()
        return out, fall, time.perf_counter() - t0
