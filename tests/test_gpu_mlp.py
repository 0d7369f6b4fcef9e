"""GPU kernels of the moment surrogate against the CPU oracle
(oracle/mlp_oracle.py), and the batched square-lattice moment generator
against the KPM oracle (oracle/kpm_oracle.py)."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _t(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def _close(got, want, tol):
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    scale = max(np.max(np.abs(want)), 1e-30)
    err = np.max(np.abs(got - want)) / scale
    assert err < tol, err


@pytest.mark.parametrize("m,n,k", [(64, 64, 32), (70, 130, 33), (5, 3, 300), (257, 96, 3)])
@pytest.mark.parametrize("split", [False, True])
def test_gemm_plain(m, n, k, split):
    """C = A B^T (+ bias) with BF16 operands and FP32 accumulation."""
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(m + n + k)
    z = 3
    a = rng.standard_normal((z, m, k)).astype(np.float32)
    b = rng.standard_normal((z, n, k)).astype(np.float32)
    bias = rng.standard_normal((z, n)).astype(np.float32)
    c = torch.full((z, m, n), np.nan, dtype=torch.float32, device="cuda")
    eng.mlp_gemm(_t(a), _t(b), c, bias=_t(bias), flags=eng.MLP_BF16X3 if split else 0)
    got = c.cpu().numpy()
    for i in range(z):
        want = mo.gemm(a[i], b[i], split) + bias[i]
        _close(got[i], want, 2e-6 if not split else 1e-5)
        if split:  # the split product is close to the FP32 product
            _close(got[i], a[i].astype(np.float64) @ b[i].T.astype(np.float64) + bias[i], 1e-4)


def test_gemm_strides_broadcast_accumulate():
    """Transposed operands (unit stride along rows), a broadcast operand
    (stride 0 over the batch) and accumulation into C."""
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(7)
    z, m, n, k = 2, 50, 40, 77
    at = rng.standard_normal((z, k, m)).astype(np.float32)     # A^T stored
    bt = rng.standard_normal((k, n)).astype(np.float32)        # B^T stored, shared
    c0 = rng.standard_normal((z, m, n)).astype(np.float32)
    c = _t(c0)
    a_view = _t(at).transpose(1, 2)
    b_view = _t(bt).t()[None].expand(z, n, k)
    eng.mlp_gemm(a_view, b_view, c, flags=eng.MLP_ACCUMULATE)
    got = c.cpu().numpy()
    for i in range(z):
        _close(got[i], c0[i] + mo.gemm(at[i].T, bt.T), 2e-6)


def test_gemm_epilogues():
    """BIAS_GELU writes Z and gelu(Z); GELU_GRAD multiplies by gelu'(aux)."""
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(11)
    z, m, n, k = 2, 33, 65, 40
    a = rng.standard_normal((z, m, k)).astype(np.float32)
    b = (rng.standard_normal((z, n, k)) / 6).astype(np.float32)
    bias = rng.standard_normal((z, n)).astype(np.float32)
    aux = rng.standard_normal((z, m, n)).astype(np.float32)
    zc = torch.empty((z, m, n), dtype=torch.float32, device="cuda")
    ac = torch.empty_like(zc)
    eng.mlp_gemm(_t(a), _t(b), zc, bias=_t(bias), flags=eng.MLP_BIAS_GELU, c2=ac)
    gc = torch.empty_like(zc)
    eng.mlp_gemm(_t(a), _t(b), gc, flags=eng.MLP_GELU_GRAD, aux=_t(aux))
    for i in range(z):
        want = mo.gemm(a[i], b[i]) + bias[i]
        _close(zc[i].cpu().numpy(), want, 2e-6)
        _close(ac[i].cpu().numpy(), mo.gelu(want), 2e-6)
        _close(gc[i].cpu().numpy(), mo.gemm(a[i], b[i]) * mo.gelu_grad(aux[i]), 2e-6)


def test_gemm_argument_errors():
    import torch
    from kwant_amd import _engine as eng
    a = torch.zeros((1, 4, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        eng.mlp_gemm(a, a, a, flags=eng.MLP_BIAS_GELU)            # no activation output
    with pytest.raises(ValueError):
        eng.mlp_gemm(a, a, a, flags=eng.MLP_GELU_GRAD)            # no pre-activation
    with pytest.raises(ValueError):
        eng.mlp_gemm(a, a, a, flags=eng.MLP_BIAS_GELU | eng.MLP_ACCUMULATE, c2=a)
    with pytest.raises(ValueError):
        eng.mlp_gemm(a, a, a, flags=64)


def test_colsum_mse_ensemble_adam():
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(5)
    g = rng.standard_normal((3, 37, 70)).astype(np.float32)
    out = torch.empty((3, 70), dtype=torch.float32, device="cuda")
    eng.mlp_colsum(_t(g), out)
    _close(out.cpu().numpy(), g.astype(np.float64).sum(axis=1), 1e-6)

    y = rng.standard_normal(5000).astype(np.float32)
    t = rng.standard_normal(5000).astype(np.float32)
    dy = torch.empty(5000, dtype=torch.float32, device="cuda")
    part = torch.empty(7, dtype=torch.float64, device="cuda")
    eng.mlp_mse(_t(y), _t(t), dy, 2.0 / 5000, part)
    loss, grad = mo.mse(y, t)
    _close(dy.cpu().numpy(), grad, 1e-6)
    assert abs(float(part.sum()) / 5000 - loss) < 1e-6 * loss

    ens = rng.standard_normal((5, 21, 130)).astype(np.float32)
    mean = torch.empty((21, 130), dtype=torch.float32, device="cuda")
    dev = torch.empty(21, dtype=torch.float32, device="cuda")
    eng.mlp_ensemble(_t(ens), mean=mean, dev=dev)
    m_ref = ens.astype(np.float64).mean(axis=0)
    _close(mean.cpu().numpy(), m_ref, 1e-6)
    _close(dev.cpu().numpy(), ens.astype(np.float64).std(axis=0, ddof=1).mean(axis=1), 1e-5)

    p = rng.standard_normal(1000).astype(np.float32)
    pm, pv = np.zeros_like(p), np.zeros_like(p)
    tp, tm, tv = _t(p), _t(pm), _t(pv)
    for step in range(1, 4):
        gr = rng.standard_normal(1000).astype(np.float32)
        eng.mlp_adam(tp, _t(gr), tm, tv, 1e-2, 0.9, 0.999, 1e-8, step)
        p, pm, pv = mo.adam(p.astype(np.float64), gr, pm, pv, 1e-2, 0.9, 0.999, 1e-8, step)
    _close(tp.cpu().numpy(), p, 1e-5)


def test_mlp_forward_backward_vs_oracle():
    """The ensemble network's output and every gradient, member by member."""
    import torch
    from kwant_amd.surrogate import MomentMLP
    layers = (3, 64, 48, 80)
    z, batch = 3, 40
    net = MomentMLP(layers, members=z, seed=4)
    rng = np.random.RandomState(9)
    x = rng.uniform(-1, 1, (z, batch, 3)).astype(np.float32)
    t = rng.standard_normal((z, batch, layers[-1])).astype(np.float32) * 0.1
    y = net.forward(_t(x))
    dy, _ = net.loss_grad(y, _t(t))
    net.backward(dy)
    y = y.cpu().numpy()
    for i in range(z):
        ws = [net.weight(l)[i].cpu().numpy() for l in range(len(layers) - 1)]
        bs = [net.bias(l)[i].cpu().numpy() for l in range(len(layers) - 1)]
        zs, acts, yo = mo.forward(ws, bs, x[i], split=net.split)
        _close(y[i], yo, 1e-3)
        # per-member loss: mean over that member's B x O entries
        _, dyo = mo.mse(yo, t[i])
        dws, dbs = mo.backward(ws, zs, acts, dyo)
        for l in range(len(layers) - 1):
            _close(net.weight(l, net.grads)[i].cpu().numpy(), dws[l], 2e-3)
            _close(net.bias(l, net.grads)[i].cpu().numpy(), dbs[l], 2e-3)


@pytest.mark.parametrize("precision,tol", [("complex128", 1e-11), ("complex64", 1e-5)])
def test_square_family_vs_oracle(precision, tol):
    """Batched block-diagonal moments equal kwant's SpectralDensity moments
    (oracle restatement) of each system with the same seeded vectors."""
    from oracle import kpm_oracle as ko
    from kwant_amd.lattices import square
    from kwant_amd.surrogate import SquareFamily
    l, m, r = 12, 40, 4
    fam = SquareFamily(l=l, num_moments=m, num_vectors=r, disorder_seed=4, vector_seed=6,
                       batch=3, precision=precision, variant=0)
    params = np.array([[1.5, 0.013, 0.3], [0.0, 0.0, -0.8], [3.9, 0.02, 1.0],
                       [2.2, 0.005, 0.0], [0.7, 0.017, -0.2]])
    got = fam.moments(params)
    for p, (w, phi, mu) in enumerate(params):
        h = square(l, disorder=w, seed=4, shift=mu, flux=phi)[0]
        sd = ko.SpectralDensity(h, num_vectors=r, num_moments=m, bounds=(-7.0, 7.0),
                                rng=6, kernel=lambda x: x)
        ref = np.real(sd._moments())
        ref = ref / ref[0]
        err = np.max(np.abs(got[p] - ref))
        assert err < tol, (p, err)


@pytest.mark.parametrize("m", [40, 64, 130, 512])
def test_moment_shift_kernel_vs_oracle(m):
    """kpm_moment_shift against the recurrence oracle on moments of random
    point spectra, with input and output scales."""
    import torch
    from oracle import shift_oracle as so
    from kwant_amd import _engine as eng
    from kwant_amd.surrogate import jackson_factors
    rng = np.random.RandomState(m)
    z, p, a_t, a_w = 2, 7, 6.15, 7.18
    n = np.arange(m)
    nu = np.empty((z, p, m))
    for i in range(z):
        for j in range(p):
            e = rng.uniform(-6, 6, 50)
            nu[i, j] = np.cos(n[:, None] * np.arccos(e[None, :] / a_t)).mean(1)
    g = jackson_factors(m)
    beta = rng.uniform(-1, 1, p) / a_w
    out = torch.empty((z, p, m), dtype=torch.float32, device="cuda")
    nu_in = (nu * g).astype(np.float32)
    eng.moment_shift(_t(nu_in), a_t / a_w, torch.from_numpy(beta).cuda(), out,
                     in_scale=torch.from_numpy(1 / g).cuda(), out_scale=torch.from_numpy(g).cuda())
    got = out.cpu().numpy()
    for i in range(z):
        want = so.moment_shift(nu_in[i].astype(np.float64), a_t / a_w, beta, 1 / g, g)
        _close(got[i], want, 2e-6)


def test_moment_shift_rejects_bad_arguments():
    import torch
    from kwant_amd import _engine as eng
    x = torch.zeros((1, 2, 8), dtype=torch.float32, device="cuda")
    beta = torch.zeros(2, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        eng.moment_shift(x, 1.5, beta, x)
    with pytest.raises(RuntimeError):
        big = torch.zeros((1, 2, 1100), dtype=torch.float32, device="cuda")
        eng.moment_shift(big, 0.5, beta, big)


def test_shift_connects_tight_and_wide_windows():
    """Moments of H(W, phi, mu) in (-7, 7) equal the exact shift of the
    moments of H(W, phi, 0) in (-6, 6) with the same random vectors."""
    import torch
    from kwant_amd import _engine as eng
    from kwant_amd.surrogate import SquareFamily
    kw = dict(l=16, num_moments=96, num_vectors=4, disorder_seed=2, vector_seed=3, batch=4,
              precision="complex128", variant=0)
    wide = SquareFamily(bounds=(-7.0, 7.0), **kw)
    tight = SquareFamily(bounds=(-6.0, 6.0), **kw)
    params = np.array([[4.0, 0.02, 1.0], [0.3, 0.011, -1.0], [2.5, 0.0, 0.4],
                       [1.0, 0.007, -0.2], [3.3, 0.015, 0.0]])
    p0 = params.copy()
    p0[:, 2] = 0
    want = wide.moments(params)
    nu = tight.moments(p0)
    beta = torch.from_numpy(params[:, 2] / wide.a).cuda()
    out = torch.empty((1,) + nu.shape, dtype=torch.float32, device="cuda")
    eng.moment_shift(_t(nu[None]), tight.a / wide.a, beta, out)
    assert np.max(np.abs(out[0].cpu().numpy() - want)) < 2e-6


def test_surrogate_gain_over_plain_kpm():
    """What the trained network gains over plain KPM on a 24 x 24 lattice:
    measured against the exact moments (dense eigendecomposition), the trained
    ensemble's DOS error is far below the untrained one's, and a gated sweep
    over 1024 points is faster than plain KPM with the same vector count."""
    import time
    import torch
    from kwant_amd.lattices import square
    from kwant_amd.surrogate import (SquareFamily, MomentSurrogate, dos_l1, PARAM_LO,
                                     PARAM_HI)
    m, r, l = 64, 16, 24
    wide = SquareFamily(l=l, num_moments=m, num_vectors=r, bounds=(-7.0, 7.0), batch=8,
                        variant=0)
    tight = SquareFamily(l=l, num_moments=m, num_vectors=r, bounds=(-6.0, 6.0), batch=8,
                         variant=0)
    sur = MomentSurrogate(wide, tight, members=5, layers=(2, 64, 64, 64, 64, m), seed=1,
                          tol=0.05)
    rng = np.random.RandomState(5)
    train = PARAM_LO + (PARAM_HI - PARAM_LO) * rng.random_sample((512, 3))
    targets = sur.training_targets(train, fresh_seed=7)
    test = PARAM_LO + (PARAM_HI - PARAM_LO) * rng.random_sample((24, 3))
    exact = []
    for w, phi, mu in test:
        h = square(l, disorder=w, seed=0, shift=mu, flux=phi)[0].toarray()
        e = np.linalg.eigvalsh(h)
        x = (e - wide.b) / wide.a
        exact.append(np.cos(np.arange(m)[:, None] * np.arccos(x[None, :])).mean(1) * wide.g)
    exact = np.array(exact)
    untrained = sur.predict(test)[0].double().cpu().numpy()
    sur.fit(train, targets, epochs=150, batch=64, lr=3e-3)
    trained = sur.predict(test)[0].double().cpu().numpy()
    kpm = wide.damped_moments(test)
    l1_untrained = np.median(dos_l1(untrained, exact))
    l1_trained = np.median(dos_l1(trained, exact))
    l1_kpm = np.median(dos_l1(kpm, exact))
    print("median DOS L1 vs exact: untrained %.3e trained %.3e plain KPM(R=%d) %.3e"
          % (l1_untrained, l1_trained, r, l1_kpm))
    assert l1_trained < 0.2 * l1_untrained
    # the gated sweep (the ensemble where its spread is within tol, plain KPM
    # elsewhere) is no worse than plain KPM with the same vector count
    gated, fall_test, _ = sur.sweep(test)
    l1_gated = np.median(dos_l1(gated, exact))
    print("gated sweep on the test points: median DOS L1 %.3e (%.0f%% fallback)"
          % (l1_gated, 100 * fall_test.mean()))
    assert l1_gated <= l1_kpm
    # points outside the surrogate's range fall back to plain KPM
    odd = np.array([[5.0, 0.01, 0.0], [1.0, 0.03, 0.0], [1.0, 0.01, 0.99 + sur.max_shift]])
    mom_odd, fall_odd, _ = sur.sweep(odd)
    assert fall_odd.all()
    assert np.allclose(mom_odd, wide.damped_moments(odd))
    sweep = PARAM_LO + (PARAM_HI - PARAM_LO) * rng.random_sample((1024, 3))
    wide.damped_moments(sweep[:8])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    wide.damped_moments(sweep)
    t_kpm = time.perf_counter() - t0
    mom, fall, t_sur = sur.sweep(sweep)
    print("sweep of 1024 points: plain KPM %.3f s, surrogate %.3f s (%.1f%% fallback)"
          % (t_kpm, t_sur, 100 * fall.mean()))
    assert mom.shape == (1024, m)
    assert t_sur < t_kpm


def test_literal_three_input_network_learns():
    """The configuration's literal network (W, flux, mu) -> moments, trained
    on moments of H(W, flux, mu) in the wide window (no shift layer): its
    DOS error against plain KPM falls well below the untrained one's, and
    its predictions have the family's shape."""
    from kwant_amd.surrogate import SquareFamily, MomentSurrogate, dos_l1, PARAM_LO, PARAM_HI
    m, r, l = 48, 8, 16
    wide = SquareFamily(l=l, num_moments=m, num_vectors=r, bounds=(-7.0, 7.0), batch=8,
                        variant=0)
    sur = MomentSurrogate(wide, wide, members=3, layers=(3, 48, 48, m), seed=2, tol=0.05)
    assert sur.literal
    rng = np.random.RandomState(3)
    train = PARAM_LO + (PARAM_HI - PARAM_LO) * rng.random_sample((256, 3))
    targets = sur.training_targets(train, fresh_seed=5)
    test = PARAM_LO + (PARAM_HI - PARAM_LO) * rng.random_sample((16, 3))
    ref = wide.damped_moments(test)
    before = np.median(dos_l1(sur.predict(test)[0].double().cpu().numpy(), ref))
    sur.fit(train, targets, epochs=80, batch=32, lr=3e-3)
    mean, spread = sur.predict(test)
    assert tuple(mean.shape) == (16, m) and tuple(spread.shape) == (16,)
    after = np.median(dos_l1(mean.double().cpu().numpy(), ref))
    assert after < 0.3 * before, (before, after)


def test_fit_graph_replay_matches_eager():
    """Training epochs replayed from one captured HIP graph give bitwise the
    parameters of the same launches run eagerly (several replays, with the
    host's reshuffle and schedule upload between them)."""
    from kwant_amd.surrogate import MomentSurrogate

    class Fam:
        num_moments = 48
        a, b = 7.2, 0.0
        g = np.linspace(1.0, 0.05, 48)

    rng = np.random.RandomState(2)
    params = rng.uniform([0, 0, -1], [4, 0.02, 1], (96, 3))
    targets = rng.standard_normal((96, 48)) * 0.1
    out = []
    for graph in (True, False):
        tight = Fam()
        tight.a = 6.1
        sur = MomentSurrogate(Fam(), tight, members=3, layers=(2, 32, 32, 48), seed=3)
        sur.fit(params, targets, epochs=4, batch=32, lr=1e-2, graph=graph)
        out.append(sur.mlp.params.cpu().numpy())
        assert sur.mlp.t == 4 * 3
    assert np.array_equal(out[0], out[1])


def test_adam_device_schedule_and_ema():
    """Adam reading (lr, bias corrections, EMA decay) from device memory,
    with the parameters' moving average, against the oracle step."""
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(8)
    n = 777
    p = rng.standard_normal(n).astype(np.float32)
    m, v, e = np.zeros(n), np.zeros(n), np.zeros(n)
    tp, tm, tv, te = _t(p), _t(0 * p), _t(0 * p), _t(0 * p)
    pp = p.astype(np.float64)
    for step, (lr, decay) in enumerate([(1e-2, 0.0), (5e-3, 0.0), (4e-3, 0.9), (1e-3, 0.99)], 1):
        gr = rng.standard_normal(n).astype(np.float32)
        sched = _t([lr, 1 / (1 - 0.9 ** step), 1 / (1 - 0.999 ** step), decay])
        eng.mlp_adam(tp, _t(gr), tm, tv, 123.0, 0.9, 0.999, 1e-8, 99, sched=sched, ema=te)
        pp, m, v = mo.adam(pp, gr, m, v, lr, 0.9, 0.999, 1e-8, step)
        e = decay * e + (1 - decay) * pp
    _close(tp.cpu().numpy(), pp, 1e-5)
    _close(te.cpu().numpy(), e, 1e-5)


def test_failed_member_is_left_out():
    """A member whose training error is far above the others' median is
    moved behind the healthy ones and excluded from mean and spread."""
    import torch
    from kwant_amd.surrogate import MomentSurrogate

    class Fam:
        num_moments = 32
        a, b = 7.2, 0.0
        g = np.linspace(1.0, 0.05, 32)

    rng = np.random.RandomState(3)
    params = rng.uniform([0, 0, -1], [4, 0.02, 1], (128, 3))
    targets = np.tile(np.linspace(0.5, 0.0, 32), (128, 1)) * 0.1
    tight = Fam()
    tight.a = 6.1
    sur = MomentSurrogate(Fam(), tight, members=4, layers=(2, 16, 32), seed=2)
    sur.fit(params, targets, epochs=30, batch=32, lr=1e-2, health=10.0)
    assert sur.active == 4
    # spoil member 1, then re-run the health check on the training data
    sur.mlp.weight(1)[1].mul_(50.0)
    x = sur._inputs(params)
    t = torch.from_numpy(targets.astype(np.float32)).cuda()
    sur._select_members(x, t, 10.0)
    assert sur.active == 3
    assert np.argmax(sur.member_errors) == 1
    mean, spread = sur.predict(params[:8])
    assert torch.isfinite(spread).all()


def test_collapsed_ensemble_falls_back_everywhere():
    """Members that all collapse to the same constant output agree perfectly
    (zero spread), so the spread gate alone would accept them; the health
    check compares each member with the best constant prediction, leaves
    them all out, and every point then falls back (infinite spread)."""
    import torch
    from kwant_amd.surrogate import MomentSurrogate

    class Fam:
        num_moments = 32
        a, b = 7.2, 0.0
        g = np.linspace(1.0, 0.05, 32)

    rng = np.random.RandomState(4)
    params = rng.uniform([0, 0, -1], [4, 0.02, 1], (256, 3))
    prof = np.linspace(0.5, 0.0, 32)
    targets = (params[:, :1] * 0.05 + 0.1) * prof[None, :] + 0.02 * np.sin(
        params[:, 1:2] * 300.0 + np.arange(32)[None, :])
    tight = Fam()
    tight.a = 6.1
    sur = MomentSurrogate(Fam(), tight, members=4, layers=(2, 32, 32), seed=5)
    sur.fit(params, targets, epochs=200, batch=32, lr=1e-2)
    assert sur.active >= 2
    assert np.all(sur.member_errors[:sur.active] < 0.5 * sur.constant_error)
    mean, spread = sur.predict(params[:16])
    assert torch.isfinite(spread).all()
    # collapse every member onto the mean target: zero weights, output bias = mean
    for li in range(len(sur.mlp.shapes)):
        sur.mlp.weight(li).zero_()
        sur.mlp.bias(li).zero_()
    last = len(sur.mlp.shapes) - 1
    sur.mlp.bias(last).copy_(torch.from_numpy(targets.mean(axis=0).astype(np.float32)).cuda()[None])
    x = sur._inputs(params)
    t = torch.from_numpy(targets.astype(np.float32)).cuda()
    sur._select_members(x, t, 10.0)
    assert sur.active == 0
    mean, spread = sur.predict(params[:16])
    assert torch.isinf(spread).all()
