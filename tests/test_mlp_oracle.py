"""Pins of the surrogate's CPU oracle (oracle/mlp_oracle.py) against
independent implementations (torch) and closed forms, plus the host-side
pieces of kwant_amd.surrogate that need no GPU."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.RandomState(0)
    bits = rng.randint(0, 2 ** 32, size=200000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    # exact ties between two bf16 values, both parities
    ties = (np.arange(1, 2000, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = mo.bf16(x)
    finite = np.isfinite(want)
    assert np.array_equal(got[finite], want[finite])


def test_split3_recovers_16_bits():
    rng = np.random.RandomState(1)
    x = rng.standard_normal(10000).astype(np.float32)
    hi, lo = mo.split3(x)
    rel = np.abs((hi.astype(np.float64) + lo) - x) / np.abs(x)
    assert rel.max() < 2.0 ** -16
    assert np.abs(hi - x).max() / np.abs(x).max() > 2.0 ** -12  # bf16 alone is coarser


def test_gelu_matches_torch():
    import torch
    x = np.linspace(-8, 8, 4001)
    t = torch.from_numpy(x).requires_grad_(True)
    y = torch.nn.functional.gelu(t)
    y.sum().backward()
    assert np.max(np.abs(mo.gelu(x) - y.detach().numpy())) < 1e-14
    assert np.max(np.abs(mo.gelu_grad(x) - t.grad.numpy())) < 1e-14


def _net(layers, seed):
    rng = np.random.RandomState(seed)
    ws = [rng.standard_normal((o, i)) / np.sqrt(i) for i, o in zip(layers[:-1], layers[1:])]
    bs = [0.1 * rng.standard_normal(o) for o in layers[1:]]
    return ws, bs


def test_backward_matches_finite_differences():
    layers = (3, 7, 5, 4)
    ws, bs = _net(layers, 2)
    rng = np.random.RandomState(3)
    x = rng.uniform(-1, 1, (6, 3))
    t = rng.standard_normal((6, 4))
    zs, acts, y = mo.forward(ws, bs, x, rounding=False)
    _, dy = mo.mse(y, t)
    dws, dbs = mo.backward(ws, zs, acts, dy, rounding=False)
    h = 1e-6
    for l in range(len(ws)):
        for idx in [(0, 0), (ws[l].shape[0] - 1, ws[l].shape[1] - 1), (1, 2 % ws[l].shape[1])]:
            wp = [w.copy() for w in ws]
            wm = [w.copy() for w in ws]
            wp[l][idx] += h
            wm[l][idx] -= h
            fd = (mo.loss(wp, bs, x, t) - mo.loss(wm, bs, x, t)) / (2 * h)
            assert abs(fd - dws[l][idx]) < 1e-7 * max(1.0, abs(fd)), (l, idx)
        for j in (0, bs[l].shape[0] - 1):
            bp = [b.copy() for b in bs]
            bm = [b.copy() for b in bs]
            bp[l][j] += h
            bm[l][j] -= h
            fd = (mo.loss(ws, bp, x, t) - mo.loss(ws, bm, x, t)) / (2 * h)
            assert abs(fd - dbs[l][j]) < 1e-7 * max(1.0, abs(fd)), (l, j)


def test_backward_matches_torch_autograd():
    import torch
    layers = (3, 16, 16, 8)
    ws, bs = _net(layers, 5)
    rng = np.random.RandomState(6)
    x = rng.uniform(-1, 1, (10, 3))
    t = rng.standard_normal((10, 8))
    zs, acts, y = mo.forward(ws, bs, x, rounding=False)
    _, dy = mo.mse(y, t)
    dws, dbs = mo.backward(ws, zs, acts, dy, rounding=False)
    tw = [torch.from_numpy(w).requires_grad_(True) for w in ws]
    tb = [torch.from_numpy(b).requires_grad_(True) for b in bs]
    a = torch.from_numpy(x)
    for l in range(len(tw)):
        a = a @ tw[l].T + tb[l]
        if l < len(tw) - 1:
            a = torch.nn.functional.gelu(a)
    torch.mean((a - torch.from_numpy(t)) ** 2).backward()
    for l in range(len(tw)):
        assert np.max(np.abs(tw[l].grad.numpy() - dws[l])) < 1e-12
        assert np.max(np.abs(tb[l].grad.numpy() - dbs[l])) < 1e-12


def test_rounded_network_is_close_to_exact():
    """BF16 operands change outputs at the 1e-2 level, not more (sanity of
    the rounding path that the GPU is compared with)."""
    layers = (3, 64, 64, 32)
    ws, bs = _net(layers, 8)
    x = np.random.RandomState(9).uniform(-1, 1, (20, 3))
    y_exact = mo.forward(ws, bs, x, rounding=False)[2]
    y_bf = mo.forward(ws, bs, x, split=[True, False, False])[2]
    err = np.max(np.abs(y_bf - y_exact)) / np.max(np.abs(y_exact))
    assert 1e-5 < err < 3e-2


def test_adam_first_step_closed_form_and_torch():
    import torch
    rng = np.random.RandomState(4)
    p0 = rng.standard_normal(50)
    g = rng.standard_normal(50)
    lr, eps = 1e-2, 1e-8
    p1, m1, v1 = mo.adam(p0, g, 0 * p0, 0 * p0, lr, 0.9, 0.999, eps, 1)
    assert np.allclose(p1, p0 - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-15)
    tp = torch.from_numpy(p0.copy()).requires_grad_(True)
    opt = torch.optim.Adam([tp], lr=lr, betas=(0.9, 0.999), eps=eps)
    p, m, v = p0.copy(), 0 * p0, 0 * p0
    for step in range(1, 6):
        gs = rng.standard_normal(50)
        tp.grad = torch.from_numpy(gs.copy())
        opt.step()
        p, m, v = mo.adam(p, gs, m, v, lr, 0.9, 0.999, eps, step)
    assert np.max(np.abs(tp.detach().numpy() - p)) < 1e-12


def test_dos_l1_closed_forms():
    """int |rho| dx of a single damped moment: |d_0| for n = 0 and
    4 |d_n| / pi for n >= 1 (int_0^pi |cos n t| dt = 2)."""
    from kwant_amd.surrogate import dos_l1
    m = 64
    for n, c in [(0, 0.7), (1, -0.3), (17, 0.25), (63, 1.0)]:
        d = np.zeros(m)
        d[n] = c
        want = abs(c) if n == 0 else 4 * abs(c) / np.pi
        assert abs(mo.dos_l1(d, 0 * d)[0] - want) < 1e-3 * want
        assert abs(dos_l1(d, 0 * d) - want) < 2e-3 * want


def test_dos_l1_product_vs_oracle():
    """The surrogate's 2M-node L1 against the oracle's fine midpoint rule."""
    from kwant_amd.surrogate import dos_l1, jackson_factors
    rng = np.random.RandomState(12)
    m = 128
    g = jackson_factors(m)
    a = rng.standard_normal((8, m)) * np.exp(-np.arange(m) / 30) * g
    b = a + 1e-3 * rng.standard_normal((8, m)) * g
    got = dos_l1(a, b)
    want = mo.dos_l1(a, b)
    assert np.max(np.abs(got - want) / want) < 1e-2


def test_square_pattern_matches_square():
    from kwant_amd.lattices import SquarePattern, square
    p = SquarePattern(9, seed=3)
    for w, phi, mu in [(1.3, 0.013, 0.4), (2.0, 0.0, -0.7), (0.5, 0.02, 1.0)]:
        a = p.matrix(w, phi, mu).toarray()
        b = square(9, disorder=w, seed=3, shift=mu, flux=phi)[0].toarray()
        assert np.array_equal(a, b)
    assert SquarePattern(512).nnz == 1308672


def test_normalize_params_maps_box_to_unit_cube():
    from kwant_amd.surrogate import normalize_params, PARAM_LO, PARAM_HI
    x = normalize_params(np.stack([PARAM_LO, PARAM_HI, (PARAM_LO + PARAM_HI) / 2]))
    assert np.allclose(x, [[-1, -1, -1], [1, 1, 1], [0, 0, 0]])


@pytest.mark.parametrize("bad", [dict(layers=(4, 8, 512)), dict(layers=(2, 8, 100)),
                                 dict(layers=(3, 8, 100))])
def test_surrogate_rejects_bad_layers(bad):
    """Shape validation happens before any GPU work."""
    from kwant_amd.surrogate import MomentSurrogate

    class Fam:
        num_moments = 512
    with pytest.raises(ValueError):
        MomentSurrogate(Fam(), Fam(), **bad)
