"""GPU parity of the Correlator / conductivity path against the CPU oracle
and the dense brute force (complex128 + FP64 GEMM within 1e-10 of the
oracle; the reference's TOL=1e-12 equivalences of kwant/tests/test_kpm.py:62-119)."""
import numpy as np
import pytest

from oracle import brute
from oracle import kpm_oracle as ko

pytestmark = pytest.mark.gpu


def kpm():
    from kwant_amd import kpm as k
    return k


def rel(x, y, scale):
    return float(np.max(np.abs(np.asarray(x) - np.asarray(y))) / scale)


def test_corr_gemm_against_numpy():
    import torch
    from kwant_amd import _engine as eng
    dev = eng.torch_device()
    rng = np.random.RandomState(0)
    for m, n, k in ((7, 5, 33), (64, 64, 1024), (100, 130, 5000), (129, 65, 70000)):
        a = rng.normal(size=(m, k)) + 1j * rng.normal(size=(m, k))
        b = rng.normal(size=(n, k)) + 1j * rng.normal(size=(n, k))
        ref = a.conj() @ b.T
        for prec, tol, dt in ((eng.GEMM_FP64, 1e-12, torch.complex128),
                              (eng.GEMM_FP32, 2e-5, torch.complex64),
                              (eng.GEMM_FP32X3, 2e-5, torch.complex128)):
            for alu in (0, eng.GEMM_VECTOR_ALU):
                mu = torch.zeros((m, n), dtype=torch.complex128, device=dev)
                ta = torch.from_numpy(a).to(dev).to(dt)
                tb = torch.from_numpy(b).to(dev).to(dt)
                eng.corr_gemm(ta, tb, mu, prec | alu)
                got = mu.cpu().numpy()
                # error relative to the natural scale sqrt(k) |a| |b| of an entry
                err = np.max(np.abs(got - ref)) / np.sqrt(k) / 2
                assert err < tol, (m, n, k, prec, alu, err)


def test_correlator_matches_oracle():
    h = brute.gaussian_unitary(16, 3)
    op1, op2 = brute.gaussian_unitary(16, 4), brute.gaussian_unitary(16, 5)
    r = float(np.max(np.abs(np.linalg.eigvalsh(h)))) * 1.05
    kw = dict(num_moments=24, num_vectors=5, bounds=(-r, r), rng=2)
    gpu = kpm().Correlator(h, operator1=op1, operator2=op2, **kw)
    ora = ko.Correlator(h, operator1=op1, operator2=op2, **kw)
    scale = np.abs(ora.moments_matrix).max()
    assert rel(gpu.moments_matrix, ora.moments_matrix, scale) < 1e-10
    for mu in (-1.0, 0.3):
        assert abs(gpu(mu, 0.1) - ora(mu, 0.1)) < 1e-9 * (1 + abs(ora(mu, 0.1)))


@pytest.mark.parametrize("gemm", ["fp64", "sliced4", "isliced5"])
def test_correlator_streamed_rings_match_oracle(monkeypatch, gemm):
    """Both expansions stream through rings shorter than the moment count
    (runs wrap around the ring's end; Omega reaches the GEMM in chunks)."""
    from kwant_amd import correlator
    monkeypatch.setattr(correlator.Correlator, "OMEGA_CHUNK", 7)
    h = brute.gaussian_unitary(18, 13)
    op1, op2 = brute.gaussian_unitary(18, 14), brute.gaussian_unitary(18, 15)
    r = float(np.max(np.abs(np.linalg.eigvalsh(h)))) * 1.05
    kw = dict(num_moments=40, num_vectors=6, bounds=(-r, r), rng=5)
    gpu = kpm().Correlator(h, operator1=op1, operator2=op2, gemm_precision=gemm,
                           max_block_width=4, **kw)
    ora = ko.Correlator(h, operator1=op1, operator2=op2, **kw)
    assert gpu.block_widths == [4, 2]
    scale = np.abs(ora.moments_matrix).max()
    # 4 BF16 slices keep 32 bits of every component against its row's
    # largest, 5 INT8 slices 35 bits
    tol = 1e-10 if gemm == "fp64" else 1e-8
    assert rel(gpu.moments_matrix, ora.moments_matrix, scale) < tol
    assert set(gpu.timings) == {"expansion_s", "operator_s", "gemm_s"}


@pytest.mark.parametrize("ops", ["matrices", "psi_identity", "psi_callable"])
def test_correlator_row_panels_match_oracle(monkeypatch, ops):
    """The FP64 GEMM reads Psi and Omega in row panels: with panels of a few
    rows (ragged last panel, a different cut per block width) and short rings,
    the moment matrix is the oracle's, for operators on the GPU (matrices), a
    copy (identity) and the host path (callable)."""
    from kwant_amd import correlator
    monkeypatch.setattr(correlator.Correlator, "OMEGA_CHUNK", 7)
    monkeypatch.setattr(correlator.Correlator, "PANEL_BYTES", 300)
    h = brute.gaussian_unitary(18, 13)
    m1, m2 = brute.gaussian_unitary(18, 14), brute.gaussian_unitary(18, 15)
    op1, op2 = {"matrices": (m1, m2), "psi_identity": (lambda v: m1 @ v, None),
                "psi_callable": (None, lambda v: m2 @ v)}[ops]
    r = float(np.max(np.abs(np.linalg.eigvalsh(h)))) * 1.05
    kw = dict(num_moments=40, num_vectors=6, bounds=(-r, r), rng=5)
    gpu = kpm().Correlator(h, operator1=op1, operator2=op2, max_block_width=4, **kw)
    # 18 rows: 4 panels of <= 5 rows at width 4, 2 of 9 at width 2
    assert gpu.block_widths == [4, 2]
    cdt = __import__("torch").complex128
    assert len(gpu._panels(18, 4, cdt)) == 4 and len(gpu._panels(18, 2, cdt)) == 2
    ora = ko.Correlator(h, operator1=op1, operator2=op2, **kw)
    scale = np.abs(ora.moments_matrix).max()
    assert rel(gpu.moments_matrix, ora.moments_matrix, scale) < 1e-10


def test_correlator_mean_false_matches_oracle():
    h = brute.gaussian_unitary(12, 6)
    r = float(np.max(np.abs(np.linalg.eigvalsh(h)))) * 1.05
    kw = dict(num_moments=10, num_vectors=3, bounds=(-r, r), rng=2, mean=False)
    gpu = kpm().Correlator(h, **kw)
    ora = ko.Correlator(h, **kw)
    assert gpu.moments_matrix.shape == ora.moments_matrix.shape == (3, 10, 10)
    assert rel(gpu.moments_matrix, ora.moments_matrix,
               np.abs(ora.moments_matrix).max()) < 1e-10


def test_conductivity_square_flake_matches_oracle():
    from kwant_amd.lattices import square
    h, pos = square(8, flux=0.05, disorder=0.5, seed=1)
    kw = dict(num_moments=40, num_vectors=6, bounds=(-4.5, 4.5), rng=0)
    for ab in (('x', 'x'), ('x', 'y')):
        gpu = kpm().conductivity(h, *ab, positions=pos, **kw)
        ora = ko.conductivity(h, *ab, positions=pos, **kw)
        for mu in (-1.0, 0.0, 1.5):
            want = ora(mu, 0.05)
            assert abs(gpu(mu, 0.05) - want) < 1e-9 * (1 + abs(want))


def test_conductivity_formats_and_add_moments():
    """kwant/tests/test_kpm.py:62-119: direction strings vs explicit velocity
    matrices; add_moments equals a fresh run."""
    import scipy.sparse
    from kwant_amd.lattices import square
    h, pos = square(7)
    kw = dict(params=None, num_vectors=10, num_moments=12, energy_resolution=None,
              vector_factory=None, bounds=None, eps=0.05, rng=0, accumulate_vectors=False)
    k = kpm()
    ref = k.conductivity(h, alpha='x', beta='x', positions=pos, **kw)
    x_op = scipy.sparse.coo_matrix(h, copy=True)
    disp = pos[x_op.col] - pos[x_op.row]
    x_op.data = x_op.data * (1j * disp[:, 0])
    for a, b in (('x', x_op), (x_op, 'x'), (x_op, x_op)):
        c = k.conductivity(h, alpha=a, beta=b, positions=pos, **kw)
        np.testing.assert_allclose(c(), ref(), rtol=0, atol=1e-12)
    kw['accumulate_vectors'] = True
    c = k.conductivity(h, alpha='x', beta='x', positions=pos, **kw)
    c.add_moments(num_moments=90)
    kw['accumulate_vectors'] = False
    kw['num_moments'] = 102
    c102 = k.conductivity(h, alpha='x', beta='x', positions=pos, **kw)
    np.testing.assert_allclose(c102(), c(), rtol=0, atol=1e-12)


@pytest.mark.parametrize("seeding", ["state", "global"])
def test_correlator_shared_generator_matches_oracle(seeding):
    """With one generator shared by both expansions (a RandomState, or None
    after numpy.random.seed) the random vectors are the ones kwant draws:
    first expansion's v0, its vectors, then the second expansion's v0
    (kwant/kpm.py:573-603)."""
    h = brute.gaussian_unitary(14, 8)
    r = float(np.max(np.abs(np.linalg.eigvalsh(h)))) * 1.05
    op = brute.gaussian_unitary(14, 9)
    kw = dict(num_moments=16, num_vectors=3, bounds=(-r, r))

    def make(module):
        if seeding == "state":
            return module.Correlator(h, operator1=op, rng=np.random.RandomState(0), **kw)
        np.random.seed(5)
        return module.Correlator(h, operator1=op, rng=None, **kw)

    gpu = make(kpm())
    after_gpu = np.random.random_sample() if seeding == "global" else None
    ora = make(ko)
    after_ora = np.random.random_sample() if seeding == "global" else None
    assert rel(gpu.moments_matrix, ora.moments_matrix,
               np.abs(ora.moments_matrix).max()) < 1e-10
    assert after_gpu == after_ora       # the same number of draws


def test_conductivity_exact_trace_antisymmetric():
    from kwant_amd.lattices import square
    h, pos = square(6, flux=0.1)
    k = kpm()
    kw = dict(num_moments=30, bounds=(-4.2, 4.2), num_vectors=None)
    sxy = k.conductivity(h, 'x', 'y', pos, vector_factory=k.LocalVectors(h), **kw)
    syx = k.conductivity(h, 'y', 'x', pos, vector_factory=k.LocalVectors(h), **kw)
    for mu in (-1.0, 0.7):
        assert abs(sxy(mu, 0.1) + syx(mu, 0.1)) < 1e-11


def _cgemm_ref(a, b):
    return np.conj(a) @ b.T


@pytest.mark.gpu
def test_sliced_gemm_exact_on_small_integers():
    """Components that are integers of at most 7 bits (times a power of two
    per row) fit in one slice: every slice count reproduces the exact
    product bitwise."""
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(0)
    m, n, k = 70, 33, 5000
    a = (rng.randint(-127, 128, (m, k)) + 1j * rng.randint(-127, 128, (m, k))) \
        * 2.0 ** rng.randint(-20, 20, (m, 1))
    b = (rng.randint(-127, 128, (n, k)) + 1j * rng.randint(-127, 128, (n, k))) \
        * 2.0 ** rng.randint(-20, 20, (n, 1))
    want = _cgemm_ref(a, b)
    for slices in (1, 2, 4):
        mu = torch.zeros((m, n), dtype=torch.complex128, device="cuda")
        eng.corr_gemm_sliced(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mu,
                             slices=slices)
        assert np.array_equal(mu.cpu().numpy(), want), slices


@pytest.mark.gpu
@pytest.mark.parametrize("slices,tol", [(2, 3e-4), (3, 2e-6), (4, 1e-8)])
def test_sliced_gemm_accuracy(slices, tol):
    """Random complex128 operands with row scales spread over 2^+-30: the
    error against the FP64 product falls by ~2^-8 per slice."""
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(slices)
    m, n, k = 130, 96, 70000
    a = (rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))) \
        * 2.0 ** rng.uniform(-30, 30, (m, 1))
    b = (rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))) \
        * 2.0 ** rng.uniform(-30, 30, (n, 1))
    want = _cgemm_ref(a, b)
    mu = torch.zeros((m, n), dtype=torch.complex128, device="cuda")
    eng.corr_gemm_sliced(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mu,
                         slices=slices)
    got = mu.cpu().numpy()
    # error per entry relative to the entry's own scale (row x column)
    scale = np.sqrt(k) * np.abs(a).max(1)[:, None] * np.abs(b).max(1)[None, :]
    err = np.max(np.abs(got - want) / scale)
    assert err < tol, err


@pytest.mark.gpu
def test_sliced_gemm_panels_and_accumulate():
    """k longer than one panel (2^20) and accumulation into an existing mu."""
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(9)
    m, n, k = 17, 5, (1 << 20) + 4099
    a = rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
    b = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    mu0 = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    mu = torch.from_numpy(mu0.copy()).cuda()
    eng.corr_gemm_sliced(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mu, slices=4)
    want = mu0 + _cgemm_ref(a, b)
    assert np.max(np.abs(mu.cpu().numpy() - want)) < 1e-8 * np.sqrt(k) * 16


@pytest.mark.gpu
def test_int8_sliced_gemm_exact_on_small_integers():
    """Components that are integers of at most 6 bits (times a power of two
    per row) fit in one 7-bit slice: every slice count gives the exact
    product bitwise, on ragged tile shapes."""
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(1)
    for m, n, k in ((70, 33, 5000), (64, 64, 32), (1, 129, 70001)):
        a = (rng.randint(-63, 64, (m, k)) + 1j * rng.randint(-63, 64, (m, k))) \
            * 2.0 ** rng.randint(-20, 20, (m, 1))
        b = (rng.randint(-63, 64, (n, k)) + 1j * rng.randint(-63, 64, (n, k))) \
            * 2.0 ** rng.randint(-20, 20, (n, 1))
        want = _cgemm_ref(a, b)
        for slices in (1, 3, 6):
            mu = torch.zeros((m, n), dtype=torch.complex128, device="cuda")
            eng.corr_gemm_sliced(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mu,
                                 slices=slices, kind="int8")
            assert np.array_equal(mu.cpu().numpy(), want), (m, n, k, slices)


@pytest.mark.gpu
@pytest.mark.parametrize("slices,tol", [(2, 1e-3), (3, 1e-5), (4, 1e-7), (5, 1e-9), (6, 1e-11)])
def test_int8_sliced_gemm_accuracy(slices, tol):
    """Random complex128 operands with row scales spread over 2^+-30: the
    error against the FP64 product falls by ~2^-7 per slice (the bound is
    about (slices + 1) 2^(-7 slices) of the entry's natural scale)."""
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(10 + slices)
    m, n, k = 130, 96, 70000
    a = (rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))) \
        * 2.0 ** rng.uniform(-30, 30, (m, 1))
    b = (rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))) \
        * 2.0 ** rng.uniform(-30, 30, (n, 1))
    want = _cgemm_ref(a, b)
    mu = torch.zeros((m, n), dtype=torch.complex128, device="cuda")
    eng.corr_gemm_sliced(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mu,
                         slices=slices, kind="int8")
    got = mu.cpu().numpy()
    scale = np.sqrt(k) * np.abs(a).max(1)[:, None] * np.abs(b).max(1)[None, :]
    err = np.max(np.abs(got - want) / scale)
    assert err < tol, err
    # a slice more must gain accuracy (a dropped level or a wrong level
    # scale would not)
    if slices < 6:
        mu2 = torch.zeros_like(mu)
        eng.corr_gemm_sliced(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mu2,
                             slices=slices + 1, kind="int8")
        err2 = np.max(np.abs(mu2.cpu().numpy() - want) / scale)
        assert err2 < err / 16, (err, err2)


@pytest.mark.gpu
@pytest.mark.parametrize("m,n,k", [(70, 33, 5001), (17, 5, (1 << 20) + 4099)])
def test_int8_prepared_operand_equals_direct(m, n, k):
    """B sliced once (its rows filled in two calls, out of order) and
    multiplied by two different A: bitwise the direct INT8 product, and
    accumulated into an existing mu."""
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(m)
    b = torch.from_numpy(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))).cuda()
    prep = eng.IsliceOperand(n, k, 5)
    cut = n // 2
    prep.fill_rows(b[cut:], cut)
    prep.fill_rows(b[:cut], 0)
    for seed in (1, 2):
        a = torch.from_numpy(rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
                             ).cuda() * (10.0 ** seed)
        mu0 = torch.from_numpy(rng.standard_normal((m, n)) + 0j).cuda()
        want = mu0.clone()
        eng.corr_gemm_sliced(a, b, want, slices=5, kind="int8")
        got = mu0.clone()
        prep.gemm(a, got)
        assert torch.equal(got, want), seed


@pytest.mark.gpu
def test_int8_sliced_gemm_panels_and_accumulate():
    """k longer than one panel (2^20) and accumulation into an existing mu."""
    import torch
    from kwant_amd import _engine as eng
    rng = np.random.RandomState(19)
    m, n, k = 17, 5, (1 << 20) + 4099
    a = rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
    b = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    mu0 = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    mu = torch.from_numpy(mu0.copy()).cuda()
    eng.corr_gemm_sliced(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), mu, slices=5,
                         kind="int8")
    want = mu0 + _cgemm_ref(a, b)
    assert np.max(np.abs(mu.cpu().numpy() - want)) < 1e-9 * np.sqrt(k) * 16


def test_conductivity_reordered_equals_caller_order():
    """A shuffled square flake: conductivity with the rows stored in RCM
    order (velocity operators permuted alike) equals the caller-order run."""
    from kwant_amd import ordering
    from kwant_amd.lattices import square
    h0, pos0 = square(10, flux=0.05, disorder=0.5, seed=1)
    order = np.random.RandomState(7).permutation(h0.shape[0])
    h = ordering.permute_csr(h0, order)
    pos = pos0[order]
    kw = dict(num_moments=30, num_vectors=4, bounds=(-4.5, 4.5), rng=0)
    a = kpm().conductivity(h, 'x', 'y', positions=pos, reorder=False, **kw)
    b = kpm().conductivity(h, 'x', 'y', positions=pos, reorder=True, **kw)
    assert b._spectrum_R.dmat.order is not None
    for mu in (-1.0, 0.0, 1.5):
        assert abs(a(mu, 0.05) - b(mu, 0.05)) < 1e-10 * (1 + abs(a(mu, 0.05)))


def test_integral_factor_on_gpu_matches_host(monkeypatch):
    """The integral factor's two products on the FP64 matrix cores equal
    the host products (kwant/kpm.py:745-760 in GEMM form) to 1e-12."""
    from kwant_amd import correlator
    rng = np.random.default_rng(4)
    m = 384
    mu = rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))
    x = np.cos(np.linspace(0.01, np.pi - 0.01, 2 * m))
    monkeypatch.setattr(correlator, 'GPU_FACTOR_MIN', 1 << 30)
    host = correlator.integral_factor_gemm(mu, x)
    monkeypatch.setattr(correlator, 'GPU_FACTOR_MIN', 1)
    gpu = correlator.integral_factor_gemm(mu, x)
    assert np.max(np.abs(gpu - host)) < 1e-12 * np.max(np.abs(host))


def test_conductivity_of_a_lattice_system():
    """conductivity takes a vectorised-assembly system as it takes a kwant
    system (velocities from its positions) and equals the matrix call with
    the same positions."""
    from kwant_amd import kpm as k
    from kwant_amd.assembly import honeycomb_system
    syst = honeycomb_system(6)
    syst.set_onsite(lambda pos, w: w * np.cos(2.3 * pos[:, 0] - 0.7 * pos[:, 1]))
    params = dict(w=0.3)
    h = syst.hamiltonian_submatrix(params=params)
    r = float(np.max(np.asarray(abs(h).sum(axis=1)).ravel()))
    kw = dict(num_moments=40, num_vectors=3, rng=2, bounds=(-r, r))
    a = k.conductivity(syst, alpha='x', beta='y', params=params, **kw)
    b = k.conductivity(h, alpha='x', beta='y', positions=syst.pos_array, **kw)
    for mu in (-0.5, 0.0, 0.7):
        assert a(mu, 0.05) == b(mu, 0.05)


def test_conductivity_qwz_matches_exact_kubo():
    """The HIP Kubo-Bastin chain (complex128 expansions, FP64 GEMM) on the
    QWZ Chern-insulator torus with the exact trace: equal to the oracle and,
    like it, converged to the eigenstate Kubo formula (in the gap at T = 0 and
    T > 0, and inside the bands)."""
    from kubo_cases import kubo_deviation, qwz_conductivity, sigma_trace
    gpu, n, exact = qwz_conductivity(kpm(), 256)
    ora, _, _ = qwz_conductivity(ko, 256)
    got, want = sigma_trace(gpu, n), sigma_trace(ora, n)
    assert np.max(np.abs(got - want)) < 1e-9 * np.max(np.abs(want))
    assert kubo_deviation(gpu, n, exact) < 4e-3


def _dirac_small():
    """The four-band spin-orbit model of the sigma_xy benchmark at 8 x 8
    (A = 1, B = -1, M = -2, Anderson W = 1, Peierls flux 1/64)."""
    from kwant_amd.lattices import dirac_4band
    h, pos = dirac_4band(8, flux=1 / 64, disorder=1.0, seed=6)
    r = float(np.abs(h).sum(axis=1).max())           # Gershgorin radius
    return h, pos, dict(num_moments=64, num_vectors=8, bounds=(-r, r), rng=7)


def test_conductivity_dirac_4band_matches_oracle():
    """sigma_xy of the four-band Dirac model (complex 4 x 4 hoppings, Peierls
    phase, disorder) through the HIP path against the oracle: complex128
    expansions with the FP64 GEMM to 1e-10; the sliced BF16 GEMM (4 slices)
    and the complex64 / FP32 path within their accuracy."""
    h, pos, kw = _dirac_small()
    ora = ko.conductivity(h, 'x', 'y', positions=pos, **kw)
    mus = (-1.5, -0.5, 0.0, 0.4, 1.2)
    want = np.array([ora(mu, 0.05) for mu in mus])
    scale = np.max(np.abs(want))
    fp64 = kpm().conductivity(h, 'x', 'y', positions=pos, **kw)
    assert rel(fp64.moments_matrix, ora.moments_matrix,
               np.abs(ora.moments_matrix).max()) < 1e-10
    assert rel([fp64(mu, 0.05) for mu in mus], want, scale) < 1e-10
    # the sliced GEMM keeps 32 bits per component against each row's largest
    # one (Gamma within ~1e-9); sigma_xy, a small difference of large
    # terms, then moves by ~1e-6 of its largest value (1.4e-6 measured)
    sl4 = kpm().conductivity(h, 'x', 'y', positions=pos, gemm_precision='sliced4', **kw)
    d_sl4 = rel([sl4(mu, 0.05) for mu in mus], want, scale)
    assert d_sl4 < 5e-6, d_sl4
    # 5 INT8 slices: 3 bits more per component
    is5 = kpm().conductivity(h, 'x', 'y', positions=pos, gemm_precision='isliced5', **kw)
    d_is5 = rel([is5(mu, 0.05) for mu in mus], want, scale)
    assert d_is5 < 2e-6 and d_is5 < d_sl4, (d_is5, d_sl4)
    f32 = kpm().conductivity(h, 'x', 'y', positions=pos, precision='complex64', **kw)
    d_f32 = rel([f32(mu, 0.05) for mu in mus], want, scale)
    assert d_f32 < 2e-3, d_f32


def test_gamma_gemm_full_shape_sliced_vs_fp64():
    """The Gamma_nm product at the benchmark's shape (1024 x 1024 moments,
    K = 2^21 = N x 2 interleaved (site, vector) pairs): the sliced BF16 GEMM
    (4 slices) against the FP64 matrix cores, and a few entries of both
    against numpy on the host."""
    import torch
    from kwant_amd import _engine as eng
    dev = eng.torch_device()
    m, k = 1024, 1 << 21
    gen = torch.Generator(device=dev)
    gen.manual_seed(11)
    a = torch.randn((m, k), dtype=torch.complex128, device=dev, generator=gen)
    b = torch.randn((m, k), dtype=torch.complex128, device=dev, generator=gen)
    mu64 = torch.zeros((m, m), dtype=torch.complex128, device=dev)
    eng.corr_gemm(a, b, mu64, eng.GEMM_FP64)
    mus = torch.zeros_like(mu64)
    eng.corr_gemm_sliced(a, b, mus, slices=4)
    # entries are sums of k products of unit-variance complex numbers
    natural = np.sqrt(2.0 * k)
    err = float((mus - mu64).abs().max()) / natural
    assert err < 1e-7, err
    # the INT8 form with 5 slices is closer still
    work = eng.corr_gemm_sliced(a, b, mus.zero_(), slices=5, kind="int8")
    del work
    err8 = float((mus - mu64).abs().max()) / natural
    assert err8 < err / 3 and err8 < 5e-8, (err, err8)
    rows = [0, 1, 517, 1023]
    ha = a[rows].cpu().numpy()
    hb = b[rows].cpu().numpy()
    ref = ha.conj() @ hb.T
    got = mu64[rows][:, rows].cpu().numpy()
    assert np.max(np.abs(got - ref)) / natural < 1e-12
    del a, b
    torch.cuda.empty_cache()
