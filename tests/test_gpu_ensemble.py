"""Device-side cubic 4-orbital generator and the batched ensemble driver
against the CPU oracles."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import lattice_oracle as lo

pytestmark = pytest.mark.gpu


def test_gen_cubic4_matches_oracle():
    """Three systems of uneven size written into one block-diagonal CSR equal
    the oracle's matrices block for block; radii bound the spectra."""
    import torch
    from kwant_amd import _engine as eng
    sizes, seeds = [3, 5, 2], [11, 12, 13]
    mats = [lo.cubic4(L, s, sigma=0.3, t=1.0) for L, s in zip(sizes, seeds)]
    big = sp.block_diag(mats).tocsr()
    n, nnz = big.shape[0], big.nnz
    dev = eng.torch_device()
    indptr = torch.full((n + 1,), -1, dtype=torch.int32, device=dev)
    indices = torch.full((nnz,), -1, dtype=torch.int32, device=dev)
    values = torch.zeros(nnz, dtype=torch.complex128, device=dev)
    radius = torch.zeros(len(sizes), dtype=torch.int32, device=dev)
    r0 = z0 = 0
    for i, (L, s) in enumerate(zip(sizes, seeds)):
        assert eng.cubic4_nnz(L) == mats[i].nnz
        eng.gen_cubic4(L, s, 0.3, 1.0, r0, z0, indptr, indices, values, radius[i:i + 1])
        r0 += 4 * L ** 3
        z0 += mats[i].nnz
    assert np.array_equal(indptr.cpu().numpy(), big.indptr)
    assert np.array_equal(indices.cpu().numpy(), big.indices)
    assert np.max(np.abs(values.cpu().numpy() - big.data)) < 1e-14
    # Hermitian bit for bit (no FMA contraction in the on-site formula)
    gpu = sp.csr_matrix((values.cpu().numpy(), indices.cpu().numpy(), indptr.cpu().numpy()),
                        shape=big.shape)
    assert (gpu - gpu.conj().T).count_nonzero() == 0
    rad = radius.cpu().numpy().view(np.float32)
    for i, m in enumerate(mats):
        g = lo.gershgorin_radius(m)
        assert g <= rad[i] <= g * (1 + 1e-5)
        assert np.max(np.abs(np.linalg.eigvalsh(m.toarray()))) <= rad[i]


@pytest.mark.parametrize("precision,tol", [("complex128", 1e-11), ("complex64", 1e-5)])
def test_ensemble_moments_vs_oracle(precision, tol):
    """Per-system moments of one batched launch equal kwant's SpectralDensity
    (oracle restatement) of each system with the same vectors and bounds."""
    from oracle import kpm_oracle as ko
    from kwant_amd.ensemble import CubicEnsemble
    sizes = [3, 4, 2]
    ens = CubicEnsemble(sizes, seed=7, num_moments=40, num_vectors=4, precision=precision,
                        variant=0)
    ens.run(graph=False)
    got = ens.moments(damped=False)
    vecs = ens.vectors.cpu().numpy().astype(np.complex128)
    for s, L in enumerate(sizes):
        m = lo.cubic4(L, ens.seeds[s])
        r0, r1 = ens.row_off[s], ens.row_off[s + 1]
        a, b = ens.scales[s]
        rad = ens.radius[s]
        sd = ko.SpectralDensity(m, num_moments=40, bounds=(-rad, rad),
                                vector_factory=[vecs[r0:r1, c] for c in range(4)],
                                num_vectors=4, kernel=lambda x: x, eps=0.05)
        assert abs(sd._a - a) < 1e-12 * a
        ref = np.real(sd._moments())
        ref = ref / ref[0]
        assert np.max(np.abs(got[s] - ref)) < tol, s


def test_ensemble_graph_replay_equals_eager():
    """Replaying the captured step sequence (with eager work in between)
    gives the eager launches' dot products bitwise, replay after replay."""
    from kwant_amd.ensemble import CubicEnsemble
    ens = CubicEnsemble([4, 6, 3, 5], seed=3, num_moments=64, num_vectors=8, variant=0)
    ens.run(graph=False)
    eager = ens.dots.cpu().numpy().copy()
    for _ in range(3):
        ens.dots.zero_()
        ens.run(graph=True)
        assert np.array_equal(ens.dots.cpu().numpy(), eager)
        ens.run(graph=False)
        assert np.array_equal(ens.dots.cpu().numpy(), eager)


def test_gen_cubic4_onsite_matches_csr_generator():
    """The stencil's on-site blocks are the CSR generator's on-site entries,
    bitwise, with the same radius."""
    import torch
    from kwant_amd import _engine as eng
    sizes, seeds = [3, 1, 4], [21, 22, 23]
    mats = [lo.cubic4(L, s, sigma=0.3, t=1.0) for L, s in zip(sizes, seeds)]
    n = sum(4 * L ** 3 for L in sizes)
    dev = eng.torch_device()
    for cdt in (torch.complex64, torch.complex128):
        onsite = torch.zeros((n, 4), dtype=cdt, device=dev)
        radius = torch.zeros(len(sizes), dtype=torch.int32, device=dev)
        csr_radius = torch.zeros(len(sizes), dtype=torch.int32, device=dev)
        r0 = 0
        for i, (L, s) in enumerate(zip(sizes, seeds)):
            eng.gen_cubic4_onsite(L, s, 0.3, 1.0, r0, onsite, radius[i:i + 1])
            nnz = eng.cubic4_nnz(L)
            indptr = torch.empty(4 * L ** 3 + 1, dtype=torch.int32, device=dev)
            indices = torch.empty(nnz, dtype=torch.int32, device=dev)
            values = torch.empty(nnz, dtype=cdt, device=dev)
            eng.gen_cubic4(L, s, 0.3, 1.0, 0, 0, indptr, indices, values,
                           csr_radius[i:i + 1])
            csr = sp.csr_matrix((values.cpu().numpy(), indices.cpu().numpy(),
                                 indptr.cpu().numpy()), shape=(4 * L ** 3,) * 2)
            got = onsite[r0:r0 + 4 * L ** 3].cpu().numpy()
            for site in range(L ** 3):
                block = csr[4 * site:4 * site + 4, 4 * site:4 * site + 4].toarray()
                assert np.array_equal(got[4 * site:4 * site + 4], block), (L, site)
            r0 += 4 * L ** 3
        assert np.array_equal(radius.cpu().numpy(), csr_radius.cpu().numpy())


@pytest.mark.parametrize("precision,width", [("complex64", 8), ("complex64", 16),
                                             ("complex128", 4), ("complex128", 16)])
@pytest.mark.parametrize("shape", range(16))
def test_stencil_ensemble_bitwise_csr(precision, width, shape):
    """The lattice-stencil steps give the CSR kernel's dots and vectors
    bitwise, for sides 1 to 11 (every boundary case; tiles that span bands of
    y rows and end inside a system), for every launch shape, eager and
    replayed from a graph."""
    from kwant_amd import _engine as eng
    from kwant_amd.ensemble import CubicEnsemble
    assert len(eng.MarchPlan.SHAPES) == 16
    _, vt = eng.complex_dtype(precision)
    assert eng.MarchPlan.shape_ok(width, vt, shape)
    sizes = [1, 2, 3, 5, 8, 4, 11, 7]
    kw = dict(seed=5, num_moments=40, num_vectors=width, precision=precision)
    csr = CubicEnsemble(sizes, variant=0, **kw)
    st = CubicEnsemble(sizes, variant=shape, matrix_format="stencil", **kw)
    assert np.array_equal(st.radius, csr.radius)
    assert np.array_equal(st.scales, csr.scales)
    csr.run(graph=False)
    st.run(graph=False)
    assert np.array_equal(st.dots.cpu().numpy(), csr.dots.cpu().numpy())
    assert np.array_equal(st.ring.cpu().numpy(), csr.ring.cpu().numpy())
    eager = st.dots.cpu().numpy().copy()
    for _ in range(2):
        st.dots.zero_()
        st.run(graph=True)
        assert np.array_equal(st.dots.cpu().numpy(), eager)


def test_stencil_ensemble_moments_vs_oracle():
    """Per-system moments of the stencil path against the oracle's
    SpectralDensity of each system (complex128, 1e-11)."""
    from oracle import kpm_oracle as ko
    from kwant_amd.ensemble import CubicEnsemble
    sizes = [3, 2, 4]
    ens = CubicEnsemble(sizes, seed=9, num_moments=40, num_vectors=4,
                        precision="complex128", matrix_format="stencil")
    ens.run(graph=False)
    got = ens.moments(damped=False)
    vecs = ens.vectors.cpu().numpy().astype(np.complex128)
    for s, L in enumerate(sizes):
        m = lo.cubic4(L, ens.seeds[s])
        r0, r1 = ens.row_off[s], ens.row_off[s + 1]
        rad = ens.radius[s]
        sd = ko.SpectralDensity(m, num_moments=40, bounds=(-rad, rad),
                                vector_factory=[vecs[r0:r1, c] for c in range(4)],
                                num_vectors=4, kernel=lambda x: x, eps=0.05)
        ref = np.real(sd._moments())
        ref = ref / ref[0]
        assert np.max(np.abs(got[s] - ref)) < 1e-11, s
