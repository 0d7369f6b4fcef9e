"""GPU parity of the site-resolved Density path: a kwant-style Density
operator given to SpectralDensity (kwant/kpm.py:186-198, 446-448, 484) is
evaluated inside the recursion kernel (KPM_DOTS_LOCAL records, w = M^H a_0)
and checked against the CPU oracle -- the oracle's SpectralDensity with the
oracle's restatement of Density._operate as the operator callable, on the
same random vectors.  Tolerances: complex128 within 1e-11 of the oracle
relative to the largest moment, complex64 within 1e-5."""
import numpy as np
import pytest
import scipy.sparse

from fake_system import FakeSystem
from oracle import kpm_oracle as ko
from oracle.operator_oracle import density as oracle_density

pytestmark = pytest.mark.gpu

RUNS = [(30, 2), (20, 1), (10, 3)]     # 60 sites, 100 orbitals
WHERE = [5, 31, 31, 55, 0, 59, 12]


def _bounds(h):
    r = float(np.max(np.asarray(abs(h).sum(axis=1)).ravel()))
    return (-r, r)


def _mats():
    rng = np.random.default_rng(5)
    out = {}
    for site in range(60):
        norbs = 2 if site < 30 else (1 if site < 50 else 3)
        a = rng.normal(size=(norbs, norbs)) + 1j * rng.normal(size=(norbs, norbs))
        out[site] = a + a.conj().T
    return out


MATS = _mats()


def _onsite_fn(site, scale):
    return scale * MATS[site]


def _pair(onsite, sum_=False, mean=True, num_vectors=10, num_moments=40,
          precision='complex128', params=None, **kw):
    from kwant_amd import kpm
    from kwant_amd.operator import Density
    syst = FakeSystem(RUNS, seed=3)
    bounds = _bounds(syst.h)
    sd = kpm.SpectralDensity(syst, params=params,
                             operator=Density(syst, onsite, where=WHERE, sum=sum_),
                             num_vectors=num_vectors, num_moments=num_moments, bounds=bounds,
                             rng=7, mean=mean, precision=precision, **kw)
    if callable(onsite):
        ref_onsite = [onsite(s, **params) for s in WHERE]
    else:
        ref_onsite = onsite
    ref = ko.SpectralDensity(syst.h, operator=oracle_density(syst.site_ranges, ref_onsite, WHERE,
                                                             sum=sum_),
                             num_vectors=num_vectors, num_moments=num_moments, bounds=bounds,
                             rng=7, mean=mean)
    return sd, ref


def _rel(x, y):
    x, y = np.asarray(x), np.asarray(y)
    assert x.shape == y.shape, (x.shape, y.shape)
    return float(np.max(np.abs(x - y)) / np.max(np.abs(y)))


@pytest.mark.parametrize("onsite", [1, 0.5, np.eye(1) * 2.0, 'callable'])
def test_density_per_site_mean_matches_oracle(onsite):
    params = None
    if isinstance(onsite, str):
        onsite, params = _onsite_fn, dict(scale=0.5)
    sd, ref = _pair(onsite, params=params)
    assert sd._moments().shape == (40, len(WHERE))
    assert _rel(sd._moments(), ref._moments()) < 1e-11
    assert _rel(sd.densities, ref.densities) < 1e-11
    assert _rel(sd(0.3), ref(0.3)) < 1e-11
    assert _rel(sd.integrate(), ref.integrate()) < 1e-11


def test_density_per_site_per_vector_matches_oracle():
    sd, ref = _pair(_onsite_fn, mean=False, params=dict(scale=1.0), num_vectors=5)
    assert sd._moments().shape == (40, 5, len(WHERE))
    assert _rel(sd._moments(), ref._moments()) < 1e-11
    assert _rel(sd.densities, ref.densities) < 1e-11


def test_density_sum_matches_oracle():
    sd, ref = _pair(_onsite_fn, sum_=True, params=dict(scale=1.0))
    assert sd._moments().shape == (40,)
    assert _rel(sd._moments(), ref._moments()) < 1e-11


def test_density_resume_bitwise():
    once, _ = _pair(0.5, num_moments=60)
    more, _ = _pair(0.5, num_moments=40)
    more.add_moments(20)
    assert np.array_equal(once._moments(), more._moments())
    per_vec, ref = _pair(0.5, mean=False, num_moments=30, num_vectors=3)
    per_vec.add_moments(10)
    ref.add_moments(10)
    per_vec.add_vectors(2)
    ref.add_vectors(2)
    assert _rel(per_vec._moments(), ref._moments()) < 1e-11


def test_density_complex64():
    sd, ref = _pair(_onsite_fn, precision='complex64', params=dict(scale=1.0))
    assert _rel(sd._moments(), ref._moments()) < 1e-5


def test_density_reorder_bitwise():
    a, _ = _pair(0.5, reorder=False)
    b, _ = _pair(0.5, reorder=True)
    assert b._dmat.order is not None
    assert np.array_equal(a._moments(), b._moments())


def test_local_records_every_variant():
    """Each kernel shape writes the same records: per row the sum over the
    block's columns of conj(w) a_n (against numpy on the ring's vectors),
    and per row and column the products."""
    import torch
    from kwant_amd import _engine as eng
    n = 3000
    syst = FakeSystem([(n, 1)], seed=4, density=0.002)
    h = syst.h
    a = _bounds(h)[1] / 0.975
    dev = eng.torch_device()
    for prec in ('complex128', 'complex64'):
        cdt, vt = eng.complex_dtype(prec)
        mat = eng.DeviceCSR(h, prec)
        tol = 1e-13 if prec == 'complex128' else 1e-6
        for width in (1, 8, 32):
            gen = torch.Generator(device=dev)
            gen.manual_seed(width)
            block = torch.polar(torch.ones((n, width), dtype=torch.float64, device=dev),
                                torch.rand((n, width), generator=gen, device=dev,
                                           dtype=torch.float64) * 6.283).to(cdt)
            w = (block * 0.5 + 0.25j).contiguous()
            base = {}
            for v in range(eng.variant_count(width, vt)):
                for mode in (eng.DOTS_LOCAL, eng.DOTS_LOCAL_COLS):
                    rec = eng.BlockRecursion(mat, (a, 0.0), block, width, mode, w=w, variant=v)
                    first = rec.local_now().cpu().numpy()
                    recs = rec.local_steps(3).cpu().numpy()
                    out = np.concatenate([first[None], recs])
                    if v == 0:
                        base[mode] = out
                        wh = w.cpu().numpy().astype(np.complex128)
                        prod = np.conj(wh) * rec.slot(3).cpu().numpy().astype(np.complex128)
                        want = prod if mode == eng.DOTS_LOCAL_COLS else prod.sum(axis=1)
                        scale = np.max(np.abs(want))
                        assert np.max(np.abs(out[3] - want)) < tol * scale, (prec, width, mode)
                    else:
                        assert np.array_equal(out, base[mode]), (prec, width, mode, v)


def test_local_rejects_missing_w():
    import torch
    from kwant_amd import _engine as eng
    h = scipy.sparse.identity(64, dtype=complex, format='csr')
    mat = eng.DeviceCSR(h, 'complex128')
    block = torch.ones((64, 1), dtype=torch.complex128, device=eng.torch_device())
    rec = eng.BlockRecursion(mat, (2.0, 0.0), block, 1, eng.DOTS_LOCAL)
    with pytest.raises(ValueError):
        rec.local_steps(1)


# -- Current (bond records) and Source ----------------------------------------

from oracle.operator_oracle import current as oracle_current  # noqa: E402
from oracle.operator_oracle import source as oracle_source  # noqa: E402


def _hops(syst):
    hops = list(syst.graph)[::6]
    return hops + hops[:2]                          # repeated hoppings


def _current_pair(onsite, sum_=False, mean=True, num_vectors=10, num_moments=30,
                  precision='complex128', **kw):
    from kwant_amd import kpm
    from kwant_amd.operator import Current
    syst = FakeSystem(RUNS, seed=3)
    hops = _hops(syst)
    bounds = _bounds(syst.h)
    sd = kpm.SpectralDensity(syst, operator=Current(syst, onsite, where=hops, sum=sum_),
                             num_vectors=num_vectors, num_moments=num_moments, bounds=bounds,
                             rng=9, mean=mean, precision=precision, **kw)
    ref_onsite = [onsite(a) for a, _ in hops] if callable(onsite) else onsite
    ref = ko.SpectralDensity(
        syst.h, operator=oracle_current(syst.site_ranges, ref_onsite, hops,
                                        [syst.hamiltonian(a, b) for a, b in hops], sum=sum_),
        num_vectors=num_vectors, num_moments=num_moments, bounds=bounds, rng=9, mean=mean)
    return sd, ref


def _onsite_plain(site):
    return MATS[site]


@pytest.mark.parametrize("onsite", [1, 0.5, 'callable'])
def test_current_mean_matches_oracle(onsite):
    onsite = _onsite_plain if onsite == 'callable' else onsite
    sd, ref = _current_pair(onsite)
    assert sd._moments().shape == ref._moments().shape
    assert _rel(sd._moments(), ref._moments()) < 1e-11
    assert _rel(sd.densities, ref.densities) < 1e-11


def test_current_per_vector_and_sum_match_oracle():
    sd, ref = _current_pair(_onsite_plain, mean=False, num_vectors=3)
    assert _rel(sd._moments(), ref._moments()) < 1e-11
    sd, ref = _current_pair(_onsite_plain, sum_=True)
    assert sd._moments().shape == (30,)
    assert _rel(sd._moments(), ref._moments()) < 1e-11


def test_current_resume_reorder_and_complex64():
    once, _ = _current_pair(0.5, num_moments=40)
    more, _ = _current_pair(0.5, num_moments=25)
    more.add_moments(15)
    assert np.array_equal(once._moments(), more._moments())
    perm, _ = _current_pair(0.5, num_moments=40, reorder=True)
    assert perm._dmat.order is not None
    assert np.array_equal(once._moments(), perm._moments())
    sd, ref = _current_pair(_onsite_plain, precision='complex64')
    assert _rel(sd._moments(), ref._moments()) < 1e-5


def test_source_matches_oracle():
    from kwant_amd import kpm
    from kwant_amd.operator import Source
    syst = FakeSystem(RUNS, seed=3)
    bounds = _bounds(syst.h)
    for mean in (True, False):
        sd = kpm.SpectralDensity(syst, operator=Source(syst, _onsite_plain, where=WHERE),
                                 num_vectors=4, num_moments=30, bounds=bounds, rng=11, mean=mean)
        ref = ko.SpectralDensity(
            syst.h, operator=oracle_source(syst.site_ranges, [MATS[s] for s in WHERE], WHERE,
                                           [syst.hamiltonian(s, s) for s in WHERE]),
            num_vectors=4, num_moments=30, bounds=bounds, rng=11, mean=mean)
        assert _rel(sd._moments(), ref._moments()) < 1e-11


def test_bond_records_against_numpy():
    import torch
    from kwant_amd import _engine as eng
    n = 3000
    syst = FakeSystem([(n, 1)], seed=4, density=0.002)
    dev = eng.torch_device()
    for prec in ('complex128', 'complex64'):
        cdt, _ = eng.complex_dtype(prec)
        npdt = np.complex64 if prec == 'complex64' else np.complex128
        bond = eng.DeviceCSR(syst.h, prec)
        tol = 1e-13 if prec == 'complex128' else 1e-6
        for width in (1, 8, 32):
            gen = torch.Generator(device=dev)
            gen.manual_seed(width)
            blocks = [torch.polar(torch.ones((n, width), dtype=torch.float64, device=dev),
                                  torch.rand((n, width), generator=gen, device=dev,
                                             dtype=torch.float64) * 6.283).to(cdt)
                      for _ in range(4)]
            bra, ket, v, w = blocks
            hb, hk, hv, hw = (b.cpu().numpy().astype(np.complex128) for b in blocks)
            csr = syst.h.tocsr()
            r = np.repeat(np.arange(n), np.diff(csr.indptr))
            c, h = csr.indices, csr.data.astype(npdt).astype(np.complex128)
            want = 1j * (np.conj(h)[:, None] * np.conj(hb[c]) * hv[r]
                         - np.conj(hw[r]) * h[:, None] * hk[c])
            for cols in (False, True):
                out = torch.empty((csr.nnz, width) if cols else (csr.nnz,),
                                  dtype=torch.complex128, device=dev)
                bond.bond_records(bra, ket, v, w, width, out, per_column=cols)
                got = out.cpu().numpy()
                ref = want if cols else want.sum(axis=1)
                assert np.max(np.abs(got - ref)) < tol * np.max(np.abs(ref)), (prec, width, cols)


def test_correlator_with_kwant_style_operators():
    """Density / Current as Correlator operators run as sparse matrices on
    the GPU; the result equals the oracle Correlator given the same
    operators as matrices built independently (dense blocks)."""
    from kwant_amd import kpm
    from kwant_amd.operator import Current, Density
    syst = FakeSystem([(40, 1)], seed=12, density=0.1)
    hops = list(syst.graph)
    bounds = _bounds(syst.h)
    dens = np.diag(np.linspace(0.5, 1.5, 40)).astype(complex)
    jmat = np.zeros((40, 40), dtype=complex)
    for a, b in hops:
        hab = syst.h[a, b]
        jmat[b, a] += 1j * np.conj(hab)
        jmat[a, b] += -1j * hab
    ours = kpm.Correlator(syst, operator1=Density(syst, lambda s: dens[s, s]),
                          operator2=Current(syst), num_moments=40, num_vectors=4, rng=13,
                          bounds=bounds)
    ref = ko.Correlator(syst.h, operator1=scipy.sparse.csr_matrix(dens),
                        operator2=scipy.sparse.csr_matrix(jmat), num_moments=40,
                        num_vectors=4, rng=13, bounds=bounds)
    for mu, temp in ((0.1, 0.0), (-0.3, 0.05)):
        assert abs(ours(mu, temp) - ref(mu, temp)) < 1e-10 * max(1.0, abs(ref(mu, temp)))


def test_lattice_system_current_spectral_density():
    """KPM straight from a vectorised-assembly system, with a Current over
    its own hoppings, against the oracle on the same matrix."""
    from kwant_amd import kpm
    from kwant_amd.assembly import honeycomb_system
    from kwant_amd.operator import Current
    syst = honeycomb_system(8, onsite=0.0)
    syst.set_onsite(lambda pos, w: w * np.cos(3.1 * pos[:, 0] + 1.7 * pos[:, 1]))
    params = dict(w=0.4)
    h = syst.hamiltonian_submatrix(params=params)
    hops = list(syst.graph)[::5]
    bounds = _bounds(h)
    sd = kpm.SpectralDensity(syst, params=params, operator=Current(syst, where=hops),
                             num_vectors=6, num_moments=30, bounds=bounds, rng=3)
    ref = ko.SpectralDensity(
        h, operator=oracle_current(syst.site_ranges, 1, hops,
                                   [syst.hamiltonian(a, b, params=params) for a, b in hops]),
        num_vectors=6, num_moments=30, bounds=bounds, rng=3)
    assert _rel(sd._moments(), ref._moments()) < 1e-11
