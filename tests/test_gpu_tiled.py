"""Two Chebyshev steps per launch (kpm_cheb_steps2): every dot and every ring
row is bitwise what one-step launches give on the same device matrix, for
every tiled launch shape, both precisions, step counts that start at 0 and
end odd, ragged rows longer than every prefetch depth, one-row tiles and
tiles whose halo is most of the matrix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def eng():
    from kwant_amd import _engine
    return _engine


def phase_block(n, width, cdt, seed):
    import torch
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    ph = torch.rand((n, width), generator=gen, device="cuda", dtype=torch.float64)
    return torch.polar(torch.ones_like(ph), 2 * np.pi * ph).to(cdt)


def run(mat, ab, block, width, mode, chunks, tiles=None):
    rec = eng().BlockRecursion(mat, ab, block, width, mode, tiles=tiles)
    dots = [rec.step(k, want_dots=mode != eng().DOTS_NONE) for k in chunks]
    last = (rec.slot(rec.n - 1).cpu().numpy(), rec.slot(rec.n).cpu().numpy())
    return dots, last


def check_same(mat, ab, width, prec, tile_begin, chunks=(5, 8, 1, 6), seed=0):
    """tile_begin: one array for every shape, or a function of the shape's
    largest tile giving (device matrix, tile boundaries)."""
    e = eng()
    cdt, vt = e.complex_dtype(prec)
    ran = 0
    for mode in (e.DOTS_DOUBLING, e.DOTS_NONE):
        for v in range(e.tiled_variant_count(width, vt)):
            m, tb = (mat, tile_begin) if not callable(tile_begin) else \
                tile_begin(e.tiled_variant_info(width, vt, v)[2])
            try:
                tiles = e.TiledPlan(m, width, tb, variant=v)
            except RuntimeError as err:   # tile + halo larger than this shape's LDS
                assert "LDS" in str(err), err
                continue
            except ValueError as err:     # a tile larger than this shape takes
                assert "this variant takes" in str(err), err
                continue
            block = phase_block(m.shape[0], width, cdt, seed)
            ref_dots, ref_last = run(m, ab, block, width, mode, chunks)
            dots, last = run(m, ab, block, width, mode, chunks, tiles=tiles)
            for a, b in zip(dots, ref_dots):
                if mode == e.DOTS_DOUBLING:
                    assert np.array_equal(a, b), (prec, width, v)
            assert np.array_equal(last[0], ref_last[0]), (prec, width, v, mode)
            assert np.array_equal(last[1], ref_last[1]), (prec, width, v, mode)
            ran += 1
    assert ran > 0
    return ran


@pytest.mark.parametrize("prec", ["complex64", "complex128"])
@pytest.mark.parametrize("width", [16, 64])
def test_tiled_graphene_bitwise(prec, width):
    from kwant_amd.lattices import honeycomb
    from kwant_amd.ordering import tile_order
    e = eng()
    h, pos = honeycomb(40, disorder=0.5, seed=3)

    def tiled(rows):
        tiling = tile_order(pos, rows)
        print("graphene 40x40, %s, width %d, %d tiles of <= %d rows"
              % (prec, width, tiling.n_tiles, rows), flush=True)
        return e.DeviceCSR(h, prec, order=tiling.order), tiling.tile_begin
    ran = check_same(None, (3.5, 0.1), width, prec, tiled)
    _, vt = e.complex_dtype(prec)
    assert ran == 2 * e.tiled_variant_count(width, vt)   # every shape, both modes


@pytest.mark.parametrize("prec", ["complex64", "complex128"])
def test_tiled_ragged_rows_and_odd_tiles(prec):
    """Rows of 0 to 40 entries in random order (halos spread over the matrix),
    tiles of 1 to 90 rows; shapes whose LDS cannot hold a tile and its halo
    refuse the tiling."""
    from test_gpu_dos import ragged_hermitian
    e = eng()
    h = ragged_hermitian(700, 11)
    r = float(np.max(np.abs(h).sum(axis=1)))
    tb = np.array([0, 1, 2, 50, 130, 200, 260, 350, 420, 500, 560, 640, 700])
    mat = e.DeviceCSR(h, prec)
    print("ragged 700 rows, %s, tiles %s" % (prec, tb.tolist()), flush=True)
    check_same(mat, (1.01 * r, 0.0), 16, prec, tb, chunks=(1, 2, 7, 3))


def test_tiled_info_and_errors():
    from kwant_amd.lattices import honeycomb
    from kwant_amd.ordering import tile_order
    e = eng()
    h, pos = honeycomb(30, disorder=0.5, seed=3)
    tiling = tile_order(pos, 512)
    mat = e.DeviceCSR(h, "complex64", order=tiling.order)
    tiles = e.TiledPlan(mat, 64, tiling.tile_begin)
    info = tiles.info()
    assert info["tiles"] == tiling.n_tiles
    assert info["max_tile_rows"] == int(np.max(np.diff(tiling.tile_begin)))
    assert 0 < info["halo_rows"] < h.shape[0]
    assert info["workgroups"] % 32 == 0  # 8 XCDs x 4 column groups of 16
    with pytest.raises(ValueError):
        e.TiledPlan(mat, 64, [0, 10, 10, h.shape[0]])      # empty tile
    with pytest.raises(ValueError):
        e.TiledPlan(mat, 64, [0, h.shape[0] - 1])          # rows left out
    cdt, _ = e.complex_dtype("complex64")
    block = phase_block(h.shape[0], 64, cdt, 1)
    with pytest.raises(ValueError):   # a ring of two slots cannot hold a_{n-1} .. a_{n+2}
        rec = e.BlockRecursion(mat, (3.5, 0.0), block, 64, e.DOTS_DOUBLING, tiles=tiles,
                               store_slots=2)
        rec.step(4)
    with pytest.raises(ValueError):
        e.BlockRecursion(mat, (3.5, 0.0), block, 64, e.DOTS_PROJECT, w=block, tiles=tiles)


def test_spectral_density_with_tiles():
    """SpectralDensity(tiles=...) gives bitwise the moments of the same row
    order without tiles (blocks of 16, 2 and 1 vectors: only the first has a
    tiled shape; an odd moment count), through add_moments, add_vectors and a
    save/load round trip; complex128 within 1e-11 of the oracle."""
    import os
    import tempfile
    from kwant_amd import kpm
    from kwant_amd.lattices import honeycomb, honeycomb_bounds
    from kwant_amd.ordering import tile_order
    from oracle import kpm_oracle as ko
    h, pos = honeycomb(24, disorder=0.5, seed=4)
    bounds = honeycomb_bounds(1.0, 0.5)
    tiling = tile_order(pos, 128)
    kw = dict(num_moments=101, num_vectors=19, bounds=bounds, rng=7)
    ora = ko.SpectralDensity(h, **kw)
    for prec in ("complex128", "complex64"):
        plain = kpm.SpectralDensity(h, precision=prec, reorder=tiling.order, **kw)
        tiled = kpm.SpectralDensity(h, precision=prec, tiles=tiling, **kw)
        assert tiled._tiled[16] is not None
        assert tiled._blocks[0].rec.tiles is not None
        assert np.array_equal(np.array(tiled._moments_list), np.array(plain._moments_list))
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "sd.npz")
            tiled.save(path)
            back = kpm.SpectralDensity.load(path, h)
        assert back._blocks[0].rec.tiles is not None
        for sd in (plain, tiled, back):
            sd.add_moments(20)
            sd.add_vectors(5)
        assert np.array_equal(np.array(tiled._moments_list), np.array(plain._moments_list))
        assert np.array_equal(np.array(back._moments_list), np.array(plain._moments_list))
        if prec == "complex128":
            err = np.max(np.abs(np.array(plain._moments_list)[:19, :101]
                                - np.array(ora._moments_list))) / h.shape[0]
            assert err < 1e-11, err


def test_tiles_true_from_site_positions():
    """tiles=True tiles a system by its site positions (here the package's own
    LatticeSystem, which has kwant's pos() and site_ranges)."""
    from kwant_amd import kpm
    from kwant_amd.assembly import LatticeSystem
    syst = LatticeSystem([(1, 0), (0, 1)], [(0, 0)], [(0, 40), (0, 40)], norbs=2)
    syst.set_onsite(lambda pos: np.tile(np.array([[0.3, 0.1j], [-0.1j, -0.2]]),
                                        (len(pos), 1, 1)))
    syst.add_hoppings((1, 0), 0, 0, 1.0)
    syst.add_hoppings((0, 1), 0, 0, 1.0)
    kw = dict(num_moments=40, num_vectors=16, bounds=(-5, 5), rng=2)
    tiled = kpm.SpectralDensity(syst, tiles=True, **kw)
    assert tiled._tiling is not None and tiled._blocks[0].rec.tiles is not None
    plain = kpm.SpectralDensity(syst, reorder=tiled._tiling.order, **kw)
    assert np.array_equal(np.array(tiled._moments_list), np.array(plain._moments_list))


def test_tiles_auto_on_large_systems(monkeypatch):
    """The default tiles='auto' tiles a kwant system with site positions from
    2^18 orbitals on (plain DOS path, default row order, a long enough
    expansion) and nothing else: not a smaller system, not an operator run, not
    an explicit row order, not a short expansion.  The moments equal those of
    the same row order without tiles, bitwise."""
    from kwant_amd import kpm
    from kwant_amd.assembly import LatticeSystem

    def square(side):
        syst = LatticeSystem([(1, 0), (0, 1)], [(0, 0)], [(0, side), (0, side)], norbs=1)
        syst.set_onsite(lambda pos: 0.3 * np.cos(1.7 * pos[:, 0] + 0.9 * pos[:, 1]))
        syst.add_hoppings((1, 0), 0, 0, 1.0)
        syst.add_hoppings((0, 1), 0, 0, 1.0)
        return syst

    kw = dict(num_moments=24, num_vectors=16, bounds=(-4.5, 4.5), rng=3)
    big = square(512)
    assert big.site_ranges[-1][2] == kpm.TILES_AUTO_ROWS
    # 2^18 x 16 x 12 orbital-vector-steps: below the default work threshold
    assert kpm.SpectralDensity(big, **kw)._tiling is None
    monkeypatch.setattr(kpm, "TILES_AUTO_WORK", 1 << 20)
    auto = kpm.SpectralDensity(big, **kw)
    assert auto._tiling is not None
    plain = kpm.SpectralDensity(big, tiles=None, reorder=auto._tiling.order, **kw)
    assert np.array_equal(np.array(auto._moments_list), np.array(plain._moments_list))
    assert kpm.SpectralDensity(square(40), **kw)._tiling is None
    assert kpm.SpectralDensity(big, reorder=False, **kw)._tiling is None
    dens = np.ones(big.site_ranges[-1][2])
    assert kpm.SpectralDensity(big, operator=__import__("scipy").sparse.diags(dens),
                               **kw)._tiling is None


def test_tiles_graph_for_a_matrix():
    """tiles='graph' runs a plain matrix (no positions) over tiles of its
    graph's hop distances: two steps per launch, bitwise the moments of the
    same row order without tiles."""
    from kwant_amd import kpm
    from kwant_amd.lattices import honeycomb, honeycomb_bounds
    h, _ = honeycomb(60, disorder=0.5, seed=4)
    kw = dict(num_moments=31, num_vectors=16, bounds=honeycomb_bounds(1.0, 0.5), rng=5)
    tiled = kpm.SpectralDensity(h, tiles='graph', **kw)
    assert tiled._tiling is not None and tiled._blocks[0].rec.tiles is not None
    plain = kpm.SpectralDensity(h, reorder=tiled._tiling.order, **kw)
    assert np.array_equal(np.array(tiled._moments_list), np.array(plain._moments_list))
