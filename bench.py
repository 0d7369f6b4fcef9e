#!/usr/bin/env python
"""Benchmark: KPM Chebyshev steps/s on one MI355X.

Workload (the headline configuration): graphene honeycomb parallelogram of
1000 x 1000 unit cells (N = 2,000,000 sites, nnz = 7,994,002), NN hopping
t = 1, Anderson disorder U[-0.25, 0.25], 64 random-phase vectors in one
N x 64 block, complex64 recursion with FP64 dot accumulation; a full
expansion is 2048 moments = 1024 Chebyshev steps.

A "step" is one Chebyshev step over the whole block:
a_{n+1} = 2 (H a_n - b a_n)/a - a_{n-1} plus the two FP64 dot products per
column (two moments per vector).  Two kernels run it: the one-step kernel (one
launch per step, natural row order) and the two-step kernel (one launch per two
steps over compact tiles of the flake: each tile's a_{n+1} and its halo's are
formed in LDS, so a_{n+1} never makes a round trip through HBM).  Both are
timed; the faster is the headline.  The timed region holds only kernel
launches on the current HIP stream, bracketed by events; inputs are resident
in HBM before it starts.

Algorithmic bytes per step (the roofline numerator, per Chebyshev step of
the path, whichever kernel runs it):
    B = 3 R N s_v + nnz (s_m + 4) + 4 (N + 1)
with s_v = s_m = 8 bytes (complex64): read a_n once, read a_{n-1}, write
a_{n+1}, read the CSR once.  The two-step kernel's own bytes per launch
(halo rows recomputed by every tile that needs them) are reported beside it.

Also reported: complex64 moments against the complex128 GPU path on the same
vectors (the complex128 path itself matches the CPU oracle to 1e-11 in
tests/), and the CPU baseline: the oracle's restatement of kwant's
per-vector loop (scipy csr @ v + numpy temporaries + vdot, complex128) timed
on a bounded sample on one host core.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILES = [os.path.join(ROOT, "profiles", f) for f in
                 ("cheb_step_c64_graphene2M_traffic.txt", "cheb2_tiled_c64_graphene2M_traffic.txt")]


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def bytes_per_step(n, nnz, width, s_v, s_m):
    return 3 * width * n * s_v + nnz * (s_m + 4) + 4 * (n + 1)


def stream_probe(log_fn):
    """Achievable HBM rates on this GPU from torch's own kernels on 1 GiB
    complex64 tensors: a copy (1 read : 1 write) and an add c = a + b
    (2 reads : 1 write, the recursion's mix), HIP-event timed."""
    import torch
    n = 1 << 27
    a = torch.ones(n, dtype=torch.complex64, device="cuda")
    b = torch.ones(n, dtype=torch.complex64, device="cuda")
    c = torch.empty_like(a)
    out = {}
    for name, fn, nbytes in (("copy", lambda: c.copy_(a), 2 * 8 * n),
                             ("add", lambda: torch.add(a, b, out=c), 3 * 8 * n)):
        fn()
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(10):
            fn()
        end.record()
        torch.cuda.synchronize()
        out[name] = nbytes * 10 / (start.elapsed_time(end) * 1e-3) / 1e9
    log_fn("stream probe (torch, 1 GiB): copy %.0f GB/s, add %.0f GB/s" % (out["copy"], out["add"]))
    del a, b, c
    torch.cuda.empty_cache()
    return out


def read_traffic(variant):
    """Bytes per launch fetched past L2 plus written, measured with
    rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration) for this launch shape ("variant V" or "tiled V"), if
    committed; else None."""
    key = "traffic_bytes_per_launch %s:" % variant
    for path in TRAFFIC_FILES:
        try:
            for line in open(path):
                if line.startswith(key):
                    return float(line.split(":", 1)[1])
        except OSError:
            pass
    return None


def two_step_bench(h, pos, ab, block, width, warmup, steps, tile_rows, variant,
                   one_step_variant, log_fn, precision="complex64"):
    """The same expansion with two Chebyshev steps per launch: the flake's
    rows in a tiled order (kwant_amd.ordering.tile_order, tiles of about
    ``tile_rows`` rows), the vectors permuted alike.  Every launch shape whose
    LDS holds a tile and its halo is timed briefly and the fastest runs the
    timed steps; its dots are checked bitwise against one-step launches on the
    same rows first."""
    import numpy as np
    import torch
    from kwant_amd import _engine as eng
    from kwant_amd.ordering import tile_order
    n, nnz = h.shape[0], h.nnz
    t0 = time.time()
    tiling = tile_order(pos, tile_rows)
    cdt, vt = eng.complex_dtype(precision)
    s_v = 16 if vt == eng.KPM_C128 else 8
    mat = eng.DeviceCSR(h, precision, order=tiling.order)
    blk = block[torch.from_numpy(tiling.order).to(block.device)].to(cdt)
    log_fn("two-step, %s: %d tiles of <= %d rows (%.1fs)"
           % (precision, tiling.n_tiles, int(np.max(np.diff(tiling.tile_begin))),
              time.time() - t0))
    shapes = [variant] if variant >= 0 else list(range(eng.tiled_variant_count(width, vt)))
    tried = {}
    best = None
    for v in shapes:
        try:
            tp = eng.TiledPlan(mat, width, tiling.tile_begin, variant=v)
        except (RuntimeError, ValueError) as err:
            log_fn("two-step shape %d: %s" % (v, err))
            continue
        rec = eng.BlockRecursion(mat, ab, blk, width, eng.DOTS_DOUBLING, tiles=tp)
        rec.step(9)
        torch.cuda.synchronize()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        rec.step(32)
        end.record()
        torch.cuda.synchronize()
        tried[v] = start.elapsed_time(end) / 32
        if best is None or tried[v] < tried[best[0]]:
            best = (v, tp)
        del rec
    if best is None:
        return None
    v, tp = best
    info = tp.info()
    # bitwise against one-step launches on the same rows
    check = 64
    ref = eng.BlockRecursion(mat, ab, blk, width, eng.DOTS_DOUBLING,
                             variant=max(one_step_variant, 0))
    d_ref = ref.step(check)
    del ref
    rec = eng.BlockRecursion(mat, ab, blk, width, eng.DOTS_DOUBLING, tiles=tp)
    d_two = rec.step(check)
    bitwise = bool(np.array_equal(d_ref, d_two))
    del rec
    rec = eng.BlockRecursion(mat, ab, blk, width, eng.DOTS_DOUBLING, tiles=tp)
    pairs = steps // 2
    dots = torch.empty((warmup + 1 + 2 * pairs, 1, width, 4), dtype=torch.float64,
                       device=block.device)
    rec.step(warmup + 1 if warmup % 2 == 0 else warmup)   # odd: step 0 alone, then pairs
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    tp.steps(rec.plan, mat, rec.ring, rec.slot_stride, rec.slots, pairs, eng.DOTS_DOUBLING,
             dots=dots, out_base=0)
    end.record()
    torch.cuda.synchronize()
    ms_launch = start.elapsed_time(end) / pairs
    assert rec.plan.get_step() == rec.n + 2 * pairs
    kb = tp.kernel_bytes(nnz, s_v, s_v)
    gw, waves, _ = eng.tiled_variant_info(width, vt, v)
    out = {"precision": precision, "ms_per_step": ms_launch / 2, "ms_per_launch": ms_launch,
           "steps": 2 * pairs,
           "variant": v, "group_width": gw, "waves": waves, "tile_rows": tile_rows,
           "tiles": info["tiles"], "max_tile_rows": info["max_tile_rows"],
           "halo_rows": info["halo_rows"], "gather_rows": info["gather_rows"],
           "workgroups": info["workgroups"], "lds_bytes": info["lds_bytes"],
           "kernel_bytes_per_launch": kb,
           "kernel_GBs": kb / (ms_launch * 1e-3) / 1e9,
           "kernel_frac_of_hbm_peak": kb / (ms_launch * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "shapes_ms_per_step": {str(k): t for k, t in tried.items()},
           "dots_bitwise_vs_one_step": bitwise, "checked_steps": check}
    log_fn("two-step shape %d, %s: %.4f ms/step (%.4f ms per launch), kernel bytes %.3f GB "
           "per launch = %.0f GB/s; dots bitwise vs one-step: %s"
           % (v, precision, ms_launch / 2, ms_launch, kb / 1e9, out["kernel_GBs"], bitwise))
    del rec, dots, mat, blk
    return out


def moments_from_dots(dots):
    """(steps, W, 4) dots of a fresh expansion -> (W, 2 * steps) moments."""
    import numpy as np
    steps, width, _ = dots.shape
    mom = np.zeros((width, 2 * steps), dtype=np.complex128)
    mu0 = dots[0, :, 0] + 0j
    mu1 = np.conj(dots[0, :, 2] + 1j * dots[0, :, 3])
    mom[:, 0], mom[:, 1] = mu0, mu1
    d0 = dots[1:, :, 0]
    d1 = dots[1:, :, 2] + 1j * dots[1:, :, 3]
    mom[:, 2::2] = (2 * d0 - mu0).T
    mom[:, 3::2] = (2 * d1 - mu1).T
    return mom


def cpu_baseline(h, a, b, vec, steps):
    """The oracle's kwant loop for one vector: seconds per vector-step."""
    from threadpoolctl import threadpool_limits
    from oracle.kpm_oracle import time_doubling_steps
    with threadpool_limits(limits=1):
        time_doubling_steps(h, a, b, vec, 2)
        t0 = time.perf_counter()
        time_doubling_steps(h, a, b, vec, steps)
        return (time.perf_counter() - t0) / (steps + 1)


_POOL_INPUT = None   # (h, a, b), set before the pool forks: inherited, not pickled


def _pool_worker(args):
    seed, steps = args
    h, a, b = _POOL_INPUT
    import numpy as np
    from threadpoolctl import threadpool_limits
    from oracle.kpm_oracle import time_doubling_steps
    vec = np.exp(2j * np.pi * np.random.RandomState(seed).random_sample(h.shape[0]))
    with threadpool_limits(limits=1):
        time_doubling_steps(h, a, b, vec, 2)
        t0 = time.perf_counter()
        time_doubling_steps(h, a, b, vec, steps)
        return (time.perf_counter() - t0) / (steps + 1)


def cpu_baseline_pool(h, a, b, steps, width, max_procs=16):
    """The oracle's per-vector loop on all host cores: one process per core
    (at most ``max_procs``), each running its own vector, 1 thread each."""
    import multiprocessing as mp
    procs = max(1, min(max_procs, len(os.sched_getaffinity(0))))
    model = "unknown CPU"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    global _POOL_INPUT
    _POOL_INPUT = (h, a, b)
    with mp.get_context("fork").Pool(procs) as pool:
        per = pool.map(_pool_worker, [(100 + i, steps) for i in range(procs)])
    _POOL_INPUT = None
    t_vs = sum(per) / len(per)
    affinity = len(os.sched_getaffinity(0))
    return {"value": procs / (t_vs * width), "unit": "steps/s", "cores": procs,
            "kind": "port", "cpu_model": model,
            "affinity_cores": affinity, "process_cap": max_procs,
            "sample": "%d processes (the %d cores this process may run on, capped at %d) x "
                      "1 vector x %d steps of the oracle loop, %.1f ms per vector-step each, "
                      "aggregate scaled to %d-vector block steps"
                      % (procs, affinity, max_procs, steps, 1e3 * t_vs, width)}


def small_square_bench(log_fn):
    """The reference's own CPU case in full: square 100 x 100, U[-0.5, 0.5]
    disorder, 10 random-phase vectors, 100 moments, Jackson kernel -- the
    GPU engine and the oracle on the same seeded input."""
    import numpy as np
    from kwant_amd import kpm
    from kwant_amd.lattices import square
    from oracle import kpm_oracle as ko
    h, _ = square(100, disorder=1.0, seed=1)
    kw = dict(num_vectors=10, num_moments=100, rng=0)
    t0 = time.perf_counter()
    ko.SpectralDensity(h, **kw)
    t_cpu = time.perf_counter() - t0
    kpm.SpectralDensity(h, **kw)           # first call: library and plan setup
    t0 = time.perf_counter()
    kpm.SpectralDensity(h, **kw)
    t_gpu = time.perf_counter() - t0
    from kwant_amd.surrogate import dos_l1
    bounds = (-4.6, 4.6)
    g = kpm.SpectralDensity(h, bounds=bounds, **kw)
    o = ko.SpectralDensity(h, bounds=bounds, **kw)
    e_g, rho_g = g()
    e_o, rho_o = o()
    err = float(np.max(np.abs(rho_g - rho_o)) / np.max(np.abs(rho_o)))
    # DOS L1 (relative to int rho) of each GPU precision against the FP64 oracle
    ref = np.real(o._moments())
    l1 = {}
    for prec in ("complex128", "complex64"):
        m = np.real(kpm.SpectralDensity(h, bounds=bounds, precision=prec, **kw)._moments())
        l1[prec] = float(dos_l1(m[None] / ref[0], ref[None] / ref[0])[0])
    log_fn("small square case (100x100, 10 vectors, 100 moments): GPU %.3f s, CPU oracle %.3f s "
           "(wall, bounds by Lanczos / ARPACK); densities with explicit bounds agree to %.1e; "
           "DOS L1 vs the FP64 oracle: complex128 %.1e, complex64 %.1e"
           % (t_gpu, t_cpu, err, l1["complex128"], l1["complex64"]))
    return {"workload": "square 100x100, U[-0.5,0.5], 10 vectors, 100 moments, Jackson, "
                        "complex128", "gpu_seconds": t_gpu, "cpu_oracle_seconds": t_cpu,
            "densities_max_rel_diff_explicit_bounds": err,
            "dos_l1_vs_fp64_oracle": l1}


def _stratified_checks(count, seed=11):
    """Parameter points for judging the surrogate: half uniform over the
    sweep box, a quarter in the clean limit W < 0.5 and a quarter at strong
    flux phi > 0.015 (where its fallback concentrates)."""
    import numpy as np
    from kwant_amd.surrogate import PARAM_LO, PARAM_HI
    rng = np.random.RandomState(seed)
    uni = PARAM_LO + (PARAM_HI - PARAM_LO) * rng.random_sample((count // 2, 3))
    lo_w = PARAM_LO + (PARAM_HI - PARAM_LO) * rng.random_sample((count // 4, 3))
    lo_w[:, 0] *= 0.5 / PARAM_HI[0]
    hi_f = PARAM_LO + (PARAM_HI - PARAM_LO) * rng.random_sample((count - count // 2 - count // 4,
                                                                 3))
    hi_f[:, 1] = 0.015 + (PARAM_HI[1] - 0.015) * rng.random_sample(len(hi_f))
    return np.concatenate([uni, lo_w, hi_f]), {"uniform": slice(0, count // 2),
                                               "W<0.5": slice(count // 2, count // 2 + count // 4),
                                               "phi>0.015": slice(count // 2 + count // 4, count)}


def _converged(wide, params, sets):
    """Damped moments from sets x 32 random vectors (seeded sets, the
    reference's v0 offset skipped), the judge of the surrogate's DOS error.
    Also returns the means of the even and the odd sets: half the DOS L1
    distance between them estimates the reference's own sampling error."""
    import numpy as np
    import torch
    halves = np.zeros((2, len(params), wide.num_moments))
    for k in range(sets):
        rng = np.random.RandomState(1000 + k)
        rng.random_sample(wide.n)
        vecs = np.stack([np.exp(2j * np.pi * rng.random_sample(wide.n))
                         for _ in range(wide.num_vectors)], 1)
        wide._vectors = torch.from_numpy(vecs).to(wide._vectors.device).to(wide.cdt)
        halves[k % 2] += wide.damped_moments(params)
    out = halves.sum(axis=0) / sets
    return out, halves[0] / ((sets + 1) // 2), halves[1] / max(sets // 2, 1)


def surrogate_bench(epochs, tol, points, check_count, check_sets, log_fn, design="shift"):
    """The moment surrogate on the square-lattice family (512 x 512 sites,
    W in [0, 4], flux in [0, 0.02], mu in [-1, 1], 512 moments, 32 vectors).

    design 'shift': the network maps (W, flux) -> 4 x 256 GELU -> 512
    tight-window moments and mu enters by the exact shift; trained on 4096 +
    1024 + 1024 GPU-KPM runs (the extra ones in the clean limit W < 0.5 and
    at strong flux > 0.015).
    design 'literal': the configuration's network (W, flux, mu) -> 4 x 256
    GELU -> 512 wide-window moments, trained on 4096 GPU-KPM runs of
    H(W, flux, mu).

    Then: a sweep of ``points`` random points with the spread gate and KPM
    fallback, timed against plain GPU KPM, and its DOS L1 error against
    converged KPM (``check_sets`` x 32 vectors) on ``check_count`` stratified
    points (half uniform, a quarter at W < 0.5, a quarter at flux > 0.015)."""
    import numpy as np
    import torch
    from scipy.stats import qmc
    from kwant_amd.surrogate import (SquareFamily, MomentSurrogate, dos_l1, PARAM_LO,
                                     PARAM_HI, TIGHT_BOUNDS, WIDE_BOUNDS)
    out = {}
    wide = SquareFamily(l=512, num_moments=512, num_vectors=32, batch=8, bounds=WIDE_BOUNDS)
    tight = SquareFamily(l=512, num_moments=512, num_vectors=32, batch=8, bounds=TIGHT_BOUNDS)
    literal = design == "literal"
    layers = (3 if literal else 2, 256, 256, 256, 256, 512)
    sur = MomentSurrogate(wide, tight, members=5, layers=layers, seed=0, tol=tol)
    if literal:
        train = qmc.scale(qmc.Sobol(3, scramble=True, seed=8).random(4096), PARAM_LO, PARAM_HI)
        batches = [(train[s:s + 512], 100 + s) for s in range(0, 4096, 512)]
    else:
        train = qmc.scale(qmc.Sobol(2, scramble=True, seed=8).random(4096), PARAM_LO[:2],
                          PARAM_HI[:2])
        clean = qmc.scale(qmc.Sobol(2, scramble=True, seed=9).random(1024), PARAM_LO[:2],
                          np.array([0.5, PARAM_HI[1]]))
        # strong flux: the Landau-level peaks need the data too
        strong = qmc.scale(qmc.Sobol(2, scramble=True, seed=10).random(1024),
                           np.array([PARAM_LO[0], 0.015]), PARAM_HI[:2])
        batches = [(train[s:s + 512], 100 + s) for s in range(0, 4096, 512)]
        batches += [(clean[s:s + 512], 9000 + s) for s in range(0, 1024, 512)]
        batches += [(strong[s:s + 512], 12000 + s) for s in range(0, 1024, 512)]
        train = np.concatenate([train, clean, strong])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    parts = []
    for prm, seed in batches:
        parts.append(sur.training_targets(prm, fresh_seed=seed))
        log_fn("surrogate (%s): training runs %d done" % (design, sum(len(p) for p in parts)))
    targets = np.concatenate(parts)
    t_data = time.perf_counter() - t0
    log_fn("surrogate (%s): %d training runs in %.1f s" % (design, len(train), t_data))
    t0 = time.perf_counter()
    # healthy members train to within ~20 % of each other's error; one at
    # 3x the median or more (measured: 4.6x, with 47 % of the sweep sent to
    # KPM by its spread) is left out
    sur.fit(train, targets, epochs=epochs, batch=64, lr=3e-3, ema_decay=0.999, health=3.0,
            log=lambda msg: log_fn("surrogate (%s) %s" % (design, msg)))
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    log_fn("surrogate (%s): %d epochs x %d Adam steps, 5 members (%d kept), in %.1f s"
           % (design, epochs, len(train) // 64, sur.active, t_fit))
    sweep = PARAM_LO + (PARAM_HI - PARAM_LO) * np.random.RandomState(7).random_sample((points, 3))
    sur.predict(sweep[:64])
    moments, fall, t_sweep = sur.sweep(sweep)
    log_fn("surrogate (%s): sweep of %d points in %.3f s (%.2f%% KPM fallback)"
           % (design, points, t_sweep, 100 * fall.mean()))
    _, spread = sur.predict(sweep)
    spread = spread.cpu().numpy()
    sample = sweep[:64]
    wide.damped_moments(sample[:8])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    wide.damped_moments(sample)
    t_point = (time.perf_counter() - t0) / len(sample)
    t_plain = t_point * points
    log_fn("plain KPM: %.2f ms per point -> %.1f s for the sweep" % (1e3 * t_point, t_plain))
    # accuracy on the stratified points, against converged KPM
    checks, strata = _stratified_checks(check_count)
    mom_c, fall_c, _ = sur.sweep(checks)
    _, spread_c = sur.predict(checks)
    spread_c = spread_c.cpu().numpy()
    kpm_c = wide.damped_moments(checks)          # plain KPM, the seeded 32 vectors
    converged, half_a, half_b = _converged(wide, checks, check_sets)
    err = dos_l1(mom_c, converged)
    err_kpm = dos_l1(kpm_c, converged)
    ref_noise = 0.5 * dos_l1(half_a, half_b)
    fall_c = np.asarray(fall_c, dtype=bool)
    per = {}
    for name, sl in strata.items():
        e, f = err[sl], fall_c[sl]
        acc = e[~f]
        per[name] = {"points": int(sl.stop - sl.start),
                     "fallback_fraction": float(f.mean()),
                     "dos_l1_median": float(np.median(e)),
                     "dos_l1_max": float(e.max()),
                     "surrogate_points": int(acc.size),
                     "surrogate_dos_l1_max": float(acc.max()) if acc.size else None,
                     "plain_kpm32_dos_l1_median": float(np.median(err_kpm[sl])),
                     "reference_noise_median": float(np.median(ref_noise[sl]))}
        log_fn("DOS L1 vs converged KPM (%d vectors), %s: sweep median %.2e max %.2e "
               "(fallback %.1f%%), max over the %d surrogate points %s; plain KPM(32) "
               "median %.2e; the reference's own noise ~%.1e"
               % (32 * check_sets, name, per[name]["dos_l1_median"], per[name]["dos_l1_max"],
                  100 * per[name]["fallback_fraction"], acc.size,
                  "%.2e" % acc.max() if acc.size else "-",
                  per[name]["plain_kpm32_dos_l1_median"], per[name]["reference_noise_median"]))
    net = ("(W, flux, mu) -> 4x256 GELU -> 512 wide-window moments, trained on 4096 runs"
           if literal else
           "(W, flux) -> 4x256 GELU -> 512 + exact shift in mu, trained on 4096 + 1024 "
           "(W < 0.5) + 1024 (flux > 0.015) runs")
    out.update({
        "workload": "square 512x512, W in [0,4], flux in [0,0.02], mu in [-1,1], 512 moments, "
                    "32 vectors; MLP %s; 5-member ensemble; %d sweep points" % (net, points),
        "design": design,
        "speedup_vs_plain_kpm": t_plain / t_sweep,
        "sweep_s": t_sweep, "plain_kpm_sweep_s": t_plain, "plain_kpm_ms_per_point": 1e3 * t_point,
        "fallback_fraction": float(fall.mean()), "spread_tol": tol,
        "dos_l1_vs_converged_median": float(np.median(err)),
        "dos_l1_vs_converged_p90": float(np.percentile(err, 90)),
        "dos_l1_vs_converged_max": float(err.max()),
        "surrogate_points_dos_l1_max": float(err[~fall_c].max()) if (~fall_c).any() else None,
        "reference_noise_median": float(np.median(ref_noise)),
        "plain_kpm32_dos_l1_vs_converged_median": float(np.median(err_kpm)),
        "dos_l1_by_stratum": per,
        "converged_reference": "KPM with %d random vectors on %d stratified points"
                               % (32 * check_sets, check_count),
        "training_data_s": t_data, "training_s": t_fit, "epochs": epochs,
        "members_kept": int(sur.active),
        "projected_other_tolerances": {
            "%.1e" % t: {"fallback_fraction": float(np.mean(spread > t)),
                         "speedup_vs_plain_kpm":
                             t_plain / (t_sweep - fall.mean() * t_plain
                                        + np.mean(spread > t) * t_plain)}
            for t in (1.5e-3, 2e-3)},
        # tighter gates, judged on the same stratified points (a point
        # accepted at a tighter gate is also a surrogate point at tol)
        "tighter_gates": {
            "%.1e" % t: {"fallback_fraction": float(np.mean(fall | (spread > t))),
                         "speedup_vs_plain_kpm":
                             t_plain / (t_sweep - fall.mean() * t_plain
                                        + np.mean(fall | (spread > t)) * t_plain),
                         "surrogate_dos_l1_max_by_stratum": {
                             name: (float(err[sl][~fall_c[sl] & (spread_c[sl] <= t)].max())
                                    if (~fall_c[sl] & (spread_c[sl] <= t)).any() else None)
                             for name, sl in strata.items()}}
            for t in (5e-4, 7e-4) if t < tol},
    })
    return out


def surrogate_train_profile(epochs, log_fn):
    """The surrogate's training and prediction kernels alone, on synthetic
    targets (smooth random moment curves of the 512-moment shape): the
    workload of the committed MLP kernel profile.  Kernel timings do not
    depend on the target values."""
    import numpy as np
    import torch
    from kwant_amd.surrogate import MomentSurrogate, PARAM_LO, PARAM_HI, jackson_factors

    class Family:
        num_moments = 512
        a, b = 7.0 / 0.975, 0.0
        g = jackson_factors(512)

    tight = Family()
    tight.a = 6.0 / 0.975
    sur = MomentSurrogate(Family(), tight, members=5, seed=0, tol=1e-3)
    rng = np.random.RandomState(0)
    params = PARAM_LO + (PARAM_HI - PARAM_LO) * rng.random_sample((4096, 3))
    n = np.arange(512)
    targets = (np.cos(np.outer(params[:, 0], n) * 0.01) * np.exp(-n / 200.0) * tight.g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sur.fit(params, targets.astype(np.float32), epochs=epochs, batch=64, lr=3e-3)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    sur.predict(params)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sur.predict(params)
    torch.cuda.synchronize()
    t_pred = time.perf_counter() - t0
    iters = epochs * (4096 // 64)
    log_fn("surrogate training: %d epochs x 64 steps, 5 members: %.3f s (%.3f ms per "
           "iteration); 4096-point prediction %.2f ms" % (epochs, t_fit, 1e3 * t_fit / iters,
                                                          1e3 * t_pred))
    return {"workload": "5-member MLP 2 -> 4x256 GELU -> 512, batch 64, synthetic targets",
            "epochs": epochs, "fit_s": t_fit, "ms_per_iteration": 1e3 * t_fit / iters,
            "predict_4096_ms": 1e3 * t_pred}


def density_bench(mat, host_h, ab, block, width, variant, dos_ms, steps, log_fn):
    """Site-resolved density (kwant.operator.Density, sum=False, mean=True)
    on the same block: each step also reads w = M^H a_0 (here M = 1, so w =
    a_0) and writes one complex128 per row, the sum over the block's
    vectors of conj(w) a_{n+1}.  Kernel time only; the per-site sums and
    the copy to the host are outside the timed region."""
    import torch
    from kwant_amd import _engine as eng
    from kwant_amd import autotune
    n, nnz = mat.shape[0], mat.nnz
    w = block.clone()
    # the launch shape SpectralDensity would tune for this mode (variant < 0)
    if variant is None or variant < 0:
        variant = autotune.pick_variant(mat, ab, width, eng.DOTS_LOCAL, force=True)
    rec = eng.BlockRecursion(mat, ab, block, width, eng.DOTS_LOCAL, w=w, variant=variant)
    chunk = 64
    rec.local_steps(chunk)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps // chunk):
        rec.local_steps(chunk)
    end.record()
    torch.cuda.synchronize()
    ms_step = start.elapsed_time(end) / (chunk * (steps // chunk))
    bstep = bytes_per_step(n, nnz, width, 8, 8) + n * width * 8 + n * 16
    gbs = bstep / (ms_step * 1e-3) / 1e9
    log_fn("site-resolved density: %.4f ms/step (DOS step %.4f ms), %.0f GB/s algorithmic "
           "(%.1f%% of 8 TB/s)" % (ms_step, dos_ms, gbs, 100 * gbs / HBM_PEAK_GBS))
    ket = rec.slot(rec.n).clone()
    del rec
    torch.cuda.empty_cache()
    # Current over every hopping (onsite 1: v = ket, w = bra = a_0): one
    # bond-record launch per moment, after the recursion step
    import numpy as np
    import scipy.sparse
    csr = scipy.sparse.csr_matrix(host_h)
    rows = np.repeat(np.arange(n), np.diff(csr.indptr))
    keep = rows != csr.indices
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows[keep], minlength=n), out=indptr[1:])
    hop = scipy.sparse.csr_matrix((csr.data[keep], csr.indices[keep], indptr), shape=csr.shape)
    bond = eng.DeviceCSR(hop, "complex64")
    out = torch.empty(bond.nnz, dtype=torch.complex128, device=block.device)
    reps = 32
    bond.bond_records(w, ket, ket, w, width, out)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        bond.bond_records(w, ket, ket, w, width, out)
    end.record()
    torch.cuda.synchronize()
    ms_bond = start.elapsed_time(end) / reps
    # distinct bytes: bra (= w) and ket (= v) once each, the entries, the records
    bbond = 2 * n * width * 8 + bond.nnz * (8 + 4) + 4 * (n + 1) + bond.nnz * 16
    gbond = bbond / (ms_bond * 1e-3) / 1e9
    log_fn("current (bond records, %d hopping entries): %.4f ms per moment, %.0f GB/s "
           "algorithmic (%.1f%% of 8 TB/s)" % (bond.nnz, ms_bond, gbond,
                                               100 * gbond / HBM_PEAK_GBS))
    del w, ket, out, bond
    torch.cuda.empty_cache()
    return {"mode": "KPM_DOTS_LOCAL (per-row sum over the 64 vectors of conj(w) a_n+1)",
            "ms_per_step": ms_step, "dos_ms_per_step": dos_ms,
            "algorithmic_bytes_per_step": bstep, "achieved_GBs": gbs,
            "frac_of_hbm_peak": gbs / HBM_PEAK_GBS, "steps": chunk * (steps // chunk),
            "variant": variant,
            "current": {"kernel": "bond_kernel (kpm_bond_records, onsite 1, every hopping)",
                        "hopping_entries": int(hop.nnz),
                        "ms_per_moment": ms_bond, "algorithmic_bytes": bbond,
                        "achieved_GBs": gbond, "frac_of_hbm_peak": gbond / HBM_PEAK_GBS,
                        "note": "per moment: this launch plus one DOS-shaped recursion step "
                                "without dots"}}



def transform_bench(moments, channels, log_fn):
    """Moment-to-energy transform of a site-resolved expansion: the DCT-III
    of (moments x channels) complex moments to 2 x moments nodes, on the GPU
    (FP64 matrix-core GEMM, host copies included) and with scipy on one host
    core (the reference's transform)."""
    import numpy as np
    import torch
    from kwant_amd import kpm
    rng = np.random.default_rng(0)
    mom = rng.normal(size=(moments, channels)) + 1j * rng.normal(size=(moments, channels))
    kpm._calc_fft_moments(mom[:, :64].copy())         # warm-up (library, GEMM)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rho_gpu, _ = kpm._calc_fft_moments(mom)
    t_gpu = time.perf_counter() - t0
    sub = max(1, channels // 16)
    t0 = time.perf_counter()
    saved = kpm.GPU_TRANSFORM_MIN
    kpm.GPU_TRANSFORM_MIN = 1 << 62
    rho_cpu, _ = kpm._calc_fft_moments(mom[:, :sub])
    kpm.GPU_TRANSFORM_MIN = saved
    t_cpu = (time.perf_counter() - t0) * channels / sub
    err = float(np.max(np.abs(rho_gpu[:, :sub] - rho_cpu)) / np.max(np.abs(rho_cpu)))
    log_fn("transform (DCT-III, %d moments x %d channels): GPU %.3f s (host copies "
           "included), scipy %.2f s (1 core, from %d channels); max rel diff %.1e"
           % (moments, channels, t_gpu, t_cpu, sub, err))
    return {"moments": moments, "channels": channels, "gpu_s": t_gpu, "scipy_s": t_cpu,
            "scipy_sample_channels": sub, "max_rel_diff": err,
            "gemm_flops": 8.0 * 2 * moments * moments * channels}


def assembly_bench(cells, sample_cells, log_fn):
    """Hamiltonian assembly of a disordered graphene flake with a Peierls
    phase: kwant_amd.assembly (value functions evaluated over arrays) for
    cells x cells unit cells, against the oracle's one-site-at-a-time loop
    (kwant's make_sparse_full) on sample_cells^2 cells, scaled by the site
    count."""
    import numpy as np
    from kwant_amd.assembly import honeycomb_system
    from oracle import assembly_oracle

    def onsite(pos, w):
        x = np.sin(pos[:, 0] * 12.9898 + pos[:, 1] * 78.233) * 43758.5453
        return w * (x - np.floor(x) - 0.5)

    def hop(to, fr, t, phi):
        return -t * np.exp(1j * phi * (to[:, 0] - fr[:, 0]) * (to[:, 1] + fr[:, 1]))

    def build(c):
        syst = honeycomb_system(c)
        syst.set_onsite(onsite)
        syst._hops = [(d, src, hop) for d, src, _ in syst._hops]
        return syst

    params = dict(w=0.5, t=1.0, phi=1e-3)
    t0 = time.perf_counter()
    big = build(cells)
    h = big.hamiltonian_submatrix(params=params)
    t_vec = time.perf_counter() - t0
    small = build(sample_cells)
    t0 = time.perf_counter()
    ref = assembly_oracle.csr(small, params)
    t_loop = (time.perf_counter() - t0) * (cells / sample_cells) ** 2
    err = float(abs(small.hamiltonian_submatrix(params=params) - ref).max())
    log_fn("assembly (graphene %dx%d cells, %d entries): vectorised %.2f s, one-at-a-time "
           "loop %.0f s (from %dx%d cells); max |diff| %.1e on the sample"
           % (cells, cells, h.nnz, t_vec, t_loop, sample_cells, sample_cells, err))
    return {"cells": cells, "nnz": int(h.nnz), "vectorised_s": t_vec, "loop_s": t_loop,
            "loop_sample_cells": sample_cells, "max_abs_diff_sample": err}


def _time_ensemble(ens, graph):
    """(eager seconds, graph seconds) of one expansion of ``ens``."""
    t_eager = ens.run(graph=False)
    if graph:
        ens.run(graph=True)                  # capture + first replay
        t_graph = min(ens.run(graph=True) for _ in range(2))
    else:
        t_graph = min(t_eager, ens.run(graph=False))
    return t_eager, t_graph


def ensemble_bench(count, moments, width, log_fn, variant=-1, graph=True,
                   formats=("stencil", "csr"), shape=-1):
    """The batched disorder ensemble: ``count`` cubic 4-orbital systems with
    L ~ U{20..60}, generated on the GPU, ``moments`` moments with ``width``
    vectors each, every step one launch over all systems, the step sequence
    replayed from a captured HIP graph.  Both matrix formats run: the CSR
    (the general recursion kernel) and the lattice stencil (on-site blocks
    only, each system marched along x); their dots are compared bitwise, and
    the faster is the configuration's figure.  ``formats`` narrows the run
    to one format (for profiling); ``variant`` and ``shape`` fix the CSR and
    stencil launch shapes (-1: timed)."""
    import numpy as np
    import torch
    from kwant_amd.ensemble import CubicEnsemble, ensemble_sizes
    sizes = ensemble_sizes(count, 20, 60, seed=0)
    res, dots = {}, {}
    for fmt in formats:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ens = CubicEnsemble(sizes, seed=0, num_moments=moments, num_vectors=width,
                            variant=variant if fmt == "csr" else shape, matrix_format=fmt)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        t_eager, t_graph = _time_ensemble(ens, graph)
        dots[fmt] = ens.dots.cpu().numpy()
        bstep = ens.bytes_per_step()
        ms_step = 1e3 * t_graph / ens.steps
        achieved = bstep / (ms_step * 1e-3) / 1e9
        res[fmt] = {"ms_per_step": ms_step, "ms_per_step_eager": 1e3 * t_eager / ens.steps,
                    "expansion_s": t_graph, "setup_s": t_setup, "achieved_GBs": achieved,
                    "frac_of_hbm_peak": achieved / HBM_PEAK_GBS, "variant": ens.variant}
        if fmt == "stencil":
            own = ens.stencil_bytes_per_step()
            res[fmt]["kernel_bytes_per_step"] = own
            res[fmt]["kernel_frac_of_hbm_peak"] = own / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS
            res[fmt]["shape_waves_depth_tile_y"] = list(ens.march.SHAPES[ens.variant])
            res[fmt]["tuning_ms_per_step"] = {str(k): round(v, 4) for k, v in
                                              getattr(ens, "shape_ms", {}).items()}
        log_fn("ensemble (%s): %d systems, N=%d, W=%d: %.3f ms/step (graph), eager %.3f "
               "ms/step, %.0f GB/s of the path's bytes (%.1f%% of 8 TB/s); setup %.1f s"
               % (fmt, count, ens.n, width, ms_step, 1e3 * t_eager / ens.steps, achieved,
                  100 * achieved / HBM_PEAK_GBS, t_setup))
        n_rows, nnz, steps = ens.n, ens.nnz, ens.steps
        del ens
        torch.cuda.empty_cache()
    bitwise = None
    if len(dots) == 2:
        bitwise = bool(np.array_equal(dots["stencil"], dots["csr"]))
        log_fn("ensemble: stencil dots bitwise those of the CSR path: %s" % bitwise)
    best = min(res, key=lambda f: res[f]["ms_per_step"])
    out = {"workload": "%d cubic 4-orbital systems, L ~ U{20..60}, sigma 0.3, t 1, %d moments, "
                       "%d vectors each, complex64, one launch per step for all systems"
                       % (count, moments, width),
           "rows": n_rows, "nnz": nnz, "steps": steps, "matrix_format": best,
           "algorithmic_bytes_per_step": bytes_per_step(n_rows, nnz, width, 8, 8),
           "dots_bitwise_stencil_vs_csr": bitwise}
    out.update(res[best])
    for fmt in res:
        out[fmt] = res[fmt]
    return out


MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0, "int8": 5000.0}


def correlator_bench(cells, moments, vectors, width, log_fn, full=True):
    """Kubo-Bastin sigma_xy of the four-band model (cells x cells sites,
    A = 1, B = -1, M = -2, flux 1/cells per plaquette, Anderson W = 1):
    the Gamma_nm contraction (moments x moments from ``vectors`` random
    vectors) timed as one MFMA GEMM per precision at the real shape, and the
    full conductivity in each precision against FP64."""
    import numpy as np
    import torch
    from kwant_amd import _engine as eng
    from kwant_amd.lattices import dirac_4band
    from kwant_amd.correlator import _velocity, conductivity
    t0 = time.perf_counter()
    h, pos = dirac_4band(cells, A=1.0, B=-1.0, M=-2.0, flux=1.0 / cells, disorder=1.0, seed=4)
    n = h.shape[0]
    log_fn("four-band model %dx%d: N=%d nnz=%d (%.1f s)" % (cells, cells, n, h.nnz,
                                                        time.perf_counter() - t0))
    bound = float(np.max(np.abs(h).sum(axis=1)))
    eps = 0.05
    a, b = 2 * bound / (2 - eps), 0.0
    vx = _velocity(h, None, 'x', pos)
    vy = _velocity(h, None, 'y', pos)
    dev = eng.torch_device()
    m128, dvx, dvy = (eng.DeviceCSR(x, "complex128") for x in (h, vx, vy))
    rng = np.random.RandomState(5)
    block = torch.from_numpy(np.exp(2j * np.pi * rng.random_sample((n, width)))).to(dev)
    # right expansion, Psi = v_y T_n|r>
    rec_r = eng.BlockRecursion(m128, (a, b), block, width, eng.DOTS_NONE, store_slots=moments)
    rec_r.step(moments - 1, want_dots=False)
    psi = torch.empty_like(rec_r.ring)
    dvy.spmm(rec_r.ring, psi, width, n_slots=moments, x_stride=n * width, y_stride=n * width)
    del rec_r
    start = torch.empty_like(block)
    dvx.spmm(block, start, width)
    rec_l = eng.BlockRecursion(m128, (a, b), start, width, eng.DOTS_NONE, store_slots=moments)
    rec_l.step(moments - 1, want_dots=False)
    omega = rec_l.ring
    k = n * width
    flops = 8.0 * moments * moments * k
    out = {"workload": "four-band Dirac model %dx%d (N=%d), flux 1/%d, W=1; Gamma_nm %dx%d "
                       "from %d vectors; GEMM timed on one %d-vector block (K = N x %d)"
                       % (cells, cells, n, cells, moments, moments, vectors, width, width),
           "gemm": {}}
    ref = None
    work = None
    for name, prec, peak in (("fp64", eng.GEMM_FP64, "fp64"),
                             ("fp32", eng.GEMM_FP32, "fp32"), ("sliced3", 3, "bf16"),
                             ("sliced4", 4, "bf16"), ("isliced4", 4, "int8"),
                             ("isliced5", 5, "int8")):
        sliced = "sliced" in name
        kind = "int8" if name.startswith("i") else "bf16"
        if sliced:
            a_op, b_op = omega.reshape(moments, k), psi.reshape(moments, k)
        elif prec == eng.GEMM_FP32:
            a_op = omega.to(torch.complex64).reshape(moments, k)
            b_op = psi.to(torch.complex64).reshape(moments, k)
        else:
            a_op, b_op = omega.reshape(moments, k), psi.reshape(moments, k)
        mu = torch.zeros((moments, moments), dtype=torch.complex128, device=dev)

        def run():
            nonlocal work
            if sliced:
                work = eng.corr_gemm_sliced(a_op, b_op, mu, slices=prec, work=work, kind=kind)
            else:
                eng.corr_gemm(a_op, b_op, mu, prec)
        run()
        torch.cuda.synchronize()
        reps = 3
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            mu.zero_()
            run()
        e.record()
        torch.cuda.synchronize()
        sec = s.elapsed_time(e) / reps / 1e3
        gam = mu.cpu().numpy()
        if ref is None:
            ref = gam
        rel = float(np.linalg.norm(gam - ref) / np.linalg.norm(ref))
        # MFMA work actually issued: one product per kept slice pair
        # (S (S + 1) / 2) for the sliced GEMMs
        mult = prec * (prec + 1) // 2 if sliced else 1
        tflops = flops * mult / sec / 1e12
        out["gemm"][name] = {"seconds": sec, "algorithmic_TFLOPs": flops / sec / 1e12,
                             "mfma_TFLOPs": tflops, "mfma_peak_TFLOPs": MFMA_PEAK_TFLOPS[peak],
                             "frac_of_mfma_peak": tflops / MFMA_PEAK_TFLOPS[peak],
                             "rel_frobenius_err_vs_fp64": rel}
        log_fn("Gamma GEMM %s: %.3f s, %.1f TFLOP/s on MFMA (%.1f%% of %.1f), rel err vs FP64 "
               "%.2e" % (name, sec, tflops, 100 * tflops / MFMA_PEAK_TFLOPS[peak],
                         MFMA_PEAK_TFLOPS[peak], rel))
        del a_op, b_op, mu
    del omega, psi, rec_l, start, block, work
    torch.cuda.empty_cache()
    if full:
        sig, breakdown = {}, {}
        for name, prec in (("fp64", "complex128"), ("isliced5", "complex128"),
                           ("sliced4", "complex128"), ("fp32", "complex64")):
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            c = conductivity(h, alpha='x', beta='y', positions=pos, num_moments=moments,
                             num_vectors=vectors, bounds=(-bound, bound), rng=6,
                             precision=prec, gemm_precision=name)
            e_grid, s_grid = c.energies, np.array([c(mu=x, temperature=0.0)
                                                    for x in (-1.0, -0.5, 0.0, 0.5, 1.0)])
            torch.cuda.synchronize()
            sig[name] = (s_grid, time.perf_counter() - t1)
            # where the time went: the two streamed expansions (2 M steps
            # per block), the velocity SpMMs, the Omega x Psi GEMM chunks
            tm = dict(c.timings)
            widths = list(c.block_widths)
            s_v = 16 if prec == "complex128" else 8
            rec_bytes = sum(3 * w * n * s_v + h.nnz * (s_v + 4) + 4 * (n + 1) for w in widths)
            rec_steps = 2 * (moments - 1)
            breakdown[name] = {
                "block_widths": widths, "total_s": sig[name][1],
                "expansion_s": tm.get("expansion_s", 0.0),
                "velocity_spmm_s": tm.get("operator_s", 0.0),
                "gemm_s": tm.get("gemm_s", 0.0),
                "host_and_integral_factor_s": sig[name][1] - sum(tm.values()),
                # one recursion step over every block: algorithmic bytes / time
                "recursion_step_GBs": rec_bytes * rec_steps / max(tm.get("expansion_s", 0), 1e-12)
                / 1e9}
            breakdown[name]["recursion_frac_of_hbm_peak"] = \
                breakdown[name]["recursion_step_GBs"] / HBM_PEAK_GBS
            log_fn("sigma_xy %s: %.1f s total, blocks %s: expansions %.2f s (%.1f%% of 8 TB/s), "
                   "velocity SpMM %.2f s, GEMM %.2f s, host %.2f s"
                   % (name, sig[name][1], widths, breakdown[name]["expansion_s"],
                      100 * breakdown[name]["recursion_frac_of_hbm_peak"],
                      breakdown[name]["velocity_spmm_s"], breakdown[name]["gemm_s"],
                      breakdown[name]["host_and_integral_factor_s"]))
            # the cached device blocks are kept for the next precision: after
            # an empty_cache() the next run spends ~4 s in fresh device
            # allocations of its ~200 GB (measured)
            del c
        s64 = sig["fp64"][0]
        out["sigma_xy"] = {"mu": [-1.0, -0.5, 0.0, 0.5, 1.0],
                           "fp64": [float(np.real(v)) for v in s64]}
        for name in ("isliced5", "sliced4", "fp32"):
            d = float(np.max(np.abs(sig[name][0] - s64)) / np.max(np.abs(s64)))
            out["sigma_xy"][name + "_max_rel_dev"] = d
        for name in sig:
            out["sigma_xy"][name + "_seconds"] = sig[name][1]
        out["sigma_xy"]["breakdown"] = breakdown
        log_fn("sigma_xy (FP64) at mu = -1..1: %s; max rel dev isliced5 %.2e, sliced4 %.2e, "
               "FP32 %.2e" % (np.round(np.real(s64), 4), out["sigma_xy"]["isliced5_max_rel_dev"],
                              out["sigma_xy"]["sliced4_max_rel_dev"],
                              out["sigma_xy"]["fp32_max_rel_dev"]))
    return out


def calibrate():
    """Kernels of known HBM traffic: copy 2 GiB (2 GiB read + 2 GiB written)
    and a sum over 2 GiB (2 GiB read), three times each."""
    import torch
    from kwant_amd import _engine as eng
    dev = eng.torch_device()
    x = torch.ones(2 ** 28, dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
        float(x.sum())
    torch.cuda.synchronize()
    log("calibration kernels done: copy 2 GiB, sum 2 GiB (x3 each)")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1024)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--cells", type=int, default=1000, help="unit cells per side")
    ap.add_argument("--width", type=int, default=64, help="vectors per block")
    ap.add_argument("--variant", type=int, default=-1, help="kernel launch shape")
    ap.add_argument("--tile-rows", type=int, default=256,
                    help="rows per tile of the two-step kernel (0: one-step kernel only)")
    ap.add_argument("--tile-variant", type=int, default=-1,
                    help="two-step launch shape (default: the fastest that fits)")
    ap.add_argument("--ens-variant", type=int, default=-1,
                    help="kernel launch shape of the ensemble (default: autotuned)")
    ap.add_argument("--ens-shape", type=int, default=-1,
                    help="lattice-stencil launch shape of the ensemble (default: autotuned)")
    ap.add_argument("--ens-format", choices=["both", "stencil", "csr"], default="both",
                    help="matrix formats of the ensemble measurement")
    ap.add_argument("--ens-eager", action="store_true",
                    help="time the ensemble with eager launches only (no HIP graph)")
    ap.add_argument("--cpu-steps", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the batched disorder-ensemble measurement")
    ap.add_argument("--no-correlator", action="store_true",
                    help="skip the Kubo-Bastin Gamma_nm measurement")
    ap.add_argument("--only", choices=["correlator", "correlator-full", "ensemble", "surrogate",
                                       "surrogate-literal", "surrogate-train"],
                    help="run only one of the secondary measurements (for profiling)")
    ap.add_argument("--no-density", action="store_true",
                    help="skip the site-resolved density step measurement")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the moment-surrogate sweep measurement")
    ap.add_argument("--sur-epochs", type=int, default=9600)
    ap.add_argument("--sur-tol", type=float, default=1e-3)
    ap.add_argument("--sur-points", type=int, default=4096)
    ap.add_argument("--calibrate", action="store_true",
                    help="only run a 2 GiB copy and a 2 GiB read (known traffic) for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.calibrate:
        calibrate()
        return
    ens_formats = ("stencil", "csr") if args.ens_format == "both" else (args.ens_format,)
    if args.only:
        fn = {"correlator": lambda: correlator_bench(512, 1024, 32, 2, log, full=False),
              "correlator-full": lambda: correlator_bench(512, 1024, 32, 2, log),
              "ensemble": lambda: ensemble_bench(256, 512, 16, log, args.ens_variant,
                                                 not args.ens_eager, ens_formats,
                                                 args.ens_shape),
              "surrogate": lambda: surrogate_bench(args.sur_epochs, args.sur_tol,
                                                   args.sur_points, 256, 16, log),
              "surrogate-literal": lambda: surrogate_bench(args.sur_epochs, args.sur_tol,
                                                           args.sur_points, 256, 16, log,
                                                           design="literal"),
              "surrogate-train": lambda: surrogate_train_profile(22, log)}[args.only]
        print(json.dumps({args.only: fn()}), flush=True)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the KPM engine in this repository runs "
              "on a single MI355X (multi-GPU sharding is a design note only); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from kwant_amd import _engine as eng
    from kwant_amd.lattices import honeycomb, honeycomb_bounds

    t0 = time.time()
    h, pos = honeycomb(args.cells, disorder=0.5, seed=2)
    n, nnz, width = h.shape[0], h.nnz, args.width
    lmin, lmax = honeycomb_bounds(1.0, 0.5)
    eps = 0.05
    a = abs(lmax - lmin) / (2 - eps)
    b = (lmax + lmin) / 2
    log("graphene %dx%d cells: N=%d nnz=%d (%.1fs)" % (args.cells, args.cells, n, nnz,
                                                      time.time() - t0))
    rng = np.random.RandomState(3)
    rng.random_sample(n)  # the v0 draw of kwant/kpm.py:203
    block = torch.empty((n, width), dtype=torch.complex128)
    for c in range(width):
        block[:, c] = torch.from_numpy(np.exp(2j * np.pi * rng.random_sample(n)))
    cpu_vec = block[:, 0].numpy().copy()
    log("vectors ready (%.1fs)" % (time.time() - t0))

    # CPU baselines first, before the GPU is initialised (the pool forks)
    cpu = cpu_all = None
    if not args.no_cpu:
        log("CPU baseline: oracle loop, 1 vector x %d steps, complex128, 1 thread"
            % args.cpu_steps)
        t_vs = cpu_baseline(h, a, b, cpu_vec, args.cpu_steps)
        cpu = {"value": 1.0 / (t_vs * width), "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": "1 of the 64 vectors x %d steps of the same recursion "
                         "(oracle restatement of kwant's per-vector loop: scipy csr @ v, "
                         "numpy temporaries, np.vdot; complex128), %.1f ms per "
                         "vector-step, scaled x%d vectors per block step"
                         % (args.cpu_steps, 1e3 * t_vs, width)}
        log("CPU: %.1f ms per vector-step -> %.4f block steps/s" % (1e3 * t_vs, cpu["value"]))
        cpu_all = cpu_baseline_pool(h, a, b, args.cpu_steps, width)
        log("CPU, %d processes: %.4f block steps/s (%s)"
            % (cpu_all["cores"], cpu_all["value"], cpu_all["cpu_model"]))
    small = small_square_bench(log)

    dev = eng.torch_device()
    block = block.to(dev)

    m64 = eng.DeviceCSR(h, "complex64")
    if args.variant < 0:
        from kwant_amd import autotune
        args.variant = autotune.pick_variant(m64, (a, b), width, eng.DOTS_DOUBLING, force=True)
        log("autotuned launch shape: variant %d (ms/step per variant: %s)" % (
            args.variant, ", ".join("%.3f" % t for t in
                                    autotune._CACHE[autotune.key_of(m64, width,
                                                                    eng.DOTS_DOUBLING)][
                                        "ms_per_step"])))
    m_steps = args.steps
    check = {}
    if not args.no_check:
        # complex64 vs complex128 on the same vectors, one full expansion
        full = 1024
        r64 = eng.BlockRecursion(m64, (a, b), block.to(torch.complex64), width,
                                 eng.DOTS_DOUBLING, variant=args.variant)
        mom64 = moments_from_dots(r64.step(full))
        del r64
        from kwant_amd import autotune
        m128 = eng.DeviceCSR(h, "complex128")
        v128 = autotune.pick_variant(m128, (a, b), width, eng.DOTS_DOUBLING, force=True)
        r128 = eng.BlockRecursion(m128, (a, b), block, width, eng.DOTS_DOUBLING, variant=v128)
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        dots128 = r128.step(full)
        t128 = time.perf_counter() - s0
        mom128 = moments_from_dots(dots128)
        del r128, m128
        err = float(np.max(np.abs(mom64 - mom128)) / n)
        # DOS L1 of the complex64 recursion against the complex128 one (which
        # matches the FP64 oracle to 1e-11 in the tests): mean over vectors,
        # Jackson damped, relative to int rho
        from kwant_amd.kpm import jackson_kernel
        from kwant_amd.surrogate import dos_l1
        d64 = jackson_kernel(np.real(mom64).mean(0)) / n
        d128 = jackson_kernel(np.real(mom128).mean(0)) / n
        l1 = float(dos_l1(d64[None], d128[None])[0])
        check = {"moments_c64_vs_c128_max_err_over_mu0": err,
                 "dos_l1_c64_vs_c128": l1,
                 "num_moments": 2 * full,
                 "c128_steps_per_s": full / t128, "c128_variant": v128}
        log("c64 vs c128 moments: max |d mu| / mu0 = %.3e, DOS L1 %.1e; c128 %.0f steps/s "
            "(variant %d)" % (err, l1, full / t128, v128))
        if args.tile_rows > 0:
            c128_two = two_step_bench(h, pos, (a, b), block, width, args.warmup, 256,
                                      args.tile_rows, -1, -1, log, precision="complex128")
            torch.cuda.empty_cache()
            if c128_two is not None:
                check["c128_two_step"] = {k: c128_two[k] for k in (
                    "ms_per_step", "variant", "group_width", "waves", "kernel_frac_of_hbm_peak",
                    "dots_bitwise_vs_one_step")}
                check["c128_two_step"]["steps_per_s"] = 1000.0 / c128_two["ms_per_step"]
                check["c128_two_step"]["frac_of_hbm_peak"] = (
                    bytes_per_step(n, nnz, width, 16, 16) / (c128_two["ms_per_step"] * 1e-3)
                    / 1e9 / HBM_PEAK_GBS)

    block64 = block.to(torch.complex64)
    del block
    rec = eng.BlockRecursion(m64, (a, b), block64, width, eng.DOTS_DOUBLING,
                             variant=args.variant)
    plan = rec.plan
    dots = torch.empty((args.warmup + m_steps, 1, width, 4), dtype=torch.float64, device=dev)
    stride = n * width
    plan.steps(m64, rec.ring, stride, 2, args.warmup, eng.DOTS_DOUBLING, dots=dots, out_base=0)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    plan.steps(m64, rec.ring, stride, 2, m_steps, eng.DOTS_DOUBLING, dots=dots, out_base=0)
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end)
    assert plan.get_step() == args.warmup + m_steps
    ms_step = ms / m_steps
    steps_s = 1000.0 / ms_step
    bstep = bytes_per_step(n, nnz, width, 8, 8)
    achieved = bstep / (ms_step * 1e-3) / 1e9
    lpr, vec, waves, rounds, unroll, chunk, spare = eng.variant_info(width, eng.KPM_C64,
                                                                     args.variant)
    log("%.4f ms/step, %.1f steps/s, %.0f GB/s algorithmic (%.1f%% of 8 TB/s)"
        % (ms_step, steps_s, achieved, 100 * achieved / HBM_PEAK_GBS))
    one_step = {"ms_per_step": ms_step, "steps_per_s": steps_s, "achieved_GBs": achieved,
                "frac": achieved / HBM_PEAK_GBS, "variant": args.variant}

    del rec, plan, dots
    torch.cuda.empty_cache()
    two_step = None
    if args.tile_rows > 0:
        two_step = two_step_bench(h, pos, (a, b), block64, width, args.warmup, m_steps,
                                  args.tile_rows, args.tile_variant, args.variant, log)
        torch.cuda.empty_cache()
    headline_two = two_step is not None and two_step["ms_per_step"] < ms_step
    if headline_two:
        ms_step = two_step["ms_per_step"]
        steps_s = 1000.0 / ms_step
        achieved = bstep / (ms_step * 1e-3) / 1e9
        log("headline: two-step kernel, %.4f ms/step, %.1f steps/s (%.1f%% of 8 TB/s)"
            % (ms_step, steps_s, 100 * achieved / HBM_PEAK_GBS))
    site_density = None
    if not args.no_density:
        site_density = density_bench(m64, h, (a, b), block64, width, -1,
                                     one_step["ms_per_step"], 256, log)
    del m64, block64
    torch.cuda.empty_cache()
    correlator = None
    if not args.no_correlator:
        correlator = correlator_bench(512, 1024, 32, 2, log)
    ensemble = None
    if not args.no_ensemble:
        ensemble = ensemble_bench(256, 512, 16, log, args.ens_variant, not args.ens_eager,
                                  ens_formats, args.ens_shape)
    transform = assembly = None
    if not args.no_density:
        transform = transform_bench(1024, 65536, log)
        assembly = assembly_bench(1000, 60, log)
    surrogate = None
    if not args.no_surrogate:
        surrogate = surrogate_bench(args.sur_epochs, args.sur_tol, args.sur_points, 256, 16, log)

    traffic = read_traffic("tiled %d" % two_step["variant"] if headline_two
                           else "variant %d" % args.variant)
    stream = stream_probe(log)
    if headline_two:
        kernel = {"two_step": True, "group_width": two_step["group_width"],
                  "waves": two_step["waves"], "tile_rows": args.tile_rows,
                  "tiled_variant": two_step["variant"], "tiles": two_step["tiles"]}
    else:
        kernel = {"two_step": False, "lanes_per_row": lpr, "complex_per_lane": vec,
                  "waves": waves, "grid_rounds": rounds, "row_slots": unroll,
                  "interleave_chunk": chunk, "spare_per_cu": spare, "variant": args.variant}
    line = {
        "metric": "KPM Chebyshev steps/s (% of HBM peak); surrogate speedup at 1e-3 DOS error",
        "value": steps_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": m_steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64",
        "data": "synthetic",
        "config": {"workload": "graphene honeycomb %dx%d cells (N=%d, nnz=%d), Anderson "
                               "U[-0.25,0.25], R=%d random-phase vectors in one block, "
                               "2048 moments per expansion, complex64 recursion + FP64 dots"
                               % (args.cells, args.cells, n, nnz, width),
                   "N": n, "nnz": nnz, "block_width": width, "kernel": kernel},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_step": bstep,
                     "steps_per_launch": 2 if headline_two else 1,
                     "measured_stream_GBs": stream,
                     "frac_of_measured_add": achieved / stream["add"]},
        "one_step": one_step,
        "two_step": two_step,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "small_square": small,
        "correlator": correlator,
        "ensemble": ensemble,
        "site_density": site_density,
        "transform": transform,
        "assembly": assembly,
        "surrogate": surrogate,
    }
    line.update(check)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
