#!/usr/bin/env python
"""Benchmark: KPM Chebyshev steps/s on one MI355X.

Workload (the headline configuration): graphene honeycomb parallelogram of
1000 x 1000 unit cells (N = 2,000,000 sites, nnz = 7,994,002), NN hopping
t = 1, Anderson disorder U[-0.25, 0.25], 64 random-phase vectors in one
N x 64 block, complex64 recursion with FP64 dot accumulation; a full
expansion is 2048 moments = 1024 Chebyshev steps.

A "step" is one launch of the fused recursion kernel over the whole block:
a_{n+1} = 2 (H a_n - b a_n)/a - a_{n-1} plus the two FP64 dot products per
column (two moments per vector).  The timed region holds only kernel
launches on the current HIP stream, bracketed by events; inputs are resident
in HBM before it starts.

Algorithmic bytes per step (the roofline numerator):
    B = 3 R N s_v + nnz (s_m + 4) + 4 (N + 1)
with s_v = s_m = 8 bytes (complex64): read a_n once, read a_{n-1}, write
a_{n+1}, read the CSR once.

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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "cheb_step_c64_graphene2M_traffic.txt")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def bytes_per_step(n, nnz, width, s_v, s_m):
    return 3 * width * n * s_v + nnz * (s_m + 4) + 4 * (n + 1)


def read_traffic():
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), if committed."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


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
    ap.add_argument("--cpu-steps", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run a 2 GiB copy and a 2 GiB read (known traffic) for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.calibrate:
        calibrate()
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
    h, _ = honeycomb(args.cells, disorder=0.5, seed=2)
    n, nnz, width = h.shape[0], h.nnz, args.width
    lmin, lmax = honeycomb_bounds(1.0, 0.5)
    eps = 0.05
    a = abs(lmax - lmin) / (2 - eps)
    b = (lmax + lmin) / 2
    log("graphene %dx%d cells: N=%d nnz=%d (%.1fs)" % (args.cells, args.cells, n, nnz,
                                                      time.time() - t0))
    rng = np.random.RandomState(3)
    rng.random_sample(n)  # the v0 draw of kwant/kpm.py:203
    dev = eng.torch_device()
    block = torch.empty((n, width), dtype=torch.complex128)
    for c in range(width):
        block[:, c] = torch.from_numpy(np.exp(2j * np.pi * rng.random_sample(n)))
    block = block.to(dev)
    cpu_vec = block[:, 0].cpu().numpy()
    log("vectors ready (%.1fs)" % (time.time() - t0))

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
        m128 = eng.DeviceCSR(h, "complex128")
        r128 = eng.BlockRecursion(m128, (a, b), block, width, eng.DOTS_DOUBLING)
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        dots128 = r128.step(full)
        t128 = time.perf_counter() - s0
        mom128 = moments_from_dots(dots128)
        del r128, m128
        err = float(np.max(np.abs(mom64 - mom128)) / n)
        check = {"moments_c64_vs_c128_max_err_over_mu0": err,
                 "num_moments": 2 * full,
                 "c128_steps_per_s": full / t128}
        log("c64 vs c128 moments: max |d mu| / mu0 = %.3e; c128 %.0f steps/s"
            % (err, full / t128))

    rec = eng.BlockRecursion(m64, (a, b), block.to(torch.complex64), width,
                             eng.DOTS_DOUBLING, variant=args.variant)
    del block
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
    lpr, vec, waves, rounds, unroll = eng.variant_info(width, eng.KPM_C64,
                                                       args.variant)
    log("%.4f ms/step, %.1f steps/s, %.0f GB/s algorithmic (%.1f%% of 8 TB/s)"
        % (ms_step, steps_s, achieved, 100 * achieved / HBM_PEAK_GBS))

    cpu = None
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
        log("CPU: %.1f ms per vector-step -> %.4f block steps/s"
            % (1e3 * t_vs, cpu["value"]))

    traffic = read_traffic()
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
                   "N": n, "nnz": nnz, "block_width": width,
                   "kernel": {"lanes_per_row": lpr, "complex_per_lane": vec, "waves": waves,
                              "grid_rounds": rounds, "row_slots": unroll}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_step": bstep},
        "cpu_baseline": cpu,
    }
    line.update(check)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
