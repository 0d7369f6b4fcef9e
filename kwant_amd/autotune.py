"""Launch-shape autotuner for the Chebyshev recursion kernel.

The library exposes several launch shapes per (block width, value type):
lanes per matrix row, complex values per lane, residency rounds and row
slots per wavefront iteration (``kpm_variant_info``).  Which is fastest
depends on the matrix (rows, nonzeros per row, neighbour distances) and the
width, so for large problems the engine times each shape for a few steps on
the actual matrix and keeps the fastest, per (rows, nnz, width, value type,
dots mode).  Results are identical bitwise whatever shape runs (the per-row
arithmetic is the same and the FP64 dot sums are order-independent), so
tuning never changes the moments.

Results are cached in memory and, when ``KWANT_AMD_TUNE_CACHE`` names a file,
in that JSON file across processes.
"""
import json
import os

from . import _engine as eng

_CACHE = {}
MIN_ELEMENTS = 1 << 20  # smaller problems use the default shape
TIMED_STEPS = 6
FINALISTS = 3


def _cache_file():
    return os.environ.get("KWANT_AMD_TUNE_CACHE")


def _load():
    path = _cache_file()
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                for k, v in json.load(f).items():
                    _CACHE.setdefault(k, v)
        except (OSError, ValueError):
            pass


def _store():
    path = _cache_file()
    if not path:
        return
    try:
        with open(path, "w") as f:
            json.dump(_CACHE, f)
    except OSError:
        pass


def key_of(dmat, width, mode):
    key = "%d:%d:%d:%d:%d" % (dmat.shape[0], dmat.nnz, width, dmat.value_type, mode)
    fmt = getattr(dmat, "format", "csr")
    return key if fmt == "csr" else key + ":" + fmt


def candidates(dmat, width):
    """The launch shapes that can run this matrix (block-CSR: 4-wave shapes
    with at most 16 lanes per row)."""
    n_var = eng.variant_count(width, dmat.value_type)
    ok = getattr(dmat, "variant_ok", None)
    return [v for v in range(n_var) if ok is None or ok(width, dmat.value_type, v)]


# Modes timed as another: the site-resolved records cost the projection
# vector's read like DOTS_PROJECT (their records are timed with it).
_TIMED_AS = {eng.DOTS_LOCAL: eng.DOTS_PROJECT, eng.DOTS_LOCAL_COLS: eng.DOTS_PROJECT}


def time_variant(dmat, ab, width, mode, variant, block, steps=TIMED_STEPS):
    """Milliseconds per step of one launch shape on ``block`` (N x width).
    The projection modes read ``block`` itself as their vector w."""
    import torch
    mode = _TIMED_AS.get(mode, mode)
    w = block if mode == eng.DOTS_PROJECT else None
    rec = eng.BlockRecursion(dmat, ab, block, width, mode, w=w, variant=variant)
    dots = None
    if mode != eng.DOTS_NONE:
        dots = torch.empty((steps + 2, 1, width, 4), dtype=torch.float64, device=block.device)
    stride = dmat.shape[0] * width
    rec.plan.steps(dmat, rec.ring, stride, 2, 2, mode, dots=dots, out_base=0, w=w)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    rec.plan.steps(dmat, rec.ring, stride, 2, steps, mode, dots=dots, out_base=0, w=w)
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


def pick_variant(dmat, ab, width, mode=eng.DOTS_DOUBLING, force=False):
    """The fastest launch shape for this matrix and width (below the size
    threshold, unless ``force``, the first shape that can run it)."""
    import torch
    cands = candidates(dmat, width)
    if not cands:
        raise ValueError("no launch shape runs this matrix at width %d" % width)
    if len(cands) == 1 or (not force and dmat.shape[0] * width < MIN_ELEMENTS):
        return cands[0]
    if not _CACHE:
        _load()
    key = key_of(dmat, width, mode)
    if key in _CACHE:
        return int(_CACHE[key]["variant"])
    cdt, _ = eng.complex_dtype(dmat.precision)
    dev = eng.torch_device()
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    phase = torch.rand((dmat.shape[0], width), generator=gen, device=dev, dtype=torch.float64)
    block = torch.exp(2j * 3.141592653589793 * phase).to(cdt)
    times = {v: time_variant(dmat, ab, width, mode, v, block) for v in cands}
    # the fastest few are timed again, three times longer, so that run-to-run
    # noise does not decide between near-equal shapes
    short = sorted(cands, key=lambda v: times[v])[:FINALISTS]
    final = {v: time_variant(dmat, ab, width, mode, v, block, 3 * TIMED_STEPS) for v in short}
    best = min(final, key=lambda v: final[v])
    _CACHE[key] = {"variant": best, "ms_per_step": [times.get(v) for v in
                                                    range(max(cands) + 1)]}
    _store()
    return best


def _tiled_key(dmat, width, tile_begin):
    import zlib
    import numpy as np
    tb = np.ascontiguousarray(tile_begin, dtype=np.int64)
    return "%s:tiled:%08x" % (key_of(dmat, width, eng.DOTS_DOUBLING), zlib.crc32(tb.tobytes()))


def time_tiled(dmat, ab, width, tiles, block, steps=TIMED_STEPS):
    """Milliseconds per Chebyshev step of two-step launches (``tiles``, a
    TiledPlan) on ``block``."""
    import torch
    rec = eng.BlockRecursion(dmat, ab, block, width, eng.DOTS_DOUBLING, tiles=tiles)
    pairs = max(1, steps // 2)
    dots = torch.empty((3 + 2 * pairs, 1, width, 4), dtype=torch.float64, device=block.device)
    rec.step(3)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    tiles.steps(rec.plan, dmat, rec.ring, rec.slot_stride, rec.slots, pairs, eng.DOTS_DOUBLING,
                dots=dots, out_base=0)
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / (2 * pairs)


def pick_tiled(dmat, ab, width, tile_begin, force=False):
    """The two-step launch (a TiledPlan of ``dmat`` for ``tile_begin``) to run
    blocks of ``width`` vectors with, or None.  Shapes whose LDS cannot hold a
    tile and its halo (or that take smaller tiles) are skipped.  For large
    problems (or ``force``) the feasible shapes are timed, and None is returned
    when the one-step kernel on the same rows is faster; small problems take
    the first feasible shape.  Results are bitwise the same either way."""
    import torch
    plans = {}
    for v in range(eng.tiled_variant_count(width, dmat.value_type)):
        try:
            plans[v] = eng.TiledPlan(dmat, width, tile_begin, variant=v)
        except (RuntimeError, ValueError):
            continue
    if not plans:
        return None
    first = min(plans)
    if not force and dmat.shape[0] * width < MIN_ELEMENTS:
        return plans[first]
    if not _CACHE:
        _load()
    key = _tiled_key(dmat, width, tile_begin)
    if key in _CACHE:
        v = int(_CACHE[key]["variant"])
        return plans.get(v) if v >= 0 else None
    cdt, _ = eng.complex_dtype(dmat.precision)
    dev = eng.torch_device()
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    phase = torch.rand((dmat.shape[0], width), generator=gen, device=dev, dtype=torch.float64)
    block = torch.exp(2j * 3.141592653589793 * phase).to(cdt)
    times = {v: time_tiled(dmat, ab, width, p, block, 3 * TIMED_STEPS) for v, p in plans.items()}
    best = min(times, key=lambda v: times[v])
    one = pick_variant(dmat, ab, width, eng.DOTS_DOUBLING, force=True)
    one_ms = time_variant(dmat, ab, width, eng.DOTS_DOUBLING, one, block, 3 * TIMED_STEPS)
    chosen = best if times[best] < one_ms else -1
    _CACHE[key] = {"variant": chosen, "ms_per_step": {str(v): t for v, t in times.items()},
                   "one_step_ms": one_ms}
    _store()
    return plans[best] if chosen >= 0 else None
