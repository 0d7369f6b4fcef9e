"""On-disk moments: save a SpectralDensity and resume it in another process.

kwant keeps an expansion's resume state in memory only (``_last_two_alphas``,
``_moments_list`` and the vector factory's saved vectors,
kwant/kpm.py:318-376, 486-490).  Here it is written to one ``.npz`` file
(numpy arrays and a JSON header; nothing is pickled):

* the rescaling (a, b), eps, the moment count, mean, precision, launch shape
  and the device row order;
* per block of vectors: its moments, the recursion's last two vectors
  (alpha_{n-1}, alpha_n) exactly as the GPU holds them (complex64 or
  complex128), and the step counter (alpha_0 is rebuilt from the saved
  vectors, the way the block was first built);
* the vectors drawn so far, and the state of the default random-vector
  generator (numpy's MT19937), so that ``add_vectors`` after a load draws the
  same vectors as an uninterrupted run.

``load`` takes the Hamiltonian (and operator) again -- they are inputs, not
state -- and checks that its size matches.  Resumed expansions continue
bitwise: ``add_moments`` after save/load equals ``add_moments`` without it.
"""
import json

import numpy as np

FORMAT = "kwant_amd.kpm.SpectralDensity"
ENSEMBLE_FORMAT = "kwant_amd.ensemble.CubicEnsemble"
CORRELATOR_FORMAT = "kwant_amd.kpm.Correlator"
VERSION = 1


def _meta_array(meta):
    return np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)


def save_arrays(path, fmt, meta, arrays):
    """Numeric arrays and a JSON header in one .npz (no pickles)."""
    meta = dict(meta, format=fmt, version=VERSION)
    with open(path, "wb") as fh:
        np.savez(fh, meta=_meta_array(meta), **arrays)


def load_arrays(path, fmt):
    """(header, arrays) of a file written by save_arrays with format ``fmt``."""
    meta = read_meta(path, fmt)
    with np.load(path, allow_pickle=False) as z:
        return meta, {k: z[k] for k in z.files if k != "meta"}


def save(sd, path):
    """Write ``sd`` (a kwant_amd.kpm.SpectralDensity) to ``path`` (.npz)."""
    from . import _engine as eng
    vf = sd._vector_factory
    arrays = {}
    blocks = []
    for i, blk in enumerate(sd._blocks):
        entry = {"count": int(blk.count), "resumable": blk.rec is not None}
        mom = np.asarray(blk.moments if not isinstance(blk.moments, list)
                         else [np.asarray(m) for m in blk.moments])
        arrays["block%d_moments" % i] = mom
        if blk.rec is not None:
            rec = blk.rec
            entry.update(width=int(rec.width), mode=int(rec.mode), step=int(rec.n),
                         variant=int(rec.variant))
            arrays["block%d_ring" % i] = rec.ring.cpu().numpy()
        blocks.append(entry)
    saved = [v for v in vf.saved_vectors if v is not None]
    if saved:
        arrays["saved_vectors"] = np.stack([np.asarray(v, dtype=np.complex128) for v in saved])
    rng_state = getattr(sd, "_rng", None)
    default_vectors = getattr(sd, "_default_vectors", False)
    if default_vectors and rng_state is not None:
        name, key, pos, has_gauss, cached = rng_state.get_state()
        arrays["rng_key"] = np.asarray(key, dtype=np.uint32)
        rng = {"name": name, "pos": int(pos), "has_gauss": int(has_gauss),
               "cached_gaussian": float(cached)}
    else:
        rng = None
    if sd._dmat.order is not None:
        arrays["order"] = np.asarray(sd._dmat.order, dtype=np.int64)
    tiling = getattr(sd, "_tiling", None)
    if tiling is not None:
        arrays["tile_begin"] = np.asarray(tiling.tile_begin, dtype=np.int64)
    from . import kpm
    kernel = {kpm.jackson_kernel: "jackson", kpm.lorentz_kernel: "lorentz"}.get(sd.kernel,
                                                                              "custom")
    meta = {
        "kernel": kernel,
        "format": FORMAT, "version": VERSION,
        "n_rows": int(sd._dmat.shape[0]), "a": float(sd._a), "b": float(sd._b),
        "eps": float(sd.eps), "num_moments": int(sd.num_moments), "mean": bool(sd.mean),
        "precision": sd._precision, "max_block_width": int(sd._max_width),
        "variant": int(sd._variant), "accumulate": bool(vf.accumulate),
        "factory_num_vectors": int(vf.num_vectors),
        "factory_max_vectors": None if vf._max_vectors == np.inf else int(vf._max_vectors),
        "num_saved_vectors": len(saved), "default_vectors": bool(default_vectors),
        "rng": rng, "blocks": blocks, "engine_abi": int(eng.load_library().kpm_abi_version()),
    }
    arrays["meta"] = _meta_array(meta)
    with open(path, "wb") as fh:
        np.savez(fh, **arrays)


def read_meta(path, fmt=FORMAT):
    """The JSON header of a saved expansion or ensemble (no GPU needed)."""
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(bytes(z["meta"]).decode())
    if meta.get("format") != fmt:
        raise ValueError("%s is not a saved %s" % (path, fmt))
    if meta.get("version") != VERSION:
        raise ValueError("unsupported format version %r" % meta.get("version"))
    return meta


def load(path, hamiltonian, params=None, operator=None, kernel=None, vector_factory=None):
    """A SpectralDensity restored from ``path``; ``hamiltonian`` (and
    ``operator``) must be the ones the expansion was made with.  With the
    default random vectors ``add_vectors`` continues the same random
    sequence; with a user ``vector_factory`` pass an iterator positioned
    after the vectors already used."""
    import torch
    from . import _engine as eng
    from . import kpm
    meta = read_meta(path)
    abi = int(eng.load_library().kpm_abi_version())
    if meta.get("engine_abi") != abi:
        raise ValueError("%s was saved by engine ABI %r; this library is ABI %d (the saved "
                         "recursion state is not portable across engine versions)"
                         % (path, meta.get("engine_abi"), abi))
    saved_kernel = meta.get("kernel", "jackson")
    if kernel is None:
        if saved_kernel == "custom":
            raise ValueError("%s was made with a custom kernel: pass it again as kernel="
                             % path)
        kernel = {"jackson": kpm.jackson_kernel, "lorentz": kpm.lorentz_kernel}[saved_kernel]
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    h = kpm._as_csr(hamiltonian, params)
    if h.shape[0] != meta["n_rows"]:
        raise ValueError("the Hamiltonian has %d rows, the saved expansion %d"
                         % (h.shape[0], meta["n_rows"]))
    sd = kpm.SpectralDensity.__new__(kpm.SpectralDensity)
    sd.eps = meta["eps"]
    sd._set_operator(operator, params, h)
    sd.mean = meta["mean"]
    sd._precision = meta["precision"]
    sd._max_width = meta["max_block_width"]
    sd._variant = meta["variant"]
    sd._a, sd._b = meta["a"], meta["b"]
    order = arrays.get("order")
    sd._tiling, sd._tiled = None, {}
    if "tile_begin" in arrays:
        from .ordering import Tiling
        sd._tiling = Tiling(order, arrays["tile_begin"])
    dmat = eng.DeviceCSR(h, sd._precision, order=order)
    sd.hamiltonian = kpm._RescaledOperator(dmat, sd._a, sd._b)
    sd._dmat = dmat
    sd.bounds = (sd._b - sd._a, sd._b + sd._a)
    sd.num_moments = meta["num_moments"]
    sd.kernel = kernel if kernel is not None else kpm.jackson_kernel

    # the vector factory: the vectors drawn so far, then the generator
    saved = list(arrays["saved_vectors"]) if "saved_vectors" in arrays else []
    sd._rng = None
    sd._default_vectors = False
    if vector_factory is None and meta["default_vectors"] and meta["rng"] is not None:
        rng = np.random.RandomState()
        r = meta["rng"]
        rng.set_state((r["name"], arrays["rng_key"], r["pos"], r["has_gauss"],
                       r["cached_gaussian"]))
        source = kpm.RandomVectors(h, rng=rng)
        sd._rng, sd._default_vectors = rng, True
    elif vector_factory is not None:
        source = vector_factory
    else:
        source = iter(())
    vf = kpm._VectorFactory.__new__(kpm._VectorFactory)
    vf._max_vectors = np.inf if meta["factory_max_vectors"] is None else \
        meta["factory_max_vectors"]
    vf._iterator = iter(source)
    vf.accumulate = meta["accumulate"]
    vf.saved_vectors = saved + [None] * (meta["factory_num_vectors"] - len(saved))
    vf.num_vectors = meta["factory_num_vectors"]
    vf._last_idx = -np.inf
    vf._last_vector = None
    sd._vector_factory = vf

    # the blocks, in vector order
    sd._blocks = []
    cdt, _ = eng.complex_dtype(sd._precision)
    dev = eng.torch_device()
    first = 0
    for i, entry in enumerate(meta["blocks"]):
        count = entry["count"]
        if entry["resumable"]:
            if first + count > len(saved):
                raise ValueError("the saved vectors do not cover block %d" % i)
            width = entry["width"]
            blk = kpm._DosBlock(dmat, (sd._a, sd._b), saved[first:first + count], width,
                                entry["mode"], variant=entry["variant"],
                                tiles=sd._tiles_for(width, entry["mode"]))
            sd._attach(blk, width)
            ring = torch.from_numpy(arrays["block%d_ring" % i]).to(dev)
            if ring.dtype != cdt or tuple(ring.shape) != tuple(blk.rec.ring.shape):
                raise ValueError("saved recursion state does not match the expansion")
            blk.rec.ring.copy_(ring)
            blk.rec.n = entry["step"]
            blk.rec.plan.set_step(entry["step"])
        else:
            blk = kpm._DosBlock.__new__(kpm._DosBlock)
            blk.count, blk.rec = count, None
        blk.moments = arrays["block%d_moments" % i]
        first += count
        sd._blocks.append(blk)
    sd._moments_list = [m for blk in sd._blocks for m in sd._entries(blk)]
    sd.densities, sd._gammas = kpm._calc_fft_moments(sd._moments())
    return sd
