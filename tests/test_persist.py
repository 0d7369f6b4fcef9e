"""The on-disk format's header checks (CPU only; the GPU round trips are in
test_gpu_persist.py)."""
import json

import numpy as np
import pytest

from kwant_amd import persist


def _write(path, meta):
    np.savez(path, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))


def test_read_meta_accepts_the_format(tmp_path):
    p = tmp_path / "ok.npz"
    _write(p, {"format": persist.FORMAT, "version": persist.VERSION, "n_rows": 5})
    assert persist.read_meta(p)["n_rows"] == 5


def test_read_meta_rejects_other_files(tmp_path):
    p = tmp_path / "other.npz"
    _write(p, {"format": "something else", "version": 1})
    with pytest.raises(ValueError):
        persist.read_meta(p)
    q = tmp_path / "future.npz"
    _write(q, {"format": persist.FORMAT, "version": persist.VERSION + 1})
    with pytest.raises(ValueError):
        persist.read_meta(q)


def test_files_hold_no_pickles(tmp_path):
    p = tmp_path / "plain.npz"
    _write(p, {"format": persist.FORMAT, "version": persist.VERSION})
    with np.load(p, allow_pickle=False) as z:
        assert all(z[k].dtype != object for k in z.files)


def test_load_rejects_another_engine_abi(tmp_path):
    """The saved recursion state is tied to the engine version that wrote it."""
    from kwant_amd import kpm
    p = tmp_path / "old.npz"
    _write(p, {"format": persist.FORMAT, "version": persist.VERSION, "engine_abi": 1,
               "n_rows": 4})
    with pytest.raises(ValueError, match="engine ABI"):
        kpm.SpectralDensity.load(p, np.eye(4))
