"""Spectral bounds of a Hermitian matrix by FP64 Lanczos on the GPU.

Replaces the two ARPACK runs of kwant/kpm.py:1169-1172
(``eigsh(k=1, which='LA'/'SA', tol=eps/2, v0=v0)``).  The Krylov space is
built from the same start vector v0 with the library's complex128 SpMM; the
three-term Lanczos update and its dot products are FP64 device operations.
Every ``check_every`` steps the extreme Ritz values of the tridiagonal matrix
(host, tiny) are tested against ARPACK's stopping rule |beta_k s_k| <=
tol |theta| for both ends.  Like ARPACK the result is an estimate that
converges from inside the spectrum; the eps/2 slack of the rescaling
(kwant/kpm.py:1164, 1174) absorbs the remaining error.
"""
import numpy as np
import scipy.linalg

from . import _engine as eng


def _ritz_extremes(alpha, beta):
    """Smallest and largest Ritz values and their residual estimates."""
    k = len(alpha)
    if k == 1:
        return alpha[0], alpha[0], abs(beta[0]), abs(beta[0])
    w, s = scipy.linalg.eigh_tridiagonal(alpha, beta[:k - 1])
    last = beta[k - 1]
    return w[0], w[-1], abs(last * s[-1, 0]), abs(last * s[-1, -1])


def lanczos_bounds(dmat, v0, tol=0.025, max_iter=None, check_every=10, min_iter=20):
    """(lmin, lmax) of the complex128 DeviceCSR ``dmat`` from start vector v0."""
    import torch
    if dmat.value_type != eng.KPM_C128:
        raise ValueError("Lanczos bounds need the complex128 matrix")
    dev = eng.torch_device()
    n = dmat.shape[0]
    if max_iter is None:
        max_iter = min(n, 1000)
    v = torch.from_numpy(np.ascontiguousarray(v0, dtype=np.complex128)).to(dev)
    v = v / torch.linalg.vector_norm(v)
    v_prev = torch.zeros_like(v)
    w = torch.empty_like(v)
    alpha, beta = [], []
    b_prev = 0.0
    lo = hi = None
    for it in range(max_iter):
        dmat.spmm(v, w, 1)
        a_j = torch.vdot(v, w).real
        w = w - a_j * v - b_prev * v_prev
        b_j = torch.linalg.vector_norm(w)
        alpha.append(a_j)
        beta.append(b_j)
        done = it + 1 == max_iter
        if (it + 1) % check_every == 0 or done:
            al = torch.stack(alpha).cpu().numpy()
            be = torch.stack(beta).cpu().numpy()
            lo, hi, r_lo, r_hi = _ritz_extremes(al, be)
            tiny = be[-1] <= 1e-12 * max(abs(lo), abs(hi), 1.0)
            if tiny or done or (it + 1 >= min_iter and r_lo <= tol * abs(lo)
                                and r_hi <= tol * abs(hi)):
                return float(lo), float(hi)
        b_prev = b_j
        v_prev, v = v, w / b_j
        w = torch.empty_like(v)
    return float(lo), float(hi)
