"""CPU oracle of the exact moment shift (Chebyshev moments of H + mu from the
moments of H in another window).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU baseline leg may import this module, and only as the
checker; the product package ``kwant_amd`` never imports it.

No counterpart in the reference: kwant/kpm.py recomputes the recursion for
every Hamiltonian.  The identity restated here is T_n(alpha t + beta) =
sum_k C_nk T_k(t), with C_n built by the three-term recurrence of T_n and
multiplication by t in the Chebyshev basis (t T_k = (T_{k+1} + T_{|k-1|})/2).

Pinning (tests/test_shift_oracle.py):
  * C_nk against numpy.polynomial's Chebyshev composition for n <= 24;
  * shifted moments against direct evaluation sum_j w_j T_n((E_j + mu)/a)
    on explicit spectra (closed form of the moments of a point measure).
"""
import numpy as np


def shift_matrix(num_moments, alpha, beta):
    """C (M, M) with T_n(alpha t + beta) = sum_k C[n, k] T_k(t)."""
    m = num_moments
    c = np.zeros((m, m + 1))
    c[0, 0] = 1.0
    if m > 1:
        c[1, 0], c[1, 1] = beta, alpha
    for n in range(1, m - 1):
        cur = c[n]
        tc = np.zeros(m + 1)
        tc[0] = cur[1] / 2
        tc[1] = cur[0] + cur[2] / 2
        tc[2:m] = (cur[1:m - 1] + cur[3:m + 1]) / 2
        c[n + 1] = 2 * (alpha * tc + beta * cur) - c[n - 1]
    return c[:, :m]


def moment_shift(nu, alpha, beta, in_scale=None, out_scale=None):
    """mu[p, n] = out_scale[n] sum_k C_nk(alpha, beta[p]) in_scale[k] nu[p, k]."""
    nu = np.atleast_2d(np.asarray(nu, dtype=np.float64))
    p, m = nu.shape
    beta = np.broadcast_to(np.asarray(beta, dtype=np.float64), (p,))
    x = nu * (1.0 if in_scale is None else np.asarray(in_scale))
    out = np.stack([shift_matrix(m, alpha, beta[i]) @ x[i] for i in range(p)])
    return out * (1.0 if out_scale is None else np.asarray(out_scale))
