"""Dense brute-force moments -- the independent check that pins the oracle.

TEST INFRASTRUCTURE ONLY (same import rule as oracle/kpm_oracle.py).

With H~ = (H - b)/a = V diag(lam) V^H (numpy eigh), the Chebyshev moments
of a start vector r are

    mu_n = <r| T_n(H~) |r> = sum_k |(V^H r)_k|^2 cos(n arccos lam_k),

the operator moments <r| O T_n(H~) |r>, and the Kubo-Bastin moment matrix

    mu_mn = <T_m(H~) O1 r | O2 T_n(H~) r>

(kwant/kpm.py:530-540 definitions) follow the same way.  No recursion and no
doubling identity is involved, so a dropped term, a wrong index or a sign in
the recursion shows up as a mismatch.
"""
import numpy as np


def _eig(h, a, b):
    h = np.asarray(h.todense() if hasattr(h, 'todense') else h, dtype=complex)
    lam, vec = np.linalg.eigh((h - b * np.eye(len(h))) / a)
    if np.any(np.abs(lam) >= 1):
        raise ValueError("rescaled spectrum leaves (-1, 1)")
    return lam, vec


def cheb_matrix(lam, n_moments):
    """T_n(lam_k) = cos(n arccos lam_k), shape (n_moments, len(lam))."""
    return np.cos(np.arange(n_moments)[:, None] * np.arccos(lam)[None, :])


def moments(h, a, b, r, n_moments):
    """mu_n = <r|T_n(H~)|r> for n < n_moments."""
    lam, vec = _eig(h, a, b)
    c = vec.conj().T @ np.asarray(r, dtype=complex)
    return cheb_matrix(lam, n_moments) @ (np.abs(c) ** 2)


def operator_moments(h, a, b, r, op, n_moments):
    """<r| O T_n(H~) |r> for a matrix O."""
    lam, vec = _eig(h, a, b)
    r = np.asarray(r, dtype=complex)
    op = np.asarray(op.todense() if hasattr(op, 'todense') else op, dtype=complex)
    c = vec.conj().T @ r
    left = (r.conj() @ op) @ vec          # <r| O V
    return cheb_matrix(lam, n_moments) @ (left * c)


def chebyshev_vectors(h, a, b, r, n_moments):
    """T_n(H~) r for n < n_moments, shape (n_moments, N)."""
    lam, vec = _eig(h, a, b)
    c = vec.conj().T @ np.asarray(r, dtype=complex)
    t = cheb_matrix(lam, n_moments)
    return (vec @ (t * c[None, :]).T).T


def correlator_moments(h, a, b, r, op1, op2, n_moments, a_left=None, b_left=None):
    """mu_mn = <T_m(H~_L) O1 r | O2 T_n(H~) r> (one random vector)."""
    dense = lambda m: np.asarray(m.todense() if hasattr(m, 'todense') else m, dtype=complex)
    r = np.asarray(r, dtype=complex)
    o1, o2 = dense(op1), dense(op2)
    psi = (o2 @ chebyshev_vectors(h, a, b, r, n_moments).T).T
    al = a if a_left is None else a_left
    bl = b if b_left is None else b_left
    omega = chebyshev_vectors(h, al, bl, o1 @ r, n_moments)
    return omega.conj() @ psi.T


def gaussian_unitary(n, rng):
    """GUE sample with kwant.rmt.gaussian(n)'s recipe for class A
    (kwant/rmt.py:62-161): h = X + X^H with X = randn + i randn, times
    1/sqrt(2)."""
    rng = np.random.RandomState(rng) if not hasattr(rng, 'randn') else rng
    h = rng.randn(n, n) + 1j * rng.randn(n, n)
    h += h.T.conj()
    return h / np.sqrt(2)
