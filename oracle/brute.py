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


# --------------------------------------------------------------------------
# Exact Kubo formula for the Hall conductivity (independent of the KPM
# expansion, the integral factor and the prefactor of kwant/kpm.py:608-630)

def kubo_hall_exact(h, v_alpha, v_beta, mu, temperature):
    """sigma_ab * volume in units of e^2/h from the eigenstates of H.

    The Kubo-Bastin formula (Garcia, Covaci and Rappoport, PRL 114, 116602),
        sigma_ab = (i hbar e^2 / V) int dE f(E) Tr[v_a delta(E-H) v_b dG+/dE
                                                 - v_a dG-/dE v_b delta(E-H)],
    written in the eigenbasis H|k> = E_k|k> (dG+-/dE = -(E - H +- i0)^-2,
    hbar = 1, so e^2/hbar = 2 pi e^2/h), is
        sigma_ab V / (e^2/h)
            = -4 pi sum_{k != l} f(E_k) Im[(v_a)_kl (v_b)_lk] / (E_k - E_l)^2.
    Degenerate pairs (E_k = E_l) have f(E_k) = f(E_l) and cancel in pairs; they
    are left out.  At T = 0 f is the step E < mu.  For a Chern insulator on a
    torus, sigma_xy / (e^2/h) per unit cell tends to the Chern number (the TKNN
    form of the same sum), which pins the normalisation."""
    dense = lambda m: np.asarray(m.todense() if hasattr(m, 'todense') else m, dtype=complex)
    e, u = np.linalg.eigh(dense(h))
    va = u.conj().T @ dense(v_alpha) @ u
    vb = u.conj().T @ dense(v_beta) @ u
    if temperature == 0:
        f = (e < mu).astype(float)
    else:
        f = 1.0 / (1.0 + np.exp((e - mu) / temperature))
    gap = e[:, None] - e[None, :]
    gap[np.abs(gap) < 1e-9 * max(1.0, np.abs(e).max())] = np.inf
    prod = va * vb.T                        # (v_a)_kl (v_b)_lk
    return -4 * np.pi * np.sum(f[:, None] * prod.imag / gap ** 2)


def qwz_torus(L, m):
    """Qi-Wu-Zhang Chern insulator on an L x L torus, two orbitals per site:
    H(k) = sin kx s_x + sin ky s_y + (m + cos kx + cos ky) s_z, so on-site m s_z
    and hoppings (s_z - i s_x)/2 along x, (s_z - i s_y)/2 along y.  Chern
    number magnitude 1 for 0 < |m| < 2, spectrum in [-(2 + |m|), 2 + |m|].
    Returns dense (H, v_x, v_y) with the velocities i (x_col - x_row) H of
    kwant/kpm.py:1079-1081 taken with minimal-image displacements."""
    sx = np.array([[0, 1], [1, 0]], complex)
    sy = np.array([[0, -1j], [1j, 0]])
    sz = np.diag([1.0, -1.0]).astype(complex)
    hops = ((1, 0, (sz - 1j * sx) / 2), (0, 1, (sz - 1j * sy) / 2))
    n = 2 * L * L
    h = np.zeros((n, n), complex)
    dx = np.zeros((n, n))
    dy = np.zeros((n, n))

    def site(x, y):
        return 2 * ((x % L) * L + (y % L))

    for x in range(L):
        for y in range(L):
            i = site(x, y)
            h[i:i + 2, i:i + 2] += m * sz
            for sx_, sy_, t in hops:
                j = site(x + sx_, y + sy_)
                h[j:j + 2, i:i + 2] += t
                h[i:i + 2, j:j + 2] += t.conj().T
                # (x_col - x_row): -step for the entry (row j, col i)
                dx[j:j + 2, i:i + 2], dx[i:i + 2, j:j + 2] = -sx_, sx_
                dy[j:j + 2, i:i + 2], dy[i:i + 2, j:j + 2] = -sy_, sy_
    return h, 1j * dx * h, 1j * dy * h
