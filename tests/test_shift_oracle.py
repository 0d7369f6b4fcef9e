"""Pins of the moment-shift oracle (oracle/shift_oracle.py)."""
import numpy as np
import pytest

from oracle import shift_oracle as so


@pytest.mark.parametrize("alpha,beta", [(0.857, 0.139), (0.5, -0.3), (1.0, 0.0), (0.2, 0.75)])
def test_shift_matrix_matches_numpy_chebyshev(alpha, beta):
    from numpy.polynomial import chebyshev as C
    m = 24
    got = so.shift_matrix(m, alpha, beta)
    for n in range(m):
        # T_n(alpha t + beta) as a Chebyshev series in t, by composition
        tn = C.Chebyshev.basis(n)
        inner = C.Chebyshev([beta, alpha])
        want = tn(inner).coef
        want = np.pad(want, (0, m - len(want)))
        assert np.max(np.abs(got[n] - want)) < 1e-12, n


def test_shift_matrix_identity_and_bound():
    c = so.shift_matrix(64, 1.0, 0.0)
    assert np.array_equal(c, np.eye(64))
    c = so.shift_matrix(512, 0.857, 0.139)
    assert np.max(np.abs(c)) <= 2.0 + 1e-9
    assert np.allclose(np.triu(c, 1), 0)


def test_shifted_moments_of_point_spectra():
    """Moments of a weighted point spectrum: shifting the window equals
    evaluating T_n at the shifted energies."""
    rng = np.random.RandomState(0)
    a_tight, a_wide, m = 6.15, 7.18, 200
    e = rng.uniform(-6.0, 6.0, 300)
    w = rng.exponential(1.0, 300)
    n = np.arange(m)
    nu = (w[None, :] * np.cos(n[:, None] * np.arccos(e[None, :] / a_tight))).sum(1)
    for mu in (-1.0, -0.37, 0.0, 0.81, 1.0):
        direct = (w[None, :] * np.cos(n[:, None] * np.arccos((e[None, :] + mu) / a_wide))).sum(1)
        got = so.moment_shift(nu, a_tight / a_wide, mu / a_wide)[0]
        assert np.max(np.abs(got - direct)) < 1e-9 * w.sum(), mu
