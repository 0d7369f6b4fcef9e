"""Failure detection on the host (no GPU): the checks that turn NaN/Inf
inputs and bounds that miss the spectrum into clear errors."""
import numpy as np
import pytest

from kwant_amd import kpm


def test_finite_vectors_pass_and_nan_inf_raise():
    good = np.exp(2j * np.pi * np.random.RandomState(0).random_sample((50, 3)))
    kpm._check_finite_vectors(good)
    for bad in (np.nan, np.inf, complex(0, np.inf)):
        v = good.copy()
        v[7, 2] = bad
        with pytest.raises(ValueError, match=r"vectors \[2\]"):
            kpm._check_finite_vectors(v)


def test_moment_growth_guard():
    """Moments of a spectrum inside the bounds stay below mu_0; moments of a
    spectrum outside grow like cosh(n arccosh x) and are caught."""
    n = np.arange(200)
    inside = np.cos(n * np.arccos(0.97))[None, :] * 10.0        # |x| < 1
    kpm._check_moment_growth(inside + 0j, (-1, 1))
    outside = np.cosh(n * np.arccosh(1.02))[None, :] * 10.0     # x = 1.02
    with pytest.raises(ValueError, match="do not contain the spectrum"):
        kpm._check_moment_growth(outside + 0j, (-1, 1))
    with pytest.raises(ValueError, match="not finite"):
        kpm._check_moment_growth(np.array([[1.0, np.nan]]), (-1, 1))
