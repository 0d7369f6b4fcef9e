"""CPU oracle of the device-side cubic 4-orbital ensemble generator.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU baseline leg may import this module, and only as the
checker; the product package ``kwant_amd`` never imports it.

The reference builds Hamiltonians on the host through
``hamiltonian_submatrix`` (kwant/kpm.py:177-184, kwant/_system.pyx:233-338);
the ensemble of this build is generated on the GPU instead.  This module
restates the generator's definition with numpy:
  * uniform u(seed, c) = (SplitMix64(seed ^ SplitMix64(c)) >> 11) * 2^-53,
  * standard normals by Box-Muller,
    g(k) = sqrt(-2 ln(1 - u(2k))) cos(2 pi u(2k + 1)),
  * on-site block A = (X + X^H)/2 with X_ij = sigma (g(2q) + i g(2q + 1)),
    q = 16 s + 4 i + j, hopping t * I_4 to the 6 nearest neighbours.

Pinning (tests/test_lattice_oracle.py):
  * SplitMix64 against its published test vector (seed 1234567);
  * the sparsity pattern and hoppings against kwant_amd.lattices.cubic_4orb,
    an independent scipy COO construction;
  * the closed-form row offsets against a cumulative count of row lengths;
  * the normals' first two moments.
"""
import numpy as np
import scipy.sparse as sp

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)


def splitmix(z):
    """SplitMix64 output function applied to z + golden gamma (uint64 arrays)."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _C1
        z = (z ^ (z >> np.uint64(27))) * _C2
    return z ^ (z >> np.uint64(31))


def unit_uniform(seed, c):
    return (splitmix(np.uint64(seed) ^ splitmix(c)) >> np.uint64(11)).astype(np.float64) \
        * (1.0 / 9007199254740992.0)


def std_normal(seed, k):
    k = np.asarray(k, dtype=np.uint64)
    u1 = 1.0 - unit_uniform(seed, np.uint64(2) * k)
    u2 = unit_uniform(seed, np.uint64(2) * k + np.uint64(1))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def entries_before(L, x, y, z):
    """Closed-form count of CSR entries in the rows of all sites before
    (x, y, z) (the generator's row offsets)."""
    s = (x * L + y) * L + z
    cols = x * L + y
    nb = (cols * (L - 1) + np.maximum(z - 1, 0) + cols * (L - 1) + z
          + x * (L - 1) * L + np.where(y > 0, (y - 1) * L + z, 0)
          + x * (L - 1) * L + y * L + np.where(y < L - 1, z, 0)
          + np.where(x > 0, (x - 1) * L * L + y * L + z, 0)
          + x * L * L + np.where(x < L - 1, y * L + z, 0))
    return 4 * (4 * s + nb)


def cubic4(L, seed, sigma=0.3, t=1.0):
    """The generator's matrix for one system, as scipy CSR (complex128)."""
    ns = L ** 3
    n = 4 * ns
    s = np.arange(ns)
    x, y, z = s // (L * L), (s // L) % L, s % L
    q = (16 * s[:, None, None] + 4 * np.arange(4)[None, :, None]
         + np.arange(4)[None, None, :]).astype(np.uint64)
    xr = sigma * std_normal(seed, np.uint64(2) * q)
    xi = sigma * std_normal(seed, np.uint64(2) * q + np.uint64(1))
    ar = 0.5 * (xr + np.swapaxes(xr, 1, 2))
    ai = 0.5 * (xi - np.swapaxes(xi, 1, 2))
    rows, cols, vals = [], [], []
    orb = np.arange(4)
    r_on = (4 * s[:, None, None] + orb[None, :, None]) + 0 * orb[None, None, :]
    c_on = 4 * s[:, None, None] + orb[None, None, :] + 0 * orb[None, :, None]
    rows.append(r_on.ravel())
    cols.append(c_on.ravel())
    vals.append((ar + 1j * ai).ravel())
    for dx, dy, dz in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
        ok = ((x + dx >= 0) & (x + dx < L) & (y + dy >= 0) & (y + dy < L)
              & (z + dz >= 0) & (z + dz < L))
        s0 = s[ok]
        s1 = ((x[ok] + dx) * L + y[ok] + dy) * L + z[ok] + dz
        rows.append((4 * s0[:, None] + orb).ravel())
        cols.append((4 * s1[:, None] + orb).ravel())
        vals.append(np.full(4 * len(s0), t, dtype=complex))
    m = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    m.sort_indices()
    return m


def gershgorin_radius(m):
    return float(np.max(np.abs(m).sum(axis=1)))
