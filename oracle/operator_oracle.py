"""CPU oracle: kwant.operator.Density evaluated as kwant evaluates it.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module, and only
as the checker; the product package ``kwant_amd`` never imports it.

``density(...)`` returns a callable op(bra, ket) with the semantics of
kwant/operator.pyx:470-560 (``_LocalOperator.__call__``) around the MAT_ELS
loop of ``Density._operate`` (kwant/operator.pyx:771-801):

    out[w] = sum_i sum_j conj(bra[s + i]) * M_w[i * norbs + j] * ket[s + j]

for every ``where`` entry w, s and norbs its site's first orbital and orbital
count (kwant/operator.pyx:114-128), M_w its onsite matrix read flattened with
the site's own norbs as the row stride; the result is real when
check_hermiticity and bra and ket are one object after ``asarray``, and
summed when ``sum``.  Onsite normalisation (kwant/operator.pyx:273-321): a
1 x 1 onsite becomes value * identity(max norbs).

Pinning (tests/test_operator_oracle.py): closed forms (identity onsite:
sum=True is vdot(bra, ket), one orbital per site gives conj(bra_i) ket_i),
a dense block-diagonal Q_w brute force for multi-orbital matrix onsites,
and a hand-computed case of the flattened read for a scalar onsite on sites
with fewer orbitals than the largest.
"""
import numpy as np


def _orbs(site_ranges, site):
    """(first orbital, norbs) of a site (kwant/operator.pyx:114-128)."""
    run = 0
    for k in range(len(site_ranges)):
        if site_ranges[k][0] <= site:
            run = k
    first, norbs, offset = site_ranges[run]
    return offset + (site - first) * norbs, norbs


def density(site_ranges, onsite, where, check_hermiticity=True, sum=False):
    """op(bra, ket) of a Density.  ``onsite``: a scalar or square matrix used
    on every site, or a list with one matrix per ``where`` entry."""
    site_ranges = [tuple(int(v) for v in r) for r in site_ranges]
    where = [int(w) for w in where]
    if isinstance(onsite, list):
        mats = [np.asarray(m, dtype=complex).reshape(-1) for m in onsite]
    else:
        mat = np.atleast_2d(np.asarray(onsite, dtype=complex))
        if mat.shape[0] == 1:
            max_norbs = max(r[1] for r in site_ranges)
            mat = mat[0, 0] * np.eye(max_norbs, dtype=complex)
        mats = [mat.reshape(-1)] * len(where)
    tot = site_ranges[-1][2]

    def op(bra, ket=None):
        bra = np.asarray(bra, dtype=complex)
        ket = bra if ket is None else np.asarray(ket, dtype=complex)
        if bra.shape != (tot,) or ket.shape != (tot,):
            raise ValueError('vector is incorrect shape')
        result = np.zeros(len(where), dtype=complex)
        for w, site in enumerate(where):
            s, norbs = _orbs(site_ranges, site)
            m = mats[w]
            tmp = 0j
            for i in range(norbs):
                for j in range(norbs):
                    tmp += bra[s + i].conjugate() * m[i * norbs + j] * ket[s + j]
            result[w] = tmp
        if check_hermiticity and bra is ket:
            result = result.real
        return np.sum(result) if sum else result

    return op
