"""A stand-in for a finalized kwant system (kwant cannot be imported here):
the attributes that kwant's operators and KPM module read."""
import types

import numpy as np
import scipy.sparse


class FakeSystem:
    """The parts of a finalized kwant system that operators and the KPM
    engine read: site_ranges, graph.num_nodes, hamiltonian_submatrix."""

    def __init__(self, runs, seed=0, density=0.1):
        ranges, site, orb = [], 0, 0
        for count, norbs in runs:
            ranges.append((site, norbs, orb))
            site += count
            orb += count * norbs
        ranges.append((site, 0, orb))
        self.site_ranges = ranges
        self.graph = types.SimpleNamespace(num_nodes=site)
        rng = np.random.default_rng(seed)
        a = scipy.sparse.random(orb, orb, density=density, random_state=seed, dtype=complex,
                                data_rvs=lambda k: rng.normal(size=k) + 1j * rng.normal(size=k))
        self.h = ((a + a.conj().T) / 2 + scipy.sparse.diags(rng.uniform(-1, 1, orb))).tocsr()

    def hamiltonian_submatrix(self, params=None, sparse=True):
        return self.h
