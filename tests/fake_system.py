"""A stand-in for a finalized kwant system (kwant cannot be imported here):
the attributes that kwant's operators and KPM module read."""
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
        rng = np.random.default_rng(seed)
        a = scipy.sparse.random(orb, orb, density=density, random_state=seed, dtype=complex,
                                data_rvs=lambda k: rng.normal(size=k) + 1j * rng.normal(size=k))
        self.h = ((a + a.conj().T) / 2 + scipy.sparse.diags(rng.uniform(-1, 1, orb))).tocsr()

        self._start = {}
        owner = np.zeros(orb, dtype=np.int64)
        for (first, norbs, offset), (nxt, _, _) in zip(ranges[:-1], ranges[1:]):
            for k in range(nxt - first):
                self._start[first + k] = (offset + k * norbs, norbs)
                owner[offset + k * norbs:offset + (k + 1) * norbs] = first + k
        coo = self.h.tocoo()
        edges = sorted({(int(owner[r]), int(owner[c])) for r, c in zip(coo.row, coo.col)
                        if owner[r] != owner[c]})
        self.graph = _Graph(site, edges)

    def hamiltonian_submatrix(self, params=None, sparse=True):
        return self.h

    def hamiltonian(self, i, j, *args, params=None):
        (a, na), (b, nb) = self._start[i], self._start[j]
        return self.h[a:a + na, b:b + nb].toarray()


class _Graph:
    def __init__(self, num_nodes, edges):
        self.num_nodes = num_nodes
        self._edges = list(edges)
        self._set = set(self._edges)

    def __iter__(self):
        return iter(self._edges)

    def has_edge(self, a, b):
        return (a, b) in self._set
