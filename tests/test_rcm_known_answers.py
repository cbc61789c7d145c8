"""Known-answer orders of reverse_cuthill_mckee (solve.rs:200-270) on tiny graphs.

Each graph isolates one semantic that fixes the exact permutation, and its expected order is
derived by hand in the docstring from the reference's definitions:

  - the seed is the *first* unvisited vertex of minimum degree (``min_by_key`` keeps the first
    minimum, solve.rs:205);
  - George-Liu: the far vertex is the first vertex of minimum degree in the last BFS level, and
    the search restarts from it while its level structure is strictly deeper (solve.rs:208-219);
  - unvisited neighbours are visited in the stable order of ``sort_by_key`` by degree, so equal
    degrees keep the adjacency order (solve.rs:226-238);
  - one ``reverse()`` over the concatenated order of every component (solve.rs:243).

Graphs are given as two-vertex elements, so ``Graph::from_elements`` (iterative.rs:38-76)
lists each vertex's neighbours in the order of the edges that contain it.

The orders are checked on the oracle and on the library's host pattern build (no device is
involved), and a restatement with each plausible mistake planted is shown to miss the expected
order on the graph built for it, so each graph pins its semantic.
"""
import numpy as np
import pytest

import oracle


def _adjacency(n, edges):
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    return adj


# path 0-1-2-3. Degrees [1, 2, 2, 1]: the seed is 0, the first of the two minima. Levels from 0
# are [0], [1], [2], [3]; the far vertex 3 gives 4 levels too, no restart. Cuthill-McKee visits
# 0, 1, 2, 3, reversed: [3, 2, 1, 0]. (With the last minimum, 3, the result is [0, 1, 2, 3].)
PATH = (4, [(0, 1), (1, 2), (2, 3)], [3, 2, 1, 0])

# path 1-2-3-4-5 with the leaf 0 on 3; neighbours(3) = [2, 4, 0] in edge order. The seed is 0
# (degree 1, first). Levels from 0: [0], [3], [2, 4], [1, 5] (4 levels); the far vertex is 1,
# the first of the minimum-degree vertices {1, 5} of the last level. Levels from 1: [1], [2],
# [3], [4, 0], [5] (5 > 4 levels): restart from 1. Its far vertex 5 gives 5 levels: stop, the
# start is 1. From 1: 2, then 3, whose unvisited neighbours [4, 0] sort by degree (2, 1) to
# [0, 4], then 5 from 4: CM = [1, 2, 3, 0, 4, 5], reversed [5, 4, 0, 3, 2, 1].
# (No restart: start 0, CM [0, 3, 2, 4, 1, 5]. Last minimum for the far vertex: start 5.)
RESTART = (6, [(1, 2), (2, 3), (3, 4), (4, 5), (0, 3)], [5, 4, 0, 3, 2, 1])

# star with centre 0 whose neighbour list is [5, 3, 4, 1] (edge order), and 2 hanging on 1.
# Degrees: 0:4, 1:2, 2..5:1. The seed is 2 (the first degree-1 vertex). Levels from 2: [2], [1],
# [0], [5, 3, 4]; the far vertex 5 gives 4 levels: no restart. CM: 2, 1, 0, then the leaves,
# all of degree 1, in the stable order [5, 3, 4]: CM = [2, 1, 0, 5, 3, 4], reversed
# [4, 3, 5, 0, 1, 2]. (Ties broken by index instead: [5, 4, 3, 0, 1, 2].)
STABLE = (6, [(0, 5), (0, 3), (0, 4), (1, 0), (1, 2)], [4, 3, 5, 0, 1, 2])

# two components, the path 0-1-2 and the edge 3-4. The first seed is 0: CM 0, 1, 2; the next
# seed is 3 (first of {3, 4}): CM 3, 4. The whole order [0, 1, 2, 3, 4] is reversed once:
# [4, 3, 2, 1, 0]. (Reversing each component on its own gives [2, 1, 0, 4, 3].)
COMPONENTS = (5, [(0, 1), (1, 2), (3, 4)], [4, 3, 2, 1, 0])

CASES = {"first_minimum": PATH, "george_liu_restart": RESTART, "stable_degree_sort": STABLE,
         "global_reverse": COMPONENTS}


def _rcm(adj, last_minimum=False, no_restart=False, index_ties=False, per_component=False):
    """solve.rs:200-270 restated, with switches that plant one plausible mistake each."""
    n = len(adj)
    deg = [len(a) for a in adj]
    visited = [False] * n
    order = []

    def pick(cands):
        best = min(deg[v] for v in cands)
        hits = [v for v in cands if deg[v] == best]
        return hits[-1] if last_minimum else hits[0]

    def levels_from(start):
        seen = visited[:]
        seen[start] = True
        levels = [[start]]
        while True:
            nxt = []
            for v in levels[-1]:
                for w in adj[v]:
                    if not seen[w]:
                        seen[w] = True
                        nxt.append(w)
            if not nxt:
                return levels
            levels.append(nxt)

    while True:
        cand = [v for v in range(n) if not visited[v]]
        if not cand:
            break
        start = pick(cand)
        levels = levels_from(start)
        while not no_restart:
            far = pick(levels[-1])
            fl = levels_from(far)
            if len(fl) <= len(levels):
                break
            start, levels = far, fl
        first = len(order)
        visited[start] = True
        order.append(start)
        head = first
        while head < len(order):
            v = order[head]
            head += 1
            nxt = [w for w in adj[v] if not visited[w]]
            nxt.sort(key=(lambda w: (deg[w], w)) if index_ties else (lambda w: deg[w]))
            for w in nxt:
                if not visited[w]:
                    visited[w] = True
                    order.append(w)
        if per_component:
            order[first:] = order[first:][::-1]
    return order if per_component else order[::-1]


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_rcm_known_answer(name):
    n, edges, expected = CASES[name]
    g = oracle.Graph.from_elements(n, [list(e) for e in edges])
    off, tgt = g.arrays()
    assert [tgt[off[v]:off[v + 1]].tolist() for v in range(n)] == _adjacency(n, edges)
    assert g.rcm().tolist() == expected


@pytest.mark.parametrize("name", sorted(CASES))
def test_library_rcm_known_answer(name):
    """The library's host pattern build (ml_pattern_build), the numbering of every model."""
    import maille_amd as ma
    n, edges, expected = CASES[name]
    with ma.Graph.from_elements(n, [list(e) for e in edges]) as g:
        assert g.reverse_cuthill_mckee().tolist() == expected
        # Csr::zeros in that numbering: row i holds the ranks of i's lower neighbours
        off, cols = g.lower_pattern()
        rank = np.empty(n, dtype=np.int64)
        rank[np.asarray(expected)] = np.arange(n)
        adj = _adjacency(n, edges)
        for i in range(n):
            v = expected[i]
            lower = sorted(int(rank[w]) for w in adj[v] if rank[w] < i) + [i]
            assert cols[off[i]:off[i + 1]].tolist() == lower


MISTAKES = {"first_minimum": {"last_minimum": True},
            "george_liu_restart": {"no_restart": True},
            "stable_degree_sort": {"index_ties": True},
            "global_reverse": {"per_component": True}}


@pytest.mark.parametrize("name", sorted(CASES))
def test_each_graph_catches_its_planted_mistake(name):
    n, edges, expected = CASES[name]
    adj = _adjacency(n, edges)
    assert _rcm(adj) == expected
    assert _rcm(adj, **MISTAKES[name]) != expected
    # the restart graph also separates the far vertex's first minimum from the last
    if name == "george_liu_restart":
        assert _rcm(adj, last_minimum=True) != expected
