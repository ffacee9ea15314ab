"""Graphs shared by the batched-form tests: a batch of edge cases (no vertex, isolated vertices only,
one edge, a star of degree 3000, a path, a 40-clique, two hubs of degree about 1200, a random sparse
graph, isolated vertices beside an edge)."""
import numpy as np

import components_amd as ca


def special_graphs():
    rng = np.random.default_rng(5)
    star = ca.Graph.from_edges(3001, [(0, i) for i in range(1, 3001)])
    path = ca.Graph.from_edges(1000, [(i, i + 1) for i in range(999)])
    clique = ca.Graph.from_edges(40, [(i, j) for i in range(40) for j in range(i + 1, 40)])
    two = ca.Graph.from_edges(9, [(0, 1), (1, 2), (5, 6), (6, 7), (7, 8)])
    hubs = ca.Graph.from_edges(2500, [(0, i) for i in range(1, 1200)] + [(1300, i) for i in range(1301, 2500)]
                               + [(5, 1305)])
    rnd = ca.Graph.from_edges(700, rng.integers(0, 700, size=(900, 2)))
    return [
        ca.Graph(np.zeros(1, np.uint64), np.zeros(0, np.uint32)),  # no vertex
        ca.Graph(np.zeros(2, np.uint64), np.zeros(0, np.uint32)),  # one vertex
        ca.Graph(np.zeros(6, np.uint64), np.zeros(0, np.uint32)),  # isolated vertices only
        ca.Graph.from_edges(2, [(0, 1)]),
        star, path, clique, two, hubs, rnd,
        ca.Graph.from_edges(5, [(3, 4)]),  # isolated vertices beside an edge
    ]
