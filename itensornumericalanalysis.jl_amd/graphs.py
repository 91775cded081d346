"""Named graphs used by quantics networks.

Mirrors the NamedGraphs generators the reference uses for its networks
(`named_grid`, `named_comb_tree`; e.g. test/test_operators.jl:17,49,
test/test_integration.jl:17) with the same vertex names and the same vertex
order: `Tuple.(CartesianIndices(dims))`, first coordinate fastest.
"""
from __future__ import annotations

import itertools


class NamedGraph:
    """An undirected graph with hashable vertex names kept in insertion order."""

    def __init__(self, vertices, edges):
        self.vertices = list(vertices)
        self._index = {v: i for i, v in enumerate(self.vertices)}
        if len(self._index) != len(self.vertices):
            raise ValueError("duplicate vertex names")
        self.edges = []
        self._adj = {v: [] for v in self.vertices}
        for a, b in edges:
            if a not in self._index or b not in self._index:
                raise ValueError(f"edge ({a}, {b}) names an unknown vertex")
            self.edges.append((a, b))
            self._adj[a].append(b)
            self._adj[b].append(a)

    def nv(self):
        return len(self.vertices)

    def neighbors(self, v):
        return sorted(self._adj[v], key=self._index.__getitem__)

    def index(self, v):
        return self._index[v]

    def is_tree(self):
        n = len(self.vertices)
        if len(self.edges) != n - 1:
            return False
        seen = {self.vertices[0]}
        stack = [self.vertices[0]]
        while stack:
            v = stack.pop()
            for u in self._adj[v]:
                if u not in seen:
                    seen.add(u)
                    stack.append(u)
        return len(seen) == n

    def induced(self, keep):
        """Subgraph on the vertices in `keep` (graph order), with the edges among them."""
        keep = set(keep)
        verts = [v for v in self.vertices if v in keep]
        return NamedGraph(verts, [(a, b) for a, b in self.edges if a in keep and b in keep])

    def union(self, other: "NamedGraph"):
        """This graph's vertices and edges, then the other's that are new."""
        verts = self.vertices + [v for v in other.vertices if v not in self._index]
        have = {frozenset(e) for e in self.edges}
        edges = list(self.edges) + [e for e in other.edges if frozenset(e) not in have]
        return NamedGraph(verts, edges)

    def __repr__(self):
        return f"NamedGraph(nv={self.nv()}, ne={len(self.edges)})"


def _cartesian(dims):
    # first coordinate fastest, 1-based, as Julia's CartesianIndices
    ranges = [range(1, d + 1) for d in dims]
    return [tuple(reversed(t)) for t in itertools.product(*reversed(ranges))]


def named_grid(dims):
    """Grid graph with vertices (i, j, ...), edges between unit-distance neighbours."""
    dims = tuple(int(d) for d in dims)
    verts = _cartesian(dims)
    vs = set(verts)
    edges = []
    for v in verts:
        for ax in range(len(dims)):
            u = list(v)
            u[ax] += 1
            u = tuple(u)
            if u in vs:
                edges.append((v, u))
    return NamedGraph(verts, edges)


def named_comb_tree(dims):
    """Comb tree: backbone (i, 1) for i = 1..nx, tooth (i, 1)-(i, 2)-...-(i, ny)."""
    nx, ny = (int(d) for d in dims)
    verts = _cartesian((nx, ny))
    edges = [((i, 1), (i + 1, 1)) for i in range(1, nx)]
    for i in range(1, nx + 1):
        edges += [((i, j), (i, j + 1)) for j in range(1, ny)]
    return NamedGraph(verts, edges)


def named_path_graph(n):
    """Chain with integer vertex names 1..n."""
    return NamedGraph(list(range(1, n + 1)), [(i, i + 1) for i in range(1, n)])
