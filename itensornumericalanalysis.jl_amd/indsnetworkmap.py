"""Digit site layout and the digit <-> coordinate map.

Restates, for base 2 and real coordinates:
  * default_dimension_vertices (src/digit_inds.jl:17-21): dimension i owns
    vertices vs[i::map_dimension] (interleaved), digit j = j-th of them;
  * continuous_siteinds / RealIndsNetworkMap (src/indsnetworkmap.jl:53-64,83);
  * calculate_ind_values + set_ind_values! (src/IndexMaps/realindexmap.jl:67-76,
    src/IndexMaps/abstractindexmap.jl:121-138): digit j of x is bit j (MSB
    first) of floor(x 2^L); x >= 1 saturates every digit to 1;
  * calculate_p / index_value_to_scalar (abstractindexmap.jl:103-119,
    realindexmap.jl:13-15) and grid_points (realindexmap.jl:78-86).
Dimensions and digits are 1-based, as in the reference API.
"""
from __future__ import annotations

import math

from .graphs import NamedGraph


class DigitMapError(ValueError):
    """Raised for coordinates the digit map does not accept (x < 0 or NaN)."""


def default_dimension_vertices(g: NamedGraph, map_dimension: int = 1):
    vs = g.vertices
    return [list(vs[i::map_dimension]) for i in range(map_dimension)]


class IndsNetworkMap:
    """A graph with one binary digit per vertex and the map from digits to coordinates."""

    def __init__(self, g: NamedGraph, dimension_vertices, base: int = 2, ambient: NamedGraph | None = None):
        if base != 2:
            raise NotImplementedError("only base 2 digits are supported by the engine")
        self.graph = g
        # the graph this layout was reduced from (its own graph otherwise):
        # products of functions on parts of it live on its induced subgraphs
        self.ambient = ambient if ambient is not None else g
        self.base = base
        self._dimvs = [list(d) for d in dimension_vertices]
        self._vertex_site = {}
        for d, verts in enumerate(self._dimvs, start=1):
            for j, v in enumerate(verts, start=1):
                if v not in g._index:
                    raise ValueError(f"dimension vertex {v} is not in the graph")
                if v in self._vertex_site:
                    raise NotImplementedError(
                        "vertices carrying more than one digit are not supported")
                self._vertex_site[v] = (d, j)

    # --- layout queries -------------------------------------------------
    def dimension(self):
        return len(self._dimvs)

    def dimensions(self):
        return list(range(1, self.dimension() + 1))

    def dimension_vertices(self, dim: int):
        return list(self._dimvs[dim - 1])

    def ndigits(self, dim: int):
        return len(self._dimvs[dim - 1])

    def vertex(self, dim: int, digit: int):
        return self._dimvs[dim - 1][digit - 1]

    def site(self, v):
        """(dimension, digit) carried by vertex v, or None."""
        return self._vertex_site.get(v)

    def vertices(self):
        return list(self.graph.vertices)

    def is_tree(self):
        return self.graph.is_tree()

    # --- digit map --------------------------------------------------------
    @staticmethod
    def digits_of(x: float, L: int):
        """Digits 1..L (MSB first) of x, as the greedy map of the reference."""
        x = float(x)
        if not (x >= 0.0):
            raise DigitMapError(f"coordinate {x!r} is negative or NaN")
        if x >= 1.0:
            k = (1 << L) - 1
        else:
            k = int(math.ldexp(x, L))
        return [(k >> (L - j)) & 1 for j in range(1, L + 1)]

    def calculate_ind_values(self, xs, dims=None):
        """Map each digit vertex of the given dimensions to its digit value."""
        if not isinstance(xs, (list, tuple)):
            xs = [xs]
        if dims is None:
            dims = list(range(1, len(xs) + 1))
        out = {}
        for d, x in zip(dims, xs):
            for j, bit in enumerate(self.digits_of(x, self.ndigits(d)), start=1):
                out[self.vertex(d, j)] = bit
        return out

    def calculate_p(self, ind_values, dims=None):
        if dims is None:
            dims = self.dimensions()
        res = []
        for d in dims:
            tot = 0.0
            for v, val in ind_values.items():
                site = self.site(v)
                if site is not None and site[0] == d:
                    tot += val * 2.0 ** (-site[1])
            res.append(tot)
        return res

    def grid_points(self, N: int, d: int):
        L = self.ndigits(d)
        a = round(2.0 ** L / N)
        pts = [i * (a / 2.0 ** L) for i in range(0, N + 2)]
        return [x for x in pts if x < 1]

    # --- derived layouts -------------------------------------------------
    @staticmethod
    def _trim(dimvs):
        # the dimension count is the largest dimension with digits
        # (src/IndexMaps/abstractindexmap.jl:72-73)
        while dimvs and not dimvs[-1]:
            dimvs = dimvs[:-1]
        return dimvs

    def reduced(self, dims):
        """The layout of the given dimensions only: the subgraph of the vertices
        carrying their digits (reduced_indsnetworkmap, src/indsnetworkmap.jl:29-46).
        Other dimensions keep their numbers with no digits."""
        dims = [dims] if isinstance(dims, int) else list(dims)
        keep = [v for d in dims for v in self._dimvs[d - 1]]
        g = self.ambient.induced(keep)
        if not g.is_tree():
            raise NotImplementedError("the vertices of these dimensions do not form a connected tree")
        dimvs = self._trim([list(vs) if d in dims else [] for d, vs in enumerate(self._dimvs, start=1)])
        return IndsNetworkMap(self.ambient.induced(keep), dimvs, self.base, self.ambient)

    def without_dimensions(self, dims):
        """Same graph, the digits of `dims` removed (their vertices carry none)."""
        dims = [dims] if isinstance(dims, int) else list(dims)
        dimvs = self._trim([[] if d in dims else list(vs) for d, vs in enumerate(self._dimvs, start=1)])
        return IndsNetworkMap(self.graph, dimvs, self.base, self.ambient)

    def union(self, other: "IndsNetworkMap"):
        """The layout of a product of functions on these two layouts: the union
        of the graphs and of the digit maps (merge of the index maps,
        src/elementary_operators.jl:261)."""
        if self.base != other.base:
            raise ValueError("different digit bases")
        if self.ambient is other.ambient:
            # parts of one layout: the induced subgraph on both vertex sets, so
            # parts that do not touch are joined by the layout's own edges
            verts = set(self.graph.vertices) | set(other.graph.vertices)
            g = self.ambient.induced(verts)
        else:
            g = self.graph.union(other.graph)
        n = max(len(self._dimvs), len(other._dimvs))
        dimvs = []
        for d in range(n):
            a = self._dimvs[d] if d < len(self._dimvs) else []
            b = other._dimvs[d] if d < len(other._dimvs) else []
            if a[:len(b)] != b[:len(a)]:
                raise ValueError(f"the layouts place the digits of dimension {d + 1} on different vertices")
            dimvs.append(list(a if len(a) >= len(b) else b))
        if not g.is_tree():
            raise NotImplementedError("the union of the two layouts is not a tree")
        return IndsNetworkMap(g, dimvs, self.base, self.ambient if self.ambient is other.ambient else None)

    def same_layout(self, other: "IndsNetworkMap"):
        return (self.graph.vertices == other.graph.vertices
                and {frozenset(e) for e in self.graph.edges} == {frozenset(e) for e in other.graph.edges}
                and self._dimvs == other._dimvs)

    def __repr__(self):
        return f"IndsNetworkMap(nv={self.graph.nv()}, dims={[len(d) for d in self._dimvs]})"


def reduced_indsnetworkmap(s: IndsNetworkMap, dims) -> IndsNetworkMap:
    """reduced_indsnetworkmap(s, dims) (src/indsnetworkmap.jl:29-46)."""
    return s.reduced(dims)


def continuous_siteinds(g: NamedGraph, dimension_vertices=None, *, map_dimension: int = 1,
                        base: int = 2) -> IndsNetworkMap:
    """`continuous_siteinds(g; map_dimension)` / `continuous_siteinds(g, dimension_vertices)`."""
    if dimension_vertices is None:
        dimension_vertices = default_dimension_vertices(g, map_dimension)
    return IndsNetworkMap(g, dimension_vertices, base)
