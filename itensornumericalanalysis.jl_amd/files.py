"""Network files: save and load function and operator networks in the engine's
interchange format (libqtt qtt_net_save / qtt_net_load / qtt_mpo_save /
qtt_mpo_load; the layout is documented in include/qtt.h).

The C library moves the cores between the file and the device; this module
writes the vertex names (Python literals, read back with ast.literal_eval) and
rebuilds the digit layout (graph and dimension vertices) from the header, the
part of an ITensorNetworkFunction (src/itensornetworkfunction.jl:18-57) the
engine itself does not keep.
"""
from __future__ import annotations

import ast
import ctypes as C
import os
import struct

import numpy as np

from . import _capi
from .graphs import NamedGraph
from .indsnetworkmap import IndsNetworkMap
from .network import ITensorNetworkFunction, Topology, TreeOperator


def _names_arg(topo: Topology):
    enc = [repr(n).encode() for n in topo.names]
    arr = (C.c_char_p * len(enc))(*enc)
    return arr, enc


def save(obj, path: str) -> None:
    """Write a function network or an operator network (with its vertex names)."""
    from .api import device_network, device_operator
    arr, keep = _names_arg(obj.topo)
    p = os.fsencode(path)
    if isinstance(obj, TreeOperator):
        d = device_operator(obj)
        d.ctx.check(_capi.lib().qtt_mpo_save(d.ctx.handle, d.handle, p, arr))
    else:
        d = device_network(obj)
        d.ctx.check(_capi.lib().qtt_net_save(d.ctx.handle, d.handle, p, arr))


def read_header(path: str):
    """(kind, topology arrays, bonds, vertex names or None) of a network file."""
    with open(path, "rb") as fh:
        head = fh.read(24)
        if len(head) < 24 or head[:4] != b"QTTN":
            raise ValueError(f"{path}: not a network file")
        version, kind, nv, ncoord, flags = struct.unpack("<5I", head[4:])
        if version != 1:
            raise ValueError(f"{path}: unknown format version {version}")
        ints = np.frombuffer(fh.read(4 * (4 * nv + ncoord)), dtype="<i4")
        parent, sdim, vcoord, vdigit = (ints[k * nv:(k + 1) * nv].astype(np.int32) for k in range(4))
        ndigits = ints[4 * nv:].astype(np.int32)
        bond = np.frombuffer(fh.read(8 * nv), dtype="<i8").astype(np.int64)
        names = None
        if flags & 1:
            names = []
            for _ in range(nv):
                (n,) = struct.unpack("<I", fh.read(4))
                names.append(fh.read(n).decode())
    return kind, dict(parent=parent, sdim=sdim, vcoord=vcoord, vdigit=vdigit, ndigits=ndigits), bond, names


def _parse_name(text: str):
    try:
        return ast.literal_eval(text)
    except (ValueError, SyntaxError):
        return text


def layout_of(path: str):
    """The digit layout (IndsNetworkMap) and vertex order stored in a file."""
    kind, t, bond, names = read_header(path)
    nv = len(t["parent"])
    verts = [_parse_name(n) for n in names] if names is not None else list(range(1, nv + 1))
    g = NamedGraph(verts, [(verts[v], verts[int(t["parent"][v])]) for v in range(1, nv)])
    dimvs = []
    for c in range(len(t["ndigits"])):
        by_digit = {int(t["vdigit"][v]): verts[v] for v in range(nv) if t["vcoord"][v] == c}
        if sorted(by_digit) != list(range(1, int(t["ndigits"][c]) + 1)):
            raise ValueError(f"{path}: the digits of coordinate {c + 1} are not 1..{int(t['ndigits'][c])}")
        dimvs.append([by_digit[j] for j in range(1, int(t["ndigits"][c]) + 1)])
    s = IndsNetworkMap(g, dimvs)
    topo = Topology(s)
    if topo.names != verts or not np.array_equal(topo.parent, t["parent"]):
        raise ValueError(f"{path}: the vertices are not in breadth-first order from vertex 0")
    return kind, s, topo


def load(path: str):
    """The network of a file, resident on the device: an ITensorNetworkFunction
    or a TreeOperator on the layout the file describes."""
    from .api import engine_context
    kind, s, topo = layout_of(path)
    ctx = engine_context()
    h = C.c_void_p()
    p = os.fsencode(path)
    if kind == 1:
        ctx.check(_capi.lib().qtt_mpo_load(ctx.handle, p, C.byref(h)))
        return TreeOperator(s, None, topo, device=_capi.DeviceOperator(ctx, topo, handle=h))
    ctx.check(_capi.lib().qtt_net_load(ctx.handle, p, C.byref(h)))
    return ITensorNetworkFunction(s, None, topo, device=_capi.DeviceNetwork(ctx, topo, h))
