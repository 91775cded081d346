"""Brute-force dense references (numpy).  TEST INFRASTRUCTURE ONLY.

`dense_tensor` contracts a whole tree network into the 2^n array of its values
(the definition behind build_full_rank_tensor, src/utils.jl:28-39); the
operator helpers build 2^n x 2^n matrices directly from operator strings, with
the digit ops of src/digit_inds.jl:35-58 read literally.
"""
from __future__ import annotations

import numpy as np

OPS = {
    "D+": np.array([[0.0, 0.0], [1.0, 0.0]]),  # o[2,1] = 1 (1-based): s_in=1, s_out=0
    "D-": np.array([[0.0, 1.0], [0.0, 0.0]]),  # o[1,2] = 1
    "Ddn": np.array([[1.0, 0.0], [0.0, 0.0]]),
    "Dup": np.array([[0.0, 0.0], [0.0, 1.0]]),
    "I": np.eye(2),
}


def _children(parent):
    ch = [[] for _ in parent]
    for v in range(1, len(parent)):
        ch[parent[v]].append(v)
    return ch


def dense_tensor(parent, cores):
    """Values of the network as an array with one axis per vertex (internal order)."""
    ch = _children(parent)
    sub = {}
    order = list(range(len(parent)))
    for v in reversed(order):
        # T[p, s_v, (subtree axes of c1), (subtree axes of c2) ...]
        T = np.asarray(cores[v])
        verts = [v]
        for k, c in enumerate(ch[v]):
            Tc, vc = sub.pop(c)  # Tc[p_c, sites of c-subtree...]
            # contract T axis (2) (the next child bond, always right after the
            # already-expanded axes) with Tc axis 0
            ax = 1 + len(verts)  # axes: p, site(s) of verts so far, then child bonds
            T = np.tensordot(T, Tc, axes=([ax], [0]))
            # tensordot appends Tc's site axes at the end; move them right after ax-1
            n_c = Tc.ndim - 1
            src = list(range(T.ndim - n_c, T.ndim))
            dst = list(range(ax, ax + n_c))
            T = np.moveaxis(T, src, dst)
            verts += vc
        sub[v] = (T, verts)
    T, verts = sub[0]
    T = T[0]
    # reorder axes to internal vertex order
    perm = np.argsort(verts)
    return np.transpose(T, perm)


def grid_values_1d(parent, cores, vdigit):
    """f on the 2^L grid of a one-dimensional network (k = 0..2^L-1)."""
    T = dense_tensor(parent, cores)
    L = len(vdigit)
    order = np.argsort(vdigit)  # axis of digit 1 (MSB) first
    return np.transpose(T, order).reshape(2 ** L)


def dense_operator_from_terms(nv, terms, op_at_vertex):
    """sum_t coef_t (x)_v O_{t,v} as a 2^nv x 2^nv matrix M[s_in, s_out] in internal
    vertex order.  terms: list of (coef, {vertex_id: opname})."""
    M = np.zeros((2 ** nv, 2 ** nv))
    for coef, ops in terms:
        K = np.array([[coef]])
        for v in range(nv):
            o = OPS[ops[v]] if v in ops else np.eye(2)
            K = np.kron(K, o)
        M += K
    return M


def dense_operator(parent, cores):
    """Contract an operator network [p][s_in][s_out][children] into M[s_in, s_out]."""
    nv = len(parent)
    # treat (s_in, s_out) as one site axis of size 4 and reuse dense_tensor
    flat = []
    for c in cores:
        c = np.asarray(c)
        flat.append(c.reshape((c.shape[0], c.shape[1] * c.shape[2]) + c.shape[3:]))
    T = dense_tensor(parent, flat)  # axes: one per vertex, size 4 = (s_in, s_out)
    T = T.reshape((2, 2) * nv)
    perm = list(range(0, 2 * nv, 2)) + list(range(1, 2 * nv, 2))
    return np.transpose(T, perm).reshape(2 ** nv, 2 ** nv)
