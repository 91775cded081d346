"""Small helpers shared by the tests (no engine or oracle logic of their own)."""
import numpy as np


def oracle_args(f):
    """(parent, sdim, cores, vcoord, vdigit, Lc) of a network for the oracle."""
    t = f.topo
    return t.parent, t.sdim, f.cores, t.vcoord, t.vdigit, t.Lc


def oracle_eval(po, f, xs):
    t = f.topo
    y, rc = po.evaluate(t.parent, t.sdim, f.cores, t.vcoord, t.vdigit, t.Lc, xs)
    assert rc == 0
    return y


def grid_of(xs, L):
    """grid values floor(x 2^L) / 2^L (what a network with L digits sees)."""
    return np.floor(np.asarray(xs) * 2.0 ** L) / 2.0 ** L


def rel_l2(y, ref):
    return float(np.linalg.norm(np.asarray(y) - np.asarray(ref)) / max(np.linalg.norm(ref), 1e-300))
