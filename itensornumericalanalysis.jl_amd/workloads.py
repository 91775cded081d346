"""Synthetic workloads (seeded, numpy PCG64) for the five benchmark shapes.

  sine_1d         1D, 12 digits on a chain; f = sin(2 pi x) + 0.5 cos(6 pi x) + exp(-x)
           (bond 5); Dirichlet second derivative, cutoff 1e-12; 1e5 points, seed 0.
  laplacian_2d    2D, 20 digits per dimension interleaved on a 40-site chain;
           f = sum_k a_k sin(w_k x) exp(-b_k y), k = 1..8, a ~ N(0,1),
           w ~ U(1,40), b ~ U(0,3) (seed 1), bond 16; Laplacian (Dirichlet),
           cutoff 1e-10, maxdim 128; 1e6 points drawn after the coefficients.
  comb_3d         3D comb tree (3 teeth of 16 digits, tooth i = dimension i, most
           significant digit on the backbone), dense N(0,1) cores normalised,
           bond 128, seed 2; first derivative in x; recompression to maxdim 128.
  fit_2d          2D, 14 digits per dimension interleaved on a 28-site chain, cores
           N(0, 0.1^2), bond 32, seed 3; target
           g = exp(-((x-.3)^2 + (y-.6)^2)/.02) + .2 sin(12 pi x y); 2^20 points.
  fredholm_batch  512 chains, digits ~ U{8..24}, bonds ~ U{4..96}, N(0,1) cores, seed 4;
           kernel operator exp(-|x-y|); cutoff 1e-10; 4096 points each.
Julia's random streams cannot be reproduced here, so the draws are this
module's own (documented order: coefficients, then points).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .builders import add_host, cos_itn, exp_itn, multiply_host, random_itn, sin_itn
from .graphs import named_comb_tree, named_grid
from .indsnetworkmap import continuous_siteinds
from .operators import (first_derivative_operator, fredholm_operator, laplacian_operator,
                        second_derivative_operator)


@dataclass
class Workload:
    name: str
    s: object
    f: object
    op: object = None
    xs: np.ndarray = None
    cutoff: float = 0.0
    maxdim: int = 0
    extra: dict = field(default_factory=dict)


def sine_1d(npts: int = 100_000, seed: int = 0) -> Workload:
    s = continuous_siteinds(named_grid((12, 1)))
    f = add_host(add_host(sin_itn(s, k=2 * math.pi), cos_itn(s, k=6 * math.pi, c=0.5)), exp_itn(s, k=-1.0))
    op = second_derivative_operator(s)
    xs = np.random.default_rng(seed).uniform(0.0, 1.0, (npts, 1))
    return Workload("sine_1d: 1D L=12 chain, D2 Dirichlet, cutoff 1e-12", s, f, op, xs, 1e-12, 0)


def sine_1d_function(x):
    return np.sin(2 * np.pi * x) + 0.5 * np.cos(6 * np.pi * x) + np.exp(-x)


def laplacian_2d(npts: int = 1_000_000, seed: int = 1) -> Workload:
    s = continuous_siteinds(named_grid((40, 1)), map_dimension=2)
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(8)
    w = rng.uniform(1.0, 40.0, 8)
    b = rng.uniform(0.0, 3.0, 8)
    f = None
    for k in range(8):
        term = multiply_host(sin_itn(s, k=w[k], c=a[k], dim=1), exp_itn(s, k=-b[k], dim=2))
        f = term if f is None else add_host(f, term)
    xs = rng.uniform(0.0, 1.0, (npts, 2))
    op = laplacian_operator(s)
    return Workload("laplacian_2d: 2D 2x20 bits interleaved chain, Laplacian, cutoff 1e-10 maxdim 128",
                    s, f, op, xs, 1e-10, 128, {"a": a, "w": w, "b": b, "compress_maxdim": 64})


def laplacian_2d_function(xs, a, w, b):
    x, y = xs[:, 0], xs[:, 1]
    return sum(a[k] * np.sin(w[k] * x) * np.exp(-b[k] * y) for k in range(len(a)))


def comb_3d(seed: int = 2, bond: int = 128, teeth: int = 3, digits: int = 16) -> Workload:
    g = named_comb_tree((teeth, digits))
    s = continuous_siteinds(g, [[(i, j) for j in range(1, digits + 1)] for i in range(1, teeth + 1)])
    f = random_itn(s, link_space=bond, rng=np.random.default_rng(seed), normalize_cores=True)
    op = first_derivative_operator(s, dim=1)
    return Workload(f"comb_3d: comb {teeth}x{digits}, bond {bond}, d/dx, maxdim {bond}", s, f, op, None,
                    0.0, bond)


def fit_2d(npts: int = 1 << 20, seed: int = 3, bond: int = 32) -> Workload:
    s = continuous_siteinds(named_grid((28, 1)), map_dimension=2)
    rng = np.random.default_rng(seed)
    f = random_itn(s, link_space=bond, rng=rng, std=0.1)
    xs = rng.uniform(0.0, 1.0, (npts, 2))
    x, y = xs[:, 0], xs[:, 1]
    target = np.exp(-((x - 0.3) ** 2 + (y - 0.6) ** 2) / 0.02) + 0.2 * np.sin(12 * np.pi * x * y)
    return Workload("fit_2d: 2D 2x14 bits chain, bond 32, MSE gradient", s, f, None, xs, 0.0, 0,
                    {"target": target})


def fredholm_batch(nfun: int = 512, npts: int = 4096, seed: int = 4):
    """List of (Workload) for the batch of independent chains."""
    rng = np.random.default_rng(seed)
    Ls = rng.integers(8, 25, nfun)
    chis = rng.integers(4, 97, nfun)
    out = []
    for i in range(nfun):
        s = continuous_siteinds(named_grid((int(Ls[i]), 1)))
        f = random_itn(s, link_space=int(chis[i]), rng=rng)
        xs = rng.uniform(0.0, 1.0, (npts, 1))
        out.append(Workload(f"fredholm_batch[{i}]", s, f, fredholm_operator(s), xs, 1e-10, 0))
    return out
