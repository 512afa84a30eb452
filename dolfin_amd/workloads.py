"""Seeded synthetic inputs of the benchmark/parity workloads.

Draw order from numpy.random.default_rng(seed): interior-vertex jitter (gdim numbers per
vertex, boundary vertices keep their position), then the P1 coefficients (kappa then f, or
mu then lmbda), then the state u when requested.
"""
import numpy as np
import torch

from .function import Function, FunctionSpace, VectorFunctionSpace
from .mesh import UnitCubeMesh, UnitSquareMesh, jitter_interior

DEFAULT_SEED = 20190


def p1_poisson_2d(n=64, seed=DEFAULT_SEED):
    """UnitSquareMesh n x n 'right', P1, kappa ~ U[0.5, 2], f ~ U[-1, 1] (no jitter)."""
    rng = np.random.default_rng(seed)
    mesh = UnitSquareMesh(n, n, "right")
    V = FunctionSpace(mesh, "P", 1)
    kappa = rng.uniform(0.5, 2.0, V.dim())
    f = rng.uniform(-1.0, 1.0, V.dim())
    return mesh, V, kappa, f


def p1_poisson_3d(n=128, seed=DEFAULT_SEED, jitter=0.2, with_u=False):
    """UnitCubeMesh n^3 Kuhn tets, interior jitter +-jitter*h, kappa ~ U[0.5,2], f ~ U[-1,1]."""
    rng = np.random.default_rng(seed)
    mesh = UnitCubeMesh(n, n, n)
    if jitter:
        jitter_interior(mesh, jitter / n, rng)
    V = FunctionSpace(mesh, "P", 1)
    kappa = rng.uniform(0.5, 2.0, V.dim())
    f = rng.uniform(-1.0, 1.0, V.dim())
    if with_u:
        u = rng.uniform(-1.0, 1.0, V.dim())
        return mesh, V, kappa, f, u
    return mesh, V, kappa, f


def vp2_elasticity_3d(n=64, seed=DEFAULT_SEED, jitter=0.15, with_u=False, node_permutation=None):
    """UnitCubeMesh n^3, vector P2, mu ~ U[0.5,1.5], lmbda ~ U[1,10] (P1), f = (0,0,-1)."""
    rng = np.random.default_rng(seed)
    mesh = UnitCubeMesh(n, n, n)
    if jitter:
        jitter_interior(mesh, jitter / n, rng)
    V = VectorFunctionSpace(mesh, "P", 2, node_permutation=node_permutation)
    Q = V.coefficient_space()
    mu = rng.uniform(0.5, 1.5, Q.dim())
    lmbda = rng.uniform(1.0, 10.0, Q.dim())
    f3 = np.array([0.0, 0.0, -1.0])
    if with_u:
        u = rng.uniform(-1.0, 1.0, V.dim())
        return mesh, V, mu, lmbda, f3, u
    return mesh, V, mu, lmbda, f3


def to_function(V, values):
    return Function(V, torch.from_numpy(np.ascontiguousarray(values)))


def p1_algorithmic_bytes(num_cells, nvc, num_vertices, gdim, nnz, num_rows, n_coeff=1,
                         with_vector=False):
    """Compulsory traffic of one P1 matrix assembly: int32 connectivity, coordinates,
    coefficient nodal values, CSR values written once, col read once, row_ptr read once
    (+ f read and b written once when the load vector is fused)."""
    b = (4 * nvc * num_cells + 8 * gdim * num_vertices + 8 * n_coeff * num_vertices
         + 8 * nnz + 4 * nnz + 4 * (num_rows + 1))
    if with_vector:
        b += 8 * num_vertices + 8 * num_rows
    return b


def vp2_algorithmic_bytes(num_cells, num_vertices, nnz, num_rows):
    """Vector P2 matrix: 10 int32 node ids per cell, coordinates, mu and lmbda, CSR values
    written once, col and row_ptr read once."""
    return (40 * num_cells + 24 * num_vertices + 16 * num_vertices + 8 * nnz + 4 * nnz
            + 4 * (num_rows + 1))
