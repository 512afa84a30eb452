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


def batch_specs(nmembers=256, seed=5):
    """The members of the batched workload as (kind, n, member seed): UnitCubeMesh(n)
    instances, n ~ U{8..48} drawn from default_rng(seed), even members (kind 0) P1
    kappa-Poisson, odd members (kind 1) vector-P2 elasticity, each with its own seed."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(8, 49, nmembers)
    return [(i % 2, int(n), int(rng.integers(1 << 30))) for i, n in enumerate(sizes)]


def batch_form(spec):
    """The bilinear form of one batch member (kind, n, seed): p1_poisson_3d or
    vp2_elasticity_3d with its seeded coefficients, on a fresh function space."""
    from .fem import elasticity_form, poisson_form
    kind, n, s = spec
    if kind == 0:
        mesh, V, kappa, f = p1_poisson_3d(n, seed=s)
        return poisson_form(V, to_function(V, kappa))
    mesh, V, mu, lmbda, f3 = vp2_elasticity_3d(n, seed=s)
    Q = V.coefficient_space()
    return elasticity_form(V, to_function(Q, mu), to_function(Q, lmbda))


def batch_members(nmembers=256, seed=5, log=None):
    """The batched workload's bilinear forms (batch_specs, batch_form)."""
    forms = []
    for i, spec in enumerate(batch_specs(nmembers, seed)):
        forms.append(batch_form(spec))
        if log is not None and (i + 1) % 32 == 0:
            log(f"batch: {i + 1}/{nmembers} members built")
    return forms


def batch_sample(specs, per_kind=6):
    """A tuning sample of a batch: per element kind, per_kind members spread evenly over the
    kind's sizes (smallest to largest)."""
    out = []
    for kind in (0, 1):
        ks = sorted((s for s in specs if s[0] == kind), key=lambda s: s[1])
        if not ks:
            continue
        idx = np.unique(np.linspace(0, len(ks) - 1, min(per_kind, len(ks))).round().astype(int))
        out += [ks[i] for i in idx]
    return out


def vp2_node_coordinates(mesh):
    """Coordinates of the vector-P2 nodes: the vertices, then the edge midpoints."""
    x = mesh.coordinates()
    ce, ne = mesh.cell_edges()
    c = mesh.cells()
    loc = np.array([[2, 3], [1, 3], [1, 2], [0, 3], [0, 2], [0, 1]])
    ex = np.zeros((ne, 3))
    ex[ce.ravel()] = (0.5 * (x[c[:, loc[:, 0]]] + x[c[:, loc[:, 1]]])).reshape(-1, 3)
    return np.vstack([x, ex])


def rigid_body_modes(V):
    """The six rigid-body displacements (three translations, three rotations about the
    cube's centre) as dof vectors of the vector-P2 space V (P2 interpolates them exactly),
    for UFC (component-major) and node-permuted numberings alike."""
    X = vp2_node_coordinates(V.mesh())
    nn = X.shape[0]
    ce, ne = V.mesh().cell_edges()
    nodes = np.concatenate([V.mesh().cells(), V.mesh().num_vertices() + ce], axis=1)
    cd = V.cell_dofs()
    modes = [np.eye(3)[k] * np.ones((nn, 3)) for k in range(3)]
    modes += [np.cross(axis, X - 0.5) for axis in np.eye(3)]
    out = []
    for m in modes:
        r = np.zeros(V.dim())
        for comp in range(3):
            r[cd[:, 10 * comp:10 * comp + 10].ravel()] = m[nodes.ravel(), comp]
        out.append(r)
    return out
