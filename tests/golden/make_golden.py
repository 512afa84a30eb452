"""Writes the golden fixtures of this directory from the CPU oracle (oracle/).

Run from the repository root: python tests/golden/make_golden.py

Each fixture holds its inputs and the oracle's outputs, so the parity tests can compare the
HIP path with fixed data without running the oracle:

  square64_right_p1.npz  UnitSquareMesh(64, 64, "right") (8,192 triangles, 4,225 vertices),
                         P1, kappa ~ U[0.5, 2] and f ~ U[-1, 1] per vertex drawn in that
                         order from numpy.random.default_rng(20190); CSR structure, values of
                         inner(kappa*grad u, grad v)*dx and the vector of f*v*dx.
  cube4_p1_bc.npz        UnitCubeMesh(4, 4, 4), P1, kappa = 1, f = 10, DirichletBC(V, 1.0,
                         "on_boundary"), SystemAssembler (python/test/unit/fem/
                         test_system_assembler.py:73-91; Frobenius 96.847818767384, l2
                         96.564760289080).
  cube2_vp2.npz          UnitCubeMesh(2, 2, 2), vector P2 (UFC numbering), mu ~ U[0.5, 1.5]
                         then lmbda ~ U[1, 10] per vertex from default_rng(20190); elasticity
                         matrix and the body force f = (0, 0, -1).
  tetrahedron.xml        the reference's one-cell mesh (python/test/unit/fem/tetrahedron.xml),
                         whose stiffness and mass matrices test_assembler.py:351-390 prints.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

SEED = 20190


def square64():
    x, cells = O.rectangle_mesh_right(64, 64)
    rng = np.random.default_rng(SEED)
    kappa = rng.uniform(0.5, 2.0, x.shape[0])
    f = rng.uniform(-1.0, 1.0, x.shape[0])
    rp, col = O.sparsity(cells, x.shape[0])
    A = O.assemble_matrix(O.P1_POISSON, x, cells, cells, rp, col, c0=kappa)
    b = O.assemble_vector(O.P1_LOAD, x, cells, cells, x.shape[0], c0=f)
    return dict(x=x, cells=cells, kappa=kappa, f=f, row_ptr=rp, col=col, A=A, b=b)


def cube4_bc():
    x, cells = O.box_mesh(4, 4, 4)
    rp, col = O.sparsity(cells, x.shape[0])
    bc = O.boundary_vertices(x).astype(np.uint8)
    ones = np.ones(x.shape[0])
    A, b = O.assemble_system(O.P1_POISSON, O.P1_LOAD, x, cells, cells, rp, col, bc, ones,
                             a_c0=ones, L_c0=np.full(x.shape[0], 10.0))
    return dict(x=x, cells=cells, bc_mask=bc, row_ptr=rp, col=col, A=A, b=b)


def cube2_vp2():
    x, cells = O.box_mesh(2, 2, 2)
    ce, ne = O.tet_edges(cells)
    cd = O.vp2_dofmap(cells, ce, x.shape[0], ne)
    ndofs = 3 * (x.shape[0] + ne)
    rng = np.random.default_rng(SEED)
    mu = rng.uniform(0.5, 1.5, x.shape[0])
    lmbda = rng.uniform(1.0, 10.0, x.shape[0])
    f3 = np.array([0.0, 0.0, -1.0])
    rp, col = O.sparsity(cd, ndofs)
    A = O.assemble_matrix(O.VP2_ELASTICITY, x, cells, cd, rp, col, coeff_dofs=cells, c0=mu,
                          c1=lmbda)
    b = O.assemble_vector(O.VP2_BODY_FORCE, x, cells, cd, ndofs, coeff_dofs=cells, f3=f3)
    return dict(x=x, cells=cells, cell_dofs=cd, mu=mu, lmbda=lmbda, f3=f3, row_ptr=rp, col=col,
                A=A, b=b)


FIXTURES = {"square64_right_p1": square64, "cube4_p1_bc": cube4_bc, "cube2_vp2": cube2_vp2}


def main():
    for name, make in FIXTURES.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **make())
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
