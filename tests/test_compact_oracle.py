"""CPU pins of the oracle's compact MIC(0)-PCG (oracle/flip_oracle.cpp,
orc_compact_pcg; src/flip_sim.cpp:858-1151), which checks the GPU comparator.

* On a single line of fluid cells the matrix is tridiagonal, MIC(0) has no
  dropped fill and is the exact Cholesky factor: M^-1 rhs equals
  numpy.linalg.solve of the independently assembled matrix, and PCG
  converges in one iteration.
* In 3D the PCG solution satisfies A x = -div/dt to the tolerance for A
  assembled independently with scipy (diag = sum of the six face betas, 0 at
  the domain boundary; -beta between fluid cells).
"""
import numpy as np
import scipy.sparse as sp

import oracle_lib as orc

EPS = 1e-6


def assemble(mass, beta):
    nz, ny, nx = mass.shape
    fl = mass > EPS
    idx = -np.ones(mass.shape, np.int64)
    idx[fl] = np.arange(fl.sum())
    rows, cols, vals, diag = [], [], [], np.zeros(fl.sum())
    for z, y, x in zip(*np.nonzero(fl)):
        i = idx[z, y, x]
        for d, (dz, dy, dx) in enumerate(((0, 0, 1), (0, 1, 0), (1, 0, 0))):
            n_d = (nx, ny, nz)[d]
            c = (x, y, z)[d]
            for s in (-1, 1):
                if not 0 <= c + s < n_d:
                    continue                               # domain boundary: beta 0
                Z, Y, X = z + s * dz, y + s * dy, x + s * dx
                b = float(beta[d][(z, y, x) if s < 0 else (Z, Y, X)])
                diag[i] += b
                if fl[Z, Y, X]:
                    rows.append(i); cols.append(idx[Z, Y, X]); vals.append(-b)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(len(diag),) * 2) + sp.diags(diag)
    return A, fl


def test_line_of_cells_is_exact_cholesky():
    rng = np.random.default_rng(1)
    nx, ny, nz = 12, 4, 4
    mass = np.zeros((nz, ny, nx), np.float32)
    mass[1, 2, :] = 1.0
    beta = [rng.uniform(0.5, 1.5, mass.shape).astype(np.float32) for _ in range(3)]
    div = rng.normal(0, 1, mass.shape).astype(np.float32)
    x, it, rr, pre, z = orc.compact_pcg([0, 0, 0], 4, mass, EPS, beta, div, tol=1e-6, want_factor=True)
    A, fl = assemble(mass, beta)
    ref = np.linalg.solve(A.toarray(), -div[fl].astype(np.float64))
    np.testing.assert_allclose(z, ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())
    assert it == 1 and rr <= 1e-6
    np.testing.assert_allclose(x[fl], ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())


def test_3d_solution_satisfies_the_system():
    rng = np.random.default_rng(2)
    n = 24
    zz, yy, xx = np.meshgrid(*(np.arange(n) + 0.5,) * 3, indexing="ij")
    mass = ((((xx - 12) ** 2 + (yy - 10) ** 2 + (zz - 13) ** 2) < 8.5 ** 2) | (yy < 4)).astype(np.float32)
    beta = [rng.uniform(0.5, 1.5, mass.shape).astype(np.float32) for _ in range(3)]
    div = rng.normal(0, 1, mass.shape).astype(np.float32)
    x, it, rr, _, _ = orc.compact_pcg([0], n, mass, EPS, beta, div, tol=1e-6, max_iter=500)
    A, fl = assemble(mass, beta)
    b = -div[fl].astype(np.float64)
    res = np.linalg.norm(b - A @ x[fl].astype(np.float64)) / np.linalg.norm(b)
    assert 1 < it < 500 and rr <= 1e-6 and res <= 3e-6, (it, rr, res)
    assert not np.any(x[~fl])
