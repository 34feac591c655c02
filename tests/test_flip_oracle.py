"""CPU pins of the oracle's FLIP caller steps (oracle/flip_oracle.cpp):
flipRhsCB and the pressure-gradient velocity update of
src/flip_sim.cpp:1317-1349.

* Hand-derived values on a two-cell liquid: which face index belongs to
  which cell pair, the pressure of an air side taken as 0, the index-0 faces
  zeroed, faces between air cells zeroed.
* The mathematics fixes the rest: with p solving A p = div/dt for the
  operator A p_i = sum_f beta_f (p_i - p_nb) (air neighbours p = 0; assembled
  here independently with scipy), the update applied to -p leaves a
  divergence-free velocity on the liquid cells, and the literal update (with
  +p, as the reference writes it after the MSBG solve) doubles the
  divergence -- the sign slip between flipRhsCB's +div/dt and the compact
  path's -div/dt (src/flip_sim.cpp:329-338, 936-939).
* Blocks without level-0 data read 0 and are never written; non-finite
  results become 0.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_lib as orc

EPS = 1e-6


def div_of(level, bsx, mass, vel, dt=1.0):
    return orc.flip_rhs(level, bsx, mass, EPS, vel, dt) * dt


def test_two_cell_liquid_by_hand():
    n = 4
    mass = np.zeros((n, n, n), np.float32)
    mass[1, 1, 1] = mass[1, 1, 2] = 1.0               # cells (x=1,2, y=1, z=1)
    p = np.zeros_like(mass)
    p[1, 1, 1], p[1, 1, 2] = 2.0, 5.0
    beta = [np.ones_like(mass) for _ in range(3)]
    vel = [np.ones_like(mass) for _ in range(3)]
    u, v, w = orc.flip_update_velocity([0], n, mass, EPS, p, beta, vel, 0.5, negate_p=False)
    # u on the -x face of (x, 1, 1): x=0 wall; x=1 air|2; x=2 2|5; x=3 5|air
    np.testing.assert_array_equal(u[1, 1], [0.0, 1 - 0.5 * 2, 1 - 0.5 * 3, 1 + 0.5 * 5])
    # v below and above each liquid cell
    assert v[1, 1, 1] == 1 - 0.5 * 2 and v[1, 1, 2] == 1 - 0.5 * 5
    assert v[1, 2, 1] == 1 + 0.5 * 2 and v[1, 2, 2] == 1 + 0.5 * 5
    assert w[1, 1, 1] == 1 - 0.5 * 2 and w[2, 1, 2] == 1 + 0.5 * 5
    # every face with air on both sides (and every index-0 face) is zero
    touched = np.zeros_like(mass, bool)
    touched[1, 1, 1:4] = True
    assert np.all(u[~touched] == 0)
    np.testing.assert_array_equal(v[:, 0], 0)
    np.testing.assert_array_equal(w[0], 0)
    # the negated update flips every gradient term
    un, _, _ = orc.flip_update_velocity([0], n, mass, EPS, p, beta, vel, 0.5, negate_p=True)
    np.testing.assert_array_equal(un[1, 1], [0.0, 1 + 0.5 * 2, 1 + 0.5 * 3, 1 - 0.5 * 5])
    # flipRhsCB: the divergence over dt on liquid cells only
    r = orc.flip_rhs([0], n, mass, EPS, vel, 0.25)
    assert r[1, 1, 1] == 0.0 and np.count_nonzero(r) == 0   # uniform velocity
    vel2 = [a.copy() for a in vel]
    vel2[0][1, 1, 2] = 3.0                             # face between the two cells
    r = orc.flip_rhs([0], n, mass, EPS, vel2, 0.25)
    assert r[1, 1, 1] == (3.0 - 1.0) / 0.25 and r[1, 1, 2] == (1.0 - 3.0) / 0.25
    assert np.count_nonzero(r) == 2


def laplacian(mass, beta):
    """A p_i = sum over the 6 faces of beta_f (p_i - p_nb), liquid cells only,
    air neighbours p = 0; beta of the face between cells c and c + e_d is
    beta[d] at c + e_d."""
    nz, ny, nx = mass.shape
    fl = mass > EPS
    idx = -np.ones(mass.shape, np.int64)
    idx[fl] = np.arange(fl.sum())
    rows, cols, vals = [], [], []
    diag = np.zeros(fl.sum())
    for z, y, x in zip(*np.nonzero(fl)):
        i = idx[z, y, x]
        for d, (dz, dy, dx) in enumerate(((0, 0, 1), (0, 1, 0), (1, 0, 0))):
            for s in (-1, 1):
                Z, Y, X = z + s * dz, y + s * dy, x + s * dx
                cz, cy, cx = (z, y, x) if s < 0 else (Z, Y, X)
                b = float(beta[d][cz, cy, cx])
                diag[i] += b
                if fl[Z, Y, X]:
                    rows.append(i)
                    cols.append(idx[Z, Y, X])
                    vals.append(-b)
    n = int(fl.sum())
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n)) + sp.diags(diag)
    return A, fl, idx


def test_negated_update_removes_the_divergence_literal_update_doubles_it():
    rng = np.random.default_rng(4)
    n = 16
    zz, yy, xx = np.meshgrid(*(np.arange(n) + 0.5,) * 3, indexing="ij")
    mass = (((xx - 8) ** 2 + (yy - 7.5) ** 2 + (zz - 8.5) ** 2) < 5.2 ** 2).astype(np.float32)
    beta = [rng.uniform(0.5, 1.5, mass.shape).astype(np.float32) for _ in range(3)]
    vel = [rng.normal(0, 1, mass.shape).astype(np.float32) for _ in range(3)]
    dt = 0.1
    b = orc.flip_rhs([0], n, mass, EPS, vel, dt)
    A, fl, idx = laplacian(mass, beta)
    x = spla.spsolve(A.tocsc(), b[fl].astype(np.float64))
    p = np.zeros_like(mass)
    p[fl] = x
    d0 = div_of([0], n, mass, vel)[fl]
    new = orc.flip_update_velocity([0], n, mass, EPS, p, beta, vel, dt, negate_p=True)
    d1 = div_of([0], n, mass, new)[fl]
    assert np.linalg.norm(d1) <= 1e-5 * np.linalg.norm(d0)
    lit = orc.flip_update_velocity([0], n, mass, EPS, p, beta, vel, dt, negate_p=False)
    d2 = div_of([0], n, mass, lit)[fl]
    assert np.linalg.norm(d2 - 2 * d0) <= 1e-5 * np.linalg.norm(d0)
    # a mistake the identity catches: beta taken from the wrong side of the face
    shifted = [np.roll(a, 1, axis=2 - k) for k, a in enumerate(beta)]
    bad = orc.flip_update_velocity([0], n, mass, EPS, p, shifted, vel, dt, negate_p=True)
    assert np.linalg.norm(div_of([0], n, mass, bad)[fl]) > 1e-2 * np.linalg.norm(d0)


def test_blocks_without_level0_data_and_non_finite_values():
    rng = np.random.default_rng(7)
    nx, ny, nz, bsx = 32, 16, 16, 16
    level = [0, 1]                                   # block 1 holds no level-0 data
    mass = np.zeros((nz, ny, nx), np.float32)
    mass[4:12, 4:12, 10:22] = 1.0                    # liquid on both sides of the block face x = 16
    p = rng.normal(0, 1, mass.shape).astype(np.float32)
    beta = [np.ones_like(mass) for _ in range(3)]
    vel = [rng.normal(0, 1, mass.shape).astype(np.float32) for _ in range(3)]
    vel[1][6, 6, 12] = np.nan
    u, v, w = orc.flip_update_velocity(level, bsx, mass, 1e-6, p, beta, vel, 0.5, negate_p=False)
    for a, a0 in zip((u, v, w), vel):
        np.testing.assert_array_equal(a[:, :, 16:], a0[:, :, 16:])   # never written
    # -x face of x = 15: both sides in block 0, liquid
    assert u[6, 6, 15] == vel[0][6, 6, 15] - 0.5 * 1.0 * (p[6, 6, 15] - p[6, 6, 14])
    assert v[6, 6, 12] == 0.0                                        # NaN in, liquid face: 0 out
    r = orc.flip_rhs(level, bsx, mass, 1e-6, vel, 2.0)
    # u at x = 16 lies in the block without data and reads 0
    e = ((((np.float32(0.0) - vel[0][6, 5, 15]) + vel[1][6, 6, 15]) - vel[1][6, 5, 15]) + vel[2][7, 5, 15]) - vel[2][6, 5, 15]
    assert r[6, 5, 15] == e / np.float32(2.0)
    np.testing.assert_array_equal(r[:, :, 16:], 0)
