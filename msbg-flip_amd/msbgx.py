"""Python host mirror of the msbgx C ABI (include/msbgx.h).

Mirrors the reference's MultiresSparseGrid pressure-solver interface
(src/msbg.h:829-940): ``MultiresSparseGrid.create`` / ``regularize_refinement_map``
/ ``set_refinement_map`` / ``set_domain_boundary_spec`` /
``prepare_pressure_solve_flip`` / ``solve_pressure_flip``, over ctypes.  The
product path is the HIP library ``libmsbgx.so`` next to this file; importing
this module fails loudly when that library has not been built.

Also holds the synthetic scene builder for the five benchmark configurations.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MSBGX_LIB") or os.path.join(HERE, "libmsbgx.so")   # MSBGX_LIB: sanitizer builds

MAX_SPHERES = 64
MAX_BOXES = 4

CH_PRESSURE, CH_DIVERGENCE, CH_R, CH_Z, CH_D, CH_Q = 0, 1, 2, 3, 4, 5
CH_FACE_U, CH_FACE_V, CH_FACE_W, CH_DIAGONAL, CH_FLAGS = 7, 8, 9, 10, 11
DBC_OPEN, DBC_SOLID = 0, 1
OPT_LITERAL_VCYCLE, OPT_NO_GRAPH, OPT_NO_COMPACT, OPT_UNFUSED_COARSE = 0x1, 0x2, 0x4, 0x8
OPT_FP32_DIRECTIONS = 0x10
PREC_FP32, PREC_FP16, PREC_BF16 = 0, 1, 2
CELL_SOLID, CELL_VOID = 0x0001, 0x1000
BLK_COARSE_FINE, BLK_FINE_COARSE, BLK_NO_FLUID, BLK_EXISTS = 1 << 2, 1 << 3, 1 << 5, 1 << 12


class Scene(C.Structure):
    """ctypes image of ``msbgx_scene`` (include/msbgx_scene.h)."""

    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("n_spheres", C.c_int32), ("n_boxes", C.c_int32), ("wall_faces_closed", C.c_int32),
        ("seed", C.c_uint64), ("dt", C.c_float), ("vel_scale", C.c_float),
        ("sphere", (C.c_double * 4) * MAX_SPHERES), ("box", (C.c_double * 6) * MAX_BOXES),
        ("particles_per_cell", C.c_int32), ("particle_sigma", C.c_float),
    ]


class SolveParams(C.Structure):
    _fields_ = [("n_pre", C.c_int), ("n_post", C.c_int), ("n_coarse", C.c_int),
                ("max_iter", C.c_int), ("tol", C.c_float)]


class Callbacks(C.Structure):
    _fields_ = [("sample_cell_type", C.CFUNCTYPE(C.c_uint16, C.c_void_p, C.c_int, C.c_int, C.c_int)),
                ("sample_face_coeff", C.CFUNCTYPE(C.c_float, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int)),
                ("sample_rhs", C.CFUNCTYPE(C.c_float, C.c_void_p, C.c_int, C.c_int, C.c_int)),
                ("user", C.c_void_p)]


def make_scene(n, spheres=(), boxes=(), seed=1234, dt=1.0, vel_scale=1.0, walls_closed=True,
               particles_per_cell=0, particle_sigma=0.5):
    """Build a Scene for an n^3 domain (n may be a 3-tuple)."""
    s = Scene()
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    s.nx, s.ny, s.nz = nx, ny, nz
    if len(spheres) > MAX_SPHERES or len(boxes) > MAX_BOXES:
        raise ValueError("too many primitives")
    s.n_spheres, s.n_boxes = len(spheres), len(boxes)
    for i, sp in enumerate(spheres):
        for k in range(4):
            s.sphere[i][k] = float(sp[k])
    for i, b in enumerate(boxes):
        for k in range(6):
            s.box[i][k] = float(b[k])
    s.seed, s.dt, s.vel_scale = int(seed), float(dt), float(vel_scale)
    s.wall_faces_closed = 1 if walls_closed else 0
    s.particles_per_cell, s.particle_sigma = int(particles_per_cell), float(particle_sigma)
    return s


@dataclass
class Config:
    """One synthetic workload: scene plus block size, levels and level policy."""
    name: str
    n: int
    bsx: int
    levels: int
    scene: Scene
    level_mode: int  # -1: every block level 0; else the flip_sim policy mode
    level_d0: int = 1


def config_c1(n=256, seed=1234):
    """Plain MG-PCG, single level: sphere r=0.3 at (0.5,0.5,0.4), 16^3 blocks."""
    s = make_scene(n, spheres=[(0.5, 0.5, 0.4, 0.3)], seed=seed)
    return Config("c1_sphere_single_level", n, 16, 3, s, level_mode=-1)


def config_c2(n=512, bsx=32, seed=1234):
    """Multires: bottom slab y<0.2 plus 64 spheres r~U[0.03,0.08], c~U[0.1,0.9]^3,
    3 levels assigned by block distance to the liquid surface (fine within 2 blocks)."""
    rng = np.random.default_rng(seed)
    sph = [(*rng.uniform(0.1, 0.9, 3), rng.uniform(0.03, 0.08)) for _ in range(64)]
    s = make_scene(n, spheres=sph, boxes=[(0.0, 1.0, 0.0, 0.2, 0.0, 1.0)], seed=seed)
    return Config("c2_slab_spheres_multires", n, bsx, 3, s, level_mode=1, level_d0=2)


def config_c3(seed=1234, n_problems=64):
    """Batched form: n_problems independent scenes, sizes from {64,128,192,256}^3,
    16^3 blocks, 2 MG levels, each problem either single-resolution (every block
    level 0) or two-resolution (level 0 within one block of liquid); 1-8 spheres
    with radius ~U[0.1,0.3] and centre ~U[0.2,0.8]^3; velocity seed per problem."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_problems):
        n = int(rng.choice([64, 128, 192, 256]))
        k = int(rng.integers(1, 9))
        sph = [(*rng.uniform(0.2, 0.8, 3), rng.uniform(0.1, 0.3)) for _ in range(k)]
        mode = -1 if rng.integers(0, 2) == 0 else 0
        s = make_scene(n, spheres=sph, seed=seed * 1000 + i)
        out.append(Config(f"c3_problem{i}_{n}", n, 16, 2, s, level_mode=mode, level_d0=1))
    return out


def config_c4(n=256, seed=1234):
    """Dam break: box x<0.4, y<0.6 plus a sphere r~U[0.1,0.2]."""
    rng = np.random.default_rng(seed)
    c = (rng.uniform(0.3, 0.7), rng.uniform(0.6, 0.9), rng.uniform(0.3, 0.7))
    r = rng.uniform(0.1, 0.2)
    s = make_scene(n, spheres=[(*c, r)], boxes=[(0.0, 0.4, 0.0, 0.6, 0.0, 1.0)], seed=seed)
    return Config("c4_dam_break", n, 16, 3, s, level_mode=0, level_d0=1)


def config_c5(n=1024, seed=1234, bsx=32, levels=4):
    """Full scale: dam-break box plus sphere r=0.15 at (0.7,0.8,0.5), 32^3 blocks, 4 levels; the
    right-hand side is the divergence of the particle-to-grid velocity of 8 seeded particles per
    liquid cell (uniform positions, velocity sigma 0.5)."""
    s = make_scene(n, spheres=[(0.7, 0.8, 0.5, 0.15)], boxes=[(0.0, 0.4, 0.0, 0.6, 0.0, 1.0)], seed=seed,
                   particles_per_cell=8, particle_sigma=0.5)
    return Config("c5_dam_break_particles", n, bsx, levels, s, level_mode=0, level_d0=1)


class P2GParams(C.Structure):
    """msbgx_p2g_params (FlipConfig fields, src/flip_sim.h:34-42)."""
    _fields_ = [("rho_l", C.c_float), ("rho_g", C.c_float), ("mass_eps", C.c_float), ("beta_min", C.c_float),
                ("alpha_phi", C.c_float), ("particles_per_voxel", C.c_int32)]


#: msbgx_particle / FlipParticle (src/flip_sim.h:48): position, velocity, phase
PARTICLE_DTYPE = np.dtype([("pos", np.float32, 3), ("vel", np.float32, 3), ("phase", np.int32)])


PDE_OPT_8_COLOR = 1 << 29          # OPT_8_COLOR_SCHEME (src/msbg.h:132)
PDE_MEAN_CURVATURE, PDE_BIHARMONIC = 4, 2
PDE_U16, PDE_F32 = 0, 1


class PdeParams(C.Structure):
    """msbgx_pde_params: applyChannelPdeFast's parameters (src/msbg.h:2176-2198)."""
    _fields_ = [("lapl_typ", C.c_int), ("do_constr_zero_one", C.c_int), ("dens_thr", C.c_float),
                ("dens_thr_tol1", C.c_float), ("dens_thr_tol2", C.c_float), ("num_iter", C.c_int),
                ("grid_spacing", C.c_float), ("time_step", C.c_float), ("taubin", C.c_float),
                ("anis_diff_alpha", C.c_float)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libmsbgx.so not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "msbgx_last_error": (C.c_char_p, []),
        "msbgx_create": (C.c_int, [C.POINTER(P), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "msbgx_destroy": (None, [P]),
        "msbgx_num_blocks": (C.c_int, [P]),
        "msbgx_set_options": (C.c_int, [P, C.c_uint]),
        "msbgx_set_domain_bc": (C.c_int, [P] + [C.c_int] * 6),
        "msbgx_regularize_refinement_map": (C.c_int, [P, P, P]),
        "msbgx_set_refinement_map": (C.c_int, [P, P]),
        "msbgx_prepare_pressure_solve": (C.c_int, [P, C.POINTER(Callbacks)]),
        "msbgx_prepare_pressure_solve_dense": (C.c_int, [P, P, P, P, P, P]),
        "msbgx_prepare_pressure_solve_scene": (C.c_int, [P, C.POINTER(Scene)]),
        "msbgx_solve_pressure": (C.c_int, [P, C.POINTER(SolveParams), C.POINTER(C.c_int), C.POINTER(C.c_float)]),
        "msbgx_residual_history": (C.c_int, [P, P, C.c_int]),
        "msbgx_last_solve_ms": (C.c_double, [P]),
        "msbgx_get_channel_block": (C.c_int, [P, C.c_int, C.c_int, C.c_int, P]),
        "msbgx_set_channel_block": (C.c_int, [P, C.c_int, C.c_int, C.c_int, P]),
        "msbgx_block_info": (C.c_int, [P, C.c_int, P, P]),
        "msbgx_relax_list": (C.c_int, [P, C.c_int, P, C.c_int]),
        "msbgx_omega": (C.c_float, [P]),
        "msbgx_relax": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "msbgx_multiply": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "msbgx_downsample": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
        "msbgx_prolongate": (C.c_int, [P, C.c_int, C.c_int]),
        "msbgx_zero_channel": (C.c_int, [P, C.c_int, C.c_int]),
        "msbgx_copy_channel": (C.c_int, [P, C.c_int, C.c_int]),
        "msbgx_vcycle": (C.c_int, [P, C.c_int, C.c_int, C.c_int]),
        "msbgx_time_kernel": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_int)]),
        "msbgx_level_map_from_scene": (C.c_int, [C.POINTER(Scene), C.c_int, C.c_int, C.c_int, C.c_int, P]),
        "msbgx_set_precision": (C.c_int, [P, C.c_int]),
        "msbgx_solve_pressure_refined": (C.c_int, [P, C.POINTER(SolveParams), C.c_double, C.c_int,
                                                   C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double), P]),
        "msbgx_level0_blocks": (C.c_int, [P, P, C.c_int]),
        "msbgx_batch_create": (C.c_int, [C.POINTER(P), P, C.c_int]),
        "msbgx_batch_size": (C.c_int, [P]),
        "msbgx_batch_block_base": (C.c_int, [P, C.c_int]),
        "msbgx_batch_solve": (C.c_int, [P, C.POINTER(SolveParams), P, P]),
        "msbgx_batch_residual_history": (C.c_int, [P, C.c_int, P, C.c_int]),
        "msbgx_autotune": (C.c_int, [P, C.POINTER(SolveParams), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "msbgx_set_launch_config": (C.c_int, [P, C.c_int, C.c_int]),
        "msbgx_net_create": (C.c_int, [C.POINTER(P), C.c_uint64]),
        "msbgx_net_destroy": (None, [P]),
        "msbgx_net_get_weights": (C.c_int, [P, P]),
        "msbgx_net_set_weights": (C.c_int, [P, P]),
        "msbgx_net_set_samples": (C.c_int, [P, C.c_int, P, P, P]),
        "msbgx_net_loss_grad": (C.c_int, [P, C.POINTER(C.c_double), P, P]),
        "msbgx_net_collect": (C.c_int, [P, P, C.POINTER(SolveParams), C.c_int, C.c_int, C.c_int, C.c_uint64,
                                        C.POINTER(C.c_int)]),
        "msbgx_net_num_samples": (C.c_int, [P]),
        "msbgx_net_set_mode": (C.c_int, [P, C.c_int]),
        "msbgx_net_get_mode": (C.c_int, [P]),
        "msbgx_net_set_prolongation": (C.c_int, [P, C.c_int]),
        "msbgx_net_get_prolongation": (C.c_int, [P]),
        "msbgx_net_set_coarse_scales": (C.c_int, [P, C.c_int]),
        "msbgx_net_get_coarse_scales": (C.c_int, [P]),
        "msbgx_net_set_stall": (C.c_int, [P, C.c_int, C.c_float]),
        "msbgx_net_train": (C.c_int, [P, C.c_int, C.c_int, C.c_float, C.c_uint64, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double)]),
        "msbgx_net_last_train_ms": (C.c_double, [P]),
        "msbgx_net_apply": (C.c_int, [P, P, C.c_int, C.c_int]),
        "msbgx_net_time_conv": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "msbgx_solve_pressure_learned": (C.c_int, [P, P, C.POINTER(SolveParams), C.POINTER(C.c_int),
                                                   C.POINTER(C.c_float), C.POINTER(C.c_int)]),
        "msbgx_default_p2g_params": (None, [C.POINTER(P2GParams)]),
        "msbgx_p2g_create": (C.c_int, [C.POINTER(P), P]),
        "msbgx_p2g_destroy": (None, [P]),
        "msbgx_p2g_run": (C.c_int, [P, C.c_int64, P, C.POINTER(P2GParams)]),
        "msbgx_p2g_run_scene": (C.c_int, [P, C.POINTER(Scene), C.POINTER(P2GParams), C.POINTER(C.c_int64)]),
        "msbgx_p2g_get": (C.c_int, [P] + [P] * 7),
        "msbgx_p2g_active_blocks": (C.c_int, [P, P, C.c_int]),
        "msbgx_p2g_last_times": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
        "msbgx_p2g_to_velocity": (C.c_int, [P, P]),
        "msbgx_solve_pressure_compact": (C.c_int, [P, C.c_int, C.c_float, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
        "msbgx_compact_mic0": (C.c_int, [P, P, P, C.c_int]),
        "msbgx_pde_mean_curvature": (C.c_int, [P, P, P, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_double)]),
        "msbgx_pde_mean_curvature_noalloc": (C.c_int, [P, P, P, C.c_int, C.c_int, C.c_float]),
        "msbgx_pde_apply": (C.c_int, [P, C.c_int, P, P, C.c_int, C.POINTER(PdeParams), C.POINTER(C.c_double)]),
        "msbgx_flip_set_velocity": (C.c_int, [P, P, P, P]),
        "msbgx_flip_set_velocity_scene": (C.c_int, [P, C.POINTER(Scene)]),
        "msbgx_flip_get_velocity": (C.c_int, [P, P, P, P]),
        "msbgx_flip_rhs_from_velocity": (C.c_int, [P, C.c_float]),
        "msbgx_flip_update_velocity": (C.c_int, [P, C.c_float, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


lib = _load()

#: every symbol declared in include/msbgx.h
EXPORTED = [
    "msbgx_last_error", "msbgx_default_solve_params", "msbgx_create", "msbgx_destroy", "msbgx_num_blocks",
    "msbgx_set_options", "msbgx_set_domain_bc", "msbgx_regularize_refinement_map", "msbgx_set_refinement_map",
    "msbgx_prepare_pressure_solve", "msbgx_prepare_pressure_solve_dense", "msbgx_prepare_pressure_solve_scene",
    "msbgx_solve_pressure", "msbgx_residual_history", "msbgx_last_solve_ms", "msbgx_get_channel_block",
    "msbgx_set_channel_block", "msbgx_block_info", "msbgx_relax_list", "msbgx_omega", "msbgx_relax",
    "msbgx_multiply", "msbgx_downsample", "msbgx_prolongate", "msbgx_zero_channel", "msbgx_copy_channel",
    "msbgx_vcycle", "msbgx_time_kernel", "msbgx_level_map_from_scene", "msbgx_set_precision",
    "msbgx_solve_pressure_refined", "msbgx_level0_blocks", "msbgx_batch_create", "msbgx_batch_size",
    "msbgx_batch_block_base", "msbgx_batch_solve", "msbgx_batch_residual_history", "msbgx_autotune",
    "msbgx_set_launch_config", "msbgx_net_create", "msbgx_net_destroy", "msbgx_net_get_weights",
    "msbgx_net_set_weights", "msbgx_net_set_samples", "msbgx_net_loss_grad", "msbgx_net_collect",
    "msbgx_net_num_samples", "msbgx_net_train", "msbgx_net_last_train_ms", "msbgx_net_apply",
    "msbgx_solve_pressure_learned", "msbgx_net_time_conv", "msbgx_net_set_mode", "msbgx_net_get_mode",
    "msbgx_net_set_stall", "msbgx_net_set_prolongation", "msbgx_net_get_prolongation",
    "msbgx_net_set_coarse_scales", "msbgx_net_get_coarse_scales", "msbgx_flip_set_velocity", "msbgx_flip_set_velocity_scene", "msbgx_flip_get_velocity",
    "msbgx_flip_rhs_from_velocity", "msbgx_flip_update_velocity", "msbgx_default_p2g_params", "msbgx_p2g_create",
    "msbgx_p2g_destroy", "msbgx_p2g_run", "msbgx_p2g_run_scene", "msbgx_p2g_get", "msbgx_p2g_active_blocks",
    "msbgx_p2g_last_times", "msbgx_p2g_to_velocity",
    "msbgx_solve_pressure_compact", "msbgx_compact_mic0", "msbgx_pde_mean_curvature",
    "msbgx_pde_mean_curvature_noalloc", "msbgx_pde_apply",
]

OK, EINVAL, ENOMEM, EHIP, ESTATE = 0, -1, -2, -3, -4
FLIP_SIGN_REFERENCE, FLIP_SIGN_CORRECTED = 0, 1
NET_PARAMS = 14688
NET_FINE, NET_COARSE = 0, 1
PROLONG_CONSTANT, PROLONG_LINEAR = 0, 1


class MsbgxError(RuntimeError):
    def __init__(self, rc, where):
        super().__init__(f"{where} failed (rc={rc}): {lib.msbgx_last_error().decode()}")
        self.rc = rc


def _ck(rc, where):
    if rc < 0:
        raise MsbgxError(rc, where)
    return rc


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def level_map_from_scene(scene, bsx, levels, mode, d0):
    nb = (scene.nx // bsx) * (scene.ny // bsx) * (scene.nz // bsx)
    out = np.zeros(nb, np.int32)
    _ck(lib.msbgx_level_map_from_scene(C.byref(scene), bsx, levels, mode, d0, _ptr(out)), "level_map")
    return out


class MultiresSparseGrid:
    """Handle over ``msbgx_grid``; method names follow src/msbg.h."""

    def __init__(self, h):
        self._h = h

    @classmethod
    def create(cls, sx0, sy0, sz0, block_size0=16, n_levels=3):
        h = C.c_void_p()
        _ck(lib.msbgx_create(C.byref(h), sx0, sy0, sz0, block_size0, n_levels), "create")
        g = cls(h)
        g.shape = (sx0, sy0, sz0)
        g.bsx, g.n_levels = block_size0, n_levels
        g.n_blocks = lib.msbgx_num_blocks(h)
        return g

    def destroy(self):
        if self._h:
            lib.msbgx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def set_options(self, opts):
        _ck(lib.msbgx_set_options(self._h, opts), "set_options")

    def set_domain_boundary_spec(self, left, right, bottom, top, front, back):
        _ck(lib.msbgx_set_domain_bc(self._h, left, right, bottom, top, front, back), "set_domain_bc")

    def regularize_refinement_map(self, levels):
        a = np.ascontiguousarray(levels, np.int32)
        out = np.empty_like(a)
        _ck(lib.msbgx_regularize_refinement_map(self._h, _ptr(a), _ptr(out)), "regularize")
        return out

    def set_refinement_map(self, levels):
        a = np.ascontiguousarray(levels, np.int32)
        _ck(lib.msbgx_set_refinement_map(self._h, _ptr(a)), "set_refinement_map")

    def prepare_pressure_solve_flip(self, scene=None, callbacks=None, dense=None):
        if scene is not None:
            self._scene = scene
            return _ck(lib.msbgx_prepare_pressure_solve_scene(self._h, C.byref(scene)), "prepare(scene)")
        if callbacks is not None:
            self._cb = callbacks
            return _ck(lib.msbgx_prepare_pressure_solve(self._h, C.byref(callbacks)), "prepare(callbacks)")
        flags, wu, wv, ww, rhs = dense
        self._dense = [np.ascontiguousarray(flags, np.uint16)] + [np.ascontiguousarray(a, np.float32)
                                                                  for a in (wu, wv, ww, rhs)]
        return _ck(lib.msbgx_prepare_pressure_solve_dense(self._h, *[_ptr(a) for a in self._dense]), "prepare(dense)")

    def solve_pressure_flip(self, n_pre=3, n_post=3, n_coarse=24, max_iter=500, tol=5e-4):
        p = SolveParams(n_pre, n_post, n_coarse, max_iter, tol)
        it, rr = C.c_int(), C.c_float()
        _ck(lib.msbgx_solve_pressure(self._h, C.byref(p), C.byref(it), C.byref(rr)), "solve")
        return it.value, rr.value

    # FLIP caller steps either side of the solve (src/flip_sim.cpp:1281-1350);
    # dense arrays are indexed [z][y][x], component c on each cell's -c face.
    def flip_set_velocity(self, u, v, w):
        n = self.shape[0] * self.shape[1] * self.shape[2]
        arrs = [np.ascontiguousarray(a, np.float32) for a in (u, v, w)]
        if any(a.size != n for a in arrs):
            raise ValueError("velocity arrays must hold sx0*sy0*sz0 values")
        _ck(lib.msbgx_flip_set_velocity(self._h, *[_ptr(a) for a in arrs]), "flip_set_velocity")

    def flip_set_velocity_scene(self, scene):
        _ck(lib.msbgx_flip_set_velocity_scene(self._h, C.byref(scene)), "flip_set_velocity_scene")

    def flip_get_velocity(self):
        sx, sy, sz = self.shape
        out = [np.zeros((sz, sy, sx), np.float32) for _ in range(3)]
        _ck(lib.msbgx_flip_get_velocity(self._h, *[_ptr(a) for a in out]), "flip_get_velocity")
        return out

    def flip_rhs_from_velocity(self, dt):
        _ck(lib.msbgx_flip_rhs_from_velocity(self._h, dt), "flip_rhs_from_velocity")

    def flip_update_velocity(self, dt, corrected=True):
        """Pressure-gradient update with the pressure channel; `corrected`
        applies it to -p (the MSBG path solves A p = +div/dt)."""
        mode = FLIP_SIGN_CORRECTED if corrected else FLIP_SIGN_REFERENCE
        _ck(lib.msbgx_flip_update_velocity(self._h, dt, mode), "flip_update_velocity")

    def pde_mean_curvature(self, density, active_blocks, num_iter=5, time_step=0.05):
        """msbg_demo's mean-curvature smoothing (src/msbg4.cpp:60-960) of a
        uint16 [z][y][x] render-density array; returns (smoothed copy, device ms)."""
        d = np.ascontiguousarray(density, np.uint16).copy()
        a = np.ascontiguousarray(active_blocks, np.int32)
        ms = C.c_double()
        _ck(lib.msbgx_pde_mean_curvature(self._h, _ptr(d), _ptr(a), len(a), num_iter, time_step, C.byref(ms)), "pde")
        return d, ms.value

    def pde_apply(self, channel, active_blocks, lapl_typ, do_constr_zero_one=1, dens_thr=0.0, dens_thr_tol1=0.0,
                  dens_thr_tol2=0.0, num_iter=5, grid_spacing=1.0, time_step=0.025, taubin=0.0, anis_diff_alpha=0.0):
        """applyChannelPdeFast (src/msbg4.cpp:60-960) on a [z][y][x] uint16
        (render density) or float32 array, `lapl_typ` in the reference's
        encoding (-(type + PDE_OPT_8_COLOR): 8 colours, -type: red-black);
        returns (smoothed copy, device ms)."""
        if channel.dtype == np.uint16:
            d, dtype = np.ascontiguousarray(channel).copy(), PDE_U16
        else:
            d, dtype = np.ascontiguousarray(channel, np.float32).copy(), PDE_F32
        a = np.ascontiguousarray(active_blocks, np.int32)
        p = PdeParams(lapl_typ, do_constr_zero_one, dens_thr, dens_thr_tol1, dens_thr_tol2, num_iter, grid_spacing,
                      time_step, taubin, anis_diff_alpha)
        ms = C.c_double()
        _ck(lib.msbgx_pde_apply(self._h, dtype, _ptr(d), _ptr(a), len(a), C.byref(p), C.byref(ms)), "pde_apply")
        return d, ms.value

    def pde_mean_curvature_noalloc(self, density, active_blocks, num_iter=5, time_step=0.05):
        """The same smoothing through the allocation-free entry point
        (static per-GPU work space, no device allocation); returns the
        smoothed copy."""
        d = np.ascontiguousarray(density, np.uint16).copy()
        a = np.ascontiguousarray(active_blocks, np.int32)
        _ck(lib.msbgx_pde_mean_curvature_noalloc(self._h, _ptr(d), _ptr(a), len(a), num_iter, time_step),
            "pde_noalloc")
        return d

    def solve_pressure_compact(self, max_iter=1000, tol=1e-6):
        """Compact MIC(0)-PCG comparator (src/flip_sim.cpp:858-1151); the
        pressure channel receives its solution (compact sign)."""
        it, rr = C.c_int(), C.c_float()
        _ck(lib.msbgx_solve_pressure_compact(self._h, max_iter, tol, C.byref(it), C.byref(rr)), "solve_compact")
        return it.value, rr.value

    def compact_mic0(self):
        """(MIC(0) factor, z = M^-1 rhs) per compact row ((z, y, x) order over the fluid cells)."""
        n = _ck(lib.msbgx_compact_mic0(self._h, None, None, 0), "compact_mic0")
        p, z = np.zeros(max(n, 1), np.float32), np.zeros(max(n, 1), np.float32)
        _ck(lib.msbgx_compact_mic0(self._h, _ptr(p), _ptr(z), n), "compact_mic0")
        return p[:n], z[:n]

    def set_precision(self, prec):
        _ck(lib.msbgx_set_precision(self._h, prec), "set_precision")

    def level0_blocks(self):
        out = np.zeros(self.n_blocks, np.int32)
        n = lib.msbgx_level0_blocks(self._h, _ptr(out), self.n_blocks)
        return out[:n].copy()

    def solve_pressure_refined(self, tol=1e-10, max_outer=20, inner_tol=1e-6, n_pre=3, n_post=3, n_coarse=24,
                               inner_max_iter=500, want_x=False):
        """FP64 iterative refinement around the MG-PCG.  Returns (outer, inner
        iterations, FP64 relative residual, x64 per level-0 block or None)."""
        p = SolveParams(n_pre, n_post, n_coarse, inner_max_iter, inner_tol)
        o, i, r = C.c_int(), C.c_int(), C.c_double()
        blocks = self.level0_blocks()
        x = np.zeros(len(blocks) * self.bsx ** 3) if want_x else None
        _ck(lib.msbgx_solve_pressure_refined(self._h, C.byref(p), tol, max_outer, C.byref(o), C.byref(i),
                                             C.byref(r), _ptr(x) if want_x else None), "solve_refined")
        xb = None
        if want_x:
            xb = {int(b): x[k * self.bsx ** 3:(k + 1) * self.bsx ** 3] for k, b in enumerate(blocks)}
        return o.value, i.value, r.value, xb

    def residual_history(self):
        n = lib.msbgx_residual_history(self._h, None, 0)
        out = np.zeros(max(n, 1))
        lib.msbgx_residual_history(self._h, _ptr(out), n)
        return out[:n]

    def last_solve_ms(self):
        return lib.msbgx_last_solve_ms(self._h)

    def block_info(self, mg=0):
        lv = np.zeros(self.n_blocks, np.int32)
        fl = np.zeros(self.n_blocks, np.uint32)
        _ck(lib.msbgx_block_info(self._h, mg, _ptr(lv), _ptr(fl)), "block_info")
        return lv, fl

    def relax_list(self, mg=0):
        out = np.zeros(self.n_blocks, np.int32)
        n = lib.msbgx_relax_list(self._h, mg, _ptr(out), self.n_blocks)
        return out[:n].copy()

    def omega(self):
        return lib.msbgx_omega(self._h)

    def get_block(self, ch, mg, bid, edge):
        dt = np.uint16 if ch == CH_FLAGS else np.float32
        out = np.zeros(edge ** 3, dt)
        b = _ck(lib.msbgx_get_channel_block(self._h, ch, mg, bid, _ptr(out)), "get_block")
        return out if b else None

    def set_block(self, ch, mg, bid, data):
        a = np.ascontiguousarray(data, np.float32)
        _ck(lib.msbgx_set_channel_block(self._h, ch, mg, bid, _ptr(a)), "set_block")

    def relax(self, mg, reverse, n_iter, ch_x, ch_b):
        _ck(lib.msbgx_relax(self._h, mg, int(reverse), n_iter, ch_x, ch_b), "relax")

    def multiply(self, mg, residual, ch_x, ch_b, ch_dst):
        _ck(lib.msbgx_multiply(self._h, mg, int(residual), ch_x, ch_b, ch_dst), "multiply")

    def downsample(self, mg_coarse, ch_src, ch_dst):
        _ck(lib.msbgx_downsample(self._h, mg_coarse, ch_src, ch_dst), "downsample")

    def prolongate(self, mg_fine, ch):
        _ck(lib.msbgx_prolongate(self._h, mg_fine, ch), "prolongate")

    def zero_channel(self, ch, mg):
        _ck(lib.msbgx_zero_channel(self._h, ch, mg), "zero_channel")

    def copy_channel(self, src, dst):
        _ck(lib.msbgx_copy_channel(self._h, src, dst), "copy_channel")

    def time_kernel(self, which, reps=20):
        """(ms per launch, algorithmic bytes per launch, cells per launch)"""
        ms, by, ce = C.c_double(), C.c_double(), C.c_int()
        _ck(lib.msbgx_time_kernel(self._h, which, reps, C.byref(ms), C.byref(by), C.byref(ce)), "time_kernel")
        return ms.value, by.value, ce.value

    def vcycle(self, n_pre=3, n_post=3, n_coarse=24):
        _ck(lib.msbgx_vcycle(self._h, n_pre, n_post, n_coarse), "vcycle")

    def autotune(self, n_pre=3, n_post=3, n_coarse=24, max_iter=500, tol=1e-6):
        """Measure and keep the fastest graph chunk and smoother kernel per block
        size; returns (graph_chunk, smoother_mask).  The pressure is zero after."""
        p = SolveParams(n_pre, n_post, n_coarse, max_iter, tol)
        c, m = C.c_int(), C.c_int()
        _ck(lib.msbgx_autotune(self._h, C.byref(p), C.byref(c), C.byref(m)), "autotune")
        return c.value, m.value

    def set_launch_config(self, graph_chunk, smoother_mask):
        _ck(lib.msbgx_set_launch_config(self._h, graph_chunk, smoother_mask), "set_launch_config")


class LearnedNet:
    """Handle over ``msbgx_net``: the learned preconditioner correction."""

    def __init__(self, seed=1, mode=0):
        self._h = C.c_void_p()
        _ck(lib.msbgx_net_create(C.byref(self._h), seed), "net_create")
        if mode:
            self.mode = mode

    @property
    def mode(self):
        return lib.msbgx_net_get_mode(self._h)

    @mode.setter
    def mode(self, m):
        _ck(lib.msbgx_net_set_mode(self._h, m), "net_set_mode")

    @property
    def prolongation(self):
        """Coarse form: PROLONG_CONSTANT or PROLONG_LINEAR extension of block values."""
        return lib.msbgx_net_get_prolongation(self._h)

    @prolongation.setter
    def prolongation(self, k):
        _ck(lib.msbgx_net_set_prolongation(self._h, k), "net_set_prolongation")

    @property
    def scales(self):
        """Coarse form: number of scales (block grid coarsened 1x, 2x, 4x)."""
        return lib.msbgx_net_get_coarse_scales(self._h)

    @scales.setter
    def scales(self, k):
        _ck(lib.msbgx_net_set_coarse_scales(self._h, k), "net_set_coarse_scales")

    def destroy(self):
        if self._h:
            lib.msbgx_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def weights(self):
        w = np.zeros(NET_PARAMS, np.float32)
        _ck(lib.msbgx_net_get_weights(self._h, _ptr(w)), "net_get_weights")
        return w

    @weights.setter
    def weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == NET_PARAMS
        _ck(lib.msbgx_net_set_weights(self._h, _ptr(w)), "net_set_weights")

    def save(self, path):
        """Write the weights (the network's only persistent state), its form,
        prolongation and scales as an .npz file (plain arrays, no pickled
        objects)."""
        np.savez(path, weights=self.weights, mode=np.int32(self.mode), prolongation=np.int32(self.prolongation),
                 scales=np.int32(self.scales))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as f:
            net = cls(seed=1, mode=int(f["mode"]))
            net.weights = f["weights"]
            if "prolongation" in f:
                net.prolongation = int(f["prolongation"])
            if "scales" in f:
                net.scales = int(f["scales"])
        return net

    def set_samples(self, x, target, mask):
        self._samples = [np.ascontiguousarray(x, np.float32), np.ascontiguousarray(target, np.float32),
                         np.ascontiguousarray(mask, np.uint8)]
        n = self._samples[0].shape[0]
        _ck(lib.msbgx_net_set_samples(self._h, n, *[_ptr(a) for a in self._samples]), "net_set_samples")

    def loss_grad(self, want_out=False):
        n = lib.msbgx_net_num_samples(self._h)
        loss = C.c_double()
        g = np.zeros(NET_PARAMS, np.float32)
        out = np.zeros((n, 4096), np.float32) if want_out else None
        _ck(lib.msbgx_net_loss_grad(self._h, C.byref(loss), _ptr(g), _ptr(out) if want_out else None), "net_loss_grad")
        return loss.value, g, out

    def collect(self, grid, every=1, max_pairs=100, blocks_per_pair=64, seed=1, tol=1e-7, max_iter=2000,
                n_pre=3, n_post=3, n_coarse=24):
        p = SolveParams(n_pre, n_post, n_coarse, max_iter, tol)
        pairs = C.c_int()
        _ck(lib.msbgx_net_collect(self._h, grid._h, C.byref(p), every, max_pairs, blocks_per_pair, seed,
                                  C.byref(pairs)), "net_collect")
        return pairs.value

    def set_stall(self, window=5, drop=0.1):
        _ck(lib.msbgx_net_set_stall(self._h, window, drop), "net_set_stall")

    @property
    def num_samples(self):
        return lib.msbgx_net_num_samples(self._h)

    def train(self, steps, batch=256, lr=1e-3, seed=1):
        a, b = C.c_double(), C.c_double()
        _ck(lib.msbgx_net_train(self._h, steps, batch, lr, seed, C.byref(a), C.byref(b)), "net_train")
        return a.value, b.value

    def last_train_ms(self):
        return lib.msbgx_net_last_train_ms(self._h)

    def time_conv(self, blocks=1024, reps=20):
        """(ms per launch, algorithmic FLOP per launch) of the forward MFMA layer."""
        ms, fl = C.c_double(), C.c_double()
        _ck(lib.msbgx_net_time_conv(self._h, blocks, reps, C.byref(ms), C.byref(fl)), "net_time_conv")
        return ms.value, fl.value

    def apply(self, grid, ch_in, ch_out):
        _ck(lib.msbgx_net_apply(self._h, grid._h, ch_in, ch_out), "net_apply")

    def solve(self, grid, n_pre=3, n_post=3, n_coarse=24, max_iter=500, tol=1e-6):
        """MG-PCG with the learned correction; returns (iters, rel, fallback_iter)."""
        p = SolveParams(n_pre, n_post, n_coarse, max_iter, tol)
        it, rr, fb = C.c_int(), C.c_float(), C.c_int()
        _ck(lib.msbgx_solve_pressure_learned(grid._h, self._h, C.byref(p), C.byref(it), C.byref(rr), C.byref(fb)),
            "solve_learned")
        return it.value, rr.value, fb.value


class BatchGrid(MultiresSparseGrid):
    """Batched block set over several prepared grids (``msbgx_batch_*``):
    one launch per kernel covers every problem; per-problem PCG state."""

    @classmethod
    def create(cls, problems):
        arr = (C.c_void_p * len(problems))(*[g._h for g in problems])
        h = C.c_void_p()
        _ck(lib.msbgx_batch_create(C.byref(h), arr, len(problems)), "batch_create")
        b = cls(h)
        b.bsx, b.n_levels = problems[0].bsx, problems[0].n_levels
        b.n_blocks = lib.msbgx_num_blocks(h)
        b.n_problems = len(problems)
        b.bases = [lib.msbgx_batch_block_base(h, q) for q in range(len(problems))]
        return b

    def solve(self, n_pre=3, n_post=3, n_coarse=24, max_iter=500, tol=5e-4):
        p = SolveParams(n_pre, n_post, n_coarse, max_iter, tol)
        it = np.zeros(self.n_problems, np.int32)
        rr = np.zeros(self.n_problems, np.float32)
        _ck(lib.msbgx_batch_solve(self._h, C.byref(p), _ptr(it), _ptr(rr)), "batch_solve")
        return it, rr

    def problem_history(self, q):
        n = lib.msbgx_batch_residual_history(self._h, q, None, 0)
        out = np.zeros(max(n, 1))
        lib.msbgx_batch_residual_history(self._h, q, _ptr(out), n)
        return out[:n]


def build_grid(cfg: Config, walls=DBC_SOLID):
    """create + boundary spec + level map (regularised) + setRefinementMap."""
    g = MultiresSparseGrid.create(cfg.n, cfg.n, cfg.n, cfg.bsx, cfg.levels)
    g.set_domain_boundary_spec(*([walls] * 6))
    lv = level_map(cfg)
    g.set_refinement_map(g.regularize_refinement_map(lv))
    return g


def level_map(cfg: Config):
    nb = (cfg.n // cfg.bsx) ** 3
    if cfg.level_mode < 0:
        return np.zeros(nb, np.int32)
    return level_map_from_scene(cfg.scene, cfg.bsx, cfg.levels, cfg.level_mode, cfg.level_d0)


class ParticleToGrid:
    """Particle-to-grid transfer (FlipSimulation::particleToGrid,
    src/flip_sim.cpp:579-828) over a grid's raster and level map."""

    def __init__(self, grid):
        h = C.c_void_p()
        _ck(lib.msbgx_p2g_create(C.byref(h), grid._h), "p2g_create")
        self._h = h
        self.shape = grid.shape
        self.n_blocks = grid.n_blocks

    def __del__(self):
        if getattr(self, "_h", None):
            lib.msbgx_p2g_destroy(self._h)
            self._h = None

    @staticmethod
    def params(**kw):
        p = P2GParams()
        lib.msbgx_default_p2g_params(C.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def run(self, particles, **kw):
        a = np.ascontiguousarray(particles, PARTICLE_DTYPE)
        p = self.params(**kw)
        _ck(lib.msbgx_p2g_run(self._h, len(a), a.ctypes.data_as(C.c_void_p), C.byref(p)), "p2g_run")

    def run_scene(self, scene, **kw):
        p = self.params(**kw)
        n = C.c_int64()
        _ck(lib.msbgx_p2g_run_scene(self._h, C.byref(scene), C.byref(p), C.byref(n)), "p2g_run_scene")
        return n.value

    def get(self):
        """dict of dense [z][y][x] arrays: mass, u, v, w, beta_u, beta_v, beta_w"""
        sx, sy, sz = self.shape
        names = ["mass", "u", "v", "w", "beta_u", "beta_v", "beta_w"]
        out = {k: np.zeros((sz, sy, sx), np.float32) for k in names}
        _ck(lib.msbgx_p2g_get(self._h, *[_ptr(out[k]) for k in names]), "p2g_get")
        return out

    def active_blocks(self):
        out = np.zeros(self.n_blocks, np.int32)
        n = _ck(lib.msbgx_p2g_active_blocks(self._h, _ptr(out), self.n_blocks), "p2g_active_blocks")
        return out[:n].copy()

    def last_times(self):
        """(ms whole transfer, ms gather kernel, particles) of the last run"""
        a, b, n = C.c_double(), C.c_double(), C.c_int64()
        _ck(lib.msbgx_p2g_last_times(self._h, C.byref(a), C.byref(b), C.byref(n)), "p2g_last_times")
        return a.value, b.value, n.value

    def to_velocity(self, grid):
        _ck(lib.msbgx_p2g_to_velocity(self._h, grid._h), "p2g_to_velocity")
