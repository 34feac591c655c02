"""ctypes binding of the CPU oracle (oracle/libmsbg_oracle.so) -- test
infrastructure only; mirrors the entry points of msbgx.MultiresSparseGrid."""
import ctypes as C
import os

import numpy as np

from msbgx import Scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.environ.get("MSBG_ORACLE_SO") or os.path.join(ROOT, "oracle", "libmsbg_oracle.so")

_lib = C.CDLL(ORACLE_PATH)
P = C.c_void_p
for name, res, args in [
    ("orc_create", P, [C.c_int] * 5),
    ("orc_destroy", None, [P]),
    ("orc_set_threads", None, [P, C.c_int]),
    ("orc_num_blocks", C.c_int, [P]),
    ("orc_set_domain_bc", None, [P] + [C.c_int] * 6),
    ("orc_regularize_refinement_map", None, [P, P, P]),
    ("orc_set_refinement_map", None, [P, P]),
    ("orc_level_map_from_scene", None, [C.POINTER(Scene), C.c_int, C.c_int, C.c_int, C.c_int, P]),
    ("orc_prepare_scene", None, [P, C.POINTER(Scene)]),
    ("orc_prepare_dense", None, [P, C.c_int, C.c_int, C.c_int, P, P, P, P, P]),
    ("orc_omega", C.c_float, [P]),
    ("orc_set_smoother_precision", None, [P, C.c_int]),
    ("orc_round16", None, [C.c_int, P, P, C.c_int64]),
    ("orc_block_info", None, [P, C.c_int, P, P]),
    ("orc_relax_list", C.c_int, [P, C.c_int, P, C.c_int]),
    ("orc_get_block", C.c_int, [P, C.c_int, C.c_int, C.c_int, P]),
    ("orc_set_block", C.c_int, [P, C.c_int, C.c_int, C.c_int, P]),
    ("orc_relax", None, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("orc_relax_coarse64", None, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("orc_multiply", None, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("orc_downsample", None, [P, C.c_int, C.c_int, C.c_int]),
    ("orc_prolongate", None, [P, C.c_int, C.c_int]),
    ("orc_zero_channel", None, [P, C.c_int, C.c_int]),
    ("orc_vcycle", None, [P, C.c_int, C.c_int, C.c_int]),
    ("orc_copy_channel", None, [P, C.c_int, C.c_int]),
    ("orc_solve", C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float), P, C.c_int]),
    ("orc_export_level0", None, [P, C.c_int, P]),
    ("orc_solve_fp64", C.c_int, [P, C.c_double, C.c_int, P, C.POINTER(C.c_double)]),
    ("orc_flip_rhs", None, [C.c_int] * 4 + [P, P, C.c_float, P, P, P, C.c_float, P]),
    ("orc_scene_particles", C.c_int64, [C.POINTER(Scene), P]),
    ("orc_pde_mean_curvature", None, [C.c_int] * 4 + [P, P, P, C.c_int, C.c_int, C.c_float]),
    ("orc_pde_apply", C.c_int, [C.c_int] * 4 + [P, C.c_int, P, P, C.c_int, C.c_int, C.c_int] + [C.c_float] * 3 +
     [C.c_int] + [C.c_float] * 3),
    ("orc_compact_pcg", C.c_int, [C.c_int] * 4 + [P, P, C.c_float, P, P, P, P, C.c_float, C.c_int, P, P, P,
                                                   C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    ("orc_p2g", None, [C.c_int] * 4 + [P, C.c_int64, P] + [C.c_float] * 5 + [C.c_int] + [P] * 8),
    ("orc_flip_update_velocity", None, [C.c_int] * 4 + [P, P, C.c_float, P, P, P, P, P, P, P, C.c_float,
                                                        C.c_int]),
]:
    f = getattr(_lib, name)
    f.restype, f.argtypes = res, args


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def round16(values, prec):
    """The oracle's round-to-nearest-even to FP16 (prec 1) / BF16 (prec 2) and back."""
    a = np.ascontiguousarray(values, np.float32)
    out = np.empty_like(a)
    _lib.orc_round16(prec, _p(a), _p(out), a.size)
    return out


def level_map_from_scene(scene, bsx, levels, mode, d0):
    nb = (scene.nx // bsx) * (scene.ny // bsx) * (scene.nz // bsx)
    out = np.zeros(nb, np.int32)
    _lib.orc_level_map_from_scene(C.byref(scene), bsx, levels, mode, d0, _p(out))
    return out


class Oracle:
    def __init__(self, sx, sy, sz, bsx, levels, threads=None):
        self._h = _lib.orc_create(sx, sy, sz, bsx, levels)
        if not self._h:
            raise ValueError("oracle create failed")
        self._h = C.c_void_p(self._h)
        self.n_blocks = _lib.orc_num_blocks(self._h)
        self.bsx, self.n_levels = bsx, levels
        _lib.orc_set_threads(self._h, threads or min(16, os.cpu_count() or 1))

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.orc_destroy(self._h)
            self._h = None

    def set_domain_boundary_spec(self, *bc):
        _lib.orc_set_domain_bc(self._h, *bc)

    def regularize_refinement_map(self, levels):
        a = np.ascontiguousarray(levels, np.int32)
        out = np.empty_like(a)
        _lib.orc_regularize_refinement_map(self._h, _p(a), _p(out))
        return out

    def set_refinement_map(self, levels):
        a = np.ascontiguousarray(levels, np.int32)
        _lib.orc_set_refinement_map(self._h, _p(a))

    def prepare_scene(self, scene):
        self._scene = scene
        _lib.orc_prepare_scene(self._h, C.byref(scene))

    def prepare_dense(self, n, flags, wu, wv, ww, rhs):
        self._dense = [np.ascontiguousarray(flags, np.uint16)] + [np.ascontiguousarray(a, np.float32)
                                                                  for a in (wu, wv, ww, rhs)]
        _lib.orc_prepare_dense(self._h, n[0], n[1], n[2], *[_p(a) for a in self._dense])

    def omega(self):
        return _lib.orc_omega(self._h)

    def set_smoother_precision(self, prec):
        """Precision form: the relax reads its coefficients rounded to 16 bits
        (0 FP32, 1 FP16, 2 BF16)."""
        _lib.orc_set_smoother_precision(self._h, prec)

    def relax_coarse64(self, L, n_iter, ch_x, ch_b):
        """Precision form: the coarsest level's relax in FP64 from x = 0
        (x rounded to FP32 into ch_x on the relaxed blocks)."""
        _lib.orc_relax_coarse64(self._h, L, n_iter, ch_x, ch_b)

    def block_info(self, mg=0):
        lv = np.zeros(self.n_blocks, np.int32)
        fl = np.zeros(self.n_blocks, np.uint32)
        _lib.orc_block_info(self._h, mg, _p(lv), _p(fl))
        return lv, fl

    def relax_list(self, mg=0):
        out = np.zeros(self.n_blocks, np.int32)
        n = _lib.orc_relax_list(self._h, mg, _p(out), self.n_blocks)
        return out[:n].copy()

    def get_block(self, ch, mg, bid, edge):
        out = np.zeros(edge ** 3, np.uint16 if ch == 11 else np.float32)
        b = _lib.orc_get_block(self._h, ch, mg, bid, _p(out))
        return out if b else None

    def set_block(self, ch, mg, bid, data):
        a = np.ascontiguousarray(data, np.float32)
        _lib.orc_set_block(self._h, ch, mg, bid, _p(a))

    def relax(self, mg, reverse, n_iter, ch_x, ch_b):
        _lib.orc_relax(self._h, mg, int(reverse), n_iter, ch_x, ch_b)

    def multiply(self, mg, residual, ch_x, ch_b, ch_dst):
        _lib.orc_multiply(self._h, mg, int(residual), ch_x, ch_b, ch_dst)

    def downsample(self, mg_coarse, ch_src, ch_dst):
        _lib.orc_downsample(self._h, mg_coarse, ch_src, ch_dst)

    def prolongate(self, mg_fine, ch):
        _lib.orc_prolongate(self._h, mg_fine, ch)

    def zero_channel(self, ch, mg):
        _lib.orc_zero_channel(self._h, ch, mg)

    def copy_channel(self, src, dst):
        _lib.orc_copy_channel(self._h, src, dst)

    def vcycle(self, n_pre=3, n_post=3, n_coarse=24):
        _lib.orc_vcycle(self._h, n_pre, n_post, n_coarse)

    def solve(self, n_pre=3, n_post=3, n_coarse=24, max_iter=500, tol=5e-4):
        rr = C.c_float()
        hist = np.zeros(max_iter + 1)
        it = _lib.orc_solve(self._h, n_pre, n_post, n_coarse, max_iter, tol, C.byref(rr), _p(hist), max_iter + 1)
        return it, rr.value, hist

    def solve_fp64(self, tol=1e-12, max_iter=5000):
        """FP64 Jacobi-PCG solution of the level-0 system; returns (iters, true
        rel. residual, {bid: x block})."""
        n = self.bsx ** 3
        out = np.zeros(self.n_blocks * n)
        rel = C.c_double()
        it = _lib.orc_solve_fp64(self._h, tol, max_iter, _p(out), C.byref(rel))
        lv, _ = self.block_info(0)
        return it, rel.value, {b: out[b * n:(b + 1) * n] for b in range(self.n_blocks) if lv[b] == 0}

    def export_level0(self, ch, shape):
        out = np.zeros(shape[0] * shape[1] * shape[2], np.float32)
        _lib.orc_export_level0(self._h, ch, _p(out))
        return out.reshape(shape[2], shape[1], shape[0])


def build_oracle(cfg, walls=1, threads=None):
    o = Oracle(cfg.n, cfg.n, cfg.n, cfg.bsx, cfg.levels, threads)
    o.set_domain_boundary_spec(*([walls] * 6))
    nb = (cfg.n // cfg.bsx) ** 3
    if cfg.level_mode < 0:
        lv = np.zeros(nb, np.int32)
    else:
        lv = level_map_from_scene(cfg.scene, cfg.bsx, cfg.levels, cfg.level_mode, cfg.level_d0)
    o.set_refinement_map(o.regularize_refinement_map(lv))
    return o


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _levels(block_level, nx, ny, nz, bsx):
    lv = np.ascontiguousarray(block_level, np.int32)
    if lv.size != (nx // bsx) * (ny // bsx) * (nz // bsx):
        raise ValueError("one level per block of the raster expected")
    return lv


def flip_rhs(block_level, bsx, mass, mass_eps, vel, dt):
    """flipRhsCB on every cell with level-0 data (arrays [z][y][x])."""
    nz, ny, nx = mass.shape
    lv = _levels(block_level, nx, ny, nz, bsx)
    m, u, v, w = (_f32(a) for a in (mass, *vel))
    out = np.zeros_like(m)
    _lib.orc_flip_rhs(nx, ny, nz, bsx, _p(lv), _p(m), mass_eps, _p(u), _p(v), _p(w), dt, _p(out))
    return out


def flip_update_velocity(block_level, bsx, mass, mass_eps, p, beta, vel, dt, negate_p):
    """Pressure-gradient update of src/flip_sim.cpp:1317-1349; returns new (u, v, w)."""
    nz, ny, nx = mass.shape
    lv = _levels(block_level, nx, ny, nz, bsx)
    m, pp = _f32(mass), _f32(p)
    b = [_f32(a) for a in beta]
    out = [_f32(a).copy() for a in vel]
    _lib.orc_flip_update_velocity(nx, ny, nz, bsx, _p(lv), _p(m), mass_eps, _p(pp), *[_p(a) for a in b],
                                  *[_p(a) for a in out], dt, int(negate_p))
    return out


def p2g(block_level, nx, ny, nz, bsx, particles, rho_l=1000.0, rho_g=1.0, mass_eps=1e-5, beta_min=1e-6,
        alpha_phi=1.0, ppv=8):
    """particleToGrid (src/flip_sim.cpp:579-828) in the fixed cell-major order;
    returns (dict of dense [z][y][x] outputs, per-block active flags)."""
    from msbgx import PARTICLE_DTYPE
    a = np.ascontiguousarray(particles, PARTICLE_DTYPE)
    lv = _levels(block_level, nx, ny, nz, bsx)
    names = ["mass", "u", "v", "w", "beta_u", "beta_v", "beta_w"]
    out = {k: np.zeros((nz, ny, nx), np.float32) for k in names}
    act = np.zeros(len(lv), np.uint8)
    _lib.orc_p2g(nx, ny, nz, bsx, _p(lv), len(a), a.ctypes.data_as(C.c_void_p), rho_l, rho_g, mass_eps, beta_min,
                 alpha_phi, ppv, *[_p(out[k]) for k in names], _p(act))
    return out, act


def scene_particles(scene):
    """The scene's particles in the GPU scene path's order (host restatement)."""
    from msbgx import PARTICLE_DTYPE
    n = _lib.orc_scene_particles(C.byref(scene), None)
    out = np.zeros(n, PARTICLE_DTYPE)
    _lib.orc_scene_particles(C.byref(scene), out.ctypes.data_as(C.c_void_p))
    return out


def compact_pcg(block_level, bsx, mass, mass_eps, beta, div, tol=1e-6, max_iter=1000, want_factor=False):
    """Compact MIC(0)-PCG (src/flip_sim.cpp:858-1151): (x dense [z][y][x],
    iterations, |r|/|b|, precon rows or None, M^-1 rhs rows or None)."""
    nz, ny, nx = mass.shape
    lv = _levels(block_level, nx, ny, nz, bsx)
    m, d = _f32(mass), _f32(div)
    b = [_f32(a) for a in beta]
    x = np.zeros_like(m)
    nrows = int(np.count_nonzero(m > mass_eps))
    pre = np.zeros(max(nrows, 1), np.float32) if want_factor else None
    z = np.zeros(max(nrows, 1), np.float32) if want_factor else None
    it, rr = C.c_int(), C.c_float()
    n = _lib.orc_compact_pcg(nx, ny, nz, bsx, _p(lv), _p(m), mass_eps, *[_p(a) for a in b], _p(d), tol, max_iter,
                             _p(x), _p(pre) if want_factor else None, _p(z) if want_factor else None,
                             C.byref(it), C.byref(rr))
    if want_factor:
        pre, z = pre[:n], z[:n]
    return x, it.value, rr.value, pre, z


def pde_mean_curvature(block_level, bsx, density, active, num_iter=5, dt=0.05):
    """applyChannelPdeFast (mean curvature, 8 colours) on a uint16 [z][y][x] array; returns a copy."""
    nz, ny, nx = density.shape
    lv = _levels(block_level, nx, ny, nz, bsx)
    d = np.ascontiguousarray(density, np.uint16).copy()
    a = np.ascontiguousarray(active, np.int32)
    _lib.orc_pde_mean_curvature(nx, ny, nz, bsx, _p(lv), _p(d), _p(a), len(a), num_iter, dt)
    return d


def pde_apply(block_level, bsx, data, active, lapl_typ, do_constr_zero_one=1, dens_thr=0.0, dens_thr_tol1=0.0,
              dens_thr_tol2=0.0, num_iter=5, time_step=0.025, taubin=0.0, anis_diff_alpha=0.0):
    """applyChannelPdeFast on a uint16 or float32 [z][y][x] array; returns a
    copy, or None where the oracle declines the combination."""
    nz, ny, nx = data.shape
    lv = _levels(block_level, nx, ny, nz, bsx)
    if data.dtype == np.uint16:
        d, dtype = np.ascontiguousarray(data).copy(), 0
    else:
        d, dtype = np.ascontiguousarray(data, np.float32).copy(), 1
    a = np.ascontiguousarray(active, np.int32)
    rc = _lib.orc_pde_apply(nx, ny, nz, bsx, _p(lv), dtype, _p(d), _p(a), len(a), lapl_typ, do_constr_zero_one,
                            dens_thr, dens_thr_tol1, dens_thr_tol2, num_iter, time_step, taubin, anis_diff_alpha)
    return d if rc == 0 else None
