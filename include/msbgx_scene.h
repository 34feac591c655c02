/*
 * msbgx_scene.h -- seeded synthetic pressure-projection inputs.
 *
 * A scene is an analytic description of a liquid body inside a box domain
 * plus a seeded MAC velocity field.  It supplies the three per-cell inputs the
 * MSBG FLIP pressure solver samples through its callbacks
 * (reference: src/msbg.h:829-835, FlipPressureCallbacks; the FLIP callers
 * flipCellTypeCB / flipFaceCoeffCB / flipRhsCB in src/flip_sim.cpp:313-338):
 *
 *   cell type  : 0 (liquid) or CELL_VOID (air, Dirichlet p = 0)
 *   face coeff : 1, or 0 on the domain wall faces at coordinate 0 when
 *                `wall_faces_closed` is set (Neumann walls; the +side walls are
 *                closed by the domain boundary spec, src/msbg3.cpp:728-742)
 *   rhs        : divergence of the MAC velocity on liquid cells, 0 elsewhere; the
 *                MAC velocity is either a seeded uniform field per face or the
 *                particle-to-grid transfer of seeded particles, `particles_per_cell`
 *                of them in every liquid cell (P2G of src/flip_sim.cpp:659-690, 771-779)
 *
 * Every function is a pure function of (scene, integer coordinates), written
 * with IEEE double/float operations only and no contraction, so the host and
 * the device evaluate it bit-identically.  Coordinates are fine-grid cell
 * indices, exactly as the FLIP callbacks interpret them; the solver calls
 * them with level-local coordinates (src/msbg3.cpp:2494-2508).
 */
#ifndef MSBGX_SCENE_H
#define MSBGX_SCENE_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define MSBGX_HD __host__ __device__
#else
#define MSBGX_HD
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define MSBGX_SCENE_MAX_SPHERES 64
#define MSBGX_SCENE_MAX_BOXES 4

/* Cell flag bits shared with the reference (src/msbg.h:600-625). */
#define MSBGX_CELL_SOLID 0x0001u
#define MSBGX_CELL_FIXED 0x0002u
#define MSBGX_CELL_PARTIAL_SOLID 0x0040u
#define MSBGX_CELL_VOID 0x1000u
/* Value read from a missing flags block: CELL_VOID|EMPTY_U|V|W|C (src/msbg.h:623). */
#define MSBGX_CELL_EMPTY_VAL 0x9830u

typedef struct msbgx_scene {
  int32_t nx, ny, nz;          /* fine grid size in cells                       */
  int32_t n_spheres;           /* liquid spheres                                */
  int32_t n_boxes;             /* liquid boxes (half-open in unit coordinates)  */
  int32_t wall_faces_closed;   /* 1: face coefficient 0 on faces at coord 0     */
  uint64_t seed;               /* velocity seed                                 */
  float dt;                    /* time step of the divergence                   */
  float vel_scale;             /* velocities uniform in [-vel_scale, vel_scale) */
  double sphere[MSBGX_SCENE_MAX_SPHERES][4]; /* cx, cy, cz, r (unit cube)       */
  double box[MSBGX_SCENE_MAX_BOXES][6];      /* x0, x1, y0, y1, z0, z1          */
  int32_t particles_per_cell;  /* > 0: the MAC velocity is the particle-to-grid
                                  transfer of seeded particles (msbgx_scene_p2g) */
  float particle_sigma;        /* particle velocity standard deviation          */
} msbgx_scene;

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

static inline MSBGX_HD uint64_t msbgx_splitmix64(uint64_t x)
{
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Liquid indicator of fine cell (i,j,k).  Cell centres in half-cell units
 * are the exact integers 2i+1, so the only rounding is in the scaled centre. */
static inline MSBGX_HD int msbgx_scene_is_liquid(const msbgx_scene *s, int i, int j, int k)
{
  const double X = (double)(2 * i + 1), Y = (double)(2 * j + 1), Z = (double)(2 * k + 1);
  const double sx = 2.0 * (double)s->nx, sy = 2.0 * (double)s->ny, sz = 2.0 * (double)s->nz;
  for (int b = 0; b < s->n_boxes; b++) {
    const double *bb = s->box[b];
    if (X >= bb[0] * sx && X < bb[1] * sx && Y >= bb[2] * sy && Y < bb[3] * sy &&
        Z >= bb[4] * sz && Z < bb[5] * sz)
      return 1;
  }
  for (int q = 0; q < s->n_spheres; q++) {
    const double *sp = s->sphere[q];
    const double dx = X - sp[0] * sx;
    const double dy = Y - sp[1] * sy;
    const double dz = Z - sp[2] * sz;
    const double r = sp[3] * sx;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < r * r) return 1;
  }
  return 0;
}

/* MAC velocity component `dir` on the face with index (i,j,k): the face of
 * cell (i,j,k) on its -dir side.  Uniform in [-vel_scale, vel_scale), with a
 * 24-bit mantissa so the value is exact in FP32. */
static inline MSBGX_HD float msbgx_scene_velocity(const msbgx_scene *s, int dir, int i, int j, int k)
{
  const uint64_t fx = (uint64_t)(s->nx + 1), fy = (uint64_t)(s->ny + 1), fz = (uint64_t)(s->nz + 1);
  const uint64_t idx = (((uint64_t)dir * fz + (uint64_t)k) * fy + (uint64_t)j) * fx + (uint64_t)i;
  const uint64_t h = msbgx_splitmix64(s->seed * 0xD1B54A32D192ED03ull + idx);
  const int32_t m = (int32_t)(h >> 40) - (1 << 23);          /* [-2^23, 2^23) */
  return ((float)m * (1.0f / 8388608.0f)) * s->vel_scale;
}

/* Particle p of liquid cell (i,j,k): position uniform in the cell (21-bit
 * fractions), velocity components sigma * (U_1 + ... + U_12 - 6) with 16-bit
 * uniforms U (Irwin-Hall: mean 0, variance sigma^2, an approximate normal
 * that needs no transcendental function, so host and device agree bit for
 * bit). */
static inline MSBGX_HD void msbgx_scene_particle(const msbgx_scene *s, int i, int j, int k, int p, float pos[3],
                                                 float vel[3])
{
  const uint64_t cell = ((uint64_t)k * (uint64_t)s->ny + (uint64_t)j) * (uint64_t)s->nx + (uint64_t)i;
  const uint64_t base = (s->seed ^ 0x5851F42D4C957F2Dull) * 0x2545F4914F6CDD1Dull +
                        (cell * (uint64_t)s->particles_per_cell + (uint64_t)p) * 16ull;
  const uint64_t h = msbgx_splitmix64(base);
  const float f21 = 1.0f / 2097152.0f;
  pos[0] = (float)i + (float)(h & 0x1FFFFFu) * f21;
  pos[1] = (float)j + (float)((h >> 21) & 0x1FFFFFu) * f21;
  pos[2] = (float)k + (float)((h >> 42) & 0x1FFFFFu) * f21;
  for (int d = 0; d < 3; d++) {
    float acc = 0.0f;
    for (int w = 0; w < 3; w++) {
      const uint64_t g = msbgx_splitmix64(base + 1u + 3u * (uint64_t)d + (uint64_t)w);
      for (int q = 0; q < 4; q++) acc = acc + (float)((g >> (16 * q)) & 0xFFFFu) * (1.0f / 65536.0f);
    }
    vel[d] = (acc - 6.0f) * s->particle_sigma;
  }
}

/* Particle-to-grid MAC velocity of component `dir` at face index (i,j,k)
 * (the -dir face of cell (i,j,k)): the trilinear-weighted mean of the
 * particle velocities, as the reference's scatter computes it (node offsets
 * (0,.5,.5) / (.5,0,.5) / (.5,.5,0), weight ((1-f or f) * ...) * ..., zero
 * when the weight sum is not above 1e-5; faces at index n_dir lie outside
 * the reference's velocity grid and read 0).  Written as a gather over the 18
 * cells whose particles reach the face, in a fixed order. */
static inline MSBGX_HD float msbgx_scene_p2g(const msbgx_scene *s, int dir, int i, int j, int k)
{
  const int n[3] = {s->nx, s->ny, s->nz};
  const int f[3] = {i, j, k};
  if (f[dir] >= n[dir]) return 0.0f;
  float sv = 0.0f, sw = 0.0f;
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = f[a] - 1;
    hi[a] = a == dir ? f[a] : f[a] + 1;
  }
  for (int cz = lo[2]; cz <= hi[2]; cz++)
    for (int cy = lo[1]; cy <= hi[1]; cy++)
      for (int cx = lo[0]; cx <= hi[0]; cx++) {
        if (cx < 0 || cy < 0 || cz < 0 || cx >= s->nx || cy >= s->ny || cz >= s->nz) continue;
        if (!msbgx_scene_is_liquid(s, cx, cy, cz)) continue;
        for (int p = 0; p < s->particles_per_cell; p++) {
          float pos[3], vel[3];
          msbgx_scene_particle(s, cx, cy, cz, p, pos, vel);
          float g[3], fr[3];
          int i0[3], dd[3], ok = 1;
          for (int a = 0; a < 3; a++) {
            g[a] = a == dir ? pos[a] : pos[a] - 0.5f;
            const float fl = floorf(g[a]);
            i0[a] = (int)fl;
            fr[a] = g[a] - (float)i0[a];
            dd[a] = f[a] - i0[a];
            if (dd[a] < 0 || dd[a] > 1) ok = 0;
          }
          if (!ok) continue;
          const float w = ((dd[0] ? fr[0] : 1.0f - fr[0]) * (dd[1] ? fr[1] : 1.0f - fr[1])) *
                          (dd[2] ? fr[2] : 1.0f - fr[2]);
          sv = sv + vel[dir] * w;
          sw = sw + w;
        }
      }
  return sw > 1e-5f ? sv / sw : 0.0f;
}

/* MAC velocity used by the right-hand side. */
static inline MSBGX_HD float msbgx_scene_mac(const msbgx_scene *s, int dir, int i, int j, int k)
{
  return s->particles_per_cell > 0 ? msbgx_scene_p2g(s, dir, i, j, k) : msbgx_scene_velocity(s, dir, i, j, k);
}

/* Cell-type callback: 0 for liquid, CELL_VOID for air. */
static inline MSBGX_HD uint16_t msbgx_scene_cell_type(const msbgx_scene *s, int x, int y, int z)
{
  return msbgx_scene_is_liquid(s, x, y, z) ? (uint16_t)0 : (uint16_t)MSBGX_CELL_VOID;
}

/* Face-coefficient callback (beta on the -dir face of cell (x,y,z)). */
static inline MSBGX_HD float msbgx_scene_face_coeff(const msbgx_scene *s, int dir, int x, int y, int z)
{
  const int c = dir == 0 ? x : (dir == 1 ? y : z);
  return (s->wall_faces_closed && c == 0) ? 0.0f : 1.0f;
}

/* Right-hand-side callback: the MAC divergence divided by dt on liquid cells,
 * with the operation order of flipRhsCB (src/flip_sim.cpp:329-338). */
static inline MSBGX_HD float msbgx_scene_rhs(const msbgx_scene *s, int x, int y, int z)
{
  if (!msbgx_scene_is_liquid(s, x, y, z)) return 0.0f;
  float d = msbgx_scene_mac(s, 0, x + 1, y, z) - msbgx_scene_mac(s, 0, x, y, z);
  d = d + msbgx_scene_mac(s, 1, x, y + 1, z);
  d = d - msbgx_scene_mac(s, 1, x, y, z);
  d = d + msbgx_scene_mac(s, 2, x, y, z + 1);
  d = d - msbgx_scene_mac(s, 2, x, y, z);
  return d / s->dt;
}

#ifdef __cplusplus
}
#endif

#endif /* MSBGX_SCENE_H */
