/*
 * solver.hip -- CDNA4 (gfx950) kernels and device orchestration of the MSBG
 * FLIP pressure projection: block-halo staging, the in-place block
 * Gauss-Seidel smoother, the stencil apply / residual, restriction,
 * prolongation, the V-cycle and the MG-preconditioned CG.
 *
 * Kernel map (one workgroup per sparse block unless noted):
 *   k_relax_col<B>  red-black block relaxation pass (the default): thread
 *                   (x, y) keeps its z-column and coefficients in registers;
 *                   5 local sweeps, the first four as hyperplane wavefronts,
 *                   the last a Jacobi pass; for B <= 16 the block and its
 *                   x/y halo ring stay in LDS (src/msbg3.cpp:419-1506,
 *                   2171-2240)
 *   k_relax<B>      the same pass with the whole block tile in LDS (an
 *                   autotuner alternative, FP32 operands)
 *   k_multiply<B>   y = A x or r = b - A x, optional FP64 partial of <x, y>;
 *                   column form for B <= 16 (src/msbg3.cpp:1282-1321,
 *                   1757-1912)
 *   k_restrict      2x2x2 mean / block copy / zero (src/msbg4.cpp:1513-1915)
 *   k_prolong       piecewise-constant copy-add (src/msbg4.cpp:1921-2029)
 *   k_update_xr, k_dot, k_axpy, k_finalize_*: PCG vector work with FP64
 *                   accumulation (src/msbg3.cpp:3098-3259)
 *
 * Arithmetic follows the oracle's operation sequence exactly (explicit fmaf,
 * FP32 storage, FP64 accumulation), and the file is compiled with
 * -ffp-contract=off, so each kernel reproduces the CPU restatement bit for
 * bit; only the FP64 reduction order differs.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>

#include "msbgx_internal.h"

namespace msbgx {

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      setError(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr);   \
      return MSBGX_EHIP;                                                               \
    }                                                                                  \
  } while (0)

/* ------------------------------------------------------------------ device */

struct LevelArgs {
  const int *level;
  const int64_t *off;
  const uint16_t *flags;
  const float *W0, *W1, *W2, *D;
  int nbx, nby, nbz, bsx0;
  float bcW[3];  // face coefficient at the +side domain boundary per direction
  /* precision form: 16-bit copies of W_x, W_y, W_z, D read by the smoother */
  const uint16_t *H[4];
  int prec;      // MSBGX_PREC_FP32 / FP16 / BF16
  /* batched block sets (several independent problems in one block set) */
  const NbDesc *nbd;     // per block: face neighbours' offsets and levels at this MG level
  const int *probOf;     // problem of each block
  const float *omegaOf;  // smoother weight of each problem
  /* MG0 with 32^3 level-0 blocks: per-cell coefficient records of the column
   * smoother, two float4 per cell {D, W-x, W+x, W-y}, {W+y, W-z, W+z, 0}, at
   * record off[bid] + slot * 1024 + (x + 32 y), slot = (z + x/8 + y) mod 32 */
  const float4 *pack;
  /* the same for 16-bit operands (precision form): one uint2 per cell with the
   * 16-bit encodings {D | W-x << 16, W-z | W+x << 16} at record
   * off[bid] + slot * 1024 + (x + 32 y), and per z-column (x, y) at
   * (off[bid] >> 5) + x + 32 y one uint2 {bit s: the column's cell at slot s
   * lies on an all-non-fluid SIMD segment, W+z of its top cell} */
  const uint2 *pack16;
  const uint2 *pack16col;
  /* {W-y | W+y << 16} of each cell again, alone, in the same slot-plane order
   * (read once per local sweep by the 32^3 16-bit smoother) */
  const uint32_t *pack16y;
};

__device__ __forceinline__ float fromBits16(uint16_t v, int prec)
{
  if (prec == MSBGX_PREC_FP16) return (float)__builtin_bit_cast(_Float16, v);
  return __uint_as_float((uint32_t)v << 16);
}
/* BF16 bits of v, rounded to nearest even; NaN stays NaN */
__device__ __forceinline__ uint16_t bf16Bits(float v)
{
  const uint32_t b = __float_as_uint(v);
  if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);
  return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}
/* element i of a vector stored in FP32 or as BF16 bits */
__device__ __forceinline__ float vload(const float *p, int64_t i) { return p[i]; }
__device__ __forceinline__ float vload(const uint16_t *p, int64_t i) { return __uint_as_float((uint32_t)p[i] << 16); }
__device__ __forceinline__ void vstore(float *p, int64_t i, float v) { p[i] = v; }
__device__ __forceinline__ void vstore(uint16_t *p, int64_t i, float v) { p[i] = bf16Bits(v); }
/* smoother coefficient k with the operand precision P fixed at compile time */
template <int P> __device__ __forceinline__ float smCoefP(const LevelArgs &a, int k, int64_t q)
{
  if constexpr (P == MSBGX_PREC_FP32) {
    const float *p = k == 0 ? a.W0 : (k == 1 ? a.W1 : (k == 2 ? a.W2 : a.D));
    return p[q];
  } else {
    return fromBits16(a.H[k][q], P);
  }
}
/* smoother coefficient k (0..2 = W_x, W_y, W_z; 3 = D) at cell q */
__device__ __forceinline__ float smCoef(const LevelArgs &a, int k, int64_t q)
{
  if (a.prec == MSBGX_PREC_FP32) {
    const float *p = k == 0 ? a.W0 : (k == 1 ? a.W1 : (k == 2 ? a.W2 : a.D));
    return p[q];
  }
  return fromBits16(a.H[k][q], a.prec);
}

constexpr int kMaxStallWindow = 32;

struct PcgState {
  double bnorm2, rho, rhoNew, dq, rr, tolSq;
  float alpha, beta;
  int done, conv, iter, brk, maxIter, pad;
  /* learned correction: active flag, restart request (beta = 0), a
   * non-finite network output seen, kernels of the correction skip while
   * skipNet is set, iteration of the fallback, the direction update takes
   * the plain V-cycle output (usePlain); r.z_old for the flexible beta and
   * r.V(r) for the fallback; |r|^2 of the last stallWin iterations */
  int learned, restart, bad, skipNet, fellBack, usePlain;
  int coarseMode, stallRestarts;   // coarse-correction form; stalls answered by a restart so far
  int stallWin, winBase;       // stall window (0: no stall test), at most kMaxStallWindow; its start

  double stallFactor;          // |r|^2 must fall below stallFactor * |r|^2 of stallWin iterations ago
  double rrBest;               // smallest |r|^2 before the stall window
  double zcr;                  // coarse form: <zc, r>
  float gamma, pad4;           // coarse form: Galerkin step <zc, r> / <zc, A zc> of the correction
  double rzOld, rzPlain;
  double rrWin[kMaxStallWindow];
  float coef[4];               // coarse form, several scales: Galerkin coefficients of the directions
};

template <int B> struct Cfg {
  static constexpr int S = B >= 8 ? 8 : B;     // SIMD segment width (src/msbg3.cpp:1700-1743)
  static constexpr int NSEG = B / S;
  static constexpr int T = NSEG + 2 * (B - 1);  // hyperplanes seg + y + z per sweep
  static constexpr int NT = B * B;
  static constexpr int NIN = B * B * B;
  static constexpr int NF = B * B;
};

enum FaceKind { FK_SAME = 0, FK_DOMAIN = 1, FK_COARSER = 2, FK_FINER = 3 };

/* Stage the block (h*x, zero on all-non-fluid segments) and its six halo
 * faces into LDS (src/msbg3.cpp:564-1044).  Returns per-thread skip mask over
 * z and the face kinds/neighbour offsets used for the +face coefficients. */
/* List index of this workgroup, XCD-aware: workgroups are dispatched to the
 * eight XCDs round robin (workgroup i on XCD i mod 8), so each XCD is given a
 * contiguous run of the block list instead of every eighth entry -- lists are
 * in block order, so spatial neighbours, whose halo faces a block reads, share
 * that XCD's L2.  A bijection on [0, gridDim.x). */
__device__ __forceinline__ int xcdListIndex()
{
  constexpr int NX = 8;
  const int i = blockIdx.x, n = gridDim.x, q = n / NX, r = n % NX, x = i % NX;
  return x * q + (x < r ? x : r) + i / NX;
}

/* True when every lane of the caller's SIMD segment (S consecutive lanes)
 * reports a SOLID/VOID cell: such segments are skipped by the reference's
 * block kernel (src/msbg3.cpp:589-617, 1201-1209). */
template <int S>
__device__ __forceinline__ bool segmentAllNonFluid(bool nonFluid)
{
  const uint64_t bal = __ballot(nonFluid);
  const uint64_t segBits = (bal >> (__lane_id() & ~(S - 1))) & ((1ull << S) - 1);
  return segBits == ((1ull << S) - 1);
}

/* Halo faces of a block into faces[6][B*B] (h*x of same-level neighbours,
 * mixed-resolution aggregates, 0 at the domain boundary) plus per-face kinds
 * and neighbour slot offsets (src/msbg3.cpp:665-1044, 46-161). */
template <int B>
__device__ __forceinline__ void stageFaces(float *faces, const LevelArgs &a, int bid, int lev, float h, int64_t off,
                                           const float *__restrict__ X, int *fkind, int64_t *foff,
                                           int tid = (int)threadIdx.x)
{
  constexpr int NF = Cfg<B>::NF;
  const int tx = tid % B, ty = tid / B;
  const float a13 = (float)(1.0 * (1. / 3.));      // DPMR_GRAD_COEFF_CF (src/msbg.h:85-87)
  const float a23 = (float)(1.0 * 2 * (1. / 3.));  // 2 * DPMR_GRAD_COEFF_CF
  const float *Wd[3] = {a.W0, a.W1, a.W2};
  const int u = tx, w = ty;
#pragma unroll
  for (int face = 0; face < 6; face++) {
    const int dir = face >> 1, side = face & 1;
    const int lev2 = a.nbd[bid].lev[face];
    float val = 0.0f;
    int kind;
    int64_t off2 = 0;
    if (lev2 < 0) {
      kind = FK_DOMAIN;
    } else {
      off2 = a.nbd[bid].off[face];
      if (lev2 == lev) {
        kind = FK_SAME;
        const int e = side ? 0 : B - 1;
        const int src = dir == 0 ? e + B * u + B * B * w : (dir == 1 ? u + B * e + B * B * w : u + B * w + B * B * e);
        val = X[off2 + src] * h;
      } else if (lev2 > lev) {
        /* fine block, coarser neighbour: one coarse cell, own -face coefficient */
        kind = FK_COARSER;
        constexpr int Bc = B / 2 > 0 ? B / 2 : 1;
        const int e = side ? 0 : Bc - 1;
        const int cu = u / 2, cw = w / 2;
        const int src = dir == 0 ? e + Bc * cu + Bc * Bc * cw
                                 : (dir == 1 ? cu + Bc * e + Bc * Bc * cw : cu + Bc * cw + Bc * Bc * e);
        float aa = a23;
        if (!side) {
          const int own = dir == 0 ? B * u + B * B * w : (dir == 1 ? u + B * B * w : u + B * w);
          aa *= Wd[dir][off + own];
        }
        const double sv = (double)(aa * X[off2 + src]);
        val = (float)(sv * (double)h);
      } else {
        /* coarse block, finer neighbour: four fine cells on the shared face */
        kind = FK_FINER;
        constexpr int Bf = 2 * B;
        const int e = side ? 0 : Bf - 1;
        double sv = 0.0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const int d1 = c & 1, d2 = c >> 1;
          const int fu = 2 * u + d1, fw = 2 * w + d2;
          const int src = dir == 0 ? e + Bf * fu + Bf * Bf * fw
                                   : (dir == 1 ? fu + Bf * e + Bf * Bf * fw : fu + Bf * fw + Bf * Bf * e);
          float aa = a13;
          if (side) aa *= Wd[dir][off2 + src];
          sv += (double)(aa * X[off2 + src]);
        }
        val = (float)(sv * (double)h);
      }
    }
    faces[face * NF + u + B * w] = val;
    fkind[face] = kind;
    foff[face] = off2;
  }
}

/* Stage the block interior (h*x, zero on all-non-fluid segments) and its six
 * halo faces into an LDS tile of B^3 + 6 B^2 floats; skipmask bit z marks the
 * caller's skipped segment at row z. */
template <int B>
__device__ __forceinline__ void stageTile(float *tile, uint32_t &skipmask, const LevelArgs &a, int bid, int lev,
                                          float h, int64_t off, const float *__restrict__ X, int *fkind,
                                          int64_t *foff)
{
  constexpr int S = Cfg<B>::S, NIN = Cfg<B>::NIN;
  const int tx = threadIdx.x % B, ty = threadIdx.x / B;
  uint32_t sm = 0;
#pragma unroll 4
  for (int z = 0; z < B; z++) {
    const int v = tx + B * ty + B * B * z;
    const bool skip = segmentAllNonFluid<S>((a.flags[off + v] & (CELL_SOLID | CELL_VOID)) != 0);
    sm |= (skip ? 1u : 0u) << z;
    tile[v] = skip ? 0.0f : X[off + v] * h;
  }
  skipmask = sm;
  stageFaces<B>(tile + NIN, a, bid, lev, h, off, X, fkind, foff);
}

/* One cell of the block operator: F0..F6 from the LDS tile, coefficients from
 * global memory with the staged face overrides (src/msbg3.cpp:1192-1335). */
template <int B>
__device__ __forceinline__ void cellStencil(const float *tile, const LevelArgs &a, int64_t off, const int *fkind,
                                            const int64_t *foff, int x, int y, int z, float &F0, float &S, float &D)
{
  constexpr int NIN = Cfg<B>::NIN, NF = Cfg<B>::NF;
  const int v = x + B * y + B * B * z;
  F0 = tile[v];
  const float F1 = x > 0 ? tile[v - 1] : tile[NIN + 0 * NF + y + B * z];
  const float F2 = x < B - 1 ? tile[v + 1] : tile[NIN + 1 * NF + y + B * z];
  const float F3 = y > 0 ? tile[v - B] : tile[NIN + 2 * NF + x + B * z];
  const float F4 = y < B - 1 ? tile[v + B] : tile[NIN + 3 * NF + x + B * z];
  const float F5 = z > 0 ? tile[v - B * B] : tile[NIN + 4 * NF + x + B * y];
  const float F6 = z < B - 1 ? tile[v + B * B] : tile[NIN + 5 * NF + x + B * y];
  const int64_t c = off + v;
  float W1 = a.W0[c], W3 = a.W1[c], W5 = a.W2[c];
  if (x == 0 && fkind[0] >= FK_COARSER) W1 = 1.0f;
  if (y == 0 && fkind[2] >= FK_COARSER) W3 = 1.0f;
  if (z == 0 && fkind[4] >= FK_COARSER) W5 = 1.0f;
  float W2, W4, W6;
  if (x < B - 1) W2 = a.W0[c + 1];
  else W2 = fkind[1] == FK_SAME ? a.W0[foff[1] + B * y + B * B * z] : (fkind[1] == FK_DOMAIN ? a.bcW[0] : 1.0f);
  if (y < B - 1) W4 = a.W1[c + B];
  else W4 = fkind[3] == FK_SAME ? a.W1[foff[3] + x + B * B * z] : (fkind[3] == FK_DOMAIN ? a.bcW[1] : 1.0f);
  if (z < B - 1) W6 = a.W2[c + B * B];
  else W6 = fkind[5] == FK_SAME ? a.W2[foff[5] + x + B * y] : (fkind[5] == FK_DOMAIN ? a.bcW[2] : 1.0f);
  float s = F1 * W1;
  s = fmaf(F2, W2, s);
  s = fmaf(F3, W3, s);
  s = fmaf(F4, W4, s);
  s = fmaf(F5, W5, s);
  s = fmaf(F6, W6, s);
  S = s;
  D = a.D[c];
}

/* Red-black block relaxation, one colour pass, one workgroup per block. */
template <int B>
__global__ void __launch_bounds__(Cfg<B>::NT) k_relax(LevelArgs a, const int *__restrict__ blocks,
                                                      float *__restrict__ X, const float *__restrict__ Bv,
                                                      float omega, int reverse, int nIter, const int *done)
{
  if (*done) return;
  constexpr int S = Cfg<B>::S, T = Cfg<B>::T;
  __shared__ float tile[Cfg<B>::NIN + 6 * Cfg<B>::NF];
  const int bid = blocks[xcdListIndex()];
  const int lev = a.level[bid];
  const int64_t off = a.off[bid];
  const float h = (float)(1 << lev);
  uint32_t skipmask;
  int fkind[6];
  int64_t foff[6];
  stageTile<B>(tile, skipmask, a, bid, lev, h, off, X, fkind, foff);
  __syncthreads();

  const int x = threadIdx.x % B, y = threadIdx.x / B;
  const int o = x / S + y;
  if (a.omegaOf) omega = a.omegaOf[a.probOf[bid]];
  const float om = omega, omm = 1.0f - omega;
  for (int k = 0; k < nIter - 1; k++) {
    const bool rev = (reverse != 0) == ((k & 1) != 0);
    for (int t = 0; t < T; t++) {
      const int hp = rev ? T - 1 - t : t;
      const int z = hp - o;
      if (z >= 0 && z < B && !((skipmask >> z) & 1u)) {
        float F0, Sm, D;
        cellStencil<B>(tile, a, off, fkind, foff, x, y, z, F0, Sm, D);
        const float bb = Bv[off + x + B * y + B * B * z];
        const float Y = D != 0.0f ? fmaf(om * (bb + Sm), D, omm * F0) : F0;
        tile[x + B * y + B * B * z] = Y;
      }
      __syncthreads();
    }
  }
  /* last local sweep: every neighbour read from the staged state, result /h
   * written to the channel (src/msbg3.cpp:1341-1349, 1436-1438) */
  const float HR = 1.0f / h;
  for (int z = 0; z < B; z++) {
    if ((skipmask >> z) & 1u) continue;
    float F0, Sm, D;
    cellStencil<B>(tile, a, off, fkind, foff, x, y, z, F0, Sm, D);
    const int64_t c = off + x + B * y + B * B * z;
    const float Y = D != 0.0f ? fmaf(om * (Bv[c] + Sm), D, omm * F0) : F0;
    X[c] = Y * HR;
  }
}

/* Precision form, coarsest MG level: the reference's coarsest "solve", nCoarse
 * relax iterations (src/msbg3.cpp:2423-2430), in FP64 -- x and the arithmetic
 * in double, b, W and D read as FP32 -- with k_relax's schedule, SIMD-segment
 * semantics and local sweeps.  At the coarsest MG level every block has the
 * coarsest level, so a halo face holds a same-level neighbour's h*x or, outside
 * the domain, zero, and a +face weight is the neighbour's (bcW outside). */
template <int B>
__global__ void __launch_bounds__(Cfg<B>::NT) k_relax64(LevelArgs a, const int *__restrict__ blocks,
                                                        double *__restrict__ X, const float *__restrict__ Bv,
                                                        float omega, int reverse, int nIter, const int *done)
{
  if (*done) return;
  constexpr int S = Cfg<B>::S, T = Cfg<B>::T, NIN = Cfg<B>::NIN, NF = Cfg<B>::NF;
  __shared__ double tile[NIN + 6 * NF];
  const int bid = blocks[xcdListIndex()];
  const int lev = a.level[bid];
  const int64_t off = a.off[bid];
  const double h = (double)(1 << lev);
  const int x = threadIdx.x % B, y = threadIdx.x / B;
  uint32_t skipmask = 0;
  for (int z = 0; z < B; z++) {
    const int v = x + B * y + B * B * z;
    const bool skip = segmentAllNonFluid<S>((a.flags[off + v] & (CELL_SOLID | CELL_VOID)) != 0);
    skipmask |= (skip ? 1u : 0u) << z;
    tile[v] = skip ? 0.0 : X[off + v] * h;
  }
  bool same[6];
  int64_t foff[6];
#pragma unroll
  for (int face = 0; face < 6; face++) {
    const int dir = face >> 1, side = face & 1;
    const int l2 = a.nbd[bid].lev[face];
    same[face] = l2 == lev;
    foff[face] = same[face] ? a.nbd[bid].off[face] : 0;
    const int e = side ? 0 : B - 1;
    const int src = dir == 0 ? e + B * x + B * B * y : (dir == 1 ? x + B * e + B * B * y : x + B * y + B * B * e);
    tile[NIN + face * NF + x + B * y] = same[face] ? X[foff[face] + src] * h : 0.0;
  }
  __syncthreads();
  if (a.omegaOf) omega = a.omegaOf[a.probOf[bid]];
  const double om = (double)omega, omm = 1.0 - om;
  const int o = x / S + y;
  auto stencil = [&](int z, double &F0, double &Sm, double &D) {
    const int v = x + B * y + B * B * z;
    F0 = tile[v];
    const double F1 = x > 0 ? tile[v - 1] : tile[NIN + 0 * NF + y + B * z];
    const double F2 = x < B - 1 ? tile[v + 1] : tile[NIN + 1 * NF + y + B * z];
    const double F3 = y > 0 ? tile[v - B] : tile[NIN + 2 * NF + x + B * z];
    const double F4 = y < B - 1 ? tile[v + B] : tile[NIN + 3 * NF + x + B * z];
    const double F5 = z > 0 ? tile[v - B * B] : tile[NIN + 4 * NF + x + B * y];
    const double F6 = z < B - 1 ? tile[v + B * B] : tile[NIN + 5 * NF + x + B * y];
    const int64_t c = off + v;
    const double W1 = a.W0[c], W3 = a.W1[c], W5 = a.W2[c];
    const double W2 = x < B - 1 ? a.W0[c + 1] : (same[1] ? a.W0[foff[1] + B * y + B * B * z] : a.bcW[0]);
    const double W4 = y < B - 1 ? a.W1[c + B] : (same[3] ? a.W1[foff[3] + x + B * B * z] : a.bcW[1]);
    const double W6 = z < B - 1 ? a.W2[c + B * B] : (same[5] ? a.W2[foff[5] + x + B * y] : a.bcW[2]);
    double sm = F1 * W1;
    sm = fma(F2, W2, sm);
    sm = fma(F3, W3, sm);
    sm = fma(F4, W4, sm);
    sm = fma(F5, W5, sm);
    sm = fma(F6, W6, sm);
    Sm = sm;
    D = a.D[c];
  };
  for (int k = 0; k < nIter - 1; k++) {
    const bool rev = (reverse != 0) == ((k & 1) != 0);
    for (int t = 0; t < T; t++) {
      const int hp = rev ? T - 1 - t : t;
      const int z = hp - o;
      if (z >= 0 && z < B && !((skipmask >> z) & 1u)) {
        double F0, Sm, D;
        stencil(z, F0, Sm, D);
        const double bb = Bv[off + x + B * y + B * B * z];
        tile[x + B * y + B * B * z] = D != 0.0 ? fma(om * (bb + Sm), D, omm * F0) : F0;
      }
      __syncthreads();
    }
  }
  const double HR = 1.0 / h;
  for (int z = 0; z < B; z++) {
    if ((skipmask >> z) & 1u) continue;
    double F0, Sm, D;
    stencil(z, F0, Sm, D);
    const int64_t c = off + x + B * y + B * B * z;
    const double Y = D != 0.0 ? fma(om * ((double)Bv[c] + Sm), D, omm * F0) : F0;
    X[c] = Y * HR;
  }
}

__global__ void k_fill64(double *__restrict__ x, int64_t n, const int *done)
{
  if (*done) return;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

__global__ void k_x64_to_f32(const double *__restrict__ x64, float *__restrict__ x, int64_t n, const int *done)
{
  if (*done) return;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = (float)x64[i];
}

/* The halo faces of stageFaces, into registers, for a block whose six face
 * neighbours all have its own level or lie outside the domain (every block
 * of a single-level grid, the interior of a level): the six neighbour levels
 * and offsets are read in one group and then the six face values in one
 * group (the block's NbDesc), instead of one dependent round trip per face.  Thread (u, w) =
 * (tx, ty) gets x of face cell u + B*w of each face; the face value stageFaces
 * writes is that times h for a same-level face and 0 at the domain boundary
 * (sameLevelFaceValue), applied only where the value is used so that the
 * loads stay in flight meanwhile.  Returns false, loading no face value,
 * when a neighbour has another level. */
template <int B>
__device__ __forceinline__ bool sameLevelFaces(const LevelArgs &a, int bid, int lev, float h, int64_t off,
                                               const float *__restrict__ X, float *fv, int *fkind, int64_t *foff,
                                               int tid = (int)threadIdx.x)
{
  const int u = tid % B, w = tid / B;
  const NbDesc &nd = a.nbd[bid];
  int lev2[6];
  int64_t off2[6];
#pragma unroll
  for (int face = 0; face < 6; face++) {
    lev2[face] = nd.lev[face];
    off2[face] = nd.off[face];
  }
  bool same = true;
#pragma unroll
  for (int face = 0; face < 6; face++) same = same && (lev2[face] < 0 || lev2[face] == lev);
  if (!same) return false;
#pragma unroll
  for (int face = 0; face < 6; face++) {
    const int dir = face >> 1, side = face & 1;
    const int e = side ? 0 : B - 1;
    const int src = dir == 0 ? e + B * u + B * B * w : (dir == 1 ? u + B * e + B * B * w : u + B * w + B * B * e);
    const bool in = lev2[face] >= 0;
    fv[face] = X[(in ? off2[face] : off) + src];   // own block stands in outside the domain
    fkind[face] = in ? FK_SAME : FK_DOMAIN;
    foff[face] = in ? off2[face] : 0;
  }
  return true;
}

__device__ __forceinline__ float sameLevelFaceValue(float x, int kind, float h)
{
  return kind == FK_SAME ? x * h : 0.0f;
}

/* Workgroup barrier that orders LDS only: outstanding global loads stay in
 * flight across it (all cross-thread exchange in these kernels is via LDS). */
__device__ __forceinline__ void ldsBarrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

/* Column-resident red-black block relaxation (one colour pass, one
 * workgroup of B*B threads per block).
 *
 * Thread (x, y) owns the z-column of its cell row.  The staged h*x values of
 * the column live in registers, and for B <= 16 so do all seven coefficients
 * (REGC); for B = 32 they are read each step from per-cell records laid
 * out by register slot (LevelArgs::pack), one contiguous plane per step.  Cell
 * z sits in register slot (z + o) mod B with o = x/S + y, so inside a loop
 * unrolled over one period of B hyperplanes the slot a thread touches at
 * hyperplane hp is the compile-time constant hp mod B.  The four sequential
 * local sweeps are hyperplane wavefronts over seg + y + z, forward and
 * reverse as in src/msbg3.cpp:1157-1175; the fifth is the Jacobi pass whose
 * result /h is stored (src/msbg3.cpp:1341-1349).  One barrier per step.
 * x/y neighbours come through LDS: for B <= 16 the whole block with its x/y
 * halo ring stays in LDS (Halo3) and each step writes back the cell it
 * updated; for B = 32 (no room for that) each thread publishes three values
 * per step -- old value at z, old value one step ahead, newest value one step
 * behind -- which is exactly what its x/y neighbours read at that step, in a
 * double-buffered LDS array. */
template <int B, bool REGC> struct ColRegs;
template <int B> struct ColRegs<B, true> {
  float xs[B], cd[B], cb[B], w1[B], w2[B], w3[B], w4[B], w5[B];
};
template <int B> struct ColRegs<B, false> {
  float xs[B];
};

struct ColCtx {
  const LevelArgs *a;
  const float *Bv;
  float *Xg;
  const float *faces;
  const float4 *pack;
  const float4 *packBase;   // the thread's record in plane 0 of its block
  int64_t off;
  int tx, ty, o, fk1, fk3, fk5, mixed0, mixed2, mixed4;
  int64_t fo1, fo3, fo5;
  float w6top, om, omm, HR;
  int h3;                   // B <= 16: the thread's cell at z = 0 in the halo array (Halo3)
  int h3o;                  // B <= 16: h3 - o * (B+2)^2, its cell at z = hp - o is h3o + hp * (B+2)^2
  float fz0, fz1;           // B <= 16: the thread's -z and +z halo face values
};

/* B <= 16: the block's current h*x values in LDS with the x and y halo faces
 * as a one-cell ring, (B+2)^2 * B floats, x fastest. */
template <int B> struct Halo3 {
  static constexpr int P = B + 2, P2 = P * P, N = P2 * B;
};

/* The seven coefficients of cell z of the thread's column; P < 0 reads the
 * operand precision at run time, P >= 0 fixes it (lets the compiler batch
 * the loads of the register-resident prologue). */
template <int B, int P = -1>
__device__ __forceinline__ void loadCoeffs(const ColCtx &c, int z, float &D, float &Bb, float &w1, float &w2, float &w3,
                                           float &w4, float &w5, float &w6)
{
  const LevelArgs &a = *c.a;
  auto co = [&](int k, int64_t q) { if constexpr (P < 0) return smCoef(a, k, q); else return smCoefP<P>(a, k, q); };
  const int64_t q = c.off + c.tx + B * c.ty + B * B * z;
  D = co(3, q);
  Bb = c.Bv[q];
  w1 = (c.tx == 0 && c.mixed0) ? 1.0f : co(0, q);
  w3 = (c.ty == 0 && c.mixed2) ? 1.0f : co(1, q);
  w5 = (z == 0 && c.mixed4) ? 1.0f : co(2, q);
  if (c.tx < B - 1) w2 = co(0, q + 1);
  else w2 = c.fk1 == FK_SAME ? co(0, c.fo1 + B * c.ty + B * B * z) : (c.fk1 == FK_DOMAIN ? a.bcW[0] : 1.0f);
  if (c.ty < B - 1) w4 = co(1, q + B);
  else w4 = c.fk3 == FK_SAME ? co(1, c.fo3 + c.tx + B * B * z) : (c.fk3 == FK_DOMAIN ? a.bcW[1] : 1.0f);
  w6 = z < B - 1 ? co(2, q + B * B) : c.w6top;
}

/* loadCoeffs without branches: every load is issued whatever the face kinds
 * (a cell of the own block stands in for a neighbour that is not read) and
 * the face rules are applied to the loaded values, so a run of these loads
 * stays in flight and each value is waited for only where it is first used. */
template <int B, int P>
__device__ __forceinline__ void loadCoeffsFlat(const ColCtx &c, int z, float &D, float &Bb, float &w1, float &w2,
                                               float &w3, float &w4, float &w5)
{
  const LevelArgs &a = *c.a;
  const int64_t q = c.off + c.tx + B * c.ty + B * B * z;
  const bool x1same = c.tx < B - 1 || c.fk1 == FK_SAME, y1same = c.ty < B - 1 || c.fk3 == FK_SAME;
  const int64_t q2 = c.tx < B - 1 ? q + 1 : (c.fk1 == FK_SAME ? c.fo1 + B * c.ty + B * B * z : q);
  const int64_t q4 = c.ty < B - 1 ? q + B : (c.fk3 == FK_SAME ? c.fo3 + c.tx + B * B * z : q);
  D = smCoefP<P>(a, 3, q);
  Bb = c.Bv[q];
  const float v1 = smCoefP<P>(a, 0, q), v2 = smCoefP<P>(a, 0, q2);
  const float v3 = smCoefP<P>(a, 1, q), v4 = smCoefP<P>(a, 1, q4);
  const float v5 = smCoefP<P>(a, 2, q);
  w1 = (c.tx == 0 && c.mixed0) ? 1.0f : v1;
  w3 = (c.ty == 0 && c.mixed2) ? 1.0f : v3;
  w5 = (z == 0 && c.mixed4) ? 1.0f : v5;
  w2 = x1same ? v2 : (c.fk1 == FK_DOMAIN ? a.bcW[0] : 1.0f);
  w4 = y1same ? v4 : (c.fk3 == FK_DOMAIN ? a.bcW[1] : 1.0f);
}

/* Register-resident prologue: h*x, skip bits and the coefficients of every
 * slot of the thread's column plus the +z face coefficient, with the operand
 * precision fixed.  Flags and x of all slots are loaded first, then the
 * coefficients, with no branch between the loads and nothing that waits for
 * one before the last is issued: the column costs about one memory round
 * trip, not one per slot. */
template <int B, int P>
__device__ __forceinline__ uint32_t loadColumn(ColRegs<B, true> &r, ColCtx &c, const float *__restrict__ X, float h)
{
  constexpr int S = Cfg<B>::S;
  const LevelArgs &a = *c.a;
  uint16_t fl[B];
#pragma unroll
  for (int sl = 0; sl < B; sl++) {
    const int z = (sl - c.o + 2 * B) & (B - 1);
    const int64_t q = c.off + c.tx + B * c.ty + B * B * z;
    fl[sl] = a.flags[q];
    r.xs[sl] = X[q];
  }
#pragma unroll
  for (int sl = 0; sl < B; sl++) {
    const int z = (sl - c.o + 2 * B) & (B - 1);
    loadCoeffsFlat<B, P>(c, z, r.cd[sl], r.cb[sl], r.w1[sl], r.w2[sl], r.w3[sl], r.w4[sl], r.w5[sl]);
  }
  const float w6v = smCoefP<P>(a, 2, (c.fk5 == FK_SAME ? c.fo5 : c.off) + c.tx + B * c.ty);
  c.w6top = c.fk5 == FK_SAME ? w6v : (c.fk5 == FK_DOMAIN ? a.bcW[2] : 1.0f);
  uint32_t skip = 0;
#pragma unroll
  for (int sl = 0; sl < B; sl++) {
    const bool sk = segmentAllNonFluid<S>((fl[sl] & (CELL_SOLID | CELL_VOID)) != 0);
    skip |= (sk ? 1u : 0u) << sl;
    r.xs[sl] = sk ? 0.0f : r.xs[sl] * h;
  }
  return skip;
}

template <int B, bool REGC, int DIR, bool UPDATE, int I>
__device__ __forceinline__ void colStep(ColRegs<B, REGC> &r, uint32_t skip, float (*pub)[Cfg<B>::NT], const ColCtx &c,
                                        int hp)
{
  constexpr int S = Cfg<B>::S, NF = Cfg<B>::NF, T = Cfg<B>::T;
  constexpr int R0 = (T - 1) % B;
  constexpr int sl = DIR > 0 ? I : ((R0 - I) % B + B) % B;
  constexpr int slA = ((sl + DIR) % B + B) % B;   // z + d
  constexpr int slC = ((sl - DIR) % B + B) % B;   // z - d
  constexpr int slM = ((sl - 1) % B + B) % B;
  constexpr int slP = ((sl + 1) % B + B) % B;
  const int tid = threadIdx.x, tx = c.tx, ty = c.ty;
  const int z = hp - c.o;
  const bool active = z >= 0 && z < B && !((skip >> sl) & 1u);
  float D = 0.f, Bb = 0.f, w1 = 0.f, w2 = 0.f, w3 = 0.f, w4 = 0.f, w5 = 0.f, w6 = 0.f;
  if constexpr (!REGC) {
    if (active) {
      /* coefficient records (B = 32): every thread of the workgroup reads
       * record plane sl, one contiguous stream per step */
      /* byte offset of plane sl from the thread's base record, made opaque so
       * the compiler forms it here instead of keeping 32 hoisted addresses */
      uint32_t o32 = (uint32_t)sl * (uint32_t)(Cfg<B>::NT * 2 * sizeof(float4));
      asm volatile("" : "+v"(o32));
      const float4 *rp = reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(c.packBase) + o32);
      const float4 r0 = rp[0], r1 = rp[1];
      D = r0.x; w1 = r0.y; w2 = r0.z; w3 = r0.w;
      w4 = r1.x; w5 = r1.y; w6 = r1.z;
      Bb = c.Bv[c.off + tx + B * ty + B * B * z];
    }
  }
  float *pA = pub[0], *pB = pub[1], *pC = pub[2];
  pA[tid] = r.xs[sl];
  pB[tid] = r.xs[slA];
  pC[tid] = r.xs[slC];
  ldsBarrier();
  if (active) {
    const int lane = tx % S;
    float F1, F2, F3, F4;
    if (tx == 0) F1 = c.faces[0 * NF + ty + B * z];
    else if (lane > 0) F1 = pA[tid - 1];
    else F1 = DIR > 0 ? pC[tid - 1] : pB[tid - 1];
    if (tx == B - 1) F2 = c.faces[1 * NF + ty + B * z];
    else if (lane < S - 1) F2 = pA[tid + 1];
    else F2 = DIR > 0 ? pB[tid + 1] : pC[tid + 1];
    if (ty == 0) F3 = c.faces[2 * NF + tx + B * z];
    else F3 = DIR > 0 ? pC[tid - B] : pB[tid - B];
    if (ty == B - 1) F4 = c.faces[3 * NF + tx + B * z];
    else F4 = DIR > 0 ? pB[tid + B] : pC[tid + B];
    const float F0 = r.xs[sl];
    const float F5 = z > 0 ? r.xs[slM] : c.faces[4 * NF + tx + B * ty];
    const float F6 = z < B - 1 ? r.xs[slP] : c.faces[5 * NF + tx + B * ty];
    if constexpr (REGC) {
      D = r.cd[sl]; Bb = r.cb[sl]; w1 = r.w1[sl]; w2 = r.w2[sl]; w3 = r.w3[sl]; w4 = r.w4[sl]; w5 = r.w5[sl];
      w6 = z < B - 1 ? r.w5[slP] : c.w6top;
    }
    float Sm = F1 * w1;
    Sm = fmaf(F2, w2, Sm);
    Sm = fmaf(F3, w3, Sm);
    Sm = fmaf(F4, w4, Sm);
    Sm = fmaf(F5, w5, Sm);
    Sm = fmaf(F6, w6, Sm);
    const float Y = D != 0.0f ? fmaf(c.om * (Bb + Sm), D, c.omm * F0) : F0;
    if (UPDATE) r.xs[sl] = Y;
    else c.Xg[c.off + tx + B * ty + B * B * z] = Y * c.HR;
  }
}

template <int B, bool REGC, int DIR, bool UPDATE, int I>
struct PeriodUnroll {
  __device__ __forceinline__ static void run(ColRegs<B, REGC> &r, uint32_t skip, float (*pub)[3][Cfg<B>::NT],
                                             const ColCtx &c, int base)
  {
    constexpr int T = Cfg<B>::T;
    const int step = base + I;
    if (step < T) {
      const int hp = DIR > 0 ? step : T - 1 - step;
      colStep<B, REGC, DIR, UPDATE, I>(r, skip, pub[I & 1], c, hp);
    }
    PeriodUnroll<B, REGC, DIR, UPDATE, I + 1>::run(r, skip, pub, c, base);
  }
};
template <int B, bool REGC, int DIR, bool UPDATE>
struct PeriodUnroll<B, REGC, DIR, UPDATE, B> {
  __device__ __forceinline__ static void run(ColRegs<B, REGC> &, uint32_t, float (*)[3][Cfg<B>::NT], const ColCtx &,
                                             int)
  {
  }
};

/* B <= 16: one hyperplane step over the halo array (Halo3).  Each active
 * thread reads the four in-plane neighbours of its cell at fixed offsets from
 * one address, updates the cell and writes the new value back; one barrier
 * per step.  The values read are the ones the wavefront order prescribes:
 * neighbours one hyperplane behind were written at the previous step, those
 * one ahead are not written yet, and the only neighbours on the same
 * hyperplane -- x -+ 1 inside the cell's SIMD segment, updated together as in
 * src/msbg3.cpp:1700-1743 -- are written by this step only after the writer
 * has read its own neighbours, which the old values it reads must precede
 * (the write depends on those reads, and both lanes are in one wave, whose
 * LDS operations execute in issue order). */
template <int B, int DIR, int I>
__device__ __forceinline__ void colStepH(ColRegs<B, true> &r, uint32_t skip, float *h3, const ColCtx &c, int hp)
{
  constexpr int T = Cfg<B>::T, P = Halo3<B>::P, P2 = Halo3<B>::P2;
  constexpr int R0 = (T - 1) % B;
  constexpr int sl = DIR > 0 ? I : ((R0 - I) % B + B) % B;
  constexpr int slM = ((sl - 1) % B + B) % B;
  constexpr int slP = ((sl + 1) % B + B) % B;
  const int z = hp - c.o;
  const bool active = z >= 0 && z < B && !((skip >> sl) & 1u);
  if (active) {
    float *p = h3 + c.h3o + hp * P2;   // cell z = hp - o; hp * P2 is uniform
    const float F1 = p[-1], F2 = p[1], F3 = p[-P], F4 = p[P];
    const float F0 = r.xs[sl];
    const float F5 = z > 0 ? r.xs[slM] : c.fz0;
    const float F6 = z < B - 1 ? r.xs[slP] : c.fz1;
    const float w6 = z < B - 1 ? r.w5[slP] : c.w6top;
    const float D = r.cd[sl];
    float Sm = F1 * r.w1[sl];
    Sm = fmaf(F2, r.w2[sl], Sm);
    Sm = fmaf(F3, r.w3[sl], Sm);
    Sm = fmaf(F4, r.w4[sl], Sm);
    Sm = fmaf(F5, r.w5[sl], Sm);
    Sm = fmaf(F6, w6, Sm);
    const float Y = D != 0.0f ? fmaf(c.om * (r.cb[sl] + Sm), D, c.omm * F0) : F0;
    r.xs[sl] = Y;
    *p = Y;
  }
  ldsBarrier();
}

template <int B, int DIR, int I>
struct PeriodUnrollH {
  __device__ __forceinline__ static void run(ColRegs<B, true> &r, uint32_t skip, float *h3, const ColCtx &c, int base)
  {
    constexpr int T = Cfg<B>::T;
    const int step = base + I;
    if (step < T) colStepH<B, DIR, I>(r, skip, h3, c, DIR > 0 ? step : T - 1 - step);
    PeriodUnrollH<B, DIR, I + 1>::run(r, skip, h3, c, base);
  }
};
template <int B, int DIR>
struct PeriodUnrollH<B, DIR, B> {
  __device__ __forceinline__ static void run(ColRegs<B, true> &, uint32_t, float *, const ColCtx &, int) {}
};

template <int B, int DIR>
__device__ __forceinline__ void colSweepH(ColRegs<B, true> &r, uint32_t skip, float *h3, const ColCtx &c)
{
  constexpr int T = Cfg<B>::T;
#pragma unroll 1
  for (int base = 0; base < T; base += B) PeriodUnrollH<B, DIR, 0>::run(r, skip, h3, c, base);
}

/* The fifth local sweep reads only the fourth sweep's values (a Jacobi
 * pass, src/msbg3.cpp:1341-1349), which the halo array holds once the fourth
 * sweep's last barrier has passed: no wavefront, each thread computes its B
 * cells independently and stores the result /h. */
template <int B>
__device__ __forceinline__ void colJacobiH(ColRegs<B, true> &r, uint32_t skip, const float *h3, const ColCtx &c)
{
  constexpr int P = Halo3<B>::P, P2 = Halo3<B>::P2;
#pragma unroll
  for (int sl = 0; sl < B; sl++) {
    if ((skip >> sl) & 1u) continue;
    const int z = (sl - c.o + 2 * B) & (B - 1);
    const int slM = (sl + B - 1) & (B - 1), slP = (sl + 1) & (B - 1);
    const float *p = h3 + c.h3 + z * P2;
    const float F0 = r.xs[sl];
    const float F1 = p[-1], F2 = p[1], F3 = p[-P], F4 = p[P];
    const float F5 = z > 0 ? r.xs[slM] : c.fz0;
    const float F6 = z < B - 1 ? r.xs[slP] : c.fz1;
    const float D = r.cd[sl], Bb = r.cb[sl];
    const float w6 = z < B - 1 ? r.w5[slP] : c.w6top;
    float Sm = F1 * r.w1[sl];
    Sm = fmaf(F2, r.w2[sl], Sm);
    Sm = fmaf(F3, r.w3[sl], Sm);
    Sm = fmaf(F4, r.w4[sl], Sm);
    Sm = fmaf(F5, r.w5[sl], Sm);
    Sm = fmaf(F6, w6, Sm);
    const float Y = D != 0.0f ? fmaf(c.om * (Bb + Sm), D, c.omm * F0) : F0;
    c.Xg[c.off + c.tx + B * c.ty + B * B * z] = Y * c.HR;
  }
}

/* B = 32: one local sweep of ceil(T/B) periods of B hyperplanes.  B is even,
 * so the double-buffer index I & 1 alternates across period boundaries too. */
template <int B, bool REGC, int DIR, bool UPDATE>
__device__ __forceinline__ void colSweep(ColRegs<B, REGC> &r, uint32_t skip, float (*pub)[3][Cfg<B>::NT],
                                         const ColCtx &c)
{
  constexpr int T = Cfg<B>::T;
#pragma unroll 1
  for (int base = 0; base < T; base += B) PeriodUnroll<B, REGC, DIR, UPDATE, 0>::run(r, skip, pub, c, base);
}

template <int B>
__global__ void __launch_bounds__(Cfg<B>::NT) k_relax_col(LevelArgs a, const int *__restrict__ blocks,
                                                          float *__restrict__ X, const float *__restrict__ Bv,
                                                          float omega, int reverse, const int *done)
{
  constexpr bool REGC = B <= 16;
  constexpr int S = Cfg<B>::S, NT = Cfg<B>::NT, NF = Cfg<B>::NF;
  constexpr int P = Halo3<B>::P, P2 = Halo3<B>::P2;
  __shared__ float faces[6 * NF];
  __shared__ float pub[REGC ? 1 : 2][3][REGC ? 1 : NT];   // B = 32: double-buffered step values
  __shared__ float h3[REGC ? Halo3<B>::N : 1];            // B <= 16: block values with x/y halo ring
  /* the block's descriptor is read before the done flag is tested (the empty
   * asm keeps the test after the loads), so that the flag's load joins the
   * descriptor's instead of preceding it */
  const int bid = blocks[xcdListIndex()];
  const int lev = a.level[bid];
  const int64_t off = a.off[bid];
  const int isDone = *done;
  asm volatile("" ::"s"(lev), "s"(isDone));
  if (isDone) return;
  const float h = (float)(1 << lev);
  int fkind[6];
  int64_t foff[6];
  float fv[6];            // B <= 16, same-level faces: x of this thread's face cells
  bool fast = false;
  if constexpr (REGC) fast = sameLevelFaces<B>(a, bid, lev, h, off, X, fv, fkind, foff);
  if (!fast) stageFaces<B>(faces, a, bid, lev, h, off, X, fkind, foff);

  ColCtx c;
  c.a = &a; c.Bv = Bv; c.Xg = X; c.faces = faces; c.off = off;
  c.pack = REGC ? nullptr : a.pack;
  c.packBase = REGC ? nullptr : a.pack + 2 * (off + (int64_t)threadIdx.x);
  c.tx = threadIdx.x % B; c.ty = threadIdx.x / B; c.o = c.tx / S + c.ty;
  c.fk1 = fkind[1]; c.fk3 = fkind[3]; c.fk5 = fkind[5];
  c.fo1 = foff[1]; c.fo3 = foff[3]; c.fo5 = foff[5];
  c.mixed0 = fkind[0] >= FK_COARSER; c.mixed2 = fkind[2] >= FK_COARSER; c.mixed4 = fkind[4] >= FK_COARSER;
  if (a.omegaOf) omega = a.omegaOf[a.probOf[bid]];
  c.om = omega; c.omm = 1.0f - omega; c.HR = 1.0f / h;

  ColRegs<B, REGC> r;
  uint32_t skip = 0;
  if constexpr (REGC) {
    if (a.prec == MSBGX_PREC_FP16) skip = loadColumn<B, MSBGX_PREC_FP16>(r, c, X, h);
    else if (a.prec == MSBGX_PREC_BF16) skip = loadColumn<B, MSBGX_PREC_BF16>(r, c, X, h);
    else skip = loadColumn<B, MSBGX_PREC_FP32>(r, c, X, h);
    c.h3 = (c.ty + 1) * P + c.tx + 1;
    c.h3o = c.h3 - c.o * P2;
#pragma unroll
    for (int sl = 0; sl < B; sl++) h3[c.h3 + ((sl - c.o + 2 * B) & (B - 1)) * P2] = r.xs[sl];
    if (fast) {
      /* x/y halo faces into the ring of the halo array: thread (u, w) = (tx, ty)
       * holds face value u + B*w of each face (u along the face, w = z) */
      const int u = c.tx, w = c.ty;
      h3[w * P2 + (u + 1) * P] = sameLevelFaceValue(fv[0], fkind[0], h);
      h3[w * P2 + (u + 1) * P + B + 1] = sameLevelFaceValue(fv[1], fkind[1], h);
      h3[w * P2 + u + 1] = sameLevelFaceValue(fv[2], fkind[2], h);
      h3[w * P2 + (B + 1) * P + u + 1] = sameLevelFaceValue(fv[3], fkind[3], h);
      c.fz0 = sameLevelFaceValue(fv[4], fkind[4], h);
      c.fz1 = sameLevelFaceValue(fv[5], fkind[5], h);
    }
  } else {
    c.w6top = fkind[5] == FK_SAME ? smCoef(a, 2, foff[5] + c.tx + B * c.ty)
                                  : (fkind[5] == FK_DOMAIN ? a.bcW[2] : 1.0f);
#pragma unroll
    for (int sl = 0; sl < B; sl++) {
      const int z = (sl - c.o + 2 * B) & (B - 1);
      const int64_t q = off + c.tx + B * c.ty + B * B * z;
      const bool sk = segmentAllNonFluid<S>((a.flags[q] & (CELL_SOLID | CELL_VOID)) != 0);
      skip |= (sk ? 1u : 0u) << sl;
      r.xs[sl] = sk ? 0.0f : X[q] * h;
    }
  }
  __syncthreads();
  if constexpr (REGC) {
    if (!fast) {
      const int u = c.tx, w = c.ty;
      h3[w * P2 + (u + 1) * P] = faces[0 * NF + u + B * w];
      h3[w * P2 + (u + 1) * P + B + 1] = faces[1 * NF + u + B * w];
      h3[w * P2 + u + 1] = faces[2 * NF + u + B * w];
      h3[w * P2 + (B + 1) * P + u + 1] = faces[3 * NF + u + B * w];
      c.fz0 = faces[4 * NF + c.tx + B * c.ty];
      c.fz1 = faces[5 * NF + c.tx + B * c.ty];
      __syncthreads();
    }
    /* sweep k is reversed iff reverse == (k & 1) (src/msbg3.cpp:1157) */
    if (reverse) {
#pragma unroll 1
      for (int pair = 0; pair < 2; pair++) {
        colSweepH<B, +1>(r, skip, h3, c);
        colSweepH<B, -1>(r, skip, h3, c);
      }
    } else {
#pragma unroll 1
      for (int pair = 0; pair < 2; pair++) {
        colSweepH<B, -1>(r, skip, h3, c);
        colSweepH<B, +1>(r, skip, h3, c);
      }
    }
    colJacobiH<B>(r, skip, h3, c);
  } else {
    if (reverse) {
#pragma unroll 1
      for (int pair = 0; pair < 2; pair++) {
        colSweep<B, REGC, +1, true>(r, skip, pub, c);
        colSweep<B, REGC, -1, true>(r, skip, pub, c);
      }
      colSweep<B, REGC, +1, false>(r, skip, pub, c);
    } else {
#pragma unroll 1
      for (int pair = 0; pair < 2; pair++) {
        colSweep<B, REGC, -1, true>(r, skip, pub, c);
        colSweep<B, REGC, +1, true>(r, skip, pub, c);
      }
      colSweep<B, REGC, -1, false>(r, skip, pub, c);
    }
  }
}

/* Coefficient records of the 32^3 column smoother (see LevelArgs::pack): the
 * seven values loadCoeffs returns for each cell, in the operand precision of
 * the level, stored at the record of the register slot the smoother visits
 * at that cell's hyperplane step. */
__global__ void __launch_bounds__(1024) k_pack_col32(LevelArgs a, const int *__restrict__ blocks,
                                                     const float *__restrict__ X, float4 *__restrict__ pack)
{
  constexpr int B = 32, S = Cfg<B>::S, NT = Cfg<B>::NT, NF = Cfg<B>::NF;
  __shared__ float faces[6 * NF];
  const int bid = blocks[blockIdx.x];
  const int lev = a.level[bid];
  const int64_t off = a.off[bid];
  int fkind[6];
  int64_t foff[6];
  stageFaces<B>(faces, a, bid, lev, 1.0f, off, X, fkind, foff);
  ColCtx c;
  c.a = &a; c.Bv = X; c.Xg = nullptr; c.faces = faces; c.off = off; c.pack = nullptr;
  c.tx = threadIdx.x % B; c.ty = threadIdx.x / B; c.o = c.tx / S + c.ty;
  c.fk1 = fkind[1]; c.fk3 = fkind[3]; c.fk5 = fkind[5];
  c.fo1 = foff[1]; c.fo3 = foff[3]; c.fo5 = foff[5];
  c.mixed0 = fkind[0] >= FK_COARSER; c.mixed2 = fkind[2] >= FK_COARSER; c.mixed4 = fkind[4] >= FK_COARSER;
  c.w6top = fkind[5] == FK_SAME ? smCoef(a, 2, foff[5] + c.tx + B * c.ty) : (fkind[5] == FK_DOMAIN ? a.bcW[2] : 1.0f);
  c.om = 0.f; c.omm = 0.f; c.HR = 1.f;
  for (int z = 0; z < B; z++) {
    float D, Bb, w1, w2, w3, w4, w5, w6;
    loadCoeffs<B>(c, z, D, Bb, w1, w2, w3, w4, w5, w6);
    const int sl = (z + c.o) & (B - 1);
    float4 *r = pack + 2 * (off + (int64_t)sl * NT + threadIdx.x);
    r[0] = make_float4(D, w1, w2, w3);
    r[1] = make_float4(w4, w5, w6, 0.0f);
  }
}

/* 16-bit encoding of a smoother operand that loadCoeffs returned in the
 * precision form: a value read from the 16-bit copies or one of the exact
 * constants 0 and 1, so the conversion back is exact. */
__device__ __forceinline__ uint32_t enc16(float v, int prec)
{
  if (prec == MSBGX_PREC_FP16) return __builtin_bit_cast(uint16_t, (_Float16)v);
  return __float_as_uint(v) >> 16;
}

/* 16-bit coefficient records of the register-resident 32^3 smoother
 * (LevelArgs::pack16): the values loadCoeffs returns for each cell, in the
 * level's 16-bit encoding, with the cell flags, at the record of the
 * register slot (z + x/8 + y) mod 32. */
__global__ void __launch_bounds__(1024) k_pack_col32h(LevelArgs a, const int *__restrict__ blocks,
                                                      const float *__restrict__ X, uint2 *__restrict__ pack,
                                                      uint2 *__restrict__ packcol, uint32_t *__restrict__ packy)
{
  constexpr int B = 32, S = Cfg<B>::S, NT = Cfg<B>::NT, NF = Cfg<B>::NF;
  __shared__ float faces[6 * NF];
  const int bid = blocks[blockIdx.x];
  const int lev = a.level[bid];
  const int64_t off = a.off[bid];
  int fkind[6];
  int64_t foff[6];
  stageFaces<B>(faces, a, bid, lev, 1.0f, off, X, fkind, foff);
  ColCtx c;
  c.a = &a; c.Bv = X; c.Xg = nullptr; c.faces = faces; c.off = off; c.pack = nullptr;
  c.tx = threadIdx.x % B; c.ty = threadIdx.x / B; c.o = c.tx / S + c.ty;
  c.fk1 = fkind[1]; c.fk3 = fkind[3]; c.fk5 = fkind[5];
  c.fo1 = foff[1]; c.fo3 = foff[3]; c.fo5 = foff[5];
  c.mixed0 = fkind[0] >= FK_COARSER; c.mixed2 = fkind[2] >= FK_COARSER; c.mixed4 = fkind[4] >= FK_COARSER;
  c.w6top = fkind[5] == FK_SAME ? smCoef(a, 2, foff[5] + c.tx + B * c.ty) : (fkind[5] == FK_DOMAIN ? a.bcW[2] : 1.0f);
  c.om = 0.f; c.omm = 0.f; c.HR = 1.f;
  const int p = a.prec;
  uint32_t skip = 0, top = 0;
  for (int z = 0; z < B; z++) {
    float D, Bb, w1, w2, w3, w4, w5, w6;
    loadCoeffs<B>(c, z, D, Bb, w1, w2, w3, w4, w5, w6);
    const int64_t q = off + c.tx + B * c.ty + B * B * z;
    const int sl = (z + c.o) & (B - 1);
    pack[off + (int64_t)sl * NT + threadIdx.x] =
        make_uint2(enc16(D, p) | (enc16(w1, p) << 16), enc16(w5, p) | (enc16(w2, p) << 16));
    packy[off + (int64_t)sl * NT + threadIdx.x] = enc16(w3, p) | (enc16(w4, p) << 16);
    /* the lanes of a SIMD segment share the slot at every z (x / 8 and y alike) */
    const bool sk = segmentAllNonFluid<S>((a.flags[q] & (CELL_SOLID | CELL_VOID)) != 0);
    skip |= (sk ? 1u : 0u) << sl;
    if (z == B - 1) top = enc16(w6, p);
  }
  packcol[(off >> 5) + threadIdx.x] = make_uint2(skip, top);
}

/* Register-resident smoother for 32^3 blocks with 16-bit operands (the
 * precision form; the same schedule and arithmetic as k_relax_col, so the
 * same bits).  The FP32 record kernel re-reads its 32-byte coefficient
 * record at every one of the five local sweeps.  Here a workgroup of 1024
 * threads takes the block, each thread its z-column (x, y) (or, built with
 * MSBGX_C32_COLUMNS=2, 512 threads with the columns (x, y) and (x, y + 16),
 * both at register slot hp mod 32 at hyperplane step hp).  A thread keeps
 * the coefficients of its 32 cells as packed 16-bit pairs in registers,
 * loaded once from LevelArgs::pack16 -- {D, W-x} per cell, W-z two cells
 * per register -- and b of the 32 cells; {W-y, W+y} of the step's cells
 * streams from LevelArgs::pack16y, one contiguous plane per step, loaded K
 * steps ahead (a ring of K registers).  W+x of a cell is W-x of its x+1
 * neighbour, which the next lane holds (a lane shuffle), except on the
 * block's +x face, whose 32^2 values sit in LDS.  The block's h*x values
 * with the x/y halo faces as a one-cell ring live in LDS (34^2 * 32 floats =
 * 145 KB of gfx950's 160 KB), so one block runs per CU: one column per
 * thread gives 16 waves (4 per SIMD, 128 registers each) to overlap each
 * step's LDS round trips, against 8 waves (2 per SIMD) with two columns,
 * which run their two cells' round trips one after the other (C5 colour
 * pass 2.74 ms against 3.57 ms).
 *
 * Registers are indexed by compile-time slot, so a sweep's 32-step period is
 * unrolled.  One unrolled body serves all four sweeps: a reverse sweep visits
 * slot (1 - s) mod 32 at its step s, so before it the register arrays are
 * mirrored (entry j <-> slot (1 - j) mod 32, an involution) and the same
 * code runs with the direction a run-time value.  That keeps the hot code
 * (about 35 KB) within the instruction cache; four unrolled copies did not
 * fit and every sweep streamed its code from L2. */
#ifndef MSBGX_C32_COLUMNS
#define MSBGX_C32_COLUMNS 1
#endif
struct C32 {
  /* NC z-columns per thread: (x, y + YS c) for c < NC */
  static constexpr int NC = MSBGX_C32_COLUMNS, YS = 32 / NC;
  static constexpr int B = 32, S = 8, NT = 1024 / NC, T = Cfg<32>::T, P = B + 2, P2 = P * P, N = P2 * B;
  static constexpr int K = 4;
};

struct C32Col {
  uint32_t c0[C32::B];       // {D, W-x} per register index (slot, or mirrored slot)
  uint32_t c5[C32::B / 2];   // {W-z of index 2k, W-z of index 2k+1}
  float bb[C32::B];          // b per register index
  uint32_t wq[C32::K];       // {W-y, W+y} of the next K steps
};

/* Buffer resource over one block's cells (gfx9 descriptor word 3): the block
 * base stays in scalar registers, each access needs one 32-bit offset. */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t blockRsrc(const void *base, int bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, bytes, 0x00020000);
}

struct C32Ctx {
  __amdgpu_buffer_rsrc_t rb, rx;   // b and x of the block
  __amdgpu_buffer_rsrc_t rw;       // {W-y, W+y} records of the block (LevelArgs::pack16y)
  const uint16_t *wxp;             // LDS: W+x of the block's +x face cells, [z][y]
  int tx, ty0;                     // x and the first column's y (the second: ty0 + 16)
  int o0;                          // first column's slot shift x/8 + y (the second: o0 + 16)
  int h30;                         // first column's halo-array index at z = 0
  uint32_t skip[C32::NC];          // skipped (all non-fluid) segments by slot
  float fz0[C32::NC], fz1[C32::NC], w6top[C32::NC];
  float om, omm, HR;
};

template <int PREC> __device__ __forceinline__ float lo16(uint32_t v)
{
  if constexpr (PREC == MSBGX_PREC_FP16) return (float)__builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu));
  else return __uint_as_float(v << 16);
}
template <int PREC> __device__ __forceinline__ float hi16(uint32_t v)
{
  if constexpr (PREC == MSBGX_PREC_FP16) return (float)__builtin_bit_cast(_Float16, (uint16_t)(v >> 16));
  else return __uint_as_float(v & 0xffff0000u);
}

/* Byte offset of cell (x, y, z) in a block's cell-ordered arrays. */
__device__ __forceinline__ uint32_t c32Cell(int x, int y, int z) { return (uint32_t)(x + 32 * y + 1024 * z) * 4u; }

/* {W-y, W+y} of column CI's cell at slot sl (uniform, taken mod 32): one
 * dword of the block's contiguous record plane sl. */
template <int CI> __device__ __forceinline__ uint32_t c32LoadWy(const C32Ctx &c, int sl)
{
  return __builtin_amdgcn_raw_buffer_load_b32(c.rw, c32Cell(c.tx, c.ty0 + C32::YS * CI, 0), ((uint32_t)sl & 31u) * 4096u,
                                              0);
}

/* W-z held at register index J (a half of c5[J / 2]). */
template <int PREC, int J> __device__ __forceinline__ float c32Wz(uint32_t k5)
{
  return (J & 1) ? hi16<PREC>(k5) : lo16<PREC>(k5);
}

/* Mirror a column's register arrays: entry j <-> entry (1 - j) mod 32. */
__device__ __forceinline__ void c32Mirror(C32Col &r)
{
  constexpr int B = C32::B;
#pragma unroll
  for (int j = 0; j < B; j++) {
    const int m = (B + 1 - j) % B;   // (1 - j) mod 32: pairs (0,1), (2,31), ..., (16,17)
    if (j < m) {
      uint32_t t = r.c0[j]; r.c0[j] = r.c0[m]; r.c0[m] = t;
      float u = r.bb[j]; r.bb[j] = r.bb[m]; r.bb[m] = u;
    }
  }
  /* W-z pairs: entry pair k holds mirrored slots of pair (16 - k) mod 16, halves swapped */
  uint32_t n5[B / 2];
#pragma unroll
  for (int k = 0; k < B / 2; k++) n5[k] = __builtin_amdgcn_alignbit(r.c5[(16 - k) & 15], r.c5[(16 - k) & 15], 16);
#pragma unroll
  for (int k = 0; k < B / 2; k++) r.c5[k] = n5[k];
}

/* Column CI's cell at hyperplane hp in a wavefront sweep of direction dir
 * (run time), at register index I of the (mirrored, for dir < 0) arrays.
 * The slot of the +z neighbour -- and of the next lane's cell when it starts
 * the next SIMD segment -- is at index I + 1 forward and I - 1 mirrored.
 * Call in uniform control flow (the shuffle reads other lanes). */
template <int PREC, int I, int CI>
__device__ __forceinline__ void c32Cell(C32Col &r, float *h3, const C32Ctx &c, int hp, bool fwd, int iring)
{
  constexpr int B = C32::B, P = C32::P, P2 = C32::P2, K = C32::K;
  constexpr int IP = (I + 1) % B, IM = (I + B - 1) % B;
  const int o = c.o0 + C32::YS * CI;
  const int z = hp - o;
  const int sl = hp & (B - 1);
  uint32_t sk = c.skip[CI];   // opaque copy: the slot test is made here, not hoisted as lane masks
  asm volatile("" : "+v"(sk));
  const bool active = z >= 0 && z < B && !((sk >> sl) & 1u);
  /* opaque copies: the conversions stay here instead of being hoisted out of
   * the sweep loop as loop-invariant floats */
  uint32_t k0 = r.c0[I], k0p = r.c0[IP], k0m = r.c0[IM], k5 = r.c5[I / 2], k5p = r.c5[IP / 2], k5m = r.c5[IM / 2];
  asm volatile("" : "+v"(k0), "+v"(k0p), "+v"(k0m), "+v"(k5), "+v"(k5p), "+v"(k5m));
  /* both neighbours read first, then selected: a select of two array
   * elements would become a load from a selected address (scratch memory) */
  const uint32_t k0n = fwd ? k0p : k0m, k5n = fwd ? k5p : k5m;
  const uint32_t send = (c.tx & 7) == 0 ? k0n : k0;
  const uint32_t got = (uint32_t)__shfl_down((int)send, 1, 64);
  const float Bb = r.bb[I];
  const uint32_t wy = r.wq[iring];
  if (active) {
    /* every LDS read of the cell issued at once, without a condition: the
     * +x face weight and the z neighbours at clamped planes, selected after */
    float *p = h3 + (c.h30 + C32::YS * CI * P) + z * P2;
    const uint16_t wxb = c.wxp[z * B + c.ty0 + C32::YS * CI];
    const float F0 = p[0], F1 = p[-1], F2 = p[1], F3 = p[-P], F4 = p[P];
    const float F5r = p[z > 0 ? -P2 : 0], F6r = p[z < B - 1 ? P2 : 0];
    const float D = lo16<PREC>(k0), w1 = hi16<PREC>(k0);
    const float w3 = lo16<PREC>(wy), w4 = hi16<PREC>(wy);
    const float w5 = c32Wz<PREC, I>(k5);
    const float w6n = fwd ? c32Wz<PREC, IP>(k5n) : c32Wz<PREC, IM>(k5n);
    const float w6 = z < B - 1 ? w6n : c.w6top[CI];
    const float w2 = c.tx == B - 1 ? fromBits16(wxb, PREC) : hi16<PREC>(got);
    const float F5 = z > 0 ? F5r : c.fz0[CI];
    const float F6 = z < B - 1 ? F6r : c.fz1[CI];
    float Sm = F1 * w1;
    Sm = fmaf(F2, w2, Sm);
    Sm = fmaf(F3, w3, Sm);
    Sm = fmaf(F4, w4, Sm);
    Sm = fmaf(F5, w5, Sm);
    Sm = fmaf(F6, w6, Sm);
    float Y = fmaf(c.om * (Bb + Sm), D, c.omm * F0);
    asm volatile("" : "+v"(Y));   // computed for every active lane: a select, not a branch, on D
    *p = D != 0.0f ? Y : F0;
  }
  /* refill the ring entry after its value's last use, so the load can target
   * the same register (a copy would wait for the load) */
  r.wq[iring] = c32LoadWy<CI>(c, hp + (fwd ? K : -K));
}

/* Steps base + I .. base + END - 1 of a sweep, unrolled, with no condition
 * between them (a guarded step would make the loop's back edge merge states
 * in which the ring was just refilled, and its loads would be waited for at
 * once).  hp is laundered through an empty asm so that the step's addresses
 * are formed at the step and not hoisted out of the unrolled period. */
template <int PREC, int I, int END> struct C32Period {
  __device__ __forceinline__ static void run(C32Col (&r)[C32::NC], float *h3, const C32Ctx &c, int base, bool fwd)
  {
    const int step = base + I;
    int hp = fwd ? step : C32::T - 1 - step;
    asm volatile("" : "+s"(hp));
    c32Cell<PREC, I, 0>(r[0], h3, c, hp, fwd, I % C32::K);
    if constexpr (C32::NC > 1) c32Cell<PREC, I, 1>(r[1], h3, c, hp, fwd, I % C32::K);
    ldsBarrier();
    C32Period<PREC, I + 1, END>::run(r, h3, c, base, fwd);
  }
};
template <int PREC, int END> struct C32Period<PREC, END, END> {
  __device__ __forceinline__ static void run(C32Col (&)[C32::NC], float *, const C32Ctx &, int, bool) {}
};

/* One local sweep (direction fwd, uniform; the arrays are mirrored for a
 * reverse sweep): {W-y, W+y} of the first K steps is loaded before the first
 * step, then two unrolled 32-step periods and the T - 64 = 2 remaining steps. */
template <int PREC>
__device__ __forceinline__ void c32Sweep(C32Col (&r)[C32::NC], float *h3, const C32Ctx &c, bool fwd)
{
  constexpr int T = C32::T, K = C32::K, B = C32::B, FULL = T / B * B;
#pragma unroll
  for (int s = 0; s < K; s++) {
    const int hp = fwd ? s : T - 1 - s;
    r[0].wq[s] = c32LoadWy<0>(c, hp);
    if constexpr (C32::NC > 1) r[1].wq[s] = c32LoadWy<1>(c, hp);
  }
#pragma unroll 1
  for (int base = 0; base < FULL; base += B) C32Period<PREC, 0, B>::run(r, h3, c, base, fwd);
  C32Period<PREC, 0, T - FULL>::run(r, h3, c, FULL, fwd);
}

/* The fifth local sweep reads only the fourth sweep's values (the halo array
 * after its last barrier): each thread computes its cells independently and
 * stores the result / h.  Register slot SL (forward order) holds cell
 * z = (SL - o) mod 32; its ring entry was loaded K slots earlier. */
template <int PREC, int SL, int CI>
__device__ __forceinline__ void c32JacobiCell(C32Col &r, const float *h3, const C32Ctx &c, int slv)
{
  constexpr int B = C32::B, P = C32::P, P2 = C32::P2, K = C32::K, SLP = (SL + 1) % B;
  const int o = c.o0 + C32::YS * CI;
  const int z = (slv - o + 4 * B) & (B - 1);
  uint32_t k0 = r.c0[SL], k0n = r.c0[SLP], k5 = r.c5[SL / 2], k5n = r.c5[SLP / 2];
  asm volatile("" : "+v"(k0), "+v"(k0n), "+v"(k5), "+v"(k5n));
  const uint32_t send = (c.tx & 7) == 0 ? k0n : k0;
  const uint32_t got = (uint32_t)__shfl_down((int)send, 1, 64);
  uint32_t sk = c.skip[CI];
  asm volatile("" : "+v"(sk));
  const uint32_t wy = r.wq[SL % K];
  if (!((sk >> SL) & 1u)) {
    const float D = lo16<PREC>(k0), w1 = hi16<PREC>(k0);
    const float w3 = lo16<PREC>(wy), w4 = hi16<PREC>(wy);
    const float w5 = c32Wz<PREC, SL>(k5);
    const float w6 = z < B - 1 ? c32Wz<PREC, SLP>(k5n) : c.w6top[CI];
    const float w2 = c.tx == B - 1 ? fromBits16(c.wxp[z * B + c.ty0 + C32::YS * CI], PREC) : hi16<PREC>(got);
    const float *p = h3 + (c.h30 + C32::YS * CI * P) + z * P2;
    const float F0 = p[0], F1 = p[-1], F2 = p[1], F3 = p[-P], F4 = p[P];
    const float F5 = z > 0 ? p[-P2] : c.fz0[CI];
    const float F6 = z < B - 1 ? p[P2] : c.fz1[CI];
    float Sm = F1 * w1;
    Sm = fmaf(F2, w2, Sm);
    Sm = fmaf(F3, w3, Sm);
    Sm = fmaf(F4, w4, Sm);
    Sm = fmaf(F5, w5, Sm);
    Sm = fmaf(F6, w6, Sm);
    const float Y = D != 0.0f ? fmaf(c.om * (r.bb[SL] + Sm), D, c.omm * F0) : F0;
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(Y * c.HR), c.rx, c32Cell(c.tx, c.ty0 + C32::YS * CI, z), 0, 0);
  }
  r.wq[SL % K] = c32LoadWy<CI>(c, slv + K);
}
template <int PREC, int SL> struct C32JacobiAll {
  __device__ __forceinline__ static void run(C32Col (&r)[C32::NC], const float *h3, const C32Ctx &c)
  {
    int slv = SL;   // laundered: z and the addresses are formed here
    asm volatile("" : "+s"(slv));
    c32JacobiCell<PREC, SL, 0>(r[0], h3, c, slv);
    if constexpr (C32::NC > 1) c32JacobiCell<PREC, SL, 1>(r[1], h3, c, slv);
    C32JacobiAll<PREC, SL + 1>::run(r, h3, c);
  }
};
template <int PREC> struct C32JacobiAll<PREC, C32::B> {
  __device__ __forceinline__ static void run(C32Col (&)[C32::NC], const float *, const C32Ctx &) {}
};

/* Records of register slots SL .. SL+7 of column CI into registers: one
 * group of loads in flight at a time keeps the prologue within the register
 * budget. */
template <int SL, int CI>
__device__ __forceinline__ void c32LoadRecords(C32Col &r, const __amdgpu_buffer_rsrc_t &rr, C32Ctx &c, uint16_t *wxp)
{
  constexpr int B = C32::B, NF = B * B, NR = 16;
  const int ty = c.ty0 + C32::YS * CI, o = c.o0 + C32::YS * CI;
  uint32_t vx[NR], vz[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(rr, (uint32_t)(c.tx + B * ty) * 8u, (uint32_t)(SL + k) * (NF * 8u), 0);
    vx[k] = v[0];
    vz[k] = v[1];
  }
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const int sl = SL + k;
    r.c0[sl] = vx[k];
    if (sl & 1) r.c5[sl / 2] |= vz[k] << 16;
    else r.c5[sl / 2] = vz[k] & 0xffffu;
    int slv = sl;   // laundered: z is formed here, not precomputed for every slot
    asm volatile("" : "+s"(slv));
    const int z = (slv - o + 4 * B) & (B - 1);
    if (c.tx == B - 1) wxp[z * B + ty] = (uint16_t)(vz[k] >> 16);
  }
  __builtin_amdgcn_sched_barrier(0);
}

/* x of column CI into the halo array (h*x, 0 on skipped segments) and b into
 * its registers, in groups of 8 slots whose loads are all issued before the
 * group's LDS writes (one memory round trip per group, not per slot). */
template <int CI>
__device__ __forceinline__ void c32StageX(C32Col &r, float *h3, const C32Ctx &c, float h)
{
  constexpr int B = C32::B, P = C32::P, P2 = C32::P2, G = 8;
  const int o = c.o0 + C32::YS * CI;
#pragma unroll
  for (int g = 0; g < B; g += G) {
    float xv[G];
    int zz[G];
#pragma unroll
    for (int k = 0; k < G; k++) {
      int slv = g + k;
      asm volatile("" : "+s"(slv));
      zz[k] = (slv - o + 4 * B) & (B - 1);
      const uint32_t q = c32Cell(c.tx, c.ty0 + C32::YS * CI, zz[k]);
      xv[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(c.rx, q, 0, 0));
      r.bb[g + k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(c.rb, q, 0, 0));
    }
#pragma unroll
    for (int k = 0; k < G; k++)
      h3[c.h30 + C32::YS * CI * P + zz[k] * P2] = ((c.skip[CI] >> (g + k)) & 1u) ? 0.0f : xv[k] * h;
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int PREC>
__global__ void __launch_bounds__(C32::NT) k_relax_c32h(LevelArgs a, const int *__restrict__ blocks,
                                                        float *__restrict__ X, const float *__restrict__ Bv,
                                                        float omega, int reverse, const int *done)
{
  constexpr int B = C32::B, S = C32::S, NF = B * B, P = C32::P, P2 = C32::P2, NT = C32::NT, NC = C32::NC;
  constexpr int YS = C32::YS;
  __shared__ float h3[C32::N];
  __shared__ uint16_t wxp[NF];
  const int bid = blocks[xcdListIndex()];
  const int lev = a.level[bid];
  const int64_t off = a.off[bid];
  const int isDone = *done;
  asm volatile("" ::"s"(lev), "s"(isDone));
  if (isDone) return;
  const float h = (float)(1 << lev);
  const int tx = threadIdx.x % B, ty = threadIdx.x / B;   // ty in [0, YS)
  /* halo face values of this thread's face cells (u, w) = (tx, ty + YS k) */
  int fkind[6];
  int64_t foff[6];
  float fv[NC][6];
  bool same = true;
#pragma unroll
  for (int k = 0; k < NC; k++)
    same = same && sameLevelFaces<B>(a, bid, lev, h, off, X, fv[k], fkind, foff, threadIdx.x + NT * k);
  if (same) {
#pragma unroll
    for (int k = 0; k < NC; k++)
#pragma unroll
      for (int f = 0; f < 6; f++) fv[k][f] = sameLevelFaceValue(fv[k][f], fkind[f], h);
  } else {
    float *scratch = h3 + (C32::N - 6 * NF);   // overwritten by x values below
#pragma unroll
    for (int k = 0; k < NC; k++) stageFaces<B>(scratch, a, bid, lev, h, off, X, fkind, foff, threadIdx.x + NT * k);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NC; k++)
#pragma unroll
      for (int f = 0; f < 6; f++) fv[k][f] = scratch[f * NF + tx + B * (ty + YS * k)];
    __syncthreads();
  }
  C32Ctx c;
  c.tx = tx;
  c.ty0 = ty;
  c.o0 = tx / S + ty;
  c.h30 = (ty + 1) * P + tx + 1;
  c.rb = blockRsrc(Bv + off, B * B * B * 4);
  c.rx = blockRsrc(X + off, B * B * B * 4);
  c.rw = blockRsrc(a.pack16y + off, B * B * B * 4);
  c.wxp = wxp;
  if (a.omegaOf) omega = a.omegaOf[a.probOf[bid]];
  c.om = omega; c.omm = 1.0f - omega; c.HR = 1.0f / h;
#pragma unroll
  for (int k = 0; k < NC; k++) {
    c.fz0[k] = fv[k][4];
    c.fz1[k] = fv[k][5];
  }
  /* per-column skip mask and top W+z, coefficient records, then x and b */
  C32Col r[NC];
#pragma unroll
  for (int k = 0; k < NC; k++) {
    const uint2 cw = a.pack16col[(off >> 5) + threadIdx.x + NT * k];
    c.skip[k] = cw.x;
    c.w6top[k] = lo16<PREC>(cw.y);
  }
  const __amdgpu_buffer_rsrc_t rr = blockRsrc(a.pack16 + off, B * B * B * 8);
  c32LoadRecords<0, 0>(r[0], rr, c, wxp);
  c32LoadRecords<16, 0>(r[0], rr, c, wxp);
  if constexpr (NC > 1) {
    c32LoadRecords<0, 1>(r[1], rr, c, wxp);
    c32LoadRecords<16, 1>(r[1], rr, c, wxp);
  }
  c32StageX<0>(r[0], h3, c, h);
  if constexpr (NC > 1) c32StageX<1>(r[1], h3, c, h);
#pragma unroll
  for (int k = 0; k < NC; k++) {
    const int w = ty + YS * k;
    h3[w * P2 + (tx + 1) * P] = fv[k][0];
    h3[w * P2 + (tx + 1) * P + B + 1] = fv[k][1];
    h3[w * P2 + tx + 1] = fv[k][2];
    h3[w * P2 + (B + 1) * P + tx + 1] = fv[k][3];
  }
  __syncthreads();
  /* sweep k is reversed iff reverse == (k & 1) (src/msbg3.cpp:1157); the
   * register arrays are mirrored around each reversed sweep */
#pragma unroll 1
  for (int k = 0; k < 4; k++) {
    const bool fwd = (reverse != 0) != ((k & 1) != 0);
    if (!fwd) {
#pragma unroll
      for (int q = 0; q < NC; q++) c32Mirror(r[q]);
    }
    c32Sweep<PREC>(r, h3, c, fwd);
    if (!fwd) {
#pragma unroll
      for (int q = 0; q < NC; q++) c32Mirror(r[q]);
    }
  }
  /* the fifth sweep's lane constants formed afresh from the thread index
   * (an opaque copy) and 1/h kept uniform: otherwise they come back from
   * spill slots, and each such reload waits for every store before it */
  C32Ctx cj = c;
  {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    cj.tx = t % B;
    cj.ty0 = t / B;
    cj.o0 = cj.tx / S + cj.ty0;
    cj.h30 = (cj.ty0 + 1) * P + cj.tx + 1;
    cj.HR = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(c.HR)));
  }
#pragma unroll
  for (int k = 0; k < C32::K; k++) {
    r[0].wq[k] = c32LoadWy<0>(cj, k);
    if constexpr (NC > 1) r[1].wq[k] = c32LoadWy<1>(cj, k);
  }
  C32JacobiAll<PREC, 0>::run(r, h3, cj);
}

__device__ __forceinline__ double blockReduceSum(double v, double *scratch)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int wid = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int nw = (blockDim.x + 63) / 64;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < nw; i++) r += scratch[i];
  return r;
}

/* k_multiply for B <= 16: thread (x, y) computes its z-column.  Every load of
 * the column (flags, x, the coefficients of the cell and of its +x/+y/+z
 * neighbours, b) is issued before any is waited for; h*x goes into the halo
 * array (Halo3) for the in-plane neighbours and stays in registers for the z
 * neighbours.  Cells, operation order and the FP64 partial are those of
 * cellStencil and the B = 32 loop (src/msbg3.cpp:1192-1335, 1282-1321). */
template <int B>
__device__ __forceinline__ double multiplyColumn(const LevelArgs &a, int bid, int lev, float h, int64_t off,
                                                 const float *__restrict__ X, const float *__restrict__ Bv,
                                                 float *__restrict__ Y, int residual, float *h3, float *faces,
                                                 int *fkind, int64_t *foff)
{
  constexpr int S = Cfg<B>::S, NF = Cfg<B>::NF, P = Halo3<B>::P, P2 = Halo3<B>::P2;
  const int x = threadIdx.x % B, y = threadIdx.x / B;
  float fv[6];
  const bool fast = sameLevelFaces<B>(a, bid, lev, h, off, X, fv, fkind, foff);
  if (!fast) stageFaces<B>(faces, a, bid, lev, h, off, X, fkind, foff);
  const int64_t q0 = off + x + B * y;
  const float *bsrc = residual ? Bv : X;   // X stands in when b is not read
  uint16_t fl[B];
  float xv[B], w0[B], w1[B], w2[B], dd[B], w0x[B], w1y[B], bb[B];
#pragma unroll
  for (int z = 0; z < B; z++) {
    fl[z] = a.flags[q0 + B * B * z];
    xv[z] = X[q0 + B * B * z];
  }
#pragma unroll
  for (int z = 0; z < B; z++) {
    const int64_t q = q0 + B * B * z;
    w0[z] = a.W0[q];
    w1[z] = a.W1[q];
    w2[z] = a.W2[q];
    dd[z] = a.D[q];
    w0x[z] = a.W0[x < B - 1 ? q + 1 : (fkind[1] == FK_SAME ? foff[1] + B * y + B * B * z : q)];
    w1y[z] = a.W1[y < B - 1 ? q + B : (fkind[3] == FK_SAME ? foff[3] + x + B * B * z : q)];
    bb[z] = bsrc[q];
  }
  const float w6t = a.W2[fkind[5] == FK_SAME ? foff[5] + x + B * y : q0];
  uint32_t skip = 0;
  const int hc = (y + 1) * P + x + 1;
#pragma unroll
  for (int z = 0; z < B; z++) {
    const bool sk = segmentAllNonFluid<S>((fl[z] & (CELL_SOLID | CELL_VOID)) != 0);
    skip |= (sk ? 1u : 0u) << z;
    xv[z] = sk ? 0.0f : xv[z] * h;
    h3[hc + z * P2] = xv[z];
  }
  float fz0, fz1;
  if (fast) {
    h3[y * P2 + (x + 1) * P] = sameLevelFaceValue(fv[0], fkind[0], h);
    h3[y * P2 + (x + 1) * P + B + 1] = sameLevelFaceValue(fv[1], fkind[1], h);
    h3[y * P2 + x + 1] = sameLevelFaceValue(fv[2], fkind[2], h);
    h3[y * P2 + (B + 1) * P + x + 1] = sameLevelFaceValue(fv[3], fkind[3], h);
    fz0 = sameLevelFaceValue(fv[4], fkind[4], h);
    fz1 = sameLevelFaceValue(fv[5], fkind[5], h);
  }
  __syncthreads();
  if (!fast) {
    h3[y * P2 + (x + 1) * P] = faces[0 * NF + x + B * y];
    h3[y * P2 + (x + 1) * P + B + 1] = faces[1 * NF + x + B * y];
    h3[y * P2 + x + 1] = faces[2 * NF + x + B * y];
    h3[y * P2 + (B + 1) * P + x + 1] = faces[3 * NF + x + B * y];
    fz0 = faces[4 * NF + x + B * y];
    fz1 = faces[5 * NF + x + B * y];
    __syncthreads();
  }
  const bool m0 = x == 0 && fkind[0] >= FK_COARSER, m2 = y == 0 && fkind[2] >= FK_COARSER;
  const bool m4 = fkind[4] >= FK_COARSER;
  const float xOut = fkind[1] == FK_DOMAIN ? a.bcW[0] : 1.0f, yOut = fkind[3] == FK_DOMAIN ? a.bcW[1] : 1.0f;
  const float zOut = fkind[5] == FK_SAME ? w6t : (fkind[5] == FK_DOMAIN ? a.bcW[2] : 1.0f);
  const bool xIn = x < B - 1 || fkind[1] == FK_SAME, yIn = y < B - 1 || fkind[3] == FK_SAME;
  double acc = 0.0;
#pragma unroll
  for (int z = 0; z < B; z++) {
    if ((skip >> z) & 1u) continue;
    const float *p = h3 + hc + z * P2;
    const float F0 = xv[z];
    const float F1 = p[-1], F2 = p[1], F3 = p[-P], F4 = p[P];
    const float F5 = z > 0 ? xv[z > 0 ? z - 1 : 0] : fz0;
    const float F6 = z < B - 1 ? xv[z < B - 1 ? z + 1 : 0] : fz1;
    const float W1 = m0 ? 1.0f : w0[z];
    const float W3 = m2 ? 1.0f : w1[z];
    const float W5 = (z == 0 && m4) ? 1.0f : w2[z];
    const float W2 = xIn ? w0x[z] : xOut;
    const float W4 = yIn ? w1y[z] : yOut;
    const float W6 = z < B - 1 ? w2[z < B - 1 ? z + 1 : 0] : zOut;
    float s = F1 * W1;
    s = fmaf(F2, W2, s);
    s = fmaf(F3, W3, s);
    s = fmaf(F4, W4, s);
    s = fmaf(F5, W5, s);
    s = fmaf(F6, W6, s);
    const float D = dd[z];
    float r = D != 0.0f ? fmaf(1.0f / D, F0, -s) : 0.0f;
    if (residual) r = bb[z] - r;
    Y[q0 + B * B * z] = r;
    acc += (double)F0 * (double)r;
  }
  return acc;
}

/* y = A x (or r = b - A x); non-fluid segments are not written. */
template <int B>
__global__ void __launch_bounds__(Cfg<B>::NT) k_multiply(LevelArgs a, const int *__restrict__ blocks,
                                                         const float *__restrict__ X, const float *__restrict__ Bv,
                                                         float *__restrict__ Y, int residual, double *partials,
                                                         int partialBase, const int *done)
{
  constexpr bool COL = B <= 16;
  __shared__ float tile[COL ? 1 : Cfg<B>::NIN + 6 * Cfg<B>::NF];
  __shared__ float h3[COL ? Halo3<B>::N : 1];
  __shared__ float faces[COL ? 6 * Cfg<B>::NF : 1];
  __shared__ double red[16];
  const int li = xcdListIndex();
  /* the block's descriptor is read before the done flag is tested (the empty
   * asm keeps the test after the loads), so that the flag's load joins the
   * descriptor's instead of preceding it */
  const int bid = blocks[li];
  const int lev = a.level[bid];
  const int64_t off = a.off[bid];
  const int isDone = *done;
  asm volatile("" ::"s"(lev), "s"(isDone));
  if (isDone) return;
  const float h = (float)(1 << lev);
  int fkind[6];
  int64_t foff[6];
  double acc = 0.0;
  if constexpr (COL) {
    acc = multiplyColumn<B>(a, bid, lev, h, off, X, Bv, Y, residual, h3, faces, fkind, foff);
  } else {
  uint32_t skipmask;
  stageTile<B>(tile, skipmask, a, bid, lev, h, off, X, fkind, foff);
  __syncthreads();
  const int x = threadIdx.x % B, y = threadIdx.x / B;
  for (int z = 0; z < B; z++) {
    if ((skipmask >> z) & 1u) continue;
    float F0, Sm, D;
    cellStencil<B>(tile, a, off, fkind, foff, x, y, z, F0, Sm, D);
    const int64_t c = off + x + B * y + B * B * z;
    float r = D != 0.0f ? fmaf(1.0f / D, F0, -Sm) : 0.0f;
    if (residual) r = Bv[c] - r;
    Y[c] = r;
    acc += (double)F0 * (double)r;
  }
  }
  if (partials) {
    const double s = blockReduceSum(lev == 0 ? acc : 0.0, red);
    if (threadIdx.x == 0) partials[partialBase + li] = s;   // list order: the fixed summation order
  }
}

/* Halo value of face `face` (kind, neighbour offset off2) at face cell
 * (u, w): the value stageFaces writes at faces[face][u + B w], by the same
 * float operations. */
template <int B, typename T>
__device__ __forceinline__ float faceCellValue(const LevelArgs &a, float h, int64_t off, const T *__restrict__ X,
                                               int face, int kind, int64_t off2, int u, int w)
{
  const float a13 = (float)(1.0 * (1. / 3.));
  const float a23 = (float)(1.0 * 2 * (1. / 3.));
  const float *Wd[3] = {a.W0, a.W1, a.W2};
  const int dir = face >> 1, side = face & 1;
  if (kind == FK_DOMAIN) return 0.0f;
  if (kind == FK_SAME) {
    const int e = side ? 0 : B - 1;
    const int src = dir == 0 ? e + B * u + B * B * w : (dir == 1 ? u + B * e + B * B * w : u + B * w + B * B * e);
    return vload(X, off2 + src) * h;
  }
  if (kind == FK_COARSER) {
    constexpr int Bc = B / 2 > 0 ? B / 2 : 1;
    const int e = side ? 0 : Bc - 1;
    const int cu = u / 2, cw = w / 2;
    const int src = dir == 0 ? e + Bc * cu + Bc * Bc * cw
                             : (dir == 1 ? cu + Bc * e + Bc * Bc * cw : cu + Bc * cw + Bc * Bc * e);
    float aa = a23;
    if (!side) {
      const int own = dir == 0 ? B * u + B * B * w : (dir == 1 ? u + B * B * w : u + B * w);
      aa *= Wd[dir][off + own];
    }
    const double sv = (double)(aa * vload(X, off2 + src));
    return (float)(sv * (double)h);
  }
  constexpr int Bf = 2 * B;
  const int e = side ? 0 : Bf - 1;
  double sv = 0.0;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int d1 = c & 1, d2 = c >> 1;
    const int fu = 2 * u + d1, fw = 2 * w + d2;
    const int src = dir == 0 ? e + Bf * fu + Bf * Bf * fw
                             : (dir == 1 ? fu + Bf * e + Bf * Bf * fw : fu + Bf * fw + Bf * Bf * e);
    float aa = a13;
    if (side) aa *= Wd[dir][off2 + src];
    sv += (double)(aa * vload(X, off2 + src));
  }
  return (float)(sv * (double)h);
}

/* Vectors of one k_multiply_z launch: y_v = A x_v (or r_v = b - A x_v) for
 * v < NV, sharing the coefficient, flag and b loads. */
template <int NV, typename T = float>
struct MulVecs {
  const T *x[NV];
  T *y[NV];
};

/* y = A x (or r = b - A x) for B = 32: one workgroup of B*B threads per
 * block marches through the block's z-planes.  Thread (x, y) owns its
 * z-column; each step publishes the plane's h*x (0 on all-non-fluid
 * segments) with its x/y halo ring into a double-buffered LDS plane (one
 * LDS-only barrier per step), takes the z neighbours from registers, and
 * issues later planes' loads (for one vector: coefficients, b and the
 * ring's face values two planes ahead, x, flags and W-z three; for several,
 * one and two) before it computes, so that they are in flight across the
 * barriers.  Cells, face
 * values, operation order and the FP64 partial are those of k_multiply.
 *
 * NV > 1 (residual form, the coarse correction's directions): NV vectors
 * in one pass, and instead of the <x, y> partial the Galerkin partials of
 * level-0 blocks in FP64, for a <= b <x_a, b - y_b> (= <x_a, A x_b>) then
 * <x_a, b>, value j at partials[j * gridDim.x + list position].
 * T = uint16_t (residual form only): x and y held as BF16 bits, x widened to
 * FP32 as it is read and y rounded to nearest even as it is stored; the
 * arithmetic and the Galerkin sums are those of the FP32 form on the stored x. */
template <int B, int NV, typename T = float>
__global__ void __launch_bounds__(B * B) k_multiply_z(LevelArgs a, const int *__restrict__ blocks, MulVecs<NV, T> mv,
                                                      const float *__restrict__ Bv, int residual, double *partials,
                                                      int partialBase, const int *done)
{
  constexpr int S = Cfg<B>::S, P = B + 2, P2 = P * P, BB = B * B;
  constexpr int NG = NV > 1 ? NV * (NV + 1) / 2 + NV : 1;   // partial sums per workgroup
  __shared__ float pl[NV][2 * P2];
  __shared__ double red[16];
  const int li = xcdListIndex();
  const int bid = blocks[li];
  const int lev = a.level[bid];
  const int64_t off = a.off[bid];
  const int isDone = *done;
  asm volatile("" ::"s"(lev), "s"(isDone));
  if (isDone) return;
  const float h = (float)(1 << lev);
  const int x = threadIdx.x % B, y = threadIdx.x / B;
  const NbDesc &nd = a.nbd[bid];
  int fkind[6];
  int64_t foff[6];
#pragma unroll
  for (int face = 0; face < 6; face++) {
    const int l2 = nd.lev[face];
    fkind[face] = l2 < 0 ? FK_DOMAIN : (l2 == lev ? FK_SAME : (l2 > lev ? FK_COARSER : FK_FINER));
    foff[face] = l2 < 0 ? 0 : nd.off[face];
  }
  /* x stands in when b is not read; BF16 vectors are always in residual form */
  const float *bsrc = Bv;
  if constexpr (std::is_same_v<T, float>) bsrc = residual ? Bv : mv.x[0];
  const int64_t q0 = off + x + B * y;
  /* ring role: threads 0 .. 4B-1 load the x/y halo of each plane, face
   * rf = t / B at face cell (ri, z) */
  const int t = threadIdx.x;
  const bool ring = t < 4 * B;
  const int rf = ring ? t / B : 0, ri = t % B;
  const int rdir = rf >> 1, rside = rf & 1;
  const int rl2 = nd.lev[rf];
  const int rkind = !ring ? FK_DOMAIN : (rl2 < 0 ? FK_DOMAIN : (rl2 == lev ? FK_SAME : (rl2 > lev ? FK_COARSER : FK_FINER)));
  const int64_t roff = rl2 < 0 ? 0 : nd.off[rf];
  const int rpos = rf == 0 ? (ri + 1) * P : (rf == 1 ? (ri + 1) * P + B + 1 : (rf == 2 ? ri + 1 : (B + 1) * P + ri + 1));
  const float *rW = rdir == 0 ? a.W0 : a.W1;
  /* raw loads of the ring value at plane z: same level x; coarser x and own
   * -face coefficient */
  auto ringLoad = [&](int z, float (&r1)[NV], float &r2) {
#pragma unroll
    for (int v = 0; v < NV; v++) r1[v] = 0.0f;
    r2 = 1.0f;
    if (rkind == FK_SAME) {
      const int e = rside ? 0 : B - 1;
      const int src = rdir == 0 ? e + B * ri + BB * z : ri + B * e + BB * z;
#pragma unroll
      for (int v = 0; v < NV; v++) r1[v] = vload(mv.x[v], roff + src);
    } else if (rkind == FK_COARSER) {
      constexpr int Bc = B / 2;
      const int e = rside ? 0 : Bc - 1;
      const int cu = ri / 2, cw = z / 2;
      const int src = rdir == 0 ? e + Bc * cu + Bc * Bc * cw : cu + Bc * e + Bc * Bc * cw;
#pragma unroll
      for (int v = 0; v < NV; v++) r1[v] = vload(mv.x[v], roff + src);
      if (!rside) r2 = rW[off + (rdir == 0 ? B * ri + BB * z : ri + BB * z)];
    }
  };
  auto ringValue = [&](int v, int z, float r1, float r2) -> float {
    if (rkind == FK_SAME) return r1 * h;
    if (rkind == FK_COARSER) {
      float aa = (float)(1.0 * 2 * (1. / 3.));
      if (!rside) aa *= r2;
      const double sv = (double)(aa * r1);
      return (float)(sv * (double)h);
    }
    if (rkind == FK_FINER) return faceCellValue<B>(a, h, off, mv.x[v], rf, rkind, roff, ri, z);
    return 0.0f;
  };
  float fz0[NV], fz1[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    fz0[v] = faceCellValue<B>(a, h, off, mv.x[v], 4, fkind[4], foff[4], x, y);
    fz1[v] = faceCellValue<B>(a, h, off, mv.x[v], 5, fkind[5], foff[5], x, y);
  }
  const float w6t = a.W2[fkind[5] == FK_SAME ? foff[5] + x + B * y : q0];
  const bool m0 = x == 0 && fkind[0] >= FK_COARSER, m2 = y == 0 && fkind[2] >= FK_COARSER;
  const bool m4 = fkind[4] >= FK_COARSER;
  const float xOut = fkind[1] == FK_DOMAIN ? a.bcW[0] : 1.0f, yOut = fkind[3] == FK_DOMAIN ? a.bcW[1] : 1.0f;
  const float zOut = fkind[5] == FK_SAME ? w6t : (fkind[5] == FK_DOMAIN ? a.bcW[2] : 1.0f);
  const bool xIn = x < B - 1 || fkind[1] == FK_SAME, yIn = y < B - 1 || fkind[3] == FK_SAME;
  auto qx = [&](int z) { return x < B - 1 ? q0 + BB * z + 1 : (fkind[1] == FK_SAME ? foff[1] + B * y + BB * z : q0 + BB * z); };
  auto qy = [&](int z) { return y < B - 1 ? q0 + BB * z + B : (fkind[3] == FK_SAME ? foff[3] + x + BB * z : q0 + BB * z); };

  /* loads run ahead of the plane being computed: coefficients, b and the
   * ring's face values (group A) PA planes, x, flags and W-z (group B) PA + 1
   * planes; index k of a group's arrays holds plane z + k */
  constexpr int PA = NV == 1 ? 2 : 1, PB = PA + 1;   // deeper did not pay (3 and 4 ahead measured)
  float aw0[PA + 1], aw1[PA + 1], add[PA + 1], aw0x[PA + 1], aw1y[PA + 1], abb[PA + 1], ar2[PA + 1];
  float ar1[PA + 1][NV];
  float bx[PB + 1][NV], bw2[PB + 1];
  uint16_t bfl[PB + 1];
  auto loadA = [&](int k, int zp) {
    const int64_t q = q0 + BB * zp;
    aw0[k] = a.W0[q];
    aw1[k] = a.W1[q];
    add[k] = a.D[q];
    aw0x[k] = a.W0[qx(zp)];
    aw1y[k] = a.W1[qy(zp)];
    abb[k] = bsrc[q];
    ringLoad(zp, ar1[k], ar2[k]);
  };
  auto loadB = [&](int k, int zp) {
    const int64_t q = q0 + BB * zp;
#pragma unroll
    for (int v = 0; v < NV; v++) bx[k][v] = vload(mv.x[v], q);
    bfl[k] = a.flags[q];
    bw2[k] = a.W2[q];
  };
#pragma unroll
  for (int k = 0; k < PA; k++) loadA(k, k);
#pragma unroll
  for (int k = 0; k < PB; k++) loadB(k, k);
  bool sk = segmentAllNonFluid<S>((bfl[0] & (CELL_SOLID | CELL_VOID)) != 0);
  float xs[NV], xsPrev[NV];   // h*x of plane z as the stencil reads it, and of plane z-1
#pragma unroll
  for (int v = 0; v < NV; v++) {
    xs[v] = sk ? 0.0f : bx[0][v] * h;
    xsPrev[v] = 0.0f;
  }
  double acc[NG];
#pragma unroll
  for (int j = 0; j < NG; j++) acc[j] = 0.0;
#pragma unroll
  for (int z = 0; z < B; z++) {
    /* this step's loads, for planes z + PA and z + PB */
    if (z + PA < B) loadA(PA, z + PA);
    if (z + PB < B) loadB(PB, z + PB);
    /* h*x of plane z+1 (the +z neighbour) and its skip state */
    bool skn = false;
    float xsn[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) xsn[v] = 0.0f;
    if (z + 1 < B) {
      skn = segmentAllNonFluid<S>((bfl[1] & (CELL_SOLID | CELL_VOID)) != 0);
#pragma unroll
      for (int v = 0; v < NV; v++) xsn[v] = skn ? 0.0f : bx[1][v] * h;
    }
#pragma unroll
    for (int v = 0; v < NV; v++) {
      float *p = pl[v] + (z & 1) * P2;
      p[(y + 1) * P + x + 1] = xs[v];
      if (ring) p[rpos] = ringValue(v, z, ar1[0][v], ar2[0]);
    }
    ldsBarrier();
    if (!sk) {
      const float dd = add[0], bb = abb[0];
      const float W1 = m0 ? 1.0f : aw0[0];
      const float W3 = m2 ? 1.0f : aw1[0];
      const float W5 = (z == 0 && m4) ? 1.0f : bw2[0];
      const float W2 = xIn ? aw0x[0] : xOut;
      const float W4 = yIn ? aw1y[0] : yOut;
      const float W6 = z < B - 1 ? bw2[1] : zOut;
      float res[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const float *c = pl[v] + (z & 1) * P2 + (y + 1) * P + x + 1;
        const float F0 = xs[v];
        const float F1 = c[-1], F2 = c[1], F3 = c[-P], F4 = c[P];
        const float F5 = z > 0 ? xsPrev[v] : fz0[v];
        const float F6 = z < B - 1 ? xsn[v] : fz1[v];
        float s = F1 * W1;
        s = fmaf(F2, W2, s);
        s = fmaf(F3, W3, s);
        s = fmaf(F4, W4, s);
        s = fmaf(F5, W5, s);
        s = fmaf(F6, W6, s);
        float r = dd != 0.0f ? fmaf(1.0f / dd, F0, -s) : 0.0f;
        if (residual) r = bb - r;
        vstore(mv.y[v], q0 + BB * z, r);
        res[v] = r;
        if constexpr (NV == 1) acc[0] += (double)F0 * (double)r;
      }
      if constexpr (NV > 1) {
        /* Galerkin sums of this cell: x_a with b - y_b, then with b */
        const double bd = bb;
        double xv[NV], ax[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          xv[v] = bx[0][v];
          ax[v] = bd - (double)res[v];
        }
        int j = 0;
#pragma unroll
        for (int u = 0; u < NV; u++)
#pragma unroll
          for (int v = u; v < NV; v++) acc[j++] += xv[u] * ax[v];
#pragma unroll
        for (int u = 0; u < NV; u++) acc[j++] += xv[u] * bd;
      }
    }
    /* rotate */
#pragma unroll
    for (int v = 0; v < NV; v++) {
      xsPrev[v] = xs[v];
      xs[v] = xsn[v];
    }
    sk = skn;
#pragma unroll
    for (int k = 0; k < PA; k++) {
      aw0[k] = aw0[k + 1];
      aw1[k] = aw1[k + 1];
      add[k] = add[k + 1];
      aw0x[k] = aw0x[k + 1];
      aw1y[k] = aw1y[k + 1];
      abb[k] = abb[k + 1];
      ar2[k] = ar2[k + 1];
#pragma unroll
      for (int v = 0; v < NV; v++) ar1[k][v] = ar1[k + 1][v];
    }
#pragma unroll
    for (int k = 0; k < PB; k++) {
#pragma unroll
      for (int v = 0; v < NV; v++) bx[k][v] = bx[k + 1][v];
      bfl[k] = bfl[k + 1];
      bw2[k] = bw2[k + 1];
    }
  }
  if (partials) {
#pragma unroll
    for (int j = 0; j < NG; j++) {
      const double sres = blockReduceSum(lev == 0 ? acc[j] : 0.0, red);
      if (threadIdx.x == 0) {
        if constexpr (NV == 1) partials[partialBase + li] = sres;   // list order: the fixed summation order
        else partials[(size_t)j * gridDim.x + li] = sres;
      }
    }
  }
}

/* Restriction into the coarse right-hand side of MG level Lc from the fine
 * residual of level Lc-1: 2x2x2 mean for coarsened blocks; for blocks with
 * the same level at both MG levels, a copy (first V-cycle on the channel
 * set) or zero (later V-cycles), which is what the reference's block
 * reference plus in-place zeroing produces. */
__global__ void k_restrict(const int *__restrict__ blocks, const int *__restrict__ level0, const int *__restrict__ levC,
                           const int64_t *__restrict__ offC, const int64_t *__restrict__ offF, int Lc, int bsx0,
                           const float *__restrict__ src, float *__restrict__ dst, int copyMode, const int *done)
{
  if (*done) return;
  const int bid = blocks[xcdListIndex()];
  const int lev = levC[bid];
  const int b = bsx0 >> lev, n = b * b * b;
  const int64_t oc = offC[bid], of = offF[bid];
  if (level0[bid] >= Lc) {
    for (int v = threadIdx.x; v < n; v += blockDim.x) dst[oc + v] = copyMode ? src[of + v] : 0.0f;
    return;
  }
  const int bf = 2 * b;
  for (int v = threadIdx.x; v < n; v += blockDim.x) {
    const int vx = v % b, vy = (v / b) % b, vz = v / (b * b);
    float val = 0.0f;
    for (int k = 0; k < 2; k++)
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
          const float f = src[of + (2 * vx + i) + bf * ((2 * vy + j) + bf * (2 * vz + k))];
          val = (float)((double)val + (double)f * 0.125);
        }
    dst[oc + v] = val;
  }
}

/* x_L += P x_{L+1} on non-SOLID/VOID cells.  The block edge is a power of
 * two, so the cell coordinates are shifts and masks; each cell's flag, x_L
 * and coarse value are loaded together (the coarse index stays inside the
 * coarse block for every cell) and only the add depends on the flag, so
 * several cells' loads are in flight at once. */
__global__ void k_prolong(const int *__restrict__ blocks, const int *__restrict__ level0, const int *__restrict__ levF,
                          const int64_t *__restrict__ offF, const int64_t *__restrict__ offC,
                          const uint16_t *__restrict__ flagsF, int Lf, int bsx0, float *__restrict__ xf,
                          const float *__restrict__ xc, const int *done)
{
  if (*done) return;
  const int bid = blocks[xcdListIndex()];
  const int lev = levF[bid];
  const int b = bsx0 >> lev, lg = __ffs(b) - 1, n = b * b * b, m = b - 1;
  const int64_t of = offF[bid], oc = offC[bid];
  const bool same = level0[bid] >= Lf + 1;
  const int bc = b / 2;
#pragma unroll 4
  for (int v = threadIdx.x; v < n; v += blockDim.x) {
    const uint16_t fl = flagsF[of + v];
    const int vx = v & m, vy = (v >> lg) & m, vz = v >> (2 * lg);
    const float e = xc[oc + (same ? v : (vx >> 1) + bc * ((vy >> 1) + bc * (vz >> 1)))];
    const float xv = xf[of + v];
    if (!(fl & (CELL_SOLID | CELL_VOID))) xf[of + v] = xv + e;
  }
}

__global__ void k_zero_blocks(const int *__restrict__ blocks, const int *__restrict__ lev, const int64_t *__restrict__ off,
                              int bsx0, float *__restrict__ x, const int *done)
{
  if (*done) return;
  const int bid = blocks[xcdListIndex()];
  const int b = bsx0 >> lev[bid], n = b * b * b;
  const int64_t o = off[bid];
#pragma unroll 4
  for (int v = threadIdx.x; v < n; v += blockDim.x) x[o + v] = 0.0f;
}

__global__ void k_copy_blocks(const int *__restrict__ blocks, const int *__restrict__ lev, const int64_t *__restrict__ off,
                              int bsx0, const float *__restrict__ s, float *__restrict__ d, const int *done)
{
  if (*done) return;
  const int bid = blocks[xcdListIndex()];
  const int b = bsx0 >> lev[bid], n = b * b * b;
  const int64_t o = off[bid];
#pragma unroll 4
  for (int v = threadIdx.x; v < n; v += blockDim.x) d[o + v] = s[o + v];
}

constexpr int kVecBlocks = 1024, kVecThreads = 256;
/* workgroups of grid-stride passes without a reduction (d update): twice as
 * many waves in flight as kVecBlocks gives */
constexpr int kVecBlocksNoSum = 2048;

/* partial <a, b> over [0, n) (FP64 products) */
__global__ void __launch_bounds__(kVecThreads) k_dot(const float *__restrict__ a, const float *__restrict__ b, int64_t n,
                                                     double *partials, const int *done)
{
  if (*done) return;
  __shared__ double red[kVecThreads / 64];
  double acc = 0.0;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc += (double)a[i] * (double)b[i];
  const double s = blockReduceSum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

/* r = b over [0, n) and partial |b|^2 */
__global__ void __launch_bounds__(kVecThreads) k_copy_norm(const float *__restrict__ b, float *__restrict__ r, int64_t n,
                                                           double *partials)
{
  __shared__ double red[kVecThreads / 64];
  double acc = 0.0;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float v = b[i];
    r[i] = v;
    acc += (double)v * (double)v;
  }
  const double s = blockReduceSum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

/* x += alpha d; r -= alpha q; partial |r|^2 (src/msbg3.cpp:3175-3195) */
__global__ void __launch_bounds__(kVecThreads) k_update_xr(float *__restrict__ x, float *__restrict__ r,
                                                           const float *__restrict__ d, const float *__restrict__ q,
                                                           int64_t n, const PcgState *st, double *partials)
{
  if (st->done) return;
  __shared__ double red[kVecThreads / 64];
  const float alpha = st->alpha;
  double acc = 0.0;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    x[i] = fmaf(alpha, d[i], x[i]);
    const float rv = fmaf(-alpha, q[i], r[i]);
    r[i] = rv;
    acc += (double)rv * (double)rv;
  }
  const double s = blockReduceSum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

/* d = z + beta d (src/msbg3.cpp:3224-3235) */
__global__ void __launch_bounds__(kVecThreads) k_axpy_d(float *__restrict__ d, const float *__restrict__ z, int64_t n,
                                                        const PcgState *st)
{
  if (st->done) return;
  const float beta = st->beta;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = fmaf(beta, d[i], z[i]);
}

/* Deterministic fixed-order sum of partials by one workgroup of 256 threads. */
__device__ double sumPartials(const double *p, int n)
{
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += p[i];
  const double s = blockReduceSum(acc, red);
  __shared__ double out;
  if (threadIdx.x == 0) out = s;
  __syncthreads();
  return out;
}

enum FinalizeOp { FIN_BNORM = 0, FIN_RHO0, FIN_ALPHA, FIN_CONV, FIN_BETA, FIN_FINAL, FIN_RZOLD, FIN_RZPLAIN,
                  FIN_ZCR, FIN_GAMMA };

__device__ void applyFinalize(PcgState *st, double s, int op, double *out);

__global__ void __launch_bounds__(256) k_finalize(PcgState *st, const double *partials, int n, int op, double *out)
{
  const bool skipIfDone = (op == FIN_ALPHA || op == FIN_CONV || op == FIN_BETA || op == FIN_RZOLD ||
                          op == FIN_RZPLAIN || op == FIN_ZCR || op == FIN_GAMMA);
  if (skipIfDone && st->done) return;
  const double s = sumPartials(partials, n);
  if (threadIdx.x != 0) return;
  applyFinalize(st, s, op, out);
}

/* Batched form: workgroup p finalizes problem p from its chunk partials. */
__global__ void __launch_bounds__(256) k_finalize_b(PcgState *st, const double *partials, const int *cb, const int *ce,
                                                   int op, double *out, int outStride)
{
  const int p = blockIdx.x;
  const bool skipIfDone = (op == FIN_ALPHA || op == FIN_CONV || op == FIN_BETA);
  if (skipIfDone && st[p].done) return;
  const double s = sumPartials(partials + cb[p], ce[p] - cb[p]);
  if (threadIdx.x != 0) return;
  applyFinalize(st + p, s, op, out ? out + (size_t)p * outStride : nullptr);
}

__device__ void applyFinalize(PcgState *st, double s, int op, double *out)
{
  switch (op) {
    case FIN_BNORM:
      st->bnorm2 = s;
      st->done = (s <= 1e-30 || !isfinite(s)) ? 1 : 0;
      st->conv = -1; st->iter = 0; st->brk = !isfinite(s) ? 3 : ((s <= 1e-30) ? 2 : 0);
      if (st->coarseMode) st->rzPlain = s;   // fallback direction r0 has r0.r0 = |b|^2
      break;
    case FIN_RHO0:
      st->rho = s;
      st->usePlain = 0;
      if (st->learned && (!(s > 1e-30) || !isfinite(s) || st->bad)) {
        /* the corrected z is unusable: start from the plain one */
        st->learned = 0; st->fellBack = 0; st->usePlain = 1; st->rho = st->rzPlain;
        /* coarse form: that first direction is r itself; the V-cycle-preconditioned
         * iteration that follows starts afresh (beta = 0) */
        if (st->coarseMode) st->restart = 1;
      }
      break;
    case FIN_RZPLAIN: if (st->learned) st->rzPlain = s; break;
    case FIN_ZCR: st->zcr = s; break;
    case FIN_GAMMA: {
      /* s = <zc, A zc>: the step along zc that minimises the A-norm error */
      const double gm = st->zcr / s;
      st->gamma = (s > 1e-300 && isfinite(gm)) ? (float)gm : 0.0f;
      break;
    }
    case FIN_ALPHA: {
      st->dq = s;
      const double rho = st->rho;
      if (!(s > 1e-30) || !(rho > 1e-30) || !isfinite(s) || !isfinite(rho)) { st->done = 1; st->brk = 1; break; }
      const double al = rho / s;
      if (!isfinite(al)) { st->done = 1; st->brk = 1; break; }
      st->alpha = (float)al;
      break;
    }
    case FIN_CONV:
      st->rr = s;
      st->iter += 1;
      if (out) out[st->iter - 1] = s;
      if (s <= st->tolSq * st->bnorm2) { st->done = 1; st->conv = st->iter; }
      else if (st->iter >= st->maxIter) st->done = 1;
      if (st->coarseMode) st->rzPlain = s;   // fallback direction r has r.r = |r|^2
      if (st->learned && st->stallWin > 0) {
        /* stall test of the learned correction (residuals of a flexible CG
         * need not decrease monotonically): no |r|^2 of the last stallWin
         * iterations fell below stallFactor times the smallest one before
         * them (since the window's start).  The coarse form answers its
         * first stall by restarting the directions (beta = 0) with the
         * correction kept and the window started afresh; otherwise, and at
         * the next stall, it falls back to the plain preconditioner and
         * restarts the directions */
        const int W = st->stallWin, k = st->iter - st->winBase;
        double &slot = st->rrWin[k % W];
        if (k > W) st->rrBest = k == W + 1 ? slot : fmin(st->rrBest, slot);
        slot = s;
        if (k > W) {
          double m = st->rrWin[0];
          for (int j = 1; j < W; j++) m = fmin(m, st->rrWin[j]);
          if (!(m < st->stallFactor * st->rrBest)) {
            st->restart = 1;
            if (st->coarseMode && st->stallRestarts == 0) {
              st->stallRestarts = 1;
              st->winBase = st->iter;
            } else {
              st->learned = 0;
              st->fellBack = st->iter;
            }
          }
        }
      }
      break;
    case FIN_RZOLD: if (st->learned) st->rzOld = s; break;
    case FIN_BETA: {
      st->usePlain = 0;
      if (st->learned) {
        /* flexible (Polak-Ribiere) beta while the learned correction is on */
        const double be = (s - st->rzOld) / st->rho;
        if (st->bad || !(s > 1e-30) || !isfinite(s) || !isfinite(be)) {
          /* unusable corrected z: restart from the plain V-cycle output */
          st->learned = 0; st->fellBack = st->iter; st->usePlain = 1; st->restart = st->coarseMode ? 1 : 0;
          st->beta = 0.0f;
          st->rho = st->rzPlain;
          if (!(st->rho > 1e-30) || !isfinite(st->rho)) { st->done = 1; st->brk = 1; }
          break;
        }
        st->beta = st->restart ? 0.0f : (float)be;
        st->restart = 0;
        st->rho = s;
        break;
      }
      if (!(s > 1e-30) || !isfinite(s)) { st->done = 1; st->brk = 1; break; }
      const double be = s / st->rho;   // Fletcher-Reeves for the fixed V-cycle preconditioner
      if (!isfinite(be)) { st->done = 1; st->brk = 1; break; }
      st->beta = st->restart ? 0.0f : (float)be;
      st->restart = 0;
      st->rho = s;
      break;
    }
    case FIN_FINAL: if (out) out[0] = s; break;
  }
  st->skipNet = (st->done || !st->learned) ? 1 : 0;
}

/* Divergence right-hand side of level-0 blocks from the scene's MAC velocity
 * (flipRhsCB, src/flip_sim.cpp:329-338; sampled as in src/msbg3.cpp:2551-2579). */
__global__ void k_scene_rhs(msbgx_scene s, const int *__restrict__ blocks, const int64_t *__restrict__ off, int nbx,
                            int nby, int bsx0, float *__restrict__ Bout, int bidShift)
{
  const int bid = blocks[blockIdx.x];
  const int lb = bid - bidShift;   // block id inside its own raster
  const int bx = lb % nbx, by = (lb / nbx) % nby, bz = lb / (nbx * nby);
  const int n = bsx0 * bsx0 * bsx0;
  const int64_t o = off[bid];
  for (int v = threadIdx.x; v < n; v += blockDim.x) {
    const int vx = v % bsx0, vy = (v / bsx0) % bsx0, vz = v / (bsx0 * bsx0);
    Bout[o + v] = msbgx_scene_rhs(&s, bx * bsx0 + vx, by * bsx0 + vy, bz * bsx0 + vz);
  }
}

/* -------------------------------------------------------------------- host */

/* Device state of a batched block set: the chunk table of the PCG vector
 * kernels (chunk c covers cells [cBeg[c], cEnd[c]) of problem cProb[c];
 * problem p owns chunks [cb[p], ce[p])) and one PCG state per problem. */
/* What a captured PCG graph bakes in besides its buffers: the iterations per
 * replay, the V-cycle's smoothing counts and the inert-coarse decision
 * (MSBGX_OPT_LITERAL_VCYCLE), and for the learned solve the network and its
 * form.  A solve whose key differs from the graph's recaptures. */
struct GraphKey {
  int chunk = 0, nPre = -1, nPost = -1, nCoarse = -1;
  unsigned opts = 0;
  uint64_t net = 0;
  int netMode = -1;
  bool operator==(const GraphKey &o) const
  {
    return chunk == o.chunk && nPre == o.nPre && nPost == o.nPost && nCoarse == o.nCoarse && opts == o.opts &&
           net == o.net && netMode == o.netMode;
  }
  bool operator!=(const GraphKey &o) const { return !(*this == o); }
};

static GraphKey graphKey(const Grid &g, const msbgx_solve_params &p)
{
  GraphKey k;
  k.chunk = std::max(1, g.graphChunk);
  k.nPre = p.n_pre;
  k.nPost = p.n_post;
  k.nCoarse = p.n_coarse;
  k.opts = g.options & (MSBGX_OPT_LITERAL_VCYCLE | MSBGX_OPT_UNFUSED_COARSE);
  return k;
}

struct BatchDev {
  int nChunks = 0;
  int *cProb = nullptr, *cb = nullptr, *ce = nullptr;
  int64_t *cBeg = nullptr, *cEnd = nullptr;
  /* the same table restricted to problems still iterating (live), rebuilt
   * with the block lists when enough problems have converged */
  int nChunksL = 0;
  int *cProbL = nullptr, *cbL = nullptr, *ceL = nullptr;
  int64_t *cBegL = nullptr, *cEndL = nullptr;
  std::vector<int> hProb, hCb, hCe;
  std::vector<int64_t> hBeg, hEnd;
  int listBlocks = 0;       // blocks in the current (possibly compacted) lists
  PcgState *states = nullptr;
  int *allDone = nullptr;   // 1 once every problem has converged or stopped
  double *hist = nullptr;
  int histStride = 0;
  hipGraphExec_t graph = nullptr;
  GraphKey graphKey;
};

struct LearnedCtx;
static void freeLearned(LearnedCtx *c);

struct Dev {
  hipStream_t stream = nullptr;
  PcgState *state = nullptr;
  double *partials = nullptr;
  int partialsCap = 0;
  double *hist = nullptr;
  int histCap = 0;
  int *zeroFlag = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraphExec_t graph = nullptr;
  GraphKey graphKey;
  /* iterative refinement (precision form) */
  double *x64 = nullptr;
  /* precision form: x of the coarsest MG level in FP64 (its pool's cells) */
  double *coarse64 = nullptr;
  int64_t coarse64Cap = 0;
  float *b0 = nullptr;
  int *dActive0 = nullptr;
  int64_t refCap = 0;
  bool b0Valid = false;
  /* batched block sets */
  int *probOf = nullptr;
  float *omegaOf = nullptr;
  BatchDev *batch = nullptr;
  LearnedCtx *learned = nullptr;   // buffers and graph of the learned-correction solve
  /* 32^3 column smoother: coefficient records over the MG0 pool (LevelArgs::pack,
   * LevelArgs::pack16 for 16-bit operands, built for precision pack16Prec) */
  float4 *pack32 = nullptr;
  uint2 *pack16 = nullptr, *pack16col = nullptr;
  uint32_t *pack16y = nullptr;
  int pack16Prec = MSBGX_PREC_FP32;
  /* FLIP caller steps: velocity channels (MG0 pool layout) and the level-0 block list */
  float *vel[3] = {nullptr, nullptr, nullptr};
  int *level0 = nullptr;
  uint8_t *level0NoFluid = nullptr;
  int nLevel0 = 0;
};

static void freeFlip(Dev &d)
{
  for (int c = 0; c < 3; c++) {
    if (d.vel[c]) (void)hipFree(d.vel[c]);
    d.vel[c] = nullptr;
  }
  if (d.level0) (void)hipFree(d.level0);
  if (d.level0NoFluid) (void)hipFree(d.level0NoFluid);
  d.level0 = nullptr;
  d.level0NoFluid = nullptr;
  d.nLevel0 = 0;
}

static Dev &dev(Grid &g) { return *(Dev *)g.dev; }

template <typename T> static int dalloc(T **p, size_t n)
{
  *p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipMalloc((void **)p, n * sizeof(T)));
  return MSBGX_OK;
}

static void freeLevel(MgLevel &m)
{
  auto f = [](void *p) { if (p) (void)hipFree(p); };
  f(m.d_level); f(m.d_off); f(m.d_nbd); f(m.d_flags); f(m.d_D); f(m.d_lists);
  for (int d = 0; d < 3; d++) { f(m.d_W[d]); m.d_W[d] = nullptr; }
  for (int k = 0; k < 4; k++) { f(m.d_H[k]); m.d_H[k] = nullptr; }
  for (int c = 0; c < N_VEC; c++) { f(m.d_ch[c]); m.d_ch[c] = nullptr; }
  m.d_level = nullptr; m.d_off = nullptr; m.d_nbd = nullptr; m.d_flags = nullptr; m.d_D = nullptr;
  m.d_lists = nullptr;
}

/* The PCG residual history (one double per iteration plus the final |r|^2).
 * Captured graphs hold its address, so growing it drops them first. */
static int ensureHist(Grid &g, int n)
{
  Dev &d = dev(g);
  if (d.histCap >= n) return MSBGX_OK;
  HIPCHK(hipStreamSynchronize(d.stream));
  deviceInvalidateGraphs(g);
  if (d.hist) (void)hipFree(d.hist);
  d.hist = nullptr;
  d.histCap = 0;
  if (int rc = dalloc(&d.hist, n)) return rc;
  d.histCap = n;
  return MSBGX_OK;
}

/* Teardown order: drain the stream, destroy every captured graph (they
 * reference the level, learned and batch buffers), then free the buffers. */
void deviceFree(Grid &g)
{
  if (g.dev) {
    Dev &d = dev(g);
    if (d.stream) (void)hipStreamSynchronize(d.stream);
    deviceInvalidateGraphs(g);
  }
  for (int L = 0; L < kMaxLevels; L++) freeLevel(g.mg[L]);
  if (g.dev) {
    Dev &d = dev(g);
    if (d.state) (void)hipFree(d.state);
    if (d.partials) (void)hipFree(d.partials);
    if (d.hist) (void)hipFree(d.hist);
    if (d.zeroFlag) (void)hipFree(d.zeroFlag);
    if (d.x64) (void)hipFree(d.x64);
    if (d.coarse64) (void)hipFree(d.coarse64);
    if (d.b0) (void)hipFree(d.b0);
    if (d.dActive0) (void)hipFree(d.dActive0);
    auto f = [](void *p) { if (p) (void)hipFree(p); };
    f(d.probOf); f(d.omegaOf);
    freeLearned(d.learned);
    freeFlip(d);
    if (d.pack32) (void)hipFree(d.pack32);
    if (d.pack16) (void)hipFree(d.pack16);
    if (d.pack16col) (void)hipFree(d.pack16col);
    if (d.pack16y) (void)hipFree(d.pack16y);
    if (BatchDev *b = d.batch) {
      f(b->cProb); f(b->cb); f(b->ce); f(b->cBeg); f(b->cEnd); f(b->states); f(b->allDone); f(b->hist);
      f(b->cProbL); f(b->cbL); f(b->ceL); f(b->cBegL); f(b->cEndL);
      delete b;
    }
    if (d.ev0) (void)hipEventDestroy(d.ev0);
    if (d.ev1) (void)hipEventDestroy(d.ev1);
    if (d.stream) (void)hipStreamDestroy(d.stream);
    delete &d;
    g.dev = nullptr;
  }
}

static LevelArgs levelArgs(const Grid &g, int L)
{
  const MgLevel &m = g.mg[L];
  LevelArgs a;
  a.level = m.d_level; a.off = m.d_off; a.flags = m.d_flags;
  a.W0 = m.d_W[0]; a.W1 = m.d_W[1]; a.W2 = m.d_W[2]; a.D = m.d_D;
  a.nbx = g.nbx; a.nby = g.nby; a.nbz = g.nbz; a.bsx0 = g.bsx0;
  for (int d = 0; d < 3; d++) a.bcW[d] = g.domainBc[d][1] != MSBGX_DBC_OPEN ? 0.0f : 1.0f;
  for (int k = 0; k < 4; k++) a.H[k] = m.d_H[k];
  a.prec = (g.prec != MSBGX_PREC_FP32 && m.d_H[0]) ? g.prec : MSBGX_PREC_FP32;
  const Dev *dd = (const Dev *)g.dev;
  a.nbd = m.d_nbd;
  a.probOf = dd ? dd->probOf : nullptr;
  a.omegaOf = dd ? dd->omegaOf : nullptr;
  a.pack = (L == 0 && g.bsx0 == 32 && dd) ? dd->pack32 : nullptr;
  a.pack16 = (L == 0 && g.bsx0 == 32 && dd && a.prec != MSBGX_PREC_FP32 && dd->pack16Prec == a.prec) ? dd->pack16
                                                                                                      : nullptr;
  a.pack16y = a.pack16 ? dd->pack16y : nullptr;
  a.pack16col = a.pack16 ? dd->pack16col : nullptr;
  return a;
}

/* Which channels exist at which MG level. */
static float *chanPtr(const Grid &g, int ch, int L)
{
  if (ch < 0 || ch >= N_VEC) return nullptr;
  return g.mg[L].d_ch[ch];
}

/* Device block lists of MG level L: act[lev] ..., relaxCL[c][lev] ... in one
 * buffer.  With `live` (batched block sets), only blocks of problems whose
 * live flag is set are kept; the buffer is sized for the full lists. */
static int uploadLists(Grid &g, int L, const std::vector<char> *live)
{
  MgLevel &m = g.mg[L];
  auto keep = [&](int b) { return !live || (*live)[g.probOf[b]]; };
  std::vector<int> all;
  std::vector<std::pair<int, int>> spans;
  size_t full = 0;
  auto add = [&](const std::vector<int> &v) {
    const int s0 = (int)all.size();
    for (int b : v)
      if (keep(b)) all.push_back(b);
    spans.push_back({s0, (int)all.size() - s0});
    full += v.size();
  };
  for (int l = 0; l < kMaxLevels; l++) add(m.act[l]);
  for (int c = 0; c < 2; c++)
    for (int l = 0; l < kMaxLevels; l++) add(m.relaxCL[c][l]);
  if (!m.d_lists)
    if (int rc = dalloc(&m.d_lists, full)) return rc;
  if (!all.empty()) HIPCHK(hipMemcpy(m.d_lists, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
  int si = 0;
  for (int l = 0; l < kMaxLevels; l++, si++) m.dAct[l] = DevList{m.d_lists + spans[si].first, spans[si].second};
  for (int c = 0; c < 2; c++)
    for (int l = 0; l < kMaxLevels; l++, si++) m.dRelax[c][l] = DevList{m.d_lists + spans[si].first, spans[si].second};
  return MSBGX_OK;
}

/* Coefficient records of the 32^3 column smoother over the MG0 pool,
 * rebuilt whenever the coefficients, their operand precision or the block
 * neighbourhood (batched block sets) change. */
static int buildPack32(Grid &g)
{
  if (g.bsx0 != 32 || !g.dev) return MSBGX_OK;
  Dev &d = dev(g);
  MgLevel &m = g.mg[0];
  if (!d.pack32)
    if (int rc = dalloc(&d.pack32, (size_t)m.ncells * 2)) return rc;
  std::vector<int> l0;
  for (int bid = 0; bid < g.nBlocks; bid++)
    if (m.level[bid] == 0) l0.push_back(bid);
  if (l0.empty()) return MSBGX_OK;
  int *dl = nullptr;
  if (int rc = dalloc(&dl, l0.size())) return rc;
  HIPCHK(hipMemcpy(dl, l0.data(), l0.size() * sizeof(int), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_pack_col32, dim3((unsigned)l0.size()), dim3(1024), 0, d.stream, levelArgs(g, 0), dl,
                     m.d_ch[C_X], d.pack32);
  /* 16-bit records for the register-resident smoother when the operands are 16-bit */
  const LevelArgs a0 = levelArgs(g, 0);
  d.pack16Prec = MSBGX_PREC_FP32;
  if (a0.prec != MSBGX_PREC_FP32) {
    const size_t nc = std::max<size_t>((size_t)m.ncells, 1);
    if (!d.pack16 && hipMalloc((void **)&d.pack16, nc * sizeof(uint2)) != hipSuccess) d.pack16 = nullptr;
    if (!d.pack16col && hipMalloc((void **)&d.pack16col, (nc / 32 + 1) * sizeof(uint2)) != hipSuccess)
      d.pack16col = nullptr;
    if (!d.pack16y && hipMalloc((void **)&d.pack16y, nc * sizeof(uint32_t)) != hipSuccess) d.pack16y = nullptr;
    if (!d.pack16 || !d.pack16col || !d.pack16y) {
      (void)hipFree(dl);
      setError("out of device memory for the 16-bit smoother records");
      return MSBGX_ENOMEM;
    }
    hipLaunchKernelGGL(k_pack_col32h, dim3((unsigned)l0.size()), dim3(1024), 0, d.stream, a0, dl, m.d_ch[C_X],
                       d.pack16, d.pack16col, d.pack16y);
    d.pack16Prec = a0.prec;
  }
  const hipError_t e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(d.stream);
  (void)hipFree(dl);
  HIPCHK(e);
  HIPCHK(e2);
  return MSBGX_OK;
}

/* Face neighbour of block bid across face (0..5 = -x, +x, -y, +y, -z, +z), -1
 * outside the domain: the batched neighbour table where the block set has
 * one, else the raster. */
static int hostFaceNeighbour(const Grid &g, int bid, int face)
{
  if (!g.nbrTable.empty()) return g.nbrTable[(size_t)bid * 6 + face];
  const int dir = face >> 1, side = face & 1;
  int x = bid % g.nbx, y = (bid / g.nbx) % g.nby, z = bid / (g.nbx * g.nby);
  if (dir == 0) x += side ? 1 : -1;
  else if (dir == 1) y += side ? 1 : -1;
  else z += side ? 1 : -1;
  return (x < 0 || y < 0 || z < 0 || x >= g.nbx || y >= g.nby || z >= g.nbz) ? -1 : x + g.nbx * (y + g.nby * z);
}

int deviceUpload(Grid &g, const std::vector<float> &rhs, const msbgx_scene *scene)
{
  if (!g.dev) {
    Dev *d = new Dev();
    g.dev = d;
    HIPCHK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&d->ev0));
    HIPCHK(hipEventCreate(&d->ev1));
    if (int rc = dalloc(&d->state, 1)) return rc;
    if (int rc = dalloc(&d->zeroFlag, 1)) return rc;
    HIPCHK(hipMemset(d->zeroFlag, 0, sizeof(int)));
  }
  Dev &d = dev(g);
  HIPCHK(hipStreamSynchronize(d.stream));
  deviceInvalidateGraphs(g);
  freeLearned(d.learned);
  d.learned = nullptr;
  freeFlip(d);
  for (int L = 0; L < kMaxLevels; L++) freeLevel(g.mg[L]);
  /* partials: one per block of a launch, per vector workgroup, and two sets per PCG block (k_update_xr_rz_mean) */
  int maxBlocksPerLaunch = std::max({g.nBlocks, 2 * kVecBlocks, 4096, 2 * g.nActive0});
  if (d.partialsCap < maxBlocksPerLaunch) {
    if (d.partials) (void)hipFree(d.partials);
    d.partials = nullptr;
    if (int rc = dalloc(&d.partials, maxBlocksPerLaunch)) return rc;
    d.partialsCap = maxBlocksPerLaunch;
  }
  for (int L = 0; L < g.nLevels; L++) {
    MgLevel &m = g.mg[L];
    const size_t nc = (size_t)m.ncells, nb = (size_t)g.nBlocks;
    if (int rc = dalloc(&m.d_level, nb)) return rc;
    if (int rc = dalloc(&m.d_off, nb)) return rc;
    if (int rc = dalloc(&m.d_flags, nc)) return rc;
    if (int rc = dalloc(&m.d_D, nc)) return rc;
    HIPCHK(hipMemcpy(m.d_level, m.level.data(), nb * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m.d_off, m.off.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice));
    {
      std::vector<NbDesc> nd(nb);
      for (int bid = 0; bid < g.nBlocks; bid++)
        for (int face = 0; face < 6; face++) {
          const int n = hostFaceNeighbour(g, bid, face);
          nd[bid].lev[face] = n < 0 ? -1 : m.level[n];
          nd[bid].off[face] = n < 0 ? 0 : m.off[n];
        }
      if (int rc = dalloc(&m.d_nbd, nb)) return rc;
      HIPCHK(hipMemcpy(m.d_nbd, nd.data(), nb * sizeof(NbDesc), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(m.d_flags, m.flags.data(), nc * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(m.d_D, m.D.data(), nc * sizeof(float), hipMemcpyHostToDevice));
    for (int k = 0; k < 3; k++) {
      if (int rc = dalloc(&m.d_W[k], nc)) return rc;
      HIPCHK(hipMemcpy(m.d_W[k], m.W[k].data(), nc * sizeof(float), hipMemcpyHostToDevice));
    }
    const int chans0[] = {C_X, C_B, C_R, C_Z, C_D, C_Q};
    const int chansL[] = {C_R, C_Z, C_Q};
    const int *cl = L == 0 ? chans0 : chansL;
    const int ncl = L == 0 ? 6 : 3;
    for (int i = 0; i < ncl; i++) {
      if (int rc = dalloc(&m.d_ch[cl[i]], nc)) return rc;
      HIPCHK(hipMemsetAsync(m.d_ch[cl[i]], 0, nc * sizeof(float), d.stream));
    }
    if (int rc = uploadLists(g, L, nullptr)) return rc;
  }
  d.b0Valid = false;
  if (d.dActive0) { (void)hipFree(d.dActive0); d.dActive0 = nullptr; }
  if (int rc = dalloc(&d.dActive0, g.active0.size())) return rc;
  if (!g.active0.empty())
    HIPCHK(hipMemcpy(d.dActive0, g.active0.data(), g.active0.size() * sizeof(int), hipMemcpyHostToDevice));
  if (g.prec != MSBGX_PREC_FP32)
    if (int rc = deviceSetPrecision(g, g.prec)) return rc;
  MgLevel &m0 = g.mg[0];
  if (scene) {
    /* level-0 blocks holding liquid are the head of the pool (act[0]) */
    if (m0.dAct[0].n > 0)
      hipLaunchKernelGGL(k_scene_rhs, dim3(m0.dAct[0].n), dim3(256), 0, d.stream, *scene, m0.dAct[0].ptr, m0.d_off,
                         g.nbx, g.nby, g.bsx0, m0.d_ch[C_B], 0);
    HIPCHK(hipGetLastError());
  } else if (!rhs.empty()) {
    HIPCHK(hipMemcpyAsync(m0.d_ch[C_B], rhs.data(), rhs.size() * sizeof(float), hipMemcpyHostToDevice, d.stream));
  }
  HIPCHK(hipStreamSynchronize(d.stream));
  if (d.pack32) { (void)hipFree(d.pack32); d.pack32 = nullptr; }   // the pool layout may have changed
  if (d.pack16) { (void)hipFree(d.pack16); d.pack16 = nullptr; }
  if (d.pack16col) { (void)hipFree(d.pack16col); d.pack16col = nullptr; }
  if (d.pack16y) { (void)hipFree(d.pack16y); d.pack16y = nullptr; }
  d.pack16Prec = MSBGX_PREC_FP32;
  return buildPack32(g);
}

/* ---------------------------------------------------------- launch helpers */

template <int B>
static void launchRelaxB(const Grid &g, int L, const DevList &l, float *X, const float *Bv, int reverse, const int *done,
                         hipStream_t s)
{
  if (l.n == 0) return;
  const LevelArgs a = levelArgs(g, L);
  if constexpr (B == 32) {
    /* 16-bit operands: the register-resident kernel, unless smoother variant 1
     * (the record-streaming kernel) was chosen for this block size */
    if (a.pack16 && g.smootherVariant[0] != 1) {
      if (a.prec == MSBGX_PREC_FP16)
        hipLaunchKernelGGL(k_relax_c32h<MSBGX_PREC_FP16>, dim3(l.n), dim3(C32::NT), 0, s, a, l.ptr, X, Bv, g.omega,
                           reverse, done);
      else
        hipLaunchKernelGGL(k_relax_c32h<MSBGX_PREC_BF16>, dim3(l.n), dim3(C32::NT), 0, s, a, l.ptr, X, Bv, g.omega,
                           reverse, done);
      return;
    }
    if (!a.pack) return;   // records are built at upload for every 32^3 MG0; never reached without them
  }
  if constexpr (B <= 16) {
    /* LDS-tile variant (FP32 operands only); the autotuner may pick it per block size */
    if (g.smootherVariant[__builtin_ctz(g.bsx0 / B)] == 1 && a.prec == MSBGX_PREC_FP32) {
      hipLaunchKernelGGL(k_relax<B>, dim3(l.n), dim3(Cfg<B>::NT), 0, s, a, l.ptr, X, Bv, g.omega, reverse, 5, done);
      return;
    }
  }
  hipLaunchKernelGGL(k_relax_col<B>, dim3(l.n), dim3(Cfg<B>::NT), 0, s, a, l.ptr, X, Bv, g.omega, reverse, done);
}

static void launchRelaxLev(const Grid &g, int L, int lev, const DevList &l, float *X, const float *Bv, int reverse,
                           const int *done, hipStream_t s)
{
  switch (g.bsx0 >> lev) {
    case 2: launchRelaxB<2>(g, L, l, X, Bv, reverse, done, s); break;
    case 4: launchRelaxB<4>(g, L, l, X, Bv, reverse, done, s); break;
    case 8: launchRelaxB<8>(g, L, l, X, Bv, reverse, done, s); break;
    case 16: launchRelaxB<16>(g, L, l, X, Bv, reverse, done, s); break;
    case 32: launchRelaxB<32>(g, L, l, X, Bv, reverse, done, s); break;
  }
}

/* relax(): nIter outer iterations of two colour passes (src/msbg3.cpp:2245-2369) */
static void relaxLaunch(const Grid &g, int L, bool reverse, int nIter, float *X, const float *Bv, const int *done,
                        hipStream_t s)
{
  const MgLevel &m = g.mg[L];
  if (!m.anyRelax) return;
  for (int it = 0; it < nIter; it++)
    for (int pass = 0; pass < 2; pass++) {
      const int parity = reverse ? pass : 1 - pass;
      for (int lev = L; lev < g.nLevels; lev++)
        launchRelaxLev(g, L, lev, m.dRelax[parity][lev], X, Bv, reverse ? 1 : 0, done, s);
    }
}

/* The coarsest level's relax in FP64 (precision form): nIter outer
 * iterations of two colour passes over the relaxation lists, x in X64. */
static void relax64Launch(const Grid &g, int L, int nIter, double *X64, const float *Bv, const int *done,
                          hipStream_t s)
{
  const MgLevel &m = g.mg[L];
  if (!m.anyRelax) return;
  const LevelArgs a = levelArgs(g, L);
  for (int it = 0; it < nIter; it++)
    for (int pass = 0; pass < 2; pass++) {
      const DevList &l = m.dRelax[1 - pass][L];
      if (l.n == 0) continue;
      switch (g.bsx0 >> L) {
        case 2: hipLaunchKernelGGL(k_relax64<2>, dim3(l.n), dim3(Cfg<2>::NT), 0, s, a, l.ptr, X64, Bv, g.omega, 0, 5, done); break;
        case 4: hipLaunchKernelGGL(k_relax64<4>, dim3(l.n), dim3(Cfg<4>::NT), 0, s, a, l.ptr, X64, Bv, g.omega, 0, 5, done); break;
        case 8: hipLaunchKernelGGL(k_relax64<8>, dim3(l.n), dim3(Cfg<8>::NT), 0, s, a, l.ptr, X64, Bv, g.omega, 0, 5, done); break;
        case 16: hipLaunchKernelGGL(k_relax64<16>, dim3(l.n), dim3(Cfg<16>::NT), 0, s, a, l.ptr, X64, Bv, g.omega, 0, 5, done); break;
      }
    }
}

template <int B>
static void launchMultiplyB(const Grid &g, int L, const DevList &l, const float *X, const float *Bv, float *Y, int residual,
                            double *partials, int base, const int *done, hipStream_t s)
{
  if (l.n == 0) return;
  if constexpr (B == 32) {
    MulVecs<1> mv;
    mv.x[0] = X;
    mv.y[0] = Y;
    hipLaunchKernelGGL((k_multiply_z<32, 1>), dim3(l.n), dim3(32 * 32), 0, s, levelArgs(g, L), l.ptr, mv, Bv,
                       residual, partials, base, done);
  } else
    hipLaunchKernelGGL(k_multiply<B>, dim3(l.n), dim3(Cfg<B>::NT), 0, s, levelArgs(g, L), l.ptr, X, Bv, Y, residual,
                       partials, base, done);
}

/* multiplyLaplacianMatrixOpt over all non-NO_FLUID blocks of MG level L;
 * when partials is given, FP64 partials of <x, y> over level-0 blocks are
 * written at partials[0 .. returned count). */
static int multiplyLaunch(const Grid &g, int L, bool residual, const float *X, const float *Bv, float *Y,
                          double *partials, const int *done, hipStream_t s)
{
  const MgLevel &m = g.mg[L];
  int base = 0;
  for (int lev = L; lev < g.nLevels; lev++) {
    const DevList &l = m.dAct[lev];
    switch (g.bsx0 >> lev) {
      case 2: launchMultiplyB<2>(g, L, l, X, Bv, Y, residual, partials, base, done, s); break;
      case 4: launchMultiplyB<4>(g, L, l, X, Bv, Y, residual, partials, base, done, s); break;
      case 8: launchMultiplyB<8>(g, L, l, X, Bv, Y, residual, partials, base, done, s); break;
      case 16: launchMultiplyB<16>(g, L, l, X, Bv, Y, residual, partials, base, done, s); break;
      case 32: launchMultiplyB<32>(g, L, l, X, Bv, Y, residual, partials, base, done, s); break;
    }
    base += l.n;
  }
  return base;
}

static void zeroActive(const Grid &g, int L, float *X, const int *done, hipStream_t s)
{
  const MgLevel &m = g.mg[L];
  for (int lev = L; lev < g.nLevels; lev++)
    if (m.dAct[lev].n)
      hipLaunchKernelGGL(k_zero_blocks, dim3(m.dAct[lev].n), dim3(256), 0, s, m.dAct[lev].ptr, m.d_level, m.d_off,
                         g.bsx0, X, done);
}

static void restrictLaunch(const Grid &g, int Lc, const float *src, float *dst, int copyMode, const int *done,
                           hipStream_t s)
{
  const MgLevel &c = g.mg[Lc], &f = g.mg[Lc - 1];
  const int *level0 = g.mg[0].d_level;
  for (int lev = Lc; lev < g.nLevels; lev++)
    if (c.dAct[lev].n)
      hipLaunchKernelGGL(k_restrict, dim3(c.dAct[lev].n), dim3(256), 0, s, c.dAct[lev].ptr, level0, c.d_level, c.d_off,
                         f.d_off, Lc, g.bsx0, src, dst, copyMode, done);
}

static void prolongLaunch(const Grid &g, int Lf, float *xf, const float *xc, const int *done, hipStream_t s)
{
  const MgLevel &f = g.mg[Lf], &c = g.mg[Lf + 1];
  const int *level0 = g.mg[0].d_level;
  for (int lev = Lf; lev < g.nLevels; lev++)
    if (f.dAct[lev].n)
      hipLaunchKernelGGL(k_prolong, dim3(f.dAct[lev].n), dim3(256), 0, s, f.dAct[lev].ptr, level0, f.d_level, f.d_off,
                         c.d_off, f.d_flags, Lf, g.bsx0, xf, xc, done);
}

/* True when every MG level below L has an empty relaxation list: the coarse
 * correction of vCycle(L+1) is then exactly zero. */
static bool coarseIsInert(const Grid &g, int L)
{
  for (int l = L + 1; l < g.nLevels; l++)
    if (g.mg[l].anyRelax) return false;
  return true;
}

/* vCycle (src/msbg3.cpp:2412-2465) on channels z (x), r (rhs), q (residual). */
static void vcycleLaunch(Grid &g, int L, int nPre, int nPost, int nCoarse, const int *done, hipStream_t s)
{
  MgLevel &m = g.mg[L];
  float *X = m.d_ch[C_Z], *Bv = m.d_ch[C_R], *R = m.d_ch[C_Q];
  if (L >= g.nLevels - 1) {
    Dev &d = dev(g);
    if (g.prec != MSBGX_PREC_FP32 && d.coarse64 && (g.bsx0 >> L) <= 16) {
      /* precision form: the coarsest solve in FP64, from x = 0 as the zeroed channel */
      hipLaunchKernelGGL(k_fill64, dim3(256), dim3(256), 0, s, d.coarse64, m.ncells, done);
      relax64Launch(g, L, nCoarse, d.coarse64, Bv, done, s);
      hipLaunchKernelGGL(k_x64_to_f32, dim3(256), dim3(256), 0, s, d.coarse64, X, m.ncells, done);
      return;
    }
    relaxLaunch(g, L, false, nCoarse, X, Bv, done, s);
    return;
  }
  relaxLaunch(g, L, false, nPre, X, Bv, done, s);
  const bool inert = coarseIsInert(g, L) && !(g.options & MSBGX_OPT_LITERAL_VCYCLE);
  if (!inert) {
    multiplyLaunch(g, L, true, X, Bv, R, nullptr, done, s);
    restrictLaunch(g, L + 1, R, g.mg[L + 1].d_ch[C_R], g.aliasEstablished[L + 1] ? 0 : 1, done, s);
    g.aliasEstablished[L + 1] = true;
    zeroActive(g, L + 1, g.mg[L + 1].d_ch[C_Z], done, s);
    vcycleLaunch(g, L + 1, nPre, nPost, nCoarse, done, s);
    prolongLaunch(g, L, X, g.mg[L + 1].d_ch[C_Z], done, s);
  } else {
    /* the reference still creates the coarse right-hand-side block references */
    for (int l = L + 1; l < g.nLevels; l++) g.aliasEstablished[l] = true;
  }
  relaxLaunch(g, L, true, nPost, X, Bv, done, s);
}

/* ------------------------------------------------------------------- PCG */

static void finalize(Dev &d, int n, int op, double *out)
{
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, d.stream, d.state, d.partials, n, op, out);
}

/* finalize over the partials at d.partials + first */
static void finalizeAt(Dev &d, int first, int n, int op, double *out)
{
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, d.stream, d.state, d.partials + first, n, op, out);
}

/* One PCG iteration (src/msbg3.cpp:3149-3236), every kernel a no-op once
 * the device-side done flag is set. */
static void pcgIteration(Grid &g, const msbgx_solve_params &p)
{
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  const int *done = &d.state->done;
  const int64_t n0 = (int64_t)g.nActive0 * g.bsx0 * g.bsx0 * g.bsx0;
  const int nPart = multiplyLaunch(g, 0, false, m0.d_ch[C_D], nullptr, m0.d_ch[C_Q], d.partials, done, d.stream);
  finalize(d, nPart, FIN_ALPHA, nullptr);
  hipLaunchKernelGGL(k_update_xr, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_X], m0.d_ch[C_R],
                     m0.d_ch[C_D], m0.d_ch[C_Q], n0, d.state, d.partials);
  finalize(d, kVecBlocks, FIN_CONV, d.hist);
  zeroActive(g, 0, m0.d_ch[C_Z], done, d.stream);
  vcycleLaunch(g, 0, p.n_pre, p.n_post, p.n_coarse, done, d.stream);
  hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], m0.d_ch[C_Z], n0, d.partials,
                     done);
  finalize(d, kVecBlocks, FIN_BETA, nullptr);
  hipLaunchKernelGGL(k_axpy_d, dim3(kVecBlocksNoSum), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_D], m0.d_ch[C_Z], n0,
                     d.state);
}

int deviceSolve(Grid &g, const msbgx_solve_params &p, int *iters, float *rel)
{
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  const int64_t n0 = (int64_t)g.nActive0 * g.bsx0 * g.bsx0 * g.bsx0;
  if (int rc = ensureHist(g, p.max_iter + 1)) return rc;
  PcgState init{};
  init.tolSq = (double)p.tol * (double)p.tol;
  init.maxIter = p.max_iter;
  HIPCHK(hipMemcpyAsync(d.state, &init, sizeof(init), hipMemcpyHostToDevice, d.stream));
  HIPCHK(hipEventRecord(d.ev0, d.stream));

  /* r = b and |b|^2 over the level-0 blocks; r = b elsewhere is zero */
  hipLaunchKernelGGL(k_copy_norm, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_B], m0.d_ch[C_R], n0,
                     d.partials);
  finalize(d, kVecBlocks, FIN_BNORM, nullptr);
  const int *done = &d.state->done;
  /* z = M^-1 r; d = z on every block of MG0; rho = <r, z> */
  zeroActive(g, 0, m0.d_ch[C_Z], done, d.stream);
  vcycleLaunch(g, 0, p.n_pre, p.n_post, p.n_coarse, done, d.stream);
  for (int lev = 0; lev < g.nLevels; lev++)
    if (m0.dAct[lev].n)
      hipLaunchKernelGGL(k_copy_blocks, dim3(m0.dAct[lev].n), dim3(256), 0, d.stream, m0.dAct[lev].ptr, m0.d_level,
                         m0.d_off, g.bsx0, m0.d_ch[C_Z], m0.d_ch[C_D], done);
  hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], m0.d_ch[C_Z], n0, d.partials,
                     done);
  finalize(d, kVecBlocks, FIN_RHO0, nullptr);
  HIPCHK(hipGetLastError());

  const bool useGraph = !(g.options & MSBGX_OPT_NO_GRAPH);
  const GraphKey key = graphKey(g, p);
  const int chunk = key.chunk;
  if (d.graph && d.graphKey != key) { (void)hipGraphExecDestroy(d.graph); d.graph = nullptr; }
  PcgState hs{};
  for (int done_it = 0; done_it < p.max_iter;) {
    const int n = std::min(chunk, p.max_iter - done_it);
    if (useGraph && n == chunk) {
      if (!d.graph) {
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(d.stream, hipStreamCaptureModeRelaxed));
        for (int i = 0; i < chunk; i++) pcgIteration(g, p);
        HIPCHK(hipStreamEndCapture(d.stream, &graph));
        HIPCHK(hipGraphInstantiate(&d.graph, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        d.graphKey = key;
      }
      HIPCHK(hipGraphLaunch(d.graph, d.stream));
    } else {
      for (int i = 0; i < n; i++) pcgIteration(g, p);
    }
    done_it += n;
    HIPCHK(hipMemcpyAsync(&hs, d.state, sizeof(hs), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    if (hs.done) break;
  }
  /* final |r| */
  hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], m0.d_ch[C_R], n0, d.partials,
                     d.zeroFlag);
  finalize(d, kVecBlocks, FIN_FINAL, d.hist + p.max_iter);
  HIPCHK(hipEventRecord(d.ev1, d.stream));
  HIPCHK(hipMemcpyAsync(&hs, d.state, sizeof(hs), hipMemcpyDeviceToHost, d.stream));
  double fin = 0.0;
  HIPCHK(hipMemcpyAsync(&fin, d.hist + p.max_iter, sizeof(double), hipMemcpyDeviceToHost, d.stream));
  HIPCHK(hipStreamSynchronize(d.stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, d.ev0, d.ev1));
  g.lastSolveMs = ms;
  g.history.assign(hs.iter, 0.0);
  if (hs.iter > 0) {
    std::vector<double> rr(hs.iter);
    HIPCHK(hipMemcpy(rr.data(), d.hist, hs.iter * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < hs.iter; i++) g.history[i] = std::sqrt(rr[i] / hs.bnorm2);
  }
  if (hs.brk == 3) {
    setError("non-finite right-hand side");
    return MSBGX_EINVAL;
  }
  if (hs.brk == 2) {  // |b|^2 <= 1e-30
    *iters = 0;
    *rel = 0.0f;
    return MSBGX_OK;
  }
  *rel = (float)std::sqrt(fin / hs.bnorm2);
  *iters = hs.conv > 0 ? hs.conv : p.max_iter;
  return MSBGX_OK;
}

/* ------------------------------------------------------ per-kernel hooks */

static int needChan(const Grid &g, int ch, int L, float **p)
{
  *p = chanPtr(g, ch, L);
  if (!*p) {
    setError("channel does not exist at this MG level");
    return MSBGX_EINVAL;
  }
  return MSBGX_OK;
}

int deviceGetBlock(const Grid &g, int ch, int L, int bid, void *dst)
{
  const MgLevel &m = g.mg[L];
  const int b = g.bsxL(m.level[bid]);
  const size_t n = (size_t)b * b * b;
  const int64_t o = m.off[bid];
  const void *src = nullptr;
  size_t es = sizeof(float);
  if (ch == MSBGX_CH_FLAGS) { src = m.d_flags + o; es = sizeof(uint16_t); }
  else if (ch >= MSBGX_CH_FACE_U && ch <= MSBGX_CH_FACE_W) src = m.d_W[ch - MSBGX_CH_FACE_U] + o;
  else if (ch == MSBGX_CH_DIAGONAL) src = m.d_D + o;
  else {
    float *p = chanPtr(g, ch, L);
    if (!p) return 0;
    src = p + o;
  }
  HIPCHK(hipMemcpy(dst, src, n * es, hipMemcpyDeviceToHost));
  return b;
}

int deviceSetBlock(Grid &g, int ch, int L, int bid, const float *src)
{
  float *p;
  if (int rc = needChan(g, ch, L, &p)) return rc;
  const MgLevel &m = g.mg[L];
  const int b = g.bsxL(m.level[bid]);
  HIPCHK(hipMemcpy(p + m.off[bid], src, (size_t)b * b * b * sizeof(float), hipMemcpyHostToDevice));
  if (ch == C_B && L == 0) dev(g).b0Valid = false;   // the refinement's saved right-hand side is stale
  return b;
}

int deviceRelax(Grid &g, int L, bool reverse, int nIter, int chX, int chB)
{
  float *X, *Bv;
  if (int rc = needChan(g, chX, L, &X)) return rc;
  if (int rc = needChan(g, chB, L, &Bv)) return rc;
  Dev &d = dev(g);
  relaxLaunch(g, L, reverse, nIter, X, Bv, d.zeroFlag, d.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

int deviceMultiply(Grid &g, int L, bool residual, int chX, int chB, int chDst)
{
  float *X, *Y, *Bv = nullptr;
  if (int rc = needChan(g, chX, L, &X)) return rc;
  if (int rc = needChan(g, chDst, L, &Y)) return rc;
  if (residual)
    if (int rc = needChan(g, chB, L, &Bv)) return rc;
  Dev &d = dev(g);
  multiplyLaunch(g, L, residual, X, Bv, Y, nullptr, d.zeroFlag, d.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

int deviceDownsample(Grid &g, int Lc, int chSrc, int chDst)
{
  float *S, *D;
  if (int rc = needChan(g, chSrc, Lc - 1, &S)) return rc;
  if (int rc = needChan(g, chDst, Lc, &D)) return rc;
  Dev &d = dev(g);
  restrictLaunch(g, Lc, S, D, g.aliasEstablished[Lc] ? 0 : 1, d.zeroFlag, d.stream);
  g.aliasEstablished[Lc] = true;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

int deviceProlongate(Grid &g, int Lf, int ch)
{
  float *F, *C;
  if (int rc = needChan(g, ch, Lf, &F)) return rc;
  if (int rc = needChan(g, ch, Lf + 1, &C)) return rc;
  Dev &d = dev(g);
  prolongLaunch(g, Lf, F, C, d.zeroFlag, d.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

int deviceZero(Grid &g, int ch, int L)
{
  float *X;
  if (int rc = needChan(g, ch, L, &X)) return rc;
  Dev &d = dev(g);
  zeroActive(g, L, X, d.zeroFlag, d.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

int deviceCopy(Grid &g, int src, int dst)
{
  float *S, *D;
  if (int rc = needChan(g, src, 0, &S)) return rc;
  if (int rc = needChan(g, dst, 0, &D)) return rc;
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  for (int lev = 0; lev < g.nLevels; lev++)
    if (m0.dAct[lev].n)
      hipLaunchKernelGGL(k_copy_blocks, dim3(m0.dAct[lev].n), dim3(256), 0, d.stream, m0.dAct[lev].ptr, m0.d_level,
                         m0.d_off, g.bsx0, S, D, d.zeroFlag);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

int deviceVCycle(Grid &g, int nPre, int nPost, int nCoarse)
{
  Dev &d = dev(g);
  zeroActive(g, 0, g.mg[0].d_ch[C_Z], d.zeroFlag, d.stream);
  vcycleLaunch(g, 0, nPre, nPost, nCoarse, d.zeroFlag, d.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

int deviceTimeKernel(Grid &g, int which, int reps, double *ms, double *bytes, int *cells)
{
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  const int *nd = d.zeroFlag;
  /* the smoother: the block size with the most relaxation cells at MG0, its
   * two colour lists launched in turn as a V-cycle does (one launch = one
   * colour pass) */
  int bestL = 0, bestN = -1;
  for (int l = 0; l < g.nLevels; l++) {
    const int b = g.bsxL(l);
    const int n = (m0.dRelax[0][l].n + m0.dRelax[1][l].n) * b * b * b;
    if (n > bestN) { bestN = n; bestL = l; }
  }
  int colour = 0;
  auto launch = [&]() {
    if (which == 0) {
      launchRelaxLev(g, 0, bestL, m0.dRelax[colour][bestL], m0.d_ch[C_Z], m0.d_ch[C_R], 0, nd, d.stream);
      colour ^= 1;
    } else if (which == 1) {
      const DevList &l = m0.dAct[0];
      switch (g.bsx0) {
        case 4: launchMultiplyB<4>(g, 0, l, m0.d_ch[C_Z], m0.d_ch[C_R], m0.d_ch[C_Q], 1, nullptr, 0, nd, d.stream); break;
        case 8: launchMultiplyB<8>(g, 0, l, m0.d_ch[C_Z], m0.d_ch[C_R], m0.d_ch[C_Q], 1, nullptr, 0, nd, d.stream); break;
        case 16: launchMultiplyB<16>(g, 0, l, m0.d_ch[C_Z], m0.d_ch[C_R], m0.d_ch[C_Q], 1, nullptr, 0, nd, d.stream); break;
        case 32: launchMultiplyB<32>(g, 0, l, m0.d_ch[C_Z], m0.d_ch[C_R], m0.d_ch[C_Q], 1, nullptr, 0, nd, d.stream); break;
      }
    } else {
      const int64_t n0 = (int64_t)g.nActive0 * g.bsx0 * g.bsx0 * g.bsx0;
      hipLaunchKernelGGL(k_update_xr, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_X], m0.d_ch[C_R],
                         m0.d_ch[C_D], m0.d_ch[C_Q], n0, d.state, d.partials);
    }
  };
  PcgState init{};
  init.alpha = 1e-3f;
  HIPCHK(hipMemcpyAsync(d.state, &init, sizeof(init), hipMemcpyHostToDevice, d.stream));
  launch();  // warm-up (both colours for the smoother)
  if (which == 0) launch();
  if (which == 0 && (reps & 1)) reps++;   // whole colour pairs
  HIPCHK(hipEventRecord(d.ev0, d.stream));
  for (int i = 0; i < reps; i++) launch();
  HIPCHK(hipEventRecord(d.ev1, d.stream));
  HIPCHK(hipEventSynchronize(d.ev1));
  HIPCHK(hipGetLastError());
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, d.ev0, d.ev1));
  *ms = (double)t / reps;
  int64_t c;
  double bpc;
  /* smoother: cells per launch averaged over the two colours; algorithmic
   * bytes per cell 30 with FP32 operands (flags 2, D and W_x,y,z 16, x read
   * and written 8, b 4), 22 with 16-bit ones (D and W 8) */
  if (which == 0) { c = bestN / 2; bpc = g.prec == MSBGX_PREC_FP32 ? 30.0 : 22.0; }
  else if (which == 1) { c = (int64_t)m0.dAct[0].n * g.bsx0 * g.bsx0 * g.bsx0; bpc = 30.0; }
  else { c = (int64_t)g.nActive0 * g.bsx0 * g.bsx0 * g.bsx0; bpc = 24.0; }
  *cells = (int)c;
  *bytes = bpc * (double)c;
  return MSBGX_OK;
}

/* -------------------------------------------------------- precision form */

/* Round-to-nearest-even conversion of FP32 coefficients to FP16 or BF16 bits. */
__global__ void k_to16(const float *__restrict__ src, uint16_t *__restrict__ dst, int64_t n, int prec)
{
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float v = src[i];
    uint16_t o;
    if (prec == MSBGX_PREC_FP16) o = __builtin_bit_cast(uint16_t, (_Float16)v);   // IEEE round-to-nearest-even
    else o = bf16Bits(v);
    dst[i] = o;
  }
}

int deviceSetPrecision(Grid &g, int prec)
{
  if (!g.dev) return MSBGX_OK;
  Dev &d = dev(g);
  for (int L = 0; L < g.nLevels; L++) {
    MgLevel &m = g.mg[L];
    if (prec == MSBGX_PREC_FP32) continue;
    const float *src[4] = {m.d_W[0], m.d_W[1], m.d_W[2], m.d_D};
    for (int k = 0; k < 4; k++) {
      if (!m.d_H[k])
        if (int rc = dalloc(&m.d_H[k], (size_t)m.ncells)) return rc;
      hipLaunchKernelGGL(k_to16, dim3(1024), dim3(256), 0, d.stream, src[k], m.d_H[k], m.ncells, prec);
    }
  }
  if (prec != MSBGX_PREC_FP32) {
    const int64_t nc = g.mg[g.nLevels - 1].ncells;
    if (d.coarse64Cap < nc) {
      if (d.coarse64) (void)hipFree(d.coarse64);
      d.coarse64 = nullptr;
      d.coarse64Cap = 0;
      if (int rc = dalloc(&d.coarse64, (size_t)std::max<int64_t>(nc, 1))) return rc;
      d.coarse64Cap = nc;
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  if (int rc = buildPack32(g)) return rc;   // records hold the operands in the new precision
  deviceInvalidateGraphs(g);   // kernel arguments changed
  return MSBGX_OK;
}

/* r = b0 - A x64 in FP64 over the level-0 unknowns (coarse-block and
 * non-PCG blocks hold 0), written as float into the right-hand-side channel,
 * with FP64 partials of |r|^2.  A x = (1/D) x - sum W x_nb over same-level
 * face neighbours (src/msbg3.cpp:1282-1291 evaluated in FP64). */
__global__ void __launch_bounds__(256) k_residual64(LevelArgs a, const int *__restrict__ blocks, int64_t n0,
                                                   const double *__restrict__ x64, const float *__restrict__ b0,
                                                   float *__restrict__ rout, double *partials)
{
  __shared__ double red[4];
  const int bid = blocks[blockIdx.x];
  const int B = a.bsx0, B2 = B * B, n = B2 * B;
  const int64_t off = a.off[bid];
  const int bx = bid % a.nbx, by = (bid / a.nbx) % a.nby, bz = bid / (a.nbx * a.nby);
  auto xAt = [&](int nb, int v) -> double {
    const int64_t o = a.off[nb];
    return (a.level[nb] == 0 && o < n0) ? x64[o + v] : 0.0;
  };
  double acc = 0.0;
  for (int v = threadIdx.x; v < n; v += blockDim.x) {
    const int vx = v % B, vy = (v / B) % B, vz = v / B2;
    const int64_t c = off + v;
    const float D = a.D[c];
    double y = 0.0;
    if (D != 0.0f) {
      y = x64[c] / (double)D;
      /* -x, +x, -y, +y, -z, +z */
      double s = 0.0;
      if (vx > 0) s += (double)a.W0[c] * x64[c - 1];
      else if (bx > 0) s += (double)a.W0[c] * xAt(bid - 1, v + B - 1);
      if (vx < B - 1) s += (double)a.W0[c + 1] * x64[c + 1];
      else if (bx < a.nbx - 1 && a.level[bid + 1] == 0) s += (double)a.W0[a.off[bid + 1] + v - (B - 1)] * xAt(bid + 1, v - (B - 1));
      if (vy > 0) s += (double)a.W1[c] * x64[c - B];
      else if (by > 0) s += (double)a.W1[c] * xAt(bid - a.nbx, v + B2 - B);
      if (vy < B - 1) s += (double)a.W1[c + B] * x64[c + B];
      else if (by < a.nby - 1 && a.level[bid + a.nbx] == 0) s += (double)a.W1[a.off[bid + a.nbx] + v - (B2 - B)] * xAt(bid + a.nbx, v - (B2 - B));
      const int nxy = a.nbx * a.nby;
      if (vz > 0) s += (double)a.W2[c] * x64[c - B2];
      else if (bz > 0) s += (double)a.W2[c] * xAt(bid - nxy, v + n - B2);
      if (vz < B - 1) s += (double)a.W2[c + B2] * x64[c + B2];
      else if (bz < a.nbz - 1 && a.level[bid + nxy] == 0) s += (double)a.W2[a.off[bid + nxy] + v - (n - B2)] * xAt(bid + nxy, v - (n - B2));
      y -= s;
    }
    const double r = (double)b0[c] - y;
    rout[c] = (float)r;
    acc += r * r;
  }
  const double t = blockReduceSum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

/* x64 += x (and x = (float) x64 when store) over [0, n) */
__global__ void k_accum64(double *__restrict__ x64, float *__restrict__ x, int64_t n, int store)
{
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (store) x[i] = (float)x64[i];
    else x64[i] += (double)x[i];
  }
}

int deviceSolveRefined(Grid &g, const msbgx_solve_params &p, double tol, int maxOuter, int *outer, int *inner,
                       double *rel, double *x64out)
{
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  const int64_t n0 = (int64_t)g.nActive0 * g.bsx0 * g.bsx0 * g.bsx0;
  if (d.refCap < n0) {
    if (d.x64) (void)hipFree(d.x64);
    if (d.b0) (void)hipFree(d.b0);
    d.x64 = nullptr; d.b0 = nullptr;
    if (int rc = dalloc(&d.x64, n0)) return rc;
    if (int rc = dalloc(&d.b0, n0)) return rc;
    d.refCap = n0;
    d.b0Valid = false;
  }
  if (!d.b0Valid) {
    HIPCHK(hipMemcpyAsync(d.b0, m0.d_ch[C_B], n0 * sizeof(float), hipMemcpyDeviceToDevice, d.stream));
    d.b0Valid = true;
  }
  HIPCHK(hipMemsetAsync(d.x64, 0, n0 * sizeof(double), d.stream));
  hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, d.b0, d.b0, n0, d.partials, d.zeroFlag);
  if (int rc = ensureHist(g, 2)) return rc;
  finalize(d, kVecBlocks, FIN_FINAL, d.hist);
  double bb = 0.0;
  HIPCHK(hipMemcpyAsync(&bb, d.hist, sizeof(double), hipMemcpyDeviceToHost, d.stream));
  HIPCHK(hipStreamSynchronize(d.stream));
  const LevelArgs a = levelArgs(g, 0);
  int nOuter = 0, nInner = 0;
  double relNow = 0.0;
  std::vector<double> outerHist;
  for (;;) {
    if (g.nActive0 > 0)
      hipLaunchKernelGGL(k_residual64, dim3(g.nActive0), dim3(256), 0, d.stream, a, d.dActive0, n0, d.x64, d.b0,
                         m0.d_ch[C_B], d.partials);
    finalize(d, std::max(g.nActive0, 1), FIN_FINAL, d.hist);
    double rr = 0.0;
    HIPCHK(hipMemcpyAsync(&rr, d.hist, sizeof(double), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    relNow = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
    outerHist.push_back(relNow);
    if (relNow <= tol || nOuter >= maxOuter || bb <= 1e-30) break;
    zeroActive(g, 0, m0.d_ch[C_X], d.zeroFlag, d.stream);
    int it = 0;
    float rIn = 0.f;
    if (int rc = deviceSolve(g, p, &it, &rIn)) return rc;
    nInner += it;
    nOuter++;
    hipLaunchKernelGGL(k_accum64, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, d.x64, m0.d_ch[C_X], n0, 0);
  }
  hipLaunchKernelGGL(k_accum64, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, d.x64, m0.d_ch[C_X], n0, 1);
  HIPCHK(hipMemcpyAsync(m0.d_ch[C_B], d.b0, n0 * sizeof(float), hipMemcpyDeviceToDevice, d.stream));
  if (x64out) HIPCHK(hipMemcpyAsync(x64out, d.x64, n0 * sizeof(double), hipMemcpyDeviceToHost, d.stream));
  HIPCHK(hipStreamSynchronize(d.stream));
  HIPCHK(hipGetLastError());
  g.history = outerHist;
  *outer = nOuter;
  *inner = nInner;
  *rel = relNow;
  return MSBGX_OK;
}

/* ------------------------------------------------------ batched block sets */

/* Chunked PCG vector kernels: workgroup c covers cells [beg[c], end[c]) of
 * problem prob[c]; one FP64 partial per chunk (fixed order per problem). */
__global__ void __launch_bounds__(kVecThreads) k_copy_norm_c(const float *__restrict__ b, float *__restrict__ r,
                                                             const int64_t *beg, const int64_t *end, double *partials)
{
  __shared__ double red[kVecThreads / 64];
  double acc = 0.0;
  for (int64_t i = beg[blockIdx.x] + threadIdx.x; i < end[blockIdx.x]; i += blockDim.x) {
    const float v = b[i];
    r[i] = v;
    acc += (double)v * (double)v;
  }
  const double t = blockReduceSum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kVecThreads) k_dot_c(const float *__restrict__ a, const float *__restrict__ b,
                                                       const int64_t *beg, const int64_t *end, double *partials)
{
  __shared__ double red[kVecThreads / 64];
  double acc = 0.0;
  for (int64_t i = beg[blockIdx.x] + threadIdx.x; i < end[blockIdx.x]; i += blockDim.x)
    acc += (double)a[i] * (double)b[i];
  const double t = blockReduceSum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kVecThreads) k_update_xr_c(float *__restrict__ x, float *__restrict__ r,
                                                             const float *__restrict__ d, const float *__restrict__ q,
                                                             const int *prob, const int64_t *beg, const int64_t *end,
                                                             const PcgState *st, double *partials)
{
  __shared__ double red[kVecThreads / 64];
  const PcgState &s = st[prob[blockIdx.x]];
  double acc = 0.0;
  if (!s.done) {
    const float alpha = s.alpha;
    for (int64_t i = beg[blockIdx.x] + threadIdx.x; i < end[blockIdx.x]; i += blockDim.x) {
      x[i] = fmaf(alpha, d[i], x[i]);
      const float rv = fmaf(-alpha, q[i], r[i]);
      r[i] = rv;
      acc += (double)rv * (double)rv;
    }
  }
  const double t = blockReduceSum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kVecThreads) k_axpy_c(float *__restrict__ d, const float *__restrict__ z,
                                                        const int *prob, const int64_t *beg, const int64_t *end,
                                                        const PcgState *st)
{
  const PcgState &s = st[prob[blockIdx.x]];
  if (s.done) return;
  const float beta = s.beta;
  for (int64_t i = beg[blockIdx.x] + threadIdx.x; i < end[blockIdx.x]; i += blockDim.x) d[i] = fmaf(beta, d[i], z[i]);
}

static BatchDev *batchDev(Grid &g) { return ((Dev *)g.dev)->batch; }

int deviceUploadBatch(Grid &m, const std::vector<const Grid *> &probs)
{
  std::vector<float> noRhs;
  if (int rc = deviceUpload(m, noRhs, nullptr)) return rc;
  Dev &d = dev(m);
  const int nb = m.nBlocks;
  if (int rc = dalloc(&d.probOf, nb)) return rc;
  if (int rc = dalloc(&d.omegaOf, m.nProb)) return rc;
  HIPCHK(hipMemcpy(d.probOf, m.probOf.data(), nb * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d.omegaOf, m.probOmega.data(), m.nProb * sizeof(float), hipMemcpyHostToDevice));
  /* right-hand sides: scene problems on the device, others from their host copy */
  MgLevel &m0 = m.mg[0];
  std::vector<float> hostB;
  bool anyHost = false;
  for (int q = 0; q < m.nProb; q++) anyHost |= !probs[q]->hasScene;
  if (anyHost) hostB.assign(m0.ncells, 0.0f);
  int *dl = nullptr;
  if (int rc = dalloc(&dl, nb)) return rc;
  for (int q = 0; q < m.nProb; q++) {
    const Grid &p = *probs[q];
    const int s = m.probBlock0[q];
    if (p.hasScene) {
      std::vector<int> l;
      for (int b : p.mg[0].act[0]) l.push_back(s + b);
      if (l.empty()) continue;
      HIPCHK(hipMemcpy(dl, l.data(), l.size() * sizeof(int), hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_scene_rhs, dim3((int)l.size()), dim3(256), 0, d.stream, p.scene, dl, m0.d_off, p.nbx, p.nby,
                         m.bsx0, m0.d_ch[C_B], s);
      HIPCHK(hipStreamSynchronize(d.stream));
    } else if (!p.hostRhs.empty()) {
      const int64_t n = (int64_t)m.bsx0 * m.bsx0 * m.bsx0;
      for (int b : p.active0)
        std::copy(p.hostRhs.begin() + p.mg[0].off[b], p.hostRhs.begin() + p.mg[0].off[b] + n,
                  hostB.begin() + m0.off[s + b]);
    }
  }
  (void)hipFree(dl);
  if (anyHost) {
    /* merge with the device-computed scene blocks: upload only host-problem blocks */
    const int64_t n = (int64_t)m.bsx0 * m.bsx0 * m.bsx0;
    for (int q = 0; q < m.nProb; q++) {
      if (probs[q]->hasScene) continue;
      for (int b : probs[q]->active0) {
        const int64_t o = m0.off[m.probBlock0[q] + b];
        HIPCHK(hipMemcpy(m0.d_ch[C_B] + o, hostB.data() + o, n * sizeof(float), hipMemcpyHostToDevice));
      }
    }
  }
  /* chunk table: each problem's PCG range split into chunks of <= 64K cells */
  BatchDev *bd = new BatchDev();
  d.batch = bd;
  std::vector<int> cp, cbv, cev;
  std::vector<int64_t> beg, end;
  const int64_t bcells = (int64_t)m.bsx0 * m.bsx0 * m.bsx0, CH = 65536;
  int64_t pos = 0;
  for (int q = 0; q < m.nProb; q++) {
    cbv.push_back((int)cp.size());
    const int64_t len = (int64_t)m.probActive0[q] * bcells;
    for (int64_t o = 0; o < len; o += CH) {
      cp.push_back(q);
      beg.push_back(pos + o);
      end.push_back(pos + std::min(len, o + CH));
    }
    if (len == 0) {  // keep one empty chunk so the problem has a partial
      cp.push_back(q);
      beg.push_back(pos);
      end.push_back(pos);
    }
    cev.push_back((int)cp.size());
    pos += len;
  }
  bd->nChunks = (int)cp.size();
  if (int rc = dalloc(&bd->cProb, cp.size())) return rc;
  if (int rc = dalloc(&bd->cBeg, beg.size())) return rc;
  if (int rc = dalloc(&bd->cEnd, end.size())) return rc;
  if (int rc = dalloc(&bd->cb, m.nProb)) return rc;
  if (int rc = dalloc(&bd->ce, m.nProb)) return rc;
  if (int rc = dalloc(&bd->cProbL, cp.size())) return rc;
  if (int rc = dalloc(&bd->cBegL, beg.size())) return rc;
  if (int rc = dalloc(&bd->cEndL, end.size())) return rc;
  if (int rc = dalloc(&bd->cbL, m.nProb)) return rc;
  if (int rc = dalloc(&bd->ceL, m.nProb)) return rc;
  bd->hProb = cp; bd->hBeg = beg; bd->hEnd = end; bd->hCb = cbv; bd->hCe = cev;
  if (int rc = dalloc(&bd->states, m.nProb)) return rc;
  if (int rc = dalloc(&bd->allDone, 1)) return rc;
  HIPCHK(hipMemcpy(bd->cProb, cp.data(), cp.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(bd->cBeg, beg.data(), beg.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(bd->cEnd, end.data(), end.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(bd->cb, cbv.data(), cbv.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(bd->ce, cev.data(), cev.size() * sizeof(int), hipMemcpyHostToDevice));
  bd->listBlocks = -1;   // forces setLive() to upload the full live table on the first solve
  if (d.partialsCap < bd->nChunks) {
    (void)hipFree(d.partials);
    d.partials = nullptr;
    if (int rc = dalloc(&d.partials, bd->nChunks)) return rc;
    d.partialsCap = bd->nChunks;
  }
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

/* allDone = every problem's PCG has stopped (one workgroup). */
__global__ void __launch_bounds__(256) k_all_done(const PcgState *st, int n, int *allDone)
{
  __shared__ int anyLive;
  if (threadIdx.x == 0) anyLive = 0;
  __syncthreads();
  for (int p = threadIdx.x; p < n; p += blockDim.x)
    if (!st[p].done) anyLive = 1;
  __syncthreads();
  if (threadIdx.x == 0) allDone[0] = anyLive ? 0 : 1;
}

static void finB(Grid &m, int op, double *out, int stride, bool live = false)
{
  Dev &d = dev(m);
  BatchDev &b = *batchDev(m);
  hipLaunchKernelGGL(k_finalize_b, dim3(m.nProb), dim3(256), 0, d.stream, b.states, d.partials, live ? b.cbL : b.cb,
                     live ? b.ceL : b.ce, op, out, stride);
}

/* Restrict the block lists of every MG level and the live chunk table to
 * the problems with live[q] set (all problems when live is empty).  The
 * captured graph holds the old list lengths, so it is dropped. */
static int setLive(Grid &m, const std::vector<char> &live)
{
  Dev &d = dev(m);
  BatchDev &b = *batchDev(m);
  HIPCHK(hipStreamSynchronize(d.stream));
  std::vector<char> all(m.nProb, 1);
  const std::vector<char> &lv = live.empty() ? all : live;
  for (int L = 0; L < m.nLevels; L++)
    if (int rc = uploadLists(m, L, &lv)) return rc;
  std::vector<int> cp, cbv(m.nProb, 0), cev(m.nProb, 0);
  std::vector<int64_t> beg, end;
  int blocks = 0;
  for (int q = 0; q < m.nProb; q++) {
    if (!lv[q]) continue;
    cbv[q] = (int)cp.size();
    for (int c = b.hCb[q]; c < b.hCe[q]; c++) {
      cp.push_back(q);
      beg.push_back(b.hBeg[c]);
      end.push_back(b.hEnd[c]);
    }
    cev[q] = (int)cp.size();
    blocks += (q + 1 < m.nProb ? m.probBlock0[q + 1] : m.nBlocks) - m.probBlock0[q];
  }
  b.nChunksL = (int)cp.size();
  if (!cp.empty()) {
    HIPCHK(hipMemcpy(b.cProbL, cp.data(), cp.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.cBegL, beg.data(), beg.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.cEndL, end.data(), end.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(b.cbL, cbv.data(), cbv.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b.ceL, cev.data(), cev.size() * sizeof(int), hipMemcpyHostToDevice));
  b.listBlocks = blocks;
  if (b.graph) { (void)hipGraphExecDestroy(b.graph); b.graph = nullptr; }
  return MSBGX_OK;
}

/* One batched PCG iteration: every launch covers all problems. */
static void pcgIterationBatch(Grid &m, const msbgx_solve_params &p)
{
  Dev &d = dev(m);
  BatchDev &b = *batchDev(m);
  MgLevel &m0 = m.mg[0];
  const int *nd = b.allDone;
  const int nc = b.nChunksL;
  multiplyLaunch(m, 0, false, m0.d_ch[C_D], nullptr, m0.d_ch[C_Q], nullptr, nd, d.stream);
  hipLaunchKernelGGL(k_dot_c, dim3(nc), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_D], m0.d_ch[C_Q], b.cBegL, b.cEndL,
                     d.partials);
  finB(m, FIN_ALPHA, nullptr, 0, true);
  hipLaunchKernelGGL(k_update_xr_c, dim3(nc), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_X], m0.d_ch[C_R],
                     m0.d_ch[C_D], m0.d_ch[C_Q], b.cProbL, b.cBegL, b.cEndL, b.states, d.partials);
  finB(m, FIN_CONV, b.hist, b.histStride, true);
  hipLaunchKernelGGL(k_all_done, dim3(1), dim3(256), 0, d.stream, b.states, m.nProb, b.allDone);
  zeroActive(m, 0, m0.d_ch[C_Z], nd, d.stream);
  vcycleLaunch(m, 0, p.n_pre, p.n_post, p.n_coarse, nd, d.stream);
  hipLaunchKernelGGL(k_dot_c, dim3(nc), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], m0.d_ch[C_Z], b.cBegL, b.cEndL,
                     d.partials);
  finB(m, FIN_BETA, nullptr, 0, true);
  hipLaunchKernelGGL(k_axpy_c, dim3(nc), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_D], m0.d_ch[C_Z], b.cProbL,
                     b.cBegL, b.cEndL, b.states);
}

int deviceSolveBatch(Grid &m, const msbgx_solve_params &p, int *iters, float *rel)
{
  Dev &d = dev(m);
  BatchDev &b = *batchDev(m);
  MgLevel &m0 = m.mg[0];
  const int P = m.nProb;
  if (b.histStride < p.max_iter + 1) {
    if (b.hist) (void)hipFree(b.hist);
    b.hist = nullptr;
    if (int rc = dalloc(&b.hist, (size_t)P * (p.max_iter + 1))) return rc;
    b.histStride = p.max_iter + 1;
    if (b.graph) { (void)hipGraphExecDestroy(b.graph); b.graph = nullptr; }
  }
  if (b.listBlocks != m.nBlocks)
    if (int rc = setLive(m, {})) return rc;   // every problem is live at the start
  std::vector<PcgState> init(P);
  for (auto &s : init) {
    s = PcgState{};
    s.tolSq = (double)p.tol * (double)p.tol;
    s.maxIter = p.max_iter;
  }
  HIPCHK(hipMemcpyAsync(b.states, init.data(), P * sizeof(PcgState), hipMemcpyHostToDevice, d.stream));
  HIPCHK(hipMemcpyAsync(b.allDone, d.zeroFlag, sizeof(int), hipMemcpyDeviceToDevice, d.stream));
  HIPCHK(hipEventRecord(d.ev0, d.stream));
  const int *nd = d.zeroFlag;
  hipLaunchKernelGGL(k_copy_norm_c, dim3(b.nChunks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_B], m0.d_ch[C_R], b.cBeg,
                     b.cEnd, d.partials);
  finB(m, FIN_BNORM, nullptr, 0);
  zeroActive(m, 0, m0.d_ch[C_Z], nd, d.stream);
  vcycleLaunch(m, 0, p.n_pre, p.n_post, p.n_coarse, nd, d.stream);
  for (int lev = 0; lev < m.nLevels; lev++)
    if (m0.dAct[lev].n)
      hipLaunchKernelGGL(k_copy_blocks, dim3(m0.dAct[lev].n), dim3(256), 0, d.stream, m0.dAct[lev].ptr, m0.d_level,
                         m0.d_off, m.bsx0, m0.d_ch[C_Z], m0.d_ch[C_D], nd);
  hipLaunchKernelGGL(k_dot_c, dim3(b.nChunks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], m0.d_ch[C_Z], b.cBeg, b.cEnd,
                     d.partials);
  finB(m, FIN_RHO0, nullptr, 0);
  hipLaunchKernelGGL(k_all_done, dim3(1), dim3(256), 0, d.stream, b.states, P, b.allDone);
  HIPCHK(hipGetLastError());
  const GraphKey key = graphKey(m, p);
  const int chunk = key.chunk;
  const bool useGraph = !(m.options & MSBGX_OPT_NO_GRAPH);
  if (b.graph && b.graphKey != key) { (void)hipGraphExecDestroy(b.graph); b.graph = nullptr; }
  std::vector<PcgState> hs(P);
  for (int it = 0; it < p.max_iter;) {
    const int n = std::min(chunk, p.max_iter - it);
    if (useGraph && n == chunk) {
      if (!b.graph) {
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(d.stream, hipStreamCaptureModeRelaxed));
        for (int i = 0; i < chunk; i++) pcgIterationBatch(m, p);
        HIPCHK(hipStreamEndCapture(d.stream, &graph));
        HIPCHK(hipGraphInstantiate(&b.graph, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        b.graphKey = key;
      }
      HIPCHK(hipGraphLaunch(b.graph, d.stream));
    } else {
      for (int i = 0; i < n; i++) pcgIterationBatch(m, p);
    }
    it += n;
    HIPCHK(hipMemcpyAsync(hs.data(), b.states, P * sizeof(PcgState), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    bool all = true;
    for (auto &s : hs) all &= s.done != 0;
    if (all) break;
    /* drop converged problems from the launches once they hold a quarter of
     * the listed blocks (their x and r are frozen already) */
    std::vector<char> live(P);
    int liveBlocks = 0;
    for (int q = 0; q < P; q++) {
      live[q] = hs[q].done ? 0 : 1;
      if (live[q]) liveBlocks += (q + 1 < P ? m.probBlock0[q + 1] : m.nBlocks) - m.probBlock0[q];
    }
    if (!(m.options & MSBGX_OPT_NO_COMPACT) && 4 * liveBlocks <= 3 * b.listBlocks)
      if (int rc = setLive(m, live)) return rc;
  }
  hipLaunchKernelGGL(k_dot_c, dim3(b.nChunks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], m0.d_ch[C_R], b.cBeg, b.cEnd,
                     d.partials);
  finB(m, FIN_FINAL, b.hist + p.max_iter, b.histStride);
  HIPCHK(hipEventRecord(d.ev1, d.stream));
  HIPCHK(hipMemcpyAsync(hs.data(), b.states, P * sizeof(PcgState), hipMemcpyDeviceToHost, d.stream));
  std::vector<double> fin(P);
  for (int q = 0; q < P; q++)
    HIPCHK(hipMemcpyAsync(&fin[q], b.hist + (size_t)q * b.histStride + p.max_iter, sizeof(double),
                          hipMemcpyDeviceToHost, d.stream));
  HIPCHK(hipStreamSynchronize(d.stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, d.ev0, d.ev1));
  m.lastSolveMs = ms;
  if (b.listBlocks != m.nBlocks)
    if (int rc = setLive(m, {})) return rc;   // leave the full lists for the other entry points
  std::vector<double> hh((size_t)P * b.histStride);
  HIPCHK(hipMemcpy(hh.data(), b.hist, hh.size() * sizeof(double), hipMemcpyDeviceToHost));
  m.probHistory.assign(P, {});
  for (int q = 0; q < P; q++) {
    for (int i = 0; i < hs[q].iter; i++)
      m.probHistory[q].push_back(std::sqrt(hh[(size_t)q * b.histStride + i] / hs[q].bnorm2));
    if (hs[q].brk == 3) { setError("non-finite right-hand side in a batched problem"); return MSBGX_EINVAL; }
    if (hs[q].brk == 2) { iters[q] = 0; rel[q] = 0.0f; continue; }
    rel[q] = (float)std::sqrt(fin[q] / hs[q].bnorm2);
    iters[q] = hs[q].conv > 0 ? hs[q].conv : p.max_iter;
  }
  return MSBGX_OK;
}

/* ------------------------------------------------------------- autotuner */

/* Picks, by measurement on this handle's own problem:
 *  - the smoother kernel per block size (column-resident or LDS tile; the
 *    tile form exists for FP32 operands and blocks up to 16^3; for 32^3
 *    blocks with 16-bit operands, register-resident or record-streaming),
 *    timing the two colour passes over every MG level's relaxation lists;
 *  - the number of PCG iterations per hipGraph replay (1, 2, 4, 8, 16),
 *    timing a whole solve (its second run, so graph capture is excluded).
 * Both choices reorder nothing inside a kernel, so the solver's results are
 * bit-identical for every choice.  The pressure channel is zero afterwards. */
int deviceAutotune(Grid &g, const msbgx_solve_params &p, int *chunkOut, int *variantMask)
{
  Dev &d = dev(g);
  const int reps = 10;
  int mask = 0;
  for (int lev = 0; lev < g.nLevels; lev++) {
    g.smootherVariant[lev] = 0;
    const int B = g.bsx0 >> lev;
    const bool tile = B <= 16 && g.prec == MSBGX_PREC_FP32;                   // column vs LDS tile
    const bool c32 = B == 32 && lev == 0 && levelArgs(g, 0).pack16 != nullptr;  // registers vs records
    if (!tile && !c32) continue;
    int nBlocks = 0;
    for (int L = 0; L <= lev; L++) nBlocks += g.mg[L].dRelax[0][lev].n + g.mg[L].dRelax[1][lev].n;
    if (nBlocks == 0) continue;
    float t[2] = {0.f, 0.f};
    for (int v = 0; v < 2; v++) {
      g.smootherVariant[lev] = v;
      for (int rep = -2; rep < reps; rep++) {
        if (rep == 0) HIPCHK(hipEventRecord(d.ev0, d.stream));
        for (int L = 0; L <= lev; L++)
          for (int c = 0; c < 2; c++)
            launchRelaxLev(g, L, lev, g.mg[L].dRelax[c][lev], g.mg[L].d_ch[C_Z], g.mg[L].d_ch[C_R], 0, d.zeroFlag,
                           d.stream);
      }
      HIPCHK(hipEventRecord(d.ev1, d.stream));
      HIPCHK(hipStreamSynchronize(d.stream));
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventElapsedTime(&t[v], d.ev0, d.ev1));
    }
    g.smootherVariant[lev] = t[1] < 0.95f * t[0] ? 1 : 0;
    if (g.smootherVariant[lev]) mask |= 1 << lev;
  }
  deviceInvalidateGraphs(g);   // the graphs hold the previous kernel choices

  const int P = g.batched ? g.nProb : 1;
  std::vector<int> it(P);
  std::vector<float> rel(P);
  static const int cands[] = {1, 2, 4, 8, 16};
  int best = g.graphChunk;
  double bestMs = 1e30;
  for (int c : cands) {
    g.graphChunk = c;
    double ms = 0.0;
    for (int run = 0; run < 2; run++) {
      if (int rc = deviceZero(g, C_X, 0)) return rc;
      const int rc = g.batched ? deviceSolveBatch(g, p, it.data(), rel.data()) : deviceSolve(g, p, it.data(), rel.data());
      if (rc) return rc;
      ms = g.lastSolveMs;
    }
    if (ms < bestMs) { bestMs = ms; best = c; }
  }
  g.graphChunk = best;
  if (int rc = deviceZero(g, C_X, 0)) return rc;
  *chunkOut = best;
  *variantMask = mask;
  return MSBGX_OK;
}

}  // namespace msbgx

/* The learned correction (network kernels, training, PCG integration). */
#include "learned.hip"
#include "flip.hip"
#include "p2g.hip"
#include "compact.hip"
#include "pde.hip"

namespace msbgx {
/* Captured graphs bake in launch shapes and kernel choices: drop them. */
void deviceInvalidateGraphs(Grid &g)
{
  if (!g.dev) return;
  Dev &d = dev(g);
  const bool any = d.graph || (d.batch && d.batch->graph) || (d.learned && d.learned->graph);
  if (any && d.stream) (void)hipStreamSynchronize(d.stream);   // no replay may still be running
  if (d.graph) { (void)hipGraphExecDestroy(d.graph); d.graph = nullptr; }
  if (d.batch && d.batch->graph) { (void)hipGraphExecDestroy(d.batch->graph); d.batch->graph = nullptr; }
  if (d.learned && d.learned->graph) { (void)hipGraphExecDestroy(d.learned->graph); d.learned->graph = nullptr; }
}
}  // namespace msbgx
