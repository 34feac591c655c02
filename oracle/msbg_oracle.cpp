/*
 * msbg_oracle.cpp -- CPU restatement of the MSBG FLIP pressure solver.
 *
 * TEST INFRASTRUCTURE.  This file is the parity oracle and the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it; the product library never links or calls it.
 *
 * It restates, in plain C++ (no SIMD intrinsics, no fast-math), the reference
 * semantics of:
 *   - block raster, levels and block flags      src/msbg.cpp:1202-1420, 1600-1736; src/msbg.h:3560-3590
 *   - preparePressureSolveFLIP and helpers      src/msbg3.cpp:2472-3095
 *   - block halo staging + smoother + multiply  src/msbg3.cpp:46-161, 419-1506, 1757-1912
 *   - red-black block relaxation                src/msbg3.cpp:2171-2369
 *   - restriction / prolongation                src/msbg4.cpp:1513-2029
 *   - V-cycle and MG-PCG                        src/msbg3.cpp:2375-2465, 3098-3271
 *
 * Storage mirrors the reference's SparseGrid model literally: one block map
 * per (channel, level, MG level); a missing block reads as the grid's empty
 * value; restriction of non-coarsened blocks installs a block reference into
 * the coarse right-hand side exactly as setExternalBlockDataPtr does
 * (src/msbg4.cpp:1642-1651), so the zeroing through that reference on later
 * V-cycles (src/msbg3.cpp:2375-2399, 2449) happens here too.
 *
 * Floating-point operation order: every float expression is written out with
 * explicit fmaf() where the reference uses mul_add or where this restatement
 * fixes a fused form, and the file is compiled with -ffp-contract=off so no
 * other fusion happens.  The PCG dot products accumulate in long double, as
 * the reference does (src/msbg3.cpp:3111-3249).
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <deque>
#include <mutex>
#include <memory>
#include <queue>
#include <vector>
#include <algorithm>
#include <omp.h>

#include "../include/msbgx_scene.h"

namespace {

typedef uint16_t CellFlags;

enum : unsigned {
  CELL_SOLID = 1u << 0, CELL_FIXED = 1u << 1, CELL_PARTIAL_SOLID = 1u << 6, CELL_VOID = 1u << 12,
  CELL_EMPTY_VAL = MSBGX_CELL_EMPTY_VAL
};
enum : unsigned {
  BLK_DOM_BORDER = 1u << 1, BLK_COARSE_FINE = 1u << 2, BLK_FINE_COARSE = 1u << 3,
  BLK_NO_FLUID = 1u << 5, BLK_FIXED = 1u << 11, BLK_EXISTS = 1u << 12, BLK_CUTS_SOLID = 1u << 15
};
enum { DBC_OPEN = 0, DBC_SOLID = 1 };

/* Channels of the pressure solve (src/msbg3.cpp:3100-3103). */
enum { CH_X = 0, CH_B, CH_R, CH_Z, CH_D, CH_Q, CH_TMP, CH_W0, CH_W1, CH_W2, CH_DIAG, N_FCH };

const int MAXL = 5;

struct BlockInfo { int level; unsigned flags; };

/* One SparseGrid<T>: a block map over the shared raster (src/sbg.h:914-1022). */
template <typename T> struct Grid {
  int bsx = 0, sx = 0, sy = 0, sz = 0, nbx = 0, nby = 0, nbz = 0;
  T empty{};
  std::vector<T *> map;
  std::deque<std::unique_ptr<T[]>> pool;
  const T *emptyBlock = nullptr;  // shared zero block (setEmptyBlock)
  int nvox() const { return bsx * bsx * bsx; }
  void init(int bsx_, int sx_, int sy_, int sz_, int nbx_, int nby_, int nbz_, T empty_, const T *zero)
  {
    bsx = bsx_; sx = sx_; sy = sy_; sz = sz_; nbx = nbx_; nby = nby_; nbz = nbz_; empty = empty_;
    emptyBlock = zero;
    map.assign((size_t)nbx * nby * nbz, nullptr);
    pool.clear();
  }
  T *get(int bid) const { return map[bid]; }
  bool isValue(int bid) const { return map[bid] && map[bid] != emptyBlock; }
  /* alloc-on-write; fresh blocks are zero.  Called from block-parallel
   * loops (each block by one thread): the shared pool is appended under a
   * lock, the block's own map slot needs none. */
  T *alloc(int bid)
  {
    if (!map[bid] || map[bid] == emptyBlock) {
      T *blk = new T[nvox()]();
      {
        std::lock_guard<std::mutex> lk(poolMutex());
        pool.emplace_back(blk);
      }
      map[bid] = blk;
    }
    return map[bid];
  }
  static std::mutex &poolMutex()
  {
    static std::mutex m;
    return m;
  }
  bool inRange(int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < sx && y < sy && z < sz; }
  T getValue(int x, int y, int z) const
  {
    const int b = (x / bsx) + nbx * ((y / bsx) + nby * (z / bsx));
    const T *d = map[b];
    if (!d) return empty;
    return d[(x % bsx) + bsx * ((y % bsx) + bsx * (z % bsx))];
  }
  T getValueGen(int x, int y, int z) const { return inRange(x, y, z) ? getValue(x, y, z) : empty; }
};

struct Callbacks {
  CellFlags (*cellType)(void *, int, int, int);
  float (*faceCoeff)(void *, int, int, int, int);
  float (*rhs)(void *, int, int, int);
  void *user;
};

static CellFlags sceneCellType(void *u, int x, int y, int z) { return msbgx_scene_cell_type((const msbgx_scene *)u, x, y, z); }
static float sceneFaceCoeff(void *u, int d, int x, int y, int z) { return msbgx_scene_face_coeff((const msbgx_scene *)u, d, x, y, z); }
static float sceneRhs(void *u, int x, int y, int z) { return msbgx_scene_rhs((const msbgx_scene *)u, x, y, z); }

struct DenseInput { int nx, ny, nz; const uint16_t *flags; const float *w[3]; const float *rhs; };
static size_t dIdx(const DenseInput *d, int x, int y, int z) { return (size_t)x + (size_t)d->nx * ((size_t)y + (size_t)d->ny * z); }
static CellFlags denseCellType(void *u, int x, int y, int z) { auto *d = (DenseInput *)u; return d->flags[dIdx(d, x, y, z)]; }
static float denseFaceCoeff(void *u, int k, int x, int y, int z) { auto *d = (DenseInput *)u; return d->w[k][dIdx(d, x, y, z)]; }
static float denseRhs(void *u, int x, int y, int z) { auto *d = (DenseInput *)u; return d->rhs[dIdx(d, x, y, z)]; }

static inline int simdWidth(int bsx) { return bsx >= 8 ? 8 : bsx; }   // src/msbg3.cpp:1700-1743

/* Precision form (a build addition, no reference counterpart): the
 * smoother's face coefficients and inverse diagonal rounded to 16 bits with
 * round-to-nearest-even and read back exactly.  prec 1: IEEE binary16,
 * prec 2: bfloat16 (NaN stays NaN). */
static float roundHalf(float f)
{
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
  uint32_t h;
  if (ax > 0x7f800000u) h = sign | 0x7e00u;                      // NaN
  else if (ax >= 0x477ff000u) h = sign | 0x7c00u;                // >= 65520: infinity
  else if (ax < 0x38800000u) {                                    // below 2^-14: subnormal or zero
    float a;
    std::memcpy(&a, &ax, 4);
    h = sign | (uint32_t)std::nearbyint((double)a * 16777216.0);  // units of 2^-24, ties to even
  } else {
    const uint32_t e = (ax >> 23) - 127 + 15, m = ax & 0x7fffffu;
    h = (e << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    h |= sign;
  }
  const uint32_t hs = h & 0x8000u, he = (h >> 10) & 0x1fu, hm = h & 0x3ffu;
  float v;
  if (he == 0) v = std::ldexp((float)hm, -24);
  else if (he == 31) v = hm ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
  else v = std::ldexp((float)(hm | 0x400u), (int)he - 25);
  return hs ? -v : v;
}
static float roundBf16(float f)
{
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t b;
  if ((x & 0x7fffffffu) > 0x7f800000u) b = (x >> 16) | 0x40u;
  else b = (x + 0x7fffu + ((x >> 16) & 1u)) >> 16;
  b <<= 16;
  float v;
  std::memcpy(&v, &b, 4);
  return v;
}
static inline float round16(float v, int prec) { return prec == 1 ? roundHalf(v) : (prec == 2 ? roundBf16(v) : v); }

struct Oracle {
  int sx0, sy0, sz0, bsx0, nLevels, nbx, nby, nbz, nBlocks;
  int domainBc[3][2];
  std::vector<int> level0;                       // block level map
  std::vector<BlockInfo> binfo[MAXL];            // per MG level
  std::vector<int> relaxList[MAXL];
  Grid<CellFlags> flags[MAXL][MAXL];             // [level][mg]
  Grid<float> fch[N_FCH][MAXL][MAXL];            // [channel][level][mg]
  std::vector<float> zeroBlk;
  std::vector<CellFlags> zeroBlkF;
  float omegaDefault = 6.f / 7.f, omega = 6.f / 7.f;
  int smPrec = 0;   // precision form: 16-bit smoother operands (0 FP32, 1 FP16, 2 BF16)
  /* precision form: x of the coarsest MG level's blocks in FP64 during its relax */
  std::vector<std::vector<double>> x64;
  int nBlockIter = 5;
  int nThreads = 1;

  int bsxL(int l) const { return bsx0 >> l; }
  void blockCoords(int bid, int &bx, int &by, int &bz) const
  {
    bx = bid % nbx; by = (bid / nbx) % nby; bz = bid / (nbx * nby);
  }
  bool blockInRange(int bx, int by, int bz) const { return bx >= 0 && by >= 0 && bz >= 0 && bx < nbx && by < nby && bz < nbz; }
  int blockIndex(int bx, int by, int bz) const { return bx + nbx * (by + nby * bz); }

  void resetChannels()
  {
    for (int l = 0; l < nLevels; l++)
      for (int m = 0; m < nLevels; m++) {
        const int b = bsxL(l);
        flags[l][m].init(b, sx0 >> l, sy0 >> l, sz0 >> l, nbx, nby, nbz, (CellFlags)CELL_EMPTY_VAL, zeroBlkF.data());
        for (int c = 0; c < N_FCH; c++)
          fch[c][l][m].init(b, sx0 >> l, sy0 >> l, sz0 >> l, nbx, nby, nbz, 0.0f, zeroBlk.data());
      }
  }

  Oracle(int sx, int sy, int sz, int bsx, int nl) : sx0(sx), sy0(sy), sz0(sz), bsx0(bsx), nLevels(nl)
  {
    nbx = sx / bsx; nby = sy / bsx; nbz = sz / bsx; nBlocks = nbx * nby * nbz;
    for (int i = 0; i < 3; i++) domainBc[i][0] = domainBc[i][1] = DBC_OPEN;  // src/msbg.cpp:1077-1080
    zeroBlk.assign((size_t)bsx * bsx * bsx + 64, 0.0f);
    zeroBlkF.assign((size_t)bsx * bsx * bsx + 64, 0);
    level0.assign(nBlocks, 0);
    for (int m = 0; m < MAXL; m++) binfo[m].assign(nBlocks, BlockInfo{0, 0});
    resetChannels();
    setRefinementMap(level0.data(), true);
  }

  /* ---------------------------------------------------------------- topology */

  /* src/msbg.cpp:1202-1420 */
  void regularizeRefinementMap(const int *in, int *out) const
  {
    std::vector<int> tmp2(in, in + nBlocks);
    std::vector<int> res(in, in + nBlocks);
    for (int ll = 1; ll <= nLevels - 2; ll++) {
      for (int bid = 0; bid < nBlocks; bid++) {
        int bx, by, bz;
        blockCoords(bid, bx, by, bz);
        int level = tmp2[bid];
        if (level >= ll) {
          int level2 = nLevels - 1;
          for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
              for (int dx = -1; dx <= 1; dx++) {
                if (!dx && !dy && !dz) continue;
                if (!blockInRange(bx + dx, by + dy, bz + dz)) continue;
                level2 = std::min(level2, tmp2[blockIndex(bx + dx, by + dy, bz + dz)]);
              }
          if (level2 == ll - 1) level = ll;
        }
        res[bid] = level;
      }
      tmp2 = res;
    }
    for (int i = 0; i < nBlocks; i++) out[i] = res[i];
  }

  /* src/msbg.cpp:1428-1736 (block part; no obstacle grid, no caller flags) */
  void setRefinementMap(const int *levels, bool doResetChannels)
  {
    for (int i = 0; i < nBlocks; i++) level0[i] = levels[i];
    if (doResetChannels) resetChannels();
    for (int L = 0; L < nLevels; L++) {
      for (int bid = 0; bid < nBlocks; bid++) {
        BlockInfo &bi = binfo[L][bid];
        bi.level = std::max(level0[bid], L);
        bi.flags = 0;
        int bx, by, bz;
        blockCoords(bid, bx, by, bz);
        if (bx == 0 || bx == nbx - 1 || by == 0 || by == nby - 1 || bz == 0 || bz == nbz - 1) bi.flags |= BLK_DOM_BORDER;
        int l2max = 0, l2min = nLevels - 1;
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              if (!dx && !dy && !dz) continue;
              if (!blockInRange(bx + dx, by + dy, bz + dz)) continue;
              const int l2 = std::max(level0[blockIndex(bx + dx, by + dy, bz + dz)], L);
              l2max = std::max(l2max, l2);
              l2min = std::min(l2min, l2);
            }
        if (l2max > bi.level) bi.flags |= BLK_FINE_COARSE;
        if (l2min < bi.level) bi.flags |= BLK_COARSE_FINE;
      }
    }
  }

  /* src/msbg.h:2488-2494 */
  bool isRelaxationBlock(int L, const BlockInfo &bi) const
  {
    return !(bi.flags & (BLK_NO_FLUID | BLK_FIXED)) &&
           (bi.level <= L || ((bi.flags & (BLK_COARSE_FINE | BLK_FINE_COARSE)) && bi.level == L + 1));
  }

  /* src/msbg3.cpp:2868-2879 */
  void buildRelaxationBlocks(int L)
  {
    relaxList[L].clear();
    for (int bid = 0; bid < nBlocks; bid++) {
      const BlockInfo &bi = binfo[L][bid];
      if (!(bi.flags & BLK_EXISTS)) continue;
      if (isRelaxationBlock(L, bi)) relaxList[L].push_back(bid);
    }
  }

  /* ----------------------------------------------------------------- prepare */

  /* src/msbg3.cpp:3010-3038: alloc + zero work channels at MG0, own level */
  void allocPressureChannels()
  {
    const int chans[] = {CH_X, CH_B, CH_R, CH_Z, CH_D, CH_Q, CH_TMP};
    for (int c : chans)
      for (int bid = 0; bid < nBlocks; bid++) {
        float *d = fch[c][binfo[0][bid].level][0].alloc(bid);
        (void)d;
      }
  }

  /* src/msbg3.cpp:2472-2583.  Blocks are allocated serially, then filled
   * block-parallel when the callbacks are pure functions (scene / dense). */
  void prepareLevel0(const Callbacks &cb, bool parallel)
  {
    for (int bid = 0; bid < nBlocks; bid++) {
      const int l = binfo[0][bid].level;
      flags[l][0].alloc(bid);
      for (int dir = 0; dir < 3; dir++) fch[CH_W0 + dir][l][0].alloc(bid);
      if (l == 0) fch[CH_B][0][0].alloc(bid);
      fch[CH_X][l][0].alloc(bid);
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(nThreads) if (parallel)
    for (int bid = 0; bid < nBlocks; bid++) {
      BlockInfo &bi = binfo[0][bid];
      const int l = bi.level, b = bsxL(l);
      CellFlags *dst = flags[l][0].get(bid);
      int bx, by, bz;
      blockCoords(bid, bx, by, bz);
      bool hasLiquid = false;
      for (int vz = 0, vid = 0; vz < b; vz++)
        for (int vy = 0; vy < b; vy++)
          for (int vx = 0; vx < b; vx++, vid++) {
            CellFlags fl = cb.cellType(cb.user, bx * b + vx, by * b + vy, bz * b + vz);
            fl &= (CELL_SOLID | CELL_VOID | CELL_FIXED);
            dst[vid] = fl;
            if (!(fl & (CELL_SOLID | CELL_VOID))) hasLiquid = true;
          }
      bi.flags |= BLK_EXISTS;
      if (hasLiquid) bi.flags &= ~BLK_NO_FLUID; else bi.flags |= BLK_NO_FLUID;
      bi.flags &= ~BLK_FIXED;
    }
    for (int dir = 0; dir < 3; dir++) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nThreads) if (parallel)
      for (int bid = 0; bid < nBlocks; bid++) {
        const int l = binfo[0][bid].level, b = bsxL(l);
        float *w = fch[CH_W0 + dir][l][0].get(bid);
        int bx, by, bz;
        blockCoords(bid, bx, by, bz);
        for (int vz = 0, vid = 0; vz < b; vz++)
          for (int vy = 0; vy < b; vy++)
            for (int vx = 0; vx < b; vx++, vid++) w[vid] = cb.faceCoeff(cb.user, dir, bx * b + vx, by * b + vy, bz * b + vz);
      }
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(nThreads) if (parallel)
    for (int bid = 0; bid < nBlocks; bid++) {
      if (binfo[0][bid].level != 0) continue;
      const int b = bsx0;
      float *r = fch[CH_B][0][0].get(bid);
      int bx, by, bz;
      blockCoords(bid, bx, by, bz);
      for (int vz = 0, vid = 0; vz < b; vz++)
        for (int vy = 0; vy < b; vy++)
          for (int vx = 0; vx < b; vx++, vid++) r[vid] = cb.rhs(cb.user, bx * b + vx, by * b + vy, bz * b + vz);
    }
    for (int bid = 0; bid < nBlocks; bid++) {  // setChannel(0, CH_PRESSURE)
      float *x = fch[CH_X][binfo[0][bid].level][0].get(bid);
      std::fill(x, x + fch[CH_X][binfo[0][bid].level][0].nvox(), 0.0f);
    }
  }

  /* src/msbg3.cpp:2586-2779 */
  void computeDiagonal(int L)
  {
    for (int level = L; level < nLevels; level++) {
      Grid<CellFlags> &sgF = flags[level][L];
      Grid<float> &sgD = fch[CH_DIAG][level][L];
      for (int bid = 0; bid < nBlocks; bid++)
        if (binfo[L][bid].level == level && (binfo[L][bid].flags & BLK_EXISTS)) sgD.alloc(bid);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nThreads)
      for (int bid = 0; bid < nBlocks; bid++) {
        BlockInfo &bi = binfo[L][bid];
        if (bi.level != level || !(bi.flags & BLK_EXISTS)) continue;
        CellFlags *f = sgF.get(bid);
        float *d = sgD.get(bid);
        if (!f) continue;
        const int b = bsxL(level);
        bi.flags &= ~BLK_CUTS_SOLID;
        bool var = false;
        for (int i = 0; i < b * b * b; i++) {
          f[i] &= ~CELL_PARTIAL_SOLID;
          if (f[i] & (CELL_SOLID | CELL_VOID | CELL_FIXED)) { d[i] = 0; continue; }
          d[i] = 1.f / 6.f;
          f[i] |= CELL_PARTIAL_SOLID;   // face-area channels always exist on this path
          var = true;
        }
        if (var) bi.flags |= BLK_CUTS_SOLID;
      }
    }
    if (L == 0) {
      for (int level = 0; level < nLevels; level++) {
        Grid<float> &U = fch[CH_W0][level][0], &V = fch[CH_W1][level][0], &W = fch[CH_W2][level][0];
#pragma omp parallel for schedule(dynamic, 4) num_threads(nThreads)
        for (int bid = 0; bid < nBlocks; bid++) {
          BlockInfo &bi = binfo[0][bid];
          if (bi.level != level || !(bi.flags & BLK_EXISTS)) continue;
          float *d = fch[CH_DIAG][level][0].get(bid);
          CellFlags *f = flags[level][0].get(bid);
          if (!d || !f) continue;
          int bx, by, bz;
          blockCoords(bid, bx, by, bz);
          const int b = bsxL(level);
          for (int vz = 0, vid = 0; vz < b; vz++)
            for (int vy = 0; vy < b; vy++)
              for (int vx = 0; vx < b; vx++, vid++) {
                if (f[vid] & (CELL_SOLID | CELL_VOID | CELL_FIXED)) { d[vid] = 0; continue; }
                const int gx = bx * b + vx, gy = by * b + vy, gz = bz * b + vz;
                float sum = U.getValueGen(gx, gy, gz);
                sum = sum + U.getValueGen(gx + 1, gy, gz);
                sum = sum + V.getValueGen(gx, gy, gz);
                sum = sum + V.getValueGen(gx, gy + 1, gz);
                sum = sum + W.getValueGen(gx, gy, gz);
                sum = sum + W.getValueGen(gx, gy, gz + 1);
                d[vid] = (sum > 1e-12f) ? (1.f / sum) : 0.0f;
              }
        }
      }
    } else {
      for (int level = L; level < nLevels; level++) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nThreads)
        for (int bid = 0; bid < nBlocks; bid++) {
          BlockInfo &bi = binfo[L][bid];
          if (bi.level != level || !(bi.flags & BLK_EXISTS)) continue;
          CellFlags *f = flags[level][L].get(bid);
          float *d = fch[CH_DIAG][level][L].get(bid);
          float *fu = fch[CH_W0][level][L].get(bid), *fv = fch[CH_W1][level][L].get(bid), *fw = fch[CH_W2][level][L].get(bid);
          if (!f || !d || !fu || !fv || !fw) continue;
          const int b = bsxL(level);
          for (int i = 0; i < b * b * b; i++) {
            if (f[i] & (CELL_SOLID | CELL_VOID | CELL_FIXED)) { d[i] = 0; continue; }
            float sum = fu[i] + fv[i];
            sum = sum + fw[i];
            sum *= 2.0f;
            d[i] = (sum > 1e-12f) ? (1.f / sum) : 0.0f;
          }
        }
      }
    }
  }

  /* src/msbg3.cpp:2882-3007 */
  void restrictPressureMeta(int Lc)
  {
    const int Lf = Lc - 1;
    for (int level = 0; level < nLevels; level++) {
      if (level < Lc) continue;   // no grid below the MG level
      Grid<CellFlags> &C = flags[level][Lc];
      Grid<CellFlags> &F = flags[level][Lf];
      for (int bid = 0; bid < nBlocks; bid++)
        if (binfo[Lc][bid].level == level) C.alloc(bid);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nThreads)
      for (int bid = 0; bid < nBlocks; bid++) {
        BlockInfo &bi = binfo[Lc][bid];
        if (bi.level != level) continue;
        CellFlags *dst = C.get(bid);
        int bx, by, bz;
        blockCoords(bid, bx, by, bz);
        const int b = bsxL(level);
        for (int vz = 0, vid = 0; vz < b; vz++)
          for (int vy = 0; vy < b; vy++)
            for (int vx = 0; vx < b; vx++, vid++) {
              const int fx = (bx * b + vx) * 2, fy = (by * b + vy) * 2, fz = (bz * b + vz) * 2;
              bool hasFluid = false, hasSolid = false;
              for (int dz = 0; dz < 2; dz++)
                for (int dy = 0; dy < 2; dy++)
                  for (int dx = 0; dx < 2; dx++) {
                    const CellFlags cf = F.getValueGen(fx + dx, fy + dy, fz + dz);
                    if (!(cf & (CELL_SOLID | CELL_VOID | CELL_FIXED))) hasFluid = true;
                    if (cf & CELL_SOLID) hasSolid = true;
                  }
              dst[vid] = hasFluid ? 0 : (hasSolid ? CELL_SOLID : CELL_VOID);
            }
        bool any = false;
        for (int i = 0; i < b * b * b; i++)
          if (!(dst[i] & (CELL_SOLID | CELL_VOID | CELL_FIXED))) { any = true; break; }
        bi.flags |= BLK_EXISTS;
        if (any) bi.flags &= ~BLK_NO_FLUID; else bi.flags |= BLK_NO_FLUID;
        bi.flags &= ~BLK_FIXED;
      }
    }
    for (int dir = 0; dir < 3; dir++)
      for (int level = Lc; level < nLevels; level++) {
        Grid<float> &C = fch[CH_W0 + dir][level][Lc];
        Grid<float> &F = fch[CH_W0 + dir][level][Lf];
        for (int bid = 0; bid < nBlocks; bid++)
          if (binfo[Lc][bid].level == level) C.alloc(bid);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nThreads)
        for (int bid = 0; bid < nBlocks; bid++) {
          if (binfo[Lc][bid].level != level) continue;
          float *w = C.get(bid);
          int bx, by, bz;
          blockCoords(bid, bx, by, bz);
          const int b = bsxL(level);
          for (int vz = 0, vid = 0; vz < b; vz++)
            for (int vy = 0; vy < b; vy++)
              for (int vx = 0; vx < b; vx++, vid++) {
                const int fgx = (bx * b + vx) * 2, fgy = (by * b + vy) * 2, fgz = (bz * b + vz) * 2;
                float sum = 0.f;
                for (int j = 0; j < 2; j++)
                  for (int k = 0; k < 2; k++) {
                    int fx = fgx, fy = fgy, fz = fgz;
                    if (dir == 0) { fy += j; fz += k; }
                    if (dir == 1) { fx += j; fz += k; }
                    if (dir == 2) { fx += j; fy += k; }
                    sum = sum + F.getValueGen(fx, fy, fz);
                  }
                w[vid] = sum * 0.25f;
              }
        }
      }
    computeDiagonal(Lc);
    buildRelaxationBlocks(Lc);
  }

  void prepare(const Callbacks &cb, bool parallel)
  {
    allocPressureChannels();
    prepareLevel0(cb, parallel);
    computeDiagonal(0);
    buildRelaxationBlocks(0);
    for (int L = 1; L < nLevels; L++) restrictPressureMeta(L);
    /* omega from the MG0 diagonal range (src/msbg3.cpp:3066-3094) */
    float mx = 0.f, mn = 1e30f;
    for (int level = 0; level < nLevels; level++)
      for (int bid = 0; bid < nBlocks; bid++) {
        const BlockInfo &bi = binfo[0][bid];
        if (bi.level != level || !(bi.flags & BLK_EXISTS)) continue;
        const float *d = fch[CH_DIAG][level][0].get(bid);
        const CellFlags *f = flags[level][0].get(bid);
        if (!d || !f) continue;
        for (int i = 0; i < fch[CH_DIAG][level][0].nvox(); i++) {
          if (f[i] & (CELL_SOLID | CELL_VOID | CELL_FIXED)) continue;
          const float v = d[i];
          if (v > 1e-12f) { if (v > mx) mx = v; if (v < mn) mn = v; }
        }
      }
    omega = omegaDefault;
    if (mx > 1e-12f && mn > 1e-12f) {
      const float ratio = mx / mn;
      omega = omegaDefault / sqrtf(std::max(ratio, 1.f));
      omega = std::max(omega, 0.02f);
      omega = std::min(omega, omegaDefault);
    }
  }

  /* ----------------------------------------------------- block Laplacian op */

  enum { OP_RELAX = 1, OP_MULTIPLY = 2, OP_RESIDUAL = 4 };

  /* One block of IprocessBlockLaplacian (src/msbg3.cpp:419-1506).
   * Halo tile H has a one-cell border in every direction. */
  void processBlock(int op, bool reverse, int L, int bid, int chX, int chB, int chDst, int nIter)
  {
    const BlockInfo &bi = binfo[L][bid];
    const int level = bi.level, b = bsxL(level), b2 = b * b, nv = b2 * b, S = simdWidth(b);
    const float h = (float)(1 << level);
    const int hs = b + 2, hs2 = hs * hs;
    auto H = [&](int x, int y, int z) { return (x + 1) + hs * (y + 1) + hs2 * (z + 1); };
    /* per-thread staging buffers, as the reference's tmpBlkData*Buf */
    thread_local std::vector<float> X, Wt[6];
    thread_local std::vector<uint8_t> segF;
    X.assign((size_t)hs * hs * hs, 0.0f);
    for (int k = 0; k < 6; k++) Wt[k].assign(nv, 0.0f);
    segF.assign(nv / S, 0);
    bool blockHasSkippable = false;
    const bool blockFC = (bi.flags & BLK_CUTS_SOLID) != 0;

    const CellFlags *F = flags[level][L].get(bid);
    const float *dataX = fch[chX][level][L].get(bid);
    const float *dataB = chB >= 0 ? fch[chB][level][L].get(bid) : nullptr;
    const float *dataD = fch[CH_DIAG][level][L].get(bid);
    const float *Wd[3] = {fch[CH_W0][level][L].get(bid), fch[CH_W1][level][L].get(bid), fch[CH_W2][level][L].get(bid)};
    float *dst = fch[chDst][level][L].alloc(bid);

    /* interior (src/msbg3.cpp:564-663) */
    for (int vz = 0, iS = 0; vz < b; vz++)
      for (int vy = 0; vy < b; vy++)
        for (int vx = 0; vx < b; vx += S, iS++) {
          const int v0 = vx + b * vy + b2 * vz;
          bool useFC = false, allNonFluid = true;
          for (int k = 0; k < S; k++) {
            const CellFlags f = F[v0 + k];
            if (f & (CELL_SOLID | CELL_PARTIAL_SOLID)) useFC = true;
            if (!(f & (CELL_SOLID | CELL_VOID))) allNonFluid = false;
          }
          segF[iS] = (uint8_t)((useFC ? 1 : 0) | (allNonFluid ? 2 : 0));
          blockHasSkippable |= allNonFluid;
          for (int k = 0; k < S; k++) X[H(vx + k, vy, vz)] = (allNonFluid || !dataX) ? 0.0f : dataX[v0 + k] * h;
          if (blockFC)
            for (int k = 0; k < S; k++) {
              const int v = v0 + k;
              Wt[0][v] = Wd[0][v];
              if (vx + k < b - 1) Wt[1][v] = Wd[0][v + 1];
              Wt[2][v] = Wd[1][v];
              if (vy < b - 1) Wt[3][v] = Wd[1][v + b];
              Wt[4][v] = Wd[2][v];
              if (vz < b - 1) Wt[5][v] = Wd[2][v + b2];
            }
        }

    /* border slabs (src/msbg3.cpp:665-1044) */
    int bx, by, bz;
    blockCoords(bid, bx, by, bz);
    static const int offs[3][2][3] = {{{-1, 0, 0}, {1, 0, 0}}, {{0, -1, 0}, {0, 1, 0}}, {{0, 0, -1}, {0, 0, 1}}};
    for (int iDir = 0; iDir < 3; iDir++)
      for (int iFace = 0; iFace < 2; iFace++) {
        const int *o = offs[iDir][iFace];
        const int bx2 = bx + o[0], by2 = by + o[1], bz2 = bz + o[2];
        const float *data2 = nullptr;
        const float *fa2 = nullptr;
        int level2 = level;
        bool outside = false;
        if (blockInRange(bx2, by2, bz2)) {
          const int bid2 = blockIndex(bx2, by2, bz2);
          level2 = binfo[L][bid2].level;
          if (level2 == level) {
            data2 = fch[chX][level][L].get(bid2);
            fa2 = fch[CH_W0 + iDir][level][L].get(bid2);
          } else if (level2 > level) {  // neighbour coarser
            data2 = level + 1 < MAXL ? fch[chX][level + 1][L].get(bid2) : nullptr;
            if (blockFC) fa2 = iFace ? fch[CH_W0 + iDir][level][L].get(bid2) : fch[CH_W0 + iDir][level][L].get(bid);
          } else {  // neighbour finer
            data2 = fch[chX][level - 1][L].get(bid2);
            if (blockFC) fa2 = iFace ? fch[CH_W0 + iDir][level2][L].get(bid2) : fch[CH_W0 + iDir][level2][L].get(bid);
          }
        } else {
          outside = true;
          const int domBc = domainBc[iDir][iFace];
          if (blockFC) fa2 = iFace ? nullptr : Wd[iDir];
          // +face W: zero block (non-open) or constant-one block (open)
          if (iFace && blockFC) {
            const float wv = domBc != DBC_OPEN ? 0.0f : 1.0f;
            for (int u = 0; u < b; u++)
              for (int v = 0; v < b; v++) {
                int cell;
                if (iDir == 0) cell = (b - 1) + b * u + b2 * v;
                else if (iDir == 1) cell = u + b * (b - 1) + b2 * v;
                else cell = u + b * v + b2 * (b - 1);
                Wt[2 * iDir + 1][cell] = wv;
              }
          }
          level2 = level;
        }

        if (level2 == level) {
          /* fine-fine (or domain ghost with zero data) */
          for (int u = 0; u < b; u++)
            for (int v = 0; v < b; v++) {
              int src, hidx, wcell;
              if (iDir == 0) {
                src = (iFace ? 0 : b - 1) + b * u + b2 * v;
                hidx = H(iFace ? b : -1, u, v);
                wcell = (b - 1) + b * u + b2 * v;
              } else if (iDir == 1) {
                src = u + b * (iFace ? 0 : b - 1) + b2 * v;
                hidx = H(u, iFace ? b : -1, v);
                wcell = u + b * (b - 1) + b2 * v;
              } else {
                src = u + b * v + b2 * (iFace ? 0 : b - 1);
                hidx = H(u, v, iFace ? b : -1);
                wcell = u + b * v + b2 * (b - 1);
              }
              X[hidx] = (outside || !data2) ? 0.0f : data2[src] * h;
              if (iFace && blockFC && !outside) Wt[2 * iDir + 1][wcell] = fa2 ? fa2[src] : 0.0f;
            }
        } else {
          /* mixed resolution (src/msbg3.cpp:46-161, 897-1044) */
          for (int u = 0; u < b; u++)
            for (int v = 0; v < b; v++) {
              int vx, vy, vz, hidx;
              if (iDir == 0) { vx = iFace ? b - 1 : 0; vy = u; vz = v; hidx = H(iFace ? b : -1, vy, vz); }
              else if (iDir == 1) { vx = u; vy = iFace ? b - 1 : 0; vz = v; hidx = H(vx, iFace ? b : -1, vz); }
              else { vx = u; vy = v; vz = iFace ? b - 1 : 0; hidx = H(vx, vy, iFace ? b : -1); }
              const int vid = vx + b * vy + b2 * vz;
              double sumVal = 0.0;
              if (level2 < level) {
                /* coarse block, finer neighbour: 4 children on the shared face */
                const int bh = 2 * b, bh2 = bh * bh;
                int bas0, bas;
                int d1[4][3];
                if (iDir == 0) {
                  bas0 = (8 * b2) * vz + (4 * b) * vy;
                  bas = iFace == 0 ? bas0 + (2 * b - 1) : bas0;
                  int t[4][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}};
                  memcpy(d1, t, sizeof t);
                } else if (iDir == 1) {
                  bas0 = (8 * b2) * vz + 2 * vx;
                  bas = iFace == 0 ? bas0 + (2 * b * (2 * b - 1)) : bas0;
                  int t[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {1, 0, 1}};
                  memcpy(d1, t, sizeof t);
                } else {
                  bas0 = (4 * b) * vy + 2 * vx;
                  bas = iFace == 0 ? bas0 + (4 * b2 * (2 * b - 1)) : bas0;
                  int t[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}};
                  memcpy(d1, t, sizeof t);
                }
                for (int c = 0; c < 4; c++) {
                  const int sb = d1[c][0] + d1[c][1] * bh + d1[c][2] * bh2;
                  float a = (float)(1.0 * (1. / 3.));
                  if (fa2) a *= fa2[bas0 + sb];
                  const float xv = data2 ? data2[bas + sb] : 0.0f;
                  sumVal += (double)(a * xv);
                }
              } else {
                /* fine block, coarser neighbour */
                int vid2;
                const int q = b / 2;
                if (iDir == 0) vid2 = iFace == 0 ? (b2 / 4) * (vz / 2) + q * (vy / 2) + q - 1 : (b2 / 4) * (vz / 2) + q * (vy / 2);
                else if (iDir == 1) vid2 = iFace == 0 ? (b2 / 4) * (vz / 2) + (vx / 2) + (q - 1) * q : (b2 / 4) * (vz / 2) + (vx / 2);
                else vid2 = iFace == 0 ? q * (vy / 2) + (vx / 2) + (q - 1) * (b2 / 4) : q * (vy / 2) + (vx / 2);
                int vidFA = vid;
                if (iFace == 1) {
                  if (iDir == 0) vidFA = 0 + b * vy + b2 * vz;
                  else if (iDir == 1) vidFA = vx + 0 + b2 * vz;
                  else vidFA = vx + b * vy;
                }
                float a = (float)(1.0 * 2 * (1. / 3.));
                if (fa2) a *= fa2[vidFA];
                const float xv = data2 ? data2[vid2] : 0.0f;
                sumVal += (double)(a * xv);
              }
              X[hidx] = (float)(sumVal * (double)h);
              if (blockFC) Wt[2 * iDir + iFace][vid] = 1.0f;
            }
        }
      }

    /* precision form: the smoother reads its face coefficients in 16 bits */
    if ((op & OP_RELAX) && smPrec && blockFC)
      for (int k = 0; k < 6; k++)
        for (int v = 0; v < nv; v++) Wt[k][v] = round16(Wt[k][v], smPrec);

    /* sweeps (src/msbg3.cpp:1109-1506) */
    const float HR = 1.f / h;
    const float om = omega, oneMinusOm = 1.0f - om;
    const int iters = (op & OP_RELAX) ? nIter : 1;
    const int nSeg = b / S;
    float Y[8];
    for (int it = 0; it < iters; it++) {
      const bool rev = (op & OP_RELAX) && (reverse == ((it & 1) != 0));
      for (int zz = 0; zz < b; zz++)
        for (int yy = 0; yy < b; yy++)
          for (int ss = 0; ss < nSeg; ss++) {
            const int vz = rev ? b - 1 - zz : zz, vy = rev ? b - 1 - yy : yy, sg = rev ? nSeg - 1 - ss : ss;
            const int vx0 = sg * S, v0 = vx0 + b * vy + b2 * vz;
            const int iS = v0 / S;
            if (blockHasSkippable && (segF[iS] & 2)) continue;
            const bool weighted = blockFC && segF[iS];
            for (int k = 0; k < S; k++) {
              const int v = v0 + k, x = vx0 + k;
              const float F0 = X[H(x, vy, vz)];
              const float F1 = X[H(x - 1, vy, vz)], F2 = X[H(x + 1, vy, vz)];
              const float F3 = X[H(x, vy - 1, vz)], F4 = X[H(x, vy + 1, vz)];
              const float F5 = X[H(x, vy, vz - 1)], F6 = X[H(x, vy, vz + 1)];
              const float D = (op & OP_RELAX) ? round16(dataD[v], smPrec) : dataD[v];
              float Sm;
              if (weighted) {
                Sm = F1 * Wt[0][v];
                Sm = fmaf(F2, Wt[1][v], Sm);
                Sm = fmaf(F3, Wt[2][v], Sm);
                Sm = fmaf(F4, Wt[3][v], Sm);
                Sm = fmaf(F5, Wt[4][v], Sm);
                Sm = fmaf(F6, Wt[5][v], Sm);
              } else {
                Sm = F1 + F2; Sm = Sm + F3; Sm = Sm + F4; Sm = Sm + F5; Sm = Sm + F6;
              }
              float y;
              if (op & OP_MULTIPLY) {
                y = (D != 0.0f) ? fmaf(1.f / D, F0, -Sm) : 0.0f;
                if (op & OP_RESIDUAL) y = (dataB ? dataB[v] : 0.0f) - y;
              } else {
                const float B = dataB ? dataB[v] : 0.0f;
                y = (D != 0.0f) ? fmaf(om * (B + Sm), D, oneMinusOm * F0) : F0;
              }
              Y[k] = y;
            }
            for (int k = 0; k < S; k++) {
              const int v = v0 + k, x = vx0 + k;
              if (op & OP_MULTIPLY) dst[v] = Y[k];
              else if (it < iters - 1) X[H(x, vy, vz)] = Y[k];
              else dst[v] = Y[k] * HR;
            }
          }
    }
  }

  /* src/msbg3.cpp:1757-1912 */
  void multiply(bool residual, int L, int chX, int chB, int chDst)
  {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nThreads)
    for (int bid = 0; bid < nBlocks; bid++) {
      const BlockInfo &bi = binfo[L][bid];
      if (bi.flags & (BLK_NO_FLUID | BLK_FIXED)) continue;
      processBlock(OP_MULTIPLY | (residual ? OP_RESIDUAL : 0), false, L, bid, chX, residual ? chB : -1, chDst, 1);
    }
  }

  /* src/msbg3.cpp:2171-2369 */
  void relax(bool reverse, int L, int chX, int chB, int nIter)
  {
    const std::vector<int> &list = relaxList[L];
    for (int it = 0; it < nIter; it++)
      for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nThreads)
        for (int i = 0; i < (int)list.size(); i++) {
          const int bid = list[i];
          const BlockInfo &bi = binfo[L][bid];
          if (bi.flags & BLK_FIXED) continue;
          int bx, by, bz;
          blockCoords(bid, bx, by, bz);
          bool skip = ((bx + by + bz) & 1) == pass;
          if (reverse) skip = !skip;
          if (skip) continue;
          processBlock(OP_RELAX, reverse, L, bid, chX, chB, chX, nBlockIter);
        }
      }
  }

  /* Precision form: one block of the coarsest MG level's relax in FP64 --
   * processBlock's OP_RELAX path with x and the arithmetic in double and b, W,
   * D read as FP32.  At the coarsest MG level every block has the coarsest
   * level, so a halo face is a same-level neighbour's h*x (zero outside the
   * domain) and a +face weight the neighbour's (0 or 1 outside, by the
   * boundary kind). */
  void processBlock64(bool reverse, int L, int bid, int chB, int nIter)
  {
    const BlockInfo &bi = binfo[L][bid];
    const int level = bi.level, b = bsxL(level), b2 = b * b, nv = b2 * b, S = simdWidth(b);
    const double h = (double)(1 << level);
    const int hs = b + 2, hs2 = hs * hs;
    auto H = [&](int x, int y, int z) { return (x + 1) + hs * (y + 1) + hs2 * (z + 1); };
    std::vector<double> X((size_t)hs * hs * hs, 0.0);
    std::vector<double> Wt[6];
    for (int k = 0; k < 6; k++) Wt[k].assign(nv, 0.0);
    std::vector<uint8_t> skipSeg(nv / S, 0);
    const CellFlags *F = flags[level][L].get(bid);
    const float *dataB = fch[chB][level][L].get(bid);
    const float *dataD = fch[CH_DIAG][level][L].get(bid);
    const float *Wd[3] = {fch[CH_W0][level][L].get(bid), fch[CH_W1][level][L].get(bid), fch[CH_W2][level][L].get(bid)};
    std::vector<double> &xb = x64[bid];
    for (int vz = 0, iS = 0; vz < b; vz++)
      for (int vy = 0; vy < b; vy++)
        for (int vx = 0; vx < b; vx += S, iS++) {
          const int v0 = vx + b * vy + b2 * vz;
          bool allNonFluid = true;
          for (int k = 0; k < S; k++)
            if (!(F[v0 + k] & (CELL_SOLID | CELL_VOID))) allNonFluid = false;
          skipSeg[iS] = allNonFluid;
          for (int k = 0; k < S; k++) {
            const int v = v0 + k;
            X[H(vx + k, vy, vz)] = allNonFluid ? 0.0 : xb[v] * h;
            Wt[0][v] = Wd[0][v];
            if (vx + k < b - 1) Wt[1][v] = Wd[0][v + 1];
            Wt[2][v] = Wd[1][v];
            if (vy < b - 1) Wt[3][v] = Wd[1][v + b];
            Wt[4][v] = Wd[2][v];
            if (vz < b - 1) Wt[5][v] = Wd[2][v + b2];
          }
        }
    int bx, by, bz;
    blockCoords(bid, bx, by, bz);
    static const int offs[3][2][3] = {{{-1, 0, 0}, {1, 0, 0}}, {{0, -1, 0}, {0, 1, 0}}, {{0, 0, -1}, {0, 0, 1}}};
    for (int iDir = 0; iDir < 3; iDir++)
      for (int iFace = 0; iFace < 2; iFace++) {
        const int *o = offs[iDir][iFace];
        const int bx2 = bx + o[0], by2 = by + o[1], bz2 = bz + o[2];
        const bool inside = blockInRange(bx2, by2, bz2);
        const int bid2 = inside ? blockIndex(bx2, by2, bz2) : -1;
        const float *fa2 = inside ? fch[CH_W0 + iDir][level][L].get(bid2) : nullptr;
        const float wOut = domainBc[iDir][iFace] != DBC_OPEN ? 0.0f : 1.0f;
        for (int u = 0; u < b; u++)
          for (int v = 0; v < b; v++) {
            int src, hidx, wcell;
            if (iDir == 0) {
              src = (iFace ? 0 : b - 1) + b * u + b2 * v;
              hidx = H(iFace ? b : -1, u, v);
              wcell = (b - 1) + b * u + b2 * v;
            } else if (iDir == 1) {
              src = u + b * (iFace ? 0 : b - 1) + b2 * v;
              hidx = H(u, iFace ? b : -1, v);
              wcell = u + b * (b - 1) + b2 * v;
            } else {
              src = u + b * v + b2 * (iFace ? 0 : b - 1);
              hidx = H(u, v, iFace ? b : -1);
              wcell = u + b * v + b2 * (b - 1);
            }
            X[hidx] = inside ? x64[bid2][src] * h : 0.0;
            if (iFace) Wt[2 * iDir + 1][wcell] = inside ? (fa2 ? fa2[src] : 0.0f) : wOut;
          }
      }
    const double om = (double)omega, omm = 1.0 - om, HR = 1.0 / h;
    const int nSeg = b / S;
    double Y[8];
    for (int it = 0; it < nIter; it++) {
      const bool rev = reverse == ((it & 1) != 0);
      for (int zz = 0; zz < b; zz++)
        for (int yy = 0; yy < b; yy++)
          for (int ss = 0; ss < nSeg; ss++) {
            const int vz = rev ? b - 1 - zz : zz, vy = rev ? b - 1 - yy : yy, sg = rev ? nSeg - 1 - ss : ss;
            const int vx0 = sg * S, v0 = vx0 + b * vy + b2 * vz;
            if (skipSeg[v0 / S]) continue;
            for (int k = 0; k < S; k++) {
              const int v = v0 + k, x = vx0 + k;
              const double F0 = X[H(x, vy, vz)];
              double Sm = X[H(x - 1, vy, vz)] * Wt[0][v];
              Sm = std::fma(X[H(x + 1, vy, vz)], Wt[1][v], Sm);
              Sm = std::fma(X[H(x, vy - 1, vz)], Wt[2][v], Sm);
              Sm = std::fma(X[H(x, vy + 1, vz)], Wt[3][v], Sm);
              Sm = std::fma(X[H(x, vy, vz - 1)], Wt[4][v], Sm);
              Sm = std::fma(X[H(x, vy, vz + 1)], Wt[5][v], Sm);
              const double D = dataD[v];
              const double B = dataB ? (double)dataB[v] : 0.0;
              Y[k] = D != 0.0 ? std::fma(om * (B + Sm), D, omm * F0) : F0;
            }
            for (int k = 0; k < S; k++) {
              if (it < nIter - 1) X[H(vx0 + k, vy, vz)] = Y[k];
              else xb[v0 + k] = Y[k] * HR;
            }
          }
    }
  }

  /* Precision form: the coarsest MG level's "solve" (nCoarse relax
   * iterations, src/msbg3.cpp:2423-2430) in FP64 from x = 0; chX receives x
   * rounded to FP32 on the relaxed blocks. */
  void relaxCoarse64(int L, int chX, int chB, int nIter)
  {
    x64.assign(nBlocks, std::vector<double>());
    for (int bid = 0; bid < nBlocks; bid++) x64[bid].assign((size_t)bsxL(binfo[L][bid].level) * bsxL(binfo[L][bid].level) *
                                                                  bsxL(binfo[L][bid].level), 0.0);
    const std::vector<int> &list = relaxList[L];
    for (int it = 0; it < nIter; it++)
      for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nThreads)
        for (int i = 0; i < (int)list.size(); i++) {
          const int bid = list[i];
          const BlockInfo &bi = binfo[L][bid];
          if (bi.flags & BLK_FIXED) continue;
          int bx, by, bz;
          blockCoords(bid, bx, by, bz);
          if (((bx + by + bz) & 1) == pass) continue;
          processBlock64(false, L, bid, chB, nBlockIter);
        }
      }
    for (int bid : list) {
      const BlockInfo &bi = binfo[L][bid];
      float *dst = fch[chX][bi.level][L].alloc(bid);
      const int nv = bsxL(bi.level) * bsxL(bi.level) * bsxL(bi.level);
      for (int v = 0; v < nv; v++) dst[v] = (float)x64[bid][v];
    }
  }

  /* src/msbg3.cpp:2375-2399 */
  void zeroChannelInPlace(int ch, int L)
  {
    for (int level = 0; level < nLevels; level++) {
      if (level < L) continue;
      Grid<float> &g = fch[ch][level][L];
      for (int bid = 0; bid < nBlocks; bid++) {
        if (binfo[L][bid].level != level) continue;
        float *d = g.get(bid);
        if (d) memset(d, 0, sizeof(float) * g.nvox());
        else g.alloc(bid);
      }
    }
  }

  /* src/msbg4.cpp:1513-1915, SIMPLE_AVERAGE | ALL_CELLS */
  void downsample(int Lc, int chSrc, int chDst)
  {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nThreads)
    for (int bid = 0; bid < nBlocks; bid++) {
      const int lev0 = level0[bid], level = std::max(lev0, Lc);
      Grid<float> &dst = fch[chDst][level][Lc];
      if (lev0 >= Lc) {
        Grid<float> &srcHi = fch[chSrc][level][Lc - 1];
        if (!srcHi.isValue(bid)) {
          dst.map[bid] = const_cast<float *>(dst.emptyBlock);   // copyConstBlock of an empty block
        } else {
          dst.map[bid] = srcHi.map[bid];                       // external block reference
        }
        continue;
      }
      Grid<float> &srcHi = fch[chSrc][level - 1][Lc - 1];
      Grid<float> &srcLo = fch[chSrc][level][Lc - 1];
      float *d = dst.alloc(bid);
      int bx, by, bz;
      blockCoords(bid, bx, by, bz);
      const int b = bsxL(level), bh = bsxL(level - 1);
      for (int vz = 0, vid = 0; vz < b; vz++)
        for (int vy = 0; vy < b; vy++)
          for (int vx = 0; vx < b; vx++, vid++) {
            const int x = bx * b + vx, y = by * b + vy, z = bz * b + vz;
            float val = 0.0f;
            for (int k = 0; k < 2; k++)
              for (int j = 0; j < 2; j++)
                for (int i = 0; i < 2; i++) {
                  const int x2 = 2 * x + i, y2 = 2 * y + j, z2 = 2 * z + k;
                  const int bid2 = (x2 / bh) + nbx * ((y2 / bh) + nby * (z2 / bh));
                  const float *blk = srcHi.get(bid2);
                  float f;
                  if (blk) f = blk[(x2 % bh) + bh * ((y2 % bh) + bh * (z2 % bh))];
                  else f = srcLo.getValue(x2 / 2, y2 / 2, z2 / 2);
                  val = (float)((double)val + (double)f * 0.125);
                }
            d[vid] = val;
          }
    }
  }

  /* src/msbg4.cpp:1921-2029 */
  void prolongate(int Lf, int ch)
  {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nThreads)
    for (int bid = 0; bid < nBlocks; bid++) {
      const int lev0 = level0[bid], level = std::max(lev0, Lf);
      const CellFlags *fl = flags[level][Lf].get(bid);
      float *dst = fch[ch][level][Lf].alloc(bid);
      const int b = bsxL(level);
      if (lev0 >= Lf + 1) {
        Grid<float> &src = fch[ch][level][Lf + 1];
        if (!src.isValue(bid)) continue;
        const float *s = src.get(bid);
        for (int v = 0; v < b * b * b; v++) {
          if (fl && (fl[v] & (CELL_SOLID | CELL_VOID))) continue;
          dst[v] = dst[v] + s[v];
        }
        continue;
      }
      Grid<float> &src = fch[ch][Lf + 1][Lf + 1];
      int bx, by, bz;
      blockCoords(bid, bx, by, bz);
      for (int vz = 0, v = 0; vz < b; vz++)
        for (int vy = 0; vy < b; vy++)
          for (int vx = 0; vx < b; vx++, v++) {
            if (fl && (fl[v] & (CELL_SOLID | CELL_VOID))) continue;
            const int x = bx * b + vx, y = by * b + vy, z = bz * b + vz;
            dst[v] = dst[v] + src.getValue(x >> 1, y >> 1, z >> 1);
          }
    }
  }

  /* src/msbg3.cpp:2412-2465 */
  void vCycle(int L, int chX, int chB, int chR, int nPre, int nPost, int nCoarse)
  {
    if (L >= nLevels - 1) {
      if (smPrec && bsxL(L) <= 16) relaxCoarse64(L, chX, chB, nCoarse);   // precision form: FP64 coarsest solve
      else relax(false, L, chX, chB, nCoarse);
      return;
    }
    relax(false, L, chX, chB, nPre);
    multiply(true, L, chX, chB, chR);
    zeroChannelInPlace(chB, L + 1);
    downsample(L + 1, chR, chB);
    zeroChannelInPlace(chX, L + 1);
    vCycle(L + 1, chX, chB, chR, nPre, nPost, nCoarse);
    prolongate(L, chX);
    relax(true, L, chX, chB, nPost);
  }

  /* copyChannel at MG0 over all levels (src/msbg.cpp:3282-3505) */
  void copyChannel(int src, int dst)
  {
    for (int level = 0; level < nLevels; level++)
      for (int bid = 0; bid < nBlocks; bid++) {
        if (binfo[0][bid].level != level) continue;
        const float *s = fch[src][level][0].get(bid);
        if (!s) continue;
        float *d = fch[dst][level][0].alloc(bid);
        memcpy(d, s, sizeof(float) * fch[dst][level][0].nvox());
      }
  }

  long double dotLevel0(int a, int b) const
  {
    long double s = 0;
    const Grid<float> &A = fch[a][0][0], &B = fch[b][0][0];
    for (int bid = 0; bid < nBlocks; bid++) {
      const float *da = A.get(bid), *db = B.get(bid);
      if (!da || !db) continue;
      for (int i = 0; i < A.nvox(); i++) s += (double)da[i] * (double)db[i];
    }
    return s;
  }

  struct SolveParams { int nPre, nPost, nCoarse, maxIter; float tol; };

  /* src/msbg3.cpp:3098-3259 */
  void solvePCG(const SolveParams &sp, int *outIters, float *outRel, std::vector<double> *hist)
  {
    const int chX = CH_X, chB = CH_B, chR = CH_R, chZ = CH_Z, chD = CH_D, chQ = CH_Q;
    copyChannel(chB, chR);
    const long double bNormSq = dotLevel0(chR, chR);
    if (bNormSq <= 1e-30) { *outIters = 0; *outRel = 0; return; }
    zeroChannelInPlace(chZ, 0);
    vCycle(0, chZ, chR, chQ, sp.nPre, sp.nPost, sp.nCoarse);
    copyChannel(chZ, chD);
    long double rho = dotLevel0(chR, chZ);
    const double tolSq = (double)sp.tol * (double)sp.tol;
    int conv = -1;
    for (int iter = 0; iter < sp.maxIter; iter++) {
      multiply(false, 0, chD, -1, chQ);
      const long double dq = dotLevel0(chD, chQ);
      if (!(dq > 1e-30) || !(rho > 1e-30) || !std::isfinite((double)dq) || !std::isfinite((double)rho)) break;
      const double alphaD = (double)rho / (double)dq;
      if (!std::isfinite(alphaD)) break;
      const float alpha = (float)alphaD;
      long double rr = 0;
      {
        Grid<float> &X = fch[chX][0][0], &R = fch[chR][0][0], &D = fch[chD][0][0], &Q = fch[chQ][0][0];
        for (int bid = 0; bid < nBlocks; bid++) {
          float *dx = X.get(bid), *dr = R.get(bid), *dd = D.get(bid), *dqq = Q.get(bid);
          if (!dx || !dr) continue;
          for (int i = 0; i < X.nvox(); i++) {
            if (dd) dx[i] = fmaf(alpha, dd[i], dx[i]);
            if (dqq) dr[i] = fmaf(-alpha, dqq[i], dr[i]);
            rr += (double)dr[i] * (double)dr[i];
          }
        }
      }
      if (hist) hist->push_back(sqrt((double)rr / (double)bNormSq));
      if (rr <= tolSq * bNormSq) { conv = iter + 1; break; }
      zeroChannelInPlace(chZ, 0);
      vCycle(0, chZ, chR, chQ, sp.nPre, sp.nPost, sp.nCoarse);
      const long double rhoNew = dotLevel0(chR, chZ);
      if (!(rhoNew > 1e-30) || !std::isfinite((double)rhoNew)) break;
      const double betaD = (double)rhoNew / (double)rho;
      if (!std::isfinite(betaD)) break;
      const float beta = (float)betaD;
      {
        Grid<float> &Z = fch[chZ][0][0], &D = fch[chD][0][0];
        for (int bid = 0; bid < nBlocks; bid++) {
          float *dz = Z.get(bid), *dd = D.get(bid);
          if (!dz || !dd) continue;
          for (int i = 0; i < Z.nvox(); i++) dd[i] = fmaf(beta, dd[i], dz[i]);
        }
      }
      rho = rhoNew;
    }
    const long double fin = dotLevel0(chR, chR);
    *outRel = (float)sqrt((double)fin / (double)bNormSq);
    *outIters = conv > 0 ? conv : sp.maxIter;
  }
};

/* FP64 reference for the precision form: the same level-0 operator as the
 * PCG of the reference applies (Q4: coarse-block unknowns held at 0), in
 * double -- (A x)_i = x_i / D_i - sum_nb W x_nb over same-level level-0
 * face neighbours (src/msbg3.cpp:1282-1291) -- solved with textbook
 * Jacobi-preconditioned CG (M^-1 = D, zero on non-fluid cells). */
struct Fp64Solver {
  Oracle &o;
  int b, n;
  std::vector<int> lvl0;        // level-0 block ids
  std::vector<int> slot;        // bid -> index in lvl0 or -1
  explicit Fp64Solver(Oracle &o_) : o(o_), b(o_.bsx0), n(o_.bsx0 * o_.bsx0 * o_.bsx0)
  {
    slot.assign(o.nBlocks, -1);
    for (int bid = 0; bid < o.nBlocks; bid++)
      if (o.binfo[0][bid].level == 0) { slot[bid] = (int)lvl0.size(); lvl0.push_back(bid); }
  }
  void apply(const std::vector<double> &x, std::vector<double> &y)
  {
    const int b2 = b * b;
#pragma omp parallel for schedule(dynamic, 4) num_threads(o.nThreads)
    for (int i = 0; i < (int)lvl0.size(); i++) {
      const int bid = lvl0[i];
      int bx, by, bz;
      o.blockCoords(bid, bx, by, bz);
      const float *D = o.fch[CH_DIAG][0][0].get(bid);
      const float *W[3] = {o.fch[CH_W0][0][0].get(bid), o.fch[CH_W1][0][0].get(bid), o.fch[CH_W2][0][0].get(bid)};
      const double *xb = &x[(size_t)i * n];
      for (int v = 0; v < n; v++) {
        double yv = 0.0;
        if (D[v] != 0.0f) {
          const int vx = v % b, vy = (v / b) % b, vz = v / b2;
          const int c[3] = {vx, vy, vz}, st[3] = {1, b, b2};
          const int nbOff[3] = {1, o.nbx, o.nbx * o.nby};
          const int bc[3] = {bx, by, bz}, nbMax[3] = {o.nbx, o.nby, o.nbz};
          double s = 0.0;
          for (int d = 0; d < 3; d++) {
            /* -d face: own W */
            if (c[d] > 0) s += (double)W[d][v] * xb[v - st[d]];
            else if (bc[d] > 0 && slot[bid - nbOff[d]] >= 0)
              s += (double)W[d][v] * x[(size_t)slot[bid - nbOff[d]] * n + v + (b - 1) * st[d]];
            /* +d face: W of the next cell */
            if (c[d] < b - 1) s += (double)W[d][v + st[d]] * xb[v + st[d]];
            else if (bc[d] < nbMax[d] - 1 && slot[bid + nbOff[d]] >= 0) {
              const int nb = bid + nbOff[d];
              const float *Wn = o.fch[CH_W0 + d][0][0].get(nb);
              s += (double)Wn[v - (b - 1) * st[d]] * x[(size_t)slot[nb] * n + v - (b - 1) * st[d]];
            }
          }
          yv = xb[v] / (double)D[v] - s;
        }
        y[(size_t)i * n + v] = yv;
      }
    }
  }
  int solve(double tol, int maxit, std::vector<double> &x, double *relOut)
  {
    const size_t N = lvl0.size() * (size_t)n;
    std::vector<double> r(N), z(N), p(N), q(N), dinv(N, 0.0);
    x.assign(N, 0.0);
    for (size_t i = 0; i < lvl0.size(); i++) {
      const float *B = o.fch[CH_B][0][0].get(lvl0[i]);
      const float *D = o.fch[CH_DIAG][0][0].get(lvl0[i]);
      for (int v = 0; v < n; v++) {
        r[i * n + v] = B ? (double)B[v] : 0.0;
        dinv[i * n + v] = (double)D[v];
      }
    }
    double bb = 0;
    for (size_t k = 0; k < N; k++) bb += r[k] * r[k];
    if (bb <= 0) { *relOut = 0; return 0; }
    for (size_t k = 0; k < N; k++) z[k] = dinv[k] * r[k];
    p = z;
    double rz = 0;
    for (size_t k = 0; k < N; k++) rz += r[k] * z[k];
    int it = 0;
    double rr = bb;
    for (; it < maxit && rr > tol * tol * bb; it++) {
      apply(p, q);
      double pq = 0;
      for (size_t k = 0; k < N; k++) pq += p[k] * q[k];
      const double al = rz / pq;
      rr = 0;
      for (size_t k = 0; k < N; k++) {
        x[k] += al * p[k];
        r[k] -= al * q[k];
        rr += r[k] * r[k];
      }
      for (size_t k = 0; k < N; k++) z[k] = dinv[k] * r[k];
      double rzn = 0;
      for (size_t k = 0; k < N; k++) rzn += r[k] * z[k];
      const double be = rzn / rz;
      rz = rzn;
      for (size_t k = 0; k < N; k++) p[k] = z[k] + be * p[k];
    }
    /* true residual */
    apply(x, q);
    double tr = 0;
    for (size_t i = 0; i < lvl0.size(); i++) {
      const float *B = o.fch[CH_B][0][0].get(lvl0[i]);
      for (int v = 0; v < n; v++) {
        const double e = (B ? (double)B[v] : 0.0) - q[i * n + v];
        tr += e * e;
      }
    }
    *relOut = sqrt(tr / bb);
    return it;
  }
};

/* Caller-side level policy (src/flip_sim.cpp:480-548): 26-neighbourhood BFS
 * block distance from seed blocks; d <= d0 -> 0, d == d0 + 1 -> 1, else the
 * coarsest level.  Seeds: blocks holding a liquid cell (mode 0) or a liquid
 * cell with a non-liquid face neighbour inside the domain (mode 1). */
static void levelMapFromScene(const msbgx_scene *s, int bsx0, int nLevels, int mode, int d0, int *out)
{
  const int nbx = s->nx / bsx0, nby = s->ny / bsx0, nbz = s->nz / bsx0, nb = nbx * nby * nbz;
  std::vector<int> dist(nb, 0x7fffffff);
#pragma omp parallel for schedule(dynamic, 1)
  for (int bid = 0; bid < nb; bid++) {
    const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
    bool seed = false;
    for (int z = bz * bsx0; z < (bz + 1) * bsx0 && !seed; z++)
      for (int y = by * bsx0; y < (by + 1) * bsx0 && !seed; y++)
        for (int x = bx * bsx0; x < (bx + 1) * bsx0 && !seed; x++) {
          if (!msbgx_scene_is_liquid(s, x, y, z)) continue;
          if (mode == 0) { seed = true; break; }
          const int nb6[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
          for (int q = 0; q < 6; q++) {
            const int x2 = x + nb6[q][0], y2 = y + nb6[q][1], z2 = z + nb6[q][2];
            if (x2 < 0 || y2 < 0 || z2 < 0 || x2 >= s->nx || y2 >= s->ny || z2 >= s->nz) continue;
            if (!msbgx_scene_is_liquid(s, x2, y2, z2)) { seed = true; break; }
          }
        }
    if (seed) dist[bid] = 0;
  }
  std::queue<int> q;
  for (int b = 0; b < nb; b++) if (dist[b] == 0) q.push(b);
  const int coarsest = nLevels - 1;
  while (!q.empty()) {
    const int bid = q.front(); q.pop();
    const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
    const int d = dist[bid];
    if (d >= d0 + coarsest) continue;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          if (!dx && !dy && !dz) continue;
          const int x2 = bx + dx, y2 = by + dy, z2 = bz + dz;
          if (x2 < 0 || y2 < 0 || z2 < 0 || x2 >= nbx || y2 >= nby || z2 >= nbz) continue;
          const int b2 = x2 + nbx * (y2 + nby * z2);
          if (dist[b2] <= d + 1) continue;
          dist[b2] = d + 1;
          q.push(b2);
        }
  }
  for (int b = 0; b < nb; b++) {
    const int d = dist[b];
    if (d <= d0) out[b] = 0;
    else if (nLevels >= 2 && d == d0 + 1) out[b] = 1;
    else out[b] = coarsest;
  }
}

}  // namespace

/* ====================================================================== C ABI
 * Entry points mirror the reference facade: create / regularize / set map /
 * prepare / solve, plus kernel-level hooks used by the parity tests.       */
extern "C" {

void *orc_create(int sx0, int sy0, int sz0, int bsx0, int nLevels)
{
  if (bsx0 < 2 || (bsx0 & (bsx0 - 1)) || bsx0 > 32 || nLevels < 1 || nLevels > MAXL) return nullptr;
  if (sx0 % bsx0 || sy0 % bsx0 || sz0 % bsx0) return nullptr;
  if ((bsx0 >> (nLevels - 1)) < 2) return nullptr;
  return new Oracle(sx0, sy0, sz0, bsx0, nLevels);
}
void orc_destroy(void *h) { delete (Oracle *)h; }
void orc_set_threads(void *h, int n) { ((Oracle *)h)->nThreads = n > 0 ? n : 1; }
int orc_num_blocks(void *h) { return ((Oracle *)h)->nBlocks; }
void orc_set_domain_bc(void *h, int l, int r, int b, int t, int f, int k)
{
  Oracle *o = (Oracle *)h;
  o->domainBc[0][0] = l; o->domainBc[0][1] = r; o->domainBc[1][0] = b;
  o->domainBc[1][1] = t; o->domainBc[2][0] = f; o->domainBc[2][1] = k;
}
void orc_regularize_refinement_map(void *h, const int *in, int *out) { ((Oracle *)h)->regularizeRefinementMap(in, out); }
void orc_set_refinement_map(void *h, const int *levels) { ((Oracle *)h)->setRefinementMap(levels, true); }
void orc_level_map_from_scene(const msbgx_scene *s, int bsx0, int nLevels, int mode, int d0, int *out)
{
  levelMapFromScene(s, bsx0, nLevels, mode, d0, out);
}

void orc_prepare_scene(void *h, const msbgx_scene *s)
{
  Callbacks cb{sceneCellType, sceneFaceCoeff, sceneRhs, (void *)s};
  ((Oracle *)h)->prepare(cb, true);
}
void orc_prepare_dense(void *h, int nx, int ny, int nz, const uint16_t *flags, const float *wu, const float *wv,
                       const float *ww, const float *rhs)
{
  DenseInput d{nx, ny, nz, flags, {wu, wv, ww}, rhs};
  Callbacks cb{denseCellType, denseFaceCoeff, denseRhs, &d};
  ((Oracle *)h)->prepare(cb, true);
}
float orc_omega(void *h) { return ((Oracle *)h)->omega; }
/* Precision form: relax with 16-bit smoother operands (0 FP32, 1 FP16, 2 BF16). */
void orc_set_smoother_precision(void *h, int prec) { ((Oracle *)h)->smPrec = prec; }
/* The rounding the precision form applies, element-wise (test hook). */
void orc_round16(int prec, const float *in, float *out, int64_t n)
{
  for (int64_t i = 0; i < n; i++) out[i] = round16(in[i], prec);
}

void orc_block_info(void *h, int L, int *level, uint32_t *flags)
{
  Oracle *o = (Oracle *)h;
  for (int b = 0; b < o->nBlocks; b++) { level[b] = o->binfo[L][b].level; flags[b] = o->binfo[L][b].flags; }
}
int orc_relax_list(void *h, int L, int *out, int cap)
{
  Oracle *o = (Oracle *)h;
  const int n = (int)o->relaxList[L].size();
  for (int i = 0; i < n && i < cap; i++) out[i] = o->relaxList[L][i];
  return n;
}
/* Copy block `bid` of channel `ch` at MG level L (its own level there).
 * Channels: 0 x, 1 b, 2 r, 3 z, 4 d, 5 q, 6 tmp, 7-9 W, 10 diag, 11 flags.
 * Returns the block edge, or 0 when the block has no data. */
int orc_get_block(void *h, int ch, int L, int bid, void *dst)
{
  Oracle *o = (Oracle *)h;
  const int level = o->binfo[L][bid].level, b = o->bsxL(level), n = b * b * b;
  if (ch == 11) {
    const CellFlags *f = o->flags[level][L].get(bid);
    if (!f) return 0;
    memcpy(dst, f, sizeof(CellFlags) * n);
    return b;
  }
  const float *d = o->fch[ch][level][L].get(bid);
  if (!d) return 0;
  memcpy(dst, d, sizeof(float) * n);
  return b;
}
int orc_set_block(void *h, int ch, int L, int bid, const float *src)
{
  Oracle *o = (Oracle *)h;
  const int level = o->binfo[L][bid].level, b = o->bsxL(level);
  float *d = o->fch[ch][level][L].alloc(bid);
  memcpy(d, src, sizeof(float) * b * b * b);
  return b;
}
void orc_relax(void *h, int L, int reverse, int nIter, int chX, int chB) { ((Oracle *)h)->relax(reverse != 0, L, chX, chB, nIter); }
/* Precision form: the coarsest level's FP64 relax (from x = 0) into chX. */
void orc_relax_coarse64(void *h, int L, int nIter, int chX, int chB) { ((Oracle *)h)->relaxCoarse64(L, chX, chB, nIter); }
void orc_multiply(void *h, int L, int residual, int chX, int chB, int chDst) { ((Oracle *)h)->multiply(residual != 0, L, chX, chB, chDst); }
void orc_downsample(void *h, int Lc, int chSrc, int chDst) { ((Oracle *)h)->downsample(Lc, chSrc, chDst); }
void orc_prolongate(void *h, int Lf, int ch) { ((Oracle *)h)->prolongate(Lf, ch); }
void orc_zero_channel(void *h, int ch, int L) { ((Oracle *)h)->zeroChannelInPlace(ch, L); }
void orc_vcycle(void *h, int nPre, int nPost, int nCoarse)
{
  Oracle *o = (Oracle *)h;
  o->zeroChannelInPlace(CH_Z, 0);
  o->vCycle(0, CH_Z, CH_R, CH_Q, nPre, nPost, nCoarse);
}
void orc_copy_channel(void *h, int src, int dst) { ((Oracle *)h)->copyChannel(src, dst); }

int orc_solve(void *h, int nPre, int nPost, int nCoarse, int maxIter, float tol, float *relRes, double *hist, int histCap)
{
  Oracle *o = (Oracle *)h;
  Oracle::SolveParams sp{nPre, nPost, nCoarse, maxIter, tol};
  int it = 0;
  float rr = 0;
  std::vector<double> hv;
  o->solvePCG(sp, &it, &rr, &hv);
  if (relRes) *relRes = rr;
  if (hist) for (int i = 0; i < (int)hv.size() && i < histCap; i++) hist[i] = hv[i];
  return it;
}

/* FP64 solution of the level-0 system (see Fp64Solver); out holds
 * nBlocks * bsx0^3 doubles, block bid at bid * bsx0^3, zero for blocks that
 * are not level 0.  Returns CG iterations; *rel = FP64 true residual. */
int orc_solve_fp64(void *h, double tol, int maxit, double *out, double *rel)
{
  Oracle *o = (Oracle *)h;
  Fp64Solver s(*o);
  std::vector<double> x;
  const int it = s.solve(tol, maxit, x, rel);
  const size_t n = (size_t)o->bsx0 * o->bsx0 * o->bsx0;
  memset(out, 0, sizeof(double) * n * o->nBlocks);
  for (size_t i = 0; i < s.lvl0.size(); i++) memcpy(out + (size_t)s.lvl0[i] * n, &x[i * n], sizeof(double) * n);
  return it;
}

/* Dense export of a channel at MG0: every block at its own level, written
 * into a fine-grid array by piecewise-constant upsampling of coarse blocks. */
void orc_export_level0(void *h, int ch, float *out)
{
  Oracle *o = (Oracle *)h;
  const int sx = o->sx0, sy = o->sy0;
  for (int bid = 0; bid < o->nBlocks; bid++) {
    const int level = o->binfo[0][bid].level, b = o->bsxL(level), r = 1 << level;
    const float *d = o->fch[ch][level][0].get(bid);
    int bx, by, bz;
    o->blockCoords(bid, bx, by, bz);
    for (int z = 0; z < o->bsx0; z++)
      for (int y = 0; y < o->bsx0; y++)
        for (int x = 0; x < o->bsx0; x++) {
          const int v = (x / r) + b * ((y / r) + b * (z / r));
          out[(size_t)(bx * o->bsx0 + x) + (size_t)sx * ((size_t)(by * o->bsx0 + y) + (size_t)sy * (bz * o->bsx0 + z))] = d ? d[v] : 0.0f;
        }
  }
}

}  // extern "C"
