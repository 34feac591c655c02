/*
 * Host-side block topology and solver metadata of the MSBG FLIP pressure
 * projection: refinement-map regularisation, per-MG-level block levels and
 * flags, level-0 sampling, diagonals, coarse-level metadata and block lists.
 * Integer outputs are bit-exact with the reference; float metadata (W, D)
 * follows the reference operation order.
 */
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <queue>

#include "msbgx_internal.h"

namespace msbgx {

/* Regularisation: passes ll = 1 .. nLevels-2; a block with level >= ll whose
 * finest 26-neighbour has level ll-1 becomes level ll.  Each pass reads the
 * previous pass's map (src/msbg.cpp:1202-1420). */
void regularize(const Grid &g, const int *in, int *out)
{
  std::vector<int> prev(in, in + g.nBlocks), cur(in, in + g.nBlocks);
  for (int ll = 1; ll <= g.nLevels - 2; ll++) {
#pragma omp parallel for schedule(static)
    for (int bid = 0; bid < g.nBlocks; bid++) {
      int lev = prev[bid];
      if (lev >= ll) {
        int bx, by, bz;
        g.blockCoords(bid, bx, by, bz);
        int finest = g.nLevels - 1;
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              if ((dx | dy | dz) == 0 || !g.inRaster(bx + dx, by + dy, bz + dz)) continue;
              finest = std::min(finest, prev[g.blockIndex(bx + dx, by + dy, bz + dz)]);
            }
        if (finest == ll - 1) lev = ll;
      }
      cur[bid] = lev;
    }
    prev = cur;
  }
  std::copy(cur.begin(), cur.end(), out);
}

/* Per-MG-level block level max(map, L), BLK_DOM_BORDER and the resolution
 * border flags over the 26-neighbourhood (src/msbg.cpp:1650-1736). */
void setRefinementMap(Grid &g, const int *levels)
{
  g.map.assign(levels, levels + g.nBlocks);
  for (int L = 0; L < g.nLevels; L++) {
    MgLevel &m = g.mg[L];
    m.level.assign(g.nBlocks, 0);
    m.bflags.assign(g.nBlocks, 0);
#pragma omp parallel for schedule(static)
    for (int bid = 0; bid < g.nBlocks; bid++) {
      int bx, by, bz;
      g.blockCoords(bid, bx, by, bz);
      const int lev = std::max(g.map[bid], L);
      uint32_t f = 0;
      if (bx == 0 || by == 0 || bz == 0 || bx == g.nbx - 1 || by == g.nby - 1 || bz == g.nbz - 1) f |= BLK_DOM_BORDER;
      int hi = 0, lo = g.nLevels - 1;
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            if ((dx | dy | dz) == 0 || !g.inRaster(bx + dx, by + dy, bz + dz)) continue;
            const int l2 = std::max(g.map[g.blockIndex(bx + dx, by + dy, bz + dz)], L);
            hi = std::max(hi, l2);
            lo = std::min(lo, l2);
          }
      if (hi > lev) f |= BLK_FINE_COARSE;
      if (lo < lev) f |= BLK_COARSE_FINE;
      m.level[bid] = lev;
      m.bflags[bid] = f;
    }
  }
  for (int L = 0; L < kMaxLevels; L++) g.aliasEstablished[L] = false;
  g.prepared = false;
}

namespace {

/* Natural slot layout of one MG level: blocks in bid order. */
void naturalLayout(const Grid &g, MgLevel &m)
{
  m.off.assign(g.nBlocks, 0);
  int64_t o = 0;
  for (int bid = 0; bid < g.nBlocks; bid++) {
    m.off[bid] = o;
    const int64_t b = g.bsxL(m.level[bid]);
    o += b * b * b;
  }
  m.ncells = o;
}

/* Value of a level-`lev` grid at MG level L at level-lev coordinates: the
 * block must live at that level there, else the grid's empty value
 * (src/sbg.h:880-895, 2626-2655). */
template <typename T>
inline T gridValue(const Grid &g, const MgLevel &m, const std::vector<T> &data, int lev, int x, int y, int z, T empty)
{
  const int b = g.bsxL(lev);
  if (x < 0 || y < 0 || z < 0 || x >= (g.sx0 >> lev) || y >= (g.sy0 >> lev) || z >= (g.sz0 >> lev)) return empty;
  const int bid = g.blockIndex(x / b, y / b, z / b);
  if (m.level[bid] != lev) return empty;
  return data[m.off[bid] + (x % b) + b * ((y % b) + b * (z % b))];
}

bool isRelaxationBlock(int L, int lev, uint32_t f)
{
  return !(f & (BLK_NO_FLUID | BLK_FIXED)) &&
         (lev <= L || ((f & (BLK_COARSE_FINE | BLK_FINE_COARSE)) && lev == L + 1));
}

/* CELL_PARTIAL_SOLID on fluid cells, BLK_CUTS_SOLID, and the inverse diagonal
 * (src/msbg3.cpp:2586-2779).  Face-area channels always exist here. */
void computeDiagonal(Grid &g, int L)
{
  MgLevel &m = g.mg[L];
  m.D.assign(m.ncells, 0.0f);
#pragma omp parallel for schedule(dynamic, 4)
  for (int bid = 0; bid < g.nBlocks; bid++) {
    if (!(m.bflags[bid] & BLK_EXISTS)) continue;
    const int lev = m.level[bid], b = g.bsxL(lev);
    const int64_t o = m.off[bid];
    bool anyFluid = false;
    int bx, by, bz;
    g.blockCoords(bid, bx, by, bz);
    for (int vz = 0, v = 0; vz < b; vz++)
      for (int vy = 0; vy < b; vy++)
        for (int vx = 0; vx < b; vx++, v++) {
          uint16_t &f = m.flags[o + v];
          f &= (uint16_t)~CELL_PARTIAL_SOLID;
          if (f & (CELL_SOLID | CELL_VOID | CELL_FIXED)) { m.D[o + v] = 0.0f; continue; }
          f |= CELL_PARTIAL_SOLID;
          anyFluid = true;
          float sum;
          if (L == 0) {
            const int gx = bx * b + vx, gy = by * b + vy, gz = bz * b + vz;
            sum = gridValue(g, m, m.W[0], lev, gx, gy, gz, 0.0f);
            sum = sum + gridValue(g, m, m.W[0], lev, gx + 1, gy, gz, 0.0f);
            sum = sum + gridValue(g, m, m.W[1], lev, gx, gy, gz, 0.0f);
            sum = sum + gridValue(g, m, m.W[1], lev, gx, gy + 1, gz, 0.0f);
            sum = sum + gridValue(g, m, m.W[2], lev, gx, gy, gz, 0.0f);
            sum = sum + gridValue(g, m, m.W[2], lev, gx, gy, gz + 1, 0.0f);
          } else {
            sum = m.W[0][o + v] + m.W[1][o + v];
            sum = sum + m.W[2][o + v];
            sum *= 2.0f;
          }
          m.D[o + v] = (sum > 1e-12f) ? 1.f / sum : 0.0f;
        }
    if (anyFluid) m.bflags[bid] |= BLK_CUTS_SOLID;
    else m.bflags[bid] &= ~BLK_CUTS_SOLID;
  }
}

/* Coarse metadata at MG level Lc sampled from the same-level grids of MG
 * level Lc-1 at doubled coordinates (src/msbg3.cpp:2882-3007). */
void restrictMeta(Grid &g, int Lc)
{
  MgLevel &c = g.mg[Lc];
  const MgLevel &f = g.mg[Lc - 1];
  c.flags.assign(c.ncells, 0);
  for (int d = 0; d < 3; d++) c.W[d].assign(c.ncells, 0.0f);
#pragma omp parallel for schedule(dynamic, 4)
  for (int bid = 0; bid < g.nBlocks; bid++) {
    const int lev = c.level[bid], b = g.bsxL(lev);
    const int64_t o = c.off[bid];
    int bx, by, bz;
    g.blockCoords(bid, bx, by, bz);
    bool any = false;
    for (int vz = 0, v = 0; vz < b; vz++)
      for (int vy = 0; vy < b; vy++)
        for (int vx = 0; vx < b; vx++, v++) {
          const int fx = (bx * b + vx) * 2, fy = (by * b + vy) * 2, fz = (bz * b + vz) * 2;
          bool fl = false, so = false;
          for (int dz = 0; dz < 2; dz++)
            for (int dy = 0; dy < 2; dy++)
              for (int dx = 0; dx < 2; dx++) {
                const uint16_t cf = gridValue(g, f, f.flags, lev, fx + dx, fy + dy, fz + dz, CELL_EMPTY_VAL);
                if (!(cf & (CELL_SOLID | CELL_VOID | CELL_FIXED))) fl = true;
                if (cf & CELL_SOLID) so = true;
              }
          const uint16_t cv = fl ? 0 : (so ? CELL_SOLID : CELL_VOID);
          c.flags[o + v] = cv;
          if (!(cv & (CELL_SOLID | CELL_VOID | CELL_FIXED))) any = true;
          for (int d = 0; d < 3; d++) {
            float sum = 0.f;
            for (int j = 0; j < 2; j++)
              for (int k = 0; k < 2; k++) {
                int x = fx, y = fy, z = fz;
                if (d == 0) { y += j; z += k; }
                if (d == 1) { x += j; z += k; }
                if (d == 2) { x += j; y += k; }
                sum = sum + gridValue(g, f, f.W[d], lev, x, y, z, 0.0f);
              }
            c.W[d][o + v] = sum * 0.25f;
          }
        }
    c.bflags[bid] |= BLK_EXISTS;
    if (any) c.bflags[bid] &= ~BLK_NO_FLUID;
    else c.bflags[bid] |= BLK_NO_FLUID;
    c.bflags[bid] &= ~BLK_FIXED;
  }
}

void buildLists(Grid &g, int L)
{
  MgLevel &m = g.mg[L];
  m.relax.clear();
  for (int l = 0; l < kMaxLevels; l++) {
    m.act[l].clear();
    m.relaxCL[0][l].clear();
    m.relaxCL[1][l].clear();
  }
  for (int bid = 0; bid < g.nBlocks; bid++) {
    const uint32_t f = m.bflags[bid];
    if (!(f & BLK_EXISTS)) continue;
    const int lev = m.level[bid];
    if (!(f & (BLK_NO_FLUID | BLK_FIXED))) m.act[lev].push_back(bid);
    if (isRelaxationBlock(L, lev, f)) {
      m.relax.push_back(bid);
      int bx, by, bz;
      g.blockCoords(bid, bx, by, bz);
      m.relaxCL[(bx + by + bz) & 1][lev].push_back(bid);
    }
  }
  m.anyRelax = !m.relax.empty();
}

}  // namespace

int buildMetadata(Grid &g, Sampler &s, std::vector<float> &rhs)
{
  for (int L = 0; L < g.nLevels; L++)
    for (int bid = 0; bid < g.nBlocks; bid++) g.mg[L].bflags[bid] &= ~(BLK_EXISTS | BLK_NO_FLUID | BLK_FIXED | BLK_CUTS_SOLID);

  /* MG0: sample cell types in a natural layout first, then choose the pool
   * layout with the PCG-active level-0 blocks first. */
  MgLevel &m0 = g.mg[0];
  naturalLayout(g, m0);
  std::vector<uint16_t> flagsNat(m0.ncells);
  std::vector<float> wNat[3], rhsNat;
  for (int d = 0; d < 3; d++) wNat[d].resize(m0.ncells);
  const bool hostRhs = !s.rhsOnDevice();
  if (hostRhs) rhsNat.assign(m0.ncells, 0.0f);
  const bool par = s.parallelSafe();
  std::vector<char> blockFluid(g.nBlocks, 0), blockRhs(g.nBlocks, 0);
  /* src/msbg3.cpp:2472-2579: cell types, then face coefficients, then the
   * right-hand side of level-0 blocks, all at level-local coordinates */
#pragma omp parallel for schedule(dynamic, 1) if (par)
  for (int bid = 0; bid < g.nBlocks; bid++) {
    const int lev = m0.level[bid], b = g.bsxL(lev);
    const int64_t o = m0.off[bid];
    int bx, by, bz;
    g.blockCoords(bid, bx, by, bz);
    bool liq = false;
    for (int vz = 0, v = 0; vz < b; vz++)
      for (int vy = 0; vy < b; vy++)
        for (int vx = 0; vx < b; vx++, v++) {
          uint16_t fl = s.cellType(bx * b + vx, by * b + vy, bz * b + vz);
          fl &= (CELL_SOLID | CELL_VOID | CELL_FIXED);
          flagsNat[o + v] = fl;
          if (!(fl & (CELL_SOLID | CELL_VOID))) liq = true;
        }
    blockFluid[bid] = liq;
  }
  for (int d = 0; d < 3; d++) {
#pragma omp parallel for schedule(dynamic, 1) if (par)
    for (int bid = 0; bid < g.nBlocks; bid++) {
      const int lev = m0.level[bid], b = g.bsxL(lev);
      const int64_t o = m0.off[bid];
      int bx, by, bz;
      g.blockCoords(bid, bx, by, bz);
      for (int vz = 0, v = 0; vz < b; vz++)
        for (int vy = 0; vy < b; vy++)
          for (int vx = 0; vx < b; vx++, v++) wNat[d][o + v] = s.faceCoeff(d, bx * b + vx, by * b + vy, bz * b + vz);
    }
  }
  if (hostRhs) {
#pragma omp parallel for schedule(dynamic, 1) if (par)
    for (int bid = 0; bid < g.nBlocks; bid++) {
      if (m0.level[bid] != 0) continue;
      const int b = g.bsx0;
      const int64_t o = m0.off[bid];
      int bx, by, bz;
      g.blockCoords(bid, bx, by, bz);
      bool nz = false;
      for (int vz = 0, v = 0; vz < b; vz++)
        for (int vy = 0; vy < b; vy++)
          for (int vx = 0; vx < b; vx++, v++) {
            const float r = s.rhs(bx * b + vx, by * b + vy, bz * b + vz);
            rhsNat[o + v] = r;
            if (r != 0.0f) nz = true;
          }
      blockRhs[bid] = nz;
    }
  }
  for (int bid = 0; bid < g.nBlocks; bid++) {
    uint32_t &f = m0.bflags[bid];
    f |= BLK_EXISTS;
    if (blockFluid[bid]) f &= ~BLK_NO_FLUID; else f |= BLK_NO_FLUID;
    f &= ~BLK_FIXED;
  }

  /* MG0 pool layout: PCG-active level-0 blocks first. */
  std::vector<int> order;
  order.reserve(g.nBlocks);
  for (int bid = 0; bid < g.nBlocks; bid++)
    if (m0.level[bid] == 0 && (blockFluid[bid] || blockRhs[bid])) order.push_back(bid);
  g.nActive0 = (int)order.size();
  g.active0 = order;
  for (int bid = 0; bid < g.nBlocks; bid++)
    if (!(m0.level[bid] == 0 && (blockFluid[bid] || blockRhs[bid]))) order.push_back(bid);
  std::vector<int64_t> offNat = m0.off;
  int64_t o = 0;
  for (int bid : order) {
    m0.off[bid] = o;
    const int64_t b = g.bsxL(m0.level[bid]);
    o += b * b * b;
  }
  m0.flags.assign(m0.ncells, 0);
  for (int d = 0; d < 3; d++) m0.W[d].assign(m0.ncells, 0.0f);
  rhs.assign(hostRhs ? m0.ncells : 0, 0.0f);
#pragma omp parallel for schedule(static)
  for (int bid = 0; bid < g.nBlocks; bid++) {
    const int64_t b = g.bsxL(m0.level[bid]), n = b * b * b, src = offNat[bid], dst = m0.off[bid];
    std::copy(flagsNat.begin() + src, flagsNat.begin() + src + n, m0.flags.begin() + dst);
    for (int d = 0; d < 3; d++) std::copy(wNat[d].begin() + src, wNat[d].begin() + src + n, m0.W[d].begin() + dst);
    if (hostRhs) std::copy(rhsNat.begin() + src, rhsNat.begin() + src + n, rhs.begin() + dst);
  }

  computeDiagonal(g, 0);
  buildLists(g, 0);
  for (int L = 1; L < g.nLevels; L++) {
    naturalLayout(g, g.mg[L]);
    restrictMeta(g, L);
    computeDiagonal(g, L);
    buildLists(g, L);
  }

  /* smoother weight from the MG0 diagonal range (src/msbg3.cpp:3066-3094) */
  float mx = 0.f, mn = 1e30f;
  for (int64_t i = 0; i < m0.ncells; i++) {
    if (m0.flags[i] & (CELL_SOLID | CELL_VOID | CELL_FIXED)) continue;
    const float v = m0.D[i];
    if (v > 1e-12f) {
      if (v > mx) mx = v;
      if (v < mn) mn = v;
    }
  }
  const float omDef = 6.f / 7.f;
  g.omega = omDef;
  if (mx > 1e-12f && mn > 1e-12f) {
    const float ratio = mx / mn;
    float om = omDef / sqrtf(std::max(ratio, 1.f));
    om = std::max(om, 0.02f);
    g.omega = std::min(om, omDef);
  }
  return MSBGX_OK;
}

/* Caller-side level policy (src/flip_sim.cpp:480-548). */
int levelMapFromScene(const msbgx_scene *s, int bsx0, int nLevels, int mode, int d0, int *out)
{
  if (!s || !out || bsx0 <= 0 || s->nx % bsx0 || s->ny % bsx0 || s->nz % bsx0 || nLevels < 1) return MSBGX_EINVAL;
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
          static const int n6[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
          for (int q = 0; q < 6 && !seed; q++) {
            const int x2 = x + n6[q][0], y2 = y + n6[q][1], z2 = z + n6[q][2];
            if (x2 < 0 || y2 < 0 || z2 < 0 || x2 >= s->nx || y2 >= s->ny || z2 >= s->nz) continue;
            if (!msbgx_scene_is_liquid(s, x2, y2, z2)) seed = true;
          }
        }
    if (seed) dist[bid] = 0;
  }
  const int coarsest = nLevels - 1;
  std::queue<int> q;
  for (int b = 0; b < nb; b++)
    if (dist[b] == 0) q.push(b);
  while (!q.empty()) {
    const int bid = q.front();
    q.pop();
    const int d = dist[bid];
    if (d >= d0 + coarsest) continue;
    const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          if ((dx | dy | dz) == 0) continue;
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
    out[b] = d <= d0 ? 0 : ((nLevels >= 2 && d == d0 + 1) ? 1 : coarsest);
  }
  return MSBGX_OK;
}

}  // namespace msbgx
