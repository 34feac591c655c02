/*
 * flip_oracle.cpp -- CPU restatement of the FLIP caller's steps either side
 * of the pressure solve.
 *
 * TEST INFRASTRUCTURE.  Only tests/ and bench.py's CPU-baseline leg load it
 * (it is linked into libmsbg_oracle.so); the product library never links or
 * calls it.
 *
 * Restated (src/flip_sim.cpp):
 *   getF / getVC / setVC / isFluid   104-136  a cell outside the grid or in a
 *                                             block without data reads 0; a
 *                                             write there is dropped
 *   flipRhsCB                        329-338
 *   pressure-gradient update         1317-1349 (three loops, literally)
 * The grids are the MSBG channels at resolution level 0
 * (src/flip_sim.cpp:363-375): only blocks whose level is 0 hold data.
 * Arrays are dense nx*ny*nz fine grids, x fastest; vector channels are three
 * such arrays (component c = velocity on the cell's -c face).
 * Compiled with -ffp-contract=off: every float operation is the one written.
 */
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Level0Grid {
  int nx, ny, nz, bsx;
  const int *blockLevel;   // per block of the (nx/bsx, ny/bsx, nz/bsx) raster
  int64_t n() const { return (int64_t)nx * ny * nz; }
  /* getBidVidF with the data test: false outside the grid or in a block
   * without level-0 data */
  bool has(int ix, int iy, int iz, int64_t &idx) const
  {
    if (ix < 0 || iy < 0 || iz < 0 || ix >= nx || iy >= ny || iz >= nz) return false;
    const int nbx = nx / bsx, nby = ny / bsx;
    const int bid = ix / bsx + (iy / bsx) * nbx + (iz / bsx) * nbx * nby;
    if (blockLevel[bid] != 0) return false;
    idx = ((int64_t)iz * ny + iy) * nx + ix;
    return true;
  }
  float getF(const float *a, int ix, int iy, int iz) const
  {
    int64_t i;
    return has(ix, iy, iz, i) ? a[i] : 0.f;
  }
  float getVC(const float *const v[3], int ix, int iy, int iz, int c) const
  {
    int64_t i;
    return has(ix, iy, iz, i) ? v[c][i] : 0.f;
  }
  void setVC(float *const v[3], int ix, int iy, int iz, int c, float val) const
  {
    int64_t i;
    if (has(ix, iy, iz, i)) v[c][i] = val;
  }
  bool isFluid(const float *mass, int ix, int iy, int iz, float eps) const { return getF(mass, ix, iy, iz) > eps; }
};

}  // namespace

extern "C" {

/* flipRhsCB at every cell with level-0 data (0 elsewhere): out = isFluid ?
 * div / dt : 0, the divergence summed in the order of src/flip_sim.cpp:333-336. */
void orc_flip_rhs(int nx, int ny, int nz, int bsx, const int *blockLevel, const float *mass, float massEps,
                  const float *u, const float *v, const float *w, float dt, float *out)
{
  const Level0Grid G{nx, ny, nz, bsx, blockLevel};
  const float *vel[3] = {u, v, w};
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const int64_t i = ((int64_t)z * ny + y) * nx + x;
        int64_t j;
        if (!G.has(x, y, z, j) || !G.isFluid(mass, x, y, z, massEps)) {
          out[i] = 0.f;
          continue;
        }
        const float div = G.getVC(vel, x + 1, y, z, 0) - G.getVC(vel, x, y, z, 0) + G.getVC(vel, x, y + 1, z, 1) -
                          G.getVC(vel, x, y, z, 1) + G.getVC(vel, x, y, z + 1, 2) - G.getVC(vel, x, y, z, 2);
        out[i] = div / dt;
      }
}

/* The pressure-gradient update of src/flip_sim.cpp:1317-1349 over the
 * velocity arrays in place; with negate_p the pressure enters as -p. */
void orc_flip_update_velocity(int nx, int ny, int nz, int bsx, const int *blockLevel, const float *mass,
                              float massEps, const float *pIn, const float *bu, const float *bv, const float *bw,
                              float *u, float *v, float *w, float dt, int negate_p)
{
  const Level0Grid G{nx, ny, nz, bsx, blockLevel};
  float *const vel[3] = {u, v, w};
  const float *const sgBeta[3] = {bu, bv, bw};
  auto getP = [&](int ix, int iy, int iz) {
    const float p = G.getF(pIn, ix, iy, iz);
    return negate_p ? -p : p;
  };
  const float MASS_EPS = massEps;
  // U
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix <= nx; ix++) {
        if (ix == 0 || ix == nx) { G.setVC(vel, ix, iy, iz, 0, 0.f); continue; }
        bool fL = G.isFluid(mass, ix - 1, iy, iz, MASS_EPS), fR = G.isFluid(mass, ix, iy, iz, MASS_EPS);
        if (!fL && !fR) { G.setVC(vel, ix, iy, iz, 0, 0.f); continue; }
        float pL = fL ? getP(ix - 1, iy, iz) : 0.f, pR = fR ? getP(ix, iy, iz) : 0.f;
        float beta = G.getVC(sgBeta, ix, iy, iz, 0);
        float nv = G.getVC(vel, ix, iy, iz, 0) - dt * beta * (pR - pL);
        G.setVC(vel, ix, iy, iz, 0, std::isfinite(nv) ? nv : 0.f);
      }
  // V (no explicit zero at iy == ny: the face lies outside the grid)
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy <= ny; iy++)
      for (int ix = 0; ix < nx; ix++) {
        if (iy == 0) { G.setVC(vel, ix, iy, iz, 1, 0.f); continue; }
        bool fB = (iy > 0) && G.isFluid(mass, ix, iy - 1, iz, MASS_EPS);
        bool fT = (iy < ny) && G.isFluid(mass, ix, iy, iz, MASS_EPS);
        if (!fB && !fT) { G.setVC(vel, ix, iy, iz, 1, 0.f); continue; }
        float pB = fB ? getP(ix, iy - 1, iz) : 0.f, pT = fT ? getP(ix, iy, iz) : 0.f;
        float beta = G.getVC(sgBeta, ix, iy, iz, 1);
        float nv = G.getVC(vel, ix, iy, iz, 1) - dt * beta * (pT - pB);
        G.setVC(vel, ix, iy, iz, 1, std::isfinite(nv) ? nv : 0.f);
      }
  // W
  for (int iz = 0; iz <= nz; iz++)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++) {
        if (iz == 0 || iz == nz) { G.setVC(vel, ix, iy, iz, 2, 0.f); continue; }
        bool fN = G.isFluid(mass, ix, iy, iz - 1, MASS_EPS), fF = G.isFluid(mass, ix, iy, iz, MASS_EPS);
        if (!fN && !fF) { G.setVC(vel, ix, iy, iz, 2, 0.f); continue; }
        float pN = fN ? getP(ix, iy, iz - 1) : 0.f, pF = fF ? getP(ix, iy, iz) : 0.f;
        float beta = G.getVC(sgBeta, ix, iy, iz, 2);
        float nv = G.getVC(vel, ix, iy, iz, 2) - dt * beta * (pF - pN);
        G.setVC(vel, ix, iy, iz, 2, std::isfinite(nv) ? nv : 0.f);
      }
}

}  // extern "C"

/* ------------------------------------------------------------------------
 * particleToGrid (src/flip_sim.cpp:579-828), serial.
 *
 * The reference scatters from TBB worker threads into thread-local block
 * buffers and merges them, so the order of its float sums follows the thread
 * partition.  This restatement is its scatter loop run by one thread over the
 * particles in cell-major order (cells of the domain plus a one-cell halo, x
 * fastest; within a cell by particle index), which fixes the order of every
 * sum: each face receives its contributions by source cell, then particle.
 * The grids are resolution-level-0 channels: accumulation covers the whole
 * raster as the reference's block buffers do, and only blocks of level 0
 * receive the results.
 * ---------------------------------------------------------------------- */
namespace {
struct OrcParticle {
  float pos[3], vel[3];
  int32_t phase;
};
}  // namespace

extern "C" {

void orc_p2g(int nx, int ny, int nz, int bsx, const int *blockLevel, int64_t np, const void *particles, float rhoL,
             float rhoG, float massEps, float betaMin, float alphaPhi, int ppv, float *outMass, float *outU,
             float *outV, float *outW, float *outBU, float *outBV, float *outBW, uint8_t *outActive)
{
  const OrcParticle *P = (const OrcParticle *)particles;
  const int64_t n = (int64_t)nx * ny * nz;
  const Level0Grid G{nx, ny, nz, bsx, blockLevel};
  /* cell-major order over the domain plus a one-cell halo */
  const int64_t hx = nx + 2, hy = ny + 2, hz = nz + 2;
  std::vector<int64_t> bin(np), cnt(hx * hy * hz + 1, 0);
  for (int64_t i = 0; i < np; i++) {
    const int ix = (int)floorf(P[i].pos[0]), iy = (int)floorf(P[i].pos[1]), iz = (int)floorf(P[i].pos[2]);
    bin[i] = (ix >= -1 && iy >= -1 && iz >= -1 && ix <= nx && iy <= ny && iz <= nz)
                 ? ((int64_t)(iz + 1) * hy + (iy + 1)) * hx + (ix + 1)
                 : -1;
    if (bin[i] >= 0) cnt[bin[i] + 1]++;
  }
  for (int64_t c = 0; c < hx * hy * hz; c++) cnt[c + 1] += cnt[c];
  std::vector<int64_t> order(cnt[hx * hy * hz]);
  {
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t i = 0; i < np; i++)
      if (bin[i] >= 0) order[fill[bin[i]]++] = i;   // ascending i within a cell
  }
  std::vector<float> vx(n, 0.f), vy(n, 0.f), vz(n, 0.f), mass(n, 0.f), wu(n, 0.f), wv(n, 0.f), ww(n, 0.f),
      ru(n, 0.f), rv(n, 0.f), rw(n, 0.f);
  std::vector<int> liq(n, 0);
  auto bv = [&](int ix, int iy, int iz, int64_t &vid) -> bool {
    if (ix < 0 || iy < 0 || iz < 0 || ix >= nx || iy >= ny || iz >= nz) return false;
    vid = ((int64_t)iz * ny + iy) * nx + ix;
    return true;
  };
  for (int64_t k : order) {
    const OrcParticle &p = P[k];
    const float mp = (p.phase == 0) ? rhoL : rhoG;
    {
      int64_t vid;
      if (bv((int)floorf(p.pos[0]), (int)floorf(p.pos[1]), (int)floorf(p.pos[2]), vid)) {
        mass[vid] += 1.f;
        if (p.phase == 0) liq[vid]++;
      }
    }
    for (int dir = 0; dir < 3; dir++) {
      float *sv = dir == 0 ? vx.data() : (dir == 1 ? vy.data() : vz.data());
      float *sw = dir == 0 ? wu.data() : (dir == 1 ? wv.data() : ww.data());
      float *sr = dir == 0 ? ru.data() : (dir == 1 ? rv.data() : rw.data());
      const float gx = dir == 0 ? p.pos[0] : p.pos[0] - 0.5f;
      const float gy = dir == 1 ? p.pos[1] : p.pos[1] - 0.5f;
      const float gz = dir == 2 ? p.pos[2] : p.pos[2] - 0.5f;
      const int ix0 = (int)floorf(gx), iy0 = (int)floorf(gy), iz0 = (int)floorf(gz);
      const float fx = gx - ix0, fy = gy - iy0, fz = gz - iz0;
      for (int dz = 0; dz <= 1; dz++)
        for (int dy = 0; dy <= 1; dy++)
          for (int dx = 0; dx <= 1; dx++) {
            const float w = (dx ? fx : 1.f - fx) * (dy ? fy : 1.f - fy) * (dz ? fz : 1.f - fz);
            int64_t vid;
            if (bv(ix0 + dx, iy0 + dy, iz0 + dz, vid)) {
              sv[vid] += p.vel[dir] * w;
              sw[vid] += w;
              sr[vid] += mp * w;
            }
          }
    }
  }
  /* write-out, normalisation and beta on the level-0 blocks (src/flip_sim.cpp:727-826) */
  const float rhoTilde0 = (float)ppv;
  const float etaPhi = logf(rhoL / rhoG);
  const float rhoTildeMin = etaPhi * rhoG * rhoTilde0;
  const float phiDenom = alphaPhi * rhoTilde0 * rhoL;
  auto phiFromRho = [&](float rt) -> float {
    if (rt < rhoTildeMin) return 0.f;
    float v = sqrtf(std::max(rt - rhoTildeMin, 0.f) / phiDenom);
    return std::min(v, 1.f);
  };
  auto betaFromPhi = [&](float phi) -> float {
    float rho = rhoG + phi * (rhoL - rhoG);
    return std::max(1.f / rho, betaMin);
  };
  auto cellPh = [&](int cx, int cy, int cz) -> int {
    int64_t v2;
    if (!G.has(cx, cy, cz, v2)) return -1;
    if (mass[v2] <= massEps) return -1;
    return (liq[v2] * 2 > (int)(mass[v2] + 0.5f)) ? 0 : 1;
  };
  const int nbx = nx / bsx, nby = ny / bsx, nbz = nz / bsx;
  float *const outVel[3] = {outU, outV, outW};
  float *const outBeta[3] = {outBU, outBV, outBW};
  for (int64_t i = 0; i < n; i++) {
    outMass[i] = 0.f;
    for (int c = 0; c < 3; c++) outVel[c][i] = outBeta[c][i] = 0.f;
  }
  for (int bid = 0; bid < nbx * nby * nbz; bid++) {
    outActive[bid] = 0;
    if (blockLevel[bid] != 0) continue;
    const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
    for (int lz = 0; lz < bsx; lz++)
      for (int ly = 0; ly < bsx; ly++)
        for (int lx = 0; lx < bsx; lx++) {
          const int ix = bx * bsx + lx, iy = by * bsx + ly, iz = bz * bsx + lz;
          const int64_t i = ((int64_t)iz * ny + iy) * nx + ix;
          float u = 0.f + vx[i], v = 0.f + vy[i], w = 0.f + vz[i];
          if (wu[i] > massEps) u /= wu[i]; else u = 0.f;
          if (wv[i] > massEps) v /= wv[i]; else v = 0.f;
          if (ww[i] > massEps) w /= ww[i]; else w = 0.f;
          outU[i] = u; outV[i] = v; outW[i] = w;
          outMass[i] = 0.f + mass[i];
          if (mass[i] > massEps) outActive[bid] = 1;
          float phiU = phiFromRho(ru[i]), phiV = phiFromRho(rv[i]), phiW = phiFromRho(rw[i]);
          { int pL = cellPh(ix - 1, iy, iz), pR = cellPh(ix, iy, iz);
            if (pL == 0 && pR == 0) phiU = 1.f; else if (pL == 1 && pR == 1) phiU = 0.f; }
          { int pB = cellPh(ix, iy - 1, iz), pT = cellPh(ix, iy, iz);
            if (pB == 0 && pT == 0) phiV = 1.f; else if (pB == 1 && pT == 1) phiV = 0.f; }
          { int pN = cellPh(ix, iy, iz - 1), pF = cellPh(ix, iy, iz);
            if (pN == 0 && pF == 0) phiW = 1.f; else if (pN == 1 && pF == 1) phiW = 0.f; }
          outBU[i] = betaFromPhi(phiU);
          outBV[i] = betaFromPhi(phiV);
          outBW[i] = betaFromPhi(phiW);
        }
  }
}

}  // extern "C"

/* Particles of a scene's liquid cells (msbgx_scene_particle), cell-major,
 * x fastest, phase 0 -- the order in which the GPU scene path generates them.
 * Returns the count; writes when out is non-NULL. */
#include "../include/msbgx_scene.h"
extern "C" int64_t orc_scene_particles(const msbgx_scene *s, void *out)
{
  OrcParticle *P = (OrcParticle *)out;
  int64_t k = 0;
  for (int z = 0; z < s->nz; z++)
    for (int y = 0; y < s->ny; y++)
      for (int x = 0; x < s->nx; x++) {
        if (!msbgx_scene_is_liquid(s, x, y, z)) continue;
        for (int p = 0; p < s->particles_per_cell; p++, k++)
          if (P) {
            msbgx_scene_particle(s, x, y, z, p, P[k].pos, P[k].vel);
            P[k].phase = 0;
          }
      }
  return k;
}

/* ------------------------------------------------------------------------
 * Compact MIC(0)-PCG (src/flip_sim.cpp:285-301, 858-1151), literally: rows
 * = fluid cells with level-0 data in (z, y, x) order, CSR with ascending
 * columns, MIC(0) with sigma = 0.25, PCG from x = 0, sequential FP64 dots.
 * div holds div/dt per cell (the reference's sgDiv); the solution returns
 * per cell in x (0 off the rows), the factor and z = M^-1 rhs in row order.
 * ---------------------------------------------------------------------- */
extern "C" int orc_compact_pcg(int nx, int ny, int nz, int bsx, const int *blockLevel, const float *mass,
                               float massEps, const float *bu, const float *bv, const float *bw, const float *div,
                               float tol, int maxIter, float *xOut, float *preconOut, float *zOut, int *itersOut,
                               float *relOut)
{
  const Level0Grid G{nx, ny, nz, bsx, blockLevel};
  const float *const sgBeta[3] = {bu, bv, bw};
  auto lin = [=](int ix, int iy, int iz) { return (int64_t)ix + (int64_t)nx * (iy + (int64_t)ny * iz); };
  const int64_t nCells = (int64_t)nx * ny * nz;
  /* buildPressureSystem: rows */
  std::vector<int> cellToRow(nCells, -1), rowIx, rowIy, rowIz;
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++) {
        if (!G.isFluid(mass, ix, iy, iz, massEps)) continue;
        cellToRow[lin(ix, iy, iz)] = (int)rowIx.size();
        rowIx.push_back(ix); rowIy.push_back(iy); rowIz.push_back(iz);
      }
  const int n = (int)rowIx.size();
  /* rebuildPressurePattern */
  enum { XM, XP, YM, YP, ZM, ZP };
  std::vector<int> rowPtr(n + 1, 0), colInd, diagPos(n, -1), pos[6];
  for (auto &v : pos) v.assign(n, -1);
  std::vector<float> val;
  for (int row = 0; row < n; row++) {
    const int ix = rowIx[row], iy = rowIy[row], iz = rowIz[row];
    struct Nbr { int col, dir; } nbrs[6];
    int nNbr = 0;
    auto addNbr = [&](int jx, int jy, int jz, int dir) {
      if (jx < 0 || jx >= nx || jy < 0 || jy >= ny || jz < 0 || jz >= nz) return;
      const int col = cellToRow[lin(jx, jy, jz)];
      if (col >= 0) nbrs[nNbr++] = {col, dir};
    };
    addNbr(ix - 1, iy, iz, XM); addNbr(ix + 1, iy, iz, XP);
    addNbr(ix, iy - 1, iz, YM); addNbr(ix, iy + 1, iz, YP);
    addNbr(ix, iy, iz - 1, ZM); addNbr(ix, iy, iz + 1, ZP);
    for (int i = 0; i < nNbr - 1; i++)
      for (int j = i + 1; j < nNbr; j++)
        if (nbrs[j].col < nbrs[i].col) std::swap(nbrs[i], nbrs[j]);
    bool diagIn = false;
    for (int i = 0; i < nNbr; i++) {
      if (!diagIn && nbrs[i].col > row) {
        diagPos[row] = (int)colInd.size();
        colInd.push_back(row); val.push_back(0.f);
        diagIn = true;
      }
      pos[nbrs[i].dir][row] = (int)colInd.size();
      colInd.push_back(nbrs[i].col); val.push_back(0.f);
    }
    if (!diagIn) {
      diagPos[row] = (int)colInd.size();
      colInd.push_back(row); val.push_back(0.f);
    }
    rowPtr[row + 1] = (int)colInd.size();
  }
  /* updatePressureNumerics */
  std::vector<float> rhs(n), betaXm(n), betaYm(n), betaZm(n);
  for (int row = 0; row < n; row++) {
    const int ix = rowIx[row], iy = rowIy[row], iz = rowIz[row];
    const float bxm = (ix > 0) ? G.getVC(sgBeta, ix, iy, iz, 0) : 0.f;
    const float bxp = (ix + 1 < nx) ? G.getVC(sgBeta, ix + 1, iy, iz, 0) : 0.f;
    const float bym = (iy > 0) ? G.getVC(sgBeta, ix, iy, iz, 1) : 0.f;
    const float byp = (iy + 1 < ny) ? G.getVC(sgBeta, ix, iy + 1, iz, 1) : 0.f;
    const float bzm = (iz > 0) ? G.getVC(sgBeta, ix, iy, iz, 2) : 0.f;
    const float bzp = (iz + 1 < nz) ? G.getVC(sgBeta, ix, iy, iz + 1, 2) : 0.f;
    rhs[row] = -G.getF(div, ix, iy, iz);
    betaXm[row] = bxm; betaYm[row] = bym; betaZm[row] = bzm;
    val[diagPos[row]] = bxm + bxp + bym + byp + bzm + bzp;
    if (pos[XM][row] >= 0) val[pos[XM][row]] = -bxm;
    if (pos[XP][row] >= 0) val[pos[XP][row]] = -bxp;
    if (pos[YM][row] >= 0) val[pos[YM][row]] = -bym;
    if (pos[YP][row] >= 0) val[pos[YP][row]] = -byp;
    if (pos[ZM][row] >= 0) val[pos[ZM][row]] = -bzm;
    if (pos[ZP][row] >= 0) val[pos[ZP][row]] = -bzp;
  }
  auto spmv = [&](const float *xv, float *q) {
    for (int i = 0; i < n; i++) {
      float sum = 0.f;
      for (int j = rowPtr[i]; j < rowPtr[i + 1]; j++) sum += val[j] * xv[colInd[j]];
      q[i] = sum;
    }
  };
  auto dot = [&](const float *a, const float *b) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += (double)a[i] * (double)b[i];
    return s;
  };
  /* buildMIC0Compact */
  std::vector<float> precon(n, 0.f);
  const float sigma = 0.25f;
  for (int row = 0; row < n; row++) {
    const float Adiag = val[diagPos[row]];
    float e = Adiag;
    const int ix = rowIx[row], iy = rowIy[row], iz = rowIz[row];
    if (ix > 0) { const int r2 = cellToRow[lin(ix - 1, iy, iz)];
                  if (r2 >= 0) { const float bx = betaXm[row], p = precon[r2]; e -= (bx * p) * (bx * p); } }
    if (iy > 0) { const int r2 = cellToRow[lin(ix, iy - 1, iz)];
                  if (r2 >= 0) { const float by = betaYm[row], p = precon[r2]; e -= (by * p) * (by * p); } }
    if (iz > 0) { const int r2 = cellToRow[lin(ix, iy, iz - 1)];
                  if (r2 >= 0) { const float bz = betaZm[row], p = precon[r2]; e -= (bz * p) * (bz * p); } }
    precon[row] = 1.f / sqrtf(std::max(e, sigma * std::max(Adiag, 1e-10f)));
  }
  /* applyMIC0Compact */
  auto apply = [&](const float *r, float *z) {
    std::vector<float> q(n);
    for (int row = 0; row < n; row++) {
      const int ix = rowIx[row], iy = rowIy[row], iz = rowIz[row];
      float t = r[row];
      if (ix > 0) { const int r2 = cellToRow[lin(ix - 1, iy, iz)]; if (r2 >= 0) t += betaXm[row] * precon[r2] * q[r2]; }
      if (iy > 0) { const int r2 = cellToRow[lin(ix, iy - 1, iz)]; if (r2 >= 0) t += betaYm[row] * precon[r2] * q[r2]; }
      if (iz > 0) { const int r2 = cellToRow[lin(ix, iy, iz - 1)]; if (r2 >= 0) t += betaZm[row] * precon[r2] * q[r2]; }
      q[row] = t * precon[row];
    }
    for (int row = n - 1; row >= 0; row--) {
      const int ix = rowIx[row], iy = rowIy[row], iz = rowIz[row];
      const float pi = precon[row];
      float t = q[row];
      if (ix + 1 < nx) { const int r2 = cellToRow[lin(ix + 1, iy, iz)]; if (r2 >= 0) t += betaXm[r2] * pi * z[r2]; }
      if (iy + 1 < ny) { const int r2 = cellToRow[lin(ix, iy + 1, iz)]; if (r2 >= 0) t += betaYm[r2] * pi * z[r2]; }
      if (iz + 1 < nz) { const int r2 = cellToRow[lin(ix, iy, iz + 1)]; if (r2 >= 0) t += betaZm[r2] * pi * z[r2]; }
      z[row] = t * pi;
    }
  };
  std::vector<float> x(n, 0.f), r(n), z(n), d(n), q(n);
  for (int i = 0; i < n; i++) r[i] = rhs[i];
  if (preconOut) std::copy(precon.begin(), precon.end(), preconOut);
  if (zOut) {
    std::vector<float> z0(n);
    apply(rhs.data(), z0.data());
    std::copy(z0.begin(), z0.end(), zOut);
  }
  /* solvePressureCompactPCG */
  double rhsNormSq = 0.0;
  for (int i = 0; i < n; i++) rhsNormSq += (double)r[i] * (double)r[i];
  int iters = 0, converged = -1;
  if (n > 0 && rhsNormSq > 1e-30) {
    apply(r.data(), z.data());
    double rho = dot(r.data(), z.data());
    for (int i = 0; i < n; i++) d[i] = z[i];
    const double tolSq = (double)tol * (double)tol;
    for (int iter = 0; iter < maxIter; iter++) {
      spmv(d.data(), q.data());
      const double dq = dot(d.data(), q.data());
      if (dq <= 1e-30 || rho <= 1e-30) break;
      const float alpha = (float)(rho / dq);
      for (int i = 0; i < n; i++) { x[i] += alpha * d[i]; r[i] -= alpha * q[i]; }
      iters = iter + 1;
      const double rNormSq = dot(r.data(), r.data());
      if (rNormSq <= tolSq * rhsNormSq) { converged = iter + 1; break; }
      apply(r.data(), z.data());
      const double rhoNew = dot(r.data(), z.data());
      if (rho <= 1e-30) break;
      const float beta = (float)(rhoNew / rho);
      for (int i = 0; i < n; i++) d[i] = z[i] + beta * d[i];
      rho = rhoNew;
    }
  }
  if (itersOut) *itersOut = converged > 0 ? converged : iters;
  if (relOut) *relOut = rhsNormSq > 0 ? (float)std::sqrt(dot(r.data(), r.data()) / rhsNormSq) : 0.f;
  if (xOut) {
    std::fill(xOut, xOut + nCells, 0.f);
    for (int row = 0; row < n; row++) xOut[lin(rowIx[row], rowIy[row], rowIz[row])] = x[row];
  }
  return n;
}

/* ------------------------------------------------------------------------
 * applyChannelPdeFast (src/msbg4.cpp:60-960) over a density channel of
 * uint16 render density (dtype 0, read as v/65535, stored rounded to
 * nearest even) or float (dtype 1), with the two in-place block schemes:
 *   laplTyp < 0 carries the scheme (src/msbg4.cpp:100-112): with
 *   OPT_8_COLOR_SCHEME (1<<29) set, 8 colour passes (colour (bx&1) +
 *   2(by&1) + 4(bz&1), src/msbg.h:642-646), pass c updates colour c; else
 *   red-black, pass p updates the blocks with (bx+by+bz)&1 != p
 *   (src/msbg4.cpp:427-434), i.e. colour 1 first;
 *   types (src/msbg4.cpp:555-800): 4 mean curvature, 2 bi-Laplacian (halo
 *   2), 1 Laplacian times |grad|, 3 simple diffusion, 14 Perona-Malik, 15
 *   its Jacobi form, 17 Jacobi diffusion; Taubin alternation for 4 and 1
 *   (src/msbg4.cpp:621-628, 710-717); hard constraints [phi0 - tol1,
 *   phi0 + tol2] within [0, 1] against the input copy, else the optional
 *   [0, 1] clamp (src/msbg4.cpp:806-820); numIter rounded up to even;
 *   halo blocks with OPT_BC_NEUMANN (src/halo.cpp:267-480,
 *   src/halo.h:267-290).
 * Returns -1 for what the reference does not define or this restatement
 * leaves out: the unordered (laplTyp >= 0) scheme, types with the fixed
 * 18-voxel row stride (1, 3, 17) at block edges other than 16
 * (src/msbg4.cpp:688-693), red-black with stencils that reach diagonal
 * neighbour blocks (2, 4: a same-colour block is then read while it is
 * written), uint16 without a clamp, anisDiffAlpha 0, negative numIter or
 * time step (src/msbg4.cpp:144-148, 178-185).
 * ---------------------------------------------------------------------- */
namespace {

/* FMA_FAST_EXP (src/fastmath.h:583-626): Cephes-style expf with the
 * reference's own coefficient sequence (its PX2 term is added twice). */
float pdeFastExp(float x0)
{
  const float MAXLOGF = 88.72283905206835f, MINLOGF = -88.f;
  const float C1F = 0.693359375f, C2F = -2.12194440e-4f;
  const float P1 = 1.9875691500E-4f, P2 = 1.3981999507E-3f, P3 = 8.3334519073E-3f, P4 = 4.1665795894E-2f,
              P5 = 1.6666665459E-1f, P6 = 5.0000001201E-1f;
  const float LOG2EF = 1.44269504088896341f;
  float x = x0;
  float z = std::floor(LOG2EF * x + 0.5f);
  x -= z * C1F;
  x -= z * C2F;
  const int n = (int)z;
  const float x2 = x * x;
  z = x * P1;
  z += P2;
  z *= x;
  z += P2;
  z *= x;
  z += P3;
  z *= x;
  z += P4;
  z *= x;
  z += P5;
  z *= x;
  z += P6;
  z *= x2;
  z += x + 1.0f;
  uint32_t bits = (uint32_t)(n + 0x7f) << 23;
  float scale;
  std::memcpy(&scale, &bits, 4);
  z *= scale;
  if (x0 < MINLOGF) z = 0.0f;
  if (x0 > MAXLOGF) z = INFINITY;
  return z;
}

float pdeConductance(float f, float alpha)
{
  if (alpha < 0) {
    const float t = std::fabs(f) * (1.0f / -alpha);
    return pdeFastExp(-(t * t));
  }
  const float t = std::fabs(f) / alpha;
  return 1.0f / (1.0f + t * t);
}

}  // namespace

extern "C" int orc_pde_apply(int nx, int ny, int nz, int bsx, const int *blockLevel, int dtype, void *data,
                             const int *active, int nActive, int laplTyp, int doConstrZeroOne, float densThr,
                             float densThrTol1, float densThrTol2, int numIter, float timeStep, float taubin,
                             float anisDiffAlpha)
{
  const int OPT_8_COLOR = 1 << 29;
  bool color8 = false, redBlack = false;
  if (laplTyp < 0) {
    laplTyp = -laplTyp;
    if (laplTyp & OPT_8_COLOR) {
      laplTyp &= ~OPT_8_COLOR;
      color8 = true;
    } else {
      redBlack = true;
    }
  }
  const int typ = laplTyp;
  if (!color8 && !redBlack) return -1;
  if (typ != 1 && typ != 2 && typ != 3 && typ != 4 && typ != 14 && typ != 15 && typ != 17) return -1;
  if ((typ == 1 || typ == 3 || typ == 17) && bsx != 16) return -1;
  if (redBlack && (typ == 2 || typ == 4)) return -1;
  if ((typ == 14 || typ == 15) && anisDiffAlpha == 0.0f) return -1;
  if (numIter < 0 || timeStep < 0) return -1;
  const bool doHard = densThr > 1e-12f || densThrTol1 > 1e-12f;
  if (dtype == 0 && !doHard && !doConstrZeroOne) return -1;
  const bool doTaubin = std::fabs(taubin) > 1e-20f;
  const int hs = typ == 2 ? 2 : 1;
  const int nbx = nx / bsx, nby = ny / bsx;
  const int hsx = bsx + 2 * hs, hsx2 = hsx * hsx;
  uint16_t *rho16 = static_cast<uint16_t *>(data);
  float *rhoF = static_cast<float *>(data);
  auto idx = [&](int x, int y, int z) { return ((int64_t)z * ny + y) * nx + x; };
  auto hasData = [&](int x, int y, int z) {
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return false;
    return blockLevel[x / bsx + nbx * (y / bsx + nby * (z / bsx))] == 0;
  };
  auto get = [&](int64_t i) { return dtype == 0 ? rho16[i] * (1.f / 65535) : rhoF[i]; };
  std::vector<float> phi0;
  if (doHard) {
    phi0.resize((size_t)nx * ny * nz);
    for (int64_t i = 0; i < (int64_t)phi0.size(); i++) phi0[i] = get(i);
  }
  int nMaxIter = numIter;
  if ((nMaxIter % 2) == 1) nMaxIter++;
  const float dt = timeStep;
  std::vector<float> halo((size_t)hsx * hsx2);
  for (int iIter = 0; iIter < nMaxIter; iIter++)
    for (int pass = 0; pass < (color8 ? 8 : 2); pass++)
      for (int t = 0; t < nActive; t++) {
        const int bid = active[t];
        const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
        if (redBlack && ((bx + by + bz) & 1) == pass) continue;
        if (color8 && (bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2) != pass) continue;
        if (blockLevel[bid] != 0) continue;
        const int x0 = bx * bsx, y0 = by * bsx, z0 = bz * bsx;
        /* fillHaloBlock_ with OPT_BC_NEUMANN */
        for (int z = z0 - hs; z < z0 + bsx + hs; z++)
          for (int y = y0 - hs; y < y0 + bsx + hs; y++)
            for (int x = x0 - hs; x < x0 + bsx + hs; x++) {
              int64_t src;
              if (hasData(x, y, z)) src = idx(x, y, z);
              else src = idx(std::min(std::max(x, x0), x0 + bsx - 1), std::min(std::max(y, y0), y0 + bsx - 1),
                             std::min(std::max(z, z0), z0 + bsx - 1));
              halo[(size_t)(x - x0 + hs) + hsx * (y - y0 + hs) + hsx2 * (z - z0 + hs)] = get(src);
            }
        for (int vz = 0; vz < bsx; vz++)
          for (int vy = 0; vy < bsx; vy++)
            for (int vx = 0; vx < bsx; vx++) {
              const float *ph = halo.data() + (vx + hs) + hsx * (vy + hs) + hsx2 * (vz + hs);
              auto F_ = [&](int dx, int dy, int dz) { return ph[dx + dy * hsx + dz * hsx2]; };
              float phi = F_(0, 0, 0);
              if (typ == 4) {
                const float f0 = F_(0, 0, 0), f1 = F_(-1, 0, 0), f2 = F_(1, 0, 0), f3 = F_(0, -1, 0),
                            f4 = F_(0, 1, 0), f5 = F_(0, 0, -1), f6 = F_(0, 0, 1);
                const float fx = 0.5f * (f2 - f1), fy = 0.5f * (f4 - f3), fz = 0.5f * (f6 - f5);
                const float fxx = f1 + f2 - 2 * f0, fyy = f3 + f4 - 2 * f0, fzz = f5 + f6 - 2 * f0;
                const float fy_right = (F_(1, 1, 0) - F_(1, -1, 0)) * 0.5f,
                            fy_left = (F_(-1, 1, 0) - F_(-1, -1, 0)) * 0.5f,
                            fz_right = (F_(1, 0, 1) - F_(1, 0, -1)) * 0.5f,
                            fz_left = (F_(-1, 0, 1) - F_(-1, 0, -1)) * 0.5f,
                            fz_up = (F_(0, 1, 1) - F_(0, 1, -1)) * 0.5f,
                            fz_down = (F_(0, -1, 1) - F_(0, -1, -1)) * 0.5f, fyx = (fy_right - fy_left) * 0.5f,
                            fzx = (fz_right - fz_left) * 0.5f, fzy = (fz_up - fz_down) * 0.5f;
                float H = (fy * fy + fz * fz) * fxx + (fx * fx + fz * fz) * fyy + (fx * fx + fy * fy) * fzz -
                          2.0f * (fx * fy * fyx + fx * fz * fzx + fy * fz * fzy);
                const float gradMagSq = fx * fx + fy * fy + fz * fz;
                H = gradMagSq > 1e-7f ? H / gradMagSq : 0.0f;
                float D = dt * H;
                D = std::min(std::max(D, -0.1f), 0.1f);
                if (doTaubin) phi = (iIter & 1) ? phi + D : phi - taubin * D;
                else phi += D;
              } else if (typ == 2) {
                const float F0 = F_(0, 0, 0);
                const float H = 42.f * F_(0, 0, 0) -
                                12.f * (F_(1, 0, 0) + F_(0, 1, 0) + F_(0, 0, 1) + F_(-1, 0, 0) + F_(0, -1, 0) +
                                        F_(0, 0, -1)) +
                                1.f * (F_(2, 0, 0) + F_(0, 2, 0) + F_(0, 0, 2) + F_(-2, 0, 0) + F_(0, -2, 0) +
                                       F_(0, 0, -2)) +
                                2.f * (F_(1, 0, 1) + F_(0, 1, 1) + F_(-1, 0, 1) + F_(0, -1, 1) + F_(1, 0, -1) +
                                       F_(0, 1, -1) + F_(-1, 0, -1) + F_(0, -1, -1) + F_(1, 1, 0) + F_(-1, 1, 0) +
                                       F_(-1, -1, 0) + F_(1, -1, 0));
                const float GX = 0.5f * (F_(1, 0, 0) - F_(-1, 0, 0)), GY = 0.5f * (F_(0, 1, 0) - F_(0, -1, 0)),
                            GZ = 0.5f * (F_(0, 0, 1) - F_(0, 0, -1));
                const float G = std::sqrt(GX * GX + GY * GY + GZ * GZ);
                float D = -dt * H * G;
                D = std::min(std::max(D, -0.1f), 0.1f);
                phi = F0 + D;
              } else if (typ == 1 || typ == 3 || typ == 17) {
                /* fixed strides of an 18-voxel halo row (bsx 16) */
                const float F0 = phi, F1 = ph[-1], F2 = ph[1], F3 = ph[-18], F4 = ph[18], F5 = ph[-18 * 18],
                            F6 = ph[18 * 18];
                if (typ == 1) {
                  const float L = F1 + F2 + F3 + F4 + F5 + F6 - 6.0f * F0;
                  const float GX = 0.5f * (F2 - F1), GY = 0.5f * (F4 - F3), GZ = 0.5f * (F6 - F5);
                  const float G = std::sqrt(GX * GX + GY * GY + GZ * GZ);
                  const float D = dt * L * G;
                  if (doTaubin) phi = (iIter & 1) ? phi + D : phi - taubin * D;
                  else phi += D;
                } else if (typ == 3) {
                  const float S = (1.0f / 6.0f) * (F1 + F2 + F3 + F4 + F5 + F6);
                  const float D = dt * (S - phi);
                  phi += D;
                } else {
                  phi = (phi + dt * (F1 + F2 + F3 + F4 + F5 + F6)) * 1.0f / (1.0f + dt * 6.0f);
                }
              } else {   // 14, 15
                const float f0 = F_(0, 0, 0), f1 = F_(-1, 0, 0), f2 = F_(1, 0, 0), f3 = F_(0, -1, 0),
                            f4 = F_(0, 1, 0), f5 = F_(0, 0, -1), f6 = F_(0, 0, 1);
                const float g1 = f1 - f0, g2 = f2 - f0, g3 = f3 - f0, g4 = f4 - f0, g5 = f5 - f0, g6 = f6 - f0;
                const float G1 = pdeConductance(g1, anisDiffAlpha), G2 = pdeConductance(g2, anisDiffAlpha),
                            G3 = pdeConductance(g3, anisDiffAlpha), G4 = pdeConductance(g4, anisDiffAlpha),
                            G5 = pdeConductance(g5, anisDiffAlpha), G6 = pdeConductance(g6, anisDiffAlpha);
                if (typ == 15) {
                  const float sum = G1 * f1 + G2 * f2 + G3 * f3 + G4 * f4 + G5 * f5 + G6 * f6,
                              wsum = G1 + G2 + G3 + G4 + G5 + G6;
                  phi = (phi + dt * sum) * 1.0f / (1.0f + dt * wsum);
                } else {
                  const float sum = G1 * g1 + G2 * g2 + G3 * g3 + G4 * g4 + G5 * g5 + G6 * g6;
                  phi += dt * sum;
                }
              }
              const int64_t o = idx(x0 + vx, y0 + vy, z0 + vz);
              if (doHard) {
                const float phiMin = std::max(0.0f, phi0[o] - densThrTol1),
                            phiMax = std::min(1.0f, phi0[o] + densThrTol2);
                phi = std::max(std::min(phi, phiMax), phiMin);
              } else if (doConstrZeroOne) {
                phi = std::min(std::max(phi, 0.0f), 1.0f);
              }
              if (dtype == 0) rho16[o] = (uint16_t)(int)std::nearbyint(phi * 65535);
              else rhoF[o] = phi;
            }
      }
  return 0;
}

/* The msbg_demo call (src/msbg_demo.cpp:716-727): mean curvature, 8
 * colours, [0, 1] clamp, uint16 render density. */
extern "C" void orc_pde_mean_curvature(int nx, int ny, int nz, int bsx, const int *blockLevel, uint16_t *rho,
                                       const int *active, int nActive, int numIter, float dt)
{
  orc_pde_apply(nx, ny, nz, bsx, blockLevel, 0, rho, active, nActive, -(4 + (1 << 29)), 1, 0.f, 0.f, 0.f, numIter,
                dt, 0.f, 0.f);
}
