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
#include <cstdint>

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
