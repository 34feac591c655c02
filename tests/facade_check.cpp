/*
 * Drives libmsbgx through include/msbgx_facade.hpp the way the reference's
 * FLIP caller drives MSBG::MultiresSparseGrid (src/flip_sim.cpp:1289-1304):
 * create, setDomainBoundarySpec_, regularizeRefinementMap, setRefinementMap,
 * preparePressureSolveFLIP with FlipPressureCallbacks, solvePressureFLIP,
 * then the pressure of every level-0 block.
 *
 * The callback and parameter structs below have the reference's layout and
 * member names (src/msbg.h:598, 829-844), so the facade compiles against the
 * same types a maintainer's build would pass it.
 *
 * Usage: facade_check <input> <output>
 *   input  (little-endian): int32 sx, sy, sz, bsx0, nLevels, nPre, nPost,
 *          nCoarse, maxIter; float32 tol; int32 level map [nBlocks];
 *          uint16 cell types [N]; float32 face_u, face_v, face_w, rhs [N]
 *          (N = sx*sy*sz, x fastest; the callbacks sample these arrays at the
 *          level-local coordinates they are given)
 *   output: int32 iterations, float32 |r|/|b|, int32 callback calls with a
 *           wrong user pointer, int32 checks of the error behaviour that
 *           failed, then float32 pressure [nBlocks][bsx0^3] (0 for blocks
 *           that are not level 0).
 * Exit status 0 when every step ran.
 */
#include <cstdint>
#include <cstdio>
#include <vector>

#include "msbgx_facade.hpp"

namespace MSBG {
typedef uint16_t CellFlags;
struct FlipPressureCallbacks {
  CellFlags (*sampleCellType)(void *user, int x, int y, int z);
  float (*sampleFaceCoeff)(void *user, int dir, int x, int y, int z);
  float (*sampleRhs)(void *user, int x, int y, int z);
  void *user;
};
struct FlipPressureSolveParams {
  int nPre = 3;
  int nPost = 3;
  int nCoarse = 24;
  int maxIter = 500;
  float tol = 5e-4f;
};
}  // namespace MSBG

namespace {

struct Fields {
  int sx, sy, sz;
  std::vector<uint16_t> flags;
  std::vector<float> w[3], rhs;
  int badUser = 0;
  size_t idx(int x, int y, int z) const { return (size_t)x + (size_t)sx * ((size_t)y + (size_t)sy * (size_t)z); }
};
Fields *g_fields = nullptr;

MSBG::CellFlags cellType(void *user, int x, int y, int z)
{
  Fields *f = static_cast<Fields *>(user);
  if (f != g_fields) { g_fields->badUser++; return 0; }
  return f->flags[f->idx(x, y, z)];
}
float faceCoeff(void *user, int dir, int x, int y, int z)
{
  Fields *f = static_cast<Fields *>(user);
  if (f != g_fields || dir < 0 || dir > 2) { g_fields->badUser++; return 0.f; }
  return f->w[dir][f->idx(x, y, z)];
}
float rhs(void *user, int x, int y, int z)
{
  Fields *f = static_cast<Fields *>(user);
  if (f != g_fields) { g_fields->badUser++; return 0.f; }
  return f->rhs[f->idx(x, y, z)];
}

template <typename T> bool readN(FILE *fp, T *dst, size_t n) { return fread(dst, sizeof(T), n, fp) == n; }

}  // namespace

int main(int argc, char **argv)
{
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
    return 2;
  }
  FILE *fp = std::fopen(argv[1], "rb");
  if (!fp) { std::perror("input"); return 2; }
  int hdr[9];
  float tol = 0.f;
  if (!readN(fp, hdr, 9) || !readN(fp, &tol, 1)) { std::fprintf(stderr, "short header\n"); return 2; }
  Fields f;
  f.sx = hdr[0]; f.sy = hdr[1]; f.sz = hdr[2];
  const int bsx = hdr[3], nLevels = hdr[4];
  const size_t n = (size_t)f.sx * f.sy * f.sz;
  const int nb = (f.sx / bsx) * (f.sy / bsx) * (f.sz / bsx);
  std::vector<int> map(nb);
  f.flags.resize(n);
  for (auto &w : f.w) w.resize(n);
  f.rhs.resize(n);
  bool ok = readN(fp, map.data(), nb) && readN(fp, f.flags.data(), n);
  for (auto &w : f.w) ok = ok && readN(fp, w.data(), n);
  ok = ok && readN(fp, f.rhs.data(), n);
  std::fclose(fp);
  if (!ok) { std::fprintf(stderr, "short input\n"); return 2; }
  g_fields = &f;

  int failedChecks = 0;
  /* create() returns nullptr on invalid arguments, as the reference's does */
  if (msbgx::MultiresSparseGrid::create(f.sx, f.sy, f.sz, 12, nLevels) != nullptr) failedChecks++;

  msbgx::MultiresSparseGrid *grid = msbgx::MultiresSparseGrid::create(f.sx, f.sy, f.sz, bsx, nLevels);
  if (!grid) { std::fprintf(stderr, "create: %s\n", msbgx_last_error()); return 1; }
  int iters = 0;
  float rel = 0.f;
  std::vector<float> pressure((size_t)nb * bsx * bsx * bsx, 0.f);
  try {
    grid->setDomainBoundarySpec_(MSBGX_DBC_SOLID, MSBGX_DBC_SOLID, MSBGX_DBC_SOLID, MSBGX_DBC_SOLID,
                                 MSBGX_DBC_SOLID, MSBGX_DBC_SOLID);
    grid->regularizeRefinementMap(map.data());
    grid->setRefinementMap(map.data());
    MSBG::FlipPressureSolveParams sp;
    sp.nPre = hdr[5]; sp.nPost = hdr[6]; sp.nCoarse = hdr[7]; sp.maxIter = hdr[8]; sp.tol = tol;
    /* a solve before the prepare is a state error */
    try {
      grid->solvePressureFLIP(sp);
      failedChecks++;
    } catch (const msbgx::Error &e) {
      if (e.code != MSBGX_ESTATE) failedChecks++;
    }
    MSBG::FlipPressureCallbacks cb;
    cb.sampleCellType = cellType;
    cb.sampleFaceCoeff = faceCoeff;
    cb.sampleRhs = rhs;
    cb.user = &f;
    grid->preparePressureSolveFLIP(cb);
    iters = grid->solvePressureFLIP(sp, &rel);
    for (int b = 0; b < nb; b++)
      if (grid->getBlockLevel(b) == 0) grid->getPressureBlock(b, pressure.data() + (size_t)b * bsx * bsx * bsx);
  } catch (const msbgx::Error &e) {
    std::fprintf(stderr, "facade error %d: %s\n", e.code, e.what());
    msbgx::MultiresSparseGrid::destroy(grid);
    return 1;
  }
  msbgx::MultiresSparseGrid::destroy(grid);
  if (grid != nullptr) failedChecks++;

  FILE *out = std::fopen(argv[2], "wb");
  if (!out) { std::perror("output"); return 2; }
  int32_t head[2] = {iters, 0};
  std::fwrite(&head[0], sizeof(int32_t), 1, out);
  std::fwrite(&rel, sizeof(float), 1, out);
  int32_t tail[2] = {f.badUser, failedChecks};
  std::fwrite(tail, sizeof(int32_t), 2, out);
  std::fwrite(pressure.data(), sizeof(float), pressure.size(), out);
  std::fclose(out);
  std::printf("facade_check: %d iterations, |r|/|b| = %.3e, bad user pointers %d, failed checks %d\n", iters,
              (double)rel, f.badUser, failedChecks);
  return 0;
}
