/*
 * msbgx_facade.hpp -- C++ facade over the msbgx C ABI with the member
 * functions the reference's FLIP caller uses on MSBG::MultiresSparseGrid
 * (src/msbg.h:829-940): create / destroy, setDomainBoundarySpec_,
 * regularizeRefinementMap, setRefinementMap, preparePressureSolveFLIP and
 * solvePressureFLIP, plus per-block pressure readback (what
 * src/flip_sim.cpp:1317-1349 reads through grid_.pressure).
 *
 * Header-only; link with -lmsbgx.  The callback and parameter structs are
 * template arguments so the header compiles both next to the reference
 * (MSBG::FlipPressureCallbacks, MSBG::FlipPressureSolveParams,
 * src/msbg.h:829-844) and without it (any struct with the same members).
 *
 * Error behaviour: the reference aborts through UT_ASSERT0 and create()
 * returns NULL on failure (src/msbg.cpp:1100-1106).  Here create() returns
 * nullptr (msbgx_last_error() says why) and every other call throws
 * msbgx::Error carrying the C ABI's status code and message.
 */
#ifndef MSBGX_FACADE_HPP
#define MSBGX_FACADE_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "msbgx.h"

namespace msbgx {

class Error : public std::runtime_error {
 public:
  Error(int rc, const std::string &where)
      : std::runtime_error(where + ": " + msbgx_last_error()), code(rc) {}
  int code;
};

class MultiresSparseGrid {
 public:
  /* create (src/msbg.h:859-875) with blockSize0 and the level count given;
   * nullptr on error. */
  static MultiresSparseGrid *create(int sx0, int sy0, int sz0, int blockSize0, int nLevels)
  {
    msbgx_grid *h = nullptr;
    if (msbgx_create(&h, sx0, sy0, sz0, blockSize0, nLevels) != MSBGX_OK) return nullptr;
    return new MultiresSparseGrid(h, blockSize0);
  }
  /* destroy (src/msbg.h:877): frees the grid and nulls the pointer. */
  static void destroy(MultiresSparseGrid *&g)
  {
    delete g;
    g = nullptr;
  }
  ~MultiresSparseGrid() { msbgx_destroy(h_); }
  MultiresSparseGrid(const MultiresSparseGrid &) = delete;
  MultiresSparseGrid &operator=(const MultiresSparseGrid &) = delete;

  /* setDomainBoundarySpec_ (src/msbg.h:935-940); kinds MSBGX_DBC_OPEN / _SOLID. */
  int setDomainBoundarySpec_(int l, int r, int b, int t, int f, int k)
  {
    check(msbgx_set_domain_bc(h_, l, r, b, t, f, k), "setDomainBoundarySpec_");
    return 0;
  }
  /* regularizeRefinementMap (src/msbg.h:922-926): 2:1 grading of `in`
   * into `out` (or in place when out is null). */
  void regularizeRefinementMap(int *in, int *out = nullptr)
  {
    check(msbgx_regularize_refinement_map(h_, in, out ? out : in), "regularizeRefinementMap");
  }
  /* setRefinementMap (src/msbg.h:928-934) with doResetChannels = true. */
  void setRefinementMap(const int *levels) { check(msbgx_set_refinement_map(h_, levels), "setRefinementMap"); }

  /* preparePressureSolveFLIP (src/msbg.h:849-852): callbacks are invoked
   * serially on this thread with level-local coordinates. */
  template <typename Callbacks> void preparePressureSolveFLIP(const Callbacks &cb, unsigned /*options*/ = 0)
  {
    msbgx_callbacks c;
    c.sample_cell_type = cb.sampleCellType;   // CellFlags is uint16_t (src/msbg.h:598)
    c.sample_face_coeff = cb.sampleFaceCoeff;
    c.sample_rhs = cb.sampleRhs;
    c.user = cb.user;
    check(msbgx_prepare_pressure_solve(h_, &c), "preparePressureSolveFLIP");
  }
  /* solvePressureFLIP (src/msbg.h:854-857): returns the iteration count
   * (maxIter when not converged); |r|/|b| into *outRelResidual. */
  template <typename Params> int solvePressureFLIP(const Params &sp, float *outRelResidual = nullptr)
  {
    msbgx_solve_params p;
    p.n_pre = sp.nPre;
    p.n_post = sp.nPost;
    p.n_coarse = sp.nCoarse;
    p.max_iter = sp.maxIter;
    p.tol = sp.tol;
    int it = 0;
    float rr = 0.f;
    check(msbgx_solve_pressure(h_, &p, &it, &rr), "solvePressureFLIP");
    if (outRelResidual) *outRelResidual = rr;
    return it;
  }

  int nBlocks() const { return msbgx_num_blocks(h_); }
  int blockSize0() const { return bsx0_; }
  /* Level of block bid at MG level 0 (getBlockLevel, src/msbg.h:3560-3590). */
  int getBlockLevel(int bid) const
  {
    const int n = nBlocks();
    if (bid < 0 || bid >= n) throw Error(MSBGX_EINVAL, "getBlockLevel: block id out of range");
    levels_.resize(n);
    flags_.resize(n);
    check(msbgx_block_info(h_, 0, levels_.data(), flags_.data()), "getBlockLevel");
    return levels_[bid];
  }
  /* Pressure of level-0 block bid: blockSize0^3 floats, x fastest. */
  void getPressureBlock(int bid, float *dst) const
  {
    const int rc = msbgx_get_channel_block(h_, MSBGX_CH_PRESSURE, 0, bid, dst);
    if (rc < 0) check(rc, "getPressureBlock");
  }
  msbgx_grid *handle() const { return h_; }

 private:
  MultiresSparseGrid(msbgx_grid *h, int bsx) : h_(h), bsx0_(bsx) {}
  static void check(int rc, const char *where)
  {
    if (rc != MSBGX_OK) throw Error(rc, where);
  }
  msbgx_grid *h_;
  int bsx0_;
  mutable std::vector<int> levels_;
  mutable std::vector<uint32_t> flags_;
};

}  // namespace msbgx

#endif
