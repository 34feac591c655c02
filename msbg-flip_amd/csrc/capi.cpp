/*
 * extern "C" entry points of libmsbgx (declared in include/msbgx.h).
 * Argument validation, error strings and the three prepare input paths
 * (callbacks, dense arrays, analytic scene) live here; the device work is in
 * solver.hip.
 */
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>

#include "msbgx_internal.h"

namespace msbgx {

static thread_local std::string g_lastError;
void setError(const std::string &msg) { g_lastError = msg; }

namespace {

struct CallbackSampler : Sampler {
  msbgx_callbacks cb;
  uint16_t cellType(int x, int y, int z) override { return cb.sample_cell_type(cb.user, x, y, z); }
  float faceCoeff(int d, int x, int y, int z) override { return cb.sample_face_coeff(cb.user, d, x, y, z); }
  float rhs(int x, int y, int z) override { return cb.sample_rhs(cb.user, x, y, z); }
};

struct DenseSampler : Sampler {
  int nx, ny;
  const uint16_t *flags;
  const float *w[3], *r;
  size_t idx(int x, int y, int z) const { return (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z); }
  uint16_t cellType(int x, int y, int z) override { return flags[idx(x, y, z)]; }
  float faceCoeff(int d, int x, int y, int z) override { return w[d][idx(x, y, z)]; }
  float rhs(int x, int y, int z) override { return r ? r[idx(x, y, z)] : 0.0f; }
  bool parallelSafe() const override { return true; }
};

/* Scene sampling on the host for cell types and face coefficients; the
 * divergence right-hand side is evaluated on the device. */
struct SceneSampler : Sampler {
  const msbgx_scene *s;
  uint16_t cellType(int x, int y, int z) override { return msbgx_scene_cell_type(s, x, y, z); }
  float faceCoeff(int d, int x, int y, int z) override { return msbgx_scene_face_coeff(s, d, x, y, z); }
  float rhs(int, int, int) override { return 0.0f; }
  bool parallelSafe() const override { return true; }
  bool rhsOnDevice() const override { return true; }
};

int checkGrid(const msbgx_grid *g)
{
  if (!g) {
    setError("null grid handle");
    return MSBGX_EINVAL;
  }
  return MSBGX_OK;
}


int prepareWith(Grid &g, Sampler &s, const msbgx_scene *scene)
{
  std::vector<float> rhs;
  int rc = buildMetadata(g, s, rhs);
  if (rc) return rc;
  /* non-finite inputs are rejected here instead of surfacing as a PCG
   * breakdown (the reference would propagate them, src/msbg3.cpp:3168-3171) */
  for (int L = 0; L < g.nLevels; L++) {
    const MgLevel &m = g.mg[L];
    for (int d = 0; d < 3; d++)
      for (float v : m.W[d])
        if (!std::isfinite(v)) { setError("non-finite face coefficient"); return MSBGX_EINVAL; }
    for (float v : m.D)
      if (!std::isfinite(v)) { setError("non-finite diagonal (face coefficients)"); return MSBGX_EINVAL; }
  }
  for (float v : rhs)
    if (!std::isfinite(v)) { setError("non-finite right-hand side"); return MSBGX_EINVAL; }
  g.hasScene = scene != nullptr;
  if (scene) g.scene = *scene;
  g.hostRhs = rhs;
  rc = deviceUpload(g, rhs, scene);
  if (rc) return rc;
  g.prepared = true;
  return MSBGX_OK;
}

}  // namespace
}  // namespace msbgx

using namespace msbgx;

struct msbgx_grid : Grid {};
struct msbgx_p2g {
  msbgx::P2gCtx *c = nullptr;
};

/* MSBGX_NO_GRAPH=1 in the environment makes new handles launch PCG
 * iterations eagerly (profiling runs whose tracer does not handle graph
 * capture). */
static unsigned defaultOptions()
{
  const char *e = getenv("MSBGX_NO_GRAPH");
  return (e && e[0] == '1') ? MSBGX_OPT_NO_GRAPH : 0u;
}

/* Single-grid operations that have no meaning on a batch handle. */
static int checkSingle(const msbgx_grid *g)
{
  if (int rc = checkGrid(g)) return rc;
  if (g->batched) {
    setError("operation not available on a batch handle");
    return MSBGX_EINVAL;
  }
  return MSBGX_OK;
}

static bool validScene(const msbgx_grid *g, const msbgx_scene *scene)
{
  if (!scene || scene->nx != g->sx0 || scene->ny != g->sy0 || scene->nz != g->sz0) {
    setError("scene size must equal the grid size");
    return false;
  }
  if (scene->n_spheres < 0 || scene->n_spheres > MSBGX_SCENE_MAX_SPHERES || scene->n_boxes < 0 ||
      scene->n_boxes > MSBGX_SCENE_MAX_BOXES || !(scene->dt != 0.0f) || scene->particles_per_cell < 0 ||
      scene->particles_per_cell > 64) {
    setError("invalid scene");
    return false;
  }
  return true;
}

extern "C" {

const char *msbgx_last_error(void) { return g_lastError.c_str(); }

void msbgx_default_solve_params(msbgx_solve_params *p)
{
  if (!p) return;
  p->n_pre = 3;
  p->n_post = 3;
  p->n_coarse = 24;
  p->max_iter = 500;
  p->tol = 5e-4f;
}

int msbgx_create(msbgx_grid **out, int sx0, int sy0, int sz0, int bsx0, int n_levels)
{
  if (!out) { setError("out is null"); return MSBGX_EINVAL; }
  *out = nullptr;
  if (bsx0 < 4 || bsx0 > 32 || (bsx0 & (bsx0 - 1))) { setError("bsx0 must be 4, 8, 16 or 32"); return MSBGX_EINVAL; }
  if (n_levels < 1 || n_levels > kMaxLevels || (bsx0 >> (n_levels - 1)) < 2) {
    setError("n_levels out of range for the block size");
    return MSBGX_EINVAL;
  }
  if (sx0 <= 0 || sy0 <= 0 || sz0 <= 0 || sx0 % bsx0 || sy0 % bsx0 || sz0 % bsx0) {
    setError("grid dimensions must be positive multiples of the block size");
    return MSBGX_EINVAL;
  }
  msbgx_grid *g = new (std::nothrow) msbgx_grid();
  if (!g) { setError("out of host memory"); return MSBGX_ENOMEM; }
  g->sx0 = sx0; g->sy0 = sy0; g->sz0 = sz0; g->bsx0 = bsx0; g->nLevels = n_levels;
  g->nbx = sx0 / bsx0; g->nby = sy0 / bsx0; g->nbz = sz0 / bsx0;
  g->nBlocks = g->nbx * g->nby * g->nbz;
  for (int i = 0; i < 3; i++) g->domainBc[i][0] = g->domainBc[i][1] = MSBGX_DBC_OPEN;
  std::vector<int> zero(g->nBlocks, 0);
  setRefinementMap(*g, zero.data());
  g->options = defaultOptions();
  *out = g;
  return MSBGX_OK;
}

void msbgx_destroy(msbgx_grid *g)
{
  if (!g) return;
  deviceFree(*g);
  delete g;
}

int msbgx_num_blocks(const msbgx_grid *g) { return g ? g->nBlocks : 0; }

int msbgx_set_options(msbgx_grid *g, unsigned options)
{
  if (int rc = checkGrid(g)) return rc;
  g->options = options;
  deviceInvalidateGraphs(*g);   // captured graphs bake in the V-cycle form
  return MSBGX_OK;
}

int msbgx_set_domain_bc(msbgx_grid *g, int l, int r, int b, int t, int f, int k)
{
  if (int rc = checkSingle(g)) return rc;
  const int v[6] = {l, r, b, t, f, k};
  for (int i = 0; i < 6; i++)
    if (v[i] != MSBGX_DBC_OPEN && v[i] != MSBGX_DBC_SOLID) { setError("unknown boundary kind"); return MSBGX_EINVAL; }
  g->domainBc[0][0] = l; g->domainBc[0][1] = r;
  g->domainBc[1][0] = b; g->domainBc[1][1] = t;
  g->domainBc[2][0] = f; g->domainBc[2][1] = k;
  g->prepared = false;
  return MSBGX_OK;
}

int msbgx_regularize_refinement_map(const msbgx_grid *g, const int *in, int *out)
{
  if (int rc = checkSingle(g)) return rc;
  if (!in || !out) { setError("null map"); return MSBGX_EINVAL; }
  if (g->nLevels < 2) { std::copy(in, in + g->nBlocks, out); return MSBGX_OK; }
  regularize(*g, in, out);
  return MSBGX_OK;
}

int msbgx_set_refinement_map(msbgx_grid *g, const int *levels)
{
  if (int rc = checkSingle(g)) return rc;
  if (!levels) { setError("null map"); return MSBGX_EINVAL; }
  for (int i = 0; i < g->nBlocks; i++)
    if (levels[i] < 0 || levels[i] >= g->nLevels) { setError("block level out of range"); return MSBGX_EINVAL; }
  /* 2:1 grading over the 26-neighbourhood is required (src/msbg.cpp:1468-1494) */
  for (int bid = 0; bid < g->nBlocks; bid++) {
    int bx, by, bz;
    g->blockCoords(bid, bx, by, bz);
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          if (!g->inRaster(bx + dx, by + dy, bz + dz)) continue;
          const int l2 = levels[g->blockIndex(bx + dx, by + dy, bz + dz)];
          if (std::abs(l2 - levels[bid]) > 1) { setError("refinement map is not 2:1 graded"); return MSBGX_EINVAL; }
        }
  }
  setRefinementMap(*g, levels);
  return MSBGX_OK;
}

int msbgx_prepare_pressure_solve(msbgx_grid *g, const msbgx_callbacks *cb)
{
  if (int rc = checkSingle(g)) return rc;
  if (!cb || !cb->sample_cell_type || !cb->sample_face_coeff || !cb->sample_rhs) { setError("incomplete callbacks"); return MSBGX_EINVAL; }
  CallbackSampler s;
  s.cb = *cb;
  return prepareWith(*g, s, nullptr);
}

int msbgx_prepare_pressure_solve_dense(msbgx_grid *g, const uint16_t *flags, const float *fu, const float *fv,
                                       const float *fw, const float *rhs)
{
  if (int rc = checkSingle(g)) return rc;
  if (!flags || !fu || !fv || !fw) { setError("null input array"); return MSBGX_EINVAL; }
  DenseSampler s;
  s.nx = g->sx0; s.ny = g->sy0; s.flags = flags; s.w[0] = fu; s.w[1] = fv; s.w[2] = fw; s.r = rhs;
  return prepareWith(*g, s, nullptr);
}

int msbgx_prepare_pressure_solve_scene(msbgx_grid *g, const msbgx_scene *scene)
{
  if (int rc = checkSingle(g)) return rc;
  if (!validScene(g, scene)) return MSBGX_EINVAL;
  SceneSampler s;
  s.s = scene;
  return prepareWith(*g, s, scene);
}

int msbgx_solve_pressure(msbgx_grid *g, const msbgx_solve_params *p, int *iters, float *rel)
{
  if (int rc = checkSingle(g)) return rc;
  if (!p || !iters || !rel) { setError("null argument"); return MSBGX_EINVAL; }
  if (!g->prepared) { setError("solve before prepare"); return MSBGX_ESTATE; }
  if (p->n_pre < 0 || p->n_post < 0 || p->n_coarse < 0 || p->max_iter < 0 || !(p->tol >= 0.0f)) {
    setError("invalid solve parameters");
    return MSBGX_EINVAL;
  }
  return deviceSolve(*g, *p, iters, rel);
}

int msbgx_set_precision(msbgx_grid *g, int prec)
{
  if (int rc = checkGrid(g)) return rc;
  if (prec != MSBGX_PREC_FP32 && prec != MSBGX_PREC_FP16 && prec != MSBGX_PREC_BF16) {
    setError("unknown precision");
    return MSBGX_EINVAL;
  }
  g->prec = prec;
  if (g->prepared) return deviceSetPrecision(*g, prec);
  return MSBGX_OK;
}

int msbgx_solve_pressure_refined(msbgx_grid *g, const msbgx_solve_params *p, double tol, int max_outer,
                                 int *outer_iters, int *inner_iters, double *rel, double *x64_out)
{
  if (int rc = checkSingle(g)) return rc;
  if (!p || !outer_iters || !inner_iters || !rel || !(tol > 0.0) || max_outer < 1) {
    setError("bad argument");
    return MSBGX_EINVAL;
  }
  if (!g->prepared) { setError("solve before prepare"); return MSBGX_ESTATE; }
  return deviceSolveRefined(*g, *p, tol, max_outer, outer_iters, inner_iters, rel, x64_out);
}

int msbgx_level0_blocks(const msbgx_grid *g, int *out, int cap)
{
  if (!g) return 0;
  for (int i = 0; i < (int)g->active0.size() && i < cap; i++) out[i] = g->active0[i];
  return (int)g->active0.size();
}

int msbgx_residual_history(const msbgx_grid *g, double *out, int cap)
{
  if (!g) return 0;
  const int n = out ? std::min<int>(cap, (int)g->history.size()) : 0;
  for (int i = 0; i < n; i++) out[i] = g->history[i];
  return (int)g->history.size();
}

double msbgx_last_solve_ms(const msbgx_grid *g) { return g ? g->lastSolveMs : 0.0; }

int msbgx_get_channel_block(const msbgx_grid *g, int ch, int mg, int bid, void *dst)
{
  if (int rc = checkGrid(g)) return rc;
  if (mg < 0 || mg >= g->nLevels || bid < 0 || bid >= g->nBlocks || !dst) { setError("bad block address"); return MSBGX_EINVAL; }
  if (!g->prepared) { setError("not prepared"); return MSBGX_ESTATE; }
  return deviceGetBlock(*g, ch, mg, bid, dst);
}

int msbgx_set_channel_block(msbgx_grid *g, int ch, int mg, int bid, const float *src)
{
  if (int rc = checkGrid(g)) return rc;
  if (mg < 0 || mg >= g->nLevels || bid < 0 || bid >= g->nBlocks || !src) { setError("bad block address"); return MSBGX_EINVAL; }
  if (!g->prepared) { setError("not prepared"); return MSBGX_ESTATE; }
  return deviceSetBlock(*g, ch, mg, bid, src);
}

int msbgx_block_info(const msbgx_grid *g, int mg, int *level, uint32_t *flags)
{
  if (int rc = checkGrid(g)) return rc;
  if (mg < 0 || mg >= g->nLevels || !level || !flags) { setError("bad argument"); return MSBGX_EINVAL; }
  for (int b = 0; b < g->nBlocks; b++) {
    level[b] = g->mg[mg].level[b];
    flags[b] = g->mg[mg].bflags[b];
  }
  return MSBGX_OK;
}

int msbgx_relax_list(const msbgx_grid *g, int mg, int *out, int cap)
{
  if (!g || mg < 0 || mg >= g->nLevels) return 0;
  const std::vector<int> &r = g->mg[mg].relax;
  for (int i = 0; i < (int)r.size() && i < cap; i++) out[i] = r[i];
  return (int)r.size();
}

float msbgx_omega(const msbgx_grid *g) { return g ? g->omega : 0.0f; }

#define MSBGX_PREPARED_OR_FAIL(g)                 \
  do {                                             \
    if (int rc_ = checkGrid(g)) return rc_;        \
    if (!(g)->prepared) {                          \
      setError("not prepared");                    \
      return MSBGX_ESTATE;                         \
    }                                              \
  } while (0)

int msbgx_relax(msbgx_grid *g, int mg, int reverse, int n_iter, int ch_x, int ch_b)
{
  MSBGX_PREPARED_OR_FAIL(g);
  if (mg < 0 || mg >= g->nLevels || n_iter < 0) { setError("bad argument"); return MSBGX_EINVAL; }
  return deviceRelax(*g, mg, reverse != 0, n_iter, ch_x, ch_b);
}
int msbgx_multiply(msbgx_grid *g, int mg, int residual, int ch_x, int ch_b, int ch_dst)
{
  MSBGX_PREPARED_OR_FAIL(g);
  if (mg < 0 || mg >= g->nLevels) { setError("bad argument"); return MSBGX_EINVAL; }
  return deviceMultiply(*g, mg, residual != 0, ch_x, ch_b, ch_dst);
}
int msbgx_downsample(msbgx_grid *g, int mg_coarse, int ch_src, int ch_dst)
{
  MSBGX_PREPARED_OR_FAIL(g);
  if (mg_coarse < 1 || mg_coarse >= g->nLevels) { setError("bad argument"); return MSBGX_EINVAL; }
  return deviceDownsample(*g, mg_coarse, ch_src, ch_dst);
}
int msbgx_prolongate(msbgx_grid *g, int mg_fine, int ch)
{
  MSBGX_PREPARED_OR_FAIL(g);
  if (mg_fine < 0 || mg_fine + 1 >= g->nLevels) { setError("bad argument"); return MSBGX_EINVAL; }
  return deviceProlongate(*g, mg_fine, ch);
}
int msbgx_zero_channel(msbgx_grid *g, int ch, int mg)
{
  MSBGX_PREPARED_OR_FAIL(g);
  if (mg < 0 || mg >= g->nLevels) { setError("bad argument"); return MSBGX_EINVAL; }
  return deviceZero(*g, ch, mg);
}
int msbgx_copy_channel(msbgx_grid *g, int ch_src, int ch_dst)
{
  MSBGX_PREPARED_OR_FAIL(g);
  return deviceCopy(*g, ch_src, ch_dst);
}
int msbgx_vcycle(msbgx_grid *g, int n_pre, int n_post, int n_coarse)
{
  MSBGX_PREPARED_OR_FAIL(g);
  return deviceVCycle(*g, n_pre, n_post, n_coarse);
}

int msbgx_time_kernel(msbgx_grid *g, int which, int reps, double *ms, double *bytes, int *cells)
{
  MSBGX_PREPARED_OR_FAIL(g);
  if (which < 0 || which > 4 || reps < 1 || !ms || !bytes || !cells) { setError("bad argument"); return MSBGX_EINVAL; }
  if (which >= 3) return deviceFlipTime(*g, which, reps, ms, bytes, cells);
  return deviceTimeKernel(*g, which, reps, ms, bytes, cells);
}

static int checkParams(const msbgx_solve_params *p)
{
  if (!p || p->n_pre < 0 || p->n_post < 0 || p->n_coarse < 0 || p->max_iter < 0 || !(p->tol >= 0.0f)) {
    setError("invalid solve parameters");
    return MSBGX_EINVAL;
  }
  return MSBGX_OK;
}

int msbgx_batch_create(msbgx_grid **out, msbgx_grid *const *problems, int n)
{
  if (!out) { setError("out is null"); return MSBGX_EINVAL; }
  *out = nullptr;
  if (!problems || n < 1) { setError("empty batch"); return MSBGX_EINVAL; }
  std::vector<const Grid *> probs;
  for (int i = 0; i < n; i++) {
    if (!problems[i] || problems[i]->batched) { setError("batch members must be single prepared grids"); return MSBGX_EINVAL; }
    probs.push_back(problems[i]);
  }
  msbgx_grid *m = new (std::nothrow) msbgx_grid();
  if (!m) { setError("out of host memory"); return MSBGX_ENOMEM; }
  int rc = buildBatchGrid(*m, probs);
  if (!rc) rc = deviceUploadBatch(*m, probs);
  if (rc) {
    deviceFree(*m);
    delete m;
    return rc;
  }
  m->prepared = true;
  m->options = defaultOptions();
  *out = m;
  return MSBGX_OK;
}

int msbgx_batch_size(const msbgx_grid *b) { return b ? (b->batched ? b->nProb : 1) : 0; }

int msbgx_batch_block_base(const msbgx_grid *b, int problem)
{
  if (int rc = checkGrid(b)) return rc;
  if (!b->batched) return problem == 0 ? 0 : (setError("problem index out of range"), MSBGX_EINVAL);
  if (problem < 0 || problem >= b->nProb) { setError("problem index out of range"); return MSBGX_EINVAL; }
  return b->probBlock0[problem];
}

int msbgx_batch_solve(msbgx_grid *b, const msbgx_solve_params *p, int *iters, float *rel)
{
  if (int rc = checkGrid(b)) return rc;
  if (!b->batched) { setError("not a batch handle"); return MSBGX_EINVAL; }
  if (!iters || !rel) { setError("null argument"); return MSBGX_EINVAL; }
  if (int rc = checkParams(p)) return rc;
  return deviceSolveBatch(*b, *p, iters, rel);
}

int msbgx_batch_residual_history(const msbgx_grid *b, int problem, double *out, int cap)
{
  if (!b || !b->batched || problem < 0 || problem >= (int)b->probHistory.size()) return 0;
  const std::vector<double> &h = b->probHistory[problem];
  for (int i = 0; out && i < (int)h.size() && i < cap; i++) out[i] = h[i];
  return (int)h.size();
}

int msbgx_autotune(msbgx_grid *g, const msbgx_solve_params *p, int *graph_chunk, int *smoother_mask)
{
  MSBGX_PREPARED_OR_FAIL(g);
  if (!graph_chunk || !smoother_mask) { setError("null argument"); return MSBGX_EINVAL; }
  if (int rc = checkParams(p)) return rc;
  return deviceAutotune(*g, *p, graph_chunk, smoother_mask);
}

int msbgx_set_launch_config(msbgx_grid *g, int graph_chunk, int smoother_mask)
{
  if (int rc = checkGrid(g)) return rc;
  if (graph_chunk < 1 || graph_chunk > 64 || smoother_mask < 0 || smoother_mask >= (1 << kMaxLevels)) {
    setError("bad launch configuration");
    return MSBGX_EINVAL;
  }
  g->graphChunk = graph_chunk;
  for (int l = 0; l < kMaxLevels; l++) g->smootherVariant[l] = (smoother_mask >> l) & 1;
  deviceInvalidateGraphs(*g);
  return MSBGX_OK;
}

int msbgx_net_create(msbgx_net **out, uint64_t seed)
{
  if (!out) { setError("out is null"); return MSBGX_EINVAL; }
  *out = nullptr;
  Net *n = nullptr;
  const int rc = netCreate(&n, seed);
  if (rc) { netDestroy(n); return rc; }
  *out = reinterpret_cast<msbgx_net *>(n);
  return MSBGX_OK;
}

void msbgx_net_destroy(msbgx_net *n) { netDestroy(reinterpret_cast<Net *>(n)); }

static Net *netOf(msbgx_net *n) { return reinterpret_cast<Net *>(n); }

int msbgx_net_get_weights(msbgx_net *n, float *w)
{
  if (!n || !w) { setError("null argument"); return MSBGX_EINVAL; }
  return netGetWeights(*netOf(n), w);
}

int msbgx_net_set_weights(msbgx_net *n, const float *w)
{
  if (!n || !w) { setError("null argument"); return MSBGX_EINVAL; }
  return netSetWeights(*netOf(n), w);
}

int msbgx_net_set_samples(msbgx_net *n, int count, const float *x, const float *target, const uint8_t *mask)
{
  if (!n || count < 1 || !x || !target || !mask) { setError("bad argument"); return MSBGX_EINVAL; }
  return netSetSamples(*netOf(n), count, x, target, mask);
}

int msbgx_net_loss_grad(msbgx_net *n, double *loss, float *grad, float *out)
{
  if (!n || !loss) { setError("null argument"); return MSBGX_EINVAL; }
  return netLossGradAll(*netOf(n), loss, grad, out);
}

static int checkLearnedGrid(msbgx_grid *g, const msbgx_net *n)
{
  if (int rc = checkSingle(g)) return rc;
  if (!g->prepared) { setError("not prepared"); return MSBGX_ESTATE; }
  /* the fine form works on 16^3 blocks; the coarse form on block means of any size */
  if (g->bsx0 != 16 && msbgx_net_get_mode(n) != MSBGX_NET_COARSE) {
    setError("the fine learned correction needs 16^3 base blocks");
    return MSBGX_EINVAL;
  }
  return MSBGX_OK;
}

int msbgx_net_collect(msbgx_net *n, msbgx_grid *g, const msbgx_solve_params *p, int every, int max_pairs,
                      int blocks_per_pair, uint64_t seed, int *pairs)
{
  if (!n || !pairs) { setError("null argument"); return MSBGX_EINVAL; }
  if (int rc = checkLearnedGrid(g, n)) return rc;
  if (int rc = checkParams(p)) return rc;
  if (every < 1 || max_pairs < 0 || blocks_per_pair < 1) { setError("bad argument"); return MSBGX_EINVAL; }
  return netCollect(*netOf(n), *g, *p, every, max_pairs, blocks_per_pair, seed, pairs);
}

int msbgx_net_set_mode(msbgx_net *n, int mode)
{
  if (!n || (mode != MSBGX_NET_FINE && mode != MSBGX_NET_COARSE)) { setError("bad argument"); return MSBGX_EINVAL; }
  return netSetMode(*netOf(n), mode);
}

int msbgx_net_set_stall(msbgx_net *n, int window, float drop)
{
  if (!n || window < 0 || window > 32 || !(drop >= 0.0f && drop < 1.0f)) { setError("bad argument"); return MSBGX_EINVAL; }
  return netSetStall(*netOf(n), window, drop);
}

int msbgx_net_get_mode(const msbgx_net *n) { return n ? netMode(*reinterpret_cast<const Net *>(n)) : -1; }

int msbgx_net_num_samples(const msbgx_net *n) { return n ? netNumSamples(*reinterpret_cast<const Net *>(n)) : 0; }

int msbgx_net_train(msbgx_net *n, int steps, int batch, float lr, uint64_t seed, double *loss_first,
                    double *loss_last)
{
  if (!n || steps < 0 || batch < 1 || !(lr > 0.0f)) { setError("bad argument"); return MSBGX_EINVAL; }
  return netTrain(*netOf(n), steps, batch, lr, seed, loss_first, loss_last);
}

int msbgx_net_set_prolongation(msbgx_net *n, int kind)
{
  if (!n || (kind != MSBGX_NET_PROLONG_CONSTANT && kind != MSBGX_NET_PROLONG_LINEAR)) {
    setError("bad argument");
    return MSBGX_EINVAL;
  }
  return netSetProlongation(*reinterpret_cast<Net *>(n), kind);
}

int msbgx_net_get_prolongation(const msbgx_net *n)
{
  return n ? netProlongation(*reinterpret_cast<const Net *>(n)) : MSBGX_EINVAL;
}

int msbgx_net_set_coarse_scales(msbgx_net *n, int scales)
{
  if (!n || scales < 1 || scales > MSBGX_NET_MAX_SCALES) {
    setError("bad argument");
    return MSBGX_EINVAL;
  }
  return netSetScales(*reinterpret_cast<Net *>(n), scales);
}

int msbgx_net_get_coarse_scales(const msbgx_net *n)
{
  return n ? netScales(*reinterpret_cast<const Net *>(n)) : MSBGX_EINVAL;
}

double msbgx_net_last_train_ms(const msbgx_net *n) { return n ? netLastTrainMs(*reinterpret_cast<const Net *>(n)) : 0.0; }

int msbgx_net_apply(msbgx_net *n, msbgx_grid *g, int ch_in, int ch_out)
{
  if (!n) { setError("null argument"); return MSBGX_EINVAL; }
  if (int rc = checkLearnedGrid(g, n)) return rc;
  if (g->bsx0 != 16 && netMode(*netOf(n)) != MSBGX_NET_COARSE) {
    setError("msbgx_net_apply: the fine form works on 16^3 blocks");
    return MSBGX_EINVAL;
  }
  if (ch_in < 0 || ch_in > MSBGX_CH_Q || ch_out < 0 || ch_out > MSBGX_CH_Q || ch_in == ch_out) {
    setError("bad channel");
    return MSBGX_EINVAL;
  }
  return netApply(*netOf(n), *g, ch_in, ch_out);
}

int msbgx_net_time_conv(msbgx_net *n, int blocks, int reps, double *ms_per_launch, double *flop_per_launch)
{
  if (!n || blocks < 1 || reps < 1 || !ms_per_launch || !flop_per_launch) { setError("bad argument"); return MSBGX_EINVAL; }
  return netTimeConv(*netOf(n), blocks, reps, ms_per_launch, flop_per_launch);
}

int msbgx_solve_pressure_learned(msbgx_grid *g, msbgx_net *n, const msbgx_solve_params *p, int *iters, float *rel,
                                 int *fallback_iter)
{
  if (!n || !iters || !rel || !fallback_iter) { setError("null argument"); return MSBGX_EINVAL; }
  if (int rc = checkLearnedGrid(g, n)) return rc;
  if (int rc = checkParams(p)) return rc;
  return deviceSolveLearned(*g, *netOf(n), *p, iters, rel, fallback_iter);
}

int msbgx_flip_set_velocity(msbgx_grid *g, const float *u, const float *v, const float *w)
{
  if (int rc = checkSingle(g)) return rc;
  if (!u || !v || !w) { setError("null velocity array"); return MSBGX_EINVAL; }
  const float *vel[3] = {u, v, w};
  return deviceFlipSetVelocity(*g, vel);
}

int msbgx_flip_set_velocity_scene(msbgx_grid *g, const msbgx_scene *scene)
{
  if (int rc = checkSingle(g)) return rc;
  if (!validScene(g, scene)) return MSBGX_EINVAL;
  return deviceFlipSetVelocityScene(*g, *scene);
}

int msbgx_flip_get_velocity(msbgx_grid *g, float *u, float *v, float *w)
{
  if (int rc = checkSingle(g)) return rc;
  if (!u || !v || !w) { setError("null velocity array"); return MSBGX_EINVAL; }
  float *vel[3] = {u, v, w};
  return deviceFlipGetVelocity(*g, vel);
}

int msbgx_flip_rhs_from_velocity(msbgx_grid *g, float dt)
{
  if (int rc = checkSingle(g)) return rc;
  if (!std::isfinite(dt) || dt == 0.0f) { setError("dt must be finite and nonzero"); return MSBGX_EINVAL; }
  return deviceFlipRhs(*g, dt);
}

int msbgx_flip_update_velocity(msbgx_grid *g, float dt, int sign_mode)
{
  if (int rc = checkSingle(g)) return rc;
  if (!std::isfinite(dt)) { setError("dt must be finite"); return MSBGX_EINVAL; }
  if (sign_mode != MSBGX_FLIP_SIGN_REFERENCE && sign_mode != MSBGX_FLIP_SIGN_CORRECTED) {
    setError("sign_mode must be MSBGX_FLIP_SIGN_REFERENCE or MSBGX_FLIP_SIGN_CORRECTED");
    return MSBGX_EINVAL;
  }
  return deviceFlipUpdate(*g, dt, sign_mode == MSBGX_FLIP_SIGN_CORRECTED ? 1 : 0);
}

void msbgx_default_p2g_params(msbgx_p2g_params *p)
{
  if (!p) return;
  p->rho_l = 1000.0f;
  p->rho_g = 1.0f;
  p->mass_eps = 1e-5f;
  p->beta_min = 1e-6f;
  p->alpha_phi = 1.0f;
  p->particles_per_voxel = 8;
}

int msbgx_p2g_create(msbgx_p2g **out, const msbgx_grid *g)
{
  if (!out) { setError("out is null"); return MSBGX_EINVAL; }
  *out = nullptr;
  if (int rc = checkSingle(g)) return rc;
  msbgx_p2g *t = new msbgx_p2g();
  const int rc = p2gCreate(&t->c, *g);
  if (rc) {
    p2gDestroy(t->c);
    delete t;
    return rc;
  }
  *out = t;
  return MSBGX_OK;
}

void msbgx_p2g_destroy(msbgx_p2g *t)
{
  if (!t) return;
  p2gDestroy(t->c);
  delete t;
}

int msbgx_p2g_run(msbgx_p2g *t, int64_t n, const msbgx_particle *particles, const msbgx_p2g_params *p)
{
  if (!t || !p || n < 0 || (n > 0 && !particles)) { setError("bad argument"); return MSBGX_EINVAL; }
  return p2gRun(*t->c, n, particles, *p);
}

int msbgx_p2g_run_scene(msbgx_p2g *t, const msbgx_scene *scene, const msbgx_p2g_params *p, int64_t *n_particles)
{
  if (!t || !p || !scene || !n_particles) { setError("bad argument"); return MSBGX_EINVAL; }
  return p2gRunScene(*t->c, *scene, *p, n_particles);
}

int msbgx_p2g_get(msbgx_p2g *t, float *mass, float *u, float *v, float *w, float *beta_u, float *beta_v, float *beta_w)
{
  if (!t) { setError("null transfer handle"); return MSBGX_EINVAL; }
  float *vel[3] = {u, v, w}, *beta[3] = {beta_u, beta_v, beta_w};
  return p2gGet(*t->c, mass, vel, beta);
}

int msbgx_p2g_active_blocks(msbgx_p2g *t, int *out, int cap)
{
  if (!t || cap < 0) { setError("bad argument"); return MSBGX_EINVAL; }
  return p2gActiveBlocks(*t->c, out, cap);
}

int msbgx_p2g_last_times(const msbgx_p2g *t, double *ms_total, double *ms_gather, int64_t *n_particles)
{
  if (!t) { setError("null transfer handle"); return MSBGX_EINVAL; }
  p2gTimes(*t->c, ms_total, ms_gather, n_particles);
  return MSBGX_OK;
}

int msbgx_p2g_to_velocity(msbgx_p2g *t, msbgx_grid *g)
{
  if (!t) { setError("null transfer handle"); return MSBGX_EINVAL; }
  if (int rc = checkSingle(g)) return rc;
  return p2gToVelocity(*t->c, *g);
}

int msbgx_solve_pressure_compact(msbgx_grid *g, int max_iter, float tol, int *iters, float *rel_residual)
{
  if (int rc = checkSingle(g)) return rc;
  if (!iters || !rel_residual || max_iter < 0 || !(tol >= 0.0f)) { setError("bad argument"); return MSBGX_EINVAL; }
  double ms = 0.0;
  const int rc = deviceSolveCompact(*g, max_iter, tol, iters, rel_residual, &ms);
  if (rc == MSBGX_OK) g->lastSolveMs = ms;
  return rc;
}

int msbgx_compact_mic0(msbgx_grid *g, float *precon, float *z, int cap)
{
  if (int rc = checkSingle(g)) return rc;
  if (cap < 0 || (cap > 0 && (!precon || !z))) { setError("bad argument"); return MSBGX_EINVAL; }
  int n = 0;
  if (int rc = deviceCompactApply(*g, precon, z, cap, &n)) return rc;
  return n;
}

int msbgx_pde_mean_curvature(msbgx_grid *g, uint16_t *density, const int *active_blocks, int n_active, int num_iter,
                             float time_step, double *ms)
{
  if (int rc = checkSingle(g)) return rc;
  if (!density || n_active < 0 || (n_active > 0 && !active_blocks) || num_iter < 0 || !std::isfinite(time_step)) {
    setError("bad argument");
    return MSBGX_EINVAL;
  }
  return devicePdeMeanCurvature(*g, density, active_blocks, n_active, num_iter, time_step, ms);
}

int msbgx_pde_apply(msbgx_grid *g, int dtype, void *channel, const int *active_blocks, int n_active,
                    const msbgx_pde_params *params, double *ms)
{
  if (int rc = checkSingle(g)) return rc;
  if (!channel || !params || n_active < 0 || (n_active > 0 && !active_blocks)) {
    setError("bad argument");
    return MSBGX_EINVAL;
  }
  return devicePdeApply(*g, dtype, channel, active_blocks, n_active, *params, ms);
}

int msbgx_pde_mean_curvature_noalloc(msbgx_grid *g, uint16_t *density, const int *active_blocks, int n_active,
                                     int num_iter, float time_step)
{
  if (int rc = checkSingle(g)) return rc;
  if (!density || n_active < 0 || (n_active > 0 && !active_blocks) || num_iter < 0 || !std::isfinite(time_step)) {
    setError("bad argument");
    return MSBGX_EINVAL;
  }
  return devicePdeMeanCurvatureNoAlloc(*g, density, active_blocks, n_active, num_iter, time_step);
}

int msbgx_level_map_from_scene(const msbgx_scene *s, int bsx0, int n_levels, int mode, int d0, int *out)
{
  const int rc = levelMapFromScene(s, bsx0, n_levels, mode, d0, out);
  if (rc) setError("invalid scene or block size");
  return rc;
}

}  // extern "C"
