/*
 * Internal state of an msbgx grid handle: host-side block topology and
 * metadata (built once per prepare) and the device pools the kernels use.
 *
 * Data layout in HBM, per MG level L:
 *   every block of the raster owns one slot of edge bsx0 >> level(L, bid)
 *   cells in each per-level pool (x fastest, then y, then z inside a block);
 *   off[bid] is the slot's first cell.  At L = 0 the slots of the blocks the
 *   PCG vector operations touch (level-0 blocks with fluid or a nonzero
 *   right-hand side) come first, so those operations stream one contiguous
 *   range.  Channels: flags (u16), face coefficients W_x/W_y/W_z, inverse
 *   diagonal D, and the work vectors (L = 0: x, b, r, z, d, q; L >= 1: r, z, q).
 */
#ifndef MSBGX_INTERNAL_H
#define MSBGX_INTERNAL_H

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/msbgx.h"

namespace msbgx {

constexpr int kMaxLevels = 5;

/* Cell flags (src/msbg.h:600-625) */
constexpr uint16_t CELL_SOLID = 1u << 0, CELL_FIXED = 1u << 1, CELL_PARTIAL_SOLID = 1u << 6,
                   CELL_VOID = 1u << 12, CELL_EMPTY_VAL = MSBGX_CELL_EMPTY_VAL;
/* Block flags (src/msbg.h:558-576) */
constexpr uint32_t BLK_DOM_BORDER = 1u << 1, BLK_COARSE_FINE = 1u << 2, BLK_FINE_COARSE = 1u << 3,
                   BLK_NO_FLUID = 1u << 5, BLK_FIXED = 1u << 11, BLK_EXISTS = 1u << 12,
                   BLK_CUTS_SOLID = 1u << 15;

enum Chan { C_X = 0, C_B = 1, C_R = 2, C_Z = 3, C_D = 4, C_Q = 5, N_VEC = 6 };

struct DevList {
  int *ptr = nullptr;  // device pointer into the level's list buffer
  int n = 0;
};

/* Face neighbours of a block at one MG level (0..5 = -x, +x, -y, +y, -z, +z):
 * their slot offsets and levels, level -1 outside the domain.  Built on the
 * host from the raster or the batched neighbour table, so a kernel reads a
 * block's halo topology with one load. */
struct NbDesc {
  int64_t off[6];
  int lev[6];
  int pad[4];
};

struct MgLevel {
  /* host topology */
  std::vector<int> level;       // block level at this MG level = max(map, L)
  std::vector<uint32_t> bflags; // BLK_* flags
  std::vector<int64_t> off;     // slot offset (cells) in this level's pools
  int64_t ncells = 0;
  std::vector<int> relax;       // relaxation blocks, ascending bid
  std::vector<int> act[kMaxLevels];            // non-NO_FLUID blocks by level
  std::vector<int> relaxCL[2][kMaxLevels];     // relax blocks by colour parity and level
  /* host metadata in this level's slot layout */
  std::vector<uint16_t> flags;
  std::vector<float> W[3], D;
  /* device */
  int *d_level = nullptr;
  int64_t *d_off = nullptr;
  NbDesc *d_nbd = nullptr;      // per block: face neighbours' offsets and levels
  uint16_t *d_flags = nullptr;
  float *d_W[3] = {nullptr, nullptr, nullptr};
  float *d_D = nullptr;
  uint16_t *d_H[4] = {nullptr, nullptr, nullptr, nullptr};  // 16-bit W_x, W_y, W_z, D (precision form)
  float *d_ch[N_VEC] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int *d_lists = nullptr;
  DevList dAct[kMaxLevels];
  DevList dRelax[2][kMaxLevels];
  bool anyRelax = false;        // relax list non-empty
};

struct Grid {
  int sx0, sy0, sz0, bsx0, nLevels, nbx, nby, nbz, nBlocks;
  int domainBc[3][2];
  unsigned options = 0;
  std::vector<int> map;         // block level map
  MgLevel mg[kMaxLevels];
  float omega = 6.f / 7.f;
  bool prepared = false;
  bool aliasEstablished[kMaxLevels] = {false, false, false, false, false};
  int nActive0 = 0;             // blocks at the head of the MG0 pool touched by PCG vector ops
  std::vector<int> active0;     // those blocks, in pool order
  int prec = MSBGX_PREC_FP32;   // smoother operand storage
  /* device state */
  void *dev = nullptr;          // solver state (device buffers, stream, graph)
  std::vector<double> history;
  double lastSolveMs = 0.0;
  /* inputs kept for building batched block sets */
  bool hasScene = false;
  msbgx_scene scene{};
  std::vector<float> hostRhs;   // MG0 pool layout (callback / dense prepare paths)
  /* batched block sets: several independent problems merged into one */
  bool batched = false;
  int nProb = 1;
  std::vector<std::vector<double>> probHistory;  // per-problem relative residuals of the last batched solve
  std::vector<int> probBlock0;  // first block of each problem
  std::vector<int> probOf;      // problem of each block
  std::vector<float> probOmega; // smoother weight per problem
  std::vector<int> nbrTable;    // 6 face neighbours per block (-1 = domain boundary)
  std::vector<int> probActive0; // PCG-active level-0 blocks per problem
  std::vector<std::vector<int>> probRaster;  // nbx, nby, nbz per problem
  int smootherVariant[kMaxLevels] = {0, 0, 0, 0, 0};  // per block level: 0 column-resident, 1 LDS tile
  int graphChunk = 4;           // PCG iterations per hipGraph replay

  int bsxL(int l) const { return bsx0 >> l; }
  void blockCoords(int bid, int &bx, int &by, int &bz) const
  {
    bx = bid % nbx;
    by = (bid / nbx) % nby;
    bz = bid / (nbx * nby);
  }
  bool inRaster(int bx, int by, int bz) const
  {
    return bx >= 0 && by >= 0 && bz >= 0 && bx < nbx && by < nby && bz < nbz;
  }
  int blockIndex(int bx, int by, int bz) const { return bx + nbx * (by + nby * bz); }
};

/* host topology (topology.cpp) */
void regularize(const Grid &g, const int *in, int *out);
void setRefinementMap(Grid &g, const int *levels);

/* Source of the level-0 inputs sampled during prepare. */
struct Sampler {
  virtual ~Sampler() {}
  virtual uint16_t cellType(int x, int y, int z) = 0;
  virtual float faceCoeff(int dir, int x, int y, int z) = 0;
  virtual float rhs(int x, int y, int z) = 0;
  virtual bool parallelSafe() const { return false; }
  virtual bool rhsOnDevice() const { return false; }
};
/* Builds all host metadata of every MG level; fills host rhs (MG0 layout)
 * unless the sampler computes it on the device. */
int buildMetadata(Grid &g, Sampler &s, std::vector<float> &rhs);
int levelMapFromScene(const msbgx_scene *s, int bsx0, int nLevels, int mode, int d0, int *out);

/* device side (solver.hip) */
int deviceUpload(Grid &g, const std::vector<float> &rhs, const msbgx_scene *sceneForRhs);
void deviceFree(Grid &g);
int deviceSolve(Grid &g, const msbgx_solve_params &p, int *iters, float *rel);
int deviceGetBlock(const Grid &g, int ch, int L, int bid, void *dst);
int deviceSetBlock(Grid &g, int ch, int L, int bid, const float *src);
int deviceRelax(Grid &g, int L, bool reverse, int nIter, int chX, int chB);
/* FLIP caller steps (flip.hip) */
int deviceFlipSetVelocity(Grid &g, const float *const vel[3]);
int deviceFlipGetVelocity(Grid &g, float *const vel[3]);
int deviceFlipSetVelocityScene(Grid &g, const msbgx_scene &s);
int deviceFlipRhs(Grid &g, float dt);
int deviceFlipUpdate(Grid &g, float dt, int negate);
int deviceFlipTime(Grid &g, int which, int reps, double *ms, double *bytes, int *cells);
/* particle-to-grid transfer (p2g.hip) */
struct P2gCtx;
int p2gCreate(P2gCtx **out, const Grid &g);
void p2gDestroy(P2gCtx *c);
int p2gRun(P2gCtx &c, int64_t n, const msbgx_particle *host, const msbgx_p2g_params &p);
int p2gRunScene(P2gCtx &c, const msbgx_scene &s, const msbgx_p2g_params &p, int64_t *nOut);
int p2gGet(P2gCtx &c, float *mass, float *const vel[3], float *const beta[3]);
int p2gActiveBlocks(P2gCtx &c, int *out, int cap);
void p2gTimes(const P2gCtx &c, double *msTotal, double *msGather, int64_t *particles);
int p2gToVelocity(P2gCtx &c, Grid &g);
/* compact MIC(0)-PCG comparator (compact.hip) */
int deviceSolveCompact(Grid &g, int maxIter, float tol, int *iters, float *rel, double *ms);
int deviceCompactApply(Grid &g, float *preconOut, float *zOut, int cap, int *nRows);
/* msbg_demo PDE smoothing (pde.hip) */
int devicePdeApply(const Grid &g, int dtype, void *dense, const int *active, int nActive, const msbgx_pde_params &p,
                   double *ms);
int devicePdeMeanCurvature(const Grid &g, uint16_t *dense, const int *active, int nActive, int numIter, float dt,
                           double *ms);
int devicePdeMeanCurvatureNoAlloc(const Grid &g, uint16_t *dense, const int *active, int nActive, int numIter,
                                  float dt);
int deviceMultiply(Grid &g, int L, bool residual, int chX, int chB, int chDst);
int deviceDownsample(Grid &g, int Lc, int chSrc, int chDst);
int deviceProlongate(Grid &g, int Lf, int ch);
int deviceZero(Grid &g, int ch, int L);
int deviceCopy(Grid &g, int src, int dst);
int deviceVCycle(Grid &g, int nPre, int nPost, int nCoarse);
int deviceTimeKernel(Grid &g, int which, int reps, double *ms, double *bytes, int *cells);
int deviceSetPrecision(Grid &g, int prec);
void deviceInvalidateGraphs(Grid &g);
/* learned correction (learned.hip) */
struct Net;
int netCreate(Net **out, uint64_t seed);
void netDestroy(Net *n);
int netSetWeights(Net &n, const float *w);
int netGetWeights(Net &n, float *w);
int netSetSamples(Net &n, int count, const float *x, const float *target, const uint8_t *mask);
int netLossGradAll(Net &n, double *loss, float *grad, float *out);
int netTrain(Net &n, int steps, int batch, float lr, uint64_t seed, double *lossFirst, double *lossLast);
int netCollect(Net &n, Grid &g, const msbgx_solve_params &p, int every, int maxPairs, int perPair, uint64_t seed,
               int *pairsOut);
int netNumSamples(const Net &n);
int netSetMode(Net &n, int mode);
int netMode(const Net &n);
int netSetProlongation(Net &n, int kind);
int netProlongation(const Net &n);
int netSetScales(Net &n, int k);
int netScales(const Net &n);
int netSetStall(Net &n, int window, float drop);
double netLastTrainMs(const Net &n);
int deviceSolveLearned(Grid &g, Net &n, const msbgx_solve_params &p, int *iters, float *rel, int *fellBack);
int netApply(Net &n, Grid &g, int chIn, int chOut);
int netTimeConv(Net &n, int nb, int reps, double *msPerLaunch, double *flopPerLaunch);
/* batched block sets (batch.cpp / solver.hip) */
int buildBatchGrid(Grid &merged, const std::vector<const Grid *> &probs);
int deviceUploadBatch(Grid &merged, const std::vector<const Grid *> &probs);
int deviceSolveBatch(Grid &merged, const msbgx_solve_params &p, int *iters, float *rel);
int deviceAutotune(Grid &g, const msbgx_solve_params &p, int *chunkOut, int *variantMask);
int deviceSolveRefined(Grid &g, const msbgx_solve_params &p, double tol, int maxOuter, int *outer, int *inner,
                       double *rel, double *x64out);

void setError(const std::string &msg);

}  // namespace msbgx

#endif
