/*
 * p2g.hip -- particle-to-grid transfer of the FLIP caller on the device
 * (compiled as part of solver.hip).
 *
 * Reference: FlipSimulation::particleToGrid (src/flip_sim.cpp:579-828).  Per
 * particle (position in cell units, velocity, phase 0 liquid / 1 air):
 *   cell mass += 1 in the particle's cell, liquid count += 1 for phase 0;
 *   on the u/v/w face grids (node offsets (0,.5,.5) / (.5,0,.5) / (.5,.5,0))
 *   with trilinear weight w = ((x-weight) * (y-weight)) * (z-weight):
 *   velocity sum += vel_c * w, weight sum += w, density sum += m_p * w
 *   (m_p = rho_L for liquid, rho_G for air).
 * Then, on the cells of level-0 blocks (the grids are resolution level 0):
 *   velocity = weight sum > massEps ? velocity sum / weight sum : 0;
 *   phi = density sum < rhoTildeMin ? 0 : min(sqrt(max(sum - rhoTildeMin, 0)
 *         / phiDenom), 1), forced to 1 between two liquid cells and to 0
 *         between two air cells (cell phase: -1 without data or mass <=
 *         massEps, 0 when 2 * liquid count > (int)(mass + 0.5), else 1);
 *   beta = max(1 / (rho_G + phi (rho_L - rho_G)), betaMin);
 *   active blocks = level-0 blocks with a cell of mass > massEps.
 *
 * The reference scatters from TBB worker threads into thread-local block
 * buffers and merges them, so its summation order follows the thread
 * partition.  Here the transfer is a gather in one fixed order, which makes
 * it deterministic: particles are binned by cell over the domain plus a
 * one-cell halo (k_p2g_key counts,
 * k_scan_* scan, k_p2g_place places with atomics, k_p2g_sort_cells orders
 * each cell's list by particle index), and k_p2g_gather visits, for each
 * face, the 18 cells whose particles can reach it in (z, y, x) order, each
 * cell's particles in index order.  The oracle sums in the same order.
 *
 * k_p2g_permute then copies the particles in bin order into SoA arrays and
 * k_p2g_gather makes one pass over the 27 cells around each cell, adding each
 * particle to all three faces it reaches.
 *
 * Memory: particle array 28 B, key + index 8 B and the bin-ordered copy 25 B
 * per particle; per bin one count and one start (8 B; 8.6 GB at 1024^3); outputs per cell of
 * the level-0 blocks: mass, liquid count, three velocities, three betas.
 */

namespace msbgx {

struct P2gCtx {
  int nx = 0, ny = 0, nz = 0, bsx = 0, lb = 0, nbx = 0, nby = 0, nbz = 0;
  std::vector<int> level0;   // level-0 blocks, ascending bid
  std::vector<int> slot;     // per block: rank in level0 or -1
  hipStream_t stream = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  int *dLevel0 = nullptr, *dSlot = nullptr;
  int64_t capP = 0;
  msbgx_particle *dPart = nullptr;
  uint32_t *dKey = nullptr, *dIdx = nullptr;
  float *dS[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};   // bin-ordered x, y, z, u, v, w
  uint8_t *dPh = nullptr;                                                  // bin-ordered phase
  uint32_t *dCount = nullptr, *dStart = nullptr, *dTile = nullptr;
  int64_t nCells = 0, nTiles = 0;
  int64_t nOut = 0;          // level0.size() * bsx^3
  float *dMass = nullptr, *dVel[3] = {nullptr, nullptr, nullptr}, *dBeta[3] = {nullptr, nullptr, nullptr};
  int *dLiq = nullptr;
  int *dActive = nullptr;
  bool valid = false;
  double msTotal = 0.0, msGather = 0.0;
  int64_t lastParticles = 0;
};

constexpr uint32_t kNoCell = 0xffffffffu;
constexpr int kScanTile = 4096;   // elements per workgroup of the scan (256 threads x 16)

/* Bin raster: the domain plus a one-cell halo (a particle in a halo cell
 * still reaches the faces of the domain's outer cells), x fastest. */
__host__ __device__ __forceinline__ int64_t p2gBin(int x, int y, int z, int nx, int ny)
{
  return ((int64_t)(z + 1) * (ny + 2) + (y + 1)) * (nx + 2) + (x + 1);
}

/* Bin of each particle, kNoCell beyond the halo; per-bin counts. */
__global__ void k_p2g_key(const msbgx_particle *__restrict__ P, int64_t n, int nx, int ny, int nz,
                          uint32_t *__restrict__ key, uint32_t *__restrict__ count)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ix = (int)floorf(P[i].pos[0]), iy = (int)floorf(P[i].pos[1]), iz = (int)floorf(P[i].pos[2]);
  uint32_t k = kNoCell;
  if (ix >= -1 && iy >= -1 && iz >= -1 && ix <= nx && iy <= ny && iz <= nz) {
    k = (uint32_t)p2gBin(ix, iy, iz, nx, ny);
    atomicAdd(&count[k], 1u);
  }
  key[i] = k;
}

/* Exclusive scan of count[0..n) into start[0..n] in three passes: tile sums,
 * a scan of the tile sums by one workgroup, and the tiles with their offsets. */
__global__ void __launch_bounds__(256) k_scan_tiles(const uint32_t *__restrict__ in, int64_t n, uint32_t *__restrict__ tile)
{
  __shared__ uint32_t part[256];
  const int64_t b = (int64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
  for (int k = 0; k < kScanTile / 256; k++) {
    const int64_t i = b + (int64_t)k * 256 + threadIdx.x;
    if (i < n) s += in[i];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) tile[blockIdx.x] = part[0];
}

__global__ void __launch_bounds__(1024) k_scan_tile_sums(uint32_t *__restrict__ tile, int64_t nt)
{
  __shared__ uint32_t part[1024];
  const int64_t per = (nt + 1023) / 1024;
  const int64_t b = (int64_t)threadIdx.x * per;
  uint32_t s = 0;
  for (int64_t i = b; i < b + per && i < nt; i++) s += tile[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const uint32_t v = (int)threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = part[threadIdx.x] - s;   // exclusive
  for (int64_t i = b; i < b + per && i < nt; i++) {
    const uint32_t t = tile[i];
    tile[i] = run;
    run += t;
  }
}

__global__ void __launch_bounds__(256) k_scan_apply(const uint32_t *__restrict__ in, int64_t n,
                                                    const uint32_t *__restrict__ tile, uint32_t *__restrict__ out)
{
  __shared__ uint32_t part[256];
  constexpr int PER = kScanTile / 256;
  const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * PER;
  uint32_t v[PER], s = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    v[k] = b + k < n ? in[b + k] : 0u;
    s += v[k];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const uint32_t t = (int)threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
    __syncthreads();
    part[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = tile[blockIdx.x] + part[threadIdx.x] - s;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    if (b + k <= n) out[b + k] = run;     // out[n] = total
    run += v[k];
  }
}

/* Slot of each particle in its cell's list; the counts return to zero. */
__global__ void k_p2g_place(const uint32_t *__restrict__ key, int64_t n, const uint32_t *__restrict__ start,
                            uint32_t *__restrict__ count, uint32_t *__restrict__ idx)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = key[i];
  if (k == kNoCell) return;
  const uint32_t old = atomicSub(&count[k], 1u);
  idx[start[k] + old - 1u] = (uint32_t)i;
}

/* Each cell's particle list in ascending particle index (insertion sort). */
__global__ void k_p2g_sort_cells(const uint32_t *__restrict__ start, int64_t nCells, uint32_t *__restrict__ idx)
{
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nCells) return;
  const uint32_t s = start[c], e = start[c + 1];
  for (uint32_t i = s + 1; i < e; i++) {
    const uint32_t v = idx[i];
    uint32_t j = i;
    while (j > s && idx[j - 1] > v) {
      idx[j] = idx[j - 1];
      j--;
    }
    idx[j] = v;
  }
}

/* Particles in bin order as SoA arrays (position, velocity, phase), so a
 * bin's particles are contiguous for the gather. */
__global__ void k_p2g_permute(const msbgx_particle *__restrict__ P, const uint32_t *__restrict__ idx, int64_t total,
                              float *__restrict__ sx, float *__restrict__ sy, float *__restrict__ sz,
                              float *__restrict__ su, float *__restrict__ sv, float *__restrict__ sw,
                              uint8_t *__restrict__ sph)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const msbgx_particle p = P[idx[k]];
  sx[k] = p.pos[0]; sy[k] = p.pos[1]; sz[k] = p.pos[2];
  su[k] = p.vel[0]; sv[k] = p.vel[1]; sw[k] = p.vel[2];
  sph[k] = (uint8_t)(p.phase == 0 ? 0 : 1);
}

struct P2gArgs {
  const float *sp[6];       // bin-ordered positions and velocities
  const uint8_t *sph;       // bin-ordered phases
  const uint32_t *start;
  const int *slot;
  int nx, ny, nz, nbx, nby, lb;
  float rhoL, rhoG, massEps, betaMin, rhoTildeMin, phiDenom;
};

/* Phase of cell (x,y,z) for the interface clamp: -1 outside, in a block
 * without level-0 data, or with mass <= massEps; 0 liquid; 1 air. */
__device__ __forceinline__ int p2gCellPhase(const P2gArgs &a, int x, int y, int z)
{
  if (x < 0 || y < 0 || z < 0 || x >= a.nx || y >= a.ny || z >= a.nz) return -1;
  const int bid = (x >> a.lb) + a.nbx * ((y >> a.lb) + a.nby * (z >> a.lb));
  if (a.slot[bid] < 0) return -1;
  const int64_t c = p2gBin(x, y, z, a.nx, a.ny);
  const uint32_t s = a.start[c], e = a.start[c + 1];
  float mass = 0.0f;
  int liq = 0;
  for (uint32_t k = s; k < e; k++) {
    mass += 1.0f;
    if (a.sph[k] == 0) liq++;
  }
  if (!(mass > a.massEps)) return -1;
  return liq * 2 > (int)(mass + 0.5f) ? 0 : 1;
}

__device__ __forceinline__ float p2gBeta(const P2gArgs &a, float rt, int pA, int pB)
{
  float phi = 0.0f;
  if (!(rt < a.rhoTildeMin)) phi = fminf(sqrtf(fmaxf(rt - a.rhoTildeMin, 0.0f) / a.phiDenom), 1.0f);
  if (pA == 0 && pB == 0) phi = 1.0f;
  else if (pA == 1 && pB == 1) phi = 0.0f;
  const float rho = a.rhoG + phi * (a.rhoL - a.rhoG);
  return fmaxf(1.0f / rho, a.betaMin);
}

/* Grid (level-0 blocks, 256-cell slices of a block), one thread per cell:
 * cell mass and liquid count, and per face the gathered sums over the 18
 * source cells.  Each cell's work is a few hundred particle visits, so every
 * cell gets its own thread. */
__global__ void __launch_bounds__(256) k_p2g_gather(P2gArgs a, const int *__restrict__ blocks, int B,
                                                    float *__restrict__ mass, int *__restrict__ liqOut,
                                                    float *__restrict__ U, float *__restrict__ V, float *__restrict__ W,
                                                    float *__restrict__ BU, float *__restrict__ BV,
                                                    float *__restrict__ BW, int *__restrict__ active)
{
  const int li = blockIdx.x, bid = blocks[li];
  const int bx = bid % a.nbx, by = (bid / a.nbx) % a.nby, bz = bid / (a.nbx * a.nby);
  const int n = B * B * B;
  const int64_t o = (int64_t)li * n;
  float *vel[3] = {U, V, W}, *bet[3] = {BU, BV, BW};
  int any = 0;
  for (int v = blockIdx.y * blockDim.x + threadIdx.x; v < n; v += gridDim.y * blockDim.x) {
    const int f[3] = {(bx << a.lb) + (v & (B - 1)), (by << a.lb) + ((v >> a.lb) & (B - 1)), (bz << a.lb) + (v >> (2 * a.lb))};
    /* the cell itself */
    {
      const int64_t c = p2gBin(f[0], f[1], f[2], a.nx, a.ny);
      const uint32_t s = a.start[c], e = a.start[c + 1];
      float m = 0.0f;
      int liq = 0;
      for (uint32_t k = s; k < e; k++) {
        m += 1.0f;
        if (a.sph[k] == 0) liq++;
      }
      mass[o + v] = m;
      liqOut[o + v] = liq;
      if (m > a.massEps) any = 1;
    }
    const int phC = p2gCellPhase(a, f[0], f[1], f[2]);
    /* one pass over the 27 cells around the cell in (z, y, x) order: each
     * particle adds to every face whose stencil it reaches, so each face's
     * sums run over its own 18 source cells in (z, y, x) order, then by
     * particle, as three separate 18-cell passes would.  Bins are x-fastest,
     * so the three cells of one (z, y) row hold one contiguous particle range,
     * visited as one loop in the same order. */
    float sv[3] = {0.0f, 0.0f, 0.0f}, sw[3] = {0.0f, 0.0f, 0.0f}, sr[3] = {0.0f, 0.0f, 0.0f};
    for (int cz = f[2] - 1; cz <= f[2] + 1; cz++)
      for (int cy = f[1] - 1; cy <= f[1] + 1; cy++) {
        const int64_t c = p2gBin(f[0] - 1, cy, cz, a.nx, a.ny);   // f - 1 >= -1, f + 1 <= n: inside the bin raster
        const uint32_t s = a.start[c], e = a.start[c + 3];
        for (uint32_t k = s; k < e; k++) {
          const float pos[3] = {a.sp[0][k], a.sp[1][k], a.sp[2][k]};
          /* per axis: the node weights on the grid whose nodes sit on the
           * axis (the face's own axis) and on the half-shifted grids */
          float frOn[3], frOff[3];
          int ddOn[3], ddOff[3];
#pragma unroll
          for (int d = 0; d < 3; d++) {
            const int i0 = (int)floorf(pos[d]);
            frOn[d] = pos[d] - (float)i0;
            ddOn[d] = f[d] - i0;
            const float g1 = pos[d] - 0.5f;
            const int i1 = (int)floorf(g1);
            frOff[d] = g1 - (float)i1;
            ddOff[d] = f[d] - i1;
          }
          const float mp = a.sph[k] == 0 ? a.rhoL : a.rhoG;
#pragma unroll
          for (int dir = 0; dir < 3; dir++) {
            float fr[3];
            int dd[3], ok = 1;
#pragma unroll
            for (int d = 0; d < 3; d++) {
              fr[d] = d == dir ? frOn[d] : frOff[d];
              dd[d] = d == dir ? ddOn[d] : ddOff[d];
              if (dd[d] < 0 || dd[d] > 1) ok = 0;
            }
            if (!ok) continue;
            const float w = ((dd[0] ? fr[0] : 1.0f - fr[0]) * (dd[1] ? fr[1] : 1.0f - fr[1])) *
                            (dd[2] ? fr[2] : 1.0f - fr[2]);
            sv[dir] = sv[dir] + a.sp[3 + dir][k] * w;
            sw[dir] = sw[dir] + w;
            sr[dir] = sr[dir] + mp * w;
          }
        }
      }
#pragma unroll
    for (int dir = 0; dir < 3; dir++) {
      vel[dir][o + v] = sw[dir] > a.massEps ? sv[dir] / sw[dir] : 0.0f;
      const int phL = p2gCellPhase(a, f[0] - (dir == 0), f[1] - (dir == 1), f[2] - (dir == 2));
      bet[dir][o + v] = p2gBeta(a, sr[dir], phL, phC);
    }
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0 && any) atomicOr(&active[li], 1);
}

/* Particles of a scene's liquid cells (msbgx_scene_particle), cell-major in
 * the order of the list of liquid cells. */
__global__ void k_p2g_scene_particles(msbgx_scene s, const uint32_t *__restrict__ cells, int64_t nCells,
                                      msbgx_particle *__restrict__ P)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int ppc = s.particles_per_cell;
  if (t >= nCells * ppc) return;
  const uint32_t c = cells[t / ppc];
  const int p = (int)(t % ppc);
  const int x = (int)(c % (uint32_t)s.nx), y = (int)((c / (uint32_t)s.nx) % (uint32_t)s.ny),
            z = (int)(c / ((uint32_t)s.nx * (uint32_t)s.ny));
  msbgx_particle q;
  msbgx_scene_particle(&s, x, y, z, p, q.pos, q.vel);
  q.phase = 0;
  P[t] = q;
}

/* Natural-layout (level-0 rank) velocity into the MG0-layout velocity channels. */
__global__ void __launch_bounds__(256) k_p2g_to_pool(const int *__restrict__ blocks, const int64_t *__restrict__ off,
                                                     int n, const float *__restrict__ U, const float *__restrict__ V,
                                                     const float *__restrict__ W, float *__restrict__ PU,
                                                     float *__restrict__ PV, float *__restrict__ PW)
{
  const int li = blockIdx.x, bid = blocks[li];
  const int64_t o = off[bid], s = (int64_t)li * n;
  for (int v = threadIdx.x; v < n; v += blockDim.x) {
    PU[o + v] = U[s + v];
    PV[o + v] = V[s + v];
    PW[o + v] = W[s + v];
  }
}

/* ------------------------------------------------------------------ host */

#define P2GCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      setError(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr);  \
      return MSBGX_EHIP;                                                              \
    }                                                                                 \
  } while (0)

void p2gDestroy(P2gCtx *c)
{
  if (!c) return;
  auto f = [](void *p) { if (p) (void)hipFree(p); };
  f(c->dLevel0); f(c->dSlot); f(c->dPart); f(c->dKey); f(c->dIdx); f(c->dCount); f(c->dStart); f(c->dTile);
  f(c->dPh);
  for (int k = 0; k < 6; k++) f(c->dS[k]);
  f(c->dMass); f(c->dLiq); f(c->dActive);
  for (int k = 0; k < 3; k++) { f(c->dVel[k]); f(c->dBeta[k]); }
  for (int k = 0; k < 3; k++)
    if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int p2gCreate(P2gCtx **out, const Grid &g)
{
  P2gCtx *c = new P2gCtx();
  *out = c;
  c->nx = g.sx0; c->ny = g.sy0; c->nz = g.sz0; c->bsx = g.bsx0; c->lb = __builtin_ctz(g.bsx0);
  c->nbx = g.nbx; c->nby = g.nby; c->nbz = g.nbz;
  c->slot.assign(g.nBlocks, -1);
  for (int b = 0; b < g.nBlocks; b++)
    if (g.map[b] == 0) {
      c->slot[b] = (int)c->level0.size();
      c->level0.push_back(b);
    }
  c->nCells = (int64_t)(g.sx0 + 2) * (g.sy0 + 2) * (g.sz0 + 2);   // bins: domain + one-cell halo
  if (c->nCells >= (int64_t)kNoCell) {
    setError("particle-to-grid transfer: domain too large for 32-bit cell indices");
    return MSBGX_EINVAL;
  }
  c->nTiles = (c->nCells + 1 + kScanTile - 1) / kScanTile;
  c->nOut = (int64_t)c->level0.size() * g.bsx0 * g.bsx0 * g.bsx0;
  P2GCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int k = 0; k < 3; k++) P2GCHK(hipEventCreate(&c->ev[k]));
  auto al = [](auto **p, size_t n) -> hipError_t { return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(**p)); };
  P2GCHK(al(&c->dLevel0, c->level0.size()));
  P2GCHK(al(&c->dSlot, c->slot.size()));
  P2GCHK(al(&c->dCount, (size_t)c->nCells));
  P2GCHK(al(&c->dStart, (size_t)c->nCells + 1));
  P2GCHK(al(&c->dTile, (size_t)c->nTiles));
  P2GCHK(al(&c->dMass, (size_t)c->nOut));
  P2GCHK(al(&c->dLiq, (size_t)c->nOut));
  P2GCHK(al(&c->dActive, c->level0.size()));
  for (int k = 0; k < 3; k++) {
    P2GCHK(al(&c->dVel[k], (size_t)c->nOut));
    P2GCHK(al(&c->dBeta[k], (size_t)c->nOut));
  }
  if (!c->level0.empty())
    P2GCHK(hipMemcpy(c->dLevel0, c->level0.data(), c->level0.size() * sizeof(int), hipMemcpyHostToDevice));
  P2GCHK(hipMemcpy(c->dSlot, c->slot.data(), c->slot.size() * sizeof(int), hipMemcpyHostToDevice));
  P2GCHK(hipMemset(c->dCount, 0, (size_t)c->nCells * sizeof(uint32_t)));
  return MSBGX_OK;
}

static int p2gReserve(P2gCtx &c, int64_t n)
{
  if (n <= c.capP) return MSBGX_OK;
  auto f = [](void *p) { if (p) (void)hipFree(p); };
  f(c.dPart); f(c.dKey); f(c.dIdx); f(c.dPh);
  for (int k = 0; k < 6; k++) { f(c.dS[k]); c.dS[k] = nullptr; }
  c.dPart = nullptr; c.dKey = nullptr; c.dIdx = nullptr; c.dPh = nullptr; c.capP = 0;
  P2GCHK(hipMalloc((void **)&c.dPart, (size_t)n * sizeof(msbgx_particle)));
  P2GCHK(hipMalloc((void **)&c.dKey, (size_t)n * sizeof(uint32_t)));
  P2GCHK(hipMalloc((void **)&c.dIdx, (size_t)n * sizeof(uint32_t)));
  for (int k = 0; k < 6; k++) P2GCHK(hipMalloc((void **)&c.dS[k], (size_t)n * sizeof(float)));
  P2GCHK(hipMalloc((void **)&c.dPh, (size_t)n));
  c.capP = n;
  return MSBGX_OK;
}

/* Binning and gather over the n particles already in c.dPart. */
static int p2gRunDevice(P2gCtx &c, int64_t n, const msbgx_p2g_params &p)
{
  if (n >= (int64_t)kNoCell) {
    setError("particle-to-grid transfer: too many particles for 32-bit indices");
    return MSBGX_EINVAL;
  }
  hipStream_t s = c.stream;
  P2GCHK(hipEventRecord(c.ev[0], s));
  /* the placement returns the counts to zero; clearing them here keeps a run
   * that stopped half-way from leaving stale counts behind */
  P2GCHK(hipMemsetAsync(c.dCount, 0, (size_t)c.nCells * sizeof(uint32_t), s));
  const int64_t gP = (n + 255) / 256;
  if (n > 0) hipLaunchKernelGGL(k_p2g_key, dim3((unsigned)gP), dim3(256), 0, s, c.dPart, n, c.nx, c.ny, c.nz, c.dKey, c.dCount);
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)c.nTiles), dim3(256), 0, s, c.dCount, c.nCells, c.dTile);
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(1024), 0, s, c.dTile, c.nTiles);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)c.nTiles), dim3(256), 0, s, c.dCount, c.nCells, c.dTile, c.dStart);
  if (n > 0) hipLaunchKernelGGL(k_p2g_place, dim3((unsigned)gP), dim3(256), 0, s, c.dKey, n, c.dStart, c.dCount, c.dIdx);
  hipLaunchKernelGGL(k_p2g_sort_cells, dim3((unsigned)((c.nCells + 255) / 256)), dim3(256), 0, s, c.dStart, c.nCells,
                     c.dIdx);
  /* binned particles: every particle except those beyond the halo, counted by the scan */
  uint32_t total = 0;
  P2GCHK(hipMemcpyAsync(&total, c.dStart + c.nCells, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  P2GCHK(hipStreamSynchronize(s));
  if (total > 0)
    hipLaunchKernelGGL(k_p2g_permute, dim3((unsigned)(((int64_t)total + 255) / 256)), dim3(256), 0, s, c.dPart, c.dIdx,
                       (int64_t)total, c.dS[0], c.dS[1], c.dS[2], c.dS[3], c.dS[4], c.dS[5], c.dPh);
  P2GCHK(hipGetLastError());
  P2GCHK(hipEventRecord(c.ev[1], s));
  P2gArgs a;
  for (int k = 0; k < 6; k++) a.sp[k] = c.dS[k];
  a.sph = c.dPh; a.start = c.dStart; a.slot = c.dSlot;
  a.nx = c.nx; a.ny = c.ny; a.nz = c.nz; a.nbx = c.nbx; a.nby = c.nby; a.lb = c.lb;
  a.rhoL = p.rho_l; a.rhoG = p.rho_g; a.massEps = p.mass_eps; a.betaMin = p.beta_min;
  /* the reference's phi parameters (src/flip_sim.cpp:744-748), same float operations */
  const float rhoTilde0 = (float)p.particles_per_voxel;
  const float etaPhi = logf(p.rho_l / p.rho_g);
  a.rhoTildeMin = etaPhi * p.rho_g * rhoTilde0;
  a.phiDenom = p.alpha_phi * rhoTilde0 * p.rho_l;
  if (!c.level0.empty()) {
    P2GCHK(hipMemsetAsync(c.dActive, 0, c.level0.size() * sizeof(int), s));
    const int slices = std::max(1, c.bsx * c.bsx * c.bsx / 256);
    hipLaunchKernelGGL(k_p2g_gather, dim3((unsigned)c.level0.size(), (unsigned)slices), dim3(256), 0, s, a, c.dLevel0,
                       c.bsx, c.dMass, c.dLiq, c.dVel[0], c.dVel[1], c.dVel[2], c.dBeta[0], c.dBeta[1], c.dBeta[2],
                       c.dActive);
  }
  P2GCHK(hipGetLastError());
  P2GCHK(hipEventRecord(c.ev[2], s));
  P2GCHK(hipEventSynchronize(c.ev[2]));
  float t0 = 0.f, t1 = 0.f;
  P2GCHK(hipEventElapsedTime(&t0, c.ev[0], c.ev[2]));
  P2GCHK(hipEventElapsedTime(&t1, c.ev[1], c.ev[2]));
  c.msTotal = t0;
  c.msGather = t1;
  c.lastParticles = n;
  c.valid = true;
  return MSBGX_OK;
}

static int p2gCheckParams(const msbgx_p2g_params &p)
{
  if (!(p.rho_l > 0.0f) || !(p.rho_g > 0.0f) || !std::isfinite(p.rho_l) || !std::isfinite(p.rho_g) ||
      !std::isfinite(p.mass_eps) || !std::isfinite(p.beta_min) || !(p.alpha_phi > 0.0f) ||
      !std::isfinite(p.alpha_phi) || p.particles_per_voxel < 1) {
    setError("invalid particle-to-grid parameters");
    return MSBGX_EINVAL;
  }
  return MSBGX_OK;
}

int p2gRun(P2gCtx &c, int64_t n, const msbgx_particle *host, const msbgx_p2g_params &p)
{
  if (int rc = p2gCheckParams(p)) return rc;
  /* floorf of the position must fit an int: finite and |x| < 2^30 */
  int bad = 0;
#pragma omp parallel for reduction(| : bad) schedule(static)
  for (int64_t i = 0; i < n; i++)
    for (int d = 0; d < 3; d++)
      if (!(std::fabs(host[i].pos[d]) < 1073741824.0f)) bad = 1;
  if (bad) {
    setError("particle positions must be finite and below 2^30 in magnitude");
    return MSBGX_EINVAL;
  }
  if (int rc = p2gReserve(c, n)) return rc;
  if (n > 0) P2GCHK(hipMemcpy(c.dPart, host, (size_t)n * sizeof(msbgx_particle), hipMemcpyHostToDevice));
  return p2gRunDevice(c, n, p);
}

int p2gRunScene(P2gCtx &c, const msbgx_scene &s, const msbgx_p2g_params &p, int64_t *nOut)
{
  if (int rc = p2gCheckParams(p)) return rc;
  if (s.nx != c.nx || s.ny != c.ny || s.nz != c.nz || s.particles_per_cell < 1) {
    setError("scene must match the grid and carry particles");
    return MSBGX_EINVAL;
  }
  /* liquid cells, x fastest (host list, per z-slice in parallel; the
   * particles themselves are generated on the GPU) */
  std::vector<std::vector<uint32_t>> slab(s.nz);
#pragma omp parallel for schedule(dynamic, 4)
  for (int z = 0; z < s.nz; z++)
    for (int y = 0; y < s.ny; y++)
      for (int x = 0; x < s.nx; x++)
        if (msbgx_scene_is_liquid(&s, x, y, z)) slab[z].push_back((uint32_t)(((int64_t)z * s.ny + y) * s.nx + x));
  std::vector<uint32_t> cells;
  for (auto &v : slab) cells.insert(cells.end(), v.begin(), v.end());
  const int64_t n = (int64_t)cells.size() * s.particles_per_cell;
  if (int rc = p2gReserve(c, n)) return rc;
  *nOut = n;
  if (n > 0) {
    /* the cell list goes through the index buffer (overwritten by the binning) */
    P2GCHK(hipMemcpy(c.dIdx, cells.data(), cells.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_p2g_scene_particles, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, s, c.dIdx,
                       (int64_t)cells.size(), c.dPart);
    P2GCHK(hipGetLastError());
  }
  return p2gRunDevice(c, n, p);
}

static void p2gToDense(const P2gCtx &c, const float *nat, float *dense)
{
  const int B = c.bsx;
  std::fill(dense, dense + (size_t)c.nx * c.ny * c.nz, 0.0f);
#pragma omp parallel for schedule(static)
  for (int li = 0; li < (int)c.level0.size(); li++) {
    const int bid = c.level0[li];
    const int bx = bid % c.nbx, by = (bid / c.nbx) % c.nby, bz = bid / (c.nbx * c.nby);
    const float *src = nat + (size_t)li * B * B * B;
    for (int vz = 0; vz < B; vz++)
      for (int vy = 0; vy < B; vy++) {
        float *d = dense + ((size_t)(bz * B + vz) * c.ny + (size_t)(by * B + vy)) * c.nx + (size_t)bx * B;
        std::copy(src + (size_t)B * (vy + (size_t)B * vz), src + (size_t)B * (vy + (size_t)B * vz) + B, d);
      }
  }
}

int p2gGet(P2gCtx &c, float *mass, float *const vel[3], float *const beta[3])
{
  if (!c.valid) {
    setError("particle-to-grid transfer has not run");
    return MSBGX_ESTATE;
  }
  std::vector<float> tmp((size_t)c.nOut);
  auto one = [&](const float *dsrc, float *dst) -> int {
    if (!dst) return MSBGX_OK;
    P2GCHK(hipMemcpy(tmp.data(), dsrc, tmp.size() * sizeof(float), hipMemcpyDeviceToHost));
    p2gToDense(c, tmp.data(), dst);
    return MSBGX_OK;
  };
  if (int rc = one(c.dMass, mass)) return rc;
  for (int k = 0; k < 3; k++) {
    if (int rc = one(c.dVel[k], vel[k])) return rc;
    if (int rc = one(c.dBeta[k], beta[k])) return rc;
  }
  return MSBGX_OK;
}

int p2gActiveBlocks(P2gCtx &c, int *out, int cap)
{
  if (!c.valid) {
    setError("particle-to-grid transfer has not run");
    return MSBGX_ESTATE;
  }
  std::vector<int> a(c.level0.size());
  if (!a.empty()) P2GCHK(hipMemcpy(a.data(), c.dActive, a.size() * sizeof(int), hipMemcpyDeviceToHost));
  int k = 0;
  for (size_t i = 0; i < a.size(); i++)
    if (a[i]) {
      if (out && k < cap) out[k] = c.level0[i];
      k++;
    }
  return k;
}

void p2gTimes(const P2gCtx &c, double *msTotal, double *msGather, int64_t *particles)
{
  if (msTotal) *msTotal = c.msTotal;
  if (msGather) *msGather = c.msGather;
  if (particles) *particles = c.lastParticles;
}

/* Velocity of the last transfer into a prepared grid's FLIP velocity
 * channels; the grid must have the same raster and level map. */
int p2gToVelocity(P2gCtx &c, Grid &g)
{
  if (!c.valid) {
    setError("particle-to-grid transfer has not run");
    return MSBGX_ESTATE;
  }
  if (int rc = flipCheck(g)) return rc;
  if (g.sx0 != c.nx || g.sy0 != c.ny || g.sz0 != c.nz || g.bsx0 != c.bsx) {
    setError("grid does not match the transfer's raster");
    return MSBGX_EINVAL;
  }
  for (int b = 0; b < g.nBlocks; b++)
    if ((g.map[b] == 0) != (c.slot[b] >= 0)) {
      setError("grid level map differs from the transfer's");
      return MSBGX_EINVAL;
    }
  if (int rc = flipEnsure(g)) return rc;
  Dev &d = dev(g);
  P2GCHK(hipStreamSynchronize(c.stream));
  if (!c.level0.empty())
    hipLaunchKernelGGL(k_p2g_to_pool, dim3((unsigned)c.level0.size()), dim3(256), 0, d.stream, c.dLevel0,
                       g.mg[0].d_off, c.bsx * c.bsx * c.bsx, c.dVel[0], c.dVel[1], c.dVel[2], d.vel[0], d.vel[1],
                       d.vel[2]);
  P2GCHK(hipGetLastError());
  P2GCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

}  // namespace msbgx
