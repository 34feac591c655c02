/*
 * pde.hip -- the msbg_demo solver entry on the GPU (compiled as part of
 * solver.hip): mean-curvature smoothing of a 16-bit render-density channel
 * with the 8-colour in-place block scheme.
 *
 * Reference: MultiresSparseGrid::applyChannelPdeFast<RenderDensity>
 * (src/msbg4.cpp:60-960) as src/msbg_demo.cpp:716-727 calls it
 * (PDE_MEAN_CURVATURE + OPT_8_COLOR_SCHEME, doConstrZeroOne, no density
 * constraints, grid spacing 1), with its halo blocks
 * (HaloBlockSet::fillHaloBlock_ with OPT_BC_NEUMANN, src/halo.cpp:267-480,
 * getOutOfBlockValue src/halo.h:267-290); RenderDensity is uint16
 * (RSURF_16_BIT, src/msbgcell.h:68):
 *   iterations: numIter rounded up to even (src/msbg4.cpp:156-157);
 *   per iteration, colours 0..7 (colour = (bx&1) + 2(by&1) + 4(bz&1),
 *   src/msbg.h:642-646), per listed level-0 block of that colour:
 *     halo (one voxel) as float v / 65535: a voxel of a level-0 block from
 *     the channel, any other voxel (outside the domain, or in a block of
 *     another level) the block's own nearest voxel (Neumann);
 *     per voxel: central differences, H = (fy^2+fz^2) fxx + (fx^2+fz^2) fyy +
 *     (fx^2+fy^2) fzz - 2 (fx fy fyx + fx fz fzx + fy fz fzy), divided by
 *     |grad|^2 when that exceeds 1e-7 (else 0), D = clamp(dt H, +-0.1),
 *     phi = clamp(phi + D, 0, 1), stored as round-to-nearest-even(phi 65535).
 * Same-colour blocks are never neighbours (not even diagonally), so one
 * colour's blocks update in place from halos that no launch of that colour
 * writes.
 *
 * Data: the channel over the level-0 blocks in ascending block order, B^3
 * uint16 per block (a level-0 block the caller leaves empty reads as its
 * values); one workgroup per block stages the (B+2)^3 halo in LDS (B <= 16).
 * Algorithmic bytes per voxel and iteration: 2 read + 2 written.
 */

#include <mutex>
#include <string>

namespace msbgx {

struct PdeArgs {
  const int *slot;   // per block: rank among the level-0 blocks, -1 otherwise
  int nbx, nby, nbz, lb;
  int n[3];
  float dt;
};

template <int B>
__global__ void __launch_bounds__(256) k_pde_mc(PdeArgs a, const int *__restrict__ blocks, uint16_t *__restrict__ rho)
{
  constexpr int H = B + 2, H2 = H * H, N = B * B * B;
  __shared__ float halo[H * H * H];
  __shared__ int nslot[27];   // slots of the 3x3x3 block neighbourhood, -1: none or outside
  const int bid = blocks[blockIdx.x];
  const int bx = bid % a.nbx, by = (bid / a.nbx) % a.nby, bz = bid / (a.nbx * a.nby);
  const int64_t own = (int64_t)a.slot[bid] * N;
  const int g0[3] = {bx * B, by * B, bz * B};
  const float inv = 1.f / 65535;   // 1.f / numeric_limits<uint16_t>::max()
  /* the 27 neighbour slots once, so that each halo voxel costs one load */
  if (threadIdx.x < 27) {
    const int nx = bx + (int)threadIdx.x % 3 - 1, ny = by + ((int)threadIdx.x / 3) % 3 - 1,
              nz = bz + (int)threadIdx.x / 9 - 1;
    const bool in = nx >= 0 && ny >= 0 && nz >= 0 && nx < a.nbx && ny < a.nby && nz < a.nbz;
    nslot[threadIdx.x] = in ? a.slot[nx + a.nbx * (ny + a.nby * nz)] : -1;
  }
  __syncthreads();
  for (int h = threadIdx.x; h < H * H * H; h += blockDim.x) {
    const int hx = h % H, hy = (h / H) % H, hz = h / H2;
    const int g[3] = {g0[0] + hx - 1, g0[1] + hy - 1, g0[2] + hz - 1};
    int64_t src = -1;
    if (g[0] >= 0 && g[1] >= 0 && g[2] >= 0 && g[0] < a.n[0] && g[1] < a.n[1] && g[2] < a.n[2]) {
      /* hx - 1 in [-1, B]: neighbour index 0, 1 or 2 along each axis */
      const int s = nslot[(hx + B - 1) / B + 3 * ((hy + B - 1) / B) + 9 * ((hz + B - 1) / B)];
      if (s >= 0) src = (int64_t)s * N + (g[0] & (B - 1)) + B * ((g[1] & (B - 1)) + B * (g[2] & (B - 1)));
    }
    if (src < 0) {   // Neumann: the block's own nearest voxel
      const int cx = min(max(hx - 1, 0), B - 1), cy = min(max(hy - 1, 0), B - 1), cz = min(max(hz - 1, 0), B - 1);
      src = own + cx + B * (cy + B * cz);
    }
    halo[h] = (float)rho[src] * inv;
  }
  __syncthreads();
  const float dt = a.dt;
  for (int v = threadIdx.x; v < N; v += blockDim.x) {
    const int vx = v % B, vy = (v / B) % B, vz = v / (B * B);
    const float *ph = halo + (vx + 1) + H * (vy + 1) + H2 * (vz + 1);
    auto F = [&](int dx, int dy, int dz) { return ph[dx + H * dy + H2 * dz]; };
    const float f0 = F(0, 0, 0), f1 = F(-1, 0, 0), f2 = F(1, 0, 0), f3 = F(0, -1, 0), f4 = F(0, 1, 0),
                f5 = F(0, 0, -1), f6 = F(0, 0, 1);
    const float fx = 0.5f * (f2 - f1), fy = 0.5f * (f4 - f3), fz = 0.5f * (f6 - f5);
    const float fxx = f1 + f2 - 2 * f0, fyy = f3 + f4 - 2 * f0, fzz = f5 + f6 - 2 * f0;
    const float fy_right = (F(1, 1, 0) - F(1, -1, 0)) * 0.5f, fy_left = (F(-1, 1, 0) - F(-1, -1, 0)) * 0.5f,
                fz_right = (F(1, 0, 1) - F(1, 0, -1)) * 0.5f, fz_left = (F(-1, 0, 1) - F(-1, 0, -1)) * 0.5f,
                fz_up = (F(0, 1, 1) - F(0, 1, -1)) * 0.5f, fz_down = (F(0, -1, 1) - F(0, -1, -1)) * 0.5f,
                fyx = (fy_right - fy_left) * 0.5f, fzx = (fz_right - fz_left) * 0.5f, fzy = (fz_up - fz_down) * 0.5f;
    float Hc = (fy * fy + fz * fz) * fxx + (fx * fx + fz * fz) * fyy + (fx * fx + fy * fy) * fzz -
               2.0f * (fx * fy * fyx + fx * fz * fzx + fy * fz * fzy);
    const float gradMagSq = fx * fx + fy * fy + fz * fz;
    Hc = gradMagSq > 1e-7f ? Hc / gradMagSq : 0.0f;
    float D = dt * Hc;
    D = fminf(fmaxf(D, -0.1f), 0.1f);
    float phi = f0 + D;
    phi = fminf(fmaxf(phi, 0.0f), 1.0f);
    rho[own + v] = (uint16_t)(int)rintf(phi * 65535);
  }
}

/* nIter iterations of the 8 colour passes: per colour one launch over that
 * colour's listed level-0 blocks (lists[start[c] .. start[c+1])). */
static void pdeColourPasses(int B, const PdeArgs &a, const int *lists, const int start[9], int nIter, uint16_t *rho,
                            hipStream_t s)
{
  for (int it = 0; it < nIter; it++)
    for (int c = 0; c < 8; c++) {
      const int cnt = start[c + 1] - start[c];
      if (cnt == 0) continue;
      if (B == 16) hipLaunchKernelGGL(k_pde_mc<16>, dim3(cnt), dim3(256), 0, s, a, lists + start[c], rho);
      else if (B == 8) hipLaunchKernelGGL(k_pde_mc<8>, dim3(cnt), dim3(256), 0, s, a, lists + start[c], rho);
      else if (B == 4) hipLaunchKernelGGL(k_pde_mc<4>, dim3(cnt), dim3(256), 0, s, a, lists + start[c], rho);
      else hipLaunchKernelGGL(k_pde_mc<2>, dim3(cnt), dim3(256), 0, s, a, lists + start[c], rho);
    }
}

/* Host: upload the dense channel, run the colour passes, download. */
int devicePdeMeanCurvature(const Grid &g, uint16_t *dense, const int *active, int nActive, int numIter, float dt,
                           double *ms)
{
  const int B = g.bsx0, lb = __builtin_ctz(B);
  if (B > 16) {
    setError("mean-curvature smoothing: block edge above 16 is not supported (the halo block is staged in LDS)");
    return MSBGX_EINVAL;
  }
  const MgLevel &m = g.mg[0];
  std::vector<int> slot(g.nBlocks, -1), level0;
  for (int b = 0; b < g.nBlocks; b++)
    if (g.map[b] == 0) {
      slot[b] = (int)level0.size();
      level0.push_back(b);
    }
  (void)m;
  /* colour lists of the listed level-0 blocks, in list order */
  std::vector<int> lists[8];
  for (int i = 0; i < nActive; i++) {
    const int b = active[i];
    if (b < 0 || b >= g.nBlocks) {
      setError("mean-curvature smoothing: active block id out of range");
      return MSBGX_EINVAL;
    }
    if (slot[b] < 0) continue;   // only level-0 blocks are smoothed (src/msbg4.cpp:447)
    int bx, by, bz;
    g.blockCoords(b, bx, by, bz);
    lists[(bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2)].push_back(b);
  }
  const int nIter = (numIter % 2 == 1) ? numIter + 1 : numIter;
  const size_t nv = level0.size() * (size_t)B * B * B;
  hipStream_t s;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct Res {
    hipStream_t s;
    std::vector<void *> p;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Res()
    {
      for (void *q : p) (void)hipFree(q);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      (void)hipStreamDestroy(s);
    }
  } res{s, {}};
  uint16_t *dRho = nullptr;
  int *dSlot = nullptr, *dLists = nullptr;
  HIPCHK(hipMalloc((void **)&dRho, std::max<size_t>(nv, 1) * sizeof(uint16_t)));
  res.p.push_back(dRho);
  HIPCHK(hipMalloc((void **)&dSlot, slot.size() * sizeof(int)));
  res.p.push_back(dSlot);
  size_t nl = 0;
  for (auto &l : lists) nl += l.size();
  HIPCHK(hipMalloc((void **)&dLists, std::max<size_t>(nl, 1) * sizeof(int)));
  res.p.push_back(dLists);
  HIPCHK(hipEventCreate(&res.e0));
  HIPCHK(hipEventCreate(&res.e1));
  /* dense -> natural level-0 layout */
  std::vector<uint16_t> nat(nv);
#pragma omp parallel for schedule(static)
  for (int li = 0; li < (int)level0.size(); li++) {
    int bx, by, bz;
    g.blockCoords(level0[li], bx, by, bz);
    for (int vz = 0; vz < B; vz++)
      for (int vy = 0; vy < B; vy++) {
        const uint16_t *src = dense + ((size_t)(bz * B + vz) * g.sy0 + (size_t)(by * B + vy)) * g.sx0 + (size_t)bx * B;
        std::copy(src, src + B, nat.data() + (size_t)li * B * B * B + (size_t)B * (vy + (size_t)B * vz));
      }
  }
  HIPCHK(hipMemcpy(dRho, nat.data(), nv * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dSlot, slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice));
  std::vector<int> all;
  int start[9] = {0};
  for (int c = 0; c < 8; c++) {
    start[c] = (int)all.size();
    all.insert(all.end(), lists[c].begin(), lists[c].end());
  }
  start[8] = (int)all.size();
  if (!all.empty()) HIPCHK(hipMemcpy(dLists, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
  PdeArgs a;
  a.slot = dSlot; a.nbx = g.nbx; a.nby = g.nby; a.nbz = g.nbz; a.lb = lb;
  a.n[0] = g.sx0; a.n[1] = g.sy0; a.n[2] = g.sz0; a.dt = dt;
  HIPCHK(hipEventRecord(res.e0, s));
  pdeColourPasses(B, a, dLists, start, nIter, dRho, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(res.e1, s));
  HIPCHK(hipEventSynchronize(res.e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, res.e0, res.e1));
  if (ms) *ms = t;
  HIPCHK(hipMemcpy(nat.data(), dRho, nv * sizeof(uint16_t), hipMemcpyDeviceToHost));
#pragma omp parallel for schedule(static)
  for (int li = 0; li < (int)level0.size(); li++) {
    int bx, by, bz;
    g.blockCoords(level0[li], bx, by, bz);
    for (int vz = 0; vz < B; vz++)
      for (int vy = 0; vy < B; vy++) {
        uint16_t *dst = dense + ((size_t)(bz * B + vz) * g.sy0 + (size_t)(by * B + vy)) * g.sx0 + (size_t)bx * B;
        const uint16_t *src = nat.data() + (size_t)li * B * B * B + (size_t)B * (vy + (size_t)B * vz);
        std::copy(src, src + B, dst);
      }
  }
  return MSBGX_OK;
}

/* ------------------------------------------------- allocation-free variant */

/* Work space of devicePdeMeanCurvatureNoAlloc, one copy per GPU, held until
 * the process ends.  In host code `scratch` names a host-side stand-in, not
 * device memory: host code only ever uses the per-device address that
 * hipGetSymbolAddress returns (scratchBase). */
static __device__ __align__(256) unsigned char scratch[MSBGX_PDE_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratchMu;
static unsigned char *g_scratchBase[kMaxScratchDevices] = {};

/* The current device's copy of `scratch`, looked up once per device with
 * that device current. */
static int scratchBase(unsigned char **out)
{
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices) {
    setError("device index beyond the scratch table");
    return MSBGX_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (!g_scratchBase[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      setError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return MSBGX_EHIP;
    }
    g_scratchBase[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratchBase[dev];
  return MSBGX_OK;
}

struct PdeStarts {
  int start[9];   // colour c's blocks are lists[start[c] .. start[c+1])
};

/* Exclusive prefix sum over the 1024 threads of a workgroup (tmp: 1024 ints). */
__device__ int blockExclusiveSum1024(int v, int *tmp)
{
  const int t = threadIdx.x;
  tmp[t] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int x = t >= o ? tmp[t - o] : 0;
    __syncthreads();
    tmp[t] += x;
    __syncthreads();
  }
  const int incl = tmp[t];
  __syncthreads();
  return incl - v;
}

/* One workgroup of 1024 threads builds the tables the host path builds on
 * the host: slot[b] = rank of b among the level-0 blocks of the map (-1
 * otherwise), level0[rank] = b, and the listed level-0 blocks per colour in
 * list order.  Thread t takes a contiguous run of blocks (and of list
 * entries), so ranks follow block order and the colour lists follow list
 * order.  Every entry of slot[0, nBlocks), level0[0, #level-0) and
 * lists[0, start[8]) is written, so nothing left from an earlier call is read. */
__global__ void __launch_bounds__(1024) k_pde_na_tables(const int *__restrict__ map, int nBlocks,
                                                        const int *__restrict__ active, int nActive, int nbx, int nby,
                                                        PdeStarts st, int *__restrict__ slot, int *__restrict__ level0,
                                                        int *__restrict__ lists)
{
  __shared__ int tmp[1024];
  const int t = threadIdx.x;
  const int kb = (nBlocks + 1023) / 1024;
  const int b0 = min(t * kb, nBlocks), b1 = min(b0 + kb, nBlocks);
  int cnt = 0;
  for (int b = b0; b < b1; b++) cnt += map[b] == 0;
  int r = blockExclusiveSum1024(cnt, tmp);
  for (int b = b0; b < b1; b++) {
    if (map[b] == 0) {
      slot[b] = r;
      level0[r] = b;
      r++;
    } else {
      slot[b] = -1;
    }
  }
  auto colourOf = [&](int b) {
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    return (bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2);
  };
  const int ka = (nActive + 1023) / 1024;
  const int a0 = min(t * ka, nActive), a1 = min(a0 + ka, nActive);
  int cc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = a0; i < a1; i++)
    if (map[active[i]] == 0) cc[colourOf(active[i])]++;
  int pos[8];
  for (int c = 0; c < 8; c++) pos[c] = st.start[c] + blockExclusiveSum1024(cc[c], tmp);
  for (int i = a0; i < a1; i++) {
    const int b = active[i];
    if (map[b] == 0) lists[pos[colourOf(b)]++] = b;
  }
}

/* The level-0 blocks between the dense channel (x fastest over the domain)
 * and the slot layout the smoothing kernels use: TO_DENSE false gathers,
 * true scatters back.  One workgroup per level-0 block. */
template <bool TO_DENSE>
__global__ void __launch_bounds__(256) k_pde_na_copy(const int *__restrict__ level0, int nbx, int nby, int lb, int sx,
                                                     int sy, uint16_t *__restrict__ dense, uint16_t *__restrict__ nat)
{
  const int B = 1 << lb, N = B * B * B;
  const int b = level0[blockIdx.x];
  const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
  uint16_t *blk = nat + (int64_t)blockIdx.x * N;
  for (int v = threadIdx.x; v < N; v += blockDim.x) {
    const int vx = v & (B - 1), vy = (v >> lb) & (B - 1), vz = v >> (2 * lb);
    const int64_t di = ((int64_t)(bz * B + vz) * sy + (by * B + vy)) * sx + (int64_t)bx * B + vx;
    if (TO_DENSE) dense[di] = blk[v];
    else blk[v] = dense[di];
  }
}

/* devicePdeMeanCurvature without any device allocation: the dense channel,
 * the slot-layout copy and the tables live in `scratch`; the host keeps only
 * counts on the stack.  Same kernels, same results. */
int devicePdeMeanCurvatureNoAlloc(const Grid &g, uint16_t *dense, const int *active, int nActive, int numIter, float dt)
{
  const int B = g.bsx0, lb = __builtin_ctz(B);
  if (B > 16) {
    setError("mean-curvature smoothing: block edge above 16 is not supported (the halo block is staged in LDS)");
    return MSBGX_EINVAL;
  }
  int nLevel0 = 0;
  for (int b = 0; b < g.nBlocks; b++) nLevel0 += g.map[b] == 0;
  int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < nActive; i++) {
    const int b = active[i];
    if (b < 0 || b >= g.nBlocks) {
      setError("mean-curvature smoothing: active block id out of range");
      return MSBGX_EINVAL;
    }
    if (g.map[b] != 0) continue;   // only level-0 blocks are smoothed (src/msbg4.cpp:447)
    int bx, by, bz;
    g.blockCoords(b, bx, by, bz);
    cnt[(bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2)]++;
  }
  PdeStarts st{};
  for (int c = 0; c < 8; c++) st.start[c + 1] = st.start[c] + cnt[c];
  /* scratch layout, every part 256-byte aligned */
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t nDense = (size_t)g.sx0 * g.sy0 * g.sz0, nv = (size_t)nLevel0 * B * B * B;
  const size_t nb = (size_t)g.nBlocks, na = (size_t)nActive;
  const size_t oNat = al(2 * nDense), oMap = oNat + al(2 * nv), oSlot = oMap + al(4 * nb), oLev0 = oSlot + al(4 * nb),
               oAct = oLev0 + al(4 * nb), oLists = oAct + al(4 * na), total = oLists + al(4 * na);
  if (total > (size_t)MSBGX_PDE_NOALLOC_SCRATCH_BYTES) {
    setError("mean-curvature smoothing without allocation: needs " + std::to_string(total) +
             " bytes of scratch, the maximum is " + std::to_string((size_t)MSBGX_PDE_NOALLOC_SCRATCH_BYTES));
    return MSBGX_EINVAL;
  }
  const int nIter = (numIter % 2 == 1) ? numIter + 1 : numIter;
  unsigned char *base = nullptr;
  if (int rc = scratchBase(&base)) return rc;
  uint16_t *dDense = reinterpret_cast<uint16_t *>(base), *dNat = reinterpret_cast<uint16_t *>(base + oNat);
  int *dMap = reinterpret_cast<int *>(base + oMap), *dSlot = reinterpret_cast<int *>(base + oSlot);
  int *dLev0 = reinterpret_cast<int *>(base + oLev0), *dAct = reinterpret_cast<int *>(base + oAct);
  int *dLists = reinterpret_cast<int *>(base + oLists);
  hipStream_t s = hipStreamPerThread;
  HIPCHK(hipMemcpyAsync(dDense, dense, 2 * nDense, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dMap, g.map.data(), 4 * nb, hipMemcpyHostToDevice, s));
  if (nActive > 0) HIPCHK(hipMemcpyAsync(dAct, active, 4 * na, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_pde_na_tables, dim3(1), dim3(1024), 0, s, dMap, g.nBlocks, dAct, nActive, g.nbx, g.nby, st, dSlot,
                     dLev0, dLists);
  if (nLevel0 > 0)
    hipLaunchKernelGGL(k_pde_na_copy<false>, dim3(nLevel0), dim3(256), 0, s, dLev0, g.nbx, g.nby, lb, g.sx0, g.sy0,
                       dDense, dNat);
  PdeArgs a;
  a.slot = dSlot; a.nbx = g.nbx; a.nby = g.nby; a.nbz = g.nbz; a.lb = lb;
  a.n[0] = g.sx0; a.n[1] = g.sy0; a.n[2] = g.sz0; a.dt = dt;
  pdeColourPasses(B, a, dLists, st.start, nIter, dNat, s);
  if (nLevel0 > 0)
    hipLaunchKernelGGL(k_pde_na_copy<true>, dim3(nLevel0), dim3(256), 0, s, dLev0, g.nbx, g.nby, lb, g.sx0, g.sy0,
                       dDense, dNat);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dense, dDense, 2 * nDense, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSBGX_OK;
}

}  // namespace msbgx

