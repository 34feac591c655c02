/*
 * flip.hip -- the FLIP caller's steps either side of the pressure solve, on
 * the device (compiled as part of solver.hip).
 *
 * FlipSimulation::pressureProjection (src/flip_sim.cpp:1281-1350) samples
 * the divergence right-hand side from the MAC velocity (flipRhsCB,
 * src/flip_sim.cpp:329-338), solves, and subtracts the pressure gradient from
 * the velocity (src/flip_sim.cpp:1317-1349).  Its grids are the MSBG
 * channels at resolution level 0 (src/flip_sim.cpp:363-375): blocks of other
 * levels hold no data there, so their cells read 0 and writes to them are
 * dropped (getF/getVC/setVC, src/flip_sim.cpp:104-133).
 *
 * Here the velocity lives in three device channels laid out like the MG0
 * pools (slot off[bid] of every level-0 block; component c of cell (x,y,z)
 * is the velocity on the cell's -c face).  The fluid test isFluid(mass) and
 * the face coefficient beta are the level-0 cell type and face coefficient
 * the prepare step sampled through flipCellTypeCB / flipFaceCoeffCB: a cell
 * is fluid iff its flags hold neither SOLID nor VOID.
 *
 *   k_flip_rhs      b = isFluid ? div(v) / dt : 0 with flipRhsCB's operation
 *                   order; +faces in a neighbour block read that block's
 *                   velocity, 0 outside the domain or in a block without data
 *   k_flip_update   per face: 0 at index 0 along its axis; 0 when neither
 *                   side is fluid; else v - dt * beta * (pR - pL) with the
 *                   pressure of a non-fluid side 0 and non-finite results 0.
 *                   With `negate` the pressures enter with the opposite sign:
 *                   the MSBG path solves A p = +div/dt while the update was
 *                   written for the compact path's A p = -div/dt, so the
 *                   literal update doubles the divergence instead of
 *                   removing it (the reference's sign slip); negated, the new
 *                   velocity's divergence on fluid cells is dt * (b - A p).
 *   k_flip_scene_vel  the velocity channels from a scene's MAC field (the
 *                   field the scene prepare's right-hand side was built from)
 *
 * All three are HBM streams over the level-0 blocks (one workgroup of 256
 * threads per block).  Dense, the update moves 42 B per cell (velocity read + write 24,
 * beta 12, pressure 4, flags 2) and the right-hand side 18 B (velocity 12,
 * flags 2, b 4); away from the liquid only the flags and the writes are
 * needed, which flipAlgorithmicBytes counts exactly.
 */

namespace msbgx {

struct FlipArgs {
  const int *level;
  const int64_t *off;
  const uint16_t *flags;
  int nbx, nby, nbz, lb;      // raster, log2 of the level-0 block edge
  int n[3];                   // fine grid size per axis
};

__device__ __forceinline__ bool flipFluid(uint16_t f) { return (f & (CELL_SOLID | CELL_VOID)) == 0; }

/* Slot offsets of the six face-neighbour blocks of block bid (order -x, +x,
 * -y, +y, -z, +z), -1 outside the raster or when the neighbour holds no
 * level-0 data.  Uniform per workgroup. */
__device__ __forceinline__ void flipNeighbourBlocks(const FlipArgs &a, int bid, int bx, int by, int bz, int64_t nbo[6])
{
  const int cb[3] = {bx, by, bz}, nr[3] = {a.nbx, a.nby, a.nbz}, bs[3] = {1, a.nbx, a.nbx * a.nby};
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const int d = k >> 1, dir = (k & 1) ? 1 : -1;
    const int c = cb[d] + dir;
    nbo[k] = -1;
    if (c >= 0 && c < nr[d]) {
      const int nb = bid + dir * bs[d];
      if (a.level[nb] == 0) nbo[k] = a.off[nb];
    }
  }
}

/* Pool index of the neighbour one step along axis d (dir -1 or +1) of cell v
 * with in-block coordinate cd along d, or -1 (outside the domain or in a
 * block without data). */
__device__ __forceinline__ int64_t flipNeighbour(int64_t o, int v, int cd, int d, int dir, int B, int stride,
                                                 const int64_t nbo[6])
{
  const int cn = cd + dir;
  if (cn >= 0 && cn < B) return o + v + dir * stride;
  const int64_t base = nbo[2 * d + (dir > 0)];
  return base < 0 ? -1 : base + v - dir * (B - 1) * stride;
}

/* flipRhsCB over the cells of level-0 blocks (src/flip_sim.cpp:329-338). */
__global__ void __launch_bounds__(256) k_flip_rhs(FlipArgs a, const int *__restrict__ blocks, const float *__restrict__ U,
                                                  const float *__restrict__ V, const float *__restrict__ W, float dt,
                                                  float *__restrict__ Bout, const uint8_t *__restrict__ noFluid)
{
  const int bid = blocks[blockIdx.x];
  const int B = 1 << a.lb, n = B * B * B;
  const int64_t o = a.off[bid];
  const int bx = bid % a.nbx, by = (bid / a.nbx) % a.nby, bz = bid / (a.nbx * a.nby);
  int64_t nbo[6];
  flipNeighbourBlocks(a, bid, bx, by, bz, nbo);
  const float *vel[3] = {U, V, W};
  const int strides[3] = {1, B, B * B};
  if (noFluid[blockIdx.x]) {
    for (int v = threadIdx.x; v < n; v += blockDim.x) Bout[o + v] = 0.0f;
    return;
  }
  for (int v = threadIdx.x; v < n; v += blockDim.x) {
    const int c[3] = {v & (B - 1), (v >> a.lb) & (B - 1), v >> (2 * a.lb)};
    float r = 0.0f;
    if (flipFluid(a.flags[o + v])) {
      float up[3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const int64_t q = flipNeighbour(o, v, c[d], d, +1, B, strides[d], nbo);
        up[d] = q >= 0 ? vel[d][q] : 0.0f;
      }
      float div = up[0] - U[o + v];
      div = div + up[1];
      div = div - V[o + v];
      div = div + up[2];
      div = div - W[o + v];
      r = div / dt;
    }
    Bout[o + v] = r;
  }
}

/* Pressure-gradient update of the three face velocities of each cell of the
 * level-0 blocks (src/flip_sim.cpp:1317-1349).  A face at index 0 along its
 * axis has no -side cell (nbo = -1 there), which the update zeroes. */
__global__ void __launch_bounds__(256) k_flip_update(FlipArgs a, const int *__restrict__ blocks, const float *__restrict__ P,
                                                     const float *__restrict__ W0, const float *__restrict__ W1,
                                                     const float *__restrict__ W2, float *__restrict__ U,
                                                     float *__restrict__ V, float *__restrict__ W, float dt, int negate,
                                                     const uint8_t *__restrict__ noFluid)
{
  const int bid = blocks[blockIdx.x];
  const int B = 1 << a.lb, n = B * B * B;
  const int64_t o = a.off[bid];
  const int bx = bid % a.nbx, by = (bid / a.nbx) % a.nby, bz = bid / (a.nbx * a.nby);
  int64_t nbo[6];
  flipNeighbourBlocks(a, bid, bx, by, bz, nbo);
  const int cb[3] = {bx, by, bz};
  float *vel[3] = {U, V, W};
  const float *beta[3] = {W0, W1, W2};
  const int strides[3] = {1, B, B * B};
  if (noFluid[blockIdx.x]) {
    /* no fluid cell in the block: every face is 0 except a -side boundary
     * face whose cell in the neighbour block is fluid */
    for (int v = threadIdx.x; v < n; v += blockDim.x) {
      const int c[3] = {v & (B - 1), (v >> a.lb) & (B - 1), v >> (2 * a.lb)};
#pragma unroll
      for (int d = 0; d < 3; d++) {
        float nv = 0.0f;
        if (c[d] == 0 && nbo[2 * d] >= 0) {
          const int64_t q = nbo[2 * d] + v + (B - 1) * strides[d];
          if (flipFluid(a.flags[q])) {
            const float pL = negate ? -P[q] : P[q], pR = 0.0f;
            const float t = vel[d][o + v] - dt * beta[d][o + v] * (pR - pL);
            nv = isfinite(t) ? t : 0.0f;
          }
        }
        vel[d][o + v] = nv;
      }
    }
    return;
  }
#pragma unroll 4
  for (int v = threadIdx.x; v < n; v += blockDim.x) {
    const int c[3] = {v & (B - 1), (v >> a.lb) & (B - 1), v >> (2 * a.lb)};
    const bool fR = flipFluid(a.flags[o + v]);
    /* every load of the cell is issued up front (addresses do not depend on
     * loaded values; a face without a -side cell reads the cell itself and
     * discards it), so a cell waits for one memory latency, not a chain */
    const float pO = P[o + v];
    int64_t q[3];
    bool has[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const int64_t qq = (cb[d] > 0 || c[d] > 0) ? flipNeighbour(o, v, c[d], d, -1, B, strides[d], nbo) : -1;
      has[d] = qq >= 0;
      q[d] = has[d] ? qq : o + v;
    }
    uint16_t fl[3];
    float pq[3], bt[3], vv[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      fl[d] = a.flags[q[d]];
      pq[d] = P[q[d]];
      bt[d] = beta[d][o + v];
      vv[d] = vel[d][o + v];
    }
    const float pOwn = fR ? (negate ? -pO : pO) : 0.0f;   // a non-fluid side reads +0
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const bool fL = has[d] && flipFluid(fl[d]);
      const float pL = fL ? (negate ? -pq[d] : pq[d]) : 0.0f;
      const float t = vv[d] - dt * bt[d] * (pOwn - pL);
      const bool face = (cb[d] > 0 || c[d] > 0) && (fL || fR);
      vel[d][o + v] = face ? (isfinite(t) ? t : 0.0f) : 0.0f;
    }
  }
}

/* Velocity channels from a scene's MAC field (msbgx_scene_mac). */
__global__ void __launch_bounds__(256) k_flip_scene_vel(msbgx_scene s, const int *__restrict__ blocks,
                                                        const int64_t *__restrict__ off, int nbx, int nby, int lb,
                                                        float *__restrict__ U, float *__restrict__ V,
                                                        float *__restrict__ W)
{
  const int bid = blocks[blockIdx.x];
  const int B = 1 << lb, n = B * B * B;
  const int64_t o = off[bid];
  const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
  for (int v = threadIdx.x; v < n; v += blockDim.x) {
    const int x = (bx << lb) + (v & (B - 1)), y = (by << lb) + ((v >> lb) & (B - 1)), z = (bz << lb) + (v >> (2 * lb));
    U[o + v] = msbgx_scene_mac(&s, 0, x, y, z);
    V[o + v] = msbgx_scene_mac(&s, 1, x, y, z);
    W[o + v] = msbgx_scene_mac(&s, 2, x, y, z);
  }
}

/* ------------------------------------------------------------------ host */

static int flipCheck(const Grid &g)
{
  if (g.batched) {
    setError("FLIP velocity steps take a single grid, not a batched block set");
    return MSBGX_EINVAL;
  }
  if (!g.prepared || !g.dev) {
    setError("FLIP velocity steps need a prepared grid");
    return MSBGX_ESTATE;
  }
  return MSBGX_OK;
}

static FlipArgs flipArgs(const Grid &g)
{
  const MgLevel &m = g.mg[0];
  FlipArgs a;
  a.level = m.d_level; a.off = m.d_off; a.flags = m.d_flags;
  a.nbx = g.nbx; a.nby = g.nby; a.nbz = g.nbz; a.lb = __builtin_ctz(g.bsx0);
  a.n[0] = g.sx0; a.n[1] = g.sy0; a.n[2] = g.sz0;
  return a;
}

/* Level-0 block list (ascending bid) and zeroed velocity channels, once per prepare. */
static int flipEnsure(Grid &g)
{
  Dev &d = dev(g);
  if (d.vel[0]) return MSBGX_OK;
  const MgLevel &m = g.mg[0];
  std::vector<int> l0;
  for (int bid = 0; bid < g.nBlocks; bid++)
    if (m.level[bid] == 0) l0.push_back(bid);
  d.nLevel0 = (int)l0.size();
  if (int rc = dalloc(&d.level0, l0.size())) return rc;
  if (!l0.empty()) HIPCHK(hipMemcpy(d.level0, l0.data(), l0.size() * sizeof(int), hipMemcpyHostToDevice));
  /* per list entry: the block holds no fluid cell (BLK_NO_FLUID of the level-0 cell types) */
  std::vector<uint8_t> nf(l0.size());
  for (size_t i = 0; i < l0.size(); i++) nf[i] = (m.bflags[l0[i]] & BLK_NO_FLUID) ? 1 : 0;
  if (int rc = dalloc(&d.level0NoFluid, nf.size())) return rc;
  if (!nf.empty()) HIPCHK(hipMemcpy(d.level0NoFluid, nf.data(), nf.size(), hipMemcpyHostToDevice));
  for (int c = 0; c < 3; c++) {
    if (int rc = dalloc(&d.vel[c], (size_t)m.ncells)) return rc;
    HIPCHK(hipMemsetAsync(d.vel[c], 0, (size_t)m.ncells * sizeof(float), d.stream));
  }
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

int deviceFlipSetVelocity(Grid &g, const float *const vel[3])
{
  if (int rc = flipCheck(g)) return rc;
  if (int rc = flipEnsure(g)) return rc;
  Dev &d = dev(g);
  const MgLevel &m = g.mg[0];
  const int B = g.bsx0;
  std::vector<float> pool((size_t)m.ncells, 0.0f);
  for (int c = 0; c < 3; c++) {
#pragma omp parallel for schedule(static)
    for (int bid = 0; bid < g.nBlocks; bid++) {
      if (m.level[bid] != 0) continue;
      int bx, by, bz;
      g.blockCoords(bid, bx, by, bz);
      float *dst = pool.data() + m.off[bid];
      for (int vz = 0; vz < B; vz++)
        for (int vy = 0; vy < B; vy++) {
          const size_t src = ((size_t)(bz * B + vz) * g.sy0 + (size_t)(by * B + vy)) * g.sx0 + (size_t)bx * B;
          std::copy(vel[c] + src, vel[c] + src + B, dst + (size_t)B * (vy + (size_t)B * vz));
        }
    }
    HIPCHK(hipMemcpy(d.vel[c], pool.data(), pool.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  return MSBGX_OK;
}

int deviceFlipGetVelocity(Grid &g, float *const vel[3])
{
  if (int rc = flipCheck(g)) return rc;
  if (int rc = flipEnsure(g)) return rc;
  Dev &d = dev(g);
  const MgLevel &m = g.mg[0];
  const int B = g.bsx0;
  std::vector<float> pool((size_t)m.ncells);
  for (int c = 0; c < 3; c++) {
    HIPCHK(hipMemcpy(pool.data(), d.vel[c], pool.size() * sizeof(float), hipMemcpyDeviceToHost));
#pragma omp parallel for schedule(static)
    for (int bid = 0; bid < g.nBlocks; bid++) {
      int bx, by, bz;
      g.blockCoords(bid, bx, by, bz);
      const bool has = m.level[bid] == 0;
      const float *src = pool.data() + (has ? m.off[bid] : 0);
      for (int vz = 0; vz < B; vz++)
        for (int vy = 0; vy < B; vy++) {
          float *dst = vel[c] + ((size_t)(bz * B + vz) * g.sy0 + (size_t)(by * B + vy)) * g.sx0 + (size_t)bx * B;
          if (has) std::copy(src + (size_t)B * (vy + (size_t)B * vz), src + (size_t)B * (vy + (size_t)B * vz) + B, dst);
          else std::fill(dst, dst + B, 0.0f);
        }
    }
  }
  return MSBGX_OK;
}

int deviceFlipSetVelocityScene(Grid &g, const msbgx_scene &s)
{
  if (int rc = flipCheck(g)) return rc;
  if (int rc = flipEnsure(g)) return rc;
  Dev &d = dev(g);
  if (d.nLevel0 > 0)
    hipLaunchKernelGGL(k_flip_scene_vel, dim3(d.nLevel0), dim3(256), 0, d.stream, s, d.level0, g.mg[0].d_off, g.nbx,
                       g.nby, __builtin_ctz(g.bsx0), d.vel[0], d.vel[1], d.vel[2]);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

static void flipLaunchRhs(Grid &g, float dt)
{
  Dev &d = dev(g);
  if (d.nLevel0 > 0)
    hipLaunchKernelGGL(k_flip_rhs, dim3(d.nLevel0), dim3(256), 0, d.stream, flipArgs(g), d.level0, d.vel[0],
                       d.vel[1], d.vel[2], dt, g.mg[0].d_ch[C_B], d.level0NoFluid);
}

static void flipLaunchUpdate(Grid &g, float dt, int negate)
{
  Dev &d = dev(g);
  const MgLevel &m = g.mg[0];
  if (d.nLevel0 > 0)
    hipLaunchKernelGGL(k_flip_update, dim3(d.nLevel0), dim3(256), 0, d.stream, flipArgs(g), d.level0, m.d_ch[C_X],
                       m.d_W[0], m.d_W[1], m.d_W[2], d.vel[0], d.vel[1], d.vel[2], dt, negate, d.level0NoFluid);
}

int deviceFlipRhs(Grid &g, float dt)
{
  if (int rc = flipCheck(g)) return rc;
  if (int rc = flipEnsure(g)) return rc;
  Dev &d = dev(g);
  flipLaunchRhs(g, dt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  d.b0Valid = false;
  return MSBGX_OK;
}

int deviceFlipUpdate(Grid &g, float dt, int negate)
{
  if (int rc = flipCheck(g)) return rc;
  if (int rc = flipEnsure(g)) return rc;
  Dev &d = dev(g);
  flipLaunchUpdate(g, dt, negate);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

/* Algorithmic HBM bytes of one launch, from the host copy of the level-0
 * cell types: every value the step must touch counted once.
 *   update: flags 2 + velocity write 12 per cell; pressure 4 per fluid cell;
 *           per face with a fluid side (index > 0): velocity read 4 + beta 4
 *   rhs:    flags 2 + b write 4 per cell; the velocity values on the six
 *           faces of each fluid cell (4 each, a face shared by two fluid
 *           cells counted once; faces outside the domain or in blocks
 *           without data are not read) */
static double flipAlgorithmicBytes(const Grid &g, int which)
{
  const MgLevel &m = g.mg[0];
  const int B = g.bsx0, lb = __builtin_ctz(B);
  const int n[3] = {g.sx0, g.sy0, g.sz0};
  auto fluidAt = [&](int x, int y, int z) -> bool {
    if (x < 0 || y < 0 || z < 0 || x >= n[0] || y >= n[1] || z >= n[2]) return false;
    const int bid = g.blockIndex(x >> lb, y >> lb, z >> lb);
    if (m.level[bid] != 0) return false;
    const int64_t q = m.off[bid] + (x & (B - 1)) + B * ((y & (B - 1)) + B * (z & (B - 1)));
    return (m.flags[q] & (CELL_SOLID | CELL_VOID)) == 0;
  };
  double total = 0.0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total)
  for (int bid = 0; bid < g.nBlocks; bid++) {
    if (m.level[bid] != 0) continue;
    int bx, by, bz;
    g.blockCoords(bid, bx, by, bz);
    double t = 0.0;
    for (int v = 0; v < B * B * B; v++) {
      const int c[3] = {(bx << lb) + (v & (B - 1)), (by << lb) + ((v >> lb) & (B - 1)), (bz << lb) + (v >> (2 * lb))};
      const bool f = fluidAt(c[0], c[1], c[2]);
      if (which == 3) {
        t += 14.0 + (f ? 4.0 : 0.0);
        for (int d = 0; d < 3; d++) {
          if (c[d] == 0) continue;
          const bool fl = fluidAt(c[0] - (d == 0), c[1] - (d == 1), c[2] - (d == 2));
          if (f || fl) t += 8.0;
        }
      } else {
        t += 6.0;
        for (int d = 0; d < 3; d++) {
          /* the -d face of this cell is read when this cell or its -d neighbour is fluid */
          const bool fl = fluidAt(c[0] - (d == 0), c[1] - (d == 1), c[2] - (d == 2));
          if (f || fl) t += 4.0;   // a fluid cell's +d face is its +d neighbour's -d face
        }
      }
    }
    total += t;
  }
  return total;
}

/* Mean device time of `reps` launches of the update (which 3) or the
 * right-hand side (which 4) between HIP events on the handle's stream. */
int deviceFlipTime(Grid &g, int which, int reps, double *ms, double *bytes, int *cells)
{
  if (int rc = flipCheck(g)) return rc;
  if (int rc = flipEnsure(g)) return rc;
  Dev &d = dev(g);
  const int64_t nc = (int64_t)d.nLevel0 * g.bsx0 * g.bsx0 * g.bsx0;
  auto launch = [&]() {
    if (which == 3) flipLaunchUpdate(g, 0.01f, 1);
    else flipLaunchRhs(g, 0.01f);
  };
  launch();
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(d.ev0, d.stream));
  for (int i = 0; i < reps; i++) launch();
  HIPCHK(hipEventRecord(d.ev1, d.stream));
  HIPCHK(hipEventSynchronize(d.ev1));
  float t = 0.0f;
  HIPCHK(hipEventElapsedTime(&t, d.ev0, d.ev1));
  *ms = (double)t / reps;
  *bytes = flipAlgorithmicBytes(g, which);
  *cells = (int)nc;
  d.b0Valid = false;
  return MSBGX_OK;
}

}  // namespace msbgx
