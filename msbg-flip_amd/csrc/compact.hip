/*
 * compact.hip -- the FLIP caller's compact MIC(0)-PCG pressure solver on the
 * GPU (compiled as part of solver.hip), the reference's "paper baseline"
 * solver kind, as a comparator for the MSBG MG-PCG.
 *
 * Reference (src/flip_sim.cpp): buildPressureSystem 953-1011 (rows = fluid
 * cells of the resolution-level-0 grid in (z, y, x) order), rebuildPressurePattern
 * 858-913 (CSR, columns ascending), updatePressureNumerics 915-951 (diag =
 * bxm + bxp + bym + byp + bzm + bzp, off-diagonals -beta, rhs = -div/dt),
 * buildMIC0Compact 1027-1059, applyMIC0Compact 1061-1097,
 * solvePressureCompactPCG 1101-1151, csrSpmv 285-294, compactDot 296-301,
 * scatterPressureToGrid 1013-1023.
 *
 * The system is taken from a prepared grid: fluid = level-0 cell type with
 * neither SOLID nor VOID, beta = level-0 face coefficient (0 where the cell
 * lies outside the domain or in a block without level-0 data), div/dt = the
 * divergence channel.  The solution goes to the pressure channel; it has the
 * compact path's sign (A p = -div/dt), for which the literal velocity update
 * is the right one.
 *
 * GPU form: the rows' six neighbour rows and seven matrix entries (zm, ym,
 * xm, diag, xp, yp, zp: the CSR column order, since rows ascend in (z, y, x))
 * are stored per row; q = A d sums them in that order, as csrSpmv does.  The
 * MIC(0) factor and both triangular solves depend on the previous hyperplane
 * x + y + z = h - 1 (forward) or h + 1 (backward) only, so they run as one
 * launch per hyperplane over that hyperplane's rows; each row's arithmetic is
 * the reference's, so factor and application are bit-identical to it.  The
 * dot products are FP64 per-workgroup partials summed in a fixed order (the
 * reference sums sequentially), alpha and beta are rounded to float as there.
 * One PCG iteration (2 (nx + ny + nz - 2) + 5 launches) is one hipGraph.
 */

namespace msbgx {

struct CompactDev {
  int n = 0, nh = 0;               // rows, hyperplanes
  std::vector<int> hpStart;        // rows of hyperplane h: hpRows[hpStart[h] .. hpStart[h+1])
  int *nbr = nullptr;              // [6][n]: xm, xp, ym, yp, zm, zp neighbour rows, -1 none
  float *val = nullptr;            // [7][n]: zm, ym, xm, diag, xp, yp, zp
  float *bm = nullptr;             // [3][n]: betaXm, betaYm, betaZm
  int *hpRows = nullptr;
  int64_t *rowCell = nullptr;      // MG0 pool index of each row's cell
  float *x = nullptr, *r = nullptr, *z = nullptr, *d = nullptr, *q = nullptr, *tq = nullptr, *precon = nullptr;
  double *part = nullptr;          // [3][kCptParts]
  double *scal = nullptr;          // rhsNormSq, rho, dq, rNormSq, rhoNew
  float *ab = nullptr;             // alpha, beta
  int *flags = nullptr;            // done, converged iteration, iteration
  hipGraphExec_t graph = nullptr;
};

constexpr int kCptParts = 1024;    // workgroups of the vector kernels

/* ------------------------------------------------------------ kernels */

__global__ void k_cpt_precon(const int *__restrict__ rows, int cnt, int n, const int *__restrict__ nbr,
                             const float *__restrict__ val, const float *__restrict__ bm, float *__restrict__ precon)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int row = rows[t];
  const float Adiag = val[3 * n + row];
  float e = Adiag;
  const int rx = nbr[0 * n + row], ry = nbr[2 * n + row], rz = nbr[4 * n + row];
  if (rx >= 0) { const float bx = bm[0 * n + row], p = precon[rx]; e -= (bx * p) * (bx * p); }
  if (ry >= 0) { const float by = bm[1 * n + row], p = precon[ry]; e -= (by * p) * (by * p); }
  if (rz >= 0) { const float bz = bm[2 * n + row], p = precon[rz]; e -= (bz * p) * (bz * p); }
  precon[row] = 1.f / sqrtf(fmaxf(e, 0.25f * fmaxf(Adiag, 1e-10f)));
}

/* forward substitution over one hyperplane: q = (r + sum_lower beta p q) * p */
__global__ void k_cpt_fwd(const int *__restrict__ rows, int cnt, int n, const int *__restrict__ nbr,
                          const float *__restrict__ bm, const float *__restrict__ precon, const float *__restrict__ r,
                          float *__restrict__ q, const int *done)
{
  if (*done) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int row = rows[t];
  float s = r[row];
  const int rx = nbr[0 * n + row], ry = nbr[2 * n + row], rz = nbr[4 * n + row];
  if (rx >= 0) s += bm[0 * n + row] * precon[rx] * q[rx];
  if (ry >= 0) s += bm[1 * n + row] * precon[ry] * q[ry];
  if (rz >= 0) s += bm[2 * n + row] * precon[rz] * q[rz];
  q[row] = s * precon[row];
}

/* backward substitution over one hyperplane: z = (q + sum_upper beta_up p z) * p */
__global__ void k_cpt_bwd(const int *__restrict__ rows, int cnt, int n, const int *__restrict__ nbr,
                          const float *__restrict__ bm, const float *__restrict__ precon, const float *__restrict__ q,
                          float *__restrict__ z, const int *done)
{
  if (*done) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int row = rows[t];
  const float pi = precon[row];
  float s = q[row];
  const int rx = nbr[1 * n + row], ry = nbr[3 * n + row], rz = nbr[5 * n + row];
  if (rx >= 0) s += bm[0 * n + rx] * pi * z[rx];
  if (ry >= 0) s += bm[1 * n + ry] * pi * z[ry];
  if (rz >= 0) s += bm[2 * n + rz] * pi * z[rz];
  z[row] = s * pi;
}

__device__ __forceinline__ double cptBlockSum(double v, double *sh)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += sh[i];
  return s;
}

/* q = A d (entries in CSR column order) and the workgroup's part of d.q */
__global__ void __launch_bounds__(256) k_cpt_spmv(int n, const int *__restrict__ nbr, const float *__restrict__ val,
                                                  const float *__restrict__ d, float *__restrict__ q,
                                                  double *__restrict__ part, const int *done)
{
  if (*done) return;
  __shared__ double sh[4];
  double acc = 0.0;
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n; row += gridDim.x * blockDim.x) {
    const int rzm = nbr[4 * n + row], rym = nbr[2 * n + row], rxm = nbr[0 * n + row];
    const int rxp = nbr[1 * n + row], ryp = nbr[3 * n + row], rzp = nbr[5 * n + row];
    float s = 0.f;
    if (rzm >= 0) s += val[0 * n + row] * d[rzm];
    if (rym >= 0) s += val[1 * n + row] * d[rym];
    if (rxm >= 0) s += val[2 * n + row] * d[rxm];
    s += val[3 * n + row] * d[row];
    if (rxp >= 0) s += val[4 * n + row] * d[rxp];
    if (ryp >= 0) s += val[5 * n + row] * d[ryp];
    if (rzp >= 0) s += val[6 * n + row] * d[rzp];
    q[row] = s;
    acc += (double)d[row] * (double)s;
  }
  const double b = cptBlockSum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

/* partial dot product a.b */
__global__ void __launch_bounds__(256) k_cpt_dot(int n, const float *__restrict__ a, const float *__restrict__ b,
                                                 double *__restrict__ part, const int *done)
{
  if (done && *done) return;
  __shared__ double sh[4];
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    acc += (double)a[i] * (double)b[i];
  const double s = cptBlockSum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

/* x += alpha d, r -= alpha q, partial r.r */
__global__ void __launch_bounds__(256) k_cpt_update(int n, const float *__restrict__ ab, const float *__restrict__ d,
                                                    const float *__restrict__ q, float *__restrict__ x,
                                                    float *__restrict__ r, double *__restrict__ part, const int *done)
{
  if (*done) return;
  __shared__ double sh[4];
  const float alpha = ab[0];
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += alpha * d[i];
    const float ri = r[i] - alpha * q[i];
    r[i] = ri;
    acc += (double)ri * (double)ri;
  }
  const double s = cptBlockSum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

/* d = z + beta d */
__global__ void k_cpt_dir(int n, const float *__restrict__ ab, const float *__restrict__ z, float *__restrict__ d,
                          const int *done)
{
  if (done && *done) return;
  const float beta = ab ? ab[1] : 0.0f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    d[i] = ab ? z[i] + beta * d[i] : z[i];
}

/* The scalar steps of one iteration (single workgroup, partials summed in
 * order).  stage 0: rho = r.z (start); 1: dq -> alpha or stop; 2: rNormSq ->
 * convergence; 3: rhoNew -> beta. */
__global__ void __launch_bounds__(64) k_cpt_scalar(int stage, const double *__restrict__ part, int np,
                                                   double *__restrict__ sc, float *__restrict__ ab,
                                                   int *__restrict__ flags, int maxIter, double tolSq)
{
  if (stage != 0 && flags[0]) return;
  /* one wave: strided partial sums, then a fixed shuffle tree */
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += 64) s += part[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (threadIdx.x != 0) return;
  switch (stage) {
  case 0:
    sc[1] = s;
    break;
  case 1:
    sc[2] = s;
    if (s <= 1e-30 || sc[1] <= 1e-30) { flags[0] = 1; break; }
    ab[0] = (float)(sc[1] / s);
    break;
  case 2:
    sc[3] = s;
    flags[2] += 1;
    if (s <= tolSq * sc[0]) { flags[0] = 1; flags[1] = flags[2]; break; }
    if (flags[2] >= maxIter) flags[0] = 1;
    break;
  case 3:
    sc[4] = s;
    if (sc[1] <= 1e-30) { flags[0] = 1; break; }
    ab[1] = (float)(s / sc[1]);
    sc[1] = s;
    break;
  }
}

/* pool <-> rows */
__global__ void k_cpt_gather(int n, const int64_t *__restrict__ rowCell, const float *__restrict__ src, float sign,
                             float *__restrict__ dst)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = sign * src[rowCell[i]];
}

__global__ void k_cpt_scatter(int n, const int64_t *__restrict__ rowCell, const float *__restrict__ src,
                              float *__restrict__ dst)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[rowCell[i]] = src[i];
}

/* ------------------------------------------------------------ host */

static void freeCompact(CompactDev *c)
{
  if (!c) return;
  auto f = [](void *p) { if (p) (void)hipFree(p); };
  f(c->nbr); f(c->val); f(c->bm); f(c->hpRows); f(c->rowCell); f(c->x); f(c->r); f(c->z); f(c->d); f(c->q);
  f(c->tq); f(c->precon); f(c->part); f(c->scal); f(c->ab); f(c->flags);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  delete c;
}

/* Rows, neighbours, entries and hyperplane lists from the prepared grid's
 * level-0 data (host copies). */
static int compactBuild(Grid &g, CompactDev &c)
{
  const MgLevel &m = g.mg[0];
  const int nx = g.sx0, ny = g.sy0, nz = g.sz0, B = g.bsx0, lb = __builtin_ctz(B);
  auto cellIdx = [&](int x, int y, int z) -> int64_t {   // MG0 pool index, -1 without level-0 data
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return -1;
    const int bid = g.blockIndex(x >> lb, y >> lb, z >> lb);
    if (m.level[bid] != 0) return -1;
    return m.off[bid] + (x & (B - 1)) + B * ((y & (B - 1)) + B * (z & (B - 1)));
  };
  auto fluid = [&](int64_t q) { return q >= 0 && (m.flags[q] & (CELL_SOLID | CELL_VOID)) == 0; };
  auto beta = [&](int d, int x, int y, int z) -> float {   // getVC(sgBeta, ...): 0 without data
    const int64_t q = cellIdx(x, y, z);
    return q >= 0 ? m.W[d][q] : 0.f;
  };
  /* rows in (z, y, x) order; cell -> row over the domain */
  std::vector<int> cellRow((size_t)nx * ny * nz, -1);
  std::vector<int> rx, ry, rz;
  std::vector<int64_t> rowCellHost;
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const int64_t q = cellIdx(x, y, z);
        if (!fluid(q)) continue;
        cellRow[((size_t)z * ny + y) * nx + x] = (int)rx.size();
        rx.push_back(x); ry.push_back(y); rz.push_back(z); rowCellHost.push_back(q);
      }
  const int n = (int)rx.size();
  c.n = n;
  auto row = [&](int x, int y, int z) -> int {
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return -1;
    return cellRow[((size_t)z * ny + y) * nx + x];
  };
  std::vector<int> nbr((size_t)6 * n);
  std::vector<float> val((size_t)7 * n), bm((size_t)3 * n);
  c.nh = nx + ny + nz - 2;
  c.hpStart.assign(c.nh + 1, 0);
  for (int i = 0; i < n; i++) c.hpStart[rx[i] + ry[i] + rz[i] + 1]++;
  for (int h = 0; h < c.nh; h++) c.hpStart[h + 1] += c.hpStart[h];
  std::vector<int> hpRows(n), fill(c.hpStart.begin(), c.hpStart.end() - 1);
  for (int i = 0; i < n; i++) {
    const int x = rx[i], y = ry[i], z = rz[i];
    const int nb[6] = {row(x - 1, y, z), row(x + 1, y, z), row(x, y - 1, z), row(x, y + 1, z), row(x, y, z - 1),
                       row(x, y, z + 1)};
    for (int k = 0; k < 6; k++) nbr[(size_t)k * n + i] = nb[k];
    const float bxm = x > 0 ? beta(0, x, y, z) : 0.f, bxp = x + 1 < nx ? beta(0, x + 1, y, z) : 0.f;
    const float bym = y > 0 ? beta(1, x, y, z) : 0.f, byp = y + 1 < ny ? beta(1, x, y + 1, z) : 0.f;
    const float bzm = z > 0 ? beta(2, x, y, z) : 0.f, bzp = z + 1 < nz ? beta(2, x, y, z + 1) : 0.f;
    bm[0 * (size_t)n + i] = bxm; bm[1 * (size_t)n + i] = bym; bm[2 * (size_t)n + i] = bzm;
    val[0 * (size_t)n + i] = nb[4] >= 0 ? -bzm : 0.f;
    val[1 * (size_t)n + i] = nb[2] >= 0 ? -bym : 0.f;
    val[2 * (size_t)n + i] = nb[0] >= 0 ? -bxm : 0.f;
    val[3 * (size_t)n + i] = bxm + bxp + bym + byp + bzm + bzp;
    val[4 * (size_t)n + i] = nb[1] >= 0 ? -bxp : 0.f;
    val[5 * (size_t)n + i] = nb[3] >= 0 ? -byp : 0.f;
    val[6 * (size_t)n + i] = nb[5] >= 0 ? -bzp : 0.f;
    hpRows[fill[x + y + z]++] = i;
  }
  auto up = [](auto **p, const auto &v) -> int {
    if (hipMalloc((void **)p, std::max<size_t>(v.size(), 1) * sizeof(v[0])) != hipSuccess) return MSBGX_ENOMEM;
    if (!v.empty() && hipMemcpy(*p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice) != hipSuccess)
      return MSBGX_EHIP;
    return MSBGX_OK;
  };
  if (int rc = up(&c.nbr, nbr)) return rc;
  if (int rc = up(&c.val, val)) return rc;
  if (int rc = up(&c.bm, bm)) return rc;
  if (int rc = up(&c.hpRows, hpRows)) return rc;
  if (int rc = up(&c.rowCell, rowCellHost)) return rc;
  for (float **p : {&c.x, &c.r, &c.z, &c.d, &c.q, &c.tq, &c.precon})
    if (int rc = dalloc(p, (size_t)n)) return rc;
  if (int rc = dalloc(&c.part, 3 * kCptParts)) return rc;
  if (int rc = dalloc(&c.scal, 5)) return rc;
  if (int rc = dalloc(&c.ab, 2)) return rc;
  if (int rc = dalloc(&c.flags, 3)) return rc;
  return MSBGX_OK;
}

static void cptLaunchPrecon(CompactDev &c, hipStream_t s)
{
  for (int h = 0; h < c.nh; h++) {
    const int cnt = c.hpStart[h + 1] - c.hpStart[h];
    if (cnt > 0)
      hipLaunchKernelGGL(k_cpt_precon, dim3((cnt + 255) / 256), dim3(256), 0, s, c.hpRows + c.hpStart[h], cnt, c.n,
                         c.nbr, c.val, c.bm, c.precon);
  }
}

/* z = M^-1 r (applyMIC0Compact) */
static void cptLaunchApply(CompactDev &c, const float *r, float *z, hipStream_t s)
{
  for (int h = 0; h < c.nh; h++) {
    const int cnt = c.hpStart[h + 1] - c.hpStart[h];
    if (cnt > 0)
      hipLaunchKernelGGL(k_cpt_fwd, dim3((cnt + 255) / 256), dim3(256), 0, s, c.hpRows + c.hpStart[h], cnt, c.n,
                         c.nbr, c.bm, c.precon, r, c.tq, c.flags);
  }
  for (int h = c.nh - 1; h >= 0; h--) {
    const int cnt = c.hpStart[h + 1] - c.hpStart[h];
    if (cnt > 0)
      hipLaunchKernelGGL(k_cpt_bwd, dim3((cnt + 255) / 256), dim3(256), 0, s, c.hpRows + c.hpStart[h], cnt, c.n,
                         c.nbr, c.bm, c.precon, c.tq, z, c.flags);
  }
}

static int cptParts(int n) { return std::max(1, std::min(kCptParts, (n + 255) / 256)); }

/* One PCG iteration of solvePressureCompactPCG (src/flip_sim.cpp:1125-1147). */
static void cptLaunchIteration(CompactDev &c, int maxIter, double tolSq, hipStream_t s)
{
  const int np = cptParts(c.n);
  hipLaunchKernelGGL(k_cpt_spmv, dim3(np), dim3(256), 0, s, c.n, c.nbr, c.val, c.d, c.q, c.part, c.flags);
  hipLaunchKernelGGL(k_cpt_scalar, dim3(1), dim3(64), 0, s, 1, c.part, np, c.scal, c.ab, c.flags, maxIter, tolSq);
  hipLaunchKernelGGL(k_cpt_update, dim3(np), dim3(256), 0, s, c.n, c.ab, c.d, c.q, c.x, c.r, c.part, c.flags);
  hipLaunchKernelGGL(k_cpt_scalar, dim3(1), dim3(64), 0, s, 2, c.part, np, c.scal, c.ab, c.flags, maxIter, tolSq);
  cptLaunchApply(c, c.r, c.z, s);
  hipLaunchKernelGGL(k_cpt_dot, dim3(np), dim3(256), 0, s, c.n, c.r, c.z, c.part, c.flags);
  hipLaunchKernelGGL(k_cpt_scalar, dim3(1), dim3(64), 0, s, 3, c.part, np, c.scal, c.ab, c.flags, maxIter, tolSq);
  hipLaunchKernelGGL(k_cpt_dir, dim3(np), dim3(256), 0, s, c.n, c.ab, c.z, c.d, c.flags);
}

/* Solve; the solution goes to the pressure channel (0 on every other cell). */
int deviceSolveCompact(Grid &g, int maxIter, float tol, int *iters, float *rel, double *ms)
{
  if (int rc = flipCheck(g)) return rc;
  Dev &d = dev(g);
  hipStream_t s = d.stream;
  CompactDev *c = new CompactDev();
  struct Guard {
    CompactDev *c;
    ~Guard() { freeCompact(c); }
  } guard{c};
  if (int rc = compactBuild(g, *c)) {
    if (rc == MSBGX_ENOMEM) setError("compact solver: device allocation failed");
    else setError("compact solver: upload failed");
    return rc;
  }
  MgLevel &m0 = g.mg[0];
  const int n = c->n;
  const int np = cptParts(n);
  HIPCHK(hipMemsetAsync(m0.d_ch[C_X], 0, (size_t)m0.ncells * sizeof(float), s));
  *iters = 0;
  *rel = 0.f;
  if (ms) *ms = 0.0;
  if (n == 0) { HIPCHK(hipStreamSynchronize(s)); return MSBGX_OK; }
  const double tolSq = (double)tol * (double)tol;
  HIPCHK(hipEventRecord(d.ev0, s));
  /* x = 0, r = rhs = -div/dt, |rhs|^2 */
  HIPCHK(hipMemsetAsync(c->x, 0, (size_t)n * sizeof(float), s));
  HIPCHK(hipMemsetAsync(c->flags, 0, 3 * sizeof(int), s));
  HIPCHK(hipMemsetAsync(c->precon, 0, (size_t)n * sizeof(float), s));
  hipLaunchKernelGGL(k_cpt_gather, dim3((n + 255) / 256), dim3(256), 0, s, n, c->rowCell, m0.d_ch[C_B], -1.0f, c->r);
  hipLaunchKernelGGL(k_cpt_dot, dim3(np), dim3(256), 0, s, n, c->r, c->r, c->part, (const int *)nullptr);
  HIPCHK(hipGetLastError());
  std::vector<double> part(np);
  HIPCHK(hipMemcpyAsync(part.data(), c->part, np * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double rhs2 = 0.0;
  for (double v : part) rhs2 += v;
  if (rhs2 <= 1e-30) {
    HIPCHK(hipEventRecord(d.ev1, s));
    HIPCHK(hipStreamSynchronize(s));
    return MSBGX_OK;
  }
  HIPCHK(hipMemcpyAsync(c->scal, &rhs2, sizeof(double), hipMemcpyHostToDevice, s));
  cptLaunchPrecon(*c, s);
  cptLaunchApply(*c, c->r, c->z, s);
  hipLaunchKernelGGL(k_cpt_dot, dim3(np), dim3(256), 0, s, n, c->r, c->z, c->part, (const int *)nullptr);
  hipLaunchKernelGGL(k_cpt_scalar, dim3(1), dim3(64), 0, s, 0, c->part, np, c->scal, c->ab, c->flags, maxIter, tolSq);
  hipLaunchKernelGGL(k_cpt_dir, dim3(np), dim3(256), 0, s, n, (const float *)nullptr, c->z, c->d, (const int *)nullptr);
  HIPCHK(hipGetLastError());
  /* one iteration per graph replay; every kernel returns at once after the done flag */
  const bool useGraph = !(g.options & MSBGX_OPT_NO_GRAPH);
  if (useGraph) {
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    cptLaunchIteration(*c, maxIter, tolSq, s);
    HIPCHK(hipStreamEndCapture(s, &graph));
    const hipError_t e = hipGraphInstantiate(&c->graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
  }
  int flags[3] = {0, 0, 0};
  for (int it = 0; it < maxIter;) {
    const int chunk = std::min(16, maxIter - it);
    for (int k = 0; k < chunk; k++) {
      if (useGraph) HIPCHK(hipGraphLaunch(c->graph, s));
      else cptLaunchIteration(*c, maxIter, tolSq, s);
    }
    it += chunk;
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(flags, c->flags, sizeof(flags), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (flags[0]) break;
  }
  hipLaunchKernelGGL(k_cpt_scatter, dim3((n + 255) / 256), dim3(256), 0, s, n, c->rowCell, c->x, m0.d_ch[C_X]);
  hipLaunchKernelGGL(k_cpt_dot, dim3(np), dim3(256), 0, s, n, c->r, c->r, c->part, (const int *)nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(d.ev1, s));
  HIPCHK(hipMemcpyAsync(part.data(), c->part, np * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double rr = 0.0;
  for (double v : part) rr += v;
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, d.ev0, d.ev1));
  if (ms) *ms = t;
  *iters = flags[1] > 0 ? flags[1] : flags[2];   // converged, else the iterations run (maxIter or a breakdown)
  *rel = (float)std::sqrt(rr / rhs2);
  return MSBGX_OK;
}

/* MIC(0) factor and one application z = M^-1 rhs (rhs = -div/dt) on the
 * compact rows; precon and z return in row order ((z, y, x) over the fluid
 * cells), up to cap rows.  With cap = 0 only the row count is returned. */
int deviceCompactApply(Grid &g, float *preconOut, float *zOut, int cap, int *nRows)
{
  if (int rc = flipCheck(g)) return rc;
  Dev &d = dev(g);
  hipStream_t s = d.stream;
  CompactDev *c = new CompactDev();
  struct Guard {
    CompactDev *c;
    ~Guard() { freeCompact(c); }
  } guard{c};
  if (int rc = compactBuild(g, *c)) return rc;
  const int n = c->n;
  *nRows = n;
  if (cap <= 0 || n == 0) return MSBGX_OK;
  if (cap < n) {
    setError("compact MIC(0): output arrays shorter than the row count");
    return MSBGX_EINVAL;
  }
  MgLevel &m0 = g.mg[0];
  {
    HIPCHK(hipMemsetAsync(c->flags, 0, 3 * sizeof(int), s));
    HIPCHK(hipMemsetAsync(c->precon, 0, (size_t)n * sizeof(float), s));
    hipLaunchKernelGGL(k_cpt_gather, dim3((n + 255) / 256), dim3(256), 0, s, n, c->rowCell, m0.d_ch[C_B], -1.0f, c->r);
    cptLaunchPrecon(*c, s);
    cptLaunchApply(*c, c->r, c->z, s);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipMemcpy(preconOut, c->precon, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(zOut, c->z, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
  return MSBGX_OK;
}

}  // namespace msbgx
