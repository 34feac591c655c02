/*
 * pde.hip -- the msbg_demo solver entry on the GPU (compiled as part of
 * solver.hip): PDE smoothing of a density channel with the in-place
 * multicolour block schemes.
 *
 * Reference: MultiresSparseGrid::applyChannelPdeFast (src/msbg4.cpp:60-960)
 * with its halo blocks (HaloBlockSet::fillHaloBlock_ with OPT_BC_NEUMANN,
 * src/halo.cpp:267-480, getOutOfBlockValue src/halo.h:267-290); msbg_demo
 * calls it for mean curvature with 8 colours on uint16 render density
 * (src/msbg_demo.cpp:716-727, RSURF_16_BIT src/msbgcell.h:68):
 *   iterations: numIter rounded up to even (src/msbg4.cpp:151-152);
 *   schemes (src/msbg4.cpp:100-112, 411-434): 8 colours, pass c updating
 *   colour (bx&1) + 2(by&1) + 4(bz&1) (src/msbg.h:642-646); or red-black,
 *   pass p updating the blocks with (bx+by+bz)&1 != p (colour 1 first);
 *   per listed level-0 block of the pass's colour:
 *     halo (one voxel, two for the bi-Laplacian) as float (uint16: v/65535):
 *     a voxel of a level-0 block from the channel, any other voxel (outside
 *     the domain, or in a block of another level) the block's own nearest
 *     voxel (Neumann);
 *     per voxel the type's update (src/msbg4.cpp:555-800): 4 mean
 *     curvature, 2 bi-Laplacian times |grad|, 1 Laplacian times |grad|, 3
 *     simple diffusion, 14 Perona-Malik, 15 its Jacobi form, 17 Jacobi
 *     diffusion; Taubin alternation for 4 and 1; then the hard constraints
 *     [max(0, phi0 - tol1), min(1, phi0 + tol2)] around the input, or the
 *     [0, 1] clamp (src/msbg4.cpp:806-820); uint16 stored as
 *     round-to-nearest-even(phi 65535).
 * In-place passes are deterministic where no stencil reaches a block that
 * the same pass writes: 8 colours for every type (same-colour blocks are
 * two blocks apart), red-black for the face-neighbour stencils (1, 3, 14,
 * 15, 17).  Red-black with the mean-curvature or bi-Laplacian stencils
 * reads diagonal blocks of the pass's own colour while they are written
 * (the reference's result then depends on thread timing) and is rejected.
 *
 * Data: the channel over the level-0 blocks in ascending block order, B^3
 * values per block; one workgroup per block stages the (B+2h)^3 halo in LDS
 * (dynamic: 157 KB for 32^3 blocks with a one-voxel halo).  Algorithmic
 * bytes per voxel and iteration: one value read and one written.
 */

#include <mutex>
#include <string>

namespace msbgx {

struct PdeArgs {
  const int *slot;   // per block: rank among the level-0 blocks, -1 otherwise
  int nbx, nby, nbz, lb;
  int n[3];
  float dt;
  int typ;            // reference laplTyp without the scheme bits
  int iIter;          // iteration index (Taubin alternation)
  int doTaubin;
  float taubin;
  int doHard;         // hard constraints around phi0
  float tol1, tol2;
  int clamp01;        // doConstrZeroOne
  float alpha;        // anisDiffAlpha
  const void *phi0;   // slot-layout copy of the input (hard constraints)
};

constexpr int kPdeOpt8Color = 1 << 29;   // OPT_8_COLOR_SCHEME (src/msbg.h:132)

template <typename T> __device__ __forceinline__ float pdeLoad(const T *p, int64_t i)
{
  if constexpr (sizeof(T) == 2) return (float)p[i] * (1.f / 65535);   // renderDensToFloat_<uint16_t>
  else return p[i];
}

/* FMA_FAST_EXP (src/fastmath.h:583-626): a Cephes-style expf with the
 * reference's coefficient sequence (its PX2 term is added twice), here in
 * IEEE single operations without contraction. */
__device__ __forceinline__ float pdeFastExp(float x0)
{
  const float MAXLOGF = 88.72283905206835f, MINLOGF = -88.f;
  const float C1F = 0.693359375f, C2F = -2.12194440e-4f;
  const float P1 = 1.9875691500E-4f, P2 = 1.3981999507E-3f, P3 = 8.3334519073E-3f, P4 = 4.1665795894E-2f,
              P5 = 1.6666665459E-1f, P6 = 5.0000001201E-1f;
  const float LOG2EF = 1.44269504088896341f;
  float x = x0;
  float z = floorf(LOG2EF * x + 0.5f);
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
  z *= __uint_as_float((uint32_t)(n + 0x7f) << 23);
  if (x0 < MINLOGF) z = 0.0f;
  if (x0 > MAXLOGF) z = __builtin_huge_valf();
  return z;
}

/* Perona-Malik conductance G_ (src/msbg4.cpp:749-753). */
__device__ __forceinline__ float pdeConductance(float f, float alpha)
{
  if (alpha < 0) {
    const float t = fabsf(f) * (1.0f / -alpha);
    return pdeFastExp(-(t * t));
  }
  const float t = fabsf(f) / alpha;
  return 1.0f / (1.0f + t * t);
}

/* One voxel's update by type (src/msbg4.cpp:555-800), F(dx, dy, dz) reading
 * the halo block; the operation order is the reference's. */
template <int HL, class FN>
__device__ __forceinline__ float pdeVoxel(const PdeArgs &a, FN F, const float *ph)
{
  const float dt = a.dt;
  float phi = F(0, 0, 0);
  if (a.typ == 4) {
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
    if (a.doTaubin) phi = (a.iIter & 1) ? phi + D : phi - a.taubin * D;
    else phi += D;
  } else if (HL == 2 && a.typ == 2) {
    const float F0 = F(0, 0, 0);
    const float Hc = 42.f * F(0, 0, 0) -
                     12.f * (F(1, 0, 0) + F(0, 1, 0) + F(0, 0, 1) + F(-1, 0, 0) + F(0, -1, 0) + F(0, 0, -1)) +
                     1.f * (F(2, 0, 0) + F(0, 2, 0) + F(0, 0, 2) + F(-2, 0, 0) + F(0, -2, 0) + F(0, 0, -2)) +
                     2.f * (F(1, 0, 1) + F(0, 1, 1) + F(-1, 0, 1) + F(0, -1, 1) + F(1, 0, -1) + F(0, 1, -1) +
                            F(-1, 0, -1) + F(0, -1, -1) + F(1, 1, 0) + F(-1, 1, 0) + F(-1, -1, 0) + F(1, -1, 0));
    const float GX = 0.5f * (F(1, 0, 0) - F(-1, 0, 0)), GY = 0.5f * (F(0, 1, 0) - F(0, -1, 0)),
                GZ = 0.5f * (F(0, 0, 1) - F(0, 0, -1));
    const float G = sqrtf(GX * GX + GY * GY + GZ * GZ);
    float D = -dt * Hc * G;
    D = fminf(fmaxf(D, -0.1f), 0.1f);
    phi = F0 + D;
  } else if (a.typ == 1 || a.typ == 3 || a.typ == 17) {
    /* the reference's fixed strides of an 18-voxel halo row: 16^3 blocks */
    const float F0 = phi, F1 = ph[-1], F2 = ph[1], F3 = ph[-18], F4 = ph[18], F5 = ph[-18 * 18], F6 = ph[18 * 18];
    if (a.typ == 1) {
      const float L = F1 + F2 + F3 + F4 + F5 + F6 - 6.0f * F0;
      const float GX = 0.5f * (F2 - F1), GY = 0.5f * (F4 - F3), GZ = 0.5f * (F6 - F5);
      const float G = sqrtf(GX * GX + GY * GY + GZ * GZ);
      const float D = dt * L * G;
      if (a.doTaubin) phi = (a.iIter & 1) ? phi + D : phi - a.taubin * D;
      else phi += D;
    } else if (a.typ == 3) {
      const float S = (1.0f / 6.0f) * (F1 + F2 + F3 + F4 + F5 + F6);
      const float D = dt * (S - phi);
      phi += D;
    } else {
      phi = (phi + dt * (F1 + F2 + F3 + F4 + F5 + F6)) * 1.0f / (1.0f + dt * 6.0f);
    }
  } else {   // 14, 15
    const float f0 = F(0, 0, 0), f1 = F(-1, 0, 0), f2 = F(1, 0, 0), f3 = F(0, -1, 0), f4 = F(0, 1, 0),
                f5 = F(0, 0, -1), f6 = F(0, 0, 1);
    const float g1 = f1 - f0, g2 = f2 - f0, g3 = f3 - f0, g4 = f4 - f0, g5 = f5 - f0, g6 = f6 - f0;
    const float al = a.alpha;
    const float G1 = pdeConductance(g1, al), G2 = pdeConductance(g2, al), G3 = pdeConductance(g3, al),
                G4 = pdeConductance(g4, al), G5 = pdeConductance(g5, al), G6 = pdeConductance(g6, al);
    if (a.typ == 15) {
      const float sum = G1 * f1 + G2 * f2 + G3 * f3 + G4 * f4 + G5 * f5 + G6 * f6,
                  wsum = G1 + G2 + G3 + G4 + G5 + G6;
      phi = (phi + dt * sum) * 1.0f / (1.0f + dt * wsum);
    } else {
      const float sum = G1 * g1 + G2 * g2 + G3 * g3 + G4 * g4 + G5 * g5 + G6 * g6;
      phi += dt * sum;
    }
  }
  return phi;
}

/* One workgroup per listed block of one colour: stage the halo block in
 * LDS, then update every voxel in place.  HL: halo width (2 for the
 * bi-Laplacian).  The face-neighbour types skip the halo's edge and corner
 * voxels, which they never read (in a red-black pass those belong to
 * blocks of the same colour). */
template <int B, int HL, typename T>
__global__ void __launch_bounds__(B == 32 ? 1024 : 256) k_pde(PdeArgs a, const int *__restrict__ blocks,
                                                              T *__restrict__ rho)
{
  constexpr int H = B + 2 * HL, H2 = H * H, N = B * B * B;
  extern __shared__ float halo[];   // H^3
  __shared__ int nslot[27];         // slots of the 3x3x3 block neighbourhood, -1: none or outside
  const int bid = blocks[blockIdx.x];
  const int bx = bid % a.nbx, by = (bid / a.nbx) % a.nby, bz = bid / (a.nbx * a.nby);
  const int64_t own = (int64_t)a.slot[bid] * N;
  const int g0[3] = {bx * B, by * B, bz * B};
  const bool faceOnly = a.typ != 4 && a.typ != 2;
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
    if (faceOnly) {
      const int outside = (hx < HL || hx >= HL + B) + (hy < HL || hy >= HL + B) + (hz < HL || hz >= HL + B);
      if (outside >= 2) continue;
    }
    const int g[3] = {g0[0] + hx - HL, g0[1] + hy - HL, g0[2] + hz - HL};
    int64_t src = -1;
    if (g[0] >= 0 && g[1] >= 0 && g[2] >= 0 && g[0] < a.n[0] && g[1] < a.n[1] && g[2] < a.n[2]) {
      /* hx - HL in [-HL, B + HL - 1]: neighbour index 0, 1 or 2 along each axis */
      const int s = nslot[(hx - HL + B) / B + 3 * ((hy - HL + B) / B) + 9 * ((hz - HL + B) / B)];
      if (s >= 0) src = (int64_t)s * N + (g[0] & (B - 1)) + B * ((g[1] & (B - 1)) + B * (g[2] & (B - 1)));
    }
    if (src < 0) {   // Neumann: the block's own nearest voxel
      const int cx = min(max(hx - HL, 0), B - 1), cy = min(max(hy - HL, 0), B - 1), cz = min(max(hz - HL, 0), B - 1);
      src = own + cx + B * (cy + B * cz);
    }
    halo[h] = pdeLoad(rho, src);
  }
  __syncthreads();
  for (int v = threadIdx.x; v < N; v += blockDim.x) {
    const int vx = v % B, vy = (v / B) % B, vz = v / (B * B);
    const float *ph = halo + (vx + HL) + H * (vy + HL) + H2 * (vz + HL);
    auto F = [&](int dx, int dy, int dz) { return ph[dx + H * dy + H2 * dz]; };
    float phi = pdeVoxel<HL>(a, F, ph);
    if (a.doHard) {
      const float p0 = pdeLoad(static_cast<const T *>(a.phi0), own + v);
      const float phiMin = fmaxf(0.0f, p0 - a.tol1), phiMax = fminf(1.0f, p0 + a.tol2);
      phi = fmaxf(fminf(phi, phiMax), phiMin);
    } else if (a.clamp01) {
      phi = fminf(fmaxf(phi, 0.0f), 1.0f);
    }
    if constexpr (sizeof(T) == 2) rho[own + v] = (uint16_t)(int)rintf(phi * 65535);
    else rho[own + v] = phi;
  }
}

template <int B, int HL, typename T> static void pdeLaunch(const PdeArgs &a, const int *list, int cnt, T *rho,
                                                           hipStream_t s)
{
  constexpr int H = B + 2 * HL;
  constexpr size_t lds = (size_t)H * H * H * sizeof(float);
  if constexpr (lds > 64 * 1024) {
    static std::once_flag once;   // the dynamic LDS limit above 64 KB, once per process
    std::call_once(once, [] {
      (void)hipFuncSetAttribute((const void *)k_pde<B, HL, T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    });
  }
  hipLaunchKernelGGL((k_pde<B, HL, T>), dim3(cnt), dim3(B == 32 ? 1024 : 256), lds, s, a, list, rho);
}

template <typename T> static void pdeLaunchB(int B, int HL, const PdeArgs &a, const int *list, int cnt, T *rho,
                                             hipStream_t s)
{
  if (HL == 2) {
    if (B == 16) pdeLaunch<16, 2, T>(a, list, cnt, rho, s);
    else if (B == 8) pdeLaunch<8, 2, T>(a, list, cnt, rho, s);
    else if (B == 4) pdeLaunch<4, 2, T>(a, list, cnt, rho, s);
    else pdeLaunch<2, 2, T>(a, list, cnt, rho, s);
    return;
  }
  if (B == 32) pdeLaunch<32, 1, T>(a, list, cnt, rho, s);
  else if (B == 16) pdeLaunch<16, 1, T>(a, list, cnt, rho, s);
  else if (B == 8) pdeLaunch<8, 1, T>(a, list, cnt, rho, s);
  else if (B == 4) pdeLaunch<4, 1, T>(a, list, cnt, rho, s);
  else pdeLaunch<2, 1, T>(a, list, cnt, rho, s);
}

/* nIter iterations of nPass colour passes: per pass one launch over that
 * pass's listed level-0 blocks (lists[start[c] .. start[c+1])). */
template <typename T>
static void pdeColourPasses(int B, PdeArgs a, const int *lists, const int *start, int nPass, int nIter, T *rho,
                            hipStream_t s)
{
  const int HL = a.typ == 2 ? 2 : 1;
  for (int it = 0; it < nIter; it++) {
    a.iIter = it;
    for (int c = 0; c < nPass; c++) {
      const int cnt = start[c + 1] - start[c];
      if (cnt > 0) pdeLaunchB<T>(B, HL, a, lists + start[c], cnt, rho, s);
    }
  }
}

/* The scheme and type of a reference laplTyp (src/msbg4.cpp:100-112) and
 * the checks of what is defined; returns MSBGX_OK and fills a.typ, *nPass. */
static int pdeDecode(int laplTyp, int B, int dtype, const msbgx_pde_params &p, PdeArgs &a, int *nPass)
{
  if (laplTyp >= 0) {
    setError("PDE smoothing: only the in-place multicolour schemes are built (laplTyp < 0)");
    return MSBGX_EINVAL;
  }
  laplTyp = -laplTyp;
  bool color8 = false;
  if (laplTyp & kPdeOpt8Color) {
    laplTyp &= ~kPdeOpt8Color;
    color8 = true;
  }
  const int t = laplTyp;
  if (t != 1 && t != 2 && t != 3 && t != 4 && t != 14 && t != 15 && t != 17) {
    setError("PDE smoothing: type must be 1, 2, 3, 4, 14, 15 or 17");
    return MSBGX_EINVAL;
  }
  if ((t == 1 || t == 3 || t == 17) && B != 16) {
    setError("PDE smoothing: types 1, 3 and 17 use the reference's fixed 18-voxel halo rows (16^3 blocks only)");
    return MSBGX_EINVAL;
  }
  if (!color8 && (t == 2 || t == 4)) {
    setError("PDE smoothing: red-black passes with the mean-curvature or bi-Laplacian stencil read same-colour "
             "diagonal blocks while they are written; use the 8-colour scheme");
    return MSBGX_EINVAL;
  }
  if (t == 2 && B > 16) {
    setError("PDE smoothing: the bi-Laplacian's two-voxel halo block of a 32^3 block exceeds the LDS");
    return MSBGX_EINVAL;
  }
  if ((t == 14 || t == 15) && !(p.anis_diff_alpha != 0.0f && std::isfinite(p.anis_diff_alpha))) {
    setError("PDE smoothing: anisotropic diffusion needs a finite nonzero anis_diff_alpha");
    return MSBGX_EINVAL;
  }
  if (p.num_iter < 0 || !(p.time_step >= 0.0f) || !std::isfinite(p.time_step) || p.grid_spacing != 1.0f ||
      !std::isfinite(p.taubin) || !std::isfinite(p.dens_thr) || !std::isfinite(p.dens_thr_tol1) ||
      !std::isfinite(p.dens_thr_tol2)) {
    setError("PDE smoothing: num_iter >= 0, finite time_step >= 0, grid_spacing 1 and finite parameters required");
    return MSBGX_EINVAL;
  }
  a.typ = t;
  a.dt = p.time_step;
  a.doTaubin = fabsf(p.taubin) > 1e-20f;
  a.taubin = p.taubin;
  a.doHard = p.dens_thr > 1e-12f || p.dens_thr_tol1 > 1e-12f;
  a.tol1 = p.dens_thr_tol1;
  a.tol2 = p.dens_thr_tol2;
  a.clamp01 = p.do_constr_zero_one != 0;
  a.alpha = p.anis_diff_alpha;
  if (dtype == MSBGX_PDE_U16 && !a.doHard && !a.clamp01) {
    setError("PDE smoothing: a uint16 channel needs its values clamped to [0, 1] (do_constr_zero_one or hard "
             "constraints)");
    return MSBGX_EINVAL;
  }
  *nPass = color8 ? 8 : 2;
  return MSBGX_OK;
}

/* Host: upload the dense channel, run the colour passes, download. */
template <typename T>
static int pdeApply(const Grid &g, T *dense, const int *active, int nActive, const msbgx_pde_params &p, int dtype,
                    double *ms)
{
  const int B = g.bsx0, lb = __builtin_ctz(B);
  if (B > 32) {
    setError("PDE smoothing: block edge above 32");
    return MSBGX_EINVAL;
  }
  PdeArgs a{};
  int nPass = 0;
  if (int rc = pdeDecode(p.lapl_typ, B, dtype, p, a, &nPass)) return rc;
  std::vector<int> slot(g.nBlocks, -1), level0;
  for (int b = 0; b < g.nBlocks; b++)
    if (g.map[b] == 0) {
      slot[b] = (int)level0.size();
      level0.push_back(b);
    }
  /* pass lists of the listed level-0 blocks, in list order */
  std::vector<int> lists[8];
  for (int i = 0; i < nActive; i++) {
    const int b = active[i];
    if (b < 0 || b >= g.nBlocks) {
      setError("PDE smoothing: active block id out of range");
      return MSBGX_EINVAL;
    }
    if (slot[b] < 0) continue;   // only level-0 blocks are smoothed (src/msbg4.cpp:447)
    int bx, by, bz;
    g.blockCoords(b, bx, by, bz);
    if (nPass == 8) lists[(bx & 1) + ((by & 1) << 1) + ((bz & 1) << 2)].push_back(b);
    else lists[((bx + by + bz) & 1) ? 0 : 1].push_back(b);   // pass p updates colour != p
  }
  const int nIter = (p.num_iter % 2 == 1) ? p.num_iter + 1 : p.num_iter;
  const size_t nv = level0.size() * (size_t)B * B * B;
  hipStream_t s;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct Res {
    hipStream_t s;
    std::vector<void *> p;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Res()
    {
      (void)hipStreamSynchronize(s);
      for (void *q : p) (void)hipFree(q);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      (void)hipStreamDestroy(s);
    }
  } res{s, {}};
  T *dRho = nullptr, *dPhi0 = nullptr;
  int *dSlot = nullptr, *dLists = nullptr;
  HIPCHK(hipMalloc((void **)&dRho, std::max<size_t>(nv, 1) * sizeof(T)));
  res.p.push_back(dRho);
  if (a.doHard) {
    HIPCHK(hipMalloc((void **)&dPhi0, std::max<size_t>(nv, 1) * sizeof(T)));
    res.p.push_back(dPhi0);
  }
  HIPCHK(hipMalloc((void **)&dSlot, slot.size() * sizeof(int)));
  res.p.push_back(dSlot);
  size_t nl = 0;
  for (auto &l : lists) nl += l.size();
  HIPCHK(hipMalloc((void **)&dLists, std::max<size_t>(nl, 1) * sizeof(int)));
  res.p.push_back(dLists);
  HIPCHK(hipEventCreate(&res.e0));
  HIPCHK(hipEventCreate(&res.e1));
  /* dense -> natural level-0 layout */
  std::vector<T> nat(nv);
#pragma omp parallel for schedule(static)
  for (int li = 0; li < (int)level0.size(); li++) {
    int bx, by, bz;
    g.blockCoords(level0[li], bx, by, bz);
    for (int vz = 0; vz < B; vz++)
      for (int vy = 0; vy < B; vy++) {
        const T *src = dense + ((size_t)(bz * B + vz) * g.sy0 + (size_t)(by * B + vy)) * g.sx0 + (size_t)bx * B;
        std::copy(src, src + B, nat.data() + (size_t)li * B * B * B + (size_t)B * (vy + (size_t)B * vz));
      }
  }
  HIPCHK(hipMemcpy(dRho, nat.data(), nv * sizeof(T), hipMemcpyHostToDevice));
  if (dPhi0) HIPCHK(hipMemcpy(dPhi0, dRho, nv * sizeof(T), hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(dSlot, slot.data(), slot.size() * sizeof(int), hipMemcpyHostToDevice));
  std::vector<int> all;
  int start[9] = {0};
  for (int c = 0; c < nPass; c++) {
    start[c] = (int)all.size();
    all.insert(all.end(), lists[c].begin(), lists[c].end());
  }
  start[nPass] = (int)all.size();
  if (!all.empty()) HIPCHK(hipMemcpy(dLists, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
  a.slot = dSlot; a.nbx = g.nbx; a.nby = g.nby; a.nbz = g.nbz; a.lb = lb;
  a.n[0] = g.sx0; a.n[1] = g.sy0; a.n[2] = g.sz0;
  a.phi0 = dPhi0;
  HIPCHK(hipEventRecord(res.e0, s));
  pdeColourPasses<T>(B, a, dLists, start, nPass, nIter, dRho, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(res.e1, s));
  HIPCHK(hipEventSynchronize(res.e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, res.e0, res.e1));
  if (ms) *ms = t;
  HIPCHK(hipMemcpy(nat.data(), dRho, nv * sizeof(T), hipMemcpyDeviceToHost));
#pragma omp parallel for schedule(static)
  for (int li = 0; li < (int)level0.size(); li++) {
    int bx, by, bz;
    g.blockCoords(level0[li], bx, by, bz);
    for (int vz = 0; vz < B; vz++)
      for (int vy = 0; vy < B; vy++) {
        T *dst = dense + ((size_t)(bz * B + vz) * g.sy0 + (size_t)(by * B + vy)) * g.sx0 + (size_t)bx * B;
        const T *src = nat.data() + (size_t)li * B * B * B + (size_t)B * (vy + (size_t)B * vz);
        std::copy(src, src + B, dst);
      }
  }
  return MSBGX_OK;
}

int devicePdeApply(const Grid &g, int dtype, void *dense, const int *active, int nActive, const msbgx_pde_params &p,
                   double *ms)
{
  if (dtype == MSBGX_PDE_U16)
    return pdeApply<uint16_t>(g, static_cast<uint16_t *>(dense), active, nActive, p, dtype, ms);
  if (dtype == MSBGX_PDE_F32) return pdeApply<float>(g, static_cast<float *>(dense), active, nActive, p, dtype, ms);
  setError("PDE smoothing: dtype must be MSBGX_PDE_U16 or MSBGX_PDE_F32");
  return MSBGX_EINVAL;
}

/* The msbg_demo call (src/msbg_demo.cpp:716-727). */
static msbgx_pde_params pdeDemoParams(int numIter, float dt)
{
  msbgx_pde_params p{};
  p.lapl_typ = -(4 + kPdeOpt8Color);
  p.do_constr_zero_one = 1;
  p.num_iter = numIter;
  p.grid_spacing = 1.0f;
  p.time_step = dt;
  return p;
}

int devicePdeMeanCurvature(const Grid &g, uint16_t *dense, const int *active, int nActive, int numIter, float dt,
                           double *ms)
{
  return devicePdeApply(g, MSBGX_PDE_U16, dense, active, nActive, pdeDemoParams(numIter, dt), ms);
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
    setError("mean-curvature smoothing without allocation: block edge above 16 is not supported");
    return MSBGX_EINVAL;
  }
  PdeArgs a{};
  int nPass = 0;
  if (int rc = pdeDecode(-(4 + kPdeOpt8Color), B, MSBGX_PDE_U16, pdeDemoParams(numIter, dt), a, &nPass)) return rc;
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
  a.slot = dSlot; a.nbx = g.nbx; a.nby = g.nby; a.nbz = g.nbz; a.lb = lb;
  a.n[0] = g.sx0; a.n[1] = g.sy0; a.n[2] = g.sz0;
  pdeColourPasses<uint16_t>(B, a, dLists, st.start, nPass, nIter, dNat, s);
  if (nLevel0 > 0)
    hipLaunchKernelGGL(k_pde_na_copy<true>, dim3(nLevel0), dim3(256), 0, s, dLev0, g.nbx, g.nby, lb, g.sx0, g.sy0,
                       dDense, dNat);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dense, dDense, 2 * nDense, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return MSBGX_OK;
}

}  // namespace msbgx

