/*
 * Learned correction of the MG-PCG preconditioner (build addition; the
 * reference has no learned component).  Compiled as part of solver.hip,
 * whose device state, PCG kernels and launch helpers it uses.
 *
 * Network: four 3x3x3 convolutions without biases, channels 1 -> 16 -> 16 ->
 * 16 -> 1, ReLU after the first three, BF16 weights and activations, FP32
 * accumulation.  It runs per level-0 16^3 block on an 18^3 patch: the block
 * plus a one-voxel halo from the 26 neighbouring blocks (zero where the
 * neighbour is not a level-0 block or outside the raster); every layer pads
 * the patch with zeros ("same" convolutions) and the last layer's output on
 * the 16^3 block cells is the correction.  Without biases the network is
 * positively homogeneous, N(a r) = a N(r) for a > 0, so residual magnitudes
 * need no normalisation at inference.
 *
 * Use in PCG: z = V(r) + N(r - A V(r)), V the plain V-cycle, i.e. the network
 * estimates the error the V-cycle leaves.  While the correction is on, beta
 * is the flexible (Polak-Ribiere) one; a non-finite network output (masked to
 * zero where it occurs) or a stall of the residual (less than 10% reduction
 * over 5 iterations) switches the correction off for the rest of the solve
 * and restarts the directions: the plain MG-PCG is the fallback.
 *
 * Training pairs come from the plain MG-PCG itself: after iteration k the
 * error is e = x* - x_k (x* the converged solve), the network input is
 * rho = r_k - A V(r_k) and its target e - V(r_k).  Each sampled block's
 * patch and target are divided by the RMS of the target over the block's
 * unknowns (harmless by the homogeneity; it weights the samples equally).
 *
 * Storage.  A patch buffer holds 20^3 positions (the 18^3 patch plus a zero
 * ring) between 32 zero slack positions on either side, so that every tap
 * of every computed position reads inside the buffer.  Activations are
 * channels-last BF16 (16 channels = 32 bytes per position).  Positions of
 * the ring and of the first and last plane are never written non-zero.
 *
 * Kernels (all positions of planes 1..18; ring positions write zeros):
 *   k_nn_gather     rho of a block + halo -> BF16 patch (inference)
 *   k_nn_conv1      1 -> 16, VALU (27 MACs x 16 per position), ReLU
 *   k_nn_conv16     16 -> 16 on MFMA (v_mfma_f32_16x16x32_bf16: 16 positions
 *                   x 16 output channels x 2 taps of 16 input channels),
 *                   input planes staged in LDS; forward (ReLU) and backward
 *                   data gradient (tap-flipped, channel-transposed weights,
 *                   times the ReLU mask of the layer input)
 *   k_nn_out        16 -> 1 on the block cells, VALU from LDS; inference adds
 *                   the masked, finite correction to z
 *   coarse form (MSBGX_NET_COARSE): k_nn_block_mean, k_nn_gather_dense,
 *                   k_nn_out_dense, k_nn_prolong_blocks, k_nn_gamma -- the same
 *                   network on the grid of block means, its output extended
 *                   over each block and scaled by the Galerkin step
 *                   <zc, r> / <zc, A zc> before z = zc + V(r - A zc)
 *   training: k_nn_loss (masked MSE gradient), k_nn_bwd4, k_nn_dw16_mfma
 *   (16 -> 16 weight gradients on MFMA, one plane per workgroup),
 *   k_nn_dw_small, k_nn_reduce / k_nn_reduce_groups (fixed-order FP64 sums
 *   of per-workgroup partials), k_nn_adam, k_nn_pack
 */

namespace msbgx {

namespace nn {
constexpr int C = 16;                       // channels of the hidden layers
constexpr int E = 20;                       // padded patch edge
constexpr int PL = E * E;                   // positions per plane
constexpr int NPOS = E * E * E;
constexpr int SL = 32;                      // slack positions on either side
constexpr int PS = NPOS + 2 * SL;           // positions per patch buffer
constexpr int T = 27;                       // taps
constexpr int NJ = 14;                      // MFMA K-steps (two taps each)
constexpr int B = 16;                       // block edge
constexpr int NCELL = B * B * B;
constexpr int NW1 = T * C, NW2 = T * C * C, NW4 = T * C;
constexpr int OFF_W1 = 0, OFF_W2 = NW1, OFF_W3 = NW1 + NW2, OFF_W4 = NW1 + 2 * NW2;
constexpr int NPARAM = 2 * NW1 + 2 * NW2;   // 14688
constexpr int ZP = 2;                       // output planes per workgroup
constexpr int NZG = 18 / ZP;                // workgroups per block along z (conv layers)
constexpr int LDS_POS = (ZP + 2) * PL + 64; // staged positions per workgroup
// tap t = 9 (dz+1) + 3 (dy+1) + (dx+1)
__host__ __device__ constexpr int tapOff(int t) { return (t % 3 - 1) + E * ((t / 3) % 3 - 1) + PL * (t / 9 - 1); }
}  // namespace nn

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t nnF2bf(float v)
{
  const uint32_t b = __float_as_uint(v);
  if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);
  return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float nnBf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bool nnRing(int pos)
{
  const int u = pos % nn::E, v = (pos / nn::E) % nn::E;
  return u == 0 || u == nn::E - 1 || v == 0 || v == nn::E - 1;
}

/* Stage the positions [S0, S0 + LDS_POS) of a 16-channel patch into LDS;
 * S0 = SL + (z0 - 1) PL - 32 for output planes starting at patch plane z0. */
__device__ __forceinline__ void nnStage16(uint4 *lds, const uint16_t *__restrict__ src, int z0)
{
  const int64_t S0 = nn::SL + (int64_t)(z0 - 1) * nn::PL - 32;
  const uint4 *g = reinterpret_cast<const uint4 *>(src + S0 * nn::C);
  for (int c = threadIdx.x; c < nn::LDS_POS * 2; c += blockDim.x) lds[c] = g[c];
}

/* Patch of rho around level-0 block blocks[i] (inference). */
__global__ void __launch_bounds__(256) k_nn_gather(const float *__restrict__ src, const int *__restrict__ blocks,
                                                   const int *__restrict__ level, const int64_t *__restrict__ off,
                                                   int nbx, int nby, int nbz, uint16_t *__restrict__ X0,
                                                   const int *skip)
{
  if (skip && *skip) return;
  const int i = blockIdx.x, bid = blocks[i];
  const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
  uint16_t *dst = X0 + (int64_t)i * nn::PS + nn::SL;
  for (int pos = threadIdx.x; pos < nn::NPOS; pos += blockDim.x) {
    const int u = pos % nn::E, v = (pos / nn::E) % nn::E, w = pos / nn::PL;
    float val = 0.0f;
    if (u >= 1 && u <= 18 && v >= 1 && v <= 18 && w >= 1 && w <= 18) {
      const int cx = u - 2, cy = v - 2, cz = w - 2;   // cell coordinates, -1..16
      const int ox = cx < 0 ? -1 : (cx > 15 ? 1 : 0), oy = cy < 0 ? -1 : (cy > 15 ? 1 : 0),
                oz = cz < 0 ? -1 : (cz > 15 ? 1 : 0);
      const int nx = bx + ox, ny = by + oy, nz = bz + oz;
      if (nx >= 0 && ny >= 0 && nz >= 0 && nx < nbx && ny < nby && nz < nbz) {
        const int nb = nx + nbx * (ny + nby * nz);
        if (level[nb] == 0) {
          const int lx = cx - 16 * ox, ly = cy - 16 * oy, lz = cz - 16 * oz;
          val = src[off[nb] + lx + 16 * ly + 256 * lz];
        }
      }
    }
    dst[pos] = nnF2bf(val);
  }
}

/* Layer 1 (1 -> 16), ReLU.  X0 patch of sample idx[i] (or i). */
__global__ void __launch_bounds__(256) k_nn_conv1(const uint16_t *__restrict__ X0, const int *__restrict__ idx,
                                                  const float *__restrict__ w1, uint16_t *__restrict__ A1,
                                                  const int *skip)
{
  if (skip && *skip) return;
  __shared__ float w[nn::T * nn::C];
  for (int k = threadIdx.x; k < nn::T * nn::C; k += blockDim.x) w[k] = w1[k];
  __syncthreads();
  const int i = blockIdx.x, s = idx ? idx[i] : i;
  const uint16_t *x = X0 + (int64_t)s * nn::PS + nn::SL;
  uint16_t *out = A1 + ((int64_t)i * nn::PS + nn::SL) * nn::C;
  const int z0 = 1 + nn::ZP * blockIdx.y;
  for (int q = threadIdx.x; q < nn::ZP * nn::PL; q += blockDim.x) {
    const int pos = z0 * nn::PL + q;
    float acc[nn::C];
#pragma unroll
    for (int c = 0; c < nn::C; c++) acc[c] = 0.0f;
    const bool ring = nnRing(pos);
    if (!ring) {
#pragma unroll
      for (int t = 0; t < nn::T; t++) {
        const float xv = nnBf2f(x[pos + nn::tapOff(t)]);
#pragma unroll
        for (int c = 0; c < nn::C; c++) acc[c] = fmaf(w[t * nn::C + c], xv, acc[c]);
      }
    }
    uint32_t packed[nn::C / 2];
#pragma unroll
    for (int c = 0; c < nn::C; c += 2)
      packed[c / 2] = (uint32_t)nnF2bf(fmaxf(acc[c], 0.0f)) | ((uint32_t)nnF2bf(fmaxf(acc[c + 1], 0.0f)) << 16);
    uint4 *o = reinterpret_cast<uint4 *>(out + (int64_t)pos * nn::C);
    o[0] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
    o[1] = make_uint4(packed[4], packed[5], packed[6], packed[7]);
  }
}

/* 16 -> 16 convolution on MFMA.  BWD = false: dst = ReLU(conv(src, W)).
 * BWD = true: dst = conv(src, Wb) * (mask > 0) (data gradient through the
 * ReLU of the layer input `mask`).  wpk: per K-step j and lane the 8 BF16
 * B-operand values (see k_nn_pack). */
template <bool BWD>
__global__ void __launch_bounds__(256) k_nn_conv16(const uint16_t *__restrict__ src, const uint4 *__restrict__ wpk,
                                                   const uint16_t *__restrict__ mask, uint16_t *__restrict__ dst,
                                                   const int *skip)
{
  if (skip && *skip) return;
  __shared__ uint4 lds[nn::LDS_POS * 2];
  const int i = blockIdx.x;
  const int z0 = 1 + nn::ZP * blockIdx.y;
  const int64_t base = (int64_t)i * nn::PS * nn::C;
  nnStage16(lds, src + base, z0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  bf16x8_t wb[nn::NJ];
#pragma unroll
  for (int j = 0; j < nn::NJ; j++) wb[j] = __builtin_bit_cast(bf16x8_t, wpk[j * 64 + lane]);
  __syncthreads();
  const int m = lane & 15, kg = lane >> 4;
  const int ldsOrigin = (z0 - 1) * nn::PL - 32;   // patch position of lds[0]
  for (int tile = wave; tile < nn::ZP * 25; tile += 4) {
    const int pb = (z0 + tile / 25) * nn::PL + (tile % 25) * 16;
    f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int j = 0; j < nn::NJ; j++) {
      const int tap = 2 * j + (kg >> 1);
      bf16x8_t a;
      if (tap < nn::T) {
        const int lp = pb + m + nn::tapOff(tap) - ldsOrigin;
        a = __builtin_bit_cast(bf16x8_t, lds[lp * 2 + (kg & 1)]);
      } else {
        a = __builtin_bit_cast(bf16x8_t, make_uint4(0u, 0u, 0u, 0u));
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, wb[j], acc, 0, 0, 0);
    }
    const int co = m;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int pos = pb + 4 * kg + r;
      float v = acc[r];
      if (BWD) {
        const uint16_t mk = mask[base + (int64_t)(nn::SL + pos) * nn::C + co];
        v = (mk & 0x7fffu) && !(mk & 0x8000u) ? v : 0.0f;
      } else {
        v = fmaxf(v, 0.0f);
      }
      if (nnRing(pos)) v = 0.0f;
      dst[base + (int64_t)(nn::SL + pos) * nn::C + co] = nnF2bf(v);
    }
  }
}

/* Layer 4 (16 -> 1) on the 16^3 block cells of two cell planes per
 * workgroup.  Inference (out == nullptr): z[off[bid] + cell] += n on cells
 * that are unknowns (not SOLID/VOID, D != 0); a non-finite n is dropped and
 * flagged in st->bad.  Training: out[i][cell] = n (0 off the mask). */
__global__ void __launch_bounds__(256) k_nn_out(const uint16_t *__restrict__ A3, const float *__restrict__ w4,
                                                const int *__restrict__ blocks, const int64_t *__restrict__ off,
                                                const uint16_t *__restrict__ flags, const float *__restrict__ D,
                                                float *__restrict__ z, PcgState *st,
                                                const uint8_t *__restrict__ tmask, const int *__restrict__ idx,
                                                float *__restrict__ out)
{
  if (st && st->skipNet) return;
  __shared__ uint4 lds[nn::LDS_POS * 2];
  __shared__ float w[nn::T * nn::C];
  const int i = blockIdx.x;
  const int cz0 = 2 * blockIdx.y;       // first cell plane
  const int z0 = cz0 + 2;               // its patch plane
  const int64_t base = (int64_t)i * nn::PS * nn::C;
  nnStage16(lds, A3 + base, z0);
  for (int k = threadIdx.x; k < nn::T * nn::C; k += blockDim.x) w[k] = w4[k];
  __syncthreads();
  const int ldsOrigin = (z0 - 1) * nn::PL - 32;
  for (int q = threadIdx.x; q < 2 * 256; q += blockDim.x) {
    const int cx = q & 15, cy = (q >> 4) & 15, cz = cz0 + (q >> 8);
    const int pos = (cz + 2) * nn::PL + (cy + 2) * nn::E + (cx + 2);
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < nn::T; t++) {
      const int lp = pos + nn::tapOff(t) - ldsOrigin;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const uint4 v = lds[lp * 2 + h];
        const uint32_t wv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
          acc = fmaf(w[t * nn::C + 8 * h + 2 * e], nnBf2f((uint16_t)(wv[e] & 0xffffu)), acc);
          acc = fmaf(w[t * nn::C + 8 * h + 2 * e + 1], nnBf2f((uint16_t)(wv[e] >> 16)), acc);
        }
      }
    }
    const int cell = cx + 16 * cy + 256 * cz;
    if (out) {
      const int s = idx ? idx[i] : i;
      out[(int64_t)i * nn::NCELL + cell] = tmask[(int64_t)s * nn::NCELL + cell] ? acc : 0.0f;
    } else {
      const int64_t c = off[blocks[i]] + cell;
      const bool unknown = !(flags[c] & (CELL_SOLID | CELL_VOID)) && D[c] != 0.0f;
      if (unknown) {
        if (isfinite(acc)) z[c] += acc;
        else if (st) st->bad = 1;
      }
    }
  }
}

/* ------------------------------------------------------------ training */

/* dY = 2 (out - target) mask * invN and per-sample partial losses. */
__global__ void __launch_bounds__(256) k_nn_loss(const float *__restrict__ out, const float *__restrict__ target,
                                                 const uint8_t *__restrict__ tmask, const int *__restrict__ idx,
                                                 float invN, float *__restrict__ dY, double *__restrict__ lossPart)
{
  __shared__ double red[4];
  const int i = blockIdx.x, s = idx ? idx[i] : i;
  double acc = 0.0;
  for (int c = threadIdx.x; c < nn::NCELL; c += blockDim.x) {
    const int64_t o = (int64_t)i * nn::NCELL + c, so = (int64_t)s * nn::NCELL + c;
    const float d = tmask[so] ? out[o] - target[so] : 0.0f;
    dY[o] = 2.0f * d * invN;
    acc += (double)d * (double)d;
  }
  const double t = blockReduceSum(acc, red);
  if (threadIdx.x == 0) lossPart[i] = t * (double)invN;
}

/* dZ3 = (sum_t W4[t][ci] dY[cell(pos - off_t)]) * (A3 > 0) on planes 1..18. */
__global__ void __launch_bounds__(256) k_nn_bwd4(const float *__restrict__ dY, const float *__restrict__ w4,
                                                 const uint16_t *__restrict__ A3, uint16_t *__restrict__ G3)
{
  __shared__ float w[nn::T * nn::C];
  for (int k = threadIdx.x; k < nn::T * nn::C; k += blockDim.x) w[k] = w4[k];
  __syncthreads();
  const int i = blockIdx.x;
  const int z0 = 1 + nn::ZP * blockIdx.y;
  const int64_t base = (int64_t)i * nn::PS * nn::C;
  const float *dy = dY + (int64_t)i * nn::NCELL;
  for (int q = threadIdx.x; q < nn::ZP * nn::PL; q += blockDim.x) {
    const int pos = z0 * nn::PL + q;
    float acc[nn::C];
#pragma unroll
    for (int c = 0; c < nn::C; c++) acc[c] = 0.0f;
    if (!nnRing(pos)) {
      const int u = pos % nn::E, v = (pos / nn::E) % nn::E, ww = pos / nn::PL;
#pragma unroll
      for (int t = 0; t < nn::T; t++) {
        const int cx = u - (t % 3 - 1) - 2, cy = v - ((t / 3) % 3 - 1) - 2, cz = ww - (t / 9 - 1) - 2;
        if (cx < 0 || cy < 0 || cz < 0 || cx > 15 || cy > 15 || cz > 15) continue;
        const float g = dy[cx + 16 * cy + 256 * cz];
#pragma unroll
        for (int c = 0; c < nn::C; c++) acc[c] = fmaf(w[t * nn::C + c], g, acc[c]);
      }
    }
    const uint16_t *mk = A3 + base + (int64_t)(nn::SL + pos) * nn::C;
    uint16_t *o = G3 + base + (int64_t)(nn::SL + pos) * nn::C;
#pragma unroll
    for (int c = 0; c < nn::C; c++) {
      const uint16_t m = mk[c];
      o[c] = nnF2bf(((m & 0x7fffu) && !(m & 0x8000u)) ? acc[c] : 0.0f);
    }
  }
}

/* Per-workgroup partial weight gradient of a 16 -> 16 layer on MFMA:
 * dW[t][ci][co] = sum_pos A[pos + off_t][ci] dZ[pos][co] over one output plane
 * (grid: samples x 18 planes).  Per tap, a 16 x 16 (ci x co) tile accumulates
 * over K = 416 plane positions in 13 v_mfma_f32_16x16x32_bf16 (the ring
 * positions and 400..415 carry dZ = 0); the wave w owns the taps t = w mod 4.  A and dZ are
 * staged channel-major in LDS so a lane's 8 consecutive positions of one
 * channel are contiguous. */
constexpr int kDwTile = 1280;   // staged A positions: three planes, slack and the K padding
constexpr int kDwGroups = 64;   // first-level groups of the weight-gradient reduction
__global__ void __launch_bounds__(256) k_nn_dw16_mfma(const uint16_t *__restrict__ A, const uint16_t *__restrict__ dZ,
                                                      float *__restrict__ partial)
{
  __shared__ uint16_t At[nn::C * kDwTile];
  __shared__ __attribute__((aligned(16))) uint16_t Gt[nn::C * 416];
  const int i = blockIdx.x;
  const int z = 1 + blockIdx.y;                            // output plane 1..18
  const int64_t base = (int64_t)i * nn::PS * nn::C;
  const int p0 = nn::SL + (z - 1) * nn::PL - 32;           // buffer position of At[.][0]
  for (int e = threadIdx.x; e < kDwTile * 2; e += blockDim.x) {
    const int q = e >> 1, h = e & 1;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (p0 + q < nn::PS) v = reinterpret_cast<const uint4 *>(A + base + (int64_t)(p0 + q) * nn::C)[h];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      At[(8 * h + 2 * k) * kDwTile + q] = (uint16_t)(w[k] & 0xffffu);
      At[(8 * h + 2 * k + 1) * kDwTile + q] = (uint16_t)(w[k] >> 16);
    }
  }
  for (int e = threadIdx.x; e < 416 * 2; e += blockDim.x) {
    const int q = e >> 1, h = e & 1;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (q < nn::PL && !nnRing(q))
      v = reinterpret_cast<const uint4 *>(dZ + base + (int64_t)(nn::SL + z * nn::PL + q) * nn::C)[h];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      Gt[(8 * h + 2 * k) * 416 + q] = (uint16_t)(w[k] & 0xffffu);
      Gt[(8 * h + 2 * k + 1) * 416 + q] = (uint16_t)(w[k] >> 16);
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kg = lane >> 4;
  float *out = partial + ((int64_t)i * gridDim.y + blockIdx.y) * nn::NW2;
  for (int t = wave; t < nn::T; t += 4) {
    f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int aoff = 32 + nn::PL + nn::tapOff(t);          // At index of plane position 0 shifted by the tap
#pragma unroll
    for (int j = 0; j < 13; j++) {
      const int k0 = 32 * j + 8 * kg;
      uint16_t av[8];
#pragma unroll
      for (int e = 0; e < 8; e++) av[e] = At[m * kDwTile + aoff + k0 + e];
      uint4 apk;
      apk.x = av[0] | ((uint32_t)av[1] << 16); apk.y = av[2] | ((uint32_t)av[3] << 16);
      apk.z = av[4] | ((uint32_t)av[5] << 16); apk.w = av[6] | ((uint32_t)av[7] << 16);
      const uint4 bpk = *reinterpret_cast<const uint4 *>(&Gt[m * 416 + k0]);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, apk), __builtin_bit_cast(bf16x8_t, bpk),
                                                    acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) out[(t * nn::C + 4 * kg + r) * nn::C + m] = acc[r];
  }
}

/* Two-level fixed-order reduction of partial rows: grid (ceil(nEnt/256),
 * groups); group g sums rows g, g + groups, ... into part2[g][e]. */
__global__ void k_nn_reduce_groups(const float *__restrict__ partial, int nPart, int nEnt, float *__restrict__ part2)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nEnt) return;
  double s = 0.0;
  for (int p = blockIdx.y; p < nPart; p += gridDim.y) s += (double)partial[(int64_t)p * nEnt + e];
  part2[(int64_t)blockIdx.y * nEnt + e] = (float)s;
}

/* Per-workgroup partials of the two 432-entry gradients:
 *   FIRST: dW1[t][co] = sum_pos X0[pos + off_t] dZ1[pos][co]   (planes 1..18)
 *   LAST:  dW4[t][ci] = sum_cell A3[pos(cell) + off_t][ci] dY[cell]
 * The workgroup's slab (and the planes its taps reach) is staged in LDS
 * first; each thread then sums its entries over the slab's positions in
 * ascending order, so the sums are those of reading them from memory. */
template <bool LAST>
__global__ void __launch_bounds__(256) k_nn_dw_small(const uint16_t *__restrict__ X0, const int *__restrict__ idx,
                                                     const uint16_t *__restrict__ G1, const uint16_t *__restrict__ A3,
                                                     const float *__restrict__ dY, float *__restrict__ partial)
{
  constexpr int NST = 4 * nn::PL;   // staged positions: the slab's two planes and one on either side
  __shared__ uint4 lds[LAST ? NST * nn::C / 8 : (NST + 2 * nn::PL * nn::C) / 8];
  const int i = blockIdx.x;
  float acc[2] = {0.0f, 0.0f};
  int ent[2];
  for (int k = 0; k < 2; k++) ent[k] = threadIdx.x + 256 * k;
  if (!LAST) {
    const int s = idx ? idx[i] : i;
    const int z0 = 1 + nn::ZP * blockIdx.y;
    const uint16_t *x = X0 + (int64_t)s * nn::PS + nn::SL + (int64_t)(z0 - 1) * nn::PL;   // planes z0-1 .. z0+2
    const uint16_t *g = G1 + ((int64_t)i * nn::PS + nn::SL + (int64_t)z0 * nn::PL) * nn::C;   // planes z0, z0+1
    const uint4 *xs4 = reinterpret_cast<const uint4 *>(x), *gs4 = reinterpret_cast<const uint4 *>(g);
    for (int c = threadIdx.x; c < NST / 8; c += blockDim.x) lds[c] = xs4[c];
    for (int c = threadIdx.x; c < 2 * nn::PL * nn::C / 8; c += blockDim.x) lds[NST / 8 + c] = gs4[c];
    __syncthreads();
    const uint16_t *xl = reinterpret_cast<const uint16_t *>(lds);
    const uint16_t *gl = xl + NST;
    int xo[2], co[2];
    for (int k = 0; k < 2; k++) {
      const int t = ent[k] < nn::NW1 ? ent[k] / nn::C : 0;
      co[k] = ent[k] % nn::C;
      xo[k] = nn::PL + nn::tapOff(t);   // staged index of the tap's source for slab position 0
    }
    /* slab positions q = plane * PL + v * E + u in ascending order, ring positions skipped */
    for (int pz = 0; pz < nn::ZP; pz++)
      for (int v = 1; v < nn::E - 1; v++)
#pragma unroll
        for (int u = 1; u < nn::E - 1; u++) {
          const int q = pz * nn::PL + v * nn::E + u;
#pragma unroll
          for (int k = 0; k < 2; k++) {
            if (ent[k] >= nn::NW1) continue;
            acc[k] = fmaf(nnBf2f(xl[q + xo[k]]), nnBf2f(gl[q * nn::C + co[k]]), acc[k]);
          }
        }
  } else {
    const int cz0 = 2 * blockIdx.y;   // cell planes cz0, cz0 + 1 = patch planes cz0 + 2, cz0 + 3
    const uint16_t *a = A3 + ((int64_t)i * nn::PS + nn::SL + (int64_t)(cz0 + 1) * nn::PL) * nn::C;
    const uint4 *as4 = reinterpret_cast<const uint4 *>(a);
    for (int c = threadIdx.x; c < NST * nn::C / 8; c += blockDim.x) lds[c] = as4[c];
    __syncthreads();
    const uint16_t *al = reinterpret_cast<const uint16_t *>(lds);
    const float *dy = dY + (int64_t)i * nn::NCELL + 256 * cz0;
    int ao[2];
    for (int k = 0; k < 2; k++) {
      const int t = ent[k] < nn::NW4 ? ent[k] / nn::C : 0, ci = ent[k] % nn::C;
      ao[k] = (nn::PL + nn::tapOff(t)) * nn::C + ci;
    }
#pragma unroll 16
    for (int q = 0; q < 512; q++) {
      const int cx = q & 15, cy = (q >> 4) & 15, cz = q >> 8;
      const float gv = dy[q];
      const int pos = cz * nn::PL + (cy + 2) * nn::E + (cx + 2);   // staged plane index from cz0 + 2
#pragma unroll
      for (int k = 0; k < 2; k++) {
        if (ent[k] >= nn::NW4) continue;
        acc[k] = fmaf(nnBf2f(al[pos * nn::C + ao[k]]), gv, acc[k]);
      }
    }
  }
  float *o = partial + ((int64_t)i * gridDim.y + blockIdx.y) * nn::NW1;
  for (int k = 0; k < 2; k++)
    if (ent[k] < nn::NW1) o[ent[k]] = acc[k];
}

/* grad[e] = sum over nPart partial rows in FP64, in a fixed order: 64
 * entries per workgroup of kReduceWaves waves, wave w summing rows w,
 * w + kReduceWaves, ... for its lane's entry, then the waves' sums in wave
 * order.  Launch: grid ceil(nEnt / 64), block 64 * kReduceWaves. */
constexpr int kReduceWaves = 16;
__global__ void __launch_bounds__(64 * kReduceWaves) k_nn_reduce(const float *__restrict__ partial, int nPart,
                                                                  int nEnt, float *__restrict__ grad)
{
  __shared__ double red[kReduceWaves][64];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int e = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (e < nEnt)
    for (int p = w; p < nPart; p += kReduceWaves) s += (double)partial[(int64_t)p * nEnt + e];
  red[w][lane] = s;
  __syncthreads();
  if (w == 0 && e < nEnt) {
    double t = red[0][lane];
    for (int k = 1; k < kReduceWaves; k++) t += red[k][lane];
    grad[e] = (float)t;
  }
}

__global__ void k_nn_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.0f - b1) * gi;
  const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  w[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

/* BF16 forms of the FP32 parameters: w1, w4 as BF16-rounded floats, and the
 * MFMA B operands of layers 2 and 3, forward and backward.  Forward: lane l
 * of K-step j holds W[tap][ci0 + e][co], co = l % 16, tap = 2j + (l / 32),
 * ci0 = 8 ((l / 16) & 1).  Backward (data gradient = conv of dZ with the
 * tap-flipped, channel-transposed weights): W[26 - tap][l % 16][ci0 + e]. */
__global__ void k_nn_pack(const float *__restrict__ w, float *__restrict__ w1f, float *__restrict__ w4f,
                          uint16_t *__restrict__ p2, uint16_t *__restrict__ p3, uint16_t *__restrict__ b2,
                          uint16_t *__restrict__ b3)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nn::NW1) {
    w1f[e] = nnBf2f(nnF2bf(w[nn::OFF_W1 + e]));
    w4f[e] = nnBf2f(nnF2bf(w[nn::OFF_W4 + e]));
  }
  if (e < nn::NJ * 64 * 8) {
    const int j = e / 512, lane = (e / 8) % 64, k = e % 8;
    const int tap = 2 * j + (lane >> 5), ci0 = 8 * ((lane >> 4) & 1), n = lane & 15;
    float f2 = 0.0f, f3 = 0.0f, g2 = 0.0f, g3 = 0.0f;
    if (tap < nn::T) {
      f2 = w[nn::OFF_W2 + (tap * nn::C + ci0 + k) * nn::C + n];
      f3 = w[nn::OFF_W3 + (tap * nn::C + ci0 + k) * nn::C + n];
      g2 = w[nn::OFF_W2 + ((26 - tap) * nn::C + n) * nn::C + ci0 + k];
      g3 = w[nn::OFF_W3 + ((26 - tap) * nn::C + n) * nn::C + ci0 + k];
    }
    p2[e] = nnF2bf(f2); p3[e] = nnF2bf(f3); b2[e] = nnF2bf(g2); b3[e] = nnF2bf(g3);
  }
}

/* Training sample extraction: patch of rho (from src) and target
 * eps = err - zv on the block cells, both divided by the patch RMS of rho;
 * mask = cell is an unknown.  One workgroup per sampled block. */
__global__ void __launch_bounds__(256) k_nn_sample(const float *__restrict__ rho, const float *__restrict__ err,
                                                   const float *__restrict__ zv, const int *__restrict__ blocks,
                                                   const int *__restrict__ level, const int64_t *__restrict__ off,
                                                   const uint16_t *__restrict__ flags, const float *__restrict__ D,
                                                   int nbx, int nby, int nbz, uint16_t *__restrict__ X0,
                                                   float *__restrict__ tgt, uint8_t *__restrict__ tmask)
{
  __shared__ double red[4];
  __shared__ float scale;
  const int i = blockIdx.x, bid = blocks[i];
  const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
  auto fetch = [&](int pos) -> float {
    const int u = pos % nn::E, v = (pos / nn::E) % nn::E, w = pos / nn::PL;
    if (u < 1 || u > 18 || v < 1 || v > 18 || w < 1 || w > 18) return 0.0f;
    const int cx = u - 2, cy = v - 2, cz = w - 2;
    const int ox = cx < 0 ? -1 : (cx > 15 ? 1 : 0), oy = cy < 0 ? -1 : (cy > 15 ? 1 : 0),
              oz = cz < 0 ? -1 : (cz > 15 ? 1 : 0);
    const int nx = bx + ox, ny = by + oy, nz = bz + oz;
    if (nx < 0 || ny < 0 || nz < 0 || nx >= nbx || ny >= nby || nz >= nbz) return 0.0f;
    const int nb = nx + nbx * (ny + nby * nz);
    if (level[nb] != 0) return 0.0f;
    return rho[off[nb] + (cx - 16 * ox) + 16 * (cy - 16 * oy) + 256 * (cz - 16 * oz)];
  };
  /* scale = 1 / RMS of the target over the block's unknowns */
  double acc = 0.0, cnt = 0.0;
  for (int c = threadIdx.x; c < nn::NCELL; c += blockDim.x) {
    const int64_t o = off[bid] + c;
    if ((flags[o] & (CELL_SOLID | CELL_VOID)) || D[o] == 0.0f) continue;
    const double e = (double)err[o] - (double)zv[o];
    acc += e * e;
    cnt += 1.0;
  }
  const double ss = blockReduceSum(acc, red);
  const double nn_ = blockReduceSum(cnt, red);
  if (threadIdx.x == 0) scale = (ss > 0.0 && nn_ > 0.0) ? (float)(1.0 / sqrt(ss / nn_)) : 0.0f;
  __syncthreads();
  const float sc = scale;
  uint16_t *x = X0 + (int64_t)i * nn::PS + nn::SL;
  for (int pos = threadIdx.x; pos < nn::NPOS; pos += blockDim.x) x[pos] = nnF2bf(fetch(pos) * sc);
  for (int c = threadIdx.x; c < nn::NCELL; c += blockDim.x) {
    const int64_t o = off[bid] + c;
    const bool unknown = sc > 0.0f && !(flags[o] & (CELL_SOLID | CELL_VOID)) && D[o] != 0.0f;
    tgt[(int64_t)i * nn::NCELL + c] = unknown ? (err[o] - zv[o]) * sc : 0.0f;
    tmask[(int64_t)i * nn::NCELL + c] = unknown ? 1 : 0;
  }
}

/* a - b over [0, n) */
__global__ void k_nn_sub(const float *__restrict__ a, const float *__restrict__ b, float *__restrict__ out, int64_t n)
{
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] - b[i];
}

/* dst = src over [0, n) when *flag == want */
__global__ void k_nn_copy(const float *__restrict__ src, float *__restrict__ dst, int64_t n, const int *flag, int want)
{
  if (*flag != want) return;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

/* d = z + beta d, z = the plain V-cycle output when st->usePlain */
__global__ void __launch_bounds__(kVecThreads) k_nn_axpy(float *__restrict__ d, const float *__restrict__ z,
                                                         const float *__restrict__ zplain, int64_t n,
                                                         const PcgState *st)
{
  if (st->done) return;
  const float beta = st->beta;
  const float *zz = st->usePlain ? zplain : z;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = fmaf(beta, d[i], zz[i]);
}

/* ------------------------------------------------ coarse-correction form */

/* dst[i] = v over [0, n) unless *skip */
__global__ void k_nn_fill(float *__restrict__ dst, int n, float v, const int *skip)
{
  if (skip && *skip) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = v;
}

/* coarse[blocks[i]] = mean of src over the block's cells (all cells, or the
 * unknowns only with `unknownOnly`; 0 when there are none) */
__global__ void __launch_bounds__(256) k_nn_block_mean(const float *__restrict__ src, const int *__restrict__ blocks,
                                                       const int64_t *__restrict__ off,
                                                       const uint16_t *__restrict__ flags,
                                                       const float *__restrict__ D, int ncell, int unknownOnly,
                                                       float *__restrict__ coarse, const int *skip)
{
  if (skip && *skip) return;
  __shared__ double red[4];
  const int bid = blocks[blockIdx.x];
  const int64_t o = off[bid];
  double acc = 0.0, cnt = 0.0;
  if (unknownOnly) {
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
      if ((flags[o + c] & (CELL_SOLID | CELL_VOID)) || D[o + c] == 0.0f) continue;
      acc += (double)src[o + c];
      cnt += 1.0;
    }
  } else {   // every cell: the flags and D are not read
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) acc += (double)src[o + c];
  }
  const double sum = blockReduceSum(acc, red);
  const double num = unknownOnly ? blockReduceSum(cnt, red) : (double)ncell;
  if (threadIdx.x == 0) coarse[bid] = num > 0.0 ? (float)(sum / num) : 0.0f;
}

/* The coarse form's PCG update, one workgroup per PCG block (blocks[w],
 * its ncell cells at off[]): x += alpha d, r -= alpha q, the partial |r|^2
 * at partials[w] and, unless st->skipNet, the partial <r_new, z> of the z
 * about to be replaced at partials[gridDim.x + w] and the block's mean of
 * r_new into coarse[block] (k_nn_block_mean's sum, in its order), so that
 * the network's input needs no second pass over r. */
__global__ void __launch_bounds__(256) k_update_xr_rz_mean(float *__restrict__ x, float *__restrict__ r,
                                                           const float *__restrict__ d, const float *__restrict__ q,
                                                           const float *__restrict__ z, const int *__restrict__ blocks,
                                                           const int64_t *__restrict__ off, int ncell,
                                                           const PcgState *st, double *partials,
                                                           float *__restrict__ coarse)
{
  if (st->done) return;
  __shared__ double red[4];
  const int bid = blocks[blockIdx.x];
  const int64_t o = off[bid];
  const float alpha = st->alpha;
  const bool withZ = !st->skipNet;
  double acc = 0.0, acz = 0.0, am = 0.0;
  if (withZ) {
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
      const int64_t i = o + c;
      x[i] = fmaf(alpha, d[i], x[i]);
      const float rv = fmaf(-alpha, q[i], r[i]);
      r[i] = rv;
      acc += (double)rv * (double)rv;
      acz += (double)rv * (double)z[i];
      am += (double)rv;
    }
  } else {
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
      const int64_t i = o + c;
      x[i] = fmaf(alpha, d[i], x[i]);
      const float rv = fmaf(-alpha, q[i], r[i]);
      r[i] = rv;
      acc += (double)rv * (double)rv;
    }
  }
  const double s = blockReduceSum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
  if (withZ) {
    const double sz = blockReduceSum(acz, red);
    const double sm = blockReduceSum(am, red);
    if (threadIdx.x == 0) {
      partials[gridDim.x + blockIdx.x] = sz;
      coarse[bid] = (float)(sm / (double)ncell);
    }
  }
}

/* Patch of tile t of the dense block-resolution grid: tile (tx, ty, tz)
 * covers coarse cells 16 t - o .. 16 t - o + 15 per axis (o: tile origin
 * offset, 0 at inference, random in training so that the domain boundary
 * is not always at a tile edge); cells outside the raster read 0.
 * Optionally divides by scale[t]. */
__global__ void __launch_bounds__(256) k_nn_gather_dense(const float *__restrict__ coarse, int nbx, int nby, int nbz,
                                                         int ntx, int nty, int ox, int oy, int oz,
                                                         const float *__restrict__ scale, uint16_t *__restrict__ X0,
                                                         const int *skip)
{
  if (skip && *skip) return;
  const int t = blockIdx.x;
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  const float sc = scale ? scale[t] : 1.0f;
  uint16_t *dst = X0 + (int64_t)t * nn::PS + nn::SL;
  for (int pos = threadIdx.x; pos < nn::NPOS; pos += blockDim.x) {
    const int u = pos % nn::E, v = (pos / nn::E) % nn::E, w = pos / nn::PL;
    float val = 0.0f;
    if (u >= 1 && u <= 18 && v >= 1 && v <= 18 && w >= 1 && w <= 18) {
      const int cx = 16 * tx + u - 2 - ox, cy = 16 * ty + v - 2 - oy, cz = 16 * tz + w - 2 - oz;
      if (cx >= 0 && cy >= 0 && cz >= 0 && cx < nbx && cy < nby && cz < nbz)
        val = coarse[cx + nbx * (cy + nby * cz)] * sc;
    }
    dst[pos] = nnF2bf(val);
  }
}

/* Last layer on the tiles of the dense grid: out[cell] for the coarse cells
 * inside the raster (grid: tiles x 8 cell-plane pairs). */
__global__ void __launch_bounds__(256) k_nn_out_dense(const uint16_t *__restrict__ A3, const float *__restrict__ w4,
                                                      int nbx, int nby, int nbz, int ntx, int nty,
                                                      float *__restrict__ out, const int *skip)
{
  if (skip && *skip) return;
  __shared__ uint4 lds[nn::LDS_POS * 2];
  __shared__ float w[nn::T * nn::C];
  const int t = blockIdx.x;
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  const int cz0 = 2 * blockIdx.y, z0 = cz0 + 2;
  nnStage16(lds, A3 + (int64_t)t * nn::PS * nn::C, z0);
  for (int k = threadIdx.x; k < nn::T * nn::C; k += blockDim.x) w[k] = w4[k];
  __syncthreads();
  const int ldsOrigin = (z0 - 1) * nn::PL - 32;
  for (int q = threadIdx.x; q < 2 * 256; q += blockDim.x) {
    const int lx = q & 15, ly = (q >> 4) & 15, lz = cz0 + (q >> 8);
    const int pos = (lz + 2) * nn::PL + (ly + 2) * nn::E + (lx + 2);
    float acc = 0.0f;
#pragma unroll 3
    for (int tp = 0; tp < nn::T; tp++) {   // a few taps at a time: all 27 at once spill
      const int lp = pos + nn::tapOff(tp) - ldsOrigin;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const uint4 v = lds[lp * 2 + h];
        const uint32_t wv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
          acc = fmaf(w[tp * nn::C + 8 * h + 2 * e], nnBf2f((uint16_t)(wv[e] & 0xffffu)), acc);
          acc = fmaf(w[tp * nn::C + 8 * h + 2 * e + 1], nnBf2f((uint16_t)(wv[e] >> 16)), acc);
        }
      }
    }
    const int cx = 16 * tx + lx, cy = 16 * ty + ly, cz = 16 * tz + lz;
    if (cx < nbx && cy < nby && cz < nbz) out[cx + nbx * (cy + nby * cz)] = acc;
  }
}

/* Coarse grid at scale f: cell C covers the f^3 blocks f C .. f C + f - 1
 * of the block grid.  wgt[C] counts the PCG blocks (level-0 blocks with
 * unknowns) among them; cells with none take no part. */

/* coarse[C] = mean of fine[] over the PCG blocks of C's 2^3 children on the
 * next finer coarse grid (weights wf), 0 without any (skipped while *skip) */
__global__ void k_nn_coarsen(const float *__restrict__ fine, const float *__restrict__ wf, int fx, int fy, int fz,
                             float *__restrict__ coarse, int cx, int cy, int cz, const int *skip)
{
  if (skip && *skip) return;
  const int n = cx * cy * cz;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int x = i % cx, y = (i / cx) % cy, z = i / (cx * cy);
    float sw = 0.0f, sv = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int X = 2 * x + (k & 1), Y = 2 * y + ((k >> 1) & 1), Z = 2 * z + (k >> 2);
      if (X >= fx || Y >= fy || Z >= fz) continue;
      const int j = X + fx * (Y + fy * Z);
      sw += wf[j];
      sv = fmaf(wf[j], fine[j], sv);
    }
    coarse[i] = sw > 0.0f ? sv / sw : 0.0f;
  }
}

/* z = coarse[C] on the unknowns of each PCG block, C the scale-f cell that
 * holds the block (0 elsewhere in the block); a non-finite value is dropped
 * and flagged (st may be null: no skip flag, nothing flagged). */
__global__ void __launch_bounds__(256) k_nn_prolong_blocks(const float *__restrict__ coarse, int f, int nbx,
                                                           int nby, int sx, int sy, const int *__restrict__ blocks,
                                                           const int64_t *__restrict__ off,
                                                           const uint16_t *__restrict__ flags,
                                                           const float *__restrict__ D, int ncell,
                                                           float *__restrict__ zc, PcgState *st)
{
  if (st && st->skipNet) return;
  const int bid = blocks[blockIdx.x];
  const int64_t o = off[bid];
  const int bx = bid % nbx, by = (bid / nbx) % nby, bz = bid / (nbx * nby);
  float v = coarse[bx / f + sx * (by / f + sy * (bz / f))];
  if (!isfinite(v)) {
    if (threadIdx.x == 0 && st) st->bad = 1;
    v = 0.0f;
  }
  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const bool unk = !(flags[o + c] & (CELL_SOLID | CELL_VOID)) && D[o + c] != 0.0f;
    zc[o + c] = unk ? v : 0.0f;
  }
}

/* z on the unknowns of each PCG block by trilinear interpolation between
 * the centres of the scale-f cells (cell-centred: a cell at offset v in a
 * block at offset j = 0..f-1 inside its scale cell sits
 * (j B + v + 0.5) / (f B) - 0.5 scale cells from that cell's centre), the
 * weights renormalised over the cells with PCG blocks (wgt > 0) among the
 * eight nearest centres (air, coarser levels and cells outside the raster
 * take no part).  A non-finite value among them is dropped and flagged. */
__global__ void __launch_bounds__(256) k_nn_prolong_lin(const float *__restrict__ coarse,
                                                        const float *__restrict__ wgt, int f, int nbx, int nby,
                                                        int sx, int sy, int sz, const int *__restrict__ blocks,
                                                        const int64_t *__restrict__ off,
                                                        const uint16_t *__restrict__ flags,
                                                        const float *__restrict__ D, int bsx,
                                                        float *__restrict__ zc, PcgState *st)
{
  if (st && st->skipNet) return;
  __shared__ float cv[27], cm[27];
  const int bid = blocks[blockIdx.x];
  const int b3[3] = {bid % nbx, (bid / nbx) % nby, bid / (nbx * nby)};
  const int C[3] = {b3[0] / f, b3[1] / f, b3[2] / f};
  if (threadIdx.x < 27) {
    const int nx = C[0] + (int)threadIdx.x % 3 - 1, ny = C[1] + ((int)threadIdx.x / 3) % 3 - 1,
              nz = C[2] + (int)threadIdx.x / 9 - 1;
    float v = 0.0f, m = 0.0f;
    if (nx >= 0 && ny >= 0 && nz >= 0 && nx < sx && ny < sy && nz < sz) {
      const int nc = nx + sx * (ny + sy * nz);
      if (wgt[nc] > 0.0f) {
        v = coarse[nc];
        m = 1.0f;
        if (!isfinite(v)) {
          if (st) st->bad = 1;
          v = 0.0f;
        }
      }
    }
    cv[threadIdx.x] = v;
    cm[threadIdx.x] = m;
  }
  __syncthreads();
  const int64_t o = off[bid];
  const int ncell = bsx * bsx * bsx;
  const float inv = 1.0f / (float)(f * bsx);
  for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
    const int v3[3] = {c % bsx, (c / bsx) % bsx, c / (bsx * bsx)};
    int lo[3];
    float t[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int pos = (b3[a] - f * C[a]) * bsx + v3[a];      // offset inside the scale cell
      const float u = ((float)pos + 0.5f) * inv - 0.5f;      // in (-0.5, 0.5)
      lo[a] = u < 0.0f ? 0 : 1;                              // 27-neighbourhood index of the lower centre
      t[a] = u < 0.0f ? u + 1.0f : u;
    }
    float sw = 0.0f, sv = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int ix = lo[0] + (k & 1), iy = lo[1] + ((k >> 1) & 1), iz = lo[2] + (k >> 2);
      const float w = ((k & 1) ? t[0] : 1.0f - t[0]) * ((k & 2) ? t[1] : 1.0f - t[1]) *
                      ((k & 4) ? t[2] : 1.0f - t[2]);
      const int q = ix + 3 * iy + 9 * iz;
      sw = fmaf(w, cm[q], sw);
      sv = fmaf(w * cm[q], cv[q], sv);
    }
    const bool unk = !(flags[o + c] & (CELL_SOLID | CELL_VOID)) && D[o + c] != 0.0f;
    zc[o + c] = unk ? (sw > 0.0f ? sv / sw : cv[13]) : 0.0f;
  }
}

/* Workgroups of the coarse form's grid-stride vector passes (combination,
 * add with <r, z>, d update): 2048 runs them 1.4 % faster per learned
 * iteration at 1024^3 than 1024, and 4096 no faster than 2048. */
constexpr int kLVecBlocks = 2048;
static_assert(kLVecBlocks <= 4096, "the device partials hold at least 4096 values");

/* Directions of the coarse form at several scales: z[a] and q[a] = r - A z[a],
 * stored in FP32 or (T = uint16_t) as BF16 bits. */
constexpr int kMaxScales = 3;
template <typename T> struct DirSetT {
  const T *z[kMaxScales];
  const T *q[kMaxScales];
};
using DirSet = DirSetT<float>;
constexpr int kGramBlocks = 512;

/* unk[o + c] = 1 on the unknowns (not SOLID/VOID, D != 0) of PCG block
 * blocks[w]'s cells, 0 elsewhere in the block. */
__global__ void __launch_bounds__(256) k_nn_unk_mask(const int *__restrict__ blocks, const int64_t *__restrict__ off,
                                                     const uint16_t *__restrict__ flags, const float *__restrict__ D,
                                                     int ncell, uint8_t *__restrict__ unk)
{
  const int64_t o = off[blocks[blockIdx.x]];
  for (int c = threadIdx.x; c < ncell; c += blockDim.x)
    unk[o + c] = (!(flags[o + c] & (CELL_SOLID | CELL_VOID)) && D[o + c] != 0.0f) ? 1 : 0;
}

/* The scales' coarse grids, PCG-block weights and outputs for one pass of
 * k_nn_prolong_lin_b. */
struct ProlongScales {
  const float *coarse[kMaxScales];
  const float *wgt[kMaxScales];
  float *z[kMaxScales];
  uint16_t *z16[kMaxScales];   // the BF16 outputs of a T = uint16_t pass
  int f[kMaxScales], sx[kMaxScales], sy[kMaxScales], sz[kMaxScales];
};

/* k_nn_prolong_lin for NS scales at once on B^3 blocks (B = 16, 32), with the
 * unknowns from a byte mask (k_nn_unk_mask) read once for all of them
 * instead of the flags and D: thread t takes the (x, y)
 * columns t, t + 256, ... and walks z.  The interpolation factors are per
 * axis, so per column and scale the x/y corners of each of the three z
 * layers of the neighbourhood are summed once (masked weight and weighted
 * value); a cell then blends the two layers around it and divides.  The same
 * weights and mask as k_nn_prolong_lin, summed in another grouping (the
 * results agree to rounding).  T = uint16_t: each value is stored at
 * ps.z16[k] as BF16 bits (rounded to nearest even) instead. */
template <int B, int NS, typename T = float>
__global__ void __launch_bounds__(256) k_nn_prolong_lin_b(ProlongScales ps, int nbx, int nby,
                                                          const int *__restrict__ blocks,
                                                          const int64_t *__restrict__ off,
                                                          const uint8_t *__restrict__ unkMask, PcgState *st)
{
  static_assert(B * B % 256 == 0, "whole columns per thread");
  if (st && st->skipNet) return;
  __shared__ float cv[NS][27], cm[NS][27];
  const int bid = blocks[blockIdx.x];
  const int b3[3] = {bid % nbx, (bid / nbx) % nby, bid / (nbx * nby)};
  if (threadIdx.x < 27 * NS) {
    const int k = threadIdx.x / 27, j = threadIdx.x % 27, f = ps.f[k];
    const int nx = b3[0] / f + j % 3 - 1, ny = b3[1] / f + (j / 3) % 3 - 1, nz = b3[2] / f + j / 9 - 1;
    float v = 0.0f, m = 0.0f;
    if (nx >= 0 && ny >= 0 && nz >= 0 && nx < ps.sx[k] && ny < ps.sy[k] && nz < ps.sz[k]) {
      const int nc = nx + ps.sx[k] * (ny + ps.sy[k] * nz);
      if (ps.wgt[k][nc] > 0.0f) {
        v = ps.coarse[k][nc];
        m = 1.0f;
        if (!isfinite(v)) {
          if (st) st->bad = 1;
          v = 0.0f;
        }
      }
    }
    cv[k][j] = v;
    cm[k][j] = m;
  }
  __syncthreads();
  const int64_t o = off[bid];
  /* per scale: offset of the block inside its scale cell along each axis,
   * and 1 / (f B) */
  int base[NS][3];
  float inv[NS];
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const int f = ps.f[k];
#pragma unroll
    for (int a = 0; a < 3; a++) base[k][a] = (b3[a] - f * (b3[a] / f)) * B;
    inv[k] = 1.0f / (float)(f * B);
  }
  auto axis = [&](int k, int a, int v, int &lo, float &t) {
    const float u = ((float)(base[k][a] + v) + 0.5f) * inv[k] - 0.5f;
    lo = u < 0.0f ? 0 : 1;
    t = u < 0.0f ? u + 1.0f : u;
  };
  if constexpr (std::is_same_v<T, float>) {
    for (int col = threadIdx.x; col < B * B; col += 256) {
      const int x = col % B, y = col / B;
      float SW[NS][3], SV[NS][3];
  #pragma unroll
      for (int k = 0; k < NS; k++) {
        int lx, ly;
        float tx, ty;
        axis(k, 0, x, lx, tx);
        axis(k, 1, y, ly, ty);
  #pragma unroll
        for (int iz = 0; iz < 3; iz++) {
          float sw = 0.0f, sv = 0.0f;
  #pragma unroll
          for (int e = 0; e < 4; e++) {
            const float w = ((e & 1) ? tx : 1.0f - tx) * ((e & 2) ? ty : 1.0f - ty);
            const int q = lx + (e & 1) + 3 * (ly + ((e >> 1) & 1)) + 9 * iz;
            sw = fmaf(w, cm[k][q], sw);
            sv = fmaf(w * cm[k][q], cv[k][q], sv);
          }
          SW[k][iz] = sw;
          SV[k][iz] = sv;
        }
      }
  #pragma unroll 4
      for (int z = 0; z < B; z++) {
        const int64_t c = o + col + (int64_t)B * B * z;
        const bool unk = unkMask[c] != 0;
  #pragma unroll
        for (int k = 0; k < NS; k++) {
          int lz;
          float tz;
          axis(k, 2, z, lz, tz);
          const float w0 = SW[k][0], w1 = SW[k][1], w2 = SW[k][2], v0 = SV[k][0], v1 = SV[k][1], v2 = SV[k][2];
          const float swl = lz ? w1 : w0, swh = lz ? w2 : w1, svl = lz ? v1 : v0, svh = lz ? v2 : v1;
          const float sw = fmaf(tz, swh, (1.0f - tz) * swl), sv = fmaf(tz, svh, (1.0f - tz) * svl);
          const float zv = unk ? (sw > 0.0f ? sv / sw : cv[k][13]) : 0.0f;
          ps.z[k][c] = zv;
        }
      }
    }
  } else {
    /* BF16 outputs: two adjacent columns per thread, each scale's two values
     * stored as one 32-bit word (column pairs start at even cells) */
    for (int col = 2 * threadIdx.x; col < B * B; col += 2 * 256) {
      const int x0 = col % B, y = col / B;
      float SW[NS][2][3], SV[NS][2][3];
#pragma unroll
      for (int k = 0; k < NS; k++) {
        int ly;
        float ty;
        axis(k, 1, y, ly, ty);
#pragma unroll
        for (int p = 0; p < 2; p++) {
          int lx;
          float tx;
          axis(k, 0, x0 + p, lx, tx);
#pragma unroll
          for (int iz = 0; iz < 3; iz++) {
            float sw = 0.0f, sv = 0.0f;
#pragma unroll
            for (int e = 0; e < 4; e++) {
              const float w = ((e & 1) ? tx : 1.0f - tx) * ((e & 2) ? ty : 1.0f - ty);
              const int q = lx + (e & 1) + 3 * (ly + ((e >> 1) & 1)) + 9 * iz;
              sw = fmaf(w, cm[k][q], sw);
              sv = fmaf(w * cm[k][q], cv[k][q], sv);
            }
            SW[k][p][iz] = sw;
            SV[k][p][iz] = sv;
          }
        }
      }
#pragma unroll 4
      for (int z = 0; z < B; z++) {
        const int64_t c = o + col + (int64_t)B * B * z;
        const uint16_t u2 = *reinterpret_cast<const uint16_t *>(unkMask + c);
        const bool unk[2] = {(u2 & 0xffu) != 0, (u2 >> 8) != 0};
#pragma unroll
        for (int k = 0; k < NS; k++) {
          int lz;
          float tz;
          axis(k, 2, z, lz, tz);
          uint32_t word = 0;
#pragma unroll
          for (int p = 0; p < 2; p++) {
            const float w0 = SW[k][p][0], w1 = SW[k][p][1], w2 = SW[k][p][2];
            const float v0 = SV[k][p][0], v1 = SV[k][p][1], v2 = SV[k][p][2];
            const float swl = lz ? w1 : w0, swh = lz ? w2 : w1, svl = lz ? v1 : v0, svh = lz ? v2 : v1;
            const float sw = fmaf(tz, swh, (1.0f - tz) * swl), sv = fmaf(tz, svh, (1.0f - tz) * svl);
            const float zv = unk[p] ? (sw > 0.0f ? sv / sw : cv[k][13]) : 0.0f;
            word |= (uint32_t)bf16Bits(zv) << (16 * p);
          }
          *reinterpret_cast<uint32_t *>(ps.z16[k] + c) = word;
        }
      }
    }
  }
}

/* Partial Galerkin sums over [0, n) in FP64, unless *skip: for a <= b,
 * <z_a, A z_b> = <z_a, r - q_b>, then <z_a, r>; value j of workgroup w at
 * partials[j * gridDim.x + w]. */
template <int K>
__global__ void __launch_bounds__(256) k_nn_gram(DirSet ds, const float *__restrict__ r, int64_t n,
                                                 double *partials, const int *skip)
{
  if (*skip) return;
  constexpr int NG = K * (K + 1) / 2;
  __shared__ double red[4];
  double acc[NG + K];
#pragma unroll
  for (int j = 0; j < NG + K; j++) acc[j] = 0.0;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double ri = r[i];
    double zv[K], aq[K];
#pragma unroll
    for (int a = 0; a < K; a++) {
      zv[a] = ds.z[a][i];
      aq[a] = ri - (double)ds.q[a][i];
    }
    int j = 0;
#pragma unroll
    for (int a = 0; a < K; a++)
#pragma unroll
      for (int b = a; b < K; b++) acc[j++] += zv[a] * aq[b];
#pragma unroll
    for (int a = 0; a < K; a++) acc[NG + a] += zv[a] * ri;
  }
#pragma unroll
  for (int j = 0; j < NG + K; j++) {
    const double t = blockReduceSum(acc[j], red);
    if (threadIdx.x == 0) partials[j * gridDim.x + blockIdx.x] = t;
  }
}

/* Galerkin coefficients c = G^-1 g of the K directions (G_ab = <z_a, A z_b>,
 * g_a = <z_a, r>) into st->coef by Cholesky in FP64; a direction whose pivot
 * is not positive beyond rounding (zero or dependent on the ones before it)
 * gets coefficient 0. */
template <int K>
__global__ void __launch_bounds__(256) k_nn_gram_solve(const double *partials, int nPart, PcgState *st)
{
  if (st->done || st->skipNet) return;
  constexpr int NG = K * (K + 1) / 2;
  __shared__ double v[NG + K], red[4];
  /* every sum in one pass over the partials (each in sumPartials' order) */
  double acc[NG + K];
#pragma unroll
  for (int j = 0; j < NG + K; j++) acc[j] = 0.0;
  for (int i = threadIdx.x; i < nPart; i += blockDim.x)
#pragma unroll
    for (int j = 0; j < NG + K; j++) acc[j] += partials[(size_t)j * nPart + i];
#pragma unroll
  for (int j = 0; j < NG + K; j++) {
    const double t = blockReduceSum(acc[j], red);
    if (threadIdx.x == 0) v[j] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double G[K][K], L[K][K], y[K], c[K];
  bool use[K];
  int j = 0;
  for (int a = 0; a < K; a++)
    for (int b = a; b < K; b++) G[a][b] = G[b][a] = v[j++];
  for (int a = 0; a < K; a++) {
    use[a] = G[a][a] > 1e-300 && isfinite(G[a][a]) && isfinite(v[NG + a]);
    for (int b = 0; b < K; b++) L[a][b] = 0.0;
  }
  for (int a = 0; a < K; a++) {
    if (!use[a]) continue;
    double d = G[a][a];
    for (int k = 0; k < a; k++)
      if (use[k]) d -= L[a][k] * L[a][k];
    if (!(d > 1e-10 * G[a][a]) || !isfinite(d)) {
      use[a] = false;
      continue;
    }
    L[a][a] = sqrt(d);
    for (int b = a + 1; b < K; b++) {
      if (!use[b]) continue;
      double e = G[b][a];
      for (int k = 0; k < a; k++)
        if (use[k]) e -= L[b][k] * L[a][k];
      L[b][a] = e / L[a][a];
    }
  }
  for (int a = 0; a < K; a++) {
    y[a] = 0.0;
    if (!use[a]) continue;
    double e = v[NG + a];
    for (int k = 0; k < a; k++)
      if (use[k]) e -= L[a][k] * y[k];
    y[a] = e / L[a][a];
  }
  for (int a = K - 1; a >= 0; a--) {
    c[a] = 0.0;
    if (!use[a]) continue;
    double e = y[a];
    for (int k = a + 1; k < K; k++)
      if (use[k]) e -= L[k][a] * c[k];
    c[a] = e / L[a][a];
  }
  for (int a = 0; a < K; a++) st->coef[a] = isfinite(c[a]) ? (float)c[a] : 0.0f;
}

/* rs = r - A zc = (1 - sum_a c_a) r + sum_a c_a q_a over [0, nDir) for
 * zc = sum_a c_a z_a, which k_nn_add_dot forms where it adds it (q[0] may be
 * rs), and rs = r over [nDir, n) (the MG0 pool has size n); with the
 * network off, rs = r exactly (see k_nn_gamma). */
template <int K, typename T = float>
__global__ void __launch_bounds__(256) k_nn_combine(DirSetT<T> ds, float *rs, const float *__restrict__ r, int64_t n,
                                                    int64_t nDir, const PcgState *st)
{
  if (st->done) return;
  const int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, di = (int64_t)gridDim.x * blockDim.x;
  if (st->skipNet) {
#pragma unroll 4
    for (int64_t i = i0; i < n; i += di) rs[i] = r[i];
    return;
  }
  float cf[K], cs = 0.0f;
#pragma unroll
  for (int a = 0; a < K; a++) {
    cf[a] = st->coef[a];
    cs += cf[a];
  }
#pragma unroll 4
  for (int64_t i = i0; i < nDir; i += di) {
    float rv = (1.0f - cs) * r[i];
#pragma unroll
    for (int a = 0; a < K; a++) rv = fmaf(cf[a], vload(ds.q[a], i), rv);
    rs[i] = rv;
  }
  for (int64_t i = nDir + i0; i < n; i += di) rs[i] = r[i];
}

/* z += zc over [0, n) unless *skip */
__global__ void k_nn_add(float *__restrict__ z, const float *__restrict__ zc, int64_t n, const int *skip)
{
  if (skip && *skip) return;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    z[i] += zc[i];
}

/* z += zc unless st->skipNet, then (unless st->done) the partial <r, z> of
 * the result, its products summed in k_dot's order.  zc: the scaled
 * direction ds.z[0] (K = 1, k_nn_gamma), else sum_a c_a z_a formed here. */
template <int K, typename T = float>
__global__ void __launch_bounds__(kVecThreads) k_nn_add_dot(float *__restrict__ z, DirSetT<T> ds,
                                                            const float *__restrict__ r, int64_t n,
                                                            const PcgState *st, double *partials)
{
  if (st->done) return;
  __shared__ double red[kVecThreads / 64];
  const bool add = !st->skipNet;
  float cf[K];
#pragma unroll
  for (int a = 0; a < K; a++) cf[a] = K == 1 ? 1.0f : st->coef[a];
  double acc = 0.0;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float zv = z[i];
    if (add) {
      float zc;
      if constexpr (K == 1) {
        zc = vload(ds.z[0], i);
      } else {
        zc = 0.0f;
#pragma unroll
        for (int a = 0; a < K; a++) zc = fmaf(cf[a], vload(ds.z[a], i), zc);
      }
      zv += zc;
      z[i] = zv;
    }
    acc += (double)r[i] * (double)zv;
  }
  const double s = blockReduceSum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

/* Scale the correction by the Galerkin step gamma over the whole MG0 pool:
 * zc *= gamma and rs = r - gamma A zc = (1 - gamma) r + gamma rs.  With the
 * network off (plain iterations, fallback) rs = r exactly, so the one V-cycle
 * of the iteration, which reads rs, gives the plain MG-PCG's z. */
__global__ void k_nn_gamma(float *__restrict__ zc, float *__restrict__ rs, const float *__restrict__ r, int64_t n,
                           const PcgState *st)
{
  if (st->done) return;
  const int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, di = (int64_t)gridDim.x * blockDim.x;
  if (st->skipNet) {
#pragma unroll 4
    for (int64_t i = i0; i < n; i += di) rs[i] = r[i];
    return;
  }
  const float gm = st->gamma;
#pragma unroll 4
  for (int64_t i = i0; i < n; i += di) {
    zc[i] = gm * zc[i];
    rs[i] = fmaf(gm, rs[i] - r[i], r[i]);
  }
}

/* partial <a, b - c> over [0, n) (FP64), unless *skip */
__global__ void __launch_bounds__(kVecThreads) k_nn_dot_diff(const float *__restrict__ a, const float *__restrict__ b,
                                                             const float *__restrict__ c, int64_t n,
                                                             double *partials, const int *skip)
{
  if (*skip) return;
  __shared__ double red[kVecThreads / 64];
  double acc = 0.0;
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc += (double)a[i] * ((double)b[i] - (double)c[i]);
  const double t = blockReduceSum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

/* Coarse-form training sample of tile t: input patch of the dense residual
 * means, target = dense error means on the cells with PCG blocks
 * (cmask > 0; mask 1 there), both divided by the target's RMS over the
 * tile. */
__global__ void __launch_bounds__(256) k_nn_tile_target(const float *__restrict__ cerr,
                                                        const float *__restrict__ cmask, int nbx, int nby, int nbz,
                                                        int ntx, int nty, int ox, int oy, int oz,
                                                        float *__restrict__ scale, float *__restrict__ tgt,
                                                        uint8_t *__restrict__ tmask)
{
  __shared__ double red[4];
  __shared__ float sc;
  const int t = blockIdx.x;
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  auto cellOf = [&](int c, int &idx) {
    const int cx = 16 * tx + (c & 15) - ox, cy = 16 * ty + ((c >> 4) & 15) - oy, cz = 16 * tz + (c >> 8) - oz;
    idx = cx + nbx * (cy + nby * cz);
    return cx >= 0 && cy >= 0 && cz >= 0 && cx < nbx && cy < nby && cz < nbz && cmask[idx] > 0.0f;
  };
  double acc = 0.0, cnt = 0.0;
  for (int c = threadIdx.x; c < nn::NCELL; c += blockDim.x) {
    int idx;
    if (!cellOf(c, idx)) continue;
    acc += (double)cerr[idx] * (double)cerr[idx];
    cnt += 1.0;
  }
  const double ss = blockReduceSum(acc, red);
  const double nm = blockReduceSum(cnt, red);
  if (threadIdx.x == 0) {
    sc = (ss > 0.0 && nm > 0.0) ? (float)(1.0 / sqrt(ss / nm)) : 0.0f;
    scale[t] = sc;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < nn::NCELL; c += blockDim.x) {
    int idx;
    const bool ok = cellOf(c, idx) && sc > 0.0f;
    tgt[(int64_t)t * nn::NCELL + c] = ok ? cerr[idx] * sc : 0.0f;
    tmask[(int64_t)t * nn::NCELL + c] = ok ? 1 : 0;
  }
}

/* ------------------------------------------------------------------ host */

struct NetBuffers {
  int cap = 0;                                   // blocks
  uint16_t *X0 = nullptr, *A1 = nullptr, *A2 = nullptr, *A3 = nullptr;
  uint16_t *G1 = nullptr, *G2 = nullptr, *G3 = nullptr;  // training only
  float *out = nullptr, *dY = nullptr;
  float *part = nullptr;                         // dW partials
  double *lossPart = nullptr;
  int *idx = nullptr;
};

struct Net {
  uint64_t serial = 0;                           // unique per created net (graph keys)
  int mode = MSBGX_NET_FINE;                     // MSBGX_NET_FINE or MSBGX_NET_COARSE
  int prolong = MSBGX_NET_PROLONG_CONSTANT;      // coarse form: extension of the block values
  int scales = 1;                                // coarse form: block grid coarsened 1x, 2x, 4x
  int stallWin = 5;                              // fallback when |r| drops by less than
  float stallDrop = 0.1f;                        // stallDrop over stallWin iterations
  hipStream_t stream = nullptr;
  std::vector<float> host;                       // FP32 parameters (host mirror)
  float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  float *w1f = nullptr, *w4f = nullptr;
  uint16_t *p2 = nullptr, *p3 = nullptr, *b2 = nullptr, *b3 = nullptr;
  int adamStep = 0;
  /* training set */
  int nS = 0, capS = 0;
  uint16_t *setX = nullptr;
  float *setT = nullptr;
  uint8_t *setM = nullptr;
  std::vector<int> maskCount;
  NetBuffers tb;                                 // training work buffers
  double lastTrainMs = 0.0;
};

static void freeBuffers(NetBuffers &b)
{
  auto f = [](void *p) { if (p) (void)hipFree(p); };
  f(b.X0); f(b.A1); f(b.A2); f(b.A3); f(b.G1); f(b.G2); f(b.G3); f(b.out); f(b.dY); f(b.part); f(b.lossPart); f(b.idx);
  b = NetBuffers{};
}

/* Patch buffers for `n` blocks, zeroed (ring, first/last plane and slack stay zero). */
static int allocBuffers(NetBuffers &b, int n, bool training, hipStream_t s)
{
  if (b.cap >= n && (!training || b.G1)) return MSBGX_OK;
  freeBuffers(b);
  const size_t act = (size_t)n * nn::PS * nn::C;
  if (int rc = dalloc(&b.X0, (size_t)n * nn::PS)) return rc;
  if (int rc = dalloc(&b.A1, act)) return rc;
  if (int rc = dalloc(&b.A2, act)) return rc;
  if (int rc = dalloc(&b.A3, act)) return rc;
  HIPCHK(hipMemsetAsync(b.X0, 0, (size_t)n * nn::PS * 2, s));
  HIPCHK(hipMemsetAsync(b.A1, 0, act * 2, s));
  HIPCHK(hipMemsetAsync(b.A2, 0, act * 2, s));
  HIPCHK(hipMemsetAsync(b.A3, 0, act * 2, s));
  if (training) {
    if (int rc = dalloc(&b.G1, act)) return rc;
    if (int rc = dalloc(&b.G2, act)) return rc;
    if (int rc = dalloc(&b.G3, act)) return rc;
    HIPCHK(hipMemsetAsync(b.G1, 0, act * 2, s));
    HIPCHK(hipMemsetAsync(b.G2, 0, act * 2, s));
    HIPCHK(hipMemsetAsync(b.G3, 0, act * 2, s));
    if (int rc = dalloc(&b.out, (size_t)n * nn::NCELL)) return rc;
    if (int rc = dalloc(&b.dY, (size_t)n * nn::NCELL)) return rc;
    if (int rc = dalloc(&b.part, (size_t)(n * 18 + kDwGroups) * nn::NW2)) return rc;
    if (int rc = dalloc(&b.lossPart, n)) return rc;
    if (int rc = dalloc(&b.idx, n)) return rc;
  }
  HIPCHK(hipStreamSynchronize(s));
  b.cap = n;
  return MSBGX_OK;
}

static int netPack(Net &n)
{
  hipLaunchKernelGGL(k_nn_pack, dim3((nn::NJ * 64 * 8 + 255) / 256), dim3(256), 0, n.stream, n.w, n.w1f, n.w4f, n.p2,
                     n.p3, n.b2, n.b3);
  HIPCHK(hipGetLastError());
  return MSBGX_OK;
}

int netCreate(Net **out, uint64_t seed)
{
  static std::atomic<uint64_t> nextSerial{1};
  Net *n = new Net();
  n->serial = nextSerial++;
  *out = n;
  HIPCHK(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
  if (int rc = dalloc(&n->w, nn::NPARAM)) return rc;
  if (int rc = dalloc(&n->m, nn::NPARAM)) return rc;
  if (int rc = dalloc(&n->v, nn::NPARAM)) return rc;
  if (int rc = dalloc(&n->grad, nn::NPARAM)) return rc;
  if (int rc = dalloc(&n->w1f, nn::NW1)) return rc;
  if (int rc = dalloc(&n->w4f, nn::NW4)) return rc;
  const size_t npk = nn::NJ * 64 * 8;
  if (int rc = dalloc(&n->p2, npk)) return rc;
  if (int rc = dalloc(&n->p3, npk)) return rc;
  if (int rc = dalloc(&n->b2, npk)) return rc;
  if (int rc = dalloc(&n->b3, npk)) return rc;
  /* He-normal initialisation from a splitmix64 stream (Box-Muller) */
  n->host.resize(nn::NPARAM);
  uint64_t st = seed;
  auto uni = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  auto normal = [&]() { return std::sqrt(-2.0 * std::log(uni())) * std::cos(6.283185307179586 * uni()); };
  const int fan[4] = {nn::T, nn::T * nn::C, nn::T * nn::C, nn::T * nn::C};
  const int beg[4] = {nn::OFF_W1, nn::OFF_W2, nn::OFF_W3, nn::OFF_W4};
  const int len[4] = {nn::NW1, nn::NW2, nn::NW2, nn::NW4};
  for (int l = 0; l < 4; l++) {
    const double sd = std::sqrt((l == 3 ? 1.0 : 2.0) / fan[l]);
    for (int e = 0; e < len[l]; e++) n->host[beg[l] + e] = (float)(sd * normal());
  }
  HIPCHK(hipMemcpy(n->w, n->host.data(), nn::NPARAM * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(n->m, 0, nn::NPARAM * sizeof(float)));
  HIPCHK(hipMemset(n->v, 0, nn::NPARAM * sizeof(float)));
  if (int rc = netPack(*n)) return rc;
  HIPCHK(hipStreamSynchronize(n->stream));
  return MSBGX_OK;
}

void netDestroy(Net *n)
{
  if (!n) return;
  auto f = [](void *p) { if (p) (void)hipFree(p); };
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  f(n->w); f(n->m); f(n->v); f(n->grad); f(n->w1f); f(n->w4f); f(n->p2); f(n->p3); f(n->b2); f(n->b3);
  f(n->setX); f(n->setT); f(n->setM);
  freeBuffers(n->tb);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

int netSetWeights(Net &n, const float *w)
{
  n.host.assign(w, w + nn::NPARAM);
  HIPCHK(hipMemcpy(n.w, w, nn::NPARAM * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(n.m, 0, nn::NPARAM * sizeof(float)));
  HIPCHK(hipMemset(n.v, 0, nn::NPARAM * sizeof(float)));
  n.adamStep = 0;
  if (int rc = netPack(n)) return rc;
  HIPCHK(hipStreamSynchronize(n.stream));
  return MSBGX_OK;
}

int netGetWeights(Net &n, float *w)
{
  HIPCHK(hipMemcpy(w, n.w, nn::NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return MSBGX_OK;
}

/* Forward of layers 1-3 for `nb` patches (X0 of sample idx[i] or i). */
static void netForward123(Net &n, const uint16_t *X0, const int *idx, NetBuffers &b, int nb, const int *skip,
                          hipStream_t s)
{
  hipLaunchKernelGGL(k_nn_conv1, dim3(nb, nn::NZG), dim3(256), 0, s, X0, idx, n.w1f, b.A1, skip);
  hipLaunchKernelGGL(k_nn_conv16<false>, dim3(nb, nn::NZG), dim3(256), 0, s, b.A1, (const uint4 *)n.p2,
                     (const uint16_t *)nullptr, b.A2, skip);
  hipLaunchKernelGGL(k_nn_conv16<false>, dim3(nb, nn::NZG), dim3(256), 0, s, b.A2, (const uint4 *)n.p3,
                     (const uint16_t *)nullptr, b.A3, skip);
}

/* dW of a 16 -> 16 layer: MFMA partials per (sample, plane), reduced in a
 * fixed order (groups of rows, then the groups). */
static void netWeightGrad16(const uint16_t *A, const uint16_t *dZ, NetBuffers &b, int nb, float *grad,
                            hipStream_t s)
{
  const int rows = nb * 18;
  float *part2 = b.part + (size_t)rows * nn::NW2;
  hipLaunchKernelGGL(k_nn_dw16_mfma, dim3(nb, 18), dim3(256), 0, s, A, dZ, b.part);
  hipLaunchKernelGGL(k_nn_reduce_groups, dim3((nn::NW2 + 255) / 256, kDwGroups), dim3(256), 0, s, b.part, rows,
                     nn::NW2, part2);
  hipLaunchKernelGGL(k_nn_reduce, dim3((nn::NW2 + 63) / 64), dim3(64 * kReduceWaves), 0, s, part2, kDwGroups, nn::NW2,
                     grad);
}

/* Loss and gradient over `nb` samples of the training set addressed by the
 * device index list b.idx (already uploaded).  Leaves grad in n.grad and the
 * per-sample loss partials in b.lossPart. */
static void netLossGrad(Net &n, NetBuffers &b, int nb, float invN)
{
  hipStream_t s = n.stream;
  netForward123(n, n.setX, b.idx, b, nb, nullptr, s);
  hipLaunchKernelGGL(k_nn_out, dim3(nb, 8), dim3(256), 0, s, b.A3, n.w4f, (const int *)nullptr,
                     (const int64_t *)nullptr, (const uint16_t *)nullptr, (const float *)nullptr, (float *)nullptr,
                     (PcgState *)nullptr, n.setM, b.idx, b.out);
  hipLaunchKernelGGL(k_nn_loss, dim3(nb), dim3(256), 0, s, b.out, n.setT, n.setM, b.idx, invN, b.dY, b.lossPart);
  /* layer 4 */
  hipLaunchKernelGGL(k_nn_dw_small<true>, dim3(nb, 8), dim3(256), 0, s, (const uint16_t *)nullptr,
                     (const int *)nullptr, (const uint16_t *)nullptr, b.A3, b.dY, b.part);
  hipLaunchKernelGGL(k_nn_reduce, dim3((nn::NW4 + 63) / 64), dim3(64 * kReduceWaves), 0, s, b.part, nb * 8, nn::NW4,
                     n.grad + nn::OFF_W4);
  hipLaunchKernelGGL(k_nn_bwd4, dim3(nb, nn::NZG), dim3(256), 0, s, b.dY, n.w4f, b.A3, b.G3);
  /* layer 3: dW3 = corr(A2, G3); G2 = conv(G3, W3b) * (A2 > 0) */
  netWeightGrad16(b.A2, b.G3, b, nb, n.grad + nn::OFF_W3, s);
  hipLaunchKernelGGL(k_nn_conv16<true>, dim3(nb, nn::NZG), dim3(256), 0, s, b.G3, (const uint4 *)n.b3, b.A2, b.G2,
                     (const int *)nullptr);
  /* layer 2 */
  netWeightGrad16(b.A1, b.G2, b, nb, n.grad + nn::OFF_W2, s);
  hipLaunchKernelGGL(k_nn_conv16<true>, dim3(nb, nn::NZG), dim3(256), 0, s, b.G2, (const uint4 *)n.b2, b.A1, b.G1,
                     (const int *)nullptr);
  /* layer 1 */
  hipLaunchKernelGGL(k_nn_dw_small<false>, dim3(nb, nn::NZG), dim3(256), 0, s, n.setX, b.idx, b.G1,
                     (const uint16_t *)nullptr, (const float *)nullptr, b.part);
  hipLaunchKernelGGL(k_nn_reduce, dim3((nn::NW1 + 63) / 64), dim3(64 * kReduceWaves), 0, s, b.part, nb * nn::NZG,
                     nn::NW1,
                     n.grad + nn::OFF_W1);
}

static int growSet(Net &n, int need)
{
  if (need <= n.capS) return MSBGX_OK;
  int cap = std::max(need, 2 * n.capS);
  uint16_t *x = nullptr;
  float *t = nullptr;
  uint8_t *m = nullptr;
  if (int rc = dalloc(&x, (size_t)cap * nn::PS)) return rc;
  if (int rc = dalloc(&t, (size_t)cap * nn::NCELL)) return rc;
  if (int rc = dalloc(&m, (size_t)cap * nn::NCELL)) return rc;
  HIPCHK(hipMemsetAsync(x, 0, (size_t)cap * nn::PS * 2, n.stream));
  if (n.nS) {
    HIPCHK(hipMemcpyAsync(x, n.setX, (size_t)n.nS * nn::PS * 2, hipMemcpyDeviceToDevice, n.stream));
    HIPCHK(hipMemcpyAsync(t, n.setT, (size_t)n.nS * nn::NCELL * 4, hipMemcpyDeviceToDevice, n.stream));
    HIPCHK(hipMemcpyAsync(m, n.setM, (size_t)n.nS * nn::NCELL, hipMemcpyDeviceToDevice, n.stream));
  }
  HIPCHK(hipStreamSynchronize(n.stream));
  auto f = [](void *p) { if (p) (void)hipFree(p); };
  f(n.setX); f(n.setT); f(n.setM);
  n.setX = x; n.setT = t; n.setM = m; n.capS = cap;
  return MSBGX_OK;
}

/* Host-provided samples: x[count][18^3] (patch, x fastest), target and mask
 * [count][16^3].  Replaces the training set. */
int netSetSamples(Net &n, int count, const float *x, const float *target, const uint8_t *mask)
{
  n.nS = 0;
  if (int rc = growSet(n, count)) return rc;
  std::vector<uint16_t> px((size_t)count * nn::PS, 0);
  auto f2bf = [](float v) {
    uint32_t b;
    memcpy(&b, &v, 4);
    if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);
    return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
  };
  for (int i = 0; i < count; i++)
    for (int w = 1; w <= 18; w++)
      for (int v = 1; v <= 18; v++)
        for (int u = 1; u <= 18; u++)
          px[(size_t)i * nn::PS + nn::SL + w * nn::PL + v * nn::E + u] =
              f2bf(x[(size_t)i * 5832 + (u - 1) + 18 * (v - 1) + 324 * (w - 1)]);
  HIPCHK(hipMemcpy(n.setX, px.data(), px.size() * 2, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(n.setT, target, (size_t)count * nn::NCELL * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(n.setM, mask, (size_t)count * nn::NCELL, hipMemcpyHostToDevice));
  n.maskCount.assign(count, 0);
  for (int i = 0; i < count; i++)
    for (int c = 0; c < nn::NCELL; c++) n.maskCount[i] += mask[(size_t)i * nn::NCELL + c] ? 1 : 0;
  n.nS = count;
  return MSBGX_OK;
}

/* Loss (mean over masked cells), gradient and outputs over the whole set. */
int netLossGradAll(Net &n, double *loss, float *grad, float *out)
{
  const int nb = n.nS;
  if (nb == 0) { setError("no samples"); return MSBGX_ESTATE; }
  if (int rc = allocBuffers(n.tb, nb, true, n.stream)) return rc;
  std::vector<int> idx(nb);
  long cnt = 0;
  for (int i = 0; i < nb; i++) { idx[i] = i; cnt += n.maskCount[i]; }
  HIPCHK(hipMemcpyAsync(n.tb.idx, idx.data(), nb * sizeof(int), hipMemcpyHostToDevice, n.stream));
  netLossGrad(n, n.tb, nb, cnt ? 1.0f / (float)cnt : 0.0f);
  HIPCHK(hipGetLastError());
  std::vector<double> lp(nb);
  HIPCHK(hipMemcpyAsync(lp.data(), n.tb.lossPart, nb * sizeof(double), hipMemcpyDeviceToHost, n.stream));
  if (grad) HIPCHK(hipMemcpyAsync(grad, n.grad, nn::NPARAM * sizeof(float), hipMemcpyDeviceToHost, n.stream));
  if (out) HIPCHK(hipMemcpyAsync(out, n.tb.out, (size_t)nb * nn::NCELL * 4, hipMemcpyDeviceToHost, n.stream));
  HIPCHK(hipStreamSynchronize(n.stream));
  double L = 0.0;
  for (double v : lp) L += v;
  *loss = L;
  return MSBGX_OK;
}

int netTrain(Net &n, int steps, int batch, float lr, uint64_t seed, double *lossFirst, double *lossLast)
{
  if (n.nS == 0) { setError("no training samples"); return MSBGX_ESTATE; }
  batch = std::min(batch, n.nS);
  if (int rc = allocBuffers(n.tb, batch, true, n.stream)) return rc;
  uint64_t st = seed ? seed : 1;
  auto next = [&]() {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return st;
  };
  std::vector<int> idx(batch);
  std::vector<double> lp(batch);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, n.stream));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double first = 0.0, last = 0.0, ema = 0.0;
  for (int it = 0; it < steps; it++) {
    long cnt = 0;
    for (int i = 0; i < batch; i++) {
      idx[i] = (int)(next() % (uint64_t)n.nS);
      cnt += n.maskCount[idx[i]];
    }
    HIPCHK(hipMemcpyAsync(n.tb.idx, idx.data(), batch * sizeof(int), hipMemcpyHostToDevice, n.stream));
    netLossGrad(n, n.tb, batch, cnt ? 1.0f / (float)cnt : 0.0f);
    n.adamStep++;
    const float c1 = 1.0f - std::pow(b1, (float)n.adamStep), c2 = 1.0f - std::pow(b2, (float)n.adamStep);
    hipLaunchKernelGGL(k_nn_adam, dim3((nn::NPARAM + 255) / 256), dim3(256), 0, n.stream, n.w, n.grad, n.m, n.v,
                       nn::NPARAM, lr, b1, b2, eps, c1, c2);
    if (int rc = netPack(n)) return rc;
    HIPCHK(hipMemcpyAsync(lp.data(), n.tb.lossPart, batch * sizeof(double), hipMemcpyDeviceToHost, n.stream));
    HIPCHK(hipStreamSynchronize(n.stream));   // idx and lp are reused next step
    double L = 0.0;
    for (double v : lp) L += v;
    ema = it == 0 ? L : 0.95 * ema + 0.05 * L;
    if (it == 0) first = L;
    last = ema;
  }
  HIPCHK(hipEventRecord(e1, n.stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  n.lastTrainMs = ms;
  HIPCHK(hipMemcpy(n.host.data(), n.w, nn::NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  if (lossFirst) *lossFirst = first;
  if (lossLast) *lossLast = last;
  return MSBGX_OK;
}

int netNumSamples(const Net &n) { return n.nS; }
int netSetMode(Net &n, int mode)
{
  if (mode != n.mode) {
    n.mode = mode;
    n.nS = 0;               // samples of the other form are not reused
    n.maskCount.clear();
    /* default stall tests: the fine correction is not step-length
     * controlled and is dropped early; the coarse one is scaled by its
     * Galerkin step and only a real stall switches it off */
    n.stallWin = mode == MSBGX_NET_COARSE ? 32 : 5;
    n.stallDrop = mode == MSBGX_NET_COARSE ? 0.02f : 0.1f;
  }
  return MSBGX_OK;
}
int netMode(const Net &n) { return n.mode; }
int netSetProlongation(Net &n, int kind)
{
  n.prolong = kind;
  return MSBGX_OK;
}
int netProlongation(const Net &n) { return n.prolong; }
int netSetScales(Net &n, int k)
{
  if (k != n.scales) {
    n.scales = k;
    n.nS = 0;               // samples are drawn at every scale the net works on
    n.maskCount.clear();
  }
  return MSBGX_OK;
}
int netScales(const Net &n) { return n.scales; }
int netSetStall(Net &n, int window, float drop)
{
  n.stallWin = window;
  n.stallDrop = drop;
  return MSBGX_OK;
}
double netLastTrainMs(const Net &n) { return n.lastTrainMs; }

/* Measurement: `reps` launches of the forward 16 -> 16 MFMA layer over `nb`
 * patches (the network's current weights, zero activations) between HIP
 * events on the network's stream; algorithmic FLOP per launch counts the
 * 16^3 cells of each patch, 27 taps x 16 x 16 channels x 2:
 * nb * 4096 * 27 * 16 * 16 * 2 (the layer computes the 18^3 positions a
 * following layer's halo needs, 1.42x that). */
int netTimeConv(Net &n, int nb, int reps, double *msPerLaunch, double *flopPerLaunch)
{
  NetBuffers b;
  if (int rc = allocBuffers(b, nb, false, n.stream)) return rc;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (int i = 0; i < 2; i++)
    hipLaunchKernelGGL(k_nn_conv16<false>, dim3(nb, nn::NZG), dim3(256), 0, n.stream, b.A1, (const uint4 *)n.p2,
                       (const uint16_t *)nullptr, b.A2, (const int *)nullptr);
  HIPCHK(hipEventRecord(e0, n.stream));
  for (int i = 0; i < reps; i++)
    hipLaunchKernelGGL(k_nn_conv16<false>, dim3(nb, nn::NZG), dim3(256), 0, n.stream, b.A1, (const uint4 *)n.p2,
                       (const uint16_t *)nullptr, b.A2, (const int *)nullptr);
  HIPCHK(hipEventRecord(e1, n.stream));
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipGetLastError());
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  freeBuffers(b);
  *msPerLaunch = ms / reps;
  *flopPerLaunch = (double)nb * 4096.0 * nn::T * nn::C * nn::C * 2.0;
  return MSBGX_OK;
}

/* ------------------------------------------------- PCG with the correction */

struct LearnedCtx {
  NetBuffers ib;                // inference buffers for the level-0 PCG blocks (or the coarse tiles)
  float *zprev = nullptr, *zplain = nullptr;
  /* coarse form: per scale f = 1, 2, 4 the dense grid of f^3-block cells
   * (dims, 16^3 tiles covering it, network input and output, PCG blocks
   * per cell) and its direction z with q = r - A z (MG0 pool size); scale
   * 0's direction and residual are the correction zc and rs themselves */
  struct Scale {
    int f = 1, sx = 0, sy = 0, sz = 0, ntx = 0, nty = 0, ntz = 0, ncoarse = 0;
    float *cin = nullptr, *cout = nullptr, *w = nullptr, *z = nullptr, *q = nullptr;
  } sc[kMaxScales];
  int nScales = 0;                 // scales allocated
  float *zc = nullptr, *rs = nullptr;
  uint8_t *unk = nullptr;          // 1 on the unknowns of the PCG blocks' cells [0, n0)
  /* two directions on 32^3 level-0 blocks: z_0, z_1, q_0, q_1 as BF16 bits,
   * each over the MG0 pool (written on [0, n0), 0 beyond) */
  uint16_t *dir16 = nullptr;
  double *gram = nullptr;          // Galerkin partials
  int64_t n0 = 0;
  hipGraphExec_t graph = nullptr;
  GraphKey graphKey;               // includes the serial of the net whose buffers the graph uses
};

static void freeLearned(LearnedCtx *c)
{
  if (!c) return;
  freeBuffers(c->ib);
  if (c->zprev) (void)hipFree(c->zprev);
  if (c->zplain) (void)hipFree(c->zplain);
  for (int k = 0; k < c->nScales; k++) {
    for (float *p : {c->sc[k].cin, c->sc[k].cout, c->sc[k].w})
      if (p) (void)hipFree(p);
    if (k > 0)
      for (float *p : {c->sc[k].z, c->sc[k].q})
        if (p) (void)hipFree(p);
  }
  for (float *p : {c->zc, c->rs})
    if (p) (void)hipFree(p);
  if (c->unk) (void)hipFree(c->unk);
  if (c->dir16) (void)hipFree(c->dir16);
  if (c->gram) (void)hipFree(c->gram);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  delete c;
}

/* z += N(r - A z) on the level-0 PCG blocks (skipped while st->skipNet). */
static void netCorrect(Grid &g, Net &n, LearnedCtx &c, hipStream_t s)
{
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  const int *skip = &d.state->skipNet;
  const int nb = g.nActive0;
  if (nb == 0) return;
  multiplyLaunch(g, 0, true, m0.d_ch[C_Z], m0.d_ch[C_R], m0.d_ch[C_Q], nullptr, skip, s);
  hipLaunchKernelGGL(k_nn_gather, dim3(nb), dim3(256), 0, s, m0.d_ch[C_Q], d.dActive0, m0.d_level, m0.d_off, g.nbx,
                     g.nby, g.nbz, c.ib.X0, skip);
  netForward123(n, c.ib.X0, nullptr, c.ib, nb, skip, s);
  hipLaunchKernelGGL(k_nn_out, dim3(nb, 8), dim3(256), 0, s, c.ib.A3, n.w4f, d.dActive0, m0.d_off, m0.d_flags, m0.d_D,
                     m0.d_ch[C_Z], d.state, (const uint8_t *)nullptr, (const int *)nullptr, (float *)nullptr);
}

static int coarseSetup(Grid &g, LearnedCtx &c, int scales, hipStream_t s);
static void coarseNetwork(Grid &g, Net &n, LearnedCtx &c, const float *src, PcgState *st, hipStream_t s,
                          bool meanDone = false, bool dir16 = false);

/* Per-kernel entry point on the level-0 PCG blocks: fine form
 * ch_out += N(ch_in); coarse form ch_out += sum over the net's scales of
 * P_f N(R_f ch_in) (R_f the mean over the PCG blocks of each f^3-block
 * cell, P_f the net's prolongation from those cells). */
int netApply(Net &n, Grid &g, int chIn, int chOut)
{
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  if (!d.learned) d.learned = new LearnedCtx();
  LearnedCtx &c = *d.learned;
  const int nb = g.nActive0;
  if (nb == 0) return MSBGX_OK;
  if (n.mode == MSBGX_NET_COARSE) {
    if (int rc = coarseSetup(g, c, n.scales, d.stream)) return rc;
    HIPCHK(hipStreamSynchronize(n.stream));
    coarseNetwork(g, n, c, m0.d_ch[chIn], nullptr, d.stream);
    for (int k = 0; k < n.scales; k++)
      hipLaunchKernelGGL(k_nn_add, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[chOut], c.sc[k].z,
                         (int64_t)m0.ncells, (const int *)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(d.stream));
    return MSBGX_OK;
  }
  if (int rc = allocBuffers(c.ib, nb, false, d.stream)) return rc;
  HIPCHK(hipStreamSynchronize(n.stream));
  hipLaunchKernelGGL(k_nn_gather, dim3(nb), dim3(256), 0, d.stream, m0.d_ch[chIn], d.dActive0, m0.d_level, m0.d_off,
                     g.nbx, g.nby, g.nbz, c.ib.X0, (const int *)nullptr);
  netForward123(n, c.ib.X0, nullptr, c.ib, nb, nullptr, d.stream);
  hipLaunchKernelGGL(k_nn_out, dim3(nb, 8), dim3(256), 0, d.stream, c.ib.A3, n.w4f, d.dActive0, m0.d_off, m0.d_flags,
                     m0.d_D, m0.d_ch[chOut], (PcgState *)nullptr, (const uint8_t *)nullptr, (const int *)nullptr,
                     (float *)nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(d.stream));
  return MSBGX_OK;
}

/* Buffers of the coarse form on grid g for the first `scales` scales: the
 * dense grids, their tiles and PCG block counts, the directions and the
 * residuals they leave, zc and rs. */
/* Whether the two-direction pass may keep its directions in BF16: 32^3
 * level-0 blocks that are all PCG blocks (the fused residual pass's case). */
static bool coarseDir16Shape(const Grid &g)
{
  return g.bsx0 == 32 && g.mg[0].dAct[0].n == g.nActive0 && g.nActive0 > 0;
}
/* Whether this iteration's K directions go through BF16 storage. */
static bool coarseDir16(const Grid &g, const LearnedCtx &c, int K)
{
  return K == 2 && c.dir16 && coarseDir16Shape(g) &&
         !(g.options & (MSBGX_OPT_UNFUSED_COARSE | MSBGX_OPT_FP32_DIRECTIONS));
}

static int coarseSetup(Grid &g, LearnedCtx &c, int scales, hipStream_t s)
{
  const size_t nc = (size_t)g.mg[0].ncells;
  if (scales >= 2 && !c.dir16 && coarseDir16Shape(g)) {
    HIPCHK(hipStreamSynchronize(s));
    if (int rc = dalloc(&c.dir16, 4 * nc)) return rc;
    HIPCHK(hipMemset(c.dir16, 0, 4 * nc * sizeof(uint16_t)));
  }
  if (c.nScales >= scales) return MSBGX_OK;
  HIPCHK(hipStreamSynchronize(s));
  if (!c.zc) {
    if (int rc = dalloc(&c.zc, nc)) return rc;
    if (int rc = dalloc(&c.rs, nc)) return rc;
    /* Galerkin partials: per workgroup of k_nn_gram, or per level-0 block of the fused residual pass */
    const size_t nPart = (size_t)std::max(kGramBlocks, g.nActive0);
    if (int rc = dalloc(&c.gram, (size_t)(kMaxScales * (kMaxScales + 3) / 2) * nPart)) return rc;
    HIPCHK(hipMemset(c.zc, 0, nc * sizeof(float)));
    HIPCHK(hipMemset(c.rs, 0, nc * sizeof(float)));
    /* the PCG blocks' unknowns as bytes, for the prolongation */
    const int64_t n0 = (int64_t)g.nActive0 * g.bsx0 * g.bsx0 * g.bsx0;
    if (int rc = dalloc(&c.unk, (size_t)std::max<int64_t>(n0, 1))) return rc;
    if (g.nActive0)
      hipLaunchKernelGGL(k_nn_unk_mask, dim3(g.nActive0), dim3(256), 0, s, dev(g).dActive0, g.mg[0].d_off,
                         g.mg[0].d_flags, g.mg[0].d_D, g.bsx0 * g.bsx0 * g.bsx0, c.unk);
  }
  /* PCG blocks per cell, scale by scale */
  std::vector<float> w((size_t)g.nbx * g.nby * g.nbz, 0.0f);
  for (int b : g.active0) w[b] = 1.0f;
  int px = g.nbx, py = g.nby, pz = g.nbz;
  for (int k = 0; k < scales; k++) {
    LearnedCtx::Scale &q = c.sc[k];
    if (k > 0) {
      const int cx = (px + 1) / 2, cy = (py + 1) / 2, cz = (pz + 1) / 2;
      std::vector<float> wc((size_t)cx * cy * cz, 0.0f);
      for (int z = 0; z < pz; z++)
        for (int y = 0; y < py; y++)
          for (int x = 0; x < px; x++) wc[x / 2 + cx * (y / 2 + cy * (z / 2))] += w[x + px * (y + py * z)];
      w.swap(wc);
      px = cx;
      py = cy;
      pz = cz;
    }
    if (k < c.nScales) continue;
    q.f = 1 << k;
    q.sx = px;
    q.sy = py;
    q.sz = pz;
    q.ntx = (px + 15) / 16;
    q.nty = (py + 15) / 16;
    q.ntz = (pz + 15) / 16;
    q.ncoarse = px * py * pz;
    if (int rc = dalloc(&q.cin, q.ncoarse)) return rc;
    if (int rc = dalloc(&q.cout, q.ncoarse)) return rc;
    if (int rc = dalloc(&q.w, q.ncoarse)) return rc;
    HIPCHK(hipMemcpy(q.w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    if (k == 0) {
      q.z = c.zc;
      q.q = c.rs;
    } else {
      if (int rc = dalloc(&q.z, nc)) return rc;
      if (int rc = dalloc(&q.q, nc)) return rc;
      HIPCHK(hipMemset(q.z, 0, nc * sizeof(float)));
      HIPCHK(hipMemset(q.q, 0, nc * sizeof(float)));
    }
    c.nScales = k + 1;
  }
  /* the network buffers are shared by the scales, one after another */
  if (int rc = allocBuffers(c.ib, c.sc[0].ntx * c.sc[0].nty * c.sc[0].ntz, false, s)) return rc;
  HIPCHK(hipStreamSynchronize(s));
  return MSBGX_OK;
}

/* Directions of the coarse form, z_f = P_f N(R_f src) for the net's scales
 * f = 1, 2, 4: R_1 the per-block mean over all the block's cells on the
 * dense block grid (0 for blocks without PCG unknowns), R_2f the mean of
 * R_f over the PCG blocks of each 2^3 cells, N the network over 16^3 tiles
 * of each grid, P_f the net's prolongation onto the PCG blocks' unknowns
 * (skipped while st->skipNet; st may be null).  meanDone: R_1 src is in
 * place already (k_update_xr_rz_mean; the grid's other cells are still 0
 * from the first call). */
static void coarseNetwork(Grid &g, Net &n, LearnedCtx &c, const float *src, PcgState *st, hipStream_t s,
                          bool meanDone, bool dir16)
{
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  const int *skip = st ? &st->skipNet : nullptr;
  const int ncell = g.bsx0 * g.bsx0 * g.bsx0;
  if (g.nActive0 == 0) return;
  LearnedCtx::Scale &s0 = c.sc[0];
  if (!meanDone) {
    hipLaunchKernelGGL(k_nn_fill, dim3((s0.ncoarse + 255) / 256), dim3(256), 0, s, s0.cin, s0.ncoarse, 0.0f, skip);
    hipLaunchKernelGGL(k_nn_block_mean, dim3(g.nActive0), dim3(256), 0, s, src, d.dActive0, m0.d_off, m0.d_flags,
                       m0.d_D, ncell, 0, s0.cin, skip);
  }
  for (int k = 0; k < n.scales; k++) {
    LearnedCtx::Scale &q = c.sc[k];
    if (k > 0) {
      const LearnedCtx::Scale &p = c.sc[k - 1];
      hipLaunchKernelGGL(k_nn_coarsen, dim3((q.ncoarse + 255) / 256), dim3(256), 0, s, p.cin, p.w, p.sx, p.sy, p.sz,
                         q.cin, q.sx, q.sy, q.sz, skip);
    }
    const int nt = q.ntx * q.nty * q.ntz;
    hipLaunchKernelGGL(k_nn_gather_dense, dim3(nt), dim3(256), 0, s, q.cin, q.sx, q.sy, q.sz, q.ntx, q.nty, 0, 0, 0,
                       (const float *)nullptr, c.ib.X0, skip);
    netForward123(n, c.ib.X0, nullptr, c.ib, nt, skip, s);
    hipLaunchKernelGGL(k_nn_out_dense, dim3(nt, 8), dim3(256), 0, s, c.ib.A3, n.w4f, q.sx, q.sy, q.sz, q.ntx, q.nty,
                       q.cout, skip);
  }
  const bool linear = n.prolong == MSBGX_NET_PROLONG_LINEAR;
  if (linear && (g.bsx0 == 16 || g.bsx0 == 32)) {   // every scale in one pass
    ProlongScales ps{};
    for (int k = 0; k < n.scales; k++) {
      const LearnedCtx::Scale &q = c.sc[k];
      ps.coarse[k] = q.cout;
      ps.wgt[k] = q.w;
      ps.z[k] = q.z;
      ps.f[k] = q.f;
      ps.sx[k] = q.sx;
      ps.sy[k] = q.sy;
      ps.sz[k] = q.sz;
    }
    auto launch = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(g.nActive0), dim3(256), 0, s, ps, g.nbx, g.nby, d.dActive0, m0.d_off, c.unk, st);
    };
    if (dir16) {   // the two-direction pass's BF16 directions (coarseDir16)
      for (int k = 0; k < 2; k++) ps.z16[k] = c.dir16 + (size_t)k * m0.ncells;
      launch(k_nn_prolong_lin_b<32, 2, uint16_t>);
      return;
    }
    const int sel = (g.bsx0 == 32 ? 3 : 0) + n.scales - 1;
    switch (sel) {
      case 0: launch(k_nn_prolong_lin_b<16, 1>); break;
      case 1: launch(k_nn_prolong_lin_b<16, 2>); break;
      case 2: launch(k_nn_prolong_lin_b<16, 3>); break;
      case 3: launch(k_nn_prolong_lin_b<32, 1>); break;
      case 4: launch(k_nn_prolong_lin_b<32, 2>); break;
      default: launch(k_nn_prolong_lin_b<32, 3>); break;
    }
    return;
  }
  for (int k = 0; k < n.scales; k++) {
    LearnedCtx::Scale &q = c.sc[k];
    if (linear)
      hipLaunchKernelGGL(k_nn_prolong_lin, dim3(g.nActive0), dim3(256), 0, s, q.cout, q.w, q.f, g.nbx, g.nby, q.sx,
                         q.sy, q.sz, d.dActive0, m0.d_off, m0.d_flags, m0.d_D, g.bsx0, q.z, st);
    else
      hipLaunchKernelGGL(k_nn_prolong_blocks, dim3(g.nActive0), dim3(256), 0, s, q.cout, q.f, g.nbx, g.nby, q.sx,
                         q.sy, d.dActive0, m0.d_off, m0.d_flags, m0.d_D, ncell, q.z, st);
  }
}

/* Galerkin combination of K > 1 directions: their residuals q_a = r - A z_a
 * and the Galerkin sums (for two directions on 32^3 level-0 blocks in the
 * same pass), then zc
 * and rs = r - A zc with the coefficients that minimise the A-norm error
 * over their span.  With BF16 directions (coarseDir16) the directions and
 * their residuals are stored as BF16 bits, and A zc is taken on the PCG
 * blocks alone: rs = r on the rest of the MG0 pool. */
template <int K> static void coarseCombine(Grid &g, LearnedCtx &c, hipStream_t s)
{
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  if constexpr (K == 2) {
    if (coarseDir16(g, c, K)) {
      const size_t nc = (size_t)m0.ncells;
      MulVecs<K, uint16_t> mv{};
      DirSetT<uint16_t> ds{};
      for (int a = 0; a < K; a++) {
        mv.x[a] = ds.z[a] = c.dir16 + (size_t)a * nc;
        mv.y[a] = c.dir16 + (size_t)(K + a) * nc;
        ds.q[a] = mv.y[a];
      }
      const DevList &l0 = m0.dAct[0];
      hipLaunchKernelGGL((k_multiply_z<32, K, uint16_t>), dim3(l0.n), dim3(32 * 32), 0, s, levelArgs(g, 0), l0.ptr,
                         mv, m0.d_ch[C_R], 1, c.gram, 0, &d.state->skipNet);
      hipLaunchKernelGGL(k_nn_gram_solve<K>, dim3(1), dim3(256), 0, s, c.gram, l0.n, d.state);
      hipLaunchKernelGGL((k_nn_combine<K, uint16_t>), dim3(kLVecBlocks), dim3(kVecThreads), 0, s, ds, c.rs,
                         m0.d_ch[C_R], (int64_t)m0.ncells, c.n0, d.state);
      return;
    }
  }
  DirSet ds{};
  for (int a = 0; a < K; a++) {
    ds.z[a] = c.sc[a].z;
    ds.q[a] = c.sc[a].q;
  }
  const int *skip = &d.state->skipNet;
  const DevList &l0 = m0.dAct[0];
  bool fused = false;
  if constexpr (K == 2) {   // three directions' state does not fit the registers of 1024 threads
    fused = g.bsx0 == 32 && l0.n == g.nActive0 && !(g.options & MSBGX_OPT_UNFUSED_COARSE);
    if (fused) {
      /* level-0 blocks: both directions' residuals and the Galerkin partials
       * in one pass; the other MG0 blocks one direction at a time */
      MulVecs<K> mv{};
      for (int a = 0; a < K; a++) {
        mv.x[a] = c.sc[a].z;
        mv.y[a] = c.sc[a].q;
      }
      if (l0.n)
        hipLaunchKernelGGL((k_multiply_z<32, K>), dim3(l0.n), dim3(32 * 32), 0, s, levelArgs(g, 0), l0.ptr, mv,
                           m0.d_ch[C_R], 1, c.gram, 0, skip);
      for (int a = 0; a < K; a++)
        for (int lev = 1; lev < g.nLevels; lev++) {
          const DevList &l = m0.dAct[lev];
          switch (g.bsx0 >> lev) {
            case 2: launchMultiplyB<2>(g, 0, l, c.sc[a].z, m0.d_ch[C_R], c.sc[a].q, 1, nullptr, 0, skip, s); break;
            case 4: launchMultiplyB<4>(g, 0, l, c.sc[a].z, m0.d_ch[C_R], c.sc[a].q, 1, nullptr, 0, skip, s); break;
            case 8: launchMultiplyB<8>(g, 0, l, c.sc[a].z, m0.d_ch[C_R], c.sc[a].q, 1, nullptr, 0, skip, s); break;
            case 16: launchMultiplyB<16>(g, 0, l, c.sc[a].z, m0.d_ch[C_R], c.sc[a].q, 1, nullptr, 0, skip, s); break;
          }
        }
      hipLaunchKernelGGL(k_nn_gram_solve<K>, dim3(1), dim3(256), 0, s, c.gram, std::max(l0.n, 1), d.state);
    }
  }
  if (!fused) {
    for (int a = 0; a < K; a++)
      multiplyLaunch(g, 0, true, c.sc[a].z, m0.d_ch[C_R], c.sc[a].q, nullptr, skip, s);
    hipLaunchKernelGGL(k_nn_gram<K>, dim3(kGramBlocks), dim3(256), 0, s, ds, m0.d_ch[C_R], c.n0, c.gram, skip);
    hipLaunchKernelGGL(k_nn_gram_solve<K>, dim3(1), dim3(256), 0, s, c.gram, kGramBlocks, d.state);
  }
  hipLaunchKernelGGL(k_nn_combine<K>, dim3(kVecBlocks), dim3(kVecThreads), 0, s, ds, c.rs, m0.d_ch[C_R],
                     (int64_t)m0.ncells, (int64_t)m0.ncells, d.state);
}

/* Coarse form, learned path (skipped while skipNet): directions
 * z_f = P_f N(R_f r) at the net's scales, zc their Galerkin combination
 * (one scale: the step gamma = <zc, r> / <zc, A zc>); then
 * z = zc + V(r - A zc), and the partials of <r, z> for the next finalize. */
static void coarseLearnedPath(Grid &g, Net &n, LearnedCtx &c, const msbgx_solve_params &p, hipStream_t s,
                              bool meanDone = false)
{
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  const int *skip = &d.state->skipNet;
  const bool dir16 = coarseDir16(g, c, n.scales);
  coarseNetwork(g, n, c, m0.d_ch[C_R], d.state, s, meanDone, dir16);
  if (n.scales == 1) {
    multiplyLaunch(g, 0, true, c.sc[0].z, m0.d_ch[C_R], c.sc[0].q, nullptr, skip, s);
    /* Galerkin step along zc: gamma = <zc, r> / <zc, A zc>, A zc = r - rs */
    hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, s, c.zc, m0.d_ch[C_R], c.n0, d.partials, skip);
    finalize(d, kVecBlocks, FIN_ZCR, nullptr);
    hipLaunchKernelGGL(k_nn_dot_diff, dim3(kVecBlocks), dim3(kVecThreads), 0, s, c.zc, m0.d_ch[C_R], c.rs, c.n0,
                       d.partials, skip);
    finalize(d, kVecBlocks, FIN_GAMMA, nullptr);
    hipLaunchKernelGGL(k_nn_gamma, dim3(kVecBlocks), dim3(kVecThreads), 0, s, c.zc, c.rs, m0.d_ch[C_R],
                       (int64_t)g.mg[0].ncells, d.state);
  } else if (n.scales == 2) {
    coarseCombine<2>(g, c, s);
  } else {
    coarseCombine<3>(g, c, s);
  }
  std::swap(m0.d_ch[C_R], c.rs);   // the V-cycle below reads its right-hand side from rs
  zeroActive(g, 0, m0.d_ch[C_Z], &d.state->done, s);
  vcycleLaunch(g, 0, p.n_pre, p.n_post, p.n_coarse, &d.state->done, s);
  std::swap(m0.d_ch[C_R], c.rs);
  if (dir16) {
    DirSetT<uint16_t> ds16{};
    for (int a = 0; a < 2; a++) ds16.z[a] = c.dir16 + (size_t)a * m0.ncells;
    hipLaunchKernelGGL((k_nn_add_dot<2, uint16_t>), dim3(kLVecBlocks), dim3(kVecThreads), 0, s, m0.d_ch[C_Z], ds16,
                       m0.d_ch[C_R], c.n0, d.state, d.partials);
    return;
  }
  DirSet ds{};
  for (int a = 0; a < n.scales; a++) ds.z[a] = c.sc[a].z;
  switch (n.scales) {
    case 1:
      hipLaunchKernelGGL(k_nn_add_dot<1>, dim3(kLVecBlocks), dim3(kVecThreads), 0, s, m0.d_ch[C_Z], ds, m0.d_ch[C_R],
                         c.n0, d.state, d.partials);
      break;
    case 2:
      hipLaunchKernelGGL(k_nn_add_dot<2>, dim3(kLVecBlocks), dim3(kVecThreads), 0, s, m0.d_ch[C_Z], ds, m0.d_ch[C_R],
                         c.n0, d.state, d.partials);
      break;
    default:
      hipLaunchKernelGGL(k_nn_add_dot<3>, dim3(kLVecBlocks), dim3(kVecThreads), 0, s, m0.d_ch[C_Z], ds, m0.d_ch[C_R],
                         c.n0, d.state, d.partials);
      break;
  }
}

static void pcgIterationLearnedCoarse(Grid &g, Net &n, LearnedCtx &c, const msbgx_solve_params &p)
{
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  const int *done = &d.state->done;
  const int64_t n0 = c.n0;
  const int nPart = multiplyLaunch(g, 0, false, m0.d_ch[C_D], nullptr, m0.d_ch[C_Q], d.partials, done, d.stream);
  finalize(d, nPart, FIN_ALPHA, nullptr);
  /* <r_k+1, z_k> of the flexible beta comes with the update, before z_k is
   * replaced, and so does the network's block-mean input */
  const int nb = g.nActive0;
  if (nb > 0)
    hipLaunchKernelGGL(k_update_xr_rz_mean, dim3(nb), dim3(256), 0, d.stream, m0.d_ch[C_X], m0.d_ch[C_R],
                       m0.d_ch[C_D], m0.d_ch[C_Q], m0.d_ch[C_Z], d.dActive0, m0.d_off, g.bsx0 * g.bsx0 * g.bsx0,
                       d.state, d.partials, c.sc[0].cin);
  finalize(d, nb, FIN_CONV, d.hist);
  finalizeAt(d, nb, nb, FIN_RZOLD, nullptr);
  coarseLearnedPath(g, n, c, p, d.stream, true);   // its V-cycle is the plain one while the network is off
  finalize(d, kLVecBlocks, FIN_BETA, nullptr);
  /* a rejected corrected z restarts along r */
  hipLaunchKernelGGL(k_nn_axpy, dim3(kLVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_D], m0.d_ch[C_Z],
                     m0.d_ch[C_R], n0, d.state);
}

static void pcgIterationLearned(Grid &g, Net &n, LearnedCtx &c, const msbgx_solve_params &p)
{
  if (n.mode == MSBGX_NET_COARSE) {
    pcgIterationLearnedCoarse(g, n, c, p);
    return;
  }
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  const int *done = &d.state->done;
  const int *skip = &d.state->skipNet;
  const int64_t n0 = c.n0;
  const int nPart = multiplyLaunch(g, 0, false, m0.d_ch[C_D], nullptr, m0.d_ch[C_Q], d.partials, done, d.stream);
  finalize(d, nPart, FIN_ALPHA, nullptr);
  hipLaunchKernelGGL(k_update_xr, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_X], m0.d_ch[C_R],
                     m0.d_ch[C_D], m0.d_ch[C_Q], n0, d.state, d.partials);
  finalize(d, kVecBlocks, FIN_CONV, d.hist);
  hipLaunchKernelGGL(k_nn_copy, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_Z], c.zprev, n0, skip, 0);
  zeroActive(g, 0, m0.d_ch[C_Z], done, d.stream);
  vcycleLaunch(g, 0, p.n_pre, p.n_post, p.n_coarse, done, d.stream);
  hipLaunchKernelGGL(k_nn_copy, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_Z], c.zplain, n0, skip, 0);
  netCorrect(g, n, c, d.stream);
  hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], c.zprev, n0, d.partials,
                     skip);
  finalize(d, kVecBlocks, FIN_RZOLD, nullptr);
  hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], c.zplain, n0, d.partials,
                     skip);
  finalize(d, kVecBlocks, FIN_RZPLAIN, nullptr);
  hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], m0.d_ch[C_Z], n0, d.partials,
                     done);
  finalize(d, kVecBlocks, FIN_BETA, nullptr);
  hipLaunchKernelGGL(k_nn_axpy, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_D], m0.d_ch[C_Z], c.zplain,
                     n0, d.state);
}

int deviceSolveLearned(Grid &g, Net &n, const msbgx_solve_params &p, int *iters, float *rel, int *fellBack)
{
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  if (!d.learned) d.learned = new LearnedCtx();
  LearnedCtx &c = *d.learned;
  c.n0 = (int64_t)g.nActive0 * g.bsx0 * g.bsx0 * g.bsx0;
  if (int rc = allocBuffers(c.ib, std::max(1, g.nActive0), false, d.stream)) return rc;
  if (!c.zprev)
    if (int rc = dalloc(&c.zprev, std::max<int64_t>(1, c.n0))) return rc;
  if (!c.zplain)
    if (int rc = dalloc(&c.zplain, std::max<int64_t>(1, c.n0))) return rc;
  if (int rc = ensureHist(g, p.max_iter + 1)) return rc;
  const bool coarse = n.mode == MSBGX_NET_COARSE;
  if (coarse)
    if (int rc = coarseSetup(g, c, n.scales, d.stream)) return rc;
  GraphKey key = graphKey(g, p);
  key.net = n.serial;
  key.netMode = n.mode + 16 * n.prolong + 64 * n.scales;
  const int chunk = key.chunk;
  if (c.graph && c.graphKey != key) {
    (void)hipGraphExecDestroy(c.graph);
    c.graph = nullptr;
  }
  PcgState init{};
  init.tolSq = (double)p.tol * (double)p.tol;
  init.maxIter = p.max_iter;
  init.learned = 1;
  init.coarseMode = coarse ? 1 : 0;
  init.stallWin = n.stallWin;
  init.stallFactor = (1.0 - (double)n.stallDrop) * (1.0 - (double)n.stallDrop);
  HIPCHK(hipMemcpyAsync(d.state, &init, sizeof(init), hipMemcpyHostToDevice, d.stream));
  HIPCHK(hipStreamSynchronize(n.stream));   // weights packed
  HIPCHK(hipEventRecord(d.ev0, d.stream));
  hipLaunchKernelGGL(k_copy_norm, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_B], m0.d_ch[C_R], c.n0,
                     d.partials);
  finalize(d, kVecBlocks, FIN_BNORM, nullptr);
  const int *done = &d.state->done;
  const int *skip = &d.state->skipNet;
  if (coarse) {
    coarseLearnedPath(g, n, c, p, d.stream);   // its V-cycle is the plain one while the network is off
  } else {
    zeroActive(g, 0, m0.d_ch[C_Z], done, d.stream);
    vcycleLaunch(g, 0, p.n_pre, p.n_post, p.n_coarse, done, d.stream);
    hipLaunchKernelGGL(k_nn_copy, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_Z], c.zplain, c.n0, skip,
                       0);
    netCorrect(g, n, c, d.stream);
    hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], c.zplain, c.n0,
                       d.partials, skip);
    finalize(d, kVecBlocks, FIN_RZPLAIN, nullptr);
    hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], m0.d_ch[C_Z], c.n0,
                       d.partials, done);
  }
  finalize(d, coarse ? kLVecBlocks : kVecBlocks, FIN_RHO0, nullptr);   // the coarse path left the partials of <r, z>
  /* d = z on every MG0 block; the PCG range takes the plain z (fine form) or
   * r (coarse form) if the corrected one was rejected */
  for (int lev = 0; lev < g.nLevels; lev++)
    if (m0.dAct[lev].n)
      hipLaunchKernelGGL(k_copy_blocks, dim3(m0.dAct[lev].n), dim3(256), 0, d.stream, m0.dAct[lev].ptr, m0.d_level,
                         m0.d_off, g.bsx0, m0.d_ch[C_Z], m0.d_ch[C_D], done);
  hipLaunchKernelGGL(k_nn_copy, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, coarse ? m0.d_ch[C_R] : c.zplain,
                     m0.d_ch[C_D], c.n0, &d.state->usePlain, 1);
  HIPCHK(hipGetLastError());
  const bool useGraph = !(g.options & MSBGX_OPT_NO_GRAPH);
  PcgState hs{};
  for (int it = 0; it < p.max_iter;) {
    const int k = std::min(chunk, p.max_iter - it);
    if (useGraph && k == chunk) {
      if (!c.graph) {
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(d.stream, hipStreamCaptureModeRelaxed));
        for (int i = 0; i < chunk; i++) pcgIterationLearned(g, n, c, p);
        HIPCHK(hipStreamEndCapture(d.stream, &graph));
        HIPCHK(hipGraphInstantiate(&c.graph, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        c.graphKey = key;
      }
      HIPCHK(hipGraphLaunch(c.graph, d.stream));
    } else {
      for (int i = 0; i < k; i++) pcgIterationLearned(g, n, c, p);
    }
    it += k;
    HIPCHK(hipMemcpyAsync(&hs, d.state, sizeof(hs), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    if (hs.done) break;
  }
  hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], m0.d_ch[C_R], c.n0,
                     d.partials, d.zeroFlag);
  finalize(d, kVecBlocks, FIN_FINAL, d.hist + p.max_iter);
  HIPCHK(hipEventRecord(d.ev1, d.stream));
  HIPCHK(hipMemcpyAsync(&hs, d.state, sizeof(hs), hipMemcpyDeviceToHost, d.stream));
  double fin = 0.0;
  HIPCHK(hipMemcpyAsync(&fin, d.hist + p.max_iter, sizeof(double), hipMemcpyDeviceToHost, d.stream));
  HIPCHK(hipStreamSynchronize(d.stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, d.ev0, d.ev1));
  g.lastSolveMs = ms;
  g.history.assign(hs.iter, 0.0);
  if (hs.iter > 0) {
    std::vector<double> rr(hs.iter);
    HIPCHK(hipMemcpy(rr.data(), d.hist, hs.iter * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < hs.iter; i++) g.history[i] = std::sqrt(rr[i] / hs.bnorm2);
  }
  *fellBack = hs.learned ? -1 : hs.fellBack;
  if (hs.brk == 3) { setError("non-finite right-hand side"); return MSBGX_EINVAL; }
  if (hs.brk == 2) { *iters = 0; *rel = 0.0f; return MSBGX_OK; }
  *rel = (float)std::sqrt(fin / hs.bnorm2);
  *iters = hs.conv > 0 ? hs.conv : p.max_iter;
  return MSBGX_OK;
}

/* Training pairs from the plain MG-PCG on a prepared grid: a solve to p.tol
 * gives x*; a second solve from 0 stops after every `every`-th iteration,
 * and `perPair` random level-0 PCG blocks of that iterate become samples. */
/* Device scratch freed on every exit path. */
struct DevScratch {
  std::vector<void *> ptrs;
  ~DevScratch()
  {
    for (void *q : ptrs) (void)hipFree(q);
  }
  template <typename T> int alloc(T **p, size_t n)
  {
    const int rc = dalloc(p, n);
    if (!rc) ptrs.push_back((void *)*p);
    return rc;
  }
};

int netCollect(Net &n, Grid &g, const msbgx_solve_params &p, int every, int maxPairs, int perPair, uint64_t seed,
               int *pairsOut)
{
  DevScratch scratch;
  Dev &d = dev(g);
  MgLevel &m0 = g.mg[0];
  const int64_t n0 = (int64_t)g.nActive0 * g.bsx0 * g.bsx0 * g.bsx0;
  *pairsOut = 0;
  if (g.nActive0 == 0 || every < 1 || maxPairs < 1 || perPair < 1) return MSBGX_OK;
  int it = 0;
  float rr = 0.f;
  if (int rc = deviceZero(g, C_X, 0)) return rc;
  if (int rc = deviceSolve(g, p, &it, &rr)) return rc;
  float *xs = nullptr, *err = nullptr;
  int *blk = nullptr;
  if (int rc = scratch.alloc(&xs, n0)) return rc;
  if (int rc = scratch.alloc(&err, n0)) return rc;
  if (int rc = scratch.alloc(&blk, perPair)) return rc;
  HIPCHK(hipMemcpyAsync(xs, m0.d_ch[C_X], n0 * sizeof(float), hipMemcpyDeviceToDevice, d.stream));
  if (int rc = deviceZero(g, C_X, 0)) return rc;
  /* second solve, eagerly, sampling after chosen iterations */
  PcgState init{};
  init.tolSq = (double)p.tol * (double)p.tol;
  init.maxIter = p.max_iter;
  HIPCHK(hipMemcpyAsync(d.state, &init, sizeof(init), hipMemcpyHostToDevice, d.stream));
  hipLaunchKernelGGL(k_copy_norm, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_B], m0.d_ch[C_R], n0,
                     d.partials);
  finalize(d, kVecBlocks, FIN_BNORM, nullptr);
  const int *done = &d.state->done;
  zeroActive(g, 0, m0.d_ch[C_Z], done, d.stream);
  vcycleLaunch(g, 0, p.n_pre, p.n_post, p.n_coarse, done, d.stream);
  for (int lev = 0; lev < g.nLevels; lev++)
    if (m0.dAct[lev].n)
      hipLaunchKernelGGL(k_copy_blocks, dim3(m0.dAct[lev].n), dim3(256), 0, d.stream, m0.dAct[lev].ptr, m0.d_level,
                         m0.d_off, g.bsx0, m0.d_ch[C_Z], m0.d_ch[C_D], done);
  hipLaunchKernelGGL(k_dot, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, m0.d_ch[C_R], m0.d_ch[C_Z], n0, d.partials,
                     done);
  finalize(d, kVecBlocks, FIN_RHO0, nullptr);
  uint64_t st = seed ? seed : 1;
  auto next = [&]() {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return st;
  };
  std::vector<int> pick(perPair);
  PcgState hs{};
  int pairs = 0;
  /* coarse form: dense grids of the residual and error means, scale by scale */
  const bool coarse = n.mode == MSBGX_NET_COARSE;
  LearnedCtx *lc = nullptr;
  float *cerr[kMaxScales] = {nullptr, nullptr, nullptr}, *tsc = nullptr;
  if (coarse) {
    if (!d.learned) d.learned = new LearnedCtx();
    lc = d.learned;
    lc->n0 = n0;
    if (int rc = coarseSetup(g, *lc, n.scales, d.stream)) return rc;
    /* tiles per iterate with any origin offset: at most one more per axis */
    const LearnedCtx::Scale &s0 = lc->sc[0];
    const int maxTiles = (s0.ntx + 1) * (s0.nty + 1) * (s0.ntz + 1);
    if (int rc = scratch.alloc(&tsc, maxTiles)) return rc;
    for (int k = 0; k < n.scales; k++)
      if (int rc = scratch.alloc(&cerr[k], lc->sc[k].ncoarse)) return rc;
  }
  for (int k = 1; k <= p.max_iter && pairs < maxPairs; k++) {
    pcgIteration(g, p);
    HIPCHK(hipMemcpyAsync(&hs, d.state, sizeof(hs), hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    if (hs.brk || hs.done) break;   // a converged iteration leaves Z unset
    if (k % every == 0 && coarse) {
      /* input: means of r_k; target: means of e_k = x* - x_k over the unknowns */
      hipLaunchKernelGGL(k_nn_sub, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, xs, m0.d_ch[C_X], err, n0);
      LearnedCtx::Scale &s0 = lc->sc[0];
      hipLaunchKernelGGL(k_nn_fill, dim3((s0.ncoarse + 255) / 256), dim3(256), 0, d.stream, s0.cin, s0.ncoarse, 0.0f,
                         (const int *)nullptr);
      hipLaunchKernelGGL(k_nn_fill, dim3((s0.ncoarse + 255) / 256), dim3(256), 0, d.stream, cerr[0], s0.ncoarse, 0.0f,
                         (const int *)nullptr);
      const int ncell = g.bsx0 * g.bsx0 * g.bsx0;
      hipLaunchKernelGGL(k_nn_block_mean, dim3(g.nActive0), dim3(256), 0, d.stream, m0.d_ch[C_R], d.dActive0,
                         m0.d_off, m0.d_flags, m0.d_D, ncell, 0, s0.cin, (const int *)nullptr);
      hipLaunchKernelGGL(k_nn_block_mean, dim3(g.nActive0), dim3(256), 0, d.stream, err, d.dActive0, m0.d_off,
                         m0.d_flags, m0.d_D, ncell, 1, cerr[0], (const int *)nullptr);
      for (int sk = 0; sk < n.scales; sk++) {
        LearnedCtx::Scale &q = lc->sc[sk];
        if (sk > 0) {
          const LearnedCtx::Scale &pq = lc->sc[sk - 1];
          hipLaunchKernelGGL(k_nn_coarsen, dim3((q.ncoarse + 255) / 256), dim3(256), 0, d.stream, pq.cin, pq.w, pq.sx,
                             pq.sy, pq.sz, q.cin, q.sx, q.sy, q.sz, (const int *)nullptr);
          hipLaunchKernelGGL(k_nn_coarsen, dim3((q.ncoarse + 255) / 256), dim3(256), 0, d.stream, cerr[sk - 1], pq.w,
                             pq.sx, pq.sy, pq.sz, cerr[sk], q.sx, q.sy, q.sz, (const int *)nullptr);
        }
        /* random tile origin: the domain boundary falls anywhere inside a tile */
        const int ox = (int)(next() % 16u), oy = (int)(next() % 16u), oz = (int)(next() % 16u);
        const int tx = (q.sx + ox + 15) / 16, ty = (q.sy + oy + 15) / 16, tz = (q.sz + oz + 15) / 16;
        const int ntp = tx * ty * tz;
        if (int rc = growSet(n, n.nS + ntp)) return rc;
        hipLaunchKernelGGL(k_nn_tile_target, dim3(ntp), dim3(256), 0, d.stream, cerr[sk], q.w, q.sx, q.sy, q.sz, tx,
                           ty, ox, oy, oz, tsc, n.setT + (size_t)n.nS * nn::NCELL,
                           n.setM + (size_t)n.nS * nn::NCELL);
        hipLaunchKernelGGL(k_nn_gather_dense, dim3(ntp), dim3(256), 0, d.stream, q.cin, q.sx, q.sy, q.sz, tx, ty, ox,
                           oy, oz, tsc, n.setX + (size_t)n.nS * nn::PS, (const int *)nullptr);
        HIPCHK(hipGetLastError());
        std::vector<uint8_t> mk((size_t)ntp * nn::NCELL);
        HIPCHK(hipMemcpyAsync(mk.data(), n.setM + (size_t)n.nS * nn::NCELL, mk.size(), hipMemcpyDeviceToHost,
                              d.stream));
        HIPCHK(hipStreamSynchronize(d.stream));
        for (int j = 0; j < ntp; j++) {
          int cnt = 0;
          for (int cc = 0; cc < nn::NCELL; cc++) cnt += mk[(size_t)j * nn::NCELL + cc];
          n.maskCount.push_back(cnt);
        }
        n.nS += ntp;
      }
      pairs++;
    } else if (k % every == 0) {
      /* z = V(r_k) is in Z */
      multiplyLaunch(g, 0, true, m0.d_ch[C_Z], m0.d_ch[C_R], m0.d_ch[C_Q], nullptr, d.zeroFlag, d.stream);
      hipLaunchKernelGGL(k_nn_sub, dim3(kVecBlocks), dim3(kVecThreads), 0, d.stream, xs, m0.d_ch[C_X], err, n0);
      for (int j = 0; j < perPair; j++) pick[j] = g.active0[next() % (uint64_t)g.nActive0];
      HIPCHK(hipMemcpyAsync(blk, pick.data(), perPair * sizeof(int), hipMemcpyHostToDevice, d.stream));
      if (int rc = growSet(n, n.nS + perPair)) return rc;
      /* err and Z are indexed by the MG0 pool; the PCG range [0, n0) is its head */
      hipLaunchKernelGGL(k_nn_sample, dim3(perPair), dim3(256), 0, d.stream, m0.d_ch[C_Q], err, m0.d_ch[C_Z], blk,
                         m0.d_level, m0.d_off, m0.d_flags, m0.d_D, g.nbx, g.nby, g.nbz, n.setX + (size_t)n.nS * nn::PS,
                         n.setT + (size_t)n.nS * nn::NCELL, n.setM + (size_t)n.nS * nn::NCELL);
      HIPCHK(hipGetLastError());
      std::vector<uint8_t> mk((size_t)perPair * nn::NCELL);
      HIPCHK(hipMemcpyAsync(mk.data(), n.setM + (size_t)n.nS * nn::NCELL, mk.size(), hipMemcpyDeviceToHost, d.stream));
      HIPCHK(hipStreamSynchronize(d.stream));
      for (int j = 0; j < perPair; j++) {
        int cnt = 0;
        for (int c = 0; c < nn::NCELL; c++) cnt += mk[(size_t)j * nn::NCELL + c];
        n.maskCount.push_back(cnt);
      }
      n.nS += perPair;
      pairs++;
    }
  }
  HIPCHK(hipStreamSynchronize(d.stream));   // scratch is freed on return
  if (int rc = deviceZero(g, C_X, 0)) return rc;
  *pairsOut = pairs;
  return MSBGX_OK;
}

}  // namespace msbgx
