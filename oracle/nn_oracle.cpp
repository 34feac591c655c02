/*
 * nn_oracle.cpp -- CPU FP64 restatement of the learned preconditioner
 * correction's network (forward, loss, backward).
 *
 * TEST INFRASTRUCTURE.  Only tests/ load it, as the checker of the GPU
 * network kernels; the product library never links or calls it.
 *
 * The reference has no learned component; this file restates the network
 * this build defines: four 3x3x3 "same" convolutions without biases on an
 * 18^3 patch (zero outside it), channels 1 -> 16 -> 16 -> 16 -> 1, ReLU after
 * the first three, output on the 16^3 interior cells (patch coordinates
 * 1..16), masked; loss = sum mask (out - target)^2 / sum mask.
 * Parameter layout: W1[27][16] (tap, out), W2, W3 [27][16][16] (tap, in,
 * out), W4[27][16] (tap, in); tap = 9 (dz+1) + 3 (dy+1) + (dx+1), a tap reads
 * the input at p + (dx, dy, dz).
 *
 * All arithmetic is FP64.  With round_bf16 != 0 the weights, the patch and
 * every stored activation and activation gradient are rounded to BF16
 * (round-to-nearest-even) at the points where the GPU kernels store BF16,
 * so the GPU result must agree to FP32-vs-FP64 accumulation error; with
 * round_bf16 == 0 it is the exact FP64 function, whose gradient the tests
 * check against finite differences.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int P = 18, P3 = P * P * P, B = 16, NC = B * B * B, C = 16, T = 27;
constexpr int NW1 = T * C, NW2 = T * C * C;
constexpr int OFF_W1 = 0, OFF_W2 = NW1, OFF_W3 = NW1 + NW2, OFF_W4 = NW1 + 2 * NW2, NPARAM = 2 * NW1 + 2 * NW2;

double bf16(double v)
{
  float f = (float)v;
  uint32_t b;
  memcpy(&b, &f, 4);
  if ((b & 0x7fffffffu) > 0x7f800000u) return v;
  b = (b + 0x7fffu + ((b >> 16) & 1u)) & 0xffff0000u;
  memcpy(&f, &b, 4);
  return f;
}

inline int idx(int x, int y, int z) { return x + P * (y + P * z); }
inline bool inside(int x, int y, int z) { return x >= 0 && y >= 0 && z >= 0 && x < P && y < P && z < P; }
inline void tapD(int t, int &dx, int &dy, int &dz) { dx = t % 3 - 1; dy = (t / 3) % 3 - 1; dz = t / 9 - 1; }

struct Net {
  std::vector<double> w;
  bool round;
  double r(double v) const { return round ? bf16(v) : v; }
};

/* Z[p][co] = sum_t sum_ci A[p + t][ci] W[t][ci][co] on the patch (A has cin channels). */
void conv(const std::vector<double> &A, int cin, const double *W, int cout, std::vector<double> &Z)
{
  Z.assign((size_t)P3 * cout, 0.0);
  for (int z = 0; z < P; z++)
    for (int y = 0; y < P; y++)
      for (int x = 0; x < P; x++) {
        double *zo = &Z[(size_t)idx(x, y, z) * cout];
        for (int t = 0; t < T; t++) {
          int dx, dy, dz;
          tapD(t, dx, dy, dz);
          if (!inside(x + dx, y + dy, z + dz)) continue;
          const double *a = &A[(size_t)idx(x + dx, y + dy, z + dz) * cin];
          for (int ci = 0; ci < cin; ci++)
            for (int co = 0; co < cout; co++) zo[co] += a[ci] * W[(t * cin + ci) * cout + co];
        }
      }
}

/* dA[q][ci] = sum_t sum_co dZ[q - t][co] W[t][ci][co] */
void convBack(const std::vector<double> &dZ, int cout, const double *W, int cin, std::vector<double> &dA)
{
  dA.assign((size_t)P3 * cin, 0.0);
  for (int z = 0; z < P; z++)
    for (int y = 0; y < P; y++)
      for (int x = 0; x < P; x++) {
        const double *g = &dZ[(size_t)idx(x, y, z) * cout];
        for (int t = 0; t < T; t++) {
          int dx, dy, dz;
          tapD(t, dx, dy, dz);
          if (!inside(x + dx, y + dy, z + dz)) continue;
          double *a = &dA[(size_t)idx(x + dx, y + dy, z + dz) * cin];
          for (int ci = 0; ci < cin; ci++)
            for (int co = 0; co < cout; co++) a[ci] += g[co] * W[(t * cin + ci) * cout + co];
        }
      }
}

/* dW[t][ci][co] += sum_p A[p + t][ci] dZ[p][co] */
void convWeightGrad(const std::vector<double> &A, int cin, const std::vector<double> &dZ, int cout, double *dW)
{
  for (int z = 0; z < P; z++)
    for (int y = 0; y < P; y++)
      for (int x = 0; x < P; x++) {
        const double *g = &dZ[(size_t)idx(x, y, z) * cout];
        for (int t = 0; t < T; t++) {
          int dx, dy, dz;
          tapD(t, dx, dy, dz);
          if (!inside(x + dx, y + dy, z + dz)) continue;
          const double *a = &A[(size_t)idx(x + dx, y + dy, z + dz) * cin];
          for (int ci = 0; ci < cin; ci++)
            for (int co = 0; co < cout; co++) dW[(t * cin + ci) * cout + co] += a[ci] * g[co];
        }
      }
}

inline int cellPos(int c) { return idx(c % B + 1, (c / B) % B + 1, c / (B * B) + 1); }

struct Acts {
  std::vector<double> a0, a1, a2, a3, out;
};

void forward(const Net &n, const float *x, const uint8_t *mask, Acts &s)
{
  s.a0.resize(P3);
  for (int i = 0; i < P3; i++) s.a0[i] = n.r(x[i]);
  std::vector<double> z;
  conv(s.a0, 1, &n.w[OFF_W1], C, z);
  for (double &v : z) v = n.r(v > 0.0 ? v : 0.0);
  s.a1 = z;
  conv(s.a1, C, &n.w[OFF_W2], C, z);
  for (double &v : z) v = n.r(v > 0.0 ? v : 0.0);
  s.a2 = z;
  conv(s.a2, C, &n.w[OFF_W3], C, z);
  for (double &v : z) v = n.r(v > 0.0 ? v : 0.0);
  s.a3 = z;
  conv(s.a3, C, &n.w[OFF_W4], 1, z);
  s.out.assign(NC, 0.0);
  for (int c = 0; c < NC; c++) s.out[c] = mask[c] ? z[cellPos(c)] : 0.0;
}

Net makeNet(const double *w, int round)
{
  Net n;
  n.round = round != 0;
  n.w.assign(w, w + NPARAM);
  if (n.round)
    for (double &v : n.w) v = bf16(v);
  return n;
}

}  // namespace

extern "C" {

int nno_num_params(void) { return NPARAM; }

/* out[count][16^3] */
void nno_forward(const double *w, int count, const float *x, const uint8_t *mask, int round_bf16, double *out)
{
  const Net n = makeNet(w, round_bf16);
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < count; i++) {
    Acts s;
    forward(n, x + (size_t)i * P3, mask + (size_t)i * NC, s);
    memcpy(out + (size_t)i * NC, s.out.data(), NC * sizeof(double));
  }
}

/* Loss over all samples and its gradient (grad may be NULL). */
void nno_loss_grad(const double *w, int count, const float *x, const float *target, const uint8_t *mask,
                   int round_bf16, double *loss, double *grad)
{
  const Net n = makeNet(w, round_bf16);
  long cnt = 0;
  for (long i = 0; i < (long)count * NC; i++) cnt += mask[i] ? 1 : 0;
  const double invN = cnt ? 1.0 / (double)cnt : 0.0;
  double L = 0.0;
  std::vector<double> G(NPARAM, 0.0);
#pragma omp parallel
  {
    std::vector<double> g(NPARAM, 0.0);
    double l = 0.0;
#pragma omp for schedule(dynamic)
    for (int i = 0; i < count; i++) {
      Acts s;
      const uint8_t *mk = mask + (size_t)i * NC;
      forward(n, x + (size_t)i * P3, mk, s);
      std::vector<double> dOut((size_t)P3, 0.0);   // on the patch, non-zero on masked cells
      for (int c = 0; c < NC; c++) {
        if (!mk[c]) continue;
        const double d = s.out[c] - (double)target[(size_t)i * NC + c];
        l += d * d * invN;
        dOut[cellPos(c)] = 2.0 * d * invN;
      }
      if (!grad) continue;
      convWeightGrad(s.a3, C, dOut, 1, &g[OFF_W4]);
      std::vector<double> dA, dZ;
      convBack(dOut, 1, &n.w[OFF_W4], C, dA);
      dZ.resize(dA.size());
      for (size_t k = 0; k < dA.size(); k++) dZ[k] = n.r(s.a3[k] > 0.0 ? dA[k] : 0.0);
      convWeightGrad(s.a2, C, dZ, C, &g[OFF_W3]);
      convBack(dZ, C, &n.w[OFF_W3], C, dA);
      for (size_t k = 0; k < dA.size(); k++) dZ[k] = n.r(s.a2[k] > 0.0 ? dA[k] : 0.0);
      convWeightGrad(s.a1, C, dZ, C, &g[OFF_W2]);
      convBack(dZ, C, &n.w[OFF_W2], C, dA);
      for (size_t k = 0; k < dA.size(); k++) dZ[k] = n.r(s.a1[k] > 0.0 ? dA[k] : 0.0);
      convWeightGrad(s.a0, 1, dZ, C, &g[OFF_W1]);
    }
#pragma omp critical
    {
      L += l;
      for (int k = 0; k < NPARAM; k++) G[k] += g[k];
    }
  }
  *loss = L;
  if (grad) memcpy(grad, G.data(), NPARAM * sizeof(double));
}

}  // extern "C"
