/* Learned c2p initial guess: the network object, its training data and its training on
 * the GPU with hand-written BF16 matrix-core kernels (FP32 accumulation, FP32 master
 * weights, Adam). The inference used inside the c2p kernel is mlp_forward_wave
 * (mlp_device.hpp); this file builds, trains, saves and loads the parameters.
 *
 * Parameter vector (FP32, this order): W1[64][6], b1[64], W2[64][64], b2[64],
 * W3[64][64], b3[64], W4[2][64], b4[2]; weights stored [out][in].
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "rmhd_device.hpp"
#include "mlp_device.hpp"
#include "../../include/maragpu.h"

using namespace maragpu;

extern "C" const double *maragpu_committed_prim(maragpu_ctx *c, int *N, int *Ng, long *T, int *n1, int *n2);

namespace {

constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + 64 * 6, OFF_W2 = OFF_B1 + 64, OFF_B2 = OFF_W2 + 64 * 64;
constexpr int OFF_W3 = OFF_B2 + 64, OFF_B3 = OFF_W3 + 64 * 64, OFF_W4 = OFF_B3 + 64;
constexpr int OFF_B4 = OFF_W4 + 2 * 64, NPARAM = OFF_B4 + 2;
constexpr uint32_t kSerialVersion = 2;   // the activation is part of the format

#define HIPOK(expr)                                                                  \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      std::fprintf(stderr, "[maragpu mlp] %s failed: %s\n", #expr, hipGetErrorString(e_)); \
      return MARAGPU_E_HIP;                                                          \
    }                                                                                \
  } while (0)

/* weight W[out][in] of a layer from the parameter vector (0 outside the layer's shape) */
__device__ __forceinline__ float weight(const float *th, int layer, int o, int i)
{
  switch (layer) {
  case 1:   // input columns i and 6 + i (hi and lo parts of feature i) share W1[o][i]
    if (i >= 2 * MLP_IN) return 0.f;
    return th[OFF_W1 + o * 6 + (i < MLP_IN ? i : i - MLP_IN)];
  case 2: return th[OFF_W2 + o * 64 + i];
  case 3: return th[OFF_W3 + o * 64 + i];
  default: return (o < 2 && i < 64) ? th[OFF_W4 + o * 64 + i] : 0.f;
  }
}

/* Re-swizzles the master weights into bf16 fragments. Forward fragments hold A = W^T
 * (rows = out, k = in); backward fragments hold A = W (rows = in, k = out). */
__global__ void k_swizzle(const float *__restrict__ th, uint16_t *__restrict__ fwd, uint16_t *__restrict__ bwd)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = (MLP_NFWD + MLP_NBWD) * MLP_FRAG;
  if (t >= total) return;
  const bool isf = t < MLP_NFWD * MLP_FRAG;
  const int u = isf ? t : t - MLP_NFWD * MLP_FRAG;
  const int f = u / MLP_FRAG, e = u % MLP_FRAG;
  const int lane = e / 4, j = e % 4;
  const int row = lane % 16, k = 4 * (lane / 16) + j;   // A[row][k] of the fragment
  float v;
  if (isf && f < MLP_F2) {
    v = weight(th, 1, 16 * f + row, k);
    fwd[u] = f32_to_bf16(v);
  }
  else if (isf) {
    // 16x32 tiles of the double-k product: lane l2, slot s holds A[l2 % 16][k2] with
    // k2 = 4 (l2 / 16) + s for s < 4, 16 + 4 (l2 / 16) + (s - 4) for s >= 4
    const int layer = f < MLP_F3 ? 2 : f < MLP_F4 ? 3 : 4;
    const int base = layer == 2 ? MLP_F2 : layer == 3 ? MLP_F3 : MLP_F4;
    const int uu = u - base * MLP_FRAG;
    const int q = uu / (2 * MLP_FRAG), e2 = uu % (2 * MLP_FRAG);
    const int l2 = e2 / 8, s = e2 % 8;
    const int mt = layer == 4 ? 0 : q / 2, kt = layer == 4 ? q : q % 2;
    const int k2 = 32 * kt + (s < 4 ? 4 * (l2 / 16) + s : 16 + 4 * (l2 / 16) + (s - 4));
    v = weight(th, layer, 16 * mt + l2 % 16, k2);
    fwd[u] = f32_to_bf16(v);
  }
  else {
    int layer, mt, kt;
    if (f < MLP_B3) { layer = 4; mt = f; kt = 0; }
    else if (f < MLP_B2) { layer = 3; mt = (f - MLP_B3) / 4; kt = (f - MLP_B3) % 4; }
    else { layer = 2; mt = (f - MLP_B2) / 4; kt = (f - MLP_B2) % 4; }
    v = weight(th, layer, 16 * kt + k, 16 * mt + row);
    bwd[u] = f32_to_bf16(v);
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t i, int k)
{
  const uint64_t z = mix64(seed * 0x9e3779b97f4a7c15ULL + i * 0xd1b54a32d192ed03ULL + (uint64_t)k * 0x632be59bd9b4e019ULL);
  return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* Synthetic training pairs: W log-uniform on [1, Wmax], magnetisation sigma = b^2/(rho h)
 * log-uniform on [smin, smax], p/rho log-uniform on [pmin, pmax], rho log-uniform on
 * [0.1, 10], isotropic velocity and field directions (random field-velocity angle). */
__global__ void k_synth_pairs(long n, uint64_t seed, double gamma, double Wmax, double smin, double smax,
                              double pmin, double pmax, float *__restrict__ X, float *__restrict__ Y,
                              int *__restrict__ ok)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double W = exp(u01(seed, i, 0) * log(Wmax));
  const double sig = exp(log(smin) + u01(seed, i, 1) * (log(smax) - log(smin)));
  const double ptr = exp(log(pmin) + u01(seed, i, 2) * (log(pmax) - log(pmin)));
  const double rh = exp(log(0.1) + u01(seed, i, 3) * (log(10.0) - log(0.1)));
  const double cv = 2.0 * u01(seed, i, 4) - 1.0, pv = 2.0 * M_PI * u01(seed, i, 5);
  const double cb = 2.0 * u01(seed, i, 6) - 1.0, pb = 2.0 * M_PI * u01(seed, i, 7);
  const double sv = sqrt(fmax(0.0, 1.0 - cv * cv)), sb = sqrt(fmax(0.0, 1.0 - cb * cb));
  const double v = sqrt(fmax(0.0, 1.0 - 1.0 / (W * W)));
  const double pre = rh * ptr;
  const double h = 1.0 + gamma / (gamma - 1.0) * ptr;
  const double b = sqrt(sig * rh * h);
  double P[8] = { rh, pre, v * sv * cos(pv), v * sv * sin(pv), v * cv,
                  b * sb * cos(pb), b * sb * sin(pb), b * cb };
  double U[8];
  const int err = prim_to_cons(P, U, gamma - 1.0);
  float x[MLP_IN];
  mlp_features(U, gamma, x);
  for (int k = 0; k < MLP_IN; ++k) X[i * MLP_IN + k] = x[k];
  const double Wt = 1.0 / sqrt(1.0 - (P[2]*P[2] + P[3]*P[3] + P[4]*P[4]));
  Y[i * 2 + 0] = (float)log(Wt);
  Y[i * 2 + 1] = (float)log(pre / U[0]);
  ok[i] = err == 0;
}

/* Pairs sampled from a grid's committed primitive state (SoA, interior zones). */
__global__ void k_state_pairs(long n, uint64_t seed, const double *__restrict__ P, long T, int N0, int N1,
                              int N2, int Ng, int n1, int n2, double gamma, float *__restrict__ X,
                              float *__restrict__ Y, int *__restrict__ ok)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long ni = (long)N0 * N1 * N2;
  long t = (long)(u01(seed, i, 9) * (double)ni);
  if (t >= ni) t = ni - 1;
  const long k = t % N2 + Ng, j = (t / N2) % N1 + Ng, ii = t / ((long)N2 * N1) + Ng;
  const long z = (ii * n1 + j) * n2 + k;
  double p[8], U[8];
  for (int q = 0; q < 8; ++q) p[q] = P[q * T + z];
  const int err = prim_to_cons(p, U, gamma - 1.0);
  float x[MLP_IN];
  mlp_features(U, gamma, x);
  for (int q = 0; q < MLP_IN; ++q) X[i * MLP_IN + q] = x[q];
  const double Wt = 1.0 / sqrt(1.0 - (p[2]*p[2] + p[3]*p[3] + p[4]*p[4]));
  Y[i * 2 + 0] = (float)log(Wt);
  Y[i * 2 + 1] = (float)log(p[1] / U[0]);
  ok[i] = err == 0;
}

struct Norms {
  float xmean[MLP_IN], xstd[MLP_IN], ymean[MLP_OUT], ystd[MLP_OUT];
};

/* Training forward: one wave per 64 samples of the minibatch. Stores the layer inputs
 * H0..H3 as bf16 [feature][batch] (H0 = standardised inputs, padded to 16 features),
 * the pre-activations per wave, and the outputs. idx == nullptr takes samples in order. */
__global__ __launch_bounds__(256) void k_train_fwd(const float *__restrict__ X, const float *__restrict__ Y,
                                                   const long *__restrict__ idx, int B, Norms nm,
                                                   const uint16_t *__restrict__ fwd, const float *__restrict__ th,
                                                   uint16_t *__restrict__ H, float *__restrict__ act,
                                                   float *__restrict__ yout, float *__restrict__ tout)
{
  __shared__ __align__(16) uint16_t xin[4][64 * MLP_KIN];
  __shared__ float ybuf[4][128];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long s = (long)blockIdx.x * 256 + threadIdx.x;        // sample within the batch
  const long src = idx ? idx[s] : s;
  {
    float xs[MLP_IN];
    for (int k = 0; k < MLP_IN; ++k) xs[k] = (X[src * MLP_IN + k] - nm.xmean[k]) * (1.0f / nm.xstd[k]);
    uint16_t *row = &xin[wave][lane * MLP_KIN];
    mlp_input_row(xs, row);
    for (int k = 0; k < MLP_KIN; ++k) H[(long)k * B + s] = row[k];
  }
  for (int o = 0; o < MLP_OUT; ++o) tout[s * 2 + o] = (Y[src * 2 + o] - nm.ymean[o]) / nm.ystd[o];
  __syncthreads();
  const long wave_global = (long)blockIdx.x * 4 + wave;
  float *a = act + wave_global * (3 * 64 * 64);
  mlp_forward_wave(xin[wave], fwd, th + OFF_B1, th + OFF_B2, th + OFF_B3, th + OFF_B4, ybuf[wave], a);
  __syncthreads();
  yout[s * 2 + 0] = ybuf[wave][lane * 2 + 0];
  yout[s * 2 + 1] = ybuf[wave][lane * 2 + 1];
  // layer inputs H1..H3 = bf16(hswish(pre-activation)), [feature][batch]
  for (int layer = 0; layer < 3; ++layer)
    for (int f = 0; f < 64; ++f)
      H[(long)(MLP_KIN + layer * 64 + f) * B + s] = f32_to_bf16(hswish(a[(layer * 64 + f) * 64 + lane]));
}

/* Training backward: loss gradient dY = 2 (y - t) / (2B), back-propagated with the
 * transposed products through layers 4, 3, 2; stores dZ_l (bf16, [feature][batch]) and
 * accumulates the loss. */
__global__ __launch_bounds__(256) void k_train_bwd(const float *__restrict__ yout, const float *__restrict__ tout,
                                                   int B, const uint16_t *__restrict__ bwd,
                                                   const float *__restrict__ act, uint16_t *__restrict__ dZ,
                                                   float *__restrict__ loss)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int grp = lane >> 4, col = lane & 15;
  const long wave_global = (long)blockIdx.x * 4 + wave;
  const long s0 = wave_global * 64;
  const float *a = act + wave_global * (3 * 64 * 64);
  // dZ4 rows 0, 1 (others 0): B fragments per n-tile
  const float scale = 2.0f / (2.0f * (float)B);
  bf16x4_t G[4][4];
  bf16x4_t D4[4];
  float lsum = 0.f;
  int g01;                                     // this lane's sample: bf16 dZ4 (out 0 | out 1 << 16)
  {
    const long s = s0 + lane;
    const float e0 = yout[s * 2 + 0] - tout[s * 2 + 0];
    const float e1 = yout[s * 2 + 1] - tout[s * 2 + 1];
    lsum = e0 * e0 + e1 * e1;
    const uint16_t g0 = f32_to_bf16(scale * e0), g1 = f32_to_bf16(scale * e1);
    dZ[(long)(3 * 64 + 0) * B + s] = g0;       // dZ4 rows after dZ1..dZ3
    dZ[(long)(3 * 64 + 1) * B + s] = g1;
    g01 = (int)((uint32_t)g0 | ((uint32_t)g1 << 16));
  }
  // B fragments of dZ4^T: lane (col, grp 0) takes sample 16 nt + col from its owner lane
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const uint32_t v = (uint32_t)__shfl(g01, 16 * nt + col, 64);
    const uint16_t g0 = grp == 0 ? (uint16_t)(v & 0xffff) : 0;
    const uint16_t g1 = grp == 0 ? (uint16_t)(v >> 16) : 0;
    D4[nt][0] = (short)g0; D4[nt][1] = (short)g1; D4[nt][2] = 0; D4[nt][3] = 0;
  }
  // dH3 = W4 dZ4^T, dZ3 = dH3 * hswish'(A3)
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const bf16x4_t w = load_frag(bwd, MLP_B4 + mt, lane);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4_t c = {0.f, 0.f, 0.f, 0.f};
      c = mfma16(w, D4[nt], c);
      float g[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int f = 16 * mt + 4 * grp + r, sl = 16 * nt + col;
        g[r] = c[r] * hswish_grad(a[(2 * 64 + f) * 64 + sl]);
      }
      G[mt][nt] = pack4(g[0], g[1], g[2], g[3]);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        dZ[(long)(2 * 64 + 16 * mt + 4 * grp + r) * B + s0 + 16 * nt + col] = (uint16_t)G[mt][nt][r];
    }
  }
  // layers 3 -> 2 and 2 -> 1
#pragma unroll
  for (int step = 0; step < 2; ++step) {
    const int f0 = step == 0 ? MLP_B3 : MLP_B2;
    const int alayer = step == 0 ? 1 : 0;     // pre-activation of the layer below
    f32x4_t C[4][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) C[mt][nt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) {
        const bf16x4_t w = load_frag(bwd, f0 + 4 * mt + kt, lane);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) C[mt][nt] = mfma16(w, G[kt][nt], C[mt][nt]);
      }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        float g[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int f = 16 * mt + 4 * grp + r, sl = 16 * nt + col;
          g[r] = C[mt][nt][r] * hswish_grad(a[(alayer * 64 + f) * 64 + sl]);
        }
        G[mt][nt] = pack4(g[0], g[1], g[2], g[3]);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          dZ[(long)(alayer * 64 + 16 * mt + 4 * grp + r) * B + s0 + 16 * nt + col] = (uint16_t)G[mt][nt][r];
      }
  }
  // loss (sum of squared standardised errors)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
  if (lane == 0) atomicAdd(loss, lsum);
}

/* Weight gradients dW^T (out x in) = sum over the batch of dZ (out x s) . H (s x in),
 * one wave per 16x16 tile and K-chunk, atomically accumulated; the tile with in-tile 0
 * also accumulates the bias gradient. Tiles: layer 1: 4 (out) x 1 (in, 16 padded);
 * layers 2, 3: 4 x 4; layer 4: 1 (2 used) x 4. */
__global__ __launch_bounds__(64) void k_train_wgrad(const uint16_t *__restrict__ H, const uint16_t *__restrict__ dZ,
                                                    int B, int kchunk, float *__restrict__ grad)
{
  const int lane = threadIdx.x;
  const int grp = lane >> 4, col = lane & 15;
  int tile = blockIdx.x;
  const int chunk = blockIdx.y;
  int layer, mt, nt;
  if (tile < 4) { layer = 1; mt = tile; nt = 0; }
  else if (tile < 20) { layer = 2; mt = (tile - 4) / 4; nt = (tile - 4) % 4; }
  else if (tile < 36) { layer = 3; mt = (tile - 20) / 4; nt = (tile - 20) % 4; }
  else { layer = 4; mt = 0; nt = tile - 36; }
  // rows of dZ (outputs of this layer) and of H (inputs of this layer)
  const long zrow = layer == 4 ? 3 * 64 : (long)(layer - 1) * 64;
  const long hrow = layer == 1 ? 0 : (long)MLP_KIN + (layer - 2) * 64;
  const int nout = layer == 4 ? 2 : 64;
  const int orow = 16 * mt + col;                  // A row of this lane
  const int irow = 16 * nt + col;                  // B column of this lane
  const bool arow_ok = orow < nout;
  const long k0 = (long)chunk * kchunk, k1 = k0 + kchunk < B ? k0 + kchunk : B;
  f32x4_t c = {0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (long k = k0; k < k1; k += 16) {
    const long s = k + 4 * grp;
    bf16x4_t a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint16_t av = arow_ok ? dZ[(zrow + orow) * B + s + j] : 0;
      a[j] = (short)av;
      b[j] = (short)H[(hrow + irow) * B + s + j];
      bsum += bf16_to_f32(av);
    }
    c = mfma16(a, b, c);
  }
  // write the 16x16 tile: lane holds rows (out) 16mt + 4grp + r, column (in) 16nt + col
  const int nin = layer == 1 ? 2 * MLP_IN : 64;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int o = 16 * mt + 4 * grp + r, i = 16 * nt + col;
    if (o < nout && i < nin) {
      long off;
      if (layer == 1) off = OFF_W1 + o * 6 + (i < MLP_IN ? i : i - MLP_IN);   // tied hi/lo columns
      else if (layer == 2) off = OFF_W2 + o * 64 + i;
      else if (layer == 3) off = OFF_W3 + o * 64 + i;
      else off = OFF_W4 + o * 64 + i;
      atomicAdd(&grad[off], c[r]);
    }
  }
  if (nt == 0) {
    bsum += __shfl_xor(bsum, 16, 64);
    bsum += __shfl_xor(bsum, 32, 64);
    if (grp == 0 && arow_ok) {
      const int boff = layer == 1 ? OFF_B1 : layer == 2 ? OFF_B2 : layer == 3 ? OFF_B3 : OFF_B4;
      atomicAdd(&grad[boff + orow], bsum);
    }
  }
}

__global__ void k_adam(float *__restrict__ th, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_gather_idx(long n, uint64_t seed, long step, int B, long *__restrict__ idx)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  const double u = u01(seed ^ 0x5bd1e995ULL, (uint64_t)step * (uint64_t)B + (uint64_t)i, 11);
  long k = (long)(u * (double)n);
  idx[i] = k < n ? k : n - 1;
}

/* Inference on host-provided standardised inputs (testing): y = f(x), raw outputs */
__global__ __launch_bounds__(256) void k_infer(const float *__restrict__ Xs, long n,
                                               const uint16_t *__restrict__ fwd, const float *__restrict__ th,
                                               float *__restrict__ Y)
{
  __shared__ __align__(16) uint16_t xin[4][64 * MLP_KIN];
  __shared__ float ybuf[4][128];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  {
    float xs[MLP_IN];
    for (int k = 0; k < MLP_IN; ++k) xs[k] = s < n ? Xs[s * MLP_IN + k] : 0.f;
    mlp_input_row(xs, &xin[wave][lane * MLP_KIN]);
  }
  __syncthreads();
  mlp_forward_wave(xin[wave], fwd, th + OFF_B1, th + OFF_B2, th + OFF_B3, th + OFF_B4, ybuf[wave], nullptr);
  __syncthreads();
  if (s < n) {
    Y[s * 2 + 0] = ybuf[wave][lane * 2 + 0];
    Y[s * 2 + 1] = ybuf[wave][lane * 2 + 1];
  }
}

}  // namespace

/* ===================================================================================== */
struct maragpu_mlp {
  float *th = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;   // device, NPARAM each
  uint16_t *fwd = nullptr, *bwd = nullptr;                          // device fragments
  Norms nm;
  int adam_t = 0;
};

namespace {

int swizzle(maragpu_mlp *net, hipStream_t s)
{
  const int total = (MLP_NFWD + MLP_NBWD) * MLP_FRAG;
  k_swizzle<<<(total + 255) / 256, 256, 0, s>>>(net->th, net->fwd, net->bwd);
  HIPOK(hipGetLastError());
  return 0;
}

int alloc_net(maragpu_mlp **out)
{
  maragpu_mlp *n = new (std::nothrow) maragpu_mlp;
  if (!n) return MARAGPU_E_NOMEM;
  if (hipMalloc(&n->th, NPARAM * 4) != hipSuccess || hipMalloc(&n->m, NPARAM * 4) != hipSuccess ||
      hipMalloc(&n->v, NPARAM * 4) != hipSuccess || hipMalloc(&n->g, NPARAM * 4) != hipSuccess ||
      hipMalloc(&n->fwd, MLP_NFWD * MLP_FRAG * 2) != hipSuccess ||
      hipMalloc(&n->bwd, MLP_NBWD * MLP_FRAG * 2) != hipSuccess) {
    maragpu_mlp_destroy(n);
    return MARAGPU_E_NOMEM;
  }
  if (hipMemset(n->m, 0, NPARAM * 4) != hipSuccess || hipMemset(n->v, 0, NPARAM * 4) != hipSuccess) {
    maragpu_mlp_destroy(n);
    return MARAGPU_E_HIP;
  }
  for (int k = 0; k < MLP_IN; ++k) { n->nm.xmean[k] = 0.f; n->nm.xstd[k] = 1.f; }
  for (int k = 0; k < MLP_OUT; ++k) { n->nm.ymean[k] = 0.f; n->nm.ystd[k] = 1.f; }
  *out = n;
  return 0;
}

}  // namespace

extern "C" {

int maragpu_mlp_create(maragpu_mlp **out, uint64_t seed)
{
  if (!out) return MARAGPU_E_INVALID;
  *out = nullptr;
  maragpu_mlp *n = nullptr;
  int rc = alloc_net(&n);
  if (rc) return rc;
  // He-style uniform initialisation from a host generator
  std::vector<float> th(NPARAM, 0.f);
  uint64_t s = seed * 0x9e3779b97f4a7c15ULL + 12345;
  auto uni = [&]() {
    s += 0x9e3779b97f4a7c15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return (float)(((z >> 11) + 0.5) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
  };
  auto fill = [&](int off, int nout, int nin) {
    const float a = std::sqrt(6.0f / (float)nin);
    for (int i = 0; i < nout * nin; ++i) th[off + i] = a * uni();
  };
  fill(OFF_W1, 64, 6);
  fill(OFF_W2, 64, 64);
  fill(OFF_W3, 64, 64);
  fill(OFF_W4, 2, 64);
  for (int i = 0; i < 2 * 64; ++i) th[OFF_W4 + i] *= 0.1f;
  if (hipMemcpy(n->th, th.data(), NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
      swizzle(n, nullptr) || hipDeviceSynchronize() != hipSuccess) {
    maragpu_mlp_destroy(n);
    return MARAGPU_E_HIP;
  }
  *out = n;
  return 0;
}

int maragpu_mlp_destroy(maragpu_mlp *n)
{
  if (!n) return MARAGPU_E_INVALID;
  hipFree(n->th); hipFree(n->m); hipFree(n->v); hipFree(n->g); hipFree(n->fwd); hipFree(n->bwd);
  delete n;
  return 0;
}

size_t maragpu_mlp_param_count(void) { return NPARAM; }

/* Serialized form: "MLPC" magic, version 2 (hard-swish activations; version 1 was SiLU),
 * NPARAM, then norms (20 floats) and the NPARAM master parameters (little-endian FP32). */
size_t maragpu_mlp_serialized_size(void) { return 12 + 20 * 4 + NPARAM * 4; }

int maragpu_mlp_save(const maragpu_mlp *n, void *buf, size_t nbytes)
{
  if (!n || !buf || nbytes < maragpu_mlp_serialized_size()) return MARAGPU_E_INVALID;
  char *p = static_cast<char *>(buf);
  const uint32_t hdr[3] = { 0x43504c4du, kSerialVersion, (uint32_t)NPARAM };
  std::memcpy(p, hdr, 12);
  float nm[20];
  for (int k = 0; k < 6; ++k) { nm[k] = n->nm.xmean[k]; nm[6 + k] = n->nm.xstd[k]; }
  for (int k = 0; k < 2; ++k) { nm[12 + k] = n->nm.ymean[k]; nm[14 + k] = n->nm.ystd[k]; }
  for (int k = 16; k < 20; ++k) nm[k] = 0.f;
  std::memcpy(p + 12, nm, 80);
  HIPOK(hipMemcpy(p + 92, n->th, NPARAM * 4, hipMemcpyDeviceToHost));
  return 0;
}

int maragpu_mlp_load(maragpu_mlp **out, const void *buf, size_t nbytes)
{
  if (!out || !buf || nbytes < maragpu_mlp_serialized_size()) return MARAGPU_E_INVALID;
  const char *p = static_cast<const char *>(buf);
  uint32_t hdr[3];
  std::memcpy(hdr, p, 12);
  if (hdr[0] != 0x43504c4du || hdr[1] != kSerialVersion || hdr[2] != (uint32_t)NPARAM) return MARAGPU_E_INVALID;
  *out = nullptr;
  maragpu_mlp *n = nullptr;
  int rc = alloc_net(&n);
  if (rc) return rc;
  float nm[20];
  std::memcpy(nm, p + 12, 80);
  for (int k = 0; k < 6; ++k) { n->nm.xmean[k] = nm[k]; n->nm.xstd[k] = nm[6 + k]; }
  for (int k = 0; k < 2; ++k) { n->nm.ymean[k] = nm[12 + k]; n->nm.ystd[k] = nm[14 + k]; }
  // a network is handed out only once its parameters and fragments are on the device
  if (hipMemcpy(n->th, p + 92, NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
      swizzle(n, nullptr) || hipDeviceSynchronize() != hipSuccess) {
    maragpu_mlp_destroy(n);
    return MARAGPU_E_HIP;
  }
  *out = n;
  return 0;
}

int maragpu_mlp_get_params(const maragpu_mlp *n, float *theta, float *norms)
{
  if (!n) return MARAGPU_E_INVALID;
  if (theta) HIPOK(hipMemcpy(theta, n->th, NPARAM * 4, hipMemcpyDeviceToHost));
  if (norms) {
    for (int k = 0; k < 6; ++k) { norms[k] = n->nm.xmean[k]; norms[6 + k] = n->nm.xstd[k]; }
    for (int k = 0; k < 2; ++k) { norms[12 + k] = n->nm.ymean[k]; norms[14 + k] = n->nm.ystd[k]; }
  }
  return 0;
}

int maragpu_mlp_set_params(maragpu_mlp *n, const float *theta, const float *norms)
{
  if (!n) return MARAGPU_E_INVALID;
  if (theta) {
    HIPOK(hipMemcpy(n->th, theta, NPARAM * 4, hipMemcpyHostToDevice));
    if (swizzle(n, nullptr)) return MARAGPU_E_HIP;
    HIPOK(hipDeviceSynchronize());
  }
  if (norms) {
    for (int k = 0; k < 6; ++k) { n->nm.xmean[k] = norms[k]; n->nm.xstd[k] = norms[6 + k]; }
    for (int k = 0; k < 2; ++k) { n->nm.ymean[k] = norms[12 + k]; n->nm.ystd[k] = norms[14 + k]; }
  }
  return 0;
}

/* Raw network on standardised inputs Xs (n x 6, host) -> raw outputs Y (n x 2, host). */
int maragpu_mlp_forward(const maragpu_mlp *net, size_t n, const float *Xs, float *Y)
{
  if (!net || !Xs || !Y) return MARAGPU_E_INVALID;
  if (n == 0) return 0;
  float *dX = nullptr, *dY = nullptr;
  HIPOK(hipMalloc(&dX, n * MLP_IN * 4));
  if (hipMalloc(&dY, n * 2 * 4) != hipSuccess) { hipFree(dX); return MARAGPU_E_NOMEM; }
  HIPOK(hipMemcpy(dX, Xs, n * MLP_IN * 4, hipMemcpyHostToDevice));
  k_infer<<<(unsigned)((n + 255) / 256), 256>>>(dX, (long)n, net->fwd, net->th, dY);
  HIPOK(hipGetLastError());
  HIPOK(hipMemcpy(Y, dY, n * 2 * 4, hipMemcpyDeviceToHost));
  hipFree(dX);
  hipFree(dY);
  return 0;
}

}  // extern "C"

namespace {

/* One optimisation step's gradient on a minibatch already laid out on the device.
 * B must be a multiple of 256. Leaves the gradient in net->g and the loss sum in *dloss. */
int batch_grad(maragpu_mlp *net, const float *dX, const float *dY, const long *didx, int B,
               uint16_t *H, uint16_t *dZ, float *act, float *yo, float *to, float *dloss, hipStream_t st)
{
  HIPOK(hipMemsetAsync(net->g, 0, NPARAM * 4, st));
  k_train_fwd<<<B / 256, 256, 0, st>>>(dX, dY, didx, B, net->nm, net->fwd, net->th, H, act, yo, to);
  k_train_bwd<<<B / 256, 256, 0, st>>>(yo, to, B, net->bwd, act, dZ, dloss);
  const int kchunk = 512;
  dim3 grid(40, (B + kchunk - 1) / kchunk);
  k_train_wgrad<<<grid, 64, 0, st>>>(H, dZ, B, kchunk, net->g);
  HIPOK(hipGetLastError());
  return 0;
}

struct Scratch {
  uint16_t *H = nullptr, *dZ = nullptr;
  float *act = nullptr, *yo = nullptr, *to = nullptr, *loss = nullptr;
  long *idx = nullptr;
  ~Scratch()
  {
    hipFree(H); hipFree(dZ); hipFree(act); hipFree(yo); hipFree(to); hipFree(loss); hipFree(idx);
  }
  int alloc(int B)
  {
    if (hipMalloc(&H, (size_t)(MLP_KIN + 3 * 64) * B * 2) != hipSuccess) return MARAGPU_E_NOMEM;
    if (hipMalloc(&dZ, (size_t)(3 * 64 + 2) * B * 2) != hipSuccess) return MARAGPU_E_NOMEM;
    if (hipMalloc(&act, (size_t)(B / 64) * 3 * 64 * 64 * 4) != hipSuccess) return MARAGPU_E_NOMEM;
    if (hipMalloc(&yo, (size_t)B * 2 * 4) != hipSuccess) return MARAGPU_E_NOMEM;
    if (hipMalloc(&to, (size_t)B * 2 * 4) != hipSuccess) return MARAGPU_E_NOMEM;
    if (hipMalloc(&loss, 4) != hipSuccess) return MARAGPU_E_NOMEM;
    if (hipMalloc(&idx, (size_t)B * 8) != hipSuccess) return MARAGPU_E_NOMEM;
    return 0;
  }
};

}  // namespace

extern "C" {

/* Gradient of the mean squared error (over n samples and both outputs, standardised
 * targets) at the current parameters: n standardised inputs and targets, host arrays.
 * n must be a multiple of 256. grad receives NPARAM floats. Used for checking the
 * training kernels against a CPU evaluation. */
int maragpu_mlp_grad(maragpu_mlp *net, size_t n, const float *Xs, const float *Ts, float *grad, float *loss)
{
  if (!net || !Xs || !Ts || !grad || n == 0 || n % 256) return MARAGPU_E_INVALID;
  const int B = (int)n;
  float *dX = nullptr, *dT = nullptr;
  Scratch sc;
  if (sc.alloc(B)) return MARAGPU_E_NOMEM;
  HIPOK(hipMalloc(&dX, n * MLP_IN * 4));
  HIPOK(hipMalloc(&dT, n * 2 * 4));
  HIPOK(hipMemcpy(dX, Xs, n * MLP_IN * 4, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(dT, Ts, n * 2 * 4, hipMemcpyHostToDevice));
  HIPOK(hipMemset(sc.loss, 0, 4));
  // identity standardisation: inputs and targets are given standardised
  const Norms keep = net->nm;
  for (int k = 0; k < 6; ++k) { net->nm.xmean[k] = 0.f; net->nm.xstd[k] = 1.f; }
  for (int k = 0; k < 2; ++k) { net->nm.ymean[k] = 0.f; net->nm.ystd[k] = 1.f; }
  int rc = batch_grad(net, dX, dT, nullptr, B, sc.H, sc.dZ, sc.act, sc.yo, sc.to, sc.loss, nullptr);
  net->nm = keep;
  if (rc) { hipFree(dX); hipFree(dT); return rc; }
  HIPOK(hipMemcpy(grad, net->g, NPARAM * 4, hipMemcpyDeviceToHost));
  float l = 0.f;
  HIPOK(hipMemcpy(&l, sc.loss, 4, hipMemcpyDeviceToHost));
  if (loss) *loss = l / (2.0f * (float)n);
  hipFree(dX);
  hipFree(dT);
  return 0;
}

}  // extern "C"

/* Training pairs on the device: standardisation-free features X[n][6] and targets
 * Y[n][2] (log W, log(p/D)); only pairs whose state passed prim_to_cons and whose values
 * are finite are kept. */
struct maragpu_mlp_data {
  float *X = nullptr, *Y = nullptr;
  long n = 0, cap = 0;
};

namespace {

/* Keeps the valid pairs among n fresh ones at (X, Y) + data->n, compacting them. */
int keep_valid(maragpu_mlp_data *d, long n, const int *dok)
{
  std::vector<float> hx((size_t)n * MLP_IN), hy((size_t)n * 2);
  std::vector<int> hok((size_t)n);
  float *X = d->X + (size_t)d->n * MLP_IN, *Y = d->Y + (size_t)d->n * 2;
  HIPOK(hipMemcpy(hx.data(), X, hx.size() * 4, hipMemcpyDeviceToHost));
  HIPOK(hipMemcpy(hy.data(), Y, hy.size() * 4, hipMemcpyDeviceToHost));
  HIPOK(hipMemcpy(hok.data(), dok, hok.size() * 4, hipMemcpyDeviceToHost));
  long m = 0;
  for (long i = 0; i < n; ++i) {
    bool good = hok[i] != 0 && std::isfinite(hy[i * 2]) && std::isfinite(hy[i * 2 + 1]);
    for (int k = 0; k < MLP_IN; ++k) good = good && std::isfinite(hx[i * MLP_IN + k]);
    if (!good) continue;
    if (m != i) {
      for (int k = 0; k < MLP_IN; ++k) hx[m * MLP_IN + k] = hx[i * MLP_IN + k];
      for (int k = 0; k < 2; ++k) hy[m * 2 + k] = hy[i * 2 + k];
    }
    ++m;
  }
  if (m != n) {
    HIPOK(hipMemcpy(X, hx.data(), (size_t)m * MLP_IN * 4, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(Y, hy.data(), (size_t)m * 2 * 4, hipMemcpyHostToDevice));
  }
  d->n += m;
  return 0;
}

struct DevInts {
  int *p = nullptr;
  ~DevInts() { hipFree(p); }
};

}  // namespace

extern "C" {

int maragpu_mlp_data_create(maragpu_mlp_data **out, int64_t capacity)
{
  if (!out || capacity <= 0) return MARAGPU_E_INVALID;
  maragpu_mlp_data *d = new (std::nothrow) maragpu_mlp_data;
  if (!d) return MARAGPU_E_NOMEM;
  if (hipMalloc(&d->X, (size_t)capacity * MLP_IN * 4) != hipSuccess ||
      hipMalloc(&d->Y, (size_t)capacity * 2 * 4) != hipSuccess) {
    hipFree(d->X);
    hipFree(d->Y);
    delete d;
    return MARAGPU_E_NOMEM;
  }
  d->cap = capacity;
  *out = d;
  return 0;
}

int maragpu_mlp_data_destroy(maragpu_mlp_data *d)
{
  if (!d) return MARAGPU_E_INVALID;
  hipFree(d->X);
  hipFree(d->Y);
  delete d;
  return 0;
}

int64_t maragpu_mlp_data_size(const maragpu_mlp_data *d) { return d ? d->n : -1; }

int maragpu_mlp_data_add_synthetic(maragpu_mlp_data *d, int64_t n, uint64_t seed, double gamma,
                                   const maragpu_mlp_sampling *r)
{
  if (!d || !r || n < 0 || d->n + n > d->cap || !(gamma > 1.0) || !(r->W_max >= 1.0) ||
      !(r->sigma_min > 0.0) || !(r->sigma_max >= r->sigma_min) || !(r->ptr_min > 0.0) ||
      !(r->ptr_max >= r->ptr_min))
    return MARAGPU_E_INVALID;
  if (n == 0) return 0;
  DevInts ok;
  HIPOK(hipMalloc(&ok.p, (size_t)n * 4));
  k_synth_pairs<<<(unsigned)((n + 255) / 256), 256>>>((long)n, seed, gamma, r->W_max, r->sigma_min, r->sigma_max,
                                                     r->ptr_min, r->ptr_max, d->X + (size_t)d->n * MLP_IN,
                                                     d->Y + (size_t)d->n * 2, ok.p);
  HIPOK(hipGetLastError());
  HIPOK(hipDeviceSynchronize());
  return keep_valid(d, (long)n, ok.p);
}

int maragpu_mlp_data_add_state(maragpu_mlp_data *d, maragpu_ctx *ctx, int64_t n, uint64_t seed, double gamma)
{
  if (!d || !ctx || n < 0 || d->n + n > d->cap || !(gamma > 1.0)) return MARAGPU_E_INVALID;
  if (n == 0) return 0;
  int N[3], Ng, n1, n2;
  long T;
  const double *P = maragpu_committed_prim(ctx, N, &Ng, &T, &n1, &n2);
  if (!P) return MARAGPU_E_INVALID;
  DevInts ok;
  HIPOK(hipMalloc(&ok.p, (size_t)n * 4));
  k_state_pairs<<<(unsigned)((n + 255) / 256), 256>>>((long)n, seed, P, T, N[0], N[1], N[2], Ng, n1, n2, gamma,
                                                     d->X + (size_t)d->n * MLP_IN, d->Y + (size_t)d->n * 2, ok.p);
  HIPOK(hipGetLastError());
  HIPOK(hipDeviceSynchronize());
  return keep_valid(d, (long)n, ok.p);
}

/* Sets the standardisation from the data (a strided subsample of at most 10^6 pairs)
 * and trains with Adam on minibatches drawn with replacement, learning rate decaying
 * from lr to lr_final on a cosine. */
int maragpu_mlp_train(maragpu_mlp *net, const maragpu_mlp_data *d, const maragpu_mlp_train_config *cfg,
                      maragpu_mlp_train_report *rep)
{
  if (!net || !d || !cfg || cfg->batch <= 0 || cfg->batch % 256 || cfg->epochs < 0 || d->n < cfg->batch)
    return MARAGPU_E_INVALID;
  const long n = d->n;
  const float *X = d->X, *Y = d->Y;
  const auto t0 = std::chrono::steady_clock::now();
  {
    const long m = n < 1000000 ? n : 1000000, stride = n / m;
    std::vector<float> hx((size_t)n * MLP_IN), hy((size_t)n * 2);
    HIPOK(hipMemcpy(hx.data(), X, hx.size() * 4, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(hy.data(), Y, hy.size() * 4, hipMemcpyDeviceToHost));
    double sx[MLP_IN] = {0}, sxx[MLP_IN] = {0}, sy[2] = {0}, syy[2] = {0};
    long cnt = 0;
    for (long i = 0; i < n; i += stride) {
      ++cnt;
      for (int k = 0; k < MLP_IN; ++k) { sx[k] += hx[i * MLP_IN + k]; sxx[k] += (double)hx[i * MLP_IN + k] * hx[i * MLP_IN + k]; }
      for (int k = 0; k < 2; ++k) { sy[k] += hy[i * 2 + k]; syy[k] += (double)hy[i * 2 + k] * hy[i * 2 + k]; }
    }
    for (int k = 0; k < MLP_IN; ++k) {
      const double mu = sx[k] / cnt, var = sxx[k] / cnt - mu * mu;
      net->nm.xmean[k] = (float)mu;
      net->nm.xstd[k] = (float)(var > 1e-12 ? std::sqrt(var) : 1.0);
    }
    for (int k = 0; k < 2; ++k) {
      const double mu = sy[k] / cnt, var = syy[k] / cnt - mu * mu;
      net->nm.ymean[k] = (float)mu;
      net->nm.ystd[k] = (float)(var > 1e-12 ? std::sqrt(var) : 1.0);
    }
  }
  const int B = cfg->batch;
  Scratch sc;
  if (sc.alloc(B)) return MARAGPU_E_NOMEM;
  const long steps = (long)cfg->epochs * (n / B);
  double first = -1.0, last = 0.0;
  const int nb = (NPARAM + 255) / 256;
  float hloss = 0.f;
  double acc = 0.0;
  long accn = 0;
  for (long s = 0; s < steps; ++s) {
    k_gather_idx<<<B / 256, 256>>>(n, cfg->seed, s, B, sc.idx);
    HIPOK(hipMemsetAsync(sc.loss, 0, 4, nullptr));
    if (batch_grad(net, X, Y, sc.idx, B, sc.H, sc.dZ, sc.act, sc.yo, sc.to, sc.loss, nullptr)) return MARAGPU_E_HIP;
    const double frac = steps > 1 ? (double)s / (double)(steps - 1) : 1.0;
    const double lr = cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1.0 + std::cos(M_PI * frac));
    ++net->adam_t;
    const float c1 = 1.f - std::pow(0.9f, (float)net->adam_t), c2 = 1.f - std::pow(0.999f, (float)net->adam_t);
    k_adam<<<nb, 256>>>(net->th, net->m, net->v, net->g, NPARAM, (float)lr, 0.9f, 0.999f, 1e-8f, c1, c2);
    if (swizzle(net, nullptr)) return MARAGPU_E_HIP;
    if (s < 20 || s >= steps - 50 || s % 200 == 0) {
      HIPOK(hipMemcpy(&hloss, sc.loss, 4, hipMemcpyDeviceToHost));
      const double l = hloss / (2.0 * B);
      if (first < 0) first = l;
      if (s >= steps - 50) { acc += l; ++accn; }
      last = l;
    }
  }
  HIPOK(hipDeviceSynchronize());
  if (rep) {
    std::memset(rep, 0, sizeof(*rep));
    rep->loss_first = first;
    rep->loss_last = accn ? acc / accn : last;
    rep->steps = steps;
    rep->samples = n;
    rep->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return 0;
}

/* Device-side parameters for the c2p kernel. */
int maragpu_mlp_device_params(const maragpu_mlp *n, void *out_params)
{
  if (!n || !out_params) return MARAGPU_E_INVALID;
  MlpDev d;
  d.fwd = n->fwd;
  d.bwd = n->bwd;
  d.b1 = n->th + OFF_B1;
  d.b2 = n->th + OFF_B2;
  d.b3 = n->th + OFF_B3;
  d.b4 = n->th + OFF_B4;
  for (int k = 0; k < MLP_IN; ++k) { d.xmean[k] = n->nm.xmean[k]; d.xinv[k] = 1.0f / n->nm.xstd[k]; }
  for (int k = 0; k < MLP_OUT; ++k) { d.ymean[k] = n->nm.ymean[k]; d.ystd[k] = n->nm.ystd[k]; }
  std::memcpy(out_params, &d, sizeof(d));
  return 0;
}

}  // extern "C"
