/* Learned conserved-to-primitive initial guess: a 6 -> 64 -> 64 -> 64 -> 2 MLP
 * evaluated with BF16 matrix cores and FP32 accumulation: v_mfma_f32_16x16x16_bf16 for the
 * first layer (16 input columns) and gfx950's v_mfma_f32_16x16x32_bf16, twice the k per
 * instruction, for the three products over the 64 hidden features.
 *
 * Activation: hard-swish, a * clamp(a/6 + 1/2, 0, 1) (Howard et al. 2019, "Searching for
 * MobileNetV3"), the piecewise-quadratic form of SiLU a * sigmoid(a). It costs two
 * full-rate VALU operations where SiLU costs an exponential and a reciprocal, both
 * quarter-rate: with 192 activations per zone that was most of the network's time.
 *
 * Layout idea. Each wave evaluates 64 samples. Every layer is computed transposed,
 *     Z^T (out x samples) = W^T (out x in) . H^T (in x samples),
 * so that the MFMA result layout (lane l holds rows 4*(l/16)+r, column l%16 of a 16x16
 * tile) is exactly the B-operand layout the next layer needs (lane l holds k = 4*(l/16)+r
 * of column l%16): activations stay in registers from layer to layer. Weights are stored
 * pre-swizzled in "fragment order" (per 16x16 tile, 64 lanes x 4 bf16; per 16x32 tile of
 * the double-k product, 64 lanes x 8 bf16), so a lane loads its A operand with one 8- or
 * 16-byte read.
 * The double-k product pairs slot s of lane group g = l/16 in A with the same slot and group
 * in B, whatever k the hardware gives that pair, so a fixed assignment of the 32 k of a tile
 * to (g, s) gives the product as long as the A fragments are stored with it. Here
 * k = 4g + s for s < 4 and 16 + 4g + (s - 4) for s >= 4: a lane's B operand for k-tile t is
 * then the two activation fragments it already holds for m-tiles 2t and 2t + 1 of the
 * previous layer.
 *
 * Features of a conserved state U = (D, tau, S, B) (all scale-free; D normalises):
 *   x0 = log(tau/D), x1 = log(1 + |S|^2/D^2), x2 = log(1 + |B|^2/D),
 *   x3 = S.B / (|S||B| + tiny), x4 = -log(1 - |S|^2/E^2) with E = tau + D, x5 = Gamma.
 * x4 is -2 log W for cold unmagnetised flow; it is formed in FP64 as
 * log(E^2) - log((E - |S|)(E + |S|)), free of the cancellation in 1 - |S|/E that would
 * otherwise hide W > 10 below the inputs' precision. (|S.B| is x3 |S| |B| and needs no
 * feature of its own.)
 * standardised with the training set's mean and standard deviation. Each standardised
 * feature enters the first layer as two bfloat16 columns, hi = bf16(x) and
 * lo = bf16(x - hi), that share one weight column: the first product then sees x to about
 * 16 mantissa bits instead of 8. Outputs (after de-standardising): y0 = log W,
 * y1 = log(p/D).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace maragpu {

constexpr int MLP_IN = 6;       // features
constexpr int MLP_KIN = 16;     // padded input width (one MFMA k-tile)
constexpr int MLP_H = 64;       // hidden width
constexpr int MLP_OUT = 2;      // outputs
constexpr int MLP_MOUT = 16;    // padded output width (one MFMA m-tile)
constexpr int MLP_FRAG = 256;   // bf16 per 16x16 fragment (64 lanes x 4)

// forward (A = W^T) fragments, in MLP_FRAG units: layer 1: 4 m-tiles x 1 k-tile of 16;
// layers 2, 3: 4 m-tiles x 2 k-tiles of 32 (two units each, tile (mt, kt) at 2 (2 mt + kt));
// layer 4: 1 x 2 k-tiles of 32
constexpr int MLP_F1 = 0, MLP_F2 = 4, MLP_F3 = 20, MLP_F4 = 36, MLP_NFWD = 40;
// backward (A = W) fragments for the hidden-state gradients: layer 4: 4 m-tiles (in) x
// 1 k-tile (out); layers 3, 2: 4 x 4
constexpr int MLP_B4 = 0, MLP_B3 = 4, MLP_B2 = 20, MLP_NBWD = 36;

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

/* Parameters on the device. */
struct MlpDev {
  const uint16_t *fwd;      // MLP_NFWD fragments (bf16 bits)
  const uint16_t *bwd;      // MLP_NBWD fragments
  const float *b1, *b2, *b3, *b4;  // biases (64, 64, 64, 2)
  float xmean[MLP_IN], xinv[MLP_IN];   // standardisation: (x - xmean) * xinv
  float ymean[MLP_OUT], ystd[MLP_OUT];
};

typedef __bf16 bf16x2v_t __attribute__((ext_vector_type(2)));
typedef float f32x2v_t __attribute__((ext_vector_type(2)));
typedef unsigned u32x2_t __attribute__((ext_vector_type(2)));

/* FP32 -> bfloat16, round to nearest even, with gfx950's v_cvt_pk_bf16_f32 (two values
 * per instruction). */
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b)
{
  const f32x2v_t v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2v_t));
}
__device__ __forceinline__ uint16_t f32_to_bf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

/* Hard-swish and its derivative (0 below -3, 1 above 3, a/3 + 1/2 between); the results
 * are rounded to bfloat16 before use. */
__device__ __forceinline__ float hswish(float a)
{
  return a * __builtin_amdgcn_fmed3f(__builtin_fmaf(a, 1.0f / 6.0f, 0.5f), 0.0f, 1.0f);
}
__device__ __forceinline__ float hswish_grad(float a)
{
  return a <= -3.0f ? 0.0f : (a >= 3.0f ? 1.0f : __builtin_fmaf(a, 1.0f / 3.0f, 0.5f));
}

/* Orders the LDS accesses of one wave: what a lane wrote before it is visible to every
 * lane of the wave after it (LDS operations of a wave complete in order). */
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* The first layer's input row of one sample: standardised features xs[MLP_IN] as
 * hi/lo bfloat16 pairs in columns k and MLP_IN + k, zero padding after 2 MLP_IN. */
__device__ __forceinline__ void mlp_input_row(const float *xs, uint16_t *row)
{
  uint32_t *w = reinterpret_cast<uint32_t *>(row);     // row is 4-byte aligned (MLP_KIN even)
  float lo[MLP_IN];
#pragma unroll
  for (int k = 0; k < MLP_IN; k += 2) {
    const uint32_t hi = pack_bf16x2(xs[k], xs[k + 1]);
    w[k / 2] = hi;
    lo[k] = xs[k] - __uint_as_float(hi << 16);
    lo[k + 1] = xs[k + 1] - __uint_as_float(hi & 0xffff0000u);
  }
#pragma unroll
  for (int k = 0; k < MLP_IN; k += 2) w[(MLP_IN + k) / 2] = pack_bf16x2(lo[k], lo[k + 1]);
#pragma unroll
  for (int k = 2 * MLP_IN; k < MLP_KIN; k += 2) w[k / 2] = 0u;
}

__device__ __forceinline__ f32x4_t mfma16(bf16x4_t a, bf16x4_t b, f32x4_t c)
{
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4_t load_frag(const uint16_t *frags, int f, int lane)
{
  const uint2 v = *reinterpret_cast<const uint2 *>(frags + (size_t)f * MLP_FRAG + lane * 4);
  bf16x4_t r;
  r[0] = (short)(v.x & 0xffff);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffff);
  r[3] = (short)(v.y >> 16);
  return r;
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short s16x8_t __attribute__((ext_vector_type(8)));

/* gfx950's 16x16x32 BF16 product (k assignment in the header comment). */
__device__ __forceinline__ f32x4_t mfma32(bf16x8_t a, bf16x8_t b, f32x4_t c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

/* A 16x32 fragment: two consecutive MLP_FRAG units starting at unit f. */
__device__ __forceinline__ bf16x8_t load_frag32(const uint16_t *frags, int f, int lane)
{
  const uint4 v = *reinterpret_cast<const uint4 *>(frags + (size_t)f * MLP_FRAG + lane * 8);
  return __builtin_bit_cast(bf16x8_t, v);
}

/* The B operand of k-tile t from the fragments of m-tiles 2t (slots 0-3) and 2t + 1. */
__device__ __forceinline__ bf16x8_t cat8(bf16x4_t lo, bf16x4_t hi)
{
  const s16x8_t v = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf16x8_t, v);
}

__device__ __forceinline__ bf16x4_t pack4(float a, float b, float c, float d)
{
  const u32x2_t v = {pack_bf16x2(a, b), pack_bf16x2(c, d)};
  return __builtin_bit_cast(bf16x4_t, v);
}

/* Features of one conserved state (FP32), before standardisation. */
__device__ __forceinline__ void mlp_features(const double *U, double gamma, float *x)
{
  const double D = U[0], tau = U[1];
  const double S2 = U[2]*U[2] + U[3]*U[3] + U[4]*U[4];
  const double B2 = U[5]*U[5] + U[6]*U[6] + U[7]*U[7];
  const double SB = U[2]*U[5] + U[3]*U[6] + U[4]*U[7];
  // ratios in FP64 (E - |S| needs it) with the hardware reciprocal (v_rcp_f64, far below
  // the FP32 precision of the features), logarithms in FP32: the features are FP32
  const double iD = __builtin_amdgcn_rcp(D);
  x[0] = logf((float)fmax(tau * iD, 1e-30));
  x[1] = log1pf((float)(S2 * iD * iD));
  x[2] = log1pf((float)(B2 * iD));
  const float q = (float)(S2 * B2);
  x[3] = q > 0.0f ? (float)SB * rsqrtf(q) : 0.0f;
  const double E = tau + D, S = sqrt(S2);
  x[4] = logf((float)(E * E * __builtin_amdgcn_rcp(fmax((E - S) * (E + S), 1e-300))));
  x[5] = (float)gamma;
}

/* One wave evaluates the network on 64 samples. xin: standardised features of the
 * wave's samples in LDS, [64][MLP_KIN] bf16 bits (columns >= MLP_IN zero); wf: forward
 * fragments (LDS or global); y: per lane, the two raw outputs of the sample whose index
 * equals the lane (returned through LDS scratch yout[64][2]).
 * If act != nullptr the pre-activations are stored for back-propagation:
 * act[layer][feature][sample] (f32, 3 x 64 x 64), samples local to the wave. */
__device__ __forceinline__ void mlp_forward_wave(const uint16_t *xin, const uint16_t *wf,
                                                 const float *b1, const float *b2, const float *b3,
                                                 const float *b4, float *yout, float *act)
{
  const int lane = threadIdx.x & 63;
  const int grp = lane >> 4, col = lane & 15;
  bf16x4_t H[4][4];       // [k-tile][n-tile] B operands of the next layer
  // layer 1 input fragments: lane holds features 4*grp..4*grp+3 of sample 16*nt + col
  bf16x4_t X[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const uint2 v = *reinterpret_cast<const uint2 *>(xin + (16 * nt + col) * MLP_KIN + 4 * grp);
    X[nt][0] = (short)(v.x & 0xffff);
    X[nt][1] = (short)(v.x >> 16);
    X[nt][2] = (short)(v.y & 0xffff);
    X[nt][3] = (short)(v.y >> 16);
  }
  // layer 1: 64 x 16 . 16 x 64; accumulators start at the bias
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const bf16x4_t a = load_frag(wf, MLP_F1 + mt, lane);
    const f32x4_t bias = *reinterpret_cast<const f32x4_t *>(b1 + 16 * mt + 4 * grp);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const f32x4_t c = mfma16(a, X[nt], bias);
      float h[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (act) act[(0 * 64 + 16 * mt + 4 * grp + r) * 64 + 16 * nt + col] = c[r];
        h[r] = hswish(c[r]);
      }
      H[mt][nt] = pack4(h[0], h[1], h[2], h[3]);
    }
  }
  // layers 2 and 3: 64 x 64 . 64 x 64, in two halves of 32 samples: a half's new
  // activations replace its old ones only after all of its products are formed, which
  // keeps 32 accumulators live instead of 64
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const int f0 = layer == 0 ? MLP_F2 : MLP_F3;
    const float *bb = layer == 0 ? b2 : b3;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      f32x4_t C[4][2];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        const f32x4_t bias = *reinterpret_cast<const f32x4_t *>(bb + 16 * mt + 4 * grp);
        C[mt][0] = bias;
        C[mt][1] = bias;
      }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) {
          const bf16x8_t a = load_frag32(wf, f0 + 2 * (2 * mt + kt), lane);
#pragma unroll
          for (int j = 0; j < 2; ++j)
            C[mt][j] = mfma32(a, cat8(H[2 * kt][2 * half + j], H[2 * kt + 1][2 * half + j]), C[mt][j]);
        }
      }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int nt = 2 * half + j;
          float h[4];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (act) act[((layer + 1) * 64 + 16 * mt + 4 * grp + r) * 64 + 16 * nt + col] = C[mt][j][r];
            h[r] = hswish(C[mt][j][r]);
          }
          H[mt][nt] = pack4(h[0], h[1], h[2], h[3]);
        }
      }
    }
  }
  // layer 4: 16 (2 used) x 64 . 64 x 64
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    f32x4_t c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
      c = mfma32(load_frag32(wf, MLP_F4 + 2 * kt, lane), cat8(H[2 * kt][nt], H[2 * kt + 1][nt]), c);
    if (grp == 0) {
      yout[(16 * nt + col) * 2 + 0] = c[0] + b4[0];
      yout[(16 * nt + col) * 2 + 1] = c[1] + b4[1];
    }
  }
}

}  // namespace maragpu
