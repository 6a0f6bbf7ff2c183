/* maragpu — device-resident 3D relativistic-MHD update for MI355X (gfx950).
 *
 * Data layout in HBM: structure of arrays, one plane of T doubles per field, T = all
 * zones of the augmented (ghost-including) grid, flat zone index z = (i*n1 + j)*n2 + k
 * as in Mara's AoS layout (Mara/valman.hpp:25-36) with the field index moved outward.
 *   Pc   committed primitives (the user's state)   8 x T doubles
 *   Pw   working primitives / c2p guesses          8 x T doubles
 *   U0   conserved state at the start of the step  8 x T doubles
 *   U    conserved state of the current RK stage   8 x T doubles (written only for a
 *        next stage that reads it: rk2 and rk3 stages after the first)
 *   F[3] face fluxes per axis (double, or float in the mixed form)
 *   A    RK4 accumulator (rk4 only)
 *
 * One step (luaC_Mara_advance, Mara/mara.cpp:297-341) is:
 *   k_p2c_c2p         per interior zone: PrimToCons of the committed state and the first
 *                     c2p pass, its result written to the zone and its periodic ghost
 *                     images (replaces Synchronize + per-zone c2p)
 * then per RK stage:
 *   k_flux<DIM> x 3   per interface over the reference's flat sweep range
 *                     (Mara/plm-split.cpp:48): reconstruct 8 fields, Riemann solve,
 *                     PCM retry, max |lambda| (for HLLD: k_flux_hlld_fast + k_flux_redo)
 *   k_update_c2p      per interior zone: constrained-transport EMF averages
 *                     (Mara/rmhd.cpp:293-337), flux divergence (plm-split.cpp:117-119) and
 *                     the RK stage combination (Mara/runge-kutta.hpp), all in FP64, then
 *                     the zone's next c2p pass from the new U in registers
 * The state is committed only if no zone failed.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <vector>

#include "rmhd_device.hpp"
#include "mlp_device.hpp"
#include "eigensystem_device.hpp"
#include "ou_field.hpp"
#include "../../include/maragpu.h"

using namespace maragpu;

namespace {

constexpr int kBlock = 256;
// block size of k_flux_hlld_fast (it ends without a block barrier): FP64 in one-wave blocks
// (33.6 ms of sweeps per plain 256^3 step, against 34.1 with 2-wave and 35.5 with 4-wave
// blocks), FP32 in blocks of 4 (8-wave blocks: 16.1 against 14.6 ms; 2-wave blocks: 14.5)
template <typename R> constexpr int kFluxBlock = sizeof(R) == 4 ? 256 : 64;
// occupancy targets of k_flux_hlld_fast (waves per SIMD), measured at 256^3 WENO5/HLLD:
// FP64 at 2 (213 VGPRs, no scratch; 4.0 ms, against 4.7 at 3 with 184 B spilled),
// FP32 at 5 (96 VGPRs, 40 B/lane scratch; at 4 waves, with none, 9 % slower)
constexpr int kFastFluxWaves64 = 2;
constexpr int kFastFluxWaves32 = 5;
constexpr int kRedoBlocks = 1024;    // grid of k_flux_redo (a grid-stride loop over its queue)

/* Device-side counters of one step; reduced per block, then atomically. */
struct DevStats {
  unsigned long long lam_bits;         // max |lambda| as the bits of a non-negative double
  unsigned long long hlld_success;
  unsigned long long hlld_failure;
  unsigned long long retries;
  unsigned long long failures;
  unsigned long long c2p_zones;
  unsigned long long c2p_iters;
  unsigned long long c2p_fallbacks;
  unsigned long long hlld_c2p_iters;
  unsigned long long hlld_newton_iters;
  unsigned long long mlp_guesses;
  unsigned long long mlp_fallbacks;
  unsigned long long redo;             // mixed form: interfaces recomputed in FP64
  unsigned long long redo_overflow;    // ... that did not fit the queue (FP32 HLLC kept)
  unsigned long long stage_failed[6];  // failed zones per c2p pass (ghost images included)
  unsigned long long stage_reset[6];   // zones reset by the emergency repair per pass (weno-split)
  unsigned long long eig_identity;     // weno-split interfaces whose eigensystem was rejected
  unsigned long long ctu_failed[4][4]; // plm-muscl: failed checks per stage and abort point
};

/* Interfaces of an HLLD sweep handed to the full FP64 solver (k_flux_redo) before the
 * update reads the fluxes: in the mixed form those whose FP32 HLLD solve failed, in the
 * plain form those outside the common path of the lean sweep (k_flux_hlld_fast). */
struct RedoQueue {
  unsigned int *count;     // device counter, cleared before each sweep
  unsigned int cap;        // slots
  unsigned int *z;         // device slots: interface index z (left zone; T < 2^32)
  uint4 *part;             // per wave of the HLLD sweep: its counters, summed by the redo
  unsigned int part_cap;   // waves
};

struct GridDesc {
  int N[3], Ng, n[3];
  long T;                  // zones including ghosts
  long sx, sy, sz;         // zone strides
  double dx, dy, dz, min_dx;
};

#define HIP_CHECK(expr)                                                            \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "[maragpu] %s failed: %s\n", #expr, hipGetErrorString(e_)); \
      return MARAGPU_E_HIP;                                                        \
    }                                                                              \
  } while (0)

/* --------------------------------------------------------------------------
 * block reductions
 * ------------------------------------------------------------------------ */
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

/* Sums NV counters over the block and adds them to dst[0..NV). */
template <int NV, int BLOCK = kBlock>
__device__ __forceinline__ void block_add(const unsigned long long (&v)[NV], unsigned long long *const (&dst)[NV])
{
  __shared__ unsigned long long red[BLOCK / 64][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const unsigned long long s = wave_sum(v[i]);
    if (lane == 0) red[wave][i] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    unsigned long long s = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) s += red[w][threadIdx.x];
    if (s) atomicAdd(dst[threadIdx.x], s);
  }
}

/* Wave reductions by DPP moves (no LDS round trip): within each row of 16 lanes by quad
 * permutations and row rotations, then across the four rows by the row broadcasts of
 * lanes 15 and 31; lane 63 ends with the wave's total, which is read into a scalar. Every
 * lane of the wave must be active. Lanes a move does not reach keep the identity (0). */
constexpr int kDppQuadXor1 = 0xB1, kDppQuadXor2 = 0x4E, kDppRowRor4 = 0x124, kDppRowRor8 = 0x128;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ unsigned dpp_u32(unsigned v)
{
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, false);
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_f64(double v)
{
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = dpp_u32<CTRL, ROW_MASK>((unsigned)b), hi = dpp_u32<CTRL, ROW_MASK>((unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ unsigned wave_sum_dpp(unsigned v)
{
  v += dpp_u32<kDppQuadXor1>(v);
  v += dpp_u32<kDppQuadXor2>(v);
  v += dpp_u32<kDppRowRor4>(v);
  v += dpp_u32<kDppRowRor8>(v);
  v += dpp_u32<kDppRowBcast15, 0xA>(v);
  v += dpp_u32<kDppRowBcast31, 0xC>(v);
  return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}
/* maximum of non-negative doubles (NaN is dropped, as fmax drops it) */
__device__ __forceinline__ double wave_max_dpp(double v)
{
  v = fmax(v, dpp_f64<kDppQuadXor1>(v));
  v = fmax(v, dpp_f64<kDppQuadXor2>(v));
  v = fmax(v, dpp_f64<kDppRowRor4>(v));
  v = fmax(v, dpp_f64<kDppRowRor8>(v));
  v = fmax(v, dpp_f64<kDppRowBcast15, 0xA>(v));
  v = fmax(v, dpp_f64<kDppRowBcast31, 0xC>(v));
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

/* The sweep's closing counters, per wave: max |lambda| (the bits of a non-negative
 * double) and the HLLD success, nested-c2p and Newton iteration counts and failed states,
 * stored by the wave's first lane in part[wave] (x, y: the bits of max |lambda|; z:
 * successes + 256 x nested-c2p iterations; w: Newton iterations + 65536 x failed states)
 * for the redo kernel to sum. No block
 * barrier and no atomic: a wave retires as soon as its own interfaces are done (a
 * per-wave atomic serialises on the counters' addresses, and the block's barrier holds
 * finished waves until the slowest of the block is done). */
__device__ __forceinline__ void sweep_counters_wave(double lam, bool success, unsigned c2p_iters,
                                                    unsigned newton_iters, bool failed, uint4 *part)
{
  const double m = wave_max_dpp(lam);
  const unsigned ns = (unsigned)__popcll(__ballot(success)), nf = (unsigned)__popcll(__ballot(failed));
  const unsigned nc = wave_sum_dpp(c2p_iters), nn = wave_sum_dpp(newton_iters);
  if ((threadIdx.x & 63) == 0) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(m);
    part[blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)] =
        make_uint4((unsigned)b, (unsigned)(b >> 32), ns + (nc << 8), nn + (nf << 16));
  }
}

/* Appends z for every lane of the wave with push set, with one atomic per wave; every
 * lane of the wave calls it. Returns false on a lane whose slot lies past the capacity. */
__device__ __forceinline__ bool queue_push(const RedoQueue &rq, bool push, long z)
{
  const unsigned long long m = __ballot(push);
  if (m == 0ull) return true;
  const int lane = (int)(threadIdx.x & 63);
  const int leader = __ffsll((unsigned long long)m) - 1;
  unsigned int base = 0u;
  if (lane == leader) base = atomicAdd(rq.count, (unsigned int)__popcll(m));
  base = (unsigned int)__shfl((int)base, leader);
  if (!push) return true;
  const unsigned int k = base + (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
  if (k >= rq.cap) return false;
  rq.z[k] = (unsigned int)z;
  return true;
}

template <int BLOCK = kBlock>
__device__ __forceinline__ void block_max_lambda(double lam, unsigned long long *dst)
{
  __shared__ double red[BLOCK / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // NaN never enters: fmax drops it, as `if (MaxLambda < ml)` does in the reference
  const double m = wave_max(lam);
  if (lane == 0) red[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) b = fmax(b, red[w]);
    if (b > 0.0) atomicMax(dst, (unsigned long long)__double_as_longlong(b));
  }
}

/* --------------------------------------------------------------------------
 * layout conversion and closed forms
 * ------------------------------------------------------------------------ */
__global__ void k_aos_to_soa(const double *__restrict__ aos, double *__restrict__ soa, long T)
{
  const long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= T) return;
#pragma unroll
  for (int q = 0; q < 8; ++q) soa[q * T + z] = aos[8 * z + q];
}

__global__ void k_soa_to_aos(const double *__restrict__ soa, double *__restrict__ aos, long T)
{
  const long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= T) return;
#pragma unroll
  for (int q = 0; q < 8; ++q) aos[8 * z + q] = soa[q * T + z];
}

__device__ __forceinline__ long interior_zone(const GridDesc &g, long t)
{
  const long nzk = g.N[2], nyk = g.N[1];
  const long k = t % nzk;
  const long j = (t / nzk) % nyk;
  const long i = t / (nzk * nyk);
  return ((i + g.Ng) * g.n[1] + (j + g.Ng)) * g.n[2] + (k + g.Ng);
}

/* one field's interior, (i, j, k) order with k fastest, out of / into its SoA plane: the
 * hyperslab DataManagerHDF5 moves between Mara's array and a prim/<field> dataset
 * (modules/unigrid.lua:243-292). Into the plane, every periodic image is written too. */
__global__ void k_field_out(const double *__restrict__ plane, double *__restrict__ out, GridDesc g)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.N[0] * g.N[1] * g.N[2]) return;
  out[t] = plane[interior_zone(g, t)];
}

__global__ void k_field_in(const double *__restrict__ in, double *__restrict__ plane, GridDesc g)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.N[0] * g.N[1] * g.N[2]) return;
  const int k = (int)(t % g.N[2]) + g.Ng;
  const int j = (int)((t / g.N[2]) % g.N[1]) + g.Ng;
  const int i = (int)(t / ((long)g.N[2] * g.N[1])) + g.Ng;
  const double v = in[t];
  int li[3], lj[3], lk[3];
  int ni = 0, nj = 0, nk = 0;
  li[ni++] = i; if (i >= g.N[0]) li[ni++] = i - g.N[0]; if (i < 2 * g.Ng) li[ni++] = i + g.N[0];
  lj[nj++] = j; if (j >= g.N[1]) lj[nj++] = j - g.N[1]; if (j < 2 * g.Ng) lj[nj++] = j + g.N[1];
  lk[nk++] = k; if (k >= g.N[2]) lk[nk++] = k - g.N[2]; if (k < 2 * g.Ng) lk[nk++] = k + g.N[2];
  for (int a = 0; a < ni; ++a)
    for (int b = 0; b < nj; ++b)
      for (int c = 0; c < nk; ++c) plane[((long)li[a] * g.n[1] + lj[b]) * g.n[2] + lk[c]] = v;
}

/* Writes the 8 fields of interior zone (i,j,k) to the zone and to each periodic ghost
 * image of it: after this every ghost holds its image, which is what
 * SimpleCartesianDomain::Synchronize produces (Mara/simple-cart.cpp:81-170). Returns the
 * number of copies (the zone's multiplicity in Mara's all-zone loops). */
__device__ __forceinline__ int scatter_images(const GridDesc &g, double *__restrict__ P, long T,
                                              int i, int j, int k, const double *v)
{
  int li[3], lj[3], lk[3];
  int ni = 0, nj = 0, nk = 0;
  li[ni++] = i; if (i >= g.N[0]) li[ni++] = i - g.N[0]; if (i < 2 * g.Ng) li[ni++] = i + g.N[0];
  lj[nj++] = j; if (j >= g.N[1]) lj[nj++] = j - g.N[1]; if (j < 2 * g.Ng) lj[nj++] = j + g.N[1];
  lk[nk++] = k; if (k >= g.N[2]) lk[nk++] = k - g.N[2]; if (k < 2 * g.Ng) lk[nk++] = k + g.N[2];
  for (int a = 0; a < ni; ++a)
    for (int b = 0; b < nj; ++b)
      for (int c = 0; c < nk; ++c) {
        const long z = ((long)li[a] * g.n[1] + lj[b]) * g.n[2] + lk[c];
#pragma unroll
        for (int q = 0; q < 8; ++q) P[q * T + z] = v[q];
      }
  return ni * nj * nk;
}

/* weno-split reads the conserved state of every zone, ghosts included, as Uglb after
 * GodunovOperator::ConsToPrim (Mara/weno-split.cpp:26-28): the c2p kernels write U to each
 * zone's images (Uw). When the emergency repair reset zones in that pass, the reference
 * rebuilds every zone's U from the repaired primitives (Mara/hydro.cpp:186-193); this
 * kernel does the same, and returns at once when no zone was reset. */
__global__ void k_conserved_after_reset(const double *__restrict__ P, double *__restrict__ Uw, GridDesc g,
                                        double gamma, const unsigned long long *__restrict__ resets)
{
  if (*resets == 0ull) return;
  const long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long T = g.T;
  if (z >= T) return;
  double p[8], u[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) p[q] = P[q * T + z];
  prim_to_cons(p, u, gamma - 1.0);
#pragma unroll
  for (int q = 0; q < 8; ++q) Uw[q * T + z] = u[q];
}

/* Newton start (Z, W) of the anton2dzw solve from the network's raw outputs: W = exp(y0)
 * clamped to [1, 1e6], p = D exp(y1), rho = D / W, h = 1 + Gamma/(Gamma-1) p / rho,
 * Z = rho h W^2. The exponentials are FP32 (the network's outputs carry a few BF16 digits:
 * the start only needs to lie in Newton's basin). */
__device__ __forceinline__ void learned_start(double D, float y0, float y1, const MlpDev &net, double gamma,
                                              double &Zg, double &Wg)
{
  double W = (double)__expf(y0 * net.ystd[0] + net.ymean[0]);
  W = W > 1.0 ? (W < 1e6 ? W : 1e6) : 1.0;
  const double pg = D * (double)__expf(y1 * net.ystd[1] + net.ymean[1]);
  const double iW = rcp(W);
  const double rg = D * iW;
  const double hg = 1.0 + gamma * rcp(gamma - 1.0) * pg * rcp(rg);
  Zg = rg * hg * W * W;
  Wg = W;
}

/* The network in LDS: forward fragments and biases, staged once per block, and per-wave
 * scratch for the layer-1 input rows and the outputs. */
struct MlpLds {
  uint16_t wf[MLP_NFWD * MLP_FRAG];
  float bias[3 * 64 + 2];
  uint16_t xin[kBlock / 64][64 * MLP_KIN];
  float yout[kBlock / 64][128];
};

/* Every thread of the block calls this, then the block synchronises. */
__device__ __forceinline__ void stage_net(MlpLds &L, const MlpDev &net)
{
  for (int i = threadIdx.x; i < MLP_NFWD * MLP_FRAG / 4; i += kBlock)
    reinterpret_cast<uint2 *>(L.wf)[i] = reinterpret_cast<const uint2 *>(net.fwd)[i];
  for (int i = threadIdx.x; i < 3 * 64 + 2; i += kBlock)
    L.bias[i] = i < 64 ? net.b1[i] : i < 128 ? net.b2[i - 64] : i < 192 ? net.b3[i - 128] : net.b4[i - 192];
}

/* The network's Newton start for each lane's conserved state u. Every lane of the wave
 * calls it together (a lane without a zone passes a placeholder state); one wave evaluates
 * its 64 states on the matrix cores. */
__device__ __forceinline__ void learned_guess(MlpLds &L, const MlpDev &net, const double *u, double gamma,
                                              double &Zg, double &Wg)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float x[MLP_IN];
  mlp_features(u, gamma, x);
#pragma unroll
  for (int q = 0; q < MLP_IN; ++q) x[q] = (x[q] - net.xmean[q]) * net.xinv[q];
  mlp_input_row(x, &L.xin[wave][lane * MLP_KIN]);
  wave_lds_sync();
  mlp_forward_wave(L.xin[wave], L.wf, L.bias, L.bias + 64, L.bias + 128, L.bias + 192, L.yout[wave], nullptr);
  wave_lds_sync();
  const float y0 = L.yout[wave][lane * 2 + 0], y1 = L.yout[wave][lane * 2 + 1];
  wave_lds_sync();   // the wave's next call rewrites xin and yout
  learned_start(u[0], y0, y1, net, gamma, Zg, Wg);
}

/* The conserved-to-primitive solve of one zone from a given Newton start (the network's,
 * or the zone's previous solution): anton2dzw from (Zg, Wg); where that fails, the plain
 * cascade (GodunovOperator::ConsToPrim, Mara/rmhd.cpp:59-131) with the previous primitive
 * state, which p holds on entry when prev_valid and which is read (prev()) only then.
 * Returns the cascade's error. With zw, the solution's (Z, W) are returned there as the
 * zone's next start (FP32: a start needs no more). */
template <typename Prev>
__device__ __forceinline__ int learned_c2p(const double *u, double *p, double Zg, double Wg, double gamma,
                                           double pfloor, bool prev_valid, Prev prev, int &iters, int &used,
                                           float2 *zw = nullptr)
{
  int which = 0;
  int err = check_cons(u);
  used = 0;
  double zwd[2];
  if (!err) {
    C2PState s;
    c2p_new_state(s, u, gamma, pfloor);
    err = c2p_anton2dzw(s, Zg, Wg, p, iters, zwd);
    used = err == 0;
  }
  if (err) {
    prev(p);
    err = cons_to_prim(u, p, gamma, pfloor, prev_valid, iters, which);
    if (!err && zw) c2p_zw_of_prim(p, gamma, zwd[0], zwd[1]);
  }
  if (!err && zw) *zw = make_float2((float)zwd[0], (float)zwd[1]);
  return err;
}

__global__ __launch_bounds__(kBlock) void kb_cons_to_prim_learned(long n, const double *__restrict__ U,
                                                                  double *__restrict__ P, double gamma,
                                                                  double pfloor, int *err_out, int *iters_out,
                                                                  int *used_out, double *start_out, MlpDev net)
{
  __shared__ __align__(16) uint16_t xin[kBlock / 64][64 * MLP_KIN];
  __shared__ float yout[kBlock / 64][128];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = t < n;
  double u[8], p[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    u[q] = valid ? U[8 * t + q] : (q < 2 ? 1.0 : 0.0);
    p[q] = valid ? P[8 * t + q] : 0.0;
  }
  float x[MLP_IN];
  mlp_features(u, gamma, x);
#pragma unroll
  for (int q = 0; q < MLP_IN; ++q) x[q] = (x[q] - net.xmean[q]) * net.xinv[q];
  mlp_input_row(x, &xin[wave][lane * MLP_KIN]);
  __syncthreads();
  mlp_forward_wave(xin[wave], net.fwd, net.b1, net.b2, net.b3, net.b4, yout[wave], nullptr);
  __syncthreads();
  if (!valid) return;
  const float y0 = yout[wave][lane * 2 + 0], y1 = yout[wave][lane * 2 + 1];
  int iters = 0, which = 0, used = 0;
  double Zg = 0.0, Wg = 0.0;
  int err = check_cons(u);
  if (!err) {
    C2PState s;
    c2p_new_state(s, u, gamma, pfloor);
    learned_start(u[0], y0, y1, net, gamma, Zg, Wg);
    err = c2p_anton2dzw(s, Zg, Wg, p, iters);
    used = err == 0;
    if (err) err = cons_to_prim(u, p, gamma, pfloor, true, iters, which);
  }
  if (start_out) {
    start_out[2 * t + 0] = Zg;
    start_out[2 * t + 1] = Wg;
  }
  err_out[t] = err;
  iters_out[t] = iters;
  used_out[t] = used;
#pragma unroll
  for (int q = 0; q < 8; ++q) P[8 * t + q] = p[q];
}

/* --------------------------------------------------------------------------
 * Flux sweep along axis DIM (MethodOfLinesSplit::intercell_flux_sweep,
 * Mara/plm-split.cpp:43-74): one thread per interface z | z+S over the reference's
 * flat range z in [2S, T-3S), S the zone stride of the axis.
 * ------------------------------------------------------------------------ */
template <typename R, int DIM>
__device__ __forceinline__ int riemann_dispatch(int solver, const R *pl, const R *pr, R *F, R gamma,
                                                double pfloor, R &lam, FluxCounters &cnt)
{
  if (solver == RIEMANN_HLLD) return riemann_hlld<R, DIM>(pl, pr, F, gamma, pfloor, lam, cnt);
  if (solver == RIEMANN_HLLC) return riemann_hllc<R, DIM>(pl, pr, F, gamma, lam);
  return riemann_hll<R, DIM>(pl, pr, F, gamma, lam);
}

/* One interface z | z+S of the sweep along DIM: reconstruction from the 6-point stencil,
 * the Riemann solve, and the retry with the zone-centred states of the two neighbours
 * (Mara/plm-split.cpp:51-71). Returns the error of the last attempt (flux 0 then). */
template <typename R, int DIM, int SOLVER>
__device__ __forceinline__ int interface_flux(const double *__restrict__ P, long T, long S, long z, int recon,
                                              const PhysParams &pp, R *F, R &lam, FluxCounters &fc,
                                              int &retried)
{
  R ql[8], qr[8];
  const R theta = (R)pp.theta;
  for (int q = 0; q < 8; ++q) {
    const double *Pq = P + q * T + z;
    const R v0 = (R)Pq[0], v1 = (R)Pq[S];
    if (recon == RECON_WENO5) {
      const R vm2 = (R)Pq[-2 * S], vm1 = (R)Pq[-S], v2 = (R)Pq[2 * S], v3 = (R)Pq[3 * S];
      ql[q] = weno5<R, true>(vm2, vm1, v0, v1, v2, pp.IS, (R)pp.sz10A);
      qr[q] = weno5<R, false>(vm1, v0, v1, v2, v3, pp.IS, (R)pp.sz10A);
    }
    else if (recon == RECON_PLM) {
      const R vm1 = (R)Pq[-S], v2 = (R)Pq[2 * S];
      ql[q] = plm<R>(vm1, v0, v1, R(+1.0), theta);
      qr[q] = plm<R>(v0, v1, v2, R(-1.0), theta);
    }
    else {
      ql[q] = v0;
      qr[q] = v1;
    }
  }
  int err = 1;
  retried = 0;
  // one call site: the reconstructed states, then the zone-centred states
#pragma unroll Shape<R>::unroll
  for (int attempt = 0; attempt < 2 && err; ++attempt) {
    if (attempt == 1) {
      retried = 1;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        ql[q] = (R)P[q * T + z];
        qr[q] = (R)P[q * T + z + S];
      }
    }
    err = riemann_dispatch<R, DIM>(SOLVER, ql, qr, F, (R)pp.gamma, pp.pfloor, lam, fc);
  }
  if (err) {
#pragma unroll
    for (int q = 0; q < 8; ++q) F[q] = R(0);
  }
  return err;
}

/* The common path of an HLLD interface: reconstruction and riemann_hlld in FAST mode.
 * Returns kHlldDefer, with F unwritten, where the full solver is needed. */
template <typename R, int DIM>
__device__ __forceinline__ int interface_flux_fast(const double *__restrict__ P, long T, long S, long z,
                                                   int recon, const PhysParams &pp, R *F, R &lam,
                                                   FluxCounters &fc)
{
  R ql[8], qr[8];
  const R theta = (R)pp.theta;
  for (int q = 0; q < 8; ++q) {
    const double *Pq = P + q * T + z;
    const R v0 = (R)Pq[0], v1 = (R)Pq[S];
    if (recon == RECON_WENO5) {
      const R vm2 = (R)Pq[-2 * S], vm1 = (R)Pq[-S], v2 = (R)Pq[2 * S], v3 = (R)Pq[3 * S];
      ql[q] = weno5<R, true>(vm2, vm1, v0, v1, v2, pp.IS, (R)pp.sz10A);
      qr[q] = weno5<R, false>(vm1, v0, v1, v2, v3, pp.IS, (R)pp.sz10A);
    }
    else if (recon == RECON_PLM) {
      ql[q] = plm<R>((R)Pq[-S], v0, v1, R(+1.0), theta);
      qr[q] = plm<R>(v0, v1, (R)Pq[2 * S], R(-1.0), theta);
    }
    else {
      ql[q] = v0;
      qr[q] = v1;
    }
  }
  return riemann_hlld<R, DIM, true>(ql, qr, F, (R)pp.gamma, pp.pfloor, lam, fc);
}

/* Flux sweep along DIM, one thread per interface. With a redo queue (mixed form, HLLD),
 * an interface whose FP32 HLLD solve fell back to HLLC or failed is queued for
 * k_flux_redo instead; its FP32 result stays only if the queue is full. */
/* Occupancy targets (waves per SIMD) of the full sweep, which runs HLL and HLLC: FP32 at
 * 3, FP64 at 2 even where 256 registers spill (the spill traffic costs less than the
 * latency that a single wave cannot hide). */
template <typename R, typename FT, int DIM, int SOLVER>
__global__ __launch_bounds__(kBlock, sizeof(R) == 4 ? 3 : 2) void k_flux(const double *__restrict__ P, FT *__restrict__ Fout,
                                                 GridDesc g, int recon, PhysParams pp, DevStats *st,
                                                 RedoQueue rq)
{
  const long T = g.T;
  const long S = DIM == 1 ? g.sx : DIM == 2 ? g.sy : g.sz;
  const long z = 2 * S + (long)blockIdx.x * blockDim.x + threadIdx.x;
  double lam_out = 0.0;
  unsigned long long cnt[7] = {0, 0, 0, 0, 0, 0, 0};
  bool want = false;
  if (z < T - 3 * S) {
    R lam = R(0);
    FluxCounters fc = {0, 0, 0, 0};
    R F[8];
    int retried;
    const int err = interface_flux<R, DIM, SOLVER>(P, T, S, z, recon, pp, F, lam, fc, retried);
    want = rq.z && (fc.hlld_fail || err);
#pragma unroll
    for (int q = 0; q < 8; ++q) Fout[q * T + z] = (FT)F[q];
    lam_out = (double)lam;
    cnt[0] = (unsigned long long)fc.hlld_ok;
    cnt[1] = (unsigned long long)fc.hlld_fail;
    cnt[2] = (unsigned long long)retried;
    cnt[3] = err ? 1ull : 0ull;
    cnt[4] = (unsigned long long)fc.c2p_iters;
    cnt[5] = (unsigned long long)fc.newton_iters;
  }
  if (rq.z) {
    if (!queue_push(rq, want, z)) cnt[6] = 1;     // full: the FP32 result stays
    else if (want) cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0;   // counted by the redo
  }
  block_max_lambda(lam_out, &st->lam_bits);
  unsigned long long *const dst[7] = { &st->hlld_success, &st->hlld_failure, &st->retries,
                                       &st->failures, &st->hlld_c2p_iters, &st->hlld_newton_iters,
                                       &st->redo_overflow };
  block_add<7>(cnt, dst);
}

/* Block order of the HLLD sweeps. The interfaces of a sweep along an axis of stride S are
 * z = a S + r over planes a in [2, T/S - 3) and offsets r in [0, S) (the reference's flat
 * range [2S, T-3S)); an interface's stencil lies at a-3..a+2 in the same column r.
 *   - The hardware hands blocks to the eight XCDs (each with its own L2) round-robin.
 *     Logical block L is given to XCD L / (nb/8): each XCD works through one contiguous
 *     eighth of the logical order, so the planes a stencil re-reads are in its own L2.
 *   - Along x (S = one i-plane, 4.4 MB of primitives at 256^3) the logical order is tiled:
 *     columns of kSweepChunk tiles of r are swept over every plane before the next
 *     column starts, so the six planes a stencil spans are a few tile-rows apart, not
 *     whole planes. Along y and z the flat order already keeps them within a few blocks.
 * The interfaces, and so every result, are those of the flat order; only their order on
 * the chip changes. */
constexpr long kSweepChunk = 8;

/* Number of blocks of BLK threads of the sweep along DIM. */
template <int DIM>
long sweep_blocks(const GridDesc &g, long BLK)
{
  const long S = DIM == 1 ? g.sx : DIM == 2 ? g.sy : g.sz;
  if (DIM == 1) return (g.T / S - 5) * ((S + BLK - 1) / BLK);
  return (g.T - 5 * S + BLK - 1) / BLK;
}

/* The calling thread's interface z of the sweep; false where it has none. The block
 * arithmetic is 32-bit: it runs once per wave on the scalar unit, and 64-bit divisions
 * there cost several hundred instructions (block counts stay far below 2^31). */
template <int DIM>
__device__ __forceinline__ bool sweep_interface(const GridDesc &g, long &z)
{
  const long S = DIM == 1 ? g.sx : DIM == 2 ? g.sy : g.sz;
  const unsigned BLK = blockDim.x, nb = gridDim.x, b = blockIdx.x;
  const unsigned q = nb >> 3;
  const unsigned L = b < 8 * q ? (b & 7) * q + (b >> 3) : b;
  if constexpr (DIM == 1) {
    constexpr unsigned C = (unsigned)kSweepChunk;
    const unsigned ntiles = (unsigned)((S + BLK - 1) / BLK);
    const unsigned nA = (unsigned)(g.T / S - 5);
    const unsigned full = ntiles / C, per_full = nA * C;
    unsigned a, tile;
    if (L < full * per_full) {
      const unsigned rem = L % per_full;
      a = rem / C;
      tile = (L / per_full) * C + rem % C;
    }
    else {
      const unsigned width = ntiles - full * C, rem = L - full * per_full;
      a = rem / width;
      tile = full * C + rem % width;
    }
    const long r = (long)tile * BLK + threadIdx.x;
    z = (2 + (long)a) * S + r;
    return r < S;
  }
  else {
    z = 2 * S + (long)L * BLK + threadIdx.x;
    return z < g.T - 3 * S;
  }
}

/* Block order of the stage update. A zone's update reads the flux planes at its own
 * index and at +-1 in i, j and k (the constrained-transport averages), so in the flat order
 * the i - 1 and i + 1 planes come back from beyond L2 (an i-plane of the three flux arrays
 * is 6.3 MB in FP32 at 256^3). As in the sweeps, XCD x takes a contiguous eighth of the
 * logical blocks, and the logical order runs in columns of kUpdateChunk tiles (a tile is a
 * block's 256 consecutive zones of an i-plane) over all i-planes before the next column
 * starts. Every zone is updated once, whatever the order. Measured against the flat order:
 * 512^3 update 44.5 -> 42.5 ms per plain step and 30.0 -> 28.6 mixed; 256^3 unchanged
 * (plain) to 1.5 % slower (mixed); columns of 4 or 64 tiles, or the XCD split alone, were
 * slower. */
constexpr unsigned kUpdateChunk = 16;

/* Blocks of the update's grid: N0 planes of ceil(N1 N2 / kBlock) tiles. */
inline unsigned update_blocks(const int *N)
{
  const long plane = (long)N[1] * N[2];
  return (unsigned)(N[0] * ((plane + kBlock - 1) / kBlock));
}

/* The calling thread's interior zone t (flat (i, j, k) order, 0 <= t < N0 N1 N2); false
 * where it has none (the last tile of a plane). */
__device__ __forceinline__ bool update_zone_index(const GridDesc &g, long &t)
{
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned q = nb >> 3;
  const unsigned L = b < 8 * q ? (b & 7) * q + (b >> 3) : b;
  const unsigned plane = (unsigned)g.N[1] * (unsigned)g.N[2];
  const unsigned ntiles = (plane + kBlock - 1) / kBlock;
  const unsigned nA = (unsigned)g.N[0];
  constexpr unsigned C = kUpdateChunk;
  const unsigned full = ntiles / C, per_full = nA * C;
  unsigned a, tile;
  if (L < full * per_full) {
    const unsigned rem = L % per_full;
    a = rem / C;
    tile = (L / per_full) * C + rem % C;
  }
  else {
    const unsigned width = ntiles - full * C, rem = L - full * per_full;
    a = rem / width;
    tile = full * C + rem % width;
  }
  const unsigned r = tile * kBlock + threadIdx.x;
  t = (long)a * plane + r;
  return r < plane;
}

/* HLLD sweep along DIM, common path only (interface_flux_fast): interfaces that need
 * the c2p cascade, a later Newton attempt, the small-normal-field branch, the HLLC
 * fallback or the retry are queued for k_flux_redo, which computes them with the full
 * FP64 solver. In the plain form that is the result the full sweep gives; in the mixed
 * form (R = float) the rare branches run in FP64 instead of FP32. Leaving them out of this
 * kernel lowers its register peak. The queue holds every interface of the sweep, so it
 * cannot overflow. */
template <typename R, int DIM>
__global__ __launch_bounds__(kFluxBlock<R>, sizeof(R) == 4 ? kFastFluxWaves32 : kFastFluxWaves64)
void k_flux_hlld_fast(const double *__restrict__ P, R *__restrict__ Fout, GridDesc g, int recon, PhysParams pp,
                      DevStats *st, RedoQueue rq)
{
  const long T = g.T;
  const long S = DIM == 1 ? g.sx : DIM == 2 ? g.sy : g.sz;
  long z;
  const bool has = sweep_interface<DIM>(g, z);
  double lam_out = 0.0;
  unsigned long long cnt[3] = {0, 0, 0};
  bool defer = false;
  if (has) {
    R lam = R(0);
    FluxCounters fc = {0, 0, 0, 0};
    R F[8];
    defer = interface_flux_fast<R, DIM>(P, T, S, z, recon, pp, F, lam, fc) != 0;
    if (!defer) {
#pragma unroll
      for (int q = 0; q < 8; ++q) Fout[q * T + z] = F[q];
      cnt[0] = 1;
      cnt[1] = (unsigned long long)fc.c2p_iters;
      cnt[2] = (unsigned long long)fc.newton_iters;
    }
    lam_out = (double)lam;
  }
  queue_push(rq, defer, z);
  sweep_counters_wave(lam_out, cnt[0] != 0, (unsigned)cnt[1], (unsigned)cnt[2], false, rq.part);
}

/* The queued interfaces of a sweep, recomputed by the full FP64 method (reconstruction,
 * HLLD attempt ladder, HLLC fallback, retry) and written over the sweep's fluxes (FP32
 * in the mixed form, FP64 in the plain one). A grid-stride loop over the queue count. */
template <typename FT, int DIM, int SOLVER>
__global__ __launch_bounds__(kBlock) void k_flux_redo(const double *__restrict__ P, FT *__restrict__ Fout,
                                                      GridDesc g, int recon, PhysParams pp, DevStats *st,
                                                      RedoQueue rq, unsigned int nparts)
{
  const unsigned int n = min(*rq.count, rq.cap);
  double lam_out = 0.0;
  unsigned long long cnt[7] = {0, 0, 0, 0, 0, 0, 0};
  // the sweep's per-wave counters (nparts waves; 0 for a pass without them)
  for (unsigned int w = blockIdx.x * blockDim.x + threadIdx.x; w < nparts; w += gridDim.x * blockDim.x) {
    const uint4 v = rq.part[w];
    lam_out = fmax(lam_out, __longlong_as_double((long long)(((unsigned long long)v.y << 32) | v.x)));
    cnt[0] += (unsigned long long)(v.z & 255u);
    cnt[4] += (unsigned long long)(v.z >> 8);
    cnt[5] += (unsigned long long)(v.w & 0xffffu);
  }
  for (unsigned int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const long T = g.T;
    const long S = DIM == 1 ? g.sx : DIM == 2 ? g.sy : g.sz;
    const long z = (long)rq.z[k];
    double lam = 0.0;
    FluxCounters fc = {0, 0, 0, 0};
    double F[8];
    int retried;
    const int err = interface_flux<double, DIM, SOLVER>(P, T, S, z, recon, pp, F, lam, fc, retried);
#pragma unroll
    for (int q = 0; q < 8; ++q) Fout[q * T + z] = (FT)F[q];
    lam_out = fmax(lam_out, lam);
    cnt[0] += (unsigned long long)fc.hlld_ok;
    cnt[1] += (unsigned long long)fc.hlld_fail;
    cnt[2] += (unsigned long long)retried;
    cnt[3] += err ? 1ull : 0ull;
    cnt[4] += (unsigned long long)fc.c2p_iters;
    cnt[5] += (unsigned long long)fc.newton_iters;
    cnt[6] += 1;
  }
  block_max_lambda(lam_out, &st->lam_bits);
  unsigned long long *const dst[7] = { &st->hlld_success, &st->hlld_failure, &st->retries,
                                       &st->failures, &st->hlld_c2p_iters, &st->hlld_newton_iters,
                                       &st->redo };
  block_add<7>(cnt, dst);
}

/* --------------------------------------------------------------------------
 * weno-split (WenoSplit, Mara/weno-split.cpp) with local Lax-Friedrichs flux splitting.
 * ------------------------------------------------------------------------ */
/* Per zone, for the sweep along DIM: the physical flux F and the fastest speed
 * A = max(|a+|, |a-|) (drive_single_sweep, Mara/weno-split.cpp:224-236), over every zone;
 * MaxLambda is their maximum. Uw holds U with its ghost images (as Uglb after ConsToPrim). */
template <typename R, int DIM>
__global__ __launch_bounds__(kBlock) void k_zone_flux(const double *__restrict__ Uw, const double *__restrict__ P,
                                                      R *__restrict__ FZ, R *__restrict__ AZ, GridDesc g,
                                                      PhysParams pp, DevStats *st)
{
  const long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long T = g.T;
  double lam_out = 0.0;
  if (z < T) {
    R u[8], p[8], f[8], ap, am;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      u[q] = (R)Uw[q * T + z];
      p[q] = (R)P[q * T + z];
    }
    flux_and_eigenvalues<R, DIM>(u, p, f, ap, am, (R)pp.gamma);
    const R a = (dfabs(ap) > dfabs(am)) ? dfabs(ap) : dfabs(am);
#pragma unroll
    for (int q = 0; q < 8; ++q) FZ[q * T + z] = f[q];
    AZ[z] = a;
    lam_out = (double)a;
  }
  block_max_lambda(lam_out, &st->lam_bits);
}

/* WenoSplit::intercell_flux_sweep, local Lax-Friedrichs branch (Mara/weno-split.cpp:163-207),
 * component-wise projection: one thread per interface z | z+S whose index along DIM lies in
 * [Ng-1, N+Ng-1] (every transverse line, ghosts included). With ml the largest A over the
 * 6-zone stencil, each field's split fluxes 0.5 (F +- ml U) are reconstructed by the
 * finite-difference WENO5 (C2R at z of the + part, C2L at z+S of the - part) and summed. The
 * reference projects with L and R of FluidEquations::Eigensystem; the identity leaves every
 * value as it is. */
template <typename R, typename FT, int DIM>
__global__ __launch_bounds__(kBlock) void k_flux_weno_split(const double *__restrict__ Uw, const R *__restrict__ FZ,
                                                            const R *__restrict__ AZ, FT *__restrict__ Fout,
                                                            GridDesc g, PhysParams pp)
{
  const long c0 = DIM == 1 ? g.N[0] + 1 : g.n[0];
  const long c1 = DIM == 2 ? g.N[1] + 1 : g.n[1];
  const long c2 = DIM == 3 ? g.N[2] + 1 : g.n[2];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c0 * c1 * c2) return;
  const long k = t % c2 + (DIM == 3 ? g.Ng - 1 : 0);
  const long j = (t / c2) % c1 + (DIM == 2 ? g.Ng - 1 : 0);
  const long i = t / (c2 * c1) + (DIM == 1 ? g.Ng - 1 : 0);
  const long T = g.T;
  const long S = DIM == 1 ? g.sx : DIM == 2 ? g.sy : g.sz;
  const long z = (i * g.n[1] + j) * g.n[2] + k;
  R ml = AZ[z - 2 * S];
#pragma unroll
  for (int m = -1; m <= 3; ++m) {
    const R a = AZ[z + m * S];
    ml = a > ml ? a : ml;
  }
  const R h = R(0.5);
#pragma unroll 1
  for (int q = 0; q < 8; ++q) {
    R fp[6], fm[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const long zm = z + (m - 2) * S;
      const R F = FZ[q * T + zm], U = (R)Uw[q * T + zm];
      fp[m] = h * (F + ml * U);
      fm[m] = h * (F - ml * U);
    }
    const R f = weno5<R, true, true>(fp[0], fp[1], fp[2], fp[3], fp[4], pp.IS, (R)pp.sz10A) +
                weno5<R, false, true>(fm[1], fm[2], fm[3], fm[4], fm[5], pp.IS, (R)pp.sz10A);
    Fout[q * T + z] = (FT)f;
  }
}

/* The same sweep with the RMHD characteristic projection (eigen = RMHD): the eigensystem
 * of the averaged primitive state 0.5 (P_z + P_z+S) in FP64 (L and R; the identity where
 * it is rejected), the 6-zone split fluxes projected with L, the finite-difference WENO5
 * of every characteristic field, and the projection of their sum back with R
 * (Mara/weno-split.cpp:176-203, the eigensystem call at :181). In the mixed form (R =
 * float) only the zone fluxes and speeds are FP32.
 * Eight lanes share one interface (eig::g8): lane q holds row q of L and R, projects and
 * reconstructs characteristic field q, and writes flux component q. */
template <typename R, int DIM>
__global__ __launch_bounds__(kBlock) void k_flux_weno_split_char(const double *__restrict__ Uw,
                                                                 const double *__restrict__ P,
                                                                 const R *__restrict__ FZ, const R *__restrict__ AZ,
                                                                 R *__restrict__ Fout, GridDesc g, PhysParams pp,
                                                                 DevStats *st)
{
#pragma clang fp contract(off)
  const long c0 = DIM == 1 ? g.N[0] + 1 : g.n[0];
  const long c1 = DIM == 2 ? g.N[1] + 1 : g.n[1];
  const long c2 = DIM == 3 ? g.N[2] + 1 : g.n[2];
  const long t = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 3;     // alike in the 8 lanes
  const int lane = (int)(threadIdx.x & 7);
  if (t >= c0 * c1 * c2) return;
  const long k = t % c2 + (DIM == 3 ? g.Ng - 1 : 0);
  const long j = (t / c2) % c1 + (DIM == 2 ? g.Ng - 1 : 0);
  const long i = t / (c2 * c1) + (DIM == 1 ? g.Ng - 1 : 0);
  const long T = g.T;
  const long S = DIM == 1 ? g.sx : DIM == 2 ? g.sy : g.sz;
  const long z = (i * g.n[1] + j) * g.n[2] + k;

  double Lrow[8], Rrow[8], lam[8];
  {
    double Ph[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) Ph[q] = 0.5 * (P[q * T + z] + P[q * T + z + S]);
    if (eig::g8::rmhd_eigensystem<DIM>(Ph, lane, Lrow, Rrow, lam, pp.gamma) != 0) {
      eig::g8::identity_rows(lane, Lrow, Rrow, lam);
      if (lane == 0) atomicAdd(&st->eig_identity, 1ull);
    }
  }
  R mlr = AZ[z - 2 * S];
#pragma unroll
  for (int m = -1; m <= 3; ++m) {
    const R a = AZ[z + m * S];
    mlr = a > mlr ? a : mlr;
  }
  // the projections and the reconstructions stay in FP64 also in the mixed form: the
  // characteristic fields carry large parts that cancel on the way back through R, so
  // their rounding is multiplied by the condition number of R
  const double ml = (double)mlr;
  double fp[6], fm[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const long zm = z + (m - 2) * S;
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const double F = (double)FZ[c * T + zm], U = Uw[c * T + zm];
      a += Lrow[c] * (0.5 * (F + ml * U));
      b += Lrow[c] * (0.5 * (F - ml * U));
    }
    fp[m] = a;
    fm[m] = b;
  }
  const double f = weno5<double, true, true>(fp[0], fp[1], fp[2], fp[3], fp[4], pp.IS, pp.sz10A) +
                   weno5<double, false, true>(fm[1], fm[2], fm[3], fm[4], fm[5], pp.IS, pp.sz10A);
  double a = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) a += Rrow[c] * eig::g8::bc(f, c);
  Fout[lane * T + z] = (R)a;
}

/* --------------------------------------------------------------------------
 * plm-muscl: PlmCtuHancockOperator::ctu_hancock (Mara/ctu-hancock.cpp:71-494), 3-D, FP64.
 * The operator's fluxes F, G, H (after CT, k_update) are those of the reference: the step-2
 * Godunov fluxes outside [sx, T - sx) and the corrector fluxes inside. Failed checks are
 * counted per abort point (ctu_failed[stage][point]); the step reports the first nonzero
 * count, as the reference throws at the first abort point with failures.
 * ------------------------------------------------------------------------ */
/* the operator's own limited slope (plm_minmod with its theta, :22-31) of field q along S at
 * zone z, zero outside the flat range [sx, T - sx) where the reference computes slopes */
__device__ __forceinline__ double ctu_slope(const double *__restrict__ P, long T, long sxz, long z, long S, int q,
                                            double theta)
{
#pragma clang fp contract(off)
  if (z < sxz || z >= T - sxz) return 0.0;
  const double ul = P[q * T + z - S], u0 = P[q * T + z], ur = P[q * T + z + S];
  const double a = theta * (u0 - ul);
  const double b = 0.5 * (ur - ul);
  const double c = theta * (ur - u0);
  const double sa = (double)((a > 0) - (a < 0)), sb = (double)((b > 0) - (b < 0)), sc = (double)((c > 0) - (c < 0));
  const double ab = fabs(a) < fabs(b) ? fabs(a) : fabs(b);
  const double m = ab < fabs(c) ? ab : fabs(c);
  return 0.25 * fabs(sa + sb) * (sa + sc) * m;
}

/* the half-slope states of face z | z+S along S: from P with its own slopes (step 2) or
 * from a predictor Ps with P's slopes (step 4) */
__device__ __forceinline__ void ctu_face_states(const double *__restrict__ P, const double *__restrict__ Ps, long T,
                                                long sxz, long S, long z, double theta, double *pl, double *pr)
{
#pragma clang fp contract(off)
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    pr[q] = Ps[q * T + z + S] - 0.5 * ctu_slope(P, T, sxz, z + S, S, q, theta);
    pl[q] = Ps[q * T + z] + 0.5 * ctu_slope(P, T, sxz, z, S, q, theta);
  }
}

/* Godunov fluxes along DIM of the half-slope states over faces z | z+S, z in [begin, T - sx),
 * with the full solver (HLL, HLLC; HLLD through the redo). An unphysical state counts a
 * failure and leaves the flux unwritten. A solver error also leaves it: then the flux is
 * what the reference's array holds there, zero in step 2 (fresh arrays) and the step-2
 * flux in step 4 (`keep`). */
template <int DIM>
__global__ __launch_bounds__(kBlock, 2) void k_ctu_faces(const double *__restrict__ P, const double *__restrict__ Ps,
                                                         double *__restrict__ F, GridDesc g, PhysParams pp, int solver,
                                                         long begin, int keep, DevStats *st,
                                                         unsigned long long *fail)
{
  const long T = g.T, sxz = g.sx;
  const long S = DIM == 1 ? g.sx : DIM == 2 ? g.sy : g.sz;
  const long z = begin + (long)blockIdx.x * blockDim.x + threadIdx.x;
  double lam_out = 0.0;
  unsigned long long cnt[5] = {0, 0, 0, 0, 0};
  if (z < T - sxz) {
    double pl[8], pr[8];
    ctu_face_states(P, Ps, T, sxz, S, z, pp.theta, pl, pr);
    if (check_prim(pl) || check_prim(pr)) cnt[0] = 1;
    else {
      double f[8], lam = 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) f[q] = keep ? F[q * T + z] : 0.0;
      FluxCounters fc = {0, 0, 0, 0};
      riemann_dispatch<double, DIM>(solver, pl, pr, f, pp.gamma, pp.pfloor, lam, fc);
#pragma unroll
      for (int q = 0; q < 8; ++q) F[q * T + z] = f[q];
      lam_out = lam;
      cnt[1] = (unsigned long long)fc.hlld_ok;
      cnt[2] = (unsigned long long)fc.hlld_fail;
      cnt[3] = (unsigned long long)fc.c2p_iters;
      cnt[4] = (unsigned long long)fc.newton_iters;
    }
  }
  block_max_lambda(lam_out, &st->lam_bits);
  unsigned long long *const dst[5] = { fail, &st->hlld_success, &st->hlld_failure, &st->hlld_c2p_iters,
                                       &st->hlld_newton_iters };
  block_add<5>(cnt, dst);
}

/* The HLLD faces of k_ctu_faces on the common path (riemann_hlld in FAST mode, a lower
 * register peak): a face that needs the c2p cascade, a later Newton attempt, the small
 * normal field branch or the HLLC fallback is queued for k_ctu_faces_redo. The queue holds
 * every face, so it cannot overflow. */
template <int DIM>
__global__ __launch_bounds__(kBlock, kFastFluxWaves64)
void k_ctu_faces_fast(const double *__restrict__ P, const double *__restrict__ Ps, double *__restrict__ F,
                      GridDesc g, PhysParams pp, long begin, DevStats *st, unsigned long long *fail, RedoQueue rq)
{
  const long T = g.T, sxz = g.sx;
  const long S = DIM == 1 ? g.sx : DIM == 2 ? g.sy : g.sz;
  const long z = begin + (long)blockIdx.x * blockDim.x + threadIdx.x;
  double lam_out = 0.0;
  unsigned long long cnt[4] = {0, 0, 0, 0};
  bool defer = false;
  if (z < T - sxz) {
    double pl[8], pr[8];
    ctu_face_states(P, Ps, T, sxz, S, z, pp.theta, pl, pr);
    if (check_prim(pl) || check_prim(pr)) cnt[0] = 1;
    else {
      double f[8], lam = 0.0;
      FluxCounters fc = {0, 0, 0, 0};
      defer = riemann_hlld<double, DIM, true>(pl, pr, f, pp.gamma, pp.pfloor, lam, fc) != 0;
      if (!defer) {
#pragma unroll
        for (int q = 0; q < 8; ++q) F[q * T + z] = f[q];
        cnt[1] = 1;
        cnt[2] = (unsigned long long)fc.c2p_iters;
        cnt[3] = (unsigned long long)fc.newton_iters;
      }
      lam_out = lam;
    }
  }
  queue_push(rq, defer, z);
  sweep_counters_wave(lam_out, cnt[1] != 0, (unsigned)cnt[2], (unsigned)cnt[3], cnt[0] != 0, rq.part);
}

/* The queued faces of k_ctu_faces_fast by the full HLLD method (attempt ladder, HLLC
 * fallback); a grid-stride loop over the queue count. Their states passed the physical
 * check in the fast pass. */
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_ctu_faces_redo(const double *__restrict__ P, const double *__restrict__ Ps,
                                                           double *__restrict__ F, GridDesc g, PhysParams pp, int keep,
                                                           DevStats *st, unsigned long long *fail, RedoQueue rq,
                                                           unsigned int nparts)
{
  const unsigned int n = min(*rq.count, rq.cap);
  const long T = g.T, sxz = g.sx;
  const long S = DIM == 1 ? g.sx : DIM == 2 ? g.sy : g.sz;
  double lam_out = 0.0;
  unsigned long long cnt[6] = {0, 0, 0, 0, 0, 0};
  // the fast pass's per-wave counters
  for (unsigned int w = blockIdx.x * blockDim.x + threadIdx.x; w < nparts; w += gridDim.x * blockDim.x) {
    const uint4 v = rq.part[w];
    lam_out = fmax(lam_out, __longlong_as_double((long long)(((unsigned long long)v.y << 32) | v.x)));
    cnt[0] += (unsigned long long)(v.z & 255u);
    cnt[2] += (unsigned long long)(v.z >> 8);
    cnt[3] += (unsigned long long)(v.w & 0xffffu);
    cnt[5] += (unsigned long long)(v.w >> 16);
  }
  for (unsigned int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const long z = (long)rq.z[k];
    double pl[8], pr[8], f[8], lam = 0.0;
    ctu_face_states(P, Ps, T, sxz, S, z, pp.theta, pl, pr);
#pragma unroll
    for (int q = 0; q < 8; ++q) f[q] = keep ? F[q * T + z] : 0.0;
    FluxCounters fc = {0, 0, 0, 0};
    riemann_hlld<double, DIM>(pl, pr, f, pp.gamma, pp.pfloor, lam, fc);
#pragma unroll
    for (int q = 0; q < 8; ++q) F[q * T + z] = f[q];
    lam_out = fmax(lam_out, lam);
    cnt[0] += (unsigned long long)fc.hlld_ok;
    cnt[1] += (unsigned long long)fc.hlld_fail;
    cnt[2] += (unsigned long long)fc.c2p_iters;
    cnt[3] += (unsigned long long)fc.newton_iters;
    cnt[4] += 1;
  }
  block_max_lambda(lam_out, &st->lam_bits);
  unsigned long long *const dst[6] = { &st->hlld_success, &st->hlld_failure, &st->hlld_c2p_iters,
                                       &st->hlld_newton_iters, &st->redo, fail };
  block_add<6>(cnt, dst);
}

/* the physical flux along DIM of the state P (U from it) */
template <int DIM>
__device__ __forceinline__ void ctu_side_flux(const double *Ps, double *F, double gamma)
{
  double u[8], a, b;
  prim_to_cons(Ps, u, gamma - 1.0);
  flux_and_eigenvalues<double, DIM, false>(u, Ps, F, a, b, gamma);
}

/* Step 3: per zone z in [sx, T), the states on the inner faces P -+ slope/2 of each axis, their
 * fluxes, and the Hancock predictor of each axis with the transverse step-2 fluxes */
__global__ __launch_bounds__(kBlock, 2) void k_ctu_predict(const double *__restrict__ P, const double *__restrict__ U,
                                                           const double *__restrict__ F, const double *__restrict__ G,
                                                           const double *__restrict__ H, double *__restrict__ Ux,
                                                           double *__restrict__ Uy, double *__restrict__ Uz, GridDesc g,
                                                           PhysParams pp, double dt, unsigned long long *fail)
{
#pragma clang fp contract(off)
  const long T = g.T, sxz = g.sx, sy = g.sy, sz = g.sz;
  const long z = sxz + (long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cnt[1] = {0};
  if (z < T) {
    double FL[8], FR[8], GL[8], GR[8], HL[8], HR[8];
    const long Sd[3] = {g.sx, g.sy, g.sz};
    double *Fo[3][2] = {{FL, FR}, {GL, GR}, {HL, HR}};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double PL[8], PR[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const double s = ctu_slope(P, T, sxz, z, Sd[d], q, pp.theta);
        PL[q] = P[q * T + z] - 0.5 * s;
        PR[q] = P[q * T + z] + 0.5 * s;
      }
      if (check_prim(PL) || check_prim(PR)) { cnt[0] += 1; continue; }
      if (d == 0) { ctu_side_flux<1>(PL, Fo[0][0], pp.gamma); ctu_side_flux<1>(PR, Fo[0][1], pp.gamma); }
      if (d == 1) { ctu_side_flux<2>(PL, Fo[1][0], pp.gamma); ctu_side_flux<2>(PR, Fo[1][1], pp.gamma); }
      if (d == 2) { ctu_side_flux<3>(PL, Fo[2][0], pp.gamma); ctu_side_flux<3>(PR, Fo[2][1], pp.gamma); }
    }
    if (cnt[0] == 0) {
      const double dx = g.dx, dy = g.dy, dz = g.dz;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const long m = q * T + z;
        const double dF = (F[m] - F[m - sxz]) / dx, dG = (G[m] - G[m - sy]) / dy, dH = (H[m] - H[m - sz]) / dz;
        Ux[m] = U[m] - ((FR[q] - FL[q]) / dx + dG + dH) * 0.5 * dt;
        Uy[m] = U[m] - ((GR[q] - GL[q]) / dy + dH + dF) * 0.5 * dt;
        Uz[m] = U[m] - ((HR[q] - HL[q]) / dz + dF + dG) * 0.5 * dt;
      }
    }
  }
  unsigned long long *const dst[1] = { fail };
  block_add<1>(cnt, dst);
}

/* the interior zone whose periodic image z is (z itself inside): each coordinate wrapped
 * once, as PeriodicBoundary::ApplyBoundaries fills the ghosts (needs N >= Ng per axis) */
__device__ __forceinline__ long image_zone(const GridDesc &g, long z)
{
  int k = (int)(z % g.n[2]), j = (int)((z / g.n[2]) % g.n[1]), i = (int)(z / ((long)g.n[1] * g.n[2]));
  i = i < g.Ng ? i + g.N[0] : (i >= g.Ng + g.N[0] ? i - g.N[0] : i);
  j = j < g.Ng ? j + g.N[1] : (j >= g.Ng + g.N[1] ? j - g.N[1] : j);
  k = k < g.Ng ? k + g.N[2] : (k >= g.Ng + g.N[2] ? k - g.N[2] : k);
  return ((long)i * g.n[1] + j) * g.n[2] + k;
}

/* the predictors, synchronised (their ghost zones are their periodic images) and inverted
 * (Rmhd::ConsToPrim) from P at every zone, ghosts included; a failure keeps P there and
 * counts once per axis. A ghost zone reads its image's predictor instead of a copy of it. */
__global__ __launch_bounds__(kBlock, 3) void k_ctu_c2p(const double *__restrict__ P, const double *__restrict__ Ux,
                                                       const double *__restrict__ Uy, const double *__restrict__ Uz,
                                                       double *__restrict__ Px, double *__restrict__ Py,
                                                       double *__restrict__ Pz, GridDesc g, double gamma, double pfloor,
                                                       DevStats *st, unsigned long long *fail)
{
  const long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long T = g.T;
  unsigned long long cnt[3] = {0, 0, 0};
  if (z < T) {
    const double *Ud[3] = {Ux, Uy, Uz};
    double *Pd[3] = {Px, Py, Pz};
    const long zi = image_zone(g, z);
#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
      double u[8], p[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) u[q] = Ud[d][q * T + zi];
      int iters = 0, which = 0;
      if (cons_to_prim(u, p, gamma, pfloor, true, iters, which, [&](double *pp) {
#pragma unroll
            for (int q = 0; q < 8; ++q) pp[q] = P[q * T + z];
          }) != 0)
        cnt[0] += 1;
#pragma unroll
      for (int q = 0; q < 8; ++q) Pd[d][q * T + z] = p[q];
      cnt[1] += 1;
      cnt[2] += (unsigned long long)iters;
    }
  }
  unsigned long long *const dst[3] = { fail, &st->c2p_zones, &st->c2p_iters };
  block_add<3>(cnt, dst);
}

/* --------------------------------------------------------------------------
 * CT + divergence + RK stage (FP64, no contraction: the same roundings as the reference)
 * ------------------------------------------------------------------------ */
enum StageKind {
  ST_EULER = 0,      // U <- U0 + dt L            (single; rk2/rk3 stage 1)
  ST_RK2_FINAL,      // U <- 0.5 (U0 + U + dt L)
  ST_RK3_2,          // U <- 3/4 U0 + 1/4 U + 1/4 dt L
  ST_RK3_3,          // U <- 1/3 U0 + 2/3 U + 2/3 dt L
  ST_RK4_1,          // A <- dt L;          U <- U0 + 0.5 A
  ST_RK4_2,          // A <- A + 2 dt L;    U <- U0 + 0.5 dt L
  ST_RK4_3,          // A <- A + 2 dt L;    U <- U0 + 1.0 dt L
  ST_RK4_4           // U <- U0 + 1/6 (A + dt L)
};

/* whether a stage of this kind reads U, the previous stage's result (the others read only
 * U0 and A): U is written only by a stage whose successor reads it */
constexpr bool stage_reads_u(int kind)
{
  return kind == ST_RK2_FINAL || kind == ST_RK3_2 || kind == ST_RK3_3;
}

template <typename FT>
__device__ __forceinline__ double ld(const FT *__restrict__ a, long i) { return (double)a[i]; }

/* One interior zone z of the stage update: CT of the three flux arrays, the flux divergence
 * and the RK combination; A written, U written when write_u, the new U returned in unew */
template <typename FT>
__device__ __forceinline__ void update_zone(const FT *__restrict__ Fx, const FT *__restrict__ Fy,
                                            const FT *__restrict__ Fz, const double *__restrict__ U0,
                                            double *__restrict__ U, double *__restrict__ A, const GridDesc &g,
                                            long z, double dt, int kind, bool write_u, double *unew)
{
#pragma clang fp contract(off)
  const long T = g.T, sx = g.sx, sy = g.sy, sz = g.sz;
  const FT *FxBy = Fx + (long)By * T, *FxBz = Fx + (long)Bz * T;
  const FT *FyBx = Fy + (long)Bx * T, *FyBz = Fy + (long)Bz * T;
  const FT *FzBx = Fz + (long)Bx * T, *FzBy = Fz + (long)By * T;

  // Rmhd::ConstrainedTransport3d, Mara/rmhd.cpp:310-329, at zone i
  auto ct_xBy = [&](long i) {
    return (2*ld(FxBy,i)+ld(FxBy,i+sy)+ld(FxBy,i-sy)-ld(FyBx,i)-ld(FyBx,i+sx)-ld(FyBx,i-sy)-ld(FyBx,i+sx-sy))*0.125;
  };
  auto ct_yBx = [&](long i) {
    return (2*ld(FyBx,i)+ld(FyBx,i+sx)+ld(FyBx,i-sx)-ld(FxBy,i)-ld(FxBy,i+sy)-ld(FxBy,i-sx)-ld(FxBy,i-sx+sy))*0.125;
  };
  auto ct_yBz = [&](long i) {
    return (2*ld(FyBz,i)+ld(FyBz,i+sz)+ld(FyBz,i-sz)-ld(FzBy,i)-ld(FzBy,i+sy)-ld(FzBy,i-sz)-ld(FzBy,i+sy-sz))*0.125;
  };
  auto ct_zBy = [&](long i) {
    return (2*ld(FzBy,i)+ld(FzBy,i+sy)+ld(FzBy,i-sy)-ld(FyBz,i)-ld(FyBz,i+sz)-ld(FyBz,i-sy)-ld(FyBz,i-sy+sz))*0.125;
  };
  auto ct_zBx = [&](long i) {
    return (2*ld(FzBx,i)+ld(FzBx,i+sx)+ld(FzBx,i-sx)-ld(FxBz,i)-ld(FxBz,i+sz)-ld(FxBz,i-sx)-ld(FxBz,i+sz-sx))*0.125;
  };
  auto ct_xBz = [&](long i) {
    return (2*ld(FxBz,i)+ld(FxBz,i+sz)+ld(FxBz,i-sz)-ld(FzBx,i)-ld(FzBx,i+sx)-ld(FzBx,i-sz)-ld(FzBx,i-sz+sx))*0.125;
  };

  double F0[8], F1[8], G0[8], G1[8], H0[8], H1[8];  // fluxes at z and at z - s
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    F0[q] = ld(Fx + q * T, z); F1[q] = ld(Fx + q * T, z - sx);
    G0[q] = ld(Fy + q * T, z); G1[q] = ld(Fy + q * T, z - sy);
    H0[q] = ld(Fz + q * T, z); H1[q] = ld(Fz + q * T, z - sz);
  }
  F0[Bx] = 0.0;           F1[Bx] = 0.0;
  F0[By] = ct_xBy(z);     F1[By] = ct_xBy(z - sx);
  F0[Bz] = ct_xBz(z);     F1[Bz] = ct_xBz(z - sx);
  G0[Bx] = ct_yBx(z);     G1[Bx] = ct_yBx(z - sy);
  G0[By] = 0.0;           G1[By] = 0.0;
  G0[Bz] = ct_yBz(z);     G1[Bz] = ct_yBz(z - sy);
  H0[Bx] = ct_zBx(z);     H1[Bx] = ct_zBx(z - sz);
  H0[By] = ct_zBy(z);     H1[By] = ct_zBy(z - sz);
  H0[Bz] = 0.0;           H1[Bz] = 0.0;

#pragma unroll
  for (int q = 0; q < 8; ++q) {
    // Mara/plm-split.cpp:117-119
    const double L = -(F0[q]-F1[q])/g.dx - (G0[q]-G1[q])/g.dy - (H0[q]-H1[q])/g.dz;
    const long m = q * T + z;
    const double u0 = U0[m];
    double un;
    switch (kind) {
    case ST_EULER: un = u0 + dt*L; break;
    case ST_RK2_FINAL: un = 0.5*(u0 + U[m] + dt*L); break;
    case ST_RK3_2: un = 3./4*u0 + 1./4*U[m] + 1./4*dt*L; break;
    case ST_RK3_3: un = 1./3*u0 + 2./3*U[m] + 2./3*dt*L; break;
    case ST_RK4_1: { const double l = dt*L; A[m] = l; un = u0 + 0.5*l; } break;
    case ST_RK4_2: { const double l = dt*L; A[m] = A[m] + 2.0*l; un = u0 + 0.5*l; } break;
    case ST_RK4_3: { const double l = dt*L; A[m] = A[m] + 2.0*l; un = u0 + 1.0*l; } break;
    default: { const double l = dt*L; un = u0 + (1.0/6.0)*(A[m] + l); } break;
    }
    if (write_u) U[m] = un;
    unew[q] = un;
  }
}

/* The zone's result of a conserved-to-primitive pass: the emergency reset of
 * Mara/hydro.cpp:165-196 where the cascade failed and ereset is on, the primitive state
 * (and with Uw the conserved state) written to the zone and its periodic images, and the
 * pass's counters: zones, Newton iterations, fallbacks, failed zones (with the zone's
 * multiplicity, as Mara's all-zone loops count them) and, in the learned form, network
 * starts and their fallbacks. */
template <int NC>
__device__ __forceinline__ void finish_c2p(const GridDesc &g, double *__restrict__ P, double *__restrict__ Uw,
                                           int i, int j, int k, const double *u, double *p, int err, int ereset,
                                           double preset, double dreset, DevStats *st, int pass, int iters,
                                           bool fallback, unsigned long long (&cnt)[NC])
{
  int failed = err != 0;
  if (failed && ereset) {
    p[0] = dreset; p[1] = preset; p[2] = 0.0; p[3] = 0.0; p[4] = 0.0;
    p[5] = u[5]; p[6] = u[6]; p[7] = u[7];
    failed = 0;
  }
  const int mult = scatter_images(g, P, g.T, i, j, k, p);
  if (Uw) {
    scatter_images(g, Uw, g.T, i, j, k, u);
    if (err != 0 && ereset) atomicAdd(&st->stage_reset[pass], 1ull);
  }
  cnt[0] += 1;
  cnt[1] += (unsigned long long)iters;
  cnt[2] += fallback ? 1ull : 0ull;
  cnt[3] += failed ? (unsigned long long)mult : 0ull;
  if constexpr (NC == 6) {
    cnt[4] += 1;
    cnt[5] += fallback ? 1ull : 0ull;
  }
}

template <int NC>
__device__ __forceinline__ void c2p_counters_out(const unsigned long long (&cnt)[NC], DevStats *st, int pass)
{
  if constexpr (NC == 6) {
    unsigned long long *const dst[6] = { &st->c2p_zones, &st->c2p_iters, &st->c2p_fallbacks,
                                         &st->stage_failed[pass], &st->mlp_guesses, &st->mlp_fallbacks };
    block_add<6>(cnt, dst);
  }
  else {
    unsigned long long *const dst[4] = { &st->c2p_zones, &st->c2p_iters, &st->c2p_fallbacks,
                                         &st->stage_failed[pass] };
    block_add<4>(cnt, dst);
  }
}

/* The pass's counters, per wave (reduced with DPP moves), stored by the wave's first lane in
 * part[wave]: x zones, y Newton iterations, z fallbacks + 65536 x failed zones (with their
 * multiplicity), w network starts + 65536 x their fallbacks; k_c2p_parts sums the slots.
 * The kernel then ends without a block barrier: a wave retires when its own zones are done
 * instead of waiting for the slowest wave of its block. */
template <int NC>
__device__ __forceinline__ void c2p_counters_wave(const unsigned long long (&cnt)[NC], uint4 *part)
{
  const unsigned nz = wave_sum_dpp((unsigned)cnt[0]), ni = wave_sum_dpp((unsigned)cnt[1]);
  const unsigned nf = wave_sum_dpp((unsigned)cnt[2]), nx = wave_sum_dpp((unsigned)cnt[3]);
  unsigned w = 0;
  if constexpr (NC == 6) w = wave_sum_dpp((unsigned)cnt[4]) | (wave_sum_dpp((unsigned)cnt[5]) << 16);
  if ((threadIdx.x & 63) == 0)
    part[blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)] = make_uint4(nz, ni, nf | (nx << 16), w);
}

template <int NC>
__global__ __launch_bounds__(kBlock) void k_c2p_parts(const uint4 *__restrict__ part, unsigned int nw, DevStats *st,
                                                      int pass)
{
  unsigned long long cnt[NC] = {};
  for (unsigned int w = blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += gridDim.x * blockDim.x) {
    const uint4 v = part[w];
    cnt[0] += v.x;
    cnt[1] += v.y;
    cnt[2] += v.z & 0xffffu;
    cnt[3] += v.z >> 16;
    if constexpr (NC == 6) {
      cnt[4] += v.w & 0xffffu;
      cnt[5] += v.w >> 16;
    }
  }
  c2p_counters_out<NC>(cnt, st, pass);
}
constexpr unsigned kPartBlocks = 128;   // grid of k_c2p_parts

/* Grid of the learned kernels: they loop over 256-zone chunks, so the network is staged
 * into LDS once per block. */
constexpr unsigned kLearnedBlocks = 4096;

/* The conserved-to-primitive pass's start: the plain cascade (C2P_CASCADE), the network on
 * the zone's U (C2P_NET, the step's first pass in the learned form), or the zone's stored
 * start, the (Z, W) of its previous solution (C2P_STORED, the stage passes of the non-plain
 * forms). The non-plain forms (mixed precision, learned c2p) store each solution's (Z, W)
 * as the zone's start for its next pass within the step. The step's first pass does not
 * start from the previous step's last solution: reading it costs that pass more time than
 * the Newton iterations it saves (0.92 against 0.88 ms at 256^3). */
enum { C2P_CASCADE = 0, C2P_NET = 1, C2P_STORED = 2 };

/* The stage update followed at once by the next conserved-to-primitive pass of the zone
 * (not the first pass): CT, the flux divergence and the RK combination, then the new U
 * from registers into the inversion (the plain cascade, or from the zone's stored start
 * with the cascade as fallback). With Uw (weno-split, plm-muscl) U's ghost images are
 * written too. */
template <typename FT, int MODE>
__global__ __launch_bounds__(kBlock, 3) void k_update_c2p(
    const FT *__restrict__ Fx, const FT *__restrict__ Fy, const FT *__restrict__ Fz,
    const double *__restrict__ U0, double *__restrict__ U, double *__restrict__ A, double *__restrict__ P,
    GridDesc g, double dt, int kind, int write_u, double gamma, double pfloor, int ereset, double preset,
    double dreset, DevStats *st, int pass, double *__restrict__ Uw, float2 *__restrict__ start, uint4 *part)
{
  static_assert(MODE == C2P_CASCADE || MODE == C2P_STORED, "stage passes start from the cascade or a stored start");
  constexpr int NC = 4;
  const long T = g.T;
  unsigned long long cnt[NC] = {};
  {
    long t = 0;
    const bool valid = update_zone_index(g, t);
    int i = 0, j = 0, k = 0;
    long z = 0;
    double u[8], p[8];
    if (valid) {
      k = (int)(t % g.N[2]) + g.Ng;
      j = (int)((t / g.N[2]) % g.N[1]) + g.Ng;
      i = (int)(t / ((long)g.N[2] * g.N[1])) + g.Ng;
      z = ((long)i * g.n[1] + j) * g.n[2] + k;
      update_zone<FT>(Fx, Fy, Fz, U0, U, A, g, z, dt, kind, write_u != 0, u);
    }
    int iters = 0, err = 0;
    bool fallback = false;
    if (valid) {
      auto prev = [&](double *q) {
#pragma unroll
        for (int f = 0; f < 8; ++f) q[f] = P[f * T + z];
      };
      if constexpr (MODE == C2P_STORED) {
        const float2 zw = start[t];
        int used;
        err = learned_c2p(u, p, (double)zw.x, (double)zw.y, gamma, pfloor, true, prev, iters, used, start + t);
        fallback = !used;
      }
      else {
        int which = 0;
        err = cons_to_prim(u, p, gamma, pfloor, true, iters, which, prev);
        fallback = which != 0;
      }
      finish_c2p<NC>(g, P, Uw, i, j, k, u, p, err, ereset, preset, dreset, st, pass, iters, fallback, cnt);
    }
  }
  c2p_counters_wave<NC>(cnt, part);
}

/* The start of a step: GodunovOperator::PrimToCons of the committed state and the first
 * conserved-to-primitive pass of the zone, whose previous state is the zero-filled
 * PrimitiveArray (Mara/mara.cpp:321), in one kernel, from the cascade or (C2P_NET) from the
 * network's start. The conserved state goes to U0 only: the first stage of every scheme
 * reads U0 and writes U. With start, the solution's (Z, W) are stored as the zone's start
 * for the first stage pass. */
// occupancy: 4 waves per SIMD for the cascade (8 % faster than 5, where it keeps 36 B per
// lane in scratch), 3 with the network (2: the same; 4: 1.9x slower)
template <int MODE>
__global__ __launch_bounds__(kBlock, MODE == C2P_NET ? 3 : 4) void k_p2c_c2p(const double *__restrict__ Pc,
                                                                    double *__restrict__ U0,
                                                                    double *__restrict__ P, GridDesc g, double gamma,
                                                                    double pfloor, int ereset, double preset,
                                                                    double dreset, DevStats *st, int pass,
                                                                    double *__restrict__ Uw, MlpDev net,
                                                                    float2 *__restrict__ start, uint4 *part)
{
  static_assert(MODE == C2P_CASCADE || MODE == C2P_NET, "the first pass starts from the cascade or the network");
  constexpr bool LEARNED = MODE == C2P_NET;
  constexpr int NC = LEARNED ? 6 : 4;
  MlpLds *L = nullptr;
  if constexpr (LEARNED) {
    __shared__ __align__(16) MlpLds lds;
    L = &lds;
    stage_net(lds, net);
    __syncthreads();
  }
  const long ni = (long)g.N[0] * g.N[1] * g.N[2];
  const long T = g.T;
  unsigned long long cnt[NC] = {};
  const long step = LEARNED ? (long)gridDim.x * kBlock : ni;
  for (long base = (long)blockIdx.x * kBlock; base < ni; base += step) {
    const long t = base + threadIdx.x;
    const bool valid = t < ni;
    int i = 0, j = 0, k = 0;
    double u[8], p[8];
    if (valid) {
      k = (int)(t % g.N[2]) + g.Ng;
      j = (int)((t / g.N[2]) % g.N[1]) + g.Ng;
      i = (int)(t / ((long)g.N[2] * g.N[1])) + g.Ng;
      const long z = ((long)i * g.n[1] + j) * g.n[2] + k;
#pragma unroll
      for (int q = 0; q < 8; ++q) p[q] = Pc[q * T + z];
      prim_to_cons(p, u, gamma - 1.0);
#pragma unroll
      for (int q = 0; q < 8; ++q) U0[q * T + z] = u[q];
    }
    else {
#pragma unroll
      for (int q = 0; q < 8; ++q) u[q] = q < 2 ? 1.0 : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) p[q] = 0.0;
    int iters = 0, err;
    bool fallback;
    if constexpr (LEARNED) {
      double Zg, Wg;
      learned_guess(*L, net, u, gamma, Zg, Wg);
      if (!valid) continue;
      int used;
      err = learned_c2p(u, p, Zg, Wg, gamma, pfloor, false, [](double *) {}, iters, used, start ? start + t : nullptr);
      fallback = !used;
    }
    else {
      if (!valid) continue;
      int which = 0;
      err = cons_to_prim(u, p, gamma, pfloor, false, iters, which);
      fallback = which != 0;
      if (start && !err) {
        double Z, W;
        c2p_zw_of_prim(p, gamma, Z, W);
        start[t] = make_float2((float)Z, (float)W);
      }
    }
    finish_c2p<NC>(g, P, Uw, i, j, k, u, p, err, ereset, preset, dreset, st, pass, iters, fallback, cnt);
  }
  c2p_counters_wave<NC>(cnt, part);
}

/* luaC_mean_max_divB stencil, Mara/mara.cpp:2298-2314, over its loop range */
/* The Mara.measure table (Mara/mara.cpp:1954-2262) over the interior zones of the committed
 * state: per zone the quantities of luaC_mean_velocity, _mean_cons, _mean_prim,
 * _mean_energies, _mean_max_temperature, _mean_max_sonic_mach, _mean_min_alfvenic_mach,
 * _mean_max_magnetic_field and _max_lorentz_factor for the RMHD fluid and the gamma-law
 * EOS, in the reference's operation order (no contraction); kMeasSums sums and kMeasExt
 * extremes per thread over a grid-stride loop, reduced per block in a fixed order into
 * one row of partials. k_measure_reduce sums the rows, also in a fixed order, so the
 * result does not change from run to run. */
constexpr int kMeasSums = 28, kMeasExt = 5, kMeasQ = kMeasSums + kMeasExt, kMeasBlocks = 1024;
constexpr int kCoolBlocks = 1024;      // grid of k_cool and k_power (grid-stride), rows of their partial sums
constexpr int kOuMaxK1 = 6;            // (2 k1 + 1)^3 modes x 48 B of amplitudes in LDS

struct MeasureUnits { double ErgPerCC, GramsPerCC, mp, kb, K2MeV; };

/* PhysicalUnits(Length, Mass, Time) (Mara/units.cpp:3-6, :92-141) and the constants of
 * Mara/eos.hpp:12-15, as TemperatureMeV / TemperatureArb use them */
MeasureUnits measure_units(const double *units)
{
  const double Length = units[0], Mass = units[1], Time = units[2];
  const double Gram = 1.0 / Mass, Centimeter = 1.0 / Length, Second = 1.0 / Time;
  const double Erg = Gram * std::pow(Centimeter / Second, 2.0);
  const double MeV = 1.60217648740e-06 * Erg;
  MeasureUnits mu;
  mu.GramsPerCC = Gram / std::pow(Centimeter, 3.0);
  mu.ErgPerCC = MeV / std::pow(1e-13 * Centimeter, 3) * 1e-33;
  mu.mp = 1.67262177774e-24 * Gram;
  mu.kb = 1.38064881300e-16 * Erg;
  mu.K2MeV = 8.621738e-11;
  return mu;
}



__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

/* Rmhd::PrimToCons (Mara/rmhd.cpp:133-163) with the reference's divisions and operation
 * order (no contraction), for the diagnostics and source steps that must round as the
 * reference does */
__device__ __forceinline__ void prim_to_cons_ref(const double *P0, double *U0, double gamma)
{
#pragma clang fp contract(off)
  const double V2 = P0[vx]*P0[vx] + P0[vy]*P0[vy] + P0[vz]*P0[vz];
  const double B2 = P0[Bx]*P0[Bx] + P0[By]*P0[By] + P0[Bz]*P0[Bz];
  const double Bv = P0[Bx]*P0[vx] + P0[By]*P0[vy] + P0[Bz]*P0[vz];
  const double W2 = 1.0 / (1.0 - V2);
  const double W = sqrt(W2);
  const double b0 = W*Bv;
  const double b2 = (B2 + b0*b0) / W2;
  const double bx = (P0[Bx] + b0 * W*P0[vx]) / W;
  const double by = (P0[By] + b0 * W*P0[vy]) / W;
  const double bz = (P0[Bz] + b0 * W*P0[vz]) / W;
  const double e = (P0[pre] / (gamma - 1)) / P0[rho];
  const double e_ = e + 0.5 * b2 / P0[rho];
  const double p_ = P0[pre] + 0.5 * b2;
  const double h_ = 1 + e_ + p_ / P0[rho];
  U0[ddd] = P0[rho] * W;
  U0[tau] = P0[rho] * h_ * W2 - p_ - b0*b0 - U0[ddd];
  U0[Sx] = P0[rho] * h_ * W2 * P0[vx] - b0*bx;
  U0[Sy] = P0[rho] * h_ * W2 * P0[vy] - b0*by;
  U0[Sz] = P0[rho] * h_ * W2 * P0[vz] - b0*bz;
  U0[Bx] = P0[Bx];
  U0[By] = P0[By];
  U0[Bz] = P0[Bz];
}

__global__ __launch_bounds__(kBlock) void k_measure(const double *__restrict__ P, GridDesc g, double gamma,
                                                    MeasureUnits mu, double *__restrict__ part)
{
#pragma clang fp contract(off)
  const long T = g.T;
  const long ni = (long)g.N[0] * g.N[1] * g.N[2];
  double sum[kMeasSums];
#pragma unroll
  for (int q = 0; q < kMeasSums; ++q) sum[q] = 0.0;
  double T_max = 0.0, Ms_max = 0.0, Ma_min = 1e20, B_max = 0.0, W_max = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ni; t += (long)gridDim.x * blockDim.x) {
    const long z = interior_zone(g, t);
    double P0[8], U0[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) P0[q] = P[q * T + z];
    prim_to_cons_ref(P0, U0, gamma);
    const double v2 = P0[vx]*P0[vx] + P0[vy]*P0[vy] + P0[vz]*P0[vz];
    const double B2 = P0[Bx]*P0[Bx] + P0[By]*P0[By] + P0[Bz]*P0[Bz];
    const double W0 = 1.0 / sqrt(1.0 - v2);
    sum[0] += W0;                                   // mean_velocity
    sum[1] += fabs(P0[vx]);
    sum[2] += fabs(P0[vy]);
    sum[3] += fabs(P0[vz]);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      sum[4 + q] += U0[q];                          // mean_cons
      sum[12 + q] += P0[q];                         // mean_prim
    }
    const double T0 = P0[pre];                      // Temperature_p; Internal = T/(Gamma-1)
    const double u0 = T0 / (gamma - 1);
    const double rhoh = P0[rho] + u0 + P0[pre];
    const double e_f = rhoh * W0*W0 - P0[pre] - U0[ddd];
    sum[20] += U0[tau] - e_f;                       // mean_energies: magnetic
    sum[21] += P0[rho] * W0 * (W0 - 1);             // kinetic
    sum[22] += W0 * u0;                             // internal
    sum[23] += U0[tau];                             // total
    {
      // AdiabaticEos::TemperatureMeV, Mara/eos.cpp:37-47
      const double p = P0[pre] / mu.ErgPerCC, r = P0[rho] / mu.GramsPerCC;
      const double TM = ((p / r) * (mu.mp / mu.kb)) * mu.K2MeV;
      sum[24] += TM;
      if (TM > T_max) T_max = TM;
    }
    {
      const double rh = P0[rho] + T0 / (gamma - 1) + T0;       // SoundSpeed2Sr
      const double cs2 = gamma * T0 / rh;
      const double M0 = sqrt((v2 / (1 - v2)) / (cs2 / (1 - cs2)));
      sum[25] += M0;
      if (M0 > Ms_max) Ms_max = M0;
    }
    {
      const double va2 = B2 / (B2 + rhoh);
      const double M0 = sqrt((v2 / (1 - v2)) / (va2 / (1 - va2)));
      sum[26] += M0;
      if (M0 < Ma_min) Ma_min = M0;
    }
    {
      const double B0 = sqrt(B2);
      sum[27] += B0;
      if (B0 > B_max) B_max = B0;
    }
    if (W0 > W_max) W_max = W0;
  }
  // fixed-order block reduction into this block's row of partials
  __shared__ double red[kBlock / 64][kMeasQ];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kMeasSums; ++q) {
    const double v = wave_sum_d(sum[q]);
    if (lane == 0) red[wave][q] = v;
  }
  double ext[kMeasExt] = {T_max, Ms_max, Ma_min, B_max, W_max};
#pragma unroll
  for (int q = 0; q < kMeasExt; ++q) {
    double v = ext[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(v, off, 64);
      v = q == 2 ? (o < v ? o : v) : (o > v ? o : v);
    }
    if (lane == 0) red[wave][kMeasSums + q] = v;
  }
  __syncthreads();
  if (threadIdx.x < kMeasQ) {
    const int q = threadIdx.x;
    double v = red[0][q];
    for (int w = 1; w < kBlock / 64; ++w) {
      const double o = red[w][q];
      if (q < kMeasSums) v += o;
      else if (q == kMeasSums + 2) v = o < v ? o : v;
      else v = o > v ? o : v;
    }
    part[(long)blockIdx.x * kMeasQ + q] = v;
  }
}

/* One block per quantity: the nb rows of partials in a fixed order */
__global__ __launch_bounds__(kBlock) void k_measure_reduce(const double *__restrict__ part, int nb,
                                                           double *__restrict__ out)
{
  const int q = blockIdx.x;
  const bool is_min = q == kMeasSums + 2, is_sum = q < kMeasSums;
  double v = is_sum ? 0.0 : (is_min ? 1e20 : 0.0);
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const double o = part[(long)b * kMeasQ + q];
    v = is_sum ? v + o : is_min ? (o < v ? o : v) : (o > v ? o : v);
  }
  __shared__ double red[kBlock];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double a = red[threadIdx.x], o = red[threadIdx.x + s];
      red[threadIdx.x] = is_sum ? a + o : is_min ? (o < a ? o : a) : (o > a ? o : a);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = red[0];
}

/* ---------------------------------------------------------------------------
 * Driving and cooling, applied to the committed state after a successful step as
 * luaC_Mara_advance does (Mara/mara.cpp:330-333)
 * ------------------------------------------------------------------------- */
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

/* DrivingProcedure::ResampleField (Mara/driving.cpp:44-85): the field at every zone centre,
 * ghosts included, F(x) = Re sum_m F_m exp(i k_m . x), summed from the last mode to the
 * first as the reference's valarray expression is. exp(i k . x) is the product of per-axis
 * factors exp(2 pi i n x) from tables (ex, ey, ez: 2 k1 + 1 per zone index), which differs
 * from the exponential of the summed phase by round-off. The amplitudes (3 numk complex
 * numbers, axis-major) are staged in LDS. */
__global__ __launch_bounds__(kBlock) void k_ou_sample(const double2 *__restrict__ ex, const double2 *__restrict__ ey,
                                                      const double2 *__restrict__ ez, const double2 *__restrict__ amp,
                                                      int k1, GridDesc g, double *__restrict__ Fd)
{
  extern __shared__ double2 sh_amp[];
  const int nk = 2 * k1 + 1, numk = nk * nk * nk;
  for (int t = threadIdx.x; t < 3 * numk; t += blockDim.x) sh_amp[t] = amp[t];
  __syncthreads();
  const long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= g.T) return;
  const long k = z % g.n[2], j = (z / g.n[2]) % g.n[1], i = z / ((long)g.n[1] * g.n[2]);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int a = nk - 1; a >= 0; --a) {
    const double2 X = ex[i * nk + a];
    for (int b = nk - 1; b >= 0; --b) {
      const double2 xy = cmul(X, ey[j * nk + b]);
      for (int c = nk - 1; c >= 0; --c) {
        const double2 e = cmul(xy, ez[k * nk + c]);
        const int m = (a * nk + b) * nk + c;
        const double2 fx = sh_amp[m], fy = sh_amp[numk + m], fz = sh_amp[2 * numk + m];
        s0 += fx.x * e.x - fx.y * e.y;
        s1 += fy.x * e.x - fy.y * e.y;
        s2 += fz.x * e.x - fz.y * e.y;
      }
    }
  }
  Fd[z] = s0;
  Fd[g.T + z] = s1;
  Fd[2 * g.T + z] = s2;
}

/* The field at n arbitrary points (xyz, 3 per point): the per-axis factors are formed per
 * point (nk <= 13 per axis), then summed as in k_ou_sample */
__global__ __launch_bounds__(kBlock) void kb_ou_sample(long n, const double *__restrict__ xyz,
                                                       const double2 *__restrict__ amp, int k1, double *__restrict__ F)
{
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int nk = 2 * k1 + 1, numk = nk * nk * nk;
  double2 e3[3][2 * kOuMaxK1 + 1];
  for (int d = 0; d < 3; ++d)
    for (int a = 0; a < nk; ++a) {
      const double ph = (2 * M_PI * (a - k1)) * xyz[3 * p + d];
      double sn, cs;
      sincos(ph, &sn, &cs);
      e3[d][a] = make_double2(cs, sn);
    }
  double s[3] = {0.0, 0.0, 0.0};
  for (int a = nk - 1; a >= 0; --a)
    for (int b = nk - 1; b >= 0; --b) {
      const double2 xy = cmul(e3[0][a], e3[1][b]);
      for (int c = nk - 1; c >= 0; --c) {
        const double2 e = cmul(xy, e3[2][c]);
        const int m = (a * nk + b) * nk + c;
#pragma unroll
        for (int q = 0; q < 3; ++q) s[q] += amp[q * numk + m].x * e.x - amp[q * numk + m].y * e.y;
      }
    }
  for (int q = 0; q < 3; ++q) F[3 * p + q] = s[q];
}

/* DrivingProcedure::Drive_rmhd (Mara/driving.cpp:154-186) on every zone: the force kicks
 * the four-velocity, u += F dt / W, and rho is rescaled so that D = rho W is unchanged */
__global__ __launch_bounds__(kBlock) void k_drive(double *__restrict__ P, const double *__restrict__ Fd, GridDesc g,
                                                  double dt)
{
#pragma clang fp contract(off)
  const long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long T = g.T;
  if (z >= T) return;
  double uf[4] = {0.0, P[vx * T + z], P[vy * T + z], P[vz * T + z]};
  const double v2 = uf[1] * uf[1] + uf[2] * uf[2] + uf[3] * uf[3];
  const double W0 = 1.0 / sqrt(1.0 - v2);
  uf[1] *= W0; uf[1] += Fd[z] * (dt / W0);
  uf[2] *= W0; uf[2] += Fd[T + z] * (dt / W0);
  uf[3] *= W0; uf[3] += Fd[2 * T + z] * (dt / W0);
  const double u2 = uf[1] * uf[1] + uf[2] * uf[2] + uf[3] * uf[3];
  const double W1 = sqrt(1.0 + u2);
  P[vx * T + z] = uf[1] / W1;
  P[vy * T + z] = uf[2] / W1;
  P[vz * T + z] = uf[3] / W1;
  P[rho * T + z] *= W0 / W1;
}

/* CoolingModuleT4 (kind 1) and CoolingModuleE4 (kind 2), Mara/cooling.cpp:13-103, on every
 * zone: the temperature (T4, in MeV) or the specific internal energy (E4) loses
 * dt (ref / t0) (x / ref)^4, the pressure is set from it, and the block's share of the
 * energy removed, (tau_before - tau_after) dV, is written to part[blockIdx.x] (fixed-order
 * block reduction) */
__global__ __launch_bounds__(kBlock) void k_cool(double *__restrict__ P, GridDesc g, double gamma, int kind,
                                                 double ref, double t0, double dt, double dV, MeasureUnits mu,
                                                 double *__restrict__ part)
{
#pragma clang fp contract(off)
  const long T = g.T;
  double removed = 0.0;
  for (long z = (long)blockIdx.x * blockDim.x + threadIdx.x; z < T; z += (long)gridDim.x * blockDim.x) {
    double P0[8], U0[8], U1[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) P0[q] = P[q * T + z];
    double T1;
    if (kind == 1) {
      const double p = P0[pre] / mu.ErgPerCC, r = P0[rho] / mu.GramsPerCC;   // TemperatureMeV
      const double T0 = ((p / r) * (mu.mp / mu.kb)) * mu.K2MeV;
      const double T1_MeV = T0 - dt * (ref / t0) * pow(T0 / ref, 4.0);
      const double rr = P0[rho] / mu.GramsPerCC;                            // TemperatureArb
      const double T_Kelvin = T1_MeV / mu.K2MeV;
      T1 = ((mu.kb * T_Kelvin) * (rr / mu.mp)) * mu.ErgPerCC;
    }
    else {
      const double e0 = (P0[pre] / (gamma - 1)) / P0[rho];
      const double e1 = e0 - dt * (ref / t0) * pow(e0 / ref, 4.0);
      T1 = (P0[rho] * e1) * (gamma - 1);
    }
    prim_to_cons_ref(P0, U0, gamma);
    P0[pre] = T1;
    prim_to_cons_ref(P0, U1, gamma);
    removed += (U0[tau] - U1[tau]) * dV;
    P[pre * T + z] = T1;
  }
  __shared__ double red[kBlock / 64];
  const double w = wave_sum_d(removed);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int i = 1; i < kBlock / 64; ++i) t += red[i];
    part[blockIdx.x] = t;
  }
}

/* AveragePowerInFields' numerator: per block, the sum of |F|^2 over its zones */
__global__ __launch_bounds__(kBlock) void k_power(const double *__restrict__ Fd, long T, double *__restrict__ part)
{
  double v = 0.0;
  for (long z = (long)blockIdx.x * blockDim.x + threadIdx.x; z < T; z += (long)gridDim.x * blockDim.x)
    v += Fd[z] * Fd[z] + Fd[T + z] * Fd[T + z] + Fd[2 * T + z] * Fd[2 * T + z];
  __shared__ double red[kBlock / 64];
  const double w = wave_sum_d(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int i = 1; i < kBlock / 64; ++i) t += red[i];
    part[blockIdx.x] = t;
  }
}

/* Adds the nb block sums, in a fixed order, to the running total *acc */
__global__ __launch_bounds__(kBlock) void k_add_partials(const double *__restrict__ part, int nb, double *acc)
{
  __shared__ double red[kBlock];
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v += part[b];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[0] += red[0];
}

__global__ void k_divB(const double *__restrict__ P, GridDesc g, double *out /* sum, max, absmax */)
{
#pragma clang fp contract(off)
  const long T = g.T;
  const int m0 = g.n[0] - 2 * g.Ng - 1, m1 = g.n[1] - 2 * g.Ng - 1, m2 = g.n[2] - 2 * g.Ng - 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long cnt = (long)m0 * m1 * m2;
  double d = 0.0;
  if (t < cnt) {
    const long k = t % m2 + g.Ng, j = (t / m2) % m1 + g.Ng, i = t / ((long)m2 * m1) + g.Ng;
    auto M = [&](long a, long b, long c) { return (a * g.n[1] + b) * g.n[2] + c; };
    const double *fx = P + (long)Bx * T, *fy = P + (long)By * T, *fz = P + (long)Bz * T;
    d = ((fx[M(i+1,j  ,k  )] + fx[M(i+1,j+1,k  )] + fx[M(i+1,j  ,k+1)] + fx[M(i+1,j+1,k+1)]) -
         (fx[M(i  ,j  ,k  )] + fx[M(i  ,j+1,k  )] + fx[M(i  ,j  ,k+1)] + fx[M(i  ,j+1,k+1)])) / 4.0
      + ((fy[M(i  ,j+1,k  )] + fy[M(i  ,j+1,k+1)] + fy[M(i+1,j+1,k  )] + fy[M(i+1,j+1,k+1)]) -
         (fy[M(i  ,j  ,k  )] + fy[M(i  ,j  ,k+1)] + fy[M(i+1,j  ,k  )] + fy[M(i+1,j  ,k+1)])) / 4.0
      + ((fz[M(i  ,j  ,k+1)] + fz[M(i+1,j  ,k+1)] + fz[M(i  ,j+1,k+1)] + fz[M(i+1,j+1,k+1)]) -
         (fz[M(i  ,j  ,k  )] + fz[M(i+1,j  ,k  )] + fz[M(i  ,j+1,k  )] + fz[M(i+1,j+1,k  )])) / 4.0;
  }
  __shared__ double rs[kBlock / 64], rm[kBlock / 64], ra[kBlock / 64];
  double s = d, mx = (t < cnt) ? d : -INFINITY, ax = fabs(d);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off, 64);
    mx = fmax(mx, __shfl_xor(mx, off, 64));
    ax = fmax(ax, __shfl_xor(ax, off, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { rs[wave] = s; rm[wave] = mx; ra[wave] = ax; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double bs = 0.0, bm = -INFINITY, ba = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) { bs += rs[w]; bm = fmax(bm, rm[w]); ba = fmax(ba, ra[w]); }
    atomicAdd(&out[0], bs);
    // signed max via the order-preserving integer image of a double
    long long bits = __double_as_longlong(bm);
    bits = bits >= 0 ? bits : (bits ^ 0x7fffffffffffffffLL);
    atomicMax((long long *)&out[1], bits);
    atomicMax((unsigned long long *)&out[2], (unsigned long long)__double_as_longlong(ba));
  }
}

/* --------------------------------------------------------------------------
 * point-function kernels for the batch entry points
 * ------------------------------------------------------------------------ */
__global__ void kb_prim_to_cons(long n, const double *P, double *U, double gamma, int *err)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p[8], u[8];
  for (int q = 0; q < 8; ++q) p[q] = P[8 * i + q];
  err[i] = prim_to_cons(p, u, gamma - 1.0);
  for (int q = 0; q < 8; ++q) U[8 * i + q] = u[q];
}

__global__ void kb_cons_to_prim(long n, const double *U, double *P, double gamma, double pfloor,
                                int *err, int *iters)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p[8], u[8];
  for (int q = 0; q < 8; ++q) { p[q] = P[8 * i + q]; u[q] = U[8 * i + q]; }
  int it = 0, which = 0;
  err[i] = cons_to_prim(u, p, gamma, pfloor, true, it, which);
  iters[i] = it;
  for (int q = 0; q < 8; ++q) P[8 * i + q] = p[q];
}

template <int DIM>
__global__ void kb_flux_eigen(long n, const double *P, double *F, double *ap, double *am, double gamma)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p[8], u[8], f[8], a, b;
  for (int q = 0; q < 8; ++q) p[q] = P[8 * i + q];
  prim_to_cons(p, u, gamma - 1.0);
  flux_and_eigenvalues<double, DIM>(u, p, f, a, b, gamma);
  for (int q = 0; q < 8; ++q) F[8 * i + q] = f[q];
  ap[i] = a;
  am[i] = b;
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void kb_eigensystem(long n, const double *P, double *L, double *R,
                                                         double *lam, int *ok, double gamma)
{
  const long i = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 3;   // one state per 8 lanes
  const int lane = (int)(threadIdx.x & 7);
  if (i >= n) return;
  double p[8], Lrow[8], Rrow[8], ev[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) p[q] = P[8 * i + q];
  const int rc = eig::g8::rmhd_eigensystem<DIM>(p, lane, Lrow, Rrow, ev, gamma);
  if (rc != 0) eig::g8::identity_rows(lane, Lrow, Rrow, ev);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    L[64 * i + 8 * lane + q] = Lrow[q];
    R[64 * i + 8 * lane + q] = Rrow[q];
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) lam[8 * i + q] = ev[q];
    ok[i] = rc == 0;
  }
}

__global__ void kb_reconstruct(long n, const double *V, double *out, int op, double theta, int IS,
                               double sz10A)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *v = V + 6 * i + 2;  // v[-2..3]
  double r = 0.0;
  switch (op) {
  case 0: r = plm<double>(v[-1], v[0], v[1], -1.0, theta); break;  // PLM_C2L
  case 1: r = plm<double>(v[-1], v[0], v[1], +1.0, theta); break;  // PLM_C2R
  case 2: r = weno5<double, true, true>(v[-2], v[-1], v[0], v[1], v[2], IS, sz10A); break;   // WENO5_FD_C2R
  case 3: r = weno5<double, false, true>(v[-2], v[-1], v[0], v[1], v[2], IS, sz10A); break;  // WENO5_FD_C2L
  case 4: r = weno5<double, true>(v[-2], v[-1], v[0], v[1], v[2], IS, sz10A); break;
  case 5: r = weno5<double, false>(v[-2], v[-1], v[0], v[1], v[2], IS, sz10A); break;
  default: r = 0.0;
  }
  out[i] = r;
}

template <typename R, int DIM>
__global__ void kb_riemann(long n, int solver, const double *PL, const double *PR, double *F,
                           double gamma, double pfloor, int *err, double *lamo)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R pl[8], pr[8], f[8];
  for (int q = 0; q < 8; ++q) { pl[q] = (R)PL[8 * i + q]; pr[q] = (R)PR[8 * i + q]; f[q] = R(0); }
  R lam = R(0);
  FluxCounters fc = {0, 0, 0, 0};
  err[i] = riemann_dispatch<R, DIM>(solver, pl, pr, f, (R)gamma, pfloor, lam, fc);
  for (int q = 0; q < 8; ++q) F[8 * i + q] = (double)f[q];
  lamo[i] = (double)lam;
}

inline unsigned nblocks(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

/* ============================================================================
 * context
 * ========================================================================== */
namespace {
/* a device buffer freed with its scope */
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) hipFree(p); }
  template <typename T> T *get() { return static_cast<T *>(p); }
};
int dalloc(DevBuf &b, size_t bytes) { return hipMalloc(&b.p, bytes ? bytes : 1) == hipSuccess ? 0 : MARAGPU_E_NOMEM; }
}  // namespace

struct maragpu_ctx {
  maragpu_config cfg;
  GridDesc g;
  int device = 0;                    // the HIP device current at create
  hipStream_t stream = nullptr;
  // HLLD sweeps: the redo of an axis runs on `side`, beside the next axis's sweep, and the
  // update waits for all three (redo_done); each axis has its own queue and counter slots
  hipStream_t side = nullptr;
  hipEvent_t sweep_done[3] = {nullptr, nullptr, nullptr}, redo_done[3] = {nullptr, nullptr, nullptr};
  unsigned redo_pending = 0;         // bit d: the redo of axis d+1 may still run on `side`
  double *Pc = nullptr, *Pw = nullptr, *U0 = nullptr, *U = nullptr, *A = nullptr;
  void *F[3] = {nullptr, nullptr, nullptr};
  DevStats *dstats = nullptr;
  DevStats *hstats = nullptr;       // pinned
  double *ddiv = nullptr;
  double max_lambda = 0.0;
  bool have_state = false;
  // per-kernel timing of the last advance
  std::vector<hipEvent_t> ev;
  double flux_ms = 0.0, c2p_ms = 0.0, upd_ms = 0.0;
  int flux_n = 0, c2p_n = 0, upd_n = 0;
  // learned c2p network (not owned; see maragpu_attach_mlp) and, in the non-plain forms, each
  // zone's Newton start for its next c2p pass within a step (the (Z, W) of its last
  // solution, in FP32: a start needs no more)
  const maragpu_mlp *mlp = nullptr;
  float2 *start = nullptr;
  // the c2p passes' per-wave counters (one slot per wave of a pass)
  uint4 *c2p_part = nullptr;
  // mixed form with HLLD: the FP64 redo queue of each sweep
  RedoQueue redo[3] = {};           // one per axis (plm-muscl's face passes use the first)
  // weno-split: U with its ghost images, and the zone fluxes (8 planes) and speeds (1 plane)
  // of the current sweep in the flux precision
  double *Uw = nullptr;
  void *FZ = nullptr;
  // the measure table: per-block partials and the reduced values
  double *dmeas = nullptr;
  // plm-muscl: the three predicted conserved states and their primitives (8 planes each)
  double *ctuU = nullptr, *ctuP = nullptr;
  // driving (DrivingProcedure): the field sampled at every zone, 3 planes of T doubles
  double *Fdrive = nullptr;
  // cooling (CoolingModuleT4 / E4): kind 0 none, 1 T4, 2 E4; per-block partials of the
  // energy removed and, after them, its running total since the last read
  int cool_kind = 0;
  double cool_ref = 0.0, cool_t0 = 0.0;
  double cool_units[3] = {1.0, 1.0, 1.0};
  double *dcool = nullptr;
};

namespace {

int fail_free(maragpu_ctx *c)
{
  maragpu_destroy(c);
  return MARAGPU_E_NOMEM;
}

GridDesc make_grid(const maragpu_config &c)
{
  // SimpleCartesianDomain (Mara/simple-cart.cpp:8-72) and prepare_integration (Mara/hydro.cpp:80-110)
  GridDesc g;
  double gdx[3];
  double min_dx = c.x1[0] - c.x0[0];
  for (int d = 0; d < 3; ++d) {
    g.N[d] = c.N[d];
    g.n[d] = c.N[d] + 2 * c.Ng;
    gdx[d] = (c.x1[d] - c.x0[d]) / c.N[d];
    if (gdx[d] < min_dx) min_dx = gdx[d];
  }
  g.Ng = c.Ng;
  g.T = (long)g.n[0] * g.n[1] * g.n[2];
  g.sz = 1;
  g.sy = g.n[2];
  g.sx = (long)g.n[1] * g.n[2];
  double lx0[3], lx1[3];
  for (int d = 0; d < 3; ++d) {
    lx0[d] = c.x0[d] + gdx[d] * 0;
    lx1[d] = c.x0[d] + gdx[d] * (0 + c.N[d]);
  }
  g.dx = (lx1[0] - lx0[0]) / (g.n[0] - 2 * c.Ng);
  g.dy = (lx1[1] - lx0[1]) / (g.n[1] - 2 * c.Ng);
  g.dz = (lx1[2] - lx0[2]) / (g.n[2] - 2 * c.Ng);
  g.min_dx = min_dx;
  return g;
}

template <typename R, typename FT, int DIM>
void launch_flux(maragpu_ctx *c, const PhysParams &pp)
{
  const long S = DIM == 1 ? c->g.sx : DIM == 2 ? c->g.sy : c->g.sz;
  const long n = c->g.T - 5 * S;
  if (n <= 0) return;
  FT *F = static_cast<FT *>(c->F[DIM - 1]);
  RedoQueue none = { nullptr, 0u, nullptr, nullptr, 0u };
  switch (c->cfg.riemann) {
  case MARAGPU_HLL:
    k_flux<R, FT, DIM, RIEMANN_HLL><<<nblocks(n), kBlock, 0, c->stream>>>(c->Pw, F, c->g, c->cfg.recon, pp, c->dstats, none);
    break;
  case MARAGPU_HLLC:
    k_flux<R, FT, DIM, RIEMANN_HLLC><<<nblocks(n), kBlock, 0, c->stream>>>(c->Pw, F, c->g, c->cfg.recon, pp, c->dstats, none);
    break;
  default:
    // the sweep, then its queued interfaces by the full FP64 solver: in the mixed form the
    // FP32 HLLD failures, in the plain form whatever left the common path
    // the redo of this axis runs on the side stream while the next axis is swept: it reads
    // only P and this axis's queue and slots, and writes only this axis's fluxes
    RedoQueue &q = c->redo[DIM - 1];
    hipMemsetAsync(q.count, 0, sizeof(unsigned int), c->stream);
    const unsigned nb = (unsigned)sweep_blocks<DIM>(c->g, kFluxBlock<R>);
    k_flux_hlld_fast<R, DIM><<<nb, kFluxBlock<R>, 0, c->stream>>>(c->Pw, F, c->g, c->cfg.recon, pp, c->dstats, q);
    hipEventRecord(c->sweep_done[DIM - 1], c->stream);
    hipStreamWaitEvent(c->side, c->sweep_done[DIM - 1], 0);
    k_flux_redo<FT, DIM, RIEMANN_HLLD><<<kRedoBlocks, kBlock, 0, c->side>>>(c->Pw, F, c->g, c->cfg.recon, pp,
                                                                            c->dstats, q, nb * (kFluxBlock<R> / 64));
    hipEventRecord(c->redo_done[DIM - 1], c->side);
    c->redo_pending |= 1u << (DIM - 1);
    break;
  }
}

/* weno-split sweep along DIM: zone fluxes and speeds, then the split interface fluxes */
template <typename R, int DIM>
void launch_weno_split(maragpu_ctx *c, const PhysParams &pp)
{
  const GridDesc &g = c->g;
  R *FZ = static_cast<R *>(c->FZ);
  R *AZ = FZ + 8 * g.T;
  k_zone_flux<R, DIM><<<nblocks(g.T), kBlock, 0, c->stream>>>(c->Uw, c->Pw, FZ, AZ, g, pp, c->dstats);
  const long cnt = (long)(DIM == 1 ? g.N[0] + 1 : g.n[0]) * (DIM == 2 ? g.N[1] + 1 : g.n[1]) *
                   (DIM == 3 ? g.N[2] + 1 : g.n[2]);
  if (c->cfg.eigen == MARAGPU_EIG_RMHD)
    k_flux_weno_split_char<R, DIM><<<nblocks(8 * cnt), kBlock, 0, c->stream>>>(
        c->Uw, c->Pw, FZ, AZ, static_cast<R *>(c->F[DIM - 1]), g, pp, c->dstats);
  else
    k_flux_weno_split<R, R, DIM><<<nblocks(cnt), kBlock, 0, c->stream>>>(c->Uw, FZ, AZ,
                                                                          static_cast<R *>(c->F[DIM - 1]), g, pp);
}

struct EventPair { int a, b; };

/* Per-block partials of the measure table for maragpu_measure_noalloc, in a file-scope
 * device array: (kMeasBlocks + 1) rows of kMeasQ doubles, 270 KB per device, held from
 * the library's load until the process ends. Its device address on each GPU is looked up
 * once, with that GPU current, when the first context on it is created. */
constexpr int kMaxDevices = 64;
__device__ __align__(256) double scratch[(kMeasBlocks + 1) * kMeasQ];
double *g_measure_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

int measure_scratch_init(int dev)
{
  if (dev < 0 || dev >= kMaxDevices) return MARAGPU_E_INVALID;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_measure_scratch[dev]) return 0;
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return MARAGPU_E_HIP;
  g_measure_scratch[dev] = static_cast<double *>(ptr);
  return 0;
}

}  // namespace

extern "C" int maragpu_mlp_device_params(const maragpu_mlp *n, void *out_params);

extern "C" {

const char *maragpu_strerror(int code)
{
  switch (code) {
  case MARAGPU_OK: return "success";
  case MARAGPU_E_INVALID: return "invalid argument or configuration";
  case MARAGPU_E_HIP: return "HIP runtime error";
  case MARAGPU_E_NOMEM: return "device allocation failed";
  case MARAGPU_E_SIZE: return "buffer size does not match the domain";
  case MARAGPU_E_NOMODEL: return "learned c2p requested but no network is loaded";
  case MARAGPU_E_IO: return "checkpoint file error";
  default: return code > 0 ? "zones failed to recover primitives (state not committed)" : "unknown error";
  }
}

int maragpu_create(const maragpu_config *cfg, maragpu_ctx **out)
{
  if (!cfg || !out) return MARAGPU_E_INVALID;
  *out = nullptr;
  for (int d = 0; d < 3; ++d)
    if (cfg->N[d] < 1 || !(cfg->x1[d] > cfg->x0[d])) return MARAGPU_E_INVALID;
  const int need_ng = cfg->recon == MARAGPU_WENO5 ? 3 : cfg->recon == MARAGPU_PLM ? 2 : 1;
  if (cfg->Ng < need_ng || cfg->Ng > 8) return MARAGPU_E_INVALID;
  // every ghost zone is the periodic image of an interior zone one period away
  for (int d = 0; d < 3; ++d)
    if (cfg->N[d] < cfg->Ng) return MARAGPU_E_INVALID;
  if (cfg->rk < 1 || cfg->rk > 4 || cfg->riemann < 0 || cfg->riemann > 2 || cfg->recon < 0 ||
      cfg->recon > 2 || cfg->precision < 0 || cfg->precision > 1 || !(cfg->gamma > 1.0))
    return MARAGPU_E_INVALID;
  if (cfg->godunov != MARAGPU_PLM_SPLIT && cfg->godunov != MARAGPU_WENO_SPLIT && cfg->godunov != MARAGPU_PLM_MUSCL)
    return MARAGPU_E_INVALID;
  // plm-muscl runs the reference's FP64 operator with the plain c2p
  if (cfg->godunov == MARAGPU_PLM_MUSCL && (cfg->precision != MARAGPU_FP64 || cfg->c2p != MARAGPU_C2P_PLAIN))
    return MARAGPU_E_INVALID;
  if (cfg->eigen != MARAGPU_EIG_COMPONENT && cfg->eigen != MARAGPU_EIG_RMHD) return MARAGPU_E_INVALID;
  if (cfg->godunov == MARAGPU_WENO_SPLIT && cfg->Ng < 3) return MARAGPU_E_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MARAGPU_E_HIP;

  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return MARAGPU_E_HIP;
  if (measure_scratch_init(dev)) return MARAGPU_E_HIP;
  maragpu_ctx *c = new (std::nothrow) maragpu_ctx;
  if (!c) return MARAGPU_E_NOMEM;
  c->device = dev;
  c->cfg = *cfg;
  c->g = make_grid(*cfg);
  if (c->g.T >= (1L << 32)) { delete c; return MARAGPU_E_INVALID; }   // 32-bit queue slots
  if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return MARAGPU_E_HIP; }
  if (hipStreamCreate(&c->side) != hipSuccess) return fail_free(c);
  for (int d = 0; d < 3; ++d)
    if (hipEventCreateWithFlags(&c->sweep_done[d], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->redo_done[d], hipEventDisableTiming) != hipSuccess)
      return fail_free(c);
  const size_t nd = 8 * (size_t)c->g.T;
  const size_t fbytes = nd * (cfg->precision == MARAGPU_MIXED ? sizeof(float) : sizeof(double));
  if (hipMalloc(&c->Pc, nd * sizeof(double)) != hipSuccess) return fail_free(c);
  if (hipMalloc(&c->Pw, nd * sizeof(double)) != hipSuccess) return fail_free(c);
  if (hipMalloc(&c->U0, nd * sizeof(double)) != hipSuccess) return fail_free(c);
  if (hipMalloc(&c->U, nd * sizeof(double)) != hipSuccess) return fail_free(c);
  if (cfg->rk == 4 && hipMalloc(&c->A, nd * sizeof(double)) != hipSuccess) return fail_free(c);
  for (int d = 0; d < 3; ++d)
    if (hipMalloc(&c->F[d], fbytes) != hipSuccess) return fail_free(c);
  if (hipMalloc(&c->dstats, sizeof(DevStats)) != hipSuccess) return fail_free(c);
  if (hipHostMalloc(&c->hstats, sizeof(DevStats), hipHostMallocDefault) != hipSuccess) return fail_free(c);
  if (hipMalloc(&c->ddiv, 4 * sizeof(double)) != hipSuccess) return fail_free(c);
  // per-wave counter slots of the c2p passes: the update's grid, or the flat grid of the
  // step start, whichever is larger
  const unsigned part_blocks = std::max(update_blocks(cfg->N), nblocks((long)cfg->N[0] * cfg->N[1] * cfg->N[2]));
  if (hipMalloc(&c->c2p_part, (size_t)part_blocks * (kBlock / 64) * sizeof(uint4)) != hipSuccess)
    return fail_free(c);
  if ((cfg->c2p == MARAGPU_C2P_LEARNED || cfg->precision == MARAGPU_MIXED) &&
      hipMalloc(&c->start, (size_t)cfg->N[0] * cfg->N[1] * cfg->N[2] * sizeof(float2)) != hipSuccess)
    return fail_free(c);
  if (cfg->godunov == MARAGPU_WENO_SPLIT) {
    const size_t rb = cfg->precision == MARAGPU_MIXED ? sizeof(float) : sizeof(double);
    if (hipMalloc(&c->Uw, nd * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->FZ, (size_t)9 * c->g.T * rb) != hipSuccess)
      return fail_free(c);
  }
  if (cfg->godunov == MARAGPU_PLM_MUSCL) {
    if (hipMalloc(&c->Uw, nd * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->ctuU, 3 * nd * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->ctuP, 3 * nd * sizeof(double)) != hipSuccess)
      return fail_free(c);
  }
  if (cfg->riemann == MARAGPU_HLLD && cfg->godunov != MARAGPU_WENO_SPLIT) {
    // a slot for every interface of a sweep or face pass (4 bytes a slot: half a state plane)
    const long cap = c->g.T;
    // the HLLD sweeps' per-wave counters: one slot per wave of the largest sweep
    long parts = 0;
    parts = std::max(parts, sweep_blocks<1>(c->g, kFluxBlock<float>) * (kFluxBlock<float> / 64));
    parts = std::max(parts, sweep_blocks<2>(c->g, kFluxBlock<float>) * (kFluxBlock<float> / 64));
    parts = std::max(parts, sweep_blocks<3>(c->g, kFluxBlock<float>) * (kFluxBlock<float> / 64));
    parts = std::max(parts, sweep_blocks<1>(c->g, kFluxBlock<double>) * (kFluxBlock<double> / 64));
    parts = std::max(parts, sweep_blocks<2>(c->g, kFluxBlock<double>) * (kFluxBlock<double> / 64));
    parts = std::max(parts, sweep_blocks<3>(c->g, kFluxBlock<double>) * (kFluxBlock<double> / 64));
    parts = std::max(parts, (long)nblocks(c->g.T) * (kBlock / 64));   // plm-muscl's face passes
    for (RedoQueue &q : c->redo) {
      q.cap = (unsigned int)cap;
      q.part_cap = (unsigned int)parts;
      if (hipMalloc(&q.count, sizeof(unsigned int)) != hipSuccess ||
          hipMalloc(&q.z, (size_t)cap * sizeof(unsigned int)) != hipSuccess ||
          hipMalloc(&q.part, (size_t)parts * sizeof(uint4)) != hipSuccess)
        return fail_free(c);
    }
  }
  // fluxes outside the sweep ranges are never written; keep them defined
  for (int d = 0; d < 3; ++d)
    if (hipMemsetAsync(c->F[d], 0, fbytes, c->stream) != hipSuccess) return fail_free(c);
  if (hipMemsetAsync(c->Pc, 0, nd * sizeof(double), c->stream) != hipSuccess) return fail_free(c);
  c->ev.resize(64);
  for (auto &e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) { e = nullptr; return fail_free(c); }
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail_free(c);
  *out = c;
  return MARAGPU_OK;
}


int maragpu_destroy(maragpu_ctx *c)
{
  if (!c) return MARAGPU_E_INVALID;
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->side) hipStreamSynchronize(c->side);
  hipFree(c->Pc); hipFree(c->Pw); hipFree(c->U0); hipFree(c->U); hipFree(c->A);
  for (int d = 0; d < 3; ++d) hipFree(c->F[d]);
  hipFree(c->dstats);
  if (c->hstats) hipHostFree(c->hstats);
  hipFree(c->ddiv);
  hipFree(c->start);
  hipFree(c->c2p_part);
  for (RedoQueue &q : c->redo) {
    hipFree(q.count);
    hipFree(q.z);
    hipFree(q.part);
  }
  hipFree(c->Uw);
  hipFree(c->FZ);
  hipFree(c->dmeas);
  hipFree(c->ctuU);
  hipFree(c->ctuP);
  hipFree(c->Fdrive);
  hipFree(c->dcool);
  for (auto &e : c->ev) if (e) hipEventDestroy(e);
  for (int d = 0; d < 3; ++d) {
    if (c->sweep_done[d]) hipEventDestroy(c->sweep_done[d]);
    if (c->redo_done[d]) hipEventDestroy(c->redo_done[d]);
  }
  if (c->side) hipStreamDestroy(c->side);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return MARAGPU_OK;
}

int maragpu_upload_prim(maragpu_ctx *c, const double *P, size_t n)
{
  if (!c || !P) return MARAGPU_E_INVALID;
  if (n != 8 * (size_t)c->g.T) return MARAGPU_E_SIZE;
  HIP_CHECK(hipMemcpyAsync(c->U0, P, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_aos_to_soa<<<nblocks(c->g.T), kBlock, 0, c->stream>>>(c->U0, c->Pc, c->g.T);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->have_state = true;
  return MARAGPU_OK;
}

int maragpu_download_prim(maragpu_ctx *c, double *P, size_t n)
{
  if (!c || !P) return MARAGPU_E_INVALID;
  if (n != 8 * (size_t)c->g.T) return MARAGPU_E_SIZE;
  k_soa_to_aos<<<nblocks(c->g.T), kBlock, 0, c->stream>>>(c->Pc, c->U0, c->g.T);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(P, c->U0, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return MARAGPU_OK;
}

int maragpu_get_config(const maragpu_ctx *c, maragpu_config *cfg)
{
  if (!c || !cfg) return MARAGPU_E_INVALID;
  *cfg = c->cfg;
  return MARAGPU_OK;
}

int maragpu_download_field(maragpu_ctx *c, int q, double *out, size_t n)
{
  if (!c || !out || q < 0 || q >= 8) return MARAGPU_E_INVALID;
  const GridDesc &g = c->g;
  const long ni = (long)g.N[0] * g.N[1] * g.N[2];
  if (n != (size_t)ni) return MARAGPU_E_SIZE;
  // U0 (8 T doubles) is staging outside an advance
  k_field_out<<<nblocks(ni), kBlock, 0, c->stream>>>(c->Pc + (size_t)q * g.T, c->U0, g);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(out, c->U0, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return MARAGPU_OK;
}

int maragpu_upload_field(maragpu_ctx *c, int q, const double *in, size_t n)
{
  if (!c || !in || q < 0 || q >= 8) return MARAGPU_E_INVALID;
  const GridDesc &g = c->g;
  const long ni = (long)g.N[0] * g.N[1] * g.N[2];
  if (n != (size_t)ni) return MARAGPU_E_SIZE;
  HIP_CHECK(hipMemcpyAsync(c->U0, in, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_field_in<<<nblocks(ni), kBlock, 0, c->stream>>>(c->U0, c->Pc + (size_t)q * g.T, g);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->have_state = true;
  return MARAGPU_OK;
}

int maragpu_advance(maragpu_ctx *c, double dt, maragpu_stats *stats)
{
  if (!c) return MARAGPU_E_INVALID;
  if (!c->have_state) return MARAGPU_E_INVALID;
  if (c->cfg.c2p == MARAGPU_C2P_LEARNED && !c->mlp) return MARAGPU_E_NOMODEL;
  const GridDesc &g = c->g;
  const long ni = (long)g.N[0] * g.N[1] * g.N[2];
  PhysParams pp;
  pp.gamma = c->cfg.gamma;
  pp.pfloor = c->cfg.pfloor;
  pp.theta = c->cfg.theta;
  pp.IS = c->cfg.IS;
  pp.sz10A = c->cfg.sz10A;
  const bool mixed = c->cfg.precision == MARAGPU_MIXED;
  int evi = 0;
  std::vector<EventPair> ef, ec, eu;
  auto mark = [&](std::vector<EventPair> &v) {
    EventPair p = { evi, evi + 1 };
    evi += 2;
    v.push_back(p);
    return p;
  };

  HIP_CHECK(hipMemsetAsync(c->dstats, 0, sizeof(DevStats), c->stream));
  HIP_CHECK(hipEventRecord(c->ev[evi], c->stream));
  const int ev_start = evi++;
  // each update runs together with the c2p pass after it, and the step's PrimToCons with
  // its first pass
  const bool learned = c->cfg.c2p == MARAGPU_C2P_LEARNED;
  MlpDev net = {};
  if (learned && maragpu_mlp_device_params(c->mlp, &net)) return MARAGPU_E_NOMODEL;
  const unsigned nbl = nblocks(ni) < kLearnedBlocks ? nblocks(ni) : kLearnedBlocks;
  // the non-plain forms chain each zone's Newton start from solution to solution
  const bool chain = c->start != nullptr;
  // the three axis sweeps, z first: their results do not depend on the order, and with HLLD
  // the redo of the last axis cannot overlap a sweep, so it is the cheapest one (x at c2)
  auto sweeps = [&]() {
    for (int d = 3; d >= 1; --d) {
      EventPair p = mark(ef);
      hipEventRecord(c->ev[p.a], c->stream);
      if (c->cfg.godunov == MARAGPU_WENO_SPLIT) {
        if (mixed) {
          if (d == 1) launch_weno_split<float, 1>(c, pp);
          if (d == 2) launch_weno_split<float, 2>(c, pp);
          if (d == 3) launch_weno_split<float, 3>(c, pp);
        }
        else {
          if (d == 1) launch_weno_split<double, 1>(c, pp);
          if (d == 2) launch_weno_split<double, 2>(c, pp);
          if (d == 3) launch_weno_split<double, 3>(c, pp);
        }
      }
      else if (mixed) {
        if (d == 1) launch_flux<float, float, 1>(c, pp);
        if (d == 2) launch_flux<float, float, 2>(c, pp);
        if (d == 3) launch_flux<float, float, 3>(c, pp);
      }
      else {
        if (d == 1) launch_flux<double, double, 1>(c, pp);
        if (d == 2) launch_flux<double, double, 2>(c, pp);
        if (d == 3) launch_flux<double, double, 3>(c, pp);
      }
      hipEventRecord(c->ev[p.b], c->stream);
    }
  };
  std::vector<int> kinds;
  switch (c->cfg.rk) {
  case 1: kinds = { ST_EULER }; break;
  case 2: kinds = { ST_EULER, ST_RK2_FINAL }; break;
  case 4: kinds = { ST_RK4_1, ST_RK4_2, ST_RK4_3, ST_RK4_4 }; break;
  default: kinds = { ST_EULER, ST_RK3_2, ST_RK3_3 }; break;
  }
  const bool ctu = c->cfg.godunov == MARAGPU_PLM_MUSCL;
  auto ctu_fluxes = [&](int s) {
    // PlmCtuHancockOperator::ctu_hancock on the stage primitives Pw and the synchronised U (Uw)
    EventPair p = mark(ef);
    hipEventRecord(c->ev[p.a], c->stream);
    const long T = g.T, sxz = g.sx;
    const size_t nd = 8 * (size_t)T;
    double *F0 = static_cast<double *>(c->F[0]), *F1 = static_cast<double *>(c->F[1]),
           *F2 = static_cast<double *>(c->F[2]);
    double *Ux = c->ctuU, *Uy = c->ctuU + nd, *Uz = c->ctuU + 2 * nd;
    double *Px = c->ctuP, *Py = c->ctuP + nd, *Pz = c->ctuP + 2 * nd;
    unsigned long long *fl = c->dstats->ctu_failed[s];
    const int rs = c->cfg.riemann;
    // the faces of one axis over [begin, T - sx): HLLD on the common path and its queue,
    // HLL and HLLC by the full kernel
    auto faces = [&](int dim, const double *Ps, double *F, long begin, int keep, unsigned long long *f) {
      const long n = T - sxz - begin;
      if (n <= 0) return;
      if (c->redo[0].z) {
        hipMemsetAsync(c->redo[0].count, 0, sizeof(unsigned int), c->stream);
        if (dim == 1) k_ctu_faces_fast<1><<<nblocks(n), kBlock, 0, c->stream>>>(c->Pw, Ps, F, g, pp, begin, c->dstats, f, c->redo[0]);
        if (dim == 2) k_ctu_faces_fast<2><<<nblocks(n), kBlock, 0, c->stream>>>(c->Pw, Ps, F, g, pp, begin, c->dstats, f, c->redo[0]);
        if (dim == 3) k_ctu_faces_fast<3><<<nblocks(n), kBlock, 0, c->stream>>>(c->Pw, Ps, F, g, pp, begin, c->dstats, f, c->redo[0]);
        const unsigned np = nblocks(n) * (kBlock / 64);
        if (dim == 1) k_ctu_faces_redo<1><<<kRedoBlocks, kBlock, 0, c->stream>>>(c->Pw, Ps, F, g, pp, keep, c->dstats, f, c->redo[0], np);
        if (dim == 2) k_ctu_faces_redo<2><<<kRedoBlocks, kBlock, 0, c->stream>>>(c->Pw, Ps, F, g, pp, keep, c->dstats, f, c->redo[0], np);
        if (dim == 3) k_ctu_faces_redo<3><<<kRedoBlocks, kBlock, 0, c->stream>>>(c->Pw, Ps, F, g, pp, keep, c->dstats, f, c->redo[0], np);
      }
      else {
        if (dim == 1) k_ctu_faces<1><<<nblocks(n), kBlock, 0, c->stream>>>(c->Pw, Ps, F, g, pp, rs, begin, keep, c->dstats, f);
        if (dim == 2) k_ctu_faces<2><<<nblocks(n), kBlock, 0, c->stream>>>(c->Pw, Ps, F, g, pp, rs, begin, keep, c->dstats, f);
        if (dim == 3) k_ctu_faces<3><<<nblocks(n), kBlock, 0, c->stream>>>(c->Pw, Ps, F, g, pp, rs, begin, keep, c->dstats, f);
      }
    };
    faces(1, c->Pw, F0, 0, 0, fl + 0);
    faces(2, c->Pw, F1, 0, 0, fl + 0);
    faces(3, c->Pw, F2, 0, 0, fl + 0);
    k_ctu_predict<<<nblocks(T - sxz), kBlock, 0, c->stream>>>(c->Pw, c->Uw, F0, F1, F2, Ux, Uy, Uz, g, pp, dt, fl + 1);
    k_ctu_c2p<<<nblocks(T), kBlock, 0, c->stream>>>(c->Pw, Ux, Uy, Uz, Px, Py, Pz, g, c->cfg.gamma, c->cfg.pfloor,
                                                    c->dstats, fl + 2);
    faces(1, Px, F0, sxz, 1, fl + 3);
    faces(2, Py, F1, sxz, 1, fl + 3);
    faces(3, Pz, F2, sxz, 1, fl + 3);
    hipEventRecord(c->ev[p.b], c->stream);
  };
  // the update reads the fluxes of all three axes: wait for the redo kernels on the side stream
  auto join_redo = [&]() {
    for (int d = 0; d < 3; ++d)
      if (c->redo_pending & (1u << d)) hipStreamWaitEvent(c->stream, c->redo_done[d], 0);
    c->redo_pending = 0;
  };
  auto update_c2p = [&](int kind, int write_u, int pass) {
    join_redo();
    EventPair p = mark(eu);
    hipEventRecord(c->ev[p.a], c->stream);
    const unsigned nb = update_blocks(g.N);
#define UPDATE_ARGS(FT) (const FT *)c->F[0], (const FT *)c->F[1], (const FT *)c->F[2], c->U0, c->U, c->A, c->Pw, \
    g, dt, kind, write_u, c->cfg.gamma, c->cfg.pfloor, c->cfg.ereset, c->cfg.preset, c->cfg.dreset, c->dstats, pass, c->Uw, \
    c->start, c->c2p_part
    if (chain && mixed) k_update_c2p<float, C2P_STORED><<<nb, kBlock, 0, c->stream>>>(UPDATE_ARGS(float));
    else if (chain) k_update_c2p<double, C2P_STORED><<<nb, kBlock, 0, c->stream>>>(UPDATE_ARGS(double));
    else if (mixed) k_update_c2p<float, C2P_CASCADE><<<nb, kBlock, 0, c->stream>>>(UPDATE_ARGS(float));
    else k_update_c2p<double, C2P_CASCADE><<<nb, kBlock, 0, c->stream>>>(UPDATE_ARGS(double));
#undef UPDATE_ARGS
    k_c2p_parts<4><<<kPartBlocks, kBlock, 0, c->stream>>>(c->c2p_part, nb * (kBlock / 64), c->dstats, pass);
    if (c->Uw && c->cfg.ereset)
      k_conserved_after_reset<<<nblocks(g.T), kBlock, 0, c->stream>>>(c->Pw, c->Uw, g, c->cfg.gamma,
                                                                      &c->dstats->stage_reset[pass]);
    hipEventRecord(c->ev[p.b], c->stream);
  };
  auto p2c_c2p = [&](int pass) {
    EventPair p = mark(ec);
    hipEventRecord(c->ev[p.a], c->stream);
#define P2C_ARGS c->Pc, c->U0, c->Pw, g, c->cfg.gamma, c->cfg.pfloor, c->cfg.ereset, c->cfg.preset, c->cfg.dreset, \
    c->dstats, pass, c->Uw, net
    if (learned) {
      k_p2c_c2p<C2P_NET><<<nbl, kBlock, 0, c->stream>>>(P2C_ARGS, c->start, c->c2p_part);
      k_c2p_parts<6><<<kPartBlocks, kBlock, 0, c->stream>>>(c->c2p_part, nbl * (kBlock / 64), c->dstats, pass);
    }
    else {
      k_p2c_c2p<C2P_CASCADE><<<nblocks(ni), kBlock, 0, c->stream>>>(P2C_ARGS, chain ? c->start : nullptr, c->c2p_part);
      k_c2p_parts<4><<<kPartBlocks, kBlock, 0, c->stream>>>(c->c2p_part, nblocks(ni) * (kBlock / 64), c->dstats,
                                                             pass);
    }
#undef P2C_ARGS
    if (c->Uw && c->cfg.ereset)
      k_conserved_after_reset<<<nblocks(g.T), kBlock, 0, c->stream>>>(c->Pw, c->Uw, g, c->cfg.gamma,
                                                                      &c->dstats->stage_reset[pass]);
    hipEventRecord(c->ev[p.b], c->stream);
  };
  int pass = 0;
  for (size_t s = 0; s < kinds.size(); ++s) {
    if (s == 0) p2c_c2p(pass++);
    if (ctu) ctu_fluxes((int)s);
    else sweeps();
    update_c2p(kinds[s], s + 1 < kinds.size() && stage_reads_u(kinds[s + 1]) ? 1 : 0, pass++);
  }
  join_redo();
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(c->ev[evi], c->stream));
  const int ev_end = evi++;
  HIP_CHECK(hipMemcpyAsync(c->hstats, c->dstats, sizeof(DevStats), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));

  const DevStats &h = *c->hstats;
  int errors = 0;
  if (ctu) {
    // PlmCtuHancockOperator::dUdt ignores the stage c2p count (Mara/ctu-hancock.cpp:52-69):
    // the first abort point with failures, else the final c2p
    for (size_t s = 0; s < kinds.size() && !errors; ++s)
      for (int a = 0; a < 4 && !errors; ++a) errors = (int)h.ctu_failed[s][a];
    if (!errors) errors = (int)h.stage_failed[pass - 1];
  }
  else {
    for (int p = 0; p < pass; ++p) {
      if (h.stage_failed[p]) { errors = (int)h.stage_failed[p]; break; }
    }
  }
  double lam;
  unsigned long long bits = h.lam_bits;
  std::memcpy(&lam, &bits, sizeof(lam));
  c->max_lambda = lam;
  if (errors == 0) {
    std::swap(c->Pc, c->Pw);
    // source steps on the committed state (luaC_Mara_advance, Mara/mara.cpp:330-333)
    if (c->Fdrive) k_drive<<<nblocks(g.T), kBlock, 0, c->stream>>>(c->Pc, c->Fdrive, g, dt);
    if (c->cool_kind) {
      MeasureUnits mu = measure_units(c->cool_units);
      const double dV = ((c->cfg.x1[0] - c->cfg.x0[0]) / g.N[0]) * ((c->cfg.x1[1] - c->cfg.x0[1]) / g.N[1]) *
                        ((c->cfg.x1[2] - c->cfg.x0[2]) / g.N[2]);
      k_cool<<<kCoolBlocks, kBlock, 0, c->stream>>>(c->Pc, g, c->cfg.gamma, c->cool_kind, c->cool_ref, c->cool_t0,
                                                    dt, dV, mu, c->dcool);
      k_add_partials<<<1, kBlock, 0, c->stream>>>(c->dcool, kCoolBlocks, c->dcool + kCoolBlocks);
    }
    if (c->Fdrive || c->cool_kind) HIP_CHECK(hipGetLastError());
  }

  auto sum_ms = [&](const std::vector<EventPair> &v) {
    double t = 0.0;
    for (auto &p : v) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, c->ev[p.a], c->ev[p.b]);
      t += ms;
    }
    return t;
  };
  c->flux_ms = sum_ms(ef); c->flux_n = (int)ef.size();
  c->c2p_ms = sum_ms(ec); c->c2p_n = (int)ec.size();
  c->upd_ms = sum_ms(eu); c->upd_n = (int)eu.size();
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->max_lambda = lam;
    stats->hlld_failure = (int64_t)h.hlld_failure;
    stats->hlld_success = (int64_t)h.hlld_success;
    stats->riemann_retries = (int64_t)h.retries;
    stats->riemann_failures = (int64_t)h.failures;
    stats->c2p_zones = (int64_t)h.c2p_zones;
    stats->c2p_iterations = (int64_t)h.c2p_iters;
    stats->c2p_fallbacks = (int64_t)h.c2p_fallbacks;
    stats->c2p_failed = errors;
    stats->hlld_c2p_iterations = (int64_t)h.hlld_c2p_iters;
    stats->hlld_newton_iterations = (int64_t)h.hlld_newton_iters;
    stats->mlp_guesses = (int64_t)h.mlp_guesses;
    stats->mlp_fallbacks = (int64_t)h.mlp_fallbacks;
    stats->fp64_redo = (int64_t)h.redo;
    stats->fp64_redo_overflow = (int64_t)h.redo_overflow;
    stats->eig_identity = (int64_t)h.eig_identity;
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev[ev_start], c->ev[ev_end]);
    stats->device_ms = ms;
  }
  return errors;
}

double maragpu_get_timestep(maragpu_ctx *c, double cfl)
{
  if (!c) return std::numeric_limits<double>::quiet_NaN();
  return cfl * c->g.min_dx / c->max_lambda;
}

int maragpu_measure_divB(maragpu_ctx *c, double *mean, double *max, double *absmax)
{
  if (!c) return MARAGPU_E_INVALID;
  const GridDesc &g = c->g;
  const long cnt = (long)(g.n[0] - 2 * g.Ng - 1) * (g.n[1] - 2 * g.Ng - 1) * (g.n[2] - 2 * g.Ng - 1);
  double init[3] = { 0.0, 0.0, 0.0 };
  {
    // signed-max slot starts at the integer image of -inf
    long long ninf = __builtin_bit_cast(long long, -std::numeric_limits<double>::infinity());
    ninf = ninf ^ 0x7fffffffffffffffLL;
    std::memcpy(&init[1], &ninf, sizeof(double));
  }
  HIP_CHECK(hipMemcpyAsync(c->ddiv, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
  if (cnt > 0) k_divB<<<nblocks(cnt), kBlock, 0, c->stream>>>(c->Pc, g, c->ddiv);
  double r[3];
  HIP_CHECK(hipMemcpyAsync(r, c->ddiv, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  long long bits;
  std::memcpy(&bits, &r[1], sizeof(bits));
  bits = bits >= 0 ? bits : (bits ^ 0x7fffffffffffffffLL);
  double smax;
  std::memcpy(&smax, &bits, sizeof(smax));
  // zones outside the stencil's loop hold 0 in the reference's div array
  if (g.T > cnt && smax < 0.0) smax = 0.0;
  const double glb = (double)g.N[0] * g.N[1] * g.N[2];
  if (mean) *mean = r[0] / (double)g.T / glb;
  if (max) *max = smax;
  if (absmax) *absmax = r[2] / g.min_dx;
  return MARAGPU_OK;
}

}  // extern "C"

namespace {
/* The measure table of the committed state with the per-block partials and the reduced
 * row in `part` ((kMeasBlocks + 1) x kMeasQ doubles, every row used written first). */
int measure_into(maragpu_ctx *c, const double *units, maragpu_measures *out, double *part)
{
  // PhysicalUnits(Length, Mass, Time), Mara/units.cpp:3-6 and :92-141; cgs by default
  // (HydroModule's units, Mara/hydro.cpp:20); constants of Mara/eos.hpp:12-15
  const double cgs[3] = {1.0, 1.0, 1.0};
  if (!units) units = cgs;
  if (!(units[0] > 0.0) || !(units[1] > 0.0) || !(units[2] > 0.0)) return MARAGPU_E_INVALID;
  const double Centimeter = 1.0 / units[0], Second = 1.0 / units[2];
  const double Erg = (1.0 / units[1]) * std::pow(Centimeter / Second, 2.0);
  const double Gauss = std::sqrt(Erg / std::pow(Centimeter, 3.0) / (4 * M_PI));
  const MeasureUnits mu = measure_units(units);
  const GridDesc &g = c->g;
  const long ni = (long)g.N[0] * g.N[1] * g.N[2];
  const int nb = (int)std::min<long>(kMeasBlocks, (ni + kBlock - 1) / kBlock);
  double *res = part + (size_t)kMeasBlocks * kMeasQ;
  k_measure<<<nb, kBlock, 0, c->stream>>>(c->Pc, g, c->cfg.gamma, mu, part);
  k_measure_reduce<<<kMeasQ, kBlock, 0, c->stream>>>(part, nb, res);
  double r[kMeasQ];
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(r, res, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  const int ttl = g.N[0] * g.N[1] * g.N[2];           // TTL_ZONES, an int in the reference
  for (int q = 0; q < 4; ++q) out->mean_velocity[q] = r[q] / ttl;
  for (int q = 0; q < 8; ++q) out->mean_cons[q] = r[4 + q] / ttl;
  for (int q = 0; q < 8; ++q) out->mean_prim[q] = r[12 + q] / ttl;
  for (int q = 0; q < 4; ++q) out->mean_energies[q] = r[20 + q] / ttl;
  out->mean_max_temperature[0] = r[24] / ttl;
  out->mean_max_temperature[1] = r[kMeasSums + 0];
  out->mean_max_sonic_mach[0] = r[25] / ttl;
  out->mean_max_sonic_mach[1] = r[kMeasSums + 1];
  out->mean_min_alfvenic_mach[0] = r[26] / ttl;
  out->mean_min_alfvenic_mach[1] = r[kMeasSums + 2];
  out->mean_max_magnetic_field[0] = (1.0 / Gauss) * r[27] / ttl;
  out->mean_max_magnetic_field[1] = (1.0 / Gauss) * r[kMeasSums + 3];
  out->max_lorentz_factor = r[kMeasSums + 4];
  return MARAGPU_OK;
}

}  // namespace

extern "C" {

int maragpu_measure(maragpu_ctx *c, const double *units, maragpu_measures *out)
{
  if (!c || !out) return MARAGPU_E_INVALID;
  if (!c->dmeas && hipMalloc(&c->dmeas, (size_t)(kMeasBlocks + 1) * kMeasQ * sizeof(double)) != hipSuccess) {
    c->dmeas = nullptr;
    return MARAGPU_E_NOMEM;
  }
  return measure_into(c, units, out, c->dmeas);
}

int maragpu_measure_noalloc(maragpu_ctx *c, const double *units, maragpu_measures *out)
{
  if (!c || !out) return MARAGPU_E_INVALID;
  if (c->device < 0 || c->device >= kMaxDevices || !g_measure_scratch[c->device]) return MARAGPU_E_INVALID;
  return measure_into(c, units, out, g_measure_scratch[c->device]);
}

namespace {
/* per-axis tables exp(2 pi i n x) at the zone centres x_at (Mara/simple-cart.hpp:42-44) for
 * n = -k1..k1, and the amplitudes (axis-major), on the device */
int ou_tables(const maragpu_host::OuField &f, const GridDesc &g, const maragpu_config &cfg, DevBuf &tab,
              DevBuf &amp)
{
  const int nk = 2 * f.k1 + 1;
  std::vector<double2> h;
  for (int d = 0; d < 3; ++d) {
    const double dx = (cfg.x1[d] - cfg.x0[d]) / cfg.N[d];
    for (int i = 0; i < g.n[d]; ++i) {
      const double x = cfg.x0[d] + dx * (i - g.Ng + 0.5);
      for (int a = 0; a < nk; ++a) {
        const double ph = (2 * M_PI * (a - f.k1)) * x;
        h.push_back(make_double2(std::cos(ph), std::sin(ph)));
      }
    }
  }
  std::vector<double2> ha(3 * (size_t)f.numk);
  for (size_t q = 0; q < ha.size(); ++q) ha[q] = make_double2(f.F[2 * q], f.F[2 * q + 1]);
  if (dalloc(tab, h.size() * sizeof(double2)) || dalloc(amp, ha.size() * sizeof(double2))) return MARAGPU_E_NOMEM;
  HIP_CHECK(hipMemcpy(tab.p, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(amp.p, ha.data(), ha.size() * sizeof(double2), hipMemcpyHostToDevice));
  return MARAGPU_OK;
}
}  // namespace

int maragpu_set_driving(maragpu_ctx *c, const maragpu_ou_field *field)
{
  if (!c) return MARAGPU_E_INVALID;
  if (!field) {
    hipFree(c->Fdrive);
    c->Fdrive = nullptr;
    return MARAGPU_OK;
  }
  const maragpu_host::OuField &f = *maragpu_ou_field_modes(field);
  if (f.k1 > kOuMaxK1) return MARAGPU_E_INVALID;
  const GridDesc &g = c->g;
  if (!c->Fdrive && hipMalloc(&c->Fdrive, 3 * (size_t)g.T * sizeof(double)) != hipSuccess) {
    c->Fdrive = nullptr;
    return MARAGPU_E_NOMEM;
  }
  DevBuf tab, amp;
  if (int rc = ou_tables(f, g, c->cfg, tab, amp)) return rc;
  const int nk = 2 * f.k1 + 1;
  const double2 *t = tab.get<double2>();
  k_ou_sample<<<nblocks(g.T), kBlock, 3 * f.numk * sizeof(double2), c->stream>>>(
      t, t + (size_t)g.n[0] * nk, t + (size_t)(g.n[0] + g.n[1]) * nk, amp.get<double2>(), f.k1, g, c->Fdrive);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return MARAGPU_OK;
}

int maragpu_driving_power(maragpu_ctx *c, double *power)
{
  if (!c || !power) return MARAGPU_E_INVALID;
  if (!c->Fdrive) { *power = 0.0; return MARAGPU_OK; }
  // AveragePowerInFields (Mara/driving.cpp:30-39): mean of |F|^2 over every zone
  DevBuf part;
  if (dalloc(part, (kCoolBlocks + 1) * sizeof(double))) return MARAGPU_E_NOMEM;
  double *acc = part.get<double>() + kCoolBlocks;
  const double zero = 0.0;
  HIP_CHECK(hipMemcpyAsync(acc, &zero, sizeof(double), hipMemcpyHostToDevice, c->stream));
  k_power<<<kCoolBlocks, kBlock, 0, c->stream>>>(c->Fdrive, c->g.T, part.get<double>());
  k_add_partials<<<1, kBlock, 0, c->stream>>>(part.get<double>(), kCoolBlocks, acc);
  double s = 0.0;
  HIP_CHECK(hipMemcpyAsync(&s, acc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  *power = s / c->g.T;
  return MARAGPU_OK;
}

int maragpu_set_cooling(maragpu_ctx *c, int kind, double ref, double t0, const double *units)
{
  if (!c || kind < 0 || kind > 2) return MARAGPU_E_INVALID;
  if (kind && (!(ref > 0.0) || !(t0 > 0.0))) return MARAGPU_E_INVALID;
  if (units && (!(units[0] > 0.0) || !(units[1] > 0.0) || !(units[2] > 0.0))) return MARAGPU_E_INVALID;
  if (kind && !c->dcool) {
    if (hipMalloc(&c->dcool, (kCoolBlocks + 1) * sizeof(double)) != hipSuccess) {
      c->dcool = nullptr;
      return MARAGPU_E_NOMEM;
    }
    const double zero = 0.0;
    HIP_CHECK(hipMemcpy(c->dcool + kCoolBlocks, &zero, sizeof(double), hipMemcpyHostToDevice));
  }
  c->cool_kind = kind;
  c->cool_ref = ref;
  c->cool_t0 = t0;
  for (int d = 0; d < 3; ++d) c->cool_units[d] = units ? units[d] : 1.0;
  return MARAGPU_OK;
}

int maragpu_cooling_energy_removed(maragpu_ctx *c, double *removed)
{
  if (!c || !removed) return MARAGPU_E_INVALID;
  *removed = 0.0;
  if (!c->dcool) return MARAGPU_OK;
  const double zero = 0.0;
  HIP_CHECK(hipMemcpyAsync(removed, c->dcool + kCoolBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipMemcpyAsync(c->dcool + kCoolBlocks, &zero, sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return MARAGPU_OK;
}

int maragpu_kernel_times(maragpu_ctx *c, double *flux_ms, int *flux_n, double *c2p_ms, int *c2p_n,
                         double *upd_ms, int *upd_n)
{
  if (!c) return MARAGPU_E_INVALID;
  if (flux_ms) *flux_ms = c->flux_ms;
  if (flux_n) *flux_n = c->flux_n;
  if (c2p_ms) *c2p_ms = c->c2p_ms;
  if (c2p_n) *c2p_n = c->c2p_n;
  if (upd_ms) *upd_ms = c->upd_ms;
  if (upd_n) *upd_n = c->upd_n;
  return MARAGPU_OK;
}

int maragpu_attach_mlp(maragpu_ctx *c, const maragpu_mlp *net)
{
  if (!c) return MARAGPU_E_INVALID;
  c->mlp = net;
  return MARAGPU_OK;
}

/* Committed primitive state on the device (SoA planes of T doubles) and the grid shape;
 * used to sample training pairs from a running simulation. */
const double *maragpu_committed_prim(maragpu_ctx *c, int *N, int *Ng, long *T, int *n1, int *n2)
{
  if (!c || !c->have_state) return nullptr;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return nullptr;
  for (int d = 0; d < 3; ++d) N[d] = c->g.N[d];
  *Ng = c->g.Ng;
  *T = c->g.T;
  *n1 = c->g.n[1];
  *n2 = c->g.n[2];
  return c->Pc;
}

}  // extern "C"

/* ---------------------------------------------------------------------------
 * batch entry points
 * ------------------------------------------------------------------------- */

extern "C" {

int maragpu_batch_prim_to_cons(size_t n, const double *P, double *U, double gamma, int *err)
{
  if (!P || !U || !err) return MARAGPU_E_INVALID;
  if (n == 0) return MARAGPU_OK;
  DevBuf dP, dU, dE;
  if (dalloc(dP, n * 64) || dalloc(dU, n * 64) || dalloc(dE, n * 4)) return MARAGPU_E_NOMEM;
  HIP_CHECK(hipMemcpy(dP.p, P, n * 64, hipMemcpyHostToDevice));
  kb_prim_to_cons<<<nblocks(n), kBlock>>>((long)n, dP.get<double>(), dU.get<double>(), gamma, dE.get<int>());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpy(U, dU.p, n * 64, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(err, dE.p, n * 4, hipMemcpyDeviceToHost));
  return MARAGPU_OK;
}

int maragpu_batch_cons_to_prim(size_t n, const double *U, double *P, double gamma, double pfloor,
                               int *err, int *iters)
{
  if (!P || !U || !err || !iters) return MARAGPU_E_INVALID;
  if (n == 0) return MARAGPU_OK;
  DevBuf dP, dU, dE, dI;
  if (dalloc(dP, n * 64) || dalloc(dU, n * 64) || dalloc(dE, n * 4) || dalloc(dI, n * 4)) return MARAGPU_E_NOMEM;
  HIP_CHECK(hipMemcpy(dP.p, P, n * 64, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(dU.p, U, n * 64, hipMemcpyHostToDevice));
  kb_cons_to_prim<<<nblocks(n), kBlock>>>((long)n, dU.get<double>(), dP.get<double>(), gamma, pfloor,
                                          dE.get<int>(), dI.get<int>());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpy(P, dP.p, n * 64, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(err, dE.p, n * 4, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(iters, dI.p, n * 4, hipMemcpyDeviceToHost));
  return MARAGPU_OK;
}

int maragpu_batch_flux_eigen(size_t n, const double *P, double *F, double *ap, double *am, int dim,
                             double gamma)
{
  if (!P || !F || !ap || !am || dim < 1 || dim > 3) return MARAGPU_E_INVALID;
  if (n == 0) return MARAGPU_OK;
  DevBuf dP, dF, dA, dB;
  if (dalloc(dP, n * 64) || dalloc(dF, n * 64) || dalloc(dA, n * 8) || dalloc(dB, n * 8)) return MARAGPU_E_NOMEM;
  HIP_CHECK(hipMemcpy(dP.p, P, n * 64, hipMemcpyHostToDevice));
  auto *p = dP.get<double>(); auto *f = dF.get<double>(); auto *a = dA.get<double>(); auto *b = dB.get<double>();
  if (dim == 1) kb_flux_eigen<1><<<nblocks(n), kBlock>>>((long)n, p, f, a, b, gamma);
  if (dim == 2) kb_flux_eigen<2><<<nblocks(n), kBlock>>>((long)n, p, f, a, b, gamma);
  if (dim == 3) kb_flux_eigen<3><<<nblocks(n), kBlock>>>((long)n, p, f, a, b, gamma);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpy(F, dF.p, n * 64, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(ap, dA.p, n * 8, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(am, dB.p, n * 8, hipMemcpyDeviceToHost));
  return MARAGPU_OK;
}

int maragpu_ou_field_sample(const maragpu_ou_field *field, size_t n, const double *xyz, double *F)
{
  if (!field || !xyz || !F) return MARAGPU_E_INVALID;
  const maragpu_host::OuField &f = *maragpu_ou_field_modes(field);
  if (f.k1 > kOuMaxK1) return MARAGPU_E_INVALID;
  if (n == 0) return MARAGPU_OK;
  std::vector<double2> ha(3 * (size_t)f.numk);
  for (size_t q = 0; q < ha.size(); ++q) ha[q] = make_double2(f.F[2 * q], f.F[2 * q + 1]);
  DevBuf dX, dF, dA;
  if (dalloc(dX, n * 24) || dalloc(dF, n * 24) || dalloc(dA, ha.size() * sizeof(double2))) return MARAGPU_E_NOMEM;
  HIP_CHECK(hipMemcpy(dX.p, xyz, n * 24, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(dA.p, ha.data(), ha.size() * sizeof(double2), hipMemcpyHostToDevice));
  kb_ou_sample<<<nblocks(n), kBlock>>>((long)n, dX.get<double>(), dA.get<double2>(), f.k1, dF.get<double>());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpy(F, dF.p, n * 24, hipMemcpyDeviceToHost));
  return MARAGPU_OK;
}

int maragpu_batch_eigensystem(size_t n, const double *P, double *L, double *R, double *lam, int *ok,
                              int dim, double gamma)
{
  if (!P || !L || !R || !lam || !ok || dim < 1 || dim > 3) return MARAGPU_E_INVALID;
  if (n == 0) return MARAGPU_OK;
  DevBuf dP, dL, dR, dV, dK;
  if (dalloc(dP, n * 64) || dalloc(dL, n * 512) || dalloc(dR, n * 512) || dalloc(dV, n * 64) ||
      dalloc(dK, n * 4))
    return MARAGPU_E_NOMEM;
  HIP_CHECK(hipMemcpy(dP.p, P, n * 64, hipMemcpyHostToDevice));
  auto *p = dP.get<double>(); auto *l = dL.get<double>(); auto *r = dR.get<double>();
  auto *v = dV.get<double>(); auto *k = dK.get<int>();
  if (dim == 1) kb_eigensystem<1><<<nblocks(8 * n), kBlock>>>((long)n, p, l, r, v, k, gamma);
  if (dim == 2) kb_eigensystem<2><<<nblocks(8 * n), kBlock>>>((long)n, p, l, r, v, k, gamma);
  if (dim == 3) kb_eigensystem<3><<<nblocks(8 * n), kBlock>>>((long)n, p, l, r, v, k, gamma);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpy(L, dL.p, n * 512, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(R, dR.p, n * 512, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(lam, dV.p, n * 64, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(ok, dK.p, n * 4, hipMemcpyDeviceToHost));
  return MARAGPU_OK;
}

int maragpu_batch_reconstruct(size_t n, const double *V, double *out, int op, double theta, int IS,
                              double sz10A)
{
  if (!V || !out || op < 0 || op > 5) return MARAGPU_E_INVALID;
  if (n == 0) return MARAGPU_OK;
  DevBuf dV, dO;
  if (dalloc(dV, n * 48) || dalloc(dO, n * 8)) return MARAGPU_E_NOMEM;
  HIP_CHECK(hipMemcpy(dV.p, V, n * 48, hipMemcpyHostToDevice));
  kb_reconstruct<<<nblocks(n), kBlock>>>((long)n, dV.get<double>(), dO.get<double>(), op, theta, IS, sz10A);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpy(out, dO.p, n * 8, hipMemcpyDeviceToHost));
  return MARAGPU_OK;
}

int maragpu_batch_cons_to_prim_learned(const maragpu_mlp *netp, size_t n, const double *U, double *P,
                                       double gamma, double pfloor, int *err, int *iters, int *used,
                                       double *start)
{
  if (!netp || !U || !P || !err || !iters || !used) return MARAGPU_E_INVALID;
  if (n == 0) return MARAGPU_OK;
  MlpDev net;
  if (maragpu_mlp_device_params(netp, &net)) return MARAGPU_E_NOMODEL;
  DevBuf dP, dU, dE, dI, dD, dS;
  if (dalloc(dP, n * 64) || dalloc(dU, n * 64) || dalloc(dE, n * 4) || dalloc(dI, n * 4) || dalloc(dD, n * 4) ||
      (start && dalloc(dS, n * 16)))
    return MARAGPU_E_NOMEM;
  HIP_CHECK(hipMemcpy(dP.p, P, n * 64, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(dU.p, U, n * 64, hipMemcpyHostToDevice));
  kb_cons_to_prim_learned<<<nblocks(n), kBlock>>>((long)n, dU.get<double>(), dP.get<double>(), gamma, pfloor,
                                                  dE.get<int>(), dI.get<int>(), dD.get<int>(),
                                                  start ? dS.get<double>() : nullptr, net);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpy(P, dP.p, n * 64, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(err, dE.p, n * 4, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(iters, dI.p, n * 4, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(used, dD.p, n * 4, hipMemcpyDeviceToHost));
  if (start) HIP_CHECK(hipMemcpy(start, dS.p, n * 16, hipMemcpyDeviceToHost));
  return MARAGPU_OK;
}

int maragpu_batch_riemann(size_t n, int solver, int precision, const double *PL, const double *PR,
                          double *F, int dim, double gamma, double pfloor, int *err, double *lam)
{
  if (!PL || !PR || !F || !err || !lam || dim < 1 || dim > 3 || solver < 0 || solver > 2)
    return MARAGPU_E_INVALID;
  if (n == 0) return MARAGPU_OK;
  DevBuf dL, dR, dF, dE, dM;
  if (dalloc(dL, n * 64) || dalloc(dR, n * 64) || dalloc(dF, n * 64) || dalloc(dE, n * 4) ||
      dalloc(dM, n * 8))
    return MARAGPU_E_NOMEM;
  HIP_CHECK(hipMemcpy(dL.p, PL, n * 64, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(dR.p, PR, n * 64, hipMemcpyHostToDevice));
  auto *l = dL.get<double>(); auto *r = dR.get<double>(); auto *f = dF.get<double>();
  auto *e = dE.get<int>(); auto *m = dM.get<double>();
  const unsigned nb = nblocks(n);
  if (precision == MARAGPU_MIXED) {
    if (dim == 1) kb_riemann<float, 1><<<nb, kBlock>>>((long)n, solver, l, r, f, gamma, pfloor, e, m);
    if (dim == 2) kb_riemann<float, 2><<<nb, kBlock>>>((long)n, solver, l, r, f, gamma, pfloor, e, m);
    if (dim == 3) kb_riemann<float, 3><<<nb, kBlock>>>((long)n, solver, l, r, f, gamma, pfloor, e, m);
  }
  else {
    if (dim == 1) kb_riemann<double, 1><<<nb, kBlock>>>((long)n, solver, l, r, f, gamma, pfloor, e, m);
    if (dim == 2) kb_riemann<double, 2><<<nb, kBlock>>>((long)n, solver, l, r, f, gamma, pfloor, e, m);
    if (dim == 3) kb_riemann<double, 3><<<nb, kBlock>>>((long)n, solver, l, r, f, gamma, pfloor, e, m);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpy(F, dF.p, n * 64, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(err, dE.p, n * 4, hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(lam, dM.p, n * 8, hipMemcpyDeviceToHost));
  return MARAGPU_OK;
}

}  // extern "C"
