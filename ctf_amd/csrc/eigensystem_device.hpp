/* Characteristic eigensystem of the RMHD flux Jacobian, FP64, one state per group of 8 lanes.
 *
 * Used by the weno-split sweep to project the split fluxes into characteristic fields
 * (WenoSplit::intercell_flux_sweep, Mara/weno-split.cpp:163-207, which calls
 * FluidEquations::Eigensystem; the reference's RMHD fluid leaves that function empty,
 * Mara/hydro.hpp:146-147). The construction:
 *   M = dU/dP and B = dF/dP by forward-mode differentiation of the state and flux formulas;
 *   A_P = M^-1 B, the Jacobian in primitive variables;
 *   eigenvalues: 0 (the normal field), v_n (entropy), the Alfven pair
 *     (b^n +- e u^n)/(b^0 +- e u^0) with e = sqrt(rho h + b^2), the fast pair by Newton on
 *     K (x-v)^4 + (1-x^2)[L (x-v)^2 + (b^0 x - b^n)^2] from outside the isotropic bound, and
 *     the slow pair from the quartic deflated by the fast pair, polished by Newton;
 *   right eigenvectors of A_P cluster by cluster (eigenvalues closer than 1e-7 share one
 *     null space of the known dimension), by complete-pivot elimination;
 *   R = M R_P with unit max-norm columns, L = R^-1; a state whose L exceeds 1e8 or whose
 *   |L R - I| exceeds 1e-8 is rejected (the caller then projects with the identity).
 * Eight lanes share one state (g8 below): each holds one column of every 8x8 matrix in
 * registers.
 */
#pragma once

#include "rmhd_device.hpp"

namespace maragpu {
namespace eig {

__device__ __forceinline__ void swapd(double &a, double &b) { const double t = a; a = b; b = t; }

/* A value and its derivative along one primitive variable: forward-mode differentiation,
 * every operation applying its derivative rule, so the Jacobians are exact to round-off
 * (a difference quotient would lose half the digits to cancellation, and the eigenvectors
 * would inherit that noise divided by the gaps of the spectrum). No operation is fused
 * into a multiply-add, so the values round as in a host build of the same formulas. */
struct D1 { double v, d; };
__device__ __forceinline__ D1 operator+(D1 a, D1 b)
{
#pragma clang fp contract(off)
  return {a.v + b.v, a.d + b.d};
}
__device__ __forceinline__ D1 operator-(D1 a, D1 b)
{
#pragma clang fp contract(off)
  return {a.v - b.v, a.d - b.d};
}
__device__ __forceinline__ D1 operator*(D1 a, D1 b)
{
#pragma clang fp contract(off)
  return {a.v * b.v, a.d * b.v + a.v * b.d};
}
__device__ __forceinline__ D1 operator/(D1 a, D1 b)
{
#pragma clang fp contract(off)
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ D1 operator+(double s, D1 a)
{
#pragma clang fp contract(off)
  return {s + a.v, a.d};
}
__device__ __forceinline__ D1 operator-(double s, D1 a)
{
#pragma clang fp contract(off)
  return {s - a.v, -a.d};
}
__device__ __forceinline__ D1 operator*(double s, D1 a)
{
#pragma clang fp contract(off)
  return {s * a.v, s * a.d};
}
__device__ __forceinline__ D1 operator/(double s, D1 b)
{
#pragma clang fp contract(off)
  const double q = s / b.v;
  return {q, -(q * b.d) / b.v};
}
__device__ __forceinline__ D1 operator/(D1 a, double s) { return {a.v / s, a.d / s}; }
__device__ __forceinline__ D1 dsqrt1(D1 a)
{
  const double r = sqrt(a.v);
  return {r, a.d / (2.0 * r)};
}

/* U and the flux F along DIM of the primitive state P, with derivatives: Rmhd::PrimToCons
 * and the flux part of Rmhd::FluxAndEigenvalues (Mara/rmhd.cpp:133-163, :164-199) in the
 * reference's operation order, with its divisions */
template <int DIM>
__device__ __forceinline__ void state_fluxes(const D1 *P, D1 *U, D1 *F, double gamma)
{
#pragma clang fp contract(off)
  constexpr int VN = DIM == 1 ? vx : DIM == 2 ? vy : vz;
  constexpr int BN = DIM == 1 ? Bx : DIM == 2 ? By : Bz;
  const D1 V2 = P[vx] * P[vx] + P[vy] * P[vy] + P[vz] * P[vz];
  const D1 B2 = P[Bx] * P[Bx] + P[By] * P[By] + P[Bz] * P[Bz];
  const D1 Bv = P[Bx] * P[vx] + P[By] * P[vy] + P[Bz] * P[vz];
  const D1 W2 = 1.0 / (1.0 - V2);
  const D1 W = dsqrt1(W2);
  const D1 b0 = W * Bv;
  const D1 b2 = (B2 + b0 * b0) / W2;
  const D1 bv[3] = {(P[Bx] + b0 * W * P[vx]) / W, (P[By] + b0 * W * P[vy]) / W,
                    (P[Bz] + b0 * W * P[vz]) / W};
  const D1 e = (P[pre] / (gamma - 1)) / P[rho];
  const D1 e_ = e + 0.5 * b2 / P[rho];
  const D1 p_ = P[pre] + 0.5 * b2;
  const D1 h_ = 1.0 + e_ + p_ / P[rho];
  U[ddd] = P[rho] * W;
  U[tau] = P[rho] * h_ * W2 - p_ - b0 * b0 - U[ddd];
#pragma unroll
  for (int a = 0; a < 3; ++a) U[Sx + a] = P[rho] * h_ * W2 * P[vx + a] - b0 * bv[a];
#pragma unroll
  for (int a = 0; a < 3; ++a) U[Bx + a] = P[Bx + a];
  F[ddd] = U[ddd] * P[VN];
  F[tau] = U[tau] * P[VN] - b0 * P[BN] / W + p_ * P[VN];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    F[Sx + a] = U[Sx + a] * P[VN] - bv[a] * P[BN] / W;
    if (a == DIM - 1) F[Sx + a] = F[Sx + a] + p_;
    F[Bx + a] = a == DIM - 1 ? D1{0.0, 0.0} : P[VN] * P[Bx + a] - P[vx + a] * P[BN];
  }
}

/* Newton on the factored characteristic quartic from x; from outside the roots it moves
 * monotonically onto the extreme one, from a deflation estimate it polishes */
__device__ inline double quartic_newton(double K, double L, double b0, double bi, double v, double x)
{
#pragma clang fp contract(off)
  for (int it = 0; it < 60; ++it) {
    const double y = x - v, y2 = y * y, c = b0 * x - bi, q1 = L * y2 + c * c, om = (1.0 - x) * (1.0 + x);
    const double Q = K * y2 * y2 + om * q1;
    const double dQ = 4.0 * K * y2 * y - 2.0 * x * q1 + om * (2.0 * L * y + 2.0 * b0 * c);
    const double step = Q / dQ;
    const double xn = x - step;
    if (!(fabs(xn) <= 1.0)) break;
    x = xn;
    if (!(fabs(step) > 1e-15)) break;
  }
  return x;
}

/* --------------------------------------------------------------------------
 * The eigensystem, computed by 8 lanes together (an aligned group of 8 lanes of a
 * wave, one state per group). Lane k of the group holds column k of every 8x8 matrix, so a
 * lane keeps 8 doubles per matrix in registers instead of 64 in private memory:
 *   column k of M = dU/dP and of B = dF/dP by one forward-mode pass along P[k];
 *   Gauss-Jordan elimination with the pivot column broadcast to the group; each lane
 *     updates its own column;
 *   the eigenvectors lane-parallel: lane i eliminates its own full copy of A_P - lambda_i I
 *     with complete pivoting (every lane of a cluster forms the same null space and keeps
 *     its own vector of it).
 * Every branch that encloses a group exchange depends only on values that all lanes of the
 * group hold alike, so the group is converged at each exchange (the lane-local null vector
 * has no exchange inside). The eigenvalues are computed redundantly by each lane. Products and sums
 * run in the order of a serial row-major implementation.
 * Outputs are row-wise: lane r returns row r of L and of R.
 * ------------------------------------------------------------------------ */
namespace g8 {

/* any lane of the group to all of it (ds_bpermute) */
__device__ __forceinline__ double bc(double v, int src) { return __shfl(v, src, 8); }

/* Partner exchanges for group reductions, as DPP moves (a few cycles instead of an LDS
 * round trip): lane i and i^1, lane i and i^2 (quad permutations), then lane i and 7-i
 * (half-row mirror), which pairs the two quads of the group. Each is an involution, so both
 * partners combine the same two values. */
constexpr int kDppXor1 = 0xB1;       // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;       // quad_perm [2, 3, 0, 1]
constexpr int kDppMirror = 0x141;    // row_half_mirror
template <int CTRL> __device__ __forceinline__ int xchg(int v)
{
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ double xchg(double v)
{
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)xchg<CTRL>((int)(unsigned)(b & 0xffffffffll));
  const unsigned hi = (unsigned)xchg<CTRL>((int)(unsigned)((unsigned long long)b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double vmax(double v)
{
  v = fmax(v, xchg<kDppXor1>(v));
  v = fmax(v, xchg<kDppXor2>(v));
  return fmax(v, xchg<kDppMirror>(v));
}
__device__ __forceinline__ int imax(int v)
{
  v = max(v, xchg<kDppXor1>(v));
  v = max(v, xchg<kDppXor2>(v));
  return max(v, xchg<kDppMirror>(v));
}
/* a value the optimizer cannot trace back to the array element it came from: without it,
 * the select chains below are folded back into an indexed load or store, which puts the
 * array in private memory */
__device__ __forceinline__ double opaque(double x)
{
  asm volatile("" : "+v"(x));
  return x;
}

/* a[i] for a run-time i, by selects (no private-memory array) */
__device__ __forceinline__ double pick(const double (&a)[8], int i)
{
  double v = opaque(a[0]);
#pragma unroll
  for (int k = 1; k < 8; ++k) v = (i == k) ? opaque(a[k]) : v;
  return v;
}
__device__ __forceinline__ void put(double (&a)[8], int i, double v)
{
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const double cur = opaque(a[k]);
    a[k] = (i == k) ? v : cur;
  }
}
__device__ __forceinline__ void swap_rows(double (&a)[8], int i, int j)
{
  const double ai = pick(a, i), aj = pick(a, j);
  put(a, i, aj);
  put(a, j, ai);
}

/* lane r receives row r of the matrix whose column k lane k holds */
__device__ __forceinline__ void transpose(const double (&col)[8], double (&row)[8], int lane)
{
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const double send = pick(col, (lane - s) & 7);      // the element the lane (lane - s) & 7 wants
    const double got = bc(send, (lane + s) & 7);
    put(row, (lane + s) & 7, got);
  }
}

/* A X = B, both held column-wise (lane k: column k), X into B: Gauss-Jordan elimination
 * with partial pivoting. Returns 1 (in every lane) on a zero pivot. */
__device__ __forceinline__ int solve(double (&A)[8], double (&B)[8], int lane)
{
#pragma clang fp contract(off)
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    double col[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) col[r] = bc(A[r], c);
    int piv = c;
    double best = fabs(col[c]);
#pragma unroll
    for (int r = c + 1; r < 8; ++r) {
      const double a = fabs(col[r]);
      if (a > best) { best = a; piv = r; }
    }
    if (!(best > 0.0)) return 1;
    swap_rows(A, c, piv);
    swap_rows(B, c, piv);
    swap_rows(col, c, piv);
    const double id = 1.0 / col[c];
    const double ac = A[c], bcv = B[c];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      if (r == c) continue;
      const double f = col[r] * id;
      if (f == 0.0) continue;
      if (lane >= c) A[r] -= f * ac;
      B[r] -= f * bcv;
    }
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const double id = 1.0 / bc(A[r], r);
    B[r] *= id;
  }
  return 0;
}

/* Step K of the complete-pivot elimination in one lane (K a compile-time constant, so every
 * index of the 8x8 array is one and the array stays in registers): the first largest |a|
 * of the trailing block in row-major order, rows K <-> pr and columns K <-> pc swapped by
 * selects, colp's nibbles K and pc exchanged, the rows below eliminated. Returns false on a
 * zero pivot. */
template <int K>
__device__ __forceinline__ bool elim_step(double (&a)[8][8], unsigned &colp)
{
#pragma clang fp contract(off)
  double best = -1.0;
  int pr = K, pc = K;
#pragma unroll
  for (int r = K; r < 8; ++r)
#pragma unroll
    for (int c = K; c < 8; ++c) {
      const double v = fabs(a[r][c]);
      if (v > best) { best = v; pr = r; pc = c; }
    }
  if (!(best > 0.0)) return false;
#pragma unroll
  for (int c = 0; c < 8; ++c) {                 // rows K <-> pr
    const double ak = opaque(a[K][c]);
    double ap = ak;
#pragma unroll
    for (int r = K + 1; r < 8; ++r) ap = (pr == r) ? opaque(a[r][c]) : ap;
#pragma unroll
    for (int r = K + 1; r < 8; ++r) {
      const double cur = opaque(a[r][c]);
      a[r][c] = (pr == r) ? ak : cur;
    }
    a[K][c] = ap;
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {                 // columns K <-> pc
    const double ak = opaque(a[r][K]);
    double ap = ak;
#pragma unroll
    for (int c = K + 1; c < 8; ++c) ap = (pc == c) ? opaque(a[r][c]) : ap;
#pragma unroll
    for (int c = K + 1; c < 8; ++c) {
      const double cur = opaque(a[r][c]);
      a[r][c] = (pc == c) ? ak : cur;
    }
    a[r][K] = ap;
  }
  const unsigned nk = (colp >> (4 * K)) & 15u, np = (colp >> (4 * pc)) & 15u;
  colp = (colp & ~((15u << (4 * K)) | (15u << (4 * pc)))) | (np << (4 * K)) | (nk << (4 * pc));
  const double piv = a[K][K];
#pragma unroll
  for (int r = K + 1; r < 8; ++r) {
    const double f = a[r][K] / piv;
#pragma unroll
    for (int c = K; c < 8; ++c) a[r][c] -= f * a[K][c];
  }
  return true;
}

/* back substitution row K (compile-time), when K < rank */
template <int K>
__device__ __forceinline__ void back_step(const double (&a)[8][8], double (&xl)[8], int rank)
{
#pragma clang fp contract(off)
  if (K < rank) {
    double s = 0.0;
#pragma unroll
    for (int c = K + 1; c < 8; ++c) s += a[K][c] * xl[c];
    xl[K] = -s / a[K][K];
  }
}

/* Null vector number fc of (a - mu I), whose null space has dimension want, in one lane:
 * a is a full 8x8 copy of the matrix in the lane's registers (destroyed). 8 - want
 * complete-pivot elimination steps, then back substitution with free column rank + fc set
 * to one and the other free columns to zero. x receives the vector in the original column
 * order. Returns 1 on a zero pivot. */
__device__ __forceinline__ int null_vector(double (&a)[8][8], double mu, int want, int fc, double (&x)[8])
{
#pragma clang fp contract(off)
#pragma unroll
  for (int r = 0; r < 8; ++r) a[r][r] = a[r][r] - mu;
  const int rank = 8 - want;
  unsigned colp = 0x76543210u;          // nibble c: the original column now at position c
  bool ok = true;
  if (ok && 0 < rank) ok = elim_step<0>(a, colp);
  if (ok && 1 < rank) ok = elim_step<1>(a, colp);
  if (ok && 2 < rank) ok = elim_step<2>(a, colp);
  if (ok && 3 < rank) ok = elim_step<3>(a, colp);
  if (ok && 4 < rank) ok = elim_step<4>(a, colp);
  if (ok && 5 < rank) ok = elim_step<5>(a, colp);
  if (ok && 6 < rank) ok = elim_step<6>(a, colp);
  if (!ok) return 1;
  double xl[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  put(xl, rank + fc, 1.0);
  back_step<6>(a, xl, rank);
  back_step<5>(a, xl, rank);
  back_step<4>(a, xl, rank);
  back_step<3>(a, xl, rank);
  back_step<2>(a, xl, rank);
  back_step<1>(a, xl, rank);
  back_step<0>(a, xl, rank);
#pragma unroll
  for (int c = 0; c < 8; ++c) x[c] = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) put(x, (int)((colp >> (4 * c)) & 15u), xl[c]);
  return 0;
}

/* L, R and the eigenvalues lam[8] (ascending) of the state P for the flux along DIM, for one
 * state per group of 8 lanes; lane r returns Lrow = L[r][:], Rrow = R[r][:] and all of
 * lam. Returns 0 when usable, 1 on a singular Jacobian, 2 when a cluster's null space could
 * not be formed, 3 when the eigenvectors are too ill-conditioned; alike in every lane. */
template <int DIM>
__device__ __forceinline__ int rmhd_eigensystem(const double *P, int lane, double (&Lrow)[8], double (&Rrow)[8], double (&lam)[8],
                                double gamma)
{
#pragma clang fp contract(off)
  constexpr int VN = DIM == 1 ? vx : DIM == 2 ? vy : vz;
  constexpr int BN = DIM == 1 ? Bx : DIM == 2 ? By : Bz;
  double Mc[8], Xc[8];
  {
    D1 Pd[8], Ud[8], Fd[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) Pd[i] = D1{P[i], i == lane ? 1.0 : 0.0};
    state_fluxes<DIM>(Pd, Ud, Fd, gamma);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      Mc[i] = Ud[i].d;
      Xc[i] = Fd[i].d;
    }
  }
  {
    double Wc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) Wc[i] = Mc[i];
    if (solve(Wc, Xc, lane)) return 1;           // Xc: column of A_P = M^-1 B
  }

  double ev[8];
  {
    const double V2 = P[vx] * P[vx] + P[vy] * P[vy] + P[vz] * P[vz];
    const double Bv = P[Bx] * P[vx] + P[By] * P[vy] + P[Bz] * P[vz];
    const double B2 = P[Bx] * P[Bx] + P[By] * P[By] + P[Bz] * P[Bz];
    const double Wl = 1.0 / sqrt(1.0 - V2);
    const double b0 = Wl * Bv;
    const double b2 = (B2 + b0 * b0) / (Wl * Wl);
    const double bi = (P[BN] + b0 * Wl * P[VN]) / Wl;
    const double v = P[VN];
    const double rhoh = P[rho] + P[pre] / (gamma - 1.0) + P[pre];
    const double cs2 = gamma * P[pre] / rhoh;
    const double K = rhoh * (1.0 / cs2 - 1.0) * Wl * Wl * Wl * Wl;
    const double Lq = -(rhoh + b2 / cs2) * Wl * Wl;
    ev[0] = 0.0;
    ev[1] = v;
    const double e = sqrt(rhoh + b2);
    ev[2] = (bi + e * Wl * v) / (b0 + e * Wl);
    ev[3] = (bi - e * Wl * v) / (b0 - e * Wl);
    const double ca2 = b2 / (rhoh + b2);
    const double w2 = cs2 + ca2 - cs2 * ca2;
    const double den = 1.0 / (1.0 - V2 * w2);
    const double rt = sqrt(w2 * (1.0 - V2) * (1.0 - V2 * w2 - v * v * (1.0 - w2)));
    const double lp = fmin((v * (1.0 - w2) + rt) * den + 1e-12, 1.0);
    const double lm = fmax((v * (1.0 - w2) - rt) * den - 1e-12, -1.0);
    // the two fast roots at once: even lanes from lp, odd lanes from lm (one Newton loop for
    // the group instead of two; each root is the same as alone)
    const double rr = quartic_newton(K, Lq, b0, bi, v, (lane & 1) ? lm : lp);
    const double r1 = bc(rr, 0), r2 = bc(rr, 1);
    const double A4 = K - Lq - b0 * b0;
    const double A3 = -4 * K * v + Lq * v * 2 + 2 * b0 * bi;
    const double A2 = 6 * K * v * v + Lq * (1.0 - v * v) + b0 * b0 - bi * bi;
    const double s = r1 + r2, pr = r1 * r2;
    const double c1 = A3 + A4 * s;
    const double c0 = A2 + c1 * s - A4 * pr;
    double disc = c1 * c1 - 4.0 * A4 * c0;
    if (disc < 0.0) disc = 0.0;
    const double sq = sqrt(disc);
    ev[4] = r1;
    ev[5] = r2;
    const double rs = quartic_newton(K, Lq, b0, bi, v, (lane & 1) ? (-c1 - sq) / (2.0 * A4) : (-c1 + sq) / (2.0 * A4));
    ev[6] = bc(rs, 0);
    ev[7] = bc(rs, 1);
  }
  // ascending, by an odd-even transposition network (fixed indices: registers only)
#pragma unroll
  for (int pass = 0; pass < 8; ++pass)
#pragma unroll
    for (int i = pass & 1; i + 1 < 8; i += 2)
      if (ev[i + 1] < ev[i]) swapd(ev[i], ev[i + 1]);

  // clusters: maximal runs of eigenvalues closer than 1e-7. Lane i forms eigenvector i, as
  // vector (i - first) of its cluster's null space, from its own copy of A_P
  double mean[8];
  int first = 0, last = 8;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int s0 = 0, s1 = 8;
#pragma unroll
    for (int j = 1; j < 8; ++j) {
      const bool cut = !(ev[j] - ev[j - 1] < 1e-7);
      if (cut && j <= k) s0 = j;
      if (cut && j > k && s1 == 8) s1 = j;
    }
    double m = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q >= s0 && q < s1) m += ev[q];
    mean[k] = m / (s1 - s0);
    if (k == lane) { first = s0; last = s1; }
  }
  double RPc[8];                                 // column `lane` of R_P
  {
    double a[8][8];
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
      for (int r = 0; r < 8; ++r) a[r][c] = bc(Xc[r], c);
    const int bad = null_vector(a, pick(mean, lane), last - first, lane - first, RPc);
    if (imax(bad)) return 2;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) ev[k] = mean[k];
  // R = M R_P, column-wise: every column of M is broadcast
  double Rc[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) Rc[r] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
#pragma unroll
    for (int r = 0; r < 8; ++r) Rc[r] += bc(Mc[r], k) * RPc[k];
  }
  {
    double m = 0.0;
#pragma unroll
    for (int r = 0; r < 8; ++r) m = fmax(m, fabs(Rc[r]));
    if (imax(!(m > 0.0) ? 1 : 0)) return 1;
#pragma unroll
    for (int r = 0; r < 8; ++r) Rc[r] /= m;
  }
  double Lc[8];
  {
    double Wc[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) { Wc[r] = Rc[r]; Lc[r] = (r == lane) ? 1.0 : 0.0; }
    if (solve(Wc, Lc, lane)) return 1;
  }
  double lmax = 0.0, err = 0.0;
  {
    double LR[8] = {0, 0, 0, 0, 0, 0, 0, 0};     // column `lane` of L R
#pragma unroll
    for (int k = 0; k < 8; ++k) {
#pragma unroll
      for (int r = 0; r < 8; ++r) LR[r] += bc(Lc[r], k) * Rc[k];
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      lmax = fmax(lmax, fabs(Lc[r]));
      err = fmax(err, fabs(LR[r] - (r == lane ? 1.0 : 0.0)));
    }
    lmax = vmax(lmax);
    err = vmax(err);
  }
  if (!(lmax < 1e8) || !(err < 1e-8)) return 3;
  transpose(Lc, Lrow, lane);
  transpose(Rc, Rrow, lane);
#pragma unroll
  for (int k = 0; k < 8; ++k) lam[k] = ev[k];
  return 0;
}

/* rows of the identity projection */
__device__ __forceinline__ void identity_rows(int lane, double (&Lrow)[8], double (&Rrow)[8], double (&lam)[8])
{
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    lam[k] = 0.0;
    Lrow[k] = Rrow[k] = (k == lane) ? 1.0 : 0.0;
  }
}

}  // namespace g8

}  // namespace eig
}  // namespace maragpu
