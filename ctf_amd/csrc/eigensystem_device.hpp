/* Characteristic eigensystem of the RMHD flux Jacobian, FP64, one state per lane.
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
 * Matrices are 8x8 row-major in the lane's private memory; every loop bound is fixed.
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

/* A X = Bm for 8x8 A and 8 right-hand sides (both row-major, overwritten, X in Bm):
 * Gauss-Jordan elimination with partial pivoting. Returns 1 on a zero pivot. */
__device__ inline int solve8(double *A, double *Bm)
{
#pragma clang fp contract(off)
  for (int c = 0; c < 8; ++c) {
    int piv = c;
    double best = fabs(A[c * 8 + c]);
    for (int r = c + 1; r < 8; ++r) {
      const double a = fabs(A[r * 8 + c]);
      if (a > best) { best = a; piv = r; }
    }
    if (!(best > 0.0)) return 1;
    if (piv != c)
      for (int k = 0; k < 8; ++k) { swapd(A[c * 8 + k], A[piv * 8 + k]); swapd(Bm[c * 8 + k], Bm[piv * 8 + k]); }
    const double id = 1.0 / A[c * 8 + c];
    for (int r = 0; r < 8; ++r) {
      if (r == c) continue;
      const double f = A[r * 8 + c] * id;
      if (f == 0.0) continue;
      for (int k = c; k < 8; ++k) A[r * 8 + k] -= f * A[c * 8 + k];
      for (int k = 0; k < 8; ++k) Bm[r * 8 + k] -= f * Bm[c * 8 + k];
    }
  }
  for (int r = 0; r < 8; ++r) {
    const double id = 1.0 / A[r * 8 + r];
    for (int k = 0; k < 8; ++k) Bm[r * 8 + k] *= id;
  }
  return 0;
}

/* `want` null vectors of the 8x8 matrix A (destroyed), whose null space has that
 * dimension: 8 - want complete-pivot elimination steps, then back substitution with each
 * free column set to one in turn, into columns col0.. of V. Returns 1 on a zero pivot. */
__device__ inline int null_space8(double *A, double *V, int col0, int want)
{
#pragma clang fp contract(off)
  int rowp[8], colp[8];
  for (int i = 0; i < 8; ++i) { rowp[i] = i; colp[i] = i; }
  const int rank = 8 - want;
  for (int k = 0; k < rank; ++k) {
    int pr = k, pc = k;
    double best = -1.0;
    for (int r = k; r < 8; ++r)
      for (int c = k; c < 8; ++c) {
        const double a = fabs(A[rowp[r] * 8 + colp[c]]);
        if (a > best) { best = a; pr = r; pc = c; }
      }
    if (!(best > 0.0)) return 1;
    { const int t = rowp[k]; rowp[k] = rowp[pr]; rowp[pr] = t; }
    { const int t = colp[k]; colp[k] = colp[pc]; colp[pc] = t; }
    const double piv = A[rowp[k] * 8 + colp[k]];
    for (int r = k + 1; r < 8; ++r) {
      const double f = A[rowp[r] * 8 + colp[k]] / piv;
      for (int c = k; c < 8; ++c) A[rowp[r] * 8 + colp[c]] -= f * A[rowp[k] * 8 + colp[c]];
    }
  }
  for (int fc = 0; fc < want; ++fc) {
    double x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    x[colp[rank + fc]] = 1.0;
    for (int k = rank - 1; k >= 0; --k) {
      double s = 0.0;
      for (int c = k + 1; c < 8; ++c) s += A[rowp[k] * 8 + colp[c]] * x[colp[c]];
      x[colp[k]] = -s / A[rowp[k] * 8 + colp[k]];
    }
    for (int i = 0; i < 8; ++i) V[i * 8 + col0 + fc] = x[i];
  }
  return 0;
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

/* L, R (64 doubles each, row-major) and the eigenvalues lam[8] (ascending) of the state P
 * for the flux along DIM. Returns 0 when usable, 1 on a singular Jacobian, 2 when a
 * cluster's null space could not be formed, 3 when the eigenvectors are too
 * ill-conditioned. */
template <int DIM>
__device__ int rmhd_eigensystem(const double *P, double *Lm, double *Rm, double *lam, double gamma)
{
#pragma clang fp contract(off)
  constexpr int VN = DIM == 1 ? vx : DIM == 2 ? vy : vz;
  constexpr int BN = DIM == 1 ? Bx : DIM == 2 ? By : Bz;
  double M[64], X[64], W[64], RP[64];
  // M = dU/dP into M, B = dF/dP into X, column k along P[k]
  for (int k = 0; k < 8; ++k) {
    D1 Pd[8], Ud[8], Fd[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) Pd[i] = D1{P[i], i == k ? 1.0 : 0.0};
    state_fluxes<DIM>(Pd, Ud, Fd, gamma);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      M[i * 8 + k] = Ud[i].d;
      X[i * 8 + k] = Fd[i].d;
    }
  }
  for (int i = 0; i < 64; ++i) W[i] = M[i];
  if (solve8(W, X)) return 1;                    // X = A_P = M^-1 B

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
    const double r1 = quartic_newton(K, Lq, b0, bi, v, lp);
    const double r2 = quartic_newton(K, Lq, b0, bi, v, lm);
    // the quartic's coefficients A4 x^4 + A3 x^3 + A2 x^2 + ..., deflated by (x-r1)(x-r2)
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
    ev[6] = quartic_newton(K, Lq, b0, bi, v, (-c1 + sq) / (2.0 * A4));
    ev[7] = quartic_newton(K, Lq, b0, bi, v, (-c1 - sq) / (2.0 * A4));
  }
  for (int i = 1; i < 8; ++i)                    // insertion sort, ascending
    for (int j = i; j > 0 && ev[j] < ev[j - 1]; --j) swapd(ev[j], ev[j - 1]);

  for (int i = 0; i < 8;) {
    int j = i + 1;
    while (j < 8 && ev[j] - ev[j - 1] < 1e-7) ++j;
    double mean = 0.0;
    for (int k = i; k < j; ++k) mean += ev[k];
    mean /= (j - i);
    for (int r = 0; r < 8; ++r)
      for (int c = 0; c < 8; ++c) W[r * 8 + c] = X[r * 8 + c] - (r == c ? mean : 0.0);
    if (null_space8(W, RP, i, j - i)) return 2;
    for (int k = i; k < j; ++k) ev[k] = mean;
    i = j;
  }
  // R = M R_P with unit max-norm columns
  for (int r = 0; r < 8; ++r)
    for (int c = 0; c < 8; ++c) {
      double s = 0.0;
      for (int k = 0; k < 8; ++k) s += M[r * 8 + k] * RP[k * 8 + c];
      Rm[r * 8 + c] = s;
    }
  for (int c = 0; c < 8; ++c) {
    double m = 0.0;
    for (int r = 0; r < 8; ++r) m = fmax(m, fabs(Rm[r * 8 + c]));
    if (!(m > 0.0)) return 1;
    for (int r = 0; r < 8; ++r) Rm[r * 8 + c] /= m;
  }
  for (int i = 0; i < 64; ++i) { W[i] = Rm[i]; Lm[i] = (i / 8 == i % 8) ? 1.0 : 0.0; }
  if (solve8(W, Lm)) return 1;
  double lmax = 0.0, err = 0.0;
  for (int r = 0; r < 8; ++r)
    for (int c = 0; c < 8; ++c) {
      lmax = fmax(lmax, fabs(Lm[r * 8 + c]));
      double s = 0.0;
      for (int k = 0; k < 8; ++k) s += Lm[r * 8 + k] * Rm[k * 8 + c];
      err = fmax(err, fabs(s - (r == c ? 1.0 : 0.0)));
    }
  if (!(lmax < 1e8) || !(err < 1e-8)) return 3;
  for (int k = 0; k < 8; ++k) lam[k] = ev[k];
  return 0;
}

/* the projection used where the eigensystem is rejected: the identity */
__device__ inline void identity_eigensystem(double *Lm, double *Rm, double *lam)
{
  for (int i = 0; i < 8; ++i) {
    lam[i] = 0.0;
    for (int j = 0; j < 8; ++j) Lm[i * 8 + j] = Rm[i * 8 + j] = (i == j) ? 1.0 : 0.0;
  }
}

}  // namespace eig
}  // namespace maragpu
