/* Device-side relativistic-MHD point physics for gfx950.
 *
 * Everything here runs per zone or per interface inside the kernels of maragpu.hip.
 * The arithmetic follows Mara's plain method (file:line cited per routine); the control
 * flow is restated for the GPU: exceptions of the reference become status returns, the
 * axis permutation is a compile-time template parameter (no dynamically indexed local
 * arrays, so nothing spills to scratch), and intermediate states that the s = 0 flux
 * does not use are never formed.
 *
 * The scalar type R of reconstruction and Riemann fluxes is a template parameter:
 * double for the plain FP64 form, float for the mixed form. Conserved-to-primitive
 * recovery (per zone and nested inside HLLD) is always FP64.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

namespace maragpu {

enum { ddd, tau, Sx, Sy, Sz, Bx, By, Bz };  // conserved, Mara/rmhd.hpp:19
enum { rho, pre, vx, vy, vz };              // primitive,  Mara/rmhd.hpp:20

// Mara/rmhd-c2p.h:10-20
enum { C2P_SUCCESS, C2P_CONS_CONTAINS_NAN, C2P_CONS_NEGATIVE_DENSITY, C2P_CONS_NEGATIVE_ENERGY,
       C2P_PRIM_CONTAINS_NAN, C2P_PRIM_NEGATIVE_PRESSURE, C2P_PRIM_NEGATIVE_RESTMASS,
       C2P_PRIM_SUPERLUMINAL, C2P_MAXITER };

enum { RECON_PCM = 0, RECON_PLM = 1, RECON_WENO5 = 2 };
enum { IS_JS96 = 0, IS_B08 = 1, IS_SZ10 = 2 };
enum { RIEMANN_HLL = 0, RIEMANN_HLLC = 1, RIEMANN_HLLD = 2 };

/* Thresholds of the Riemann solvers. FP64 uses the reference's constants verbatim;
 * FP32 scales the ones that sit below single-precision resolution. */
template <typename R> struct Tol;
template <> struct Tol<double> {
  // the c2p of the HLL state tries anton2dzw from the average of the two states (the
  // cascade's second guess) before the cascade: the same root to round-off, and the same
  // success set, since the full cascade runs whenever that attempt fails
  static constexpr bool nested_c2p_from_average = true;
  static constexpr bool nested_c2p_fp32 = false;  // that attempt in FP64
  static constexpr double newton_ftol = 1e-12;   // NewtonRaphesonSolver(15, 1e-12), riemann_hlld-rmhd.cpp:220
  static constexpr int newton_fast_max_it = 15;  // the common path's p* solve: the reference's
  static constexpr double fd_dp = 1e-8;          // HlldEquation48::Jacobian, :313
  static constexpr bool fd_relative = false;
  static constexpr bool fd_one_sided = false;    // the reference's centred difference
  static constexpr double newton_accept = 0.0;   // always at least one step, as the reference
  static constexpr double X_floor = 1e-8;        // :328,346
  static constexpr double small_delK = 1e-10;    // :379,404-405
  static constexpr double small_Bc = 1e-10;      // :408
  static constexpr double small_denom_c = 1e-10; // :428-429
  static constexpr double tol54 = 1e-14;         // :434
  static constexpr double contact = 1e-8;        // :448
  static constexpr double small_denom = 1e-8;    // :456,471,487,502
  static constexpr double small_bn = 1e-8;       // :230
  static constexpr double hllc_small_bx = 1e-12; // riemann_hllc-rmhd.cpp:22
  static constexpr double hllc_small_a = 1e-10;  // riemann_hllc-rmhd.cpp:23
};
template <> struct Tol<float> {
  static constexpr bool nested_c2p_from_average = true;
  static constexpr bool nested_c2p_fp32 = true;   // that attempt in FP32 (hll_c2p_f32)
  static constexpr float newton_ftol = 2e-6f;
  // on the sweep's common path a p* solve that has not converged in 5 steps is handed to
  // the FP64 redo: where FP32 cannot reach the tolerance (strong normal fields) the full 15
  // steps held the whole wave (c5 512^3: 142 -> 136 ms per mixed step, L1 5.2e-7 -> 5.0e-7).
  // The full FP32 solver keeps the reference's 15.
  static constexpr int newton_fast_max_it = 5;
  static constexpr float fd_dp = 1e-4f;          // difference step, relative to max(|p|, 1e-3)
  static constexpr bool fd_relative = true;
  // the derivative by a one-sided difference (one more residual evaluation, not two)
  static constexpr bool fd_one_sided = true;
  // a start whose residual is below this is the root, with no step (20x below the stopping
  // tolerance; at 3e-7 the L1 error from FP64 grows, at 1e-6 it passes 1e-6)
  static constexpr float newton_accept = 1e-7f;
  static constexpr float X_floor = 1e-8f;
  static constexpr float small_delK = 1e-6f;
  static constexpr float small_Bc = 1e-6f;
  static constexpr float small_denom_c = 1e-6f;
  static constexpr float tol54 = 1e-6f;
  static constexpr float contact = 2e-5f;
  static constexpr float small_denom = 1e-6f;
  static constexpr float small_bn = 1e-8f;
  static constexpr float hllc_small_bx = 1e-12f;
  static constexpr float hllc_small_a = 1e-6f;
};

/* Code shape per precision: the FP64 solver is register-bound, so its repeated pieces
 * (the two fast-speed evaluations, the three function evaluations of a Newton step, the
 * c2p attempts) run as loops over one copy of the code; the FP32 solver fits the
 * registers and keeps them unrolled for instruction-level parallelism. */
template <typename R> struct Shape { static constexpr int unroll = sizeof(R) == 8 ? 1 : 8; };

struct PhysParams {
  double gamma;
  double pfloor;
  double theta;
  int IS;
  double sz10A;
};

__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
// FP32: the hardware square root (within 1 ulp; the correctly rounded sqrtf adds about
// fifteen instructions of scaling and correction)
__device__ __forceinline__ float dsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double dfabs(double x) { return fabs(x); }
__device__ __forceinline__ float dfabs(float x) { return fabsf(x); }
__device__ __forceinline__ double dcbrt_pow(double x) { return cbrt(x); }
__device__ __forceinline__ float dcbrt_pow(float x) { return cbrtf(x); }
__device__ __forceinline__ double dacos(double x) { return acos(x); }
__device__ __forceinline__ float dacos(float x) { return acosf(x); }
__device__ __forceinline__ double dcos(double x) { return cos(x); }
__device__ __forceinline__ float dcos(float x) { return cosf(x); }
__device__ __forceinline__ double dsin(double x) { return sin(x); }
__device__ __forceinline__ float dsin(float x) { return sinf(x); }

/* Reciprocals. A correctly rounded FP64 division costs about ten instructions on
 * gfx950 (div_scale x2, rcp, five fma, div_fmas, div_fixup); the hardware reciprocal
 * refined by two Newton steps costs five and is within an ulp. The refinement is not
 * finite where x is 0 or inf (0 x inf); v_div_fixup_f64 then gives the IEEE quotient
 * (1/±0 = ±inf, 1/±inf = ±0, NaN for NaN) and leaves every other result unchanged, in one
 * instruction where a NaN test and select take three (4 % of the FP64 sweep). */
__device__ __forceinline__ double rcp(double x)
{
  const double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  double r = fma(y, e, y);
  e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  return __builtin_amdgcn_div_fixup(r, x, 1.0);
}
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename R> __device__ __forceinline__ bool is_nan(R x) { return x != x; }
template <typename R> __device__ __forceinline__ bool is_inf(R x) { return dfabs(x) == R(INFINITY); }

/* ---------------------------------------------------------------------------
 * Closed forms
 * ------------------------------------------------------------------------- */
template <typename R> __device__ __forceinline__ int check_cons(const R *U)
{
  // Mara/rmhd-c2p.c:153-164
  if (U[ddd] < R(0)) return C2P_CONS_NEGATIVE_DENSITY;
  if (U[tau] < R(0)) return C2P_CONS_NEGATIVE_ENERGY;
#pragma unroll
  for (int i = 0; i < 8; ++i) if (is_nan(U[i])) return C2P_CONS_CONTAINS_NAN;
  return C2P_SUCCESS;
}
template <typename R> __device__ __forceinline__ int check_prim(const R *P)
{
  // Mara/rmhd-c2p.c:165-178
  const R v2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
  if (v2 >= R(1)) return C2P_PRIM_SUPERLUMINAL;
  if (P[pre] < R(0)) return C2P_PRIM_NEGATIVE_PRESSURE;
  if (P[rho] < R(0)) return C2P_PRIM_NEGATIVE_RESTMASS;
#pragma unroll
  for (int i = 0; i < 8; ++i) if (is_nan(P[i])) return C2P_PRIM_CONTAINS_NAN;
  return C2P_SUCCESS;
}

/* Rmhd::PrimToCons, Mara/rmhd.cpp:133-163 (gamma-law: e = p/((Gamma-1) rho)) */
template <typename R> __device__ __forceinline__ int prim_to_cons(const R *P, R *U, R gm1)
{
  const R V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
  const R B2 = P[Bx]*P[Bx] + P[By]*P[By] + P[Bz]*P[Bz];
  const R Bv = P[Bx]*P[vx] + P[By]*P[vy] + P[Bz]*P[vz];
  const R W2 = rcp(R(1) - V2);
  const R W = dsqrt(W2);
  const R b0 = W*Bv;
  const R b2 = (B2 + b0*b0) * (R(1) - V2);
  const R iW = rcp(W), irho = rcp(P[rho]);
  const R bx = (P[Bx] + b0 * W*P[vx]) * iW;
  const R by = (P[By] + b0 * W*P[vy]) * iW;
  const R bz = (P[Bz] + b0 * W*P[vz]) * iW;
  const R e = (P[pre] * rcp(gm1)) * irho;
  const R e_ = e + R(0.5) * b2 * irho;
  const R p_ = P[pre] + R(0.5) * b2;
  const R h_ = R(1) + e_ + p_ * irho;
  U[ddd] = P[rho] * W;
  U[tau] = P[rho] * h_ * W2 - p_ - b0*b0 - U[ddd];
  U[Sx] = P[rho] * h_ * W2 * P[vx] - b0*bx;
  U[Sy] = P[rho] * h_ * W2 * P[vy] - b0*by;
  U[Sz] = P[rho] * h_ * W2 * P[vz] - b0*bz;
  U[Bx] = P[Bx];
  U[By] = P[By];
  U[Bz] = P[Bz];
  if (V2 >= R(1)) return C2P_PRIM_SUPERLUMINAL;
  return check_cons(U);
}

/* QuarticEquation::SolveOld and solve_cubic_equation, Mara/quartic.cpp:28-148.
 * Returns the extreme real roots (the only thing FluxAndEigenvalues keeps). */
template <typename R> __device__ __forceinline__ void quartic_extremes(R d4, R d3, R d2, R d1, R d0,
                                                                      R &rmax, R &rmin, int &nr)
{
  const R id4 = rcp(d4);
  const R a3 = d3*id4, a2 = d2*id4, a1 = d1*id4, a0 = d0*id4;
  const R au2 = -a2;
  const R au1 = (a1*a3 - R(4.0)*a0);
  const R au0 = R(4.0)*a0*a2 - a1*a1 - a0*a3*a3;

  // resolvent cubic with c3 = 1
  R x1, x3;
  int ncub;
  {
    const R ca2 = au2, ca1 = au1, ca0 = au0;
    const R q = ca1*R(1.0/3.0) - ca2*ca2*R(1.0/9.0);
    const R r = (ca1*ca2 - R(3.0)*ca0)*R(1.0/6.0) - ca2*ca2*ca2*R(1.0/27.0);
    const R delta = q*q*q + r*r;
    if (delta > R(0)) {
      R s1 = r + dsqrt(delta);
      s1 = (s1 >= R(0)) ? dcbrt_pow(s1) : -dcbrt_pow(-s1);
      R s2 = r - dsqrt(delta);
      s2 = (s2 >= R(0)) ? dcbrt_pow(s2) : -dcbrt_pow(-s2);
      x1 = (s1+s2) - ca2*R(1.0/3.0);
      x3 = R(-0.5) * (s1+s2) - ca2*R(1.0/3.0);
      ncub = 1;
    }
    else if (delta < R(0)) {
      const R theta = dacos(r*rcp(dsqrt(-q*q*q))) * R(1.0/3.0);
      const R costh = dcos(theta);
      const R sinth = dsin(theta);
      const R sq = dsqrt(-q);
      x1 = R(2.0)*sq*costh - ca2*R(1.0/3.0);
      x3 = -sq*costh - ca2*R(1.0/3.0) + dsqrt(R(3.)) * sq * sinth;
      ncub = 3;
    }
    else {
      const R s = (r >= R(0)) ? dcbrt_pow(r) : -dcbrt_pow(-r);
      x1 = R(2.0)*s - ca2*R(1.0/3.0);
      x3 = -s - ca2*R(1.0/3.0);
      ncub = 3;
    }
  }
  const R u1 = (ncub == 1) ? x1 : ((x1 > x3) ? x1 : x3);
  const R R2 = R(0.25)*a3*a3 + u1 - a2;
  const R Rr = (R2 > R(0)) ? dsqrt(R2) : R(0);
  R D2, E2;
  if (Rr != R(0)) {
    const R foo1 = R(0.75)*a3*a3 - R2 - R(2.0)*a2;
    const R foo2 = R(0.25)*(R(4.0)*a3*a2 - R(8.0)*a1 - a3*a3*a3) * rcp(Rr);
    D2 = foo1 + foo2;
    E2 = foo1 - foo2;
  }
  else {
    const R foo1 = R(0.75)*a3*a3 - R(2.0)*a2;
    const R foo2 = R(2.0) * dsqrt(u1*u1 - R(4.0)*a0);
    D2 = foo1 + foo2;
    E2 = foo1 - foo2;
  }
  // roots in the reference's order: r1 r2 (if D2 >= 0) then r3 r4 (if E2 >= 0)
  rmax = R(-1.0);
  rmin = R(1.0);
  nr = 0;
  if (D2 >= R(0)) {
    const R D = dsqrt(D2);
    const R r1 = R(-0.25)*a3 + R(0.5)*Rr - R(0.5)*D;
    const R r2 = R(-0.25)*a3 + R(0.5)*Rr + R(0.5)*D;
    if (r1 > rmax) rmax = r1;
    if (r1 < rmin) rmin = r1;
    if (r2 > rmax) rmax = r2;
    if (r2 < rmin) rmin = r2;
    nr += 2;
  }
  if (E2 >= R(0)) {
    const R E = dsqrt(E2);
    const R r3 = R(-0.25)*a3 - R(0.5)*Rr - R(0.5)*E;
    const R r4 = R(-0.25)*a3 - R(0.5)*Rr + R(0.5)*E;
    if (r3 > rmax) rmax = r3;
    if (r3 < rmin) rmin = r3;
    if (r4 > rmax) rmax = r4;
    if (r4 < rmin) rmin = r4;
    nr += 2;
  }
}

/* Extreme root of the characteristic quartic of FluxAndEigenvalues in its factored form
 *     Q(x) = K (x - v)^4 + (1 - x^2) [L (x - v)^2 + (b0 x - bi)^2],
 * which is the reference's A4 x^4 + ... + A0 (Mara/rmhd.cpp:240-258) before expansion.
 * All four roots are real and Q has a positive leading coefficient, so Newton's method
 * started outside the roots moves monotonically onto the extreme one. Evaluated in this
 * form, Q keeps its relative accuracy near the roots even when they crowd within 1/W^2 of
 * the flow speed, which is where single-precision Ferrari loses every digit.
 * Stopping: a step below 1e-7, or a next step predicted below 1e-7. Once convergence is
 * quadratic, s_{n+1} ~ M s_n^2 with M ~ s_n / s_{n-1}^2, so the step after s_n is about
 * s_n^3 / s_{n-1}^2; near a double root, where convergence is linear (s_{n+1} ~ s_n / 2),
 * the estimate is s_n / 4, so there the loop ends at a step of 4e-7 instead of 1e-7. Every
 * iterate stays outside the root, so an early stop can only overstate the speed. */
template <typename R>
__device__ __forceinline__ R quartic_root_from_outside(R K, R L, R b0, R bi, R v, R x)
{
  R sp = R(0);   // the previous |step|: no prediction from the first step
#pragma unroll 1
  for (int it = 0; it < 24; ++it) {
    const R y = x - v, y2 = y*y, c = b0*x - bi;
    const R q1 = L*y2 + c*c;
    const R om = (R(1) - x)*(R(1) + x);
    const R Q = K*y2*y2 + om*q1;
    const R dQ = R(4)*K*y2*y - R(2)*x*q1 + om*(R(2)*L*y + R(2)*b0*c);
    const R step = Q * rcp(dQ);
    const R xn = x - step;
    if (!(dfabs(xn) <= R(1))) break;      // not finite, or out of the light cone: keep x
    x = xn;
    const R s = dfabs(step);
    if (!(s > R(1e-7)) || s*s*s < R(1e-7)*sp*sp) break;
    sp = s;
  }
  return x;
}

/* Rmhd::FluxAndEigenvalues, Mara/rmhd.cpp:164-265, for normal axis DIM (1, 2, 3).
 * NEED_EIG = false skips the eigenvalues. */
template <typename R, int DIM, bool NEED_EIG = true>
__device__ __forceinline__ void flux_and_eigenvalues(const R *U, const R *P, R *F, R &ap, R &am, R gamma)
{
  const R gm1 = gamma - R(1);
  const R V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
  const R B2 = P[Bx]*P[Bx] + P[By]*P[By] + P[Bz]*P[Bz];
  const R Bv = P[Bx]*P[vx] + P[By]*P[vy] + P[Bz]*P[vz];
  const R W2 = rcp(R(1) - V2);
  const R W = dsqrt(W2);
  const R b0 = W*Bv;
  const R b2 = (B2 + b0*b0) * (R(1) - V2);
  const R iW = rcp(W), irho = rcp(P[rho]), igm1 = rcp(gm1);
  const R bx = (P[Bx] + b0 * W*P[vx]) * iW;
  const R by = (P[By] + b0 * W*P[vy]) * iW;
  const R bz = (P[Bz] + b0 * W*P[vz]) * iW;
  const R e = (P[pre] * igm1) * irho;
  const R p_ = P[pre] + R(0.5) * b2;
  const R h = R(1) + e + P[pre]*irho;

  constexpr int VN = DIM == 1 ? vx : DIM == 2 ? vy : vz;
  constexpr int BN = DIM == 1 ? Bx : DIM == 2 ? By : Bz;
  const R vn = P[VN];
  const R Bn = P[BN];
  F[ddd] = U[ddd] * vn;
  const R BniW = Bn * iW;
  F[tau] = U[tau] * vn - b0*BniW + p_*vn;
  F[Sx] = U[Sx] * vn - bx*BniW;
  F[Sy] = U[Sy] * vn - by*BniW;
  F[Sz] = U[Sz] * vn - bz*BniW;
  if (DIM == 1) F[Sx] = F[Sx] + p_;
  if (DIM == 2) F[Sy] = F[Sy] + p_;
  if (DIM == 3) F[Sz] = F[Sz] + p_;
  F[Bx] = DIM == 1 ? R(0) : vn*P[Bx] - P[vx]*Bn;
  F[By] = DIM == 2 ? R(0) : vn*P[By] - P[vy]*Bn;
  F[Bz] = DIM == 3 ? R(0) : vn*P[Bz] - P[vz]*Bn;
  if (!NEED_EIG) return;

  const R vi = vn;
  const R bi = DIM == 1 ? bx : DIM == 2 ? by : bz;
  const R rhoh = P[rho] + P[pre] * igm1 + P[pre];
  const R W4 = W2*W2;
  const R v2 = vi*vi;
  const R v3 = vi*v2;
  const R v4 = vi*v3;
  const R ics2 = rhoh * rcp(gamma*P[pre]);
  const R K = P[rho]*h * (ics2 - R(1)) * W4;
  const R L = -(P[rho]*h + b2*ics2) * W2;
  const R A4 = K - L - b0*b0;
  const R A3 = R(-4)*K*vi + L*vi*R(2) + R(2)*b0*bi;
  const R A2 = R(6)*K*v2 + L*(R(1.0)-v2) + b0*b0 - bi*bi;
  const R A1 = R(-4)*K*v3 - L*vi*R(2) - R(2)*b0*bi;
  const R A0 = K*v4 + L*v2 + bi*bi;
  if constexpr (sizeof(R) == 4) {
    // FP32 (mixed form): Newton on the factored quartic, started from the lab-frame
    // speeds of an isotropic wave at the largest comoving fast speed,
    // w^2 = cs^2 + ca^2 - cs^2 ca^2 (perpendicular propagation), which bound the extreme
    // roots from outside
    const R cs2 = gamma*P[pre]*rcp(rhoh);
    const R ca2 = b2*rcp(rhoh + b2);
    const R w2 = cs2 + ca2 - cs2*ca2;
    const R den = rcp(R(1) - V2*w2);
    const R rt = dsqrt(w2*(R(1) - V2)*(R(1) - V2*w2 - vi*vi*(R(1) - w2)));
    R lp = (vi*(R(1) - w2) + rt)*den + R(4e-6);
    R lm = (vi*(R(1) - w2) - rt)*den - R(4e-6);
    lp = lp < R(1) ? lp : R(1);
    lm = lm > R(-1) ? lm : R(-1);
    ap = quartic_root_from_outside(K, L, b0, bi, vi, lp);
    am = quartic_root_from_outside(K, L, b0, bi, vi, lm);
    (void)A4; (void)A3; (void)A2; (void)A1; (void)A0; (void)v3; (void)v4;
  }
  else {
    R rmax, rmin;
    int nr;
    quartic_extremes(A4, A3, A2, A1, A0, rmax, rmin, nr);
    ap = rmax;
    am = rmin;
  }
  if (dfabs(ap) > R(1.0) || dfabs(am) > R(1.0)) {
    ap = R(1.0);
    am = R(-1.0);
  }
}

/* Fluxes and fast speeds of both sides of an interface; one copy of the code runs
 * twice (the quartic is the bulkiest piece of a Riemann solve). */
template <typename R, int DIM>
__device__ __forceinline__ void flux_eig_pair(const R *Pl, const R *Pr, const R *Ul, const R *Ur,
                                              R *Fl, R *Fr, R &epl, R &eml, R &epr, R &emr, R gamma)
{
  // the four speeds are selected into place, not stored through a side-dependent address:
  // in the rolled FP64 loop that address would put them in private memory
  R sp0 = R(0), sm0 = R(0), sp1 = R(0), sm1 = R(0);
#pragma unroll Shape<R>::unroll
  for (int side = 0; side < 2; ++side) {
    R P[8], U[8], F[8], ep, em;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      P[i] = side ? Pr[i] : Pl[i];
      U[i] = side ? Ur[i] : Ul[i];
    }
    flux_and_eigenvalues<R, DIM>(U, P, F, ep, em, gamma);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (side) Fr[i] = F[i];
      else Fl[i] = F[i];
    }
    sp1 = side ? ep : sp1;
    sm1 = side ? em : sm1;
    sp0 = side ? sp0 : ep;
    sm0 = side ? sm0 : em;
  }
  epl = sp0; eml = sm0; epr = sp1; emr = sm1;
}

/* ---------------------------------------------------------------------------
 * Conserved -> primitive, FP64, Mara/rmhd-c2p.c and the cascade of Mara/rmhd.cpp:104-131
 * ------------------------------------------------------------------------- */
struct C2PState {
  double D, Tau, S2, B2, BS, BS2;
  double Cons[8];
  double gamma, gamf, pfloor;
};

__device__ __forceinline__ void c2p_new_state(C2PState &s, const double *U, double gamma, double pfloor)
{
  s.D = U[ddd];
  s.Tau = U[tau];
  s.S2 = U[Sx]*U[Sx] + U[Sy]*U[Sy] + U[Sz]*U[Sz];
  s.B2 = U[Bx]*U[Bx] + U[By]*U[By] + U[Bz]*U[Bz];
  s.BS = U[Bx]*U[Sx] + U[By]*U[Sy] + U[Bz]*U[Sz];
  s.BS2 = s.BS*s.BS;
#pragma unroll
  for (int i = 0; i < 8; ++i) s.Cons[i] = U[i];
  s.gamma = gamma;
  s.gamf = (gamma - 1.0) / gamma;
  s.pfloor = pfloor;
}

__device__ __forceinline__ void c2p_estimate_from_cons(const C2PState &s, double &Z, double &W)
{
  Z = sqrt(s.S2 + s.D*s.D);   // rmhd-c2p.c:97-101
  W = Z * rcp(s.D);
}

__device__ __forceinline__ void c2p_starting_prim(const C2PState &s, const double *P, double &Z, double &W)
{
  // rmhd-c2p.c:108-119
  const double V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
  const double W2 = 1.0 / (1.0 - V2);
  const double e = P[pre] / (P[rho] * (s.gamma - 1.0));
  const double h = 1.0 + e + P[pre] / P[rho];
  Z = P[rho] * h * W2;
  W = sqrt(W2);
}

__device__ __forceinline__ int c2p_reconstruct(const C2PState &s, double Z, double W, double *Pout)
{
  // rmhd-c2p.c:121-150; Pout is written only on success
  double P[8];
  const double iW = rcp(W), iZB = rcp(Z + s.B2);
  const double b0 = s.BS * W * rcp(Z);
  const double b0iW = b0 * iW;
  P[rho] = s.D*iW;
  P[pre] = (s.D*iW) * (Z*iW*rcp(s.D) - 1.0) * s.gamf;
  P[vx] = (s.Cons[Sx] + b0iW*s.Cons[Bx]) * iZB;
  P[vy] = (s.Cons[Sy] + b0iW*s.Cons[By]) * iZB;
  P[vz] = (s.Cons[Sz] + b0iW*s.Cons[Bz]) * iZB;
  P[Bx] = s.Cons[Bx];
  P[By] = s.Cons[By];
  P[Bz] = s.Cons[Bz];
  if (P[pre] < s.pfloor && s.pfloor > 0.0) P[pre] = s.pfloor;
  const int err = check_prim(P);
  if (err) return err;
#pragma unroll
  for (int i = 0; i < 8; ++i) Pout[i] = P[i];
  return C2P_SUCCESS;
}

/* rmhd_c2p_solve_anton2dzw, Mara/rmhd-c2p.c:182-253 (Anton et al. 2006, eqs. 84-85).
 * The residuals and Jacobian are the reference's; the powers of 1/Z and 1/W are formed
 * from one reciprocal each per iteration (the reference divides a dozen times). With zw,
 * the converged (Z, W) are returned there too (a start for the zone's next solve). */
__device__ __forceinline__ int c2p_anton2dzw(const C2PState &s, double Z, double W, double *P, int &iters,
                                             double *zw = nullptr)
{
  double error = 1.0;
  int it = 0;
  const double D = s.D, B2 = s.B2, BS2 = s.BS2, S2 = s.S2, Tau = s.Tau, gamf = s.gamf;
  const double iD = rcp(D);
  double iZ = rcp(Z), iW = rcp(W);
  while (error > 1e-12) {
    const double Z2 = Z*Z;
    const double Z3 = Z*Z2;
    const double W2 = W*W;
    const double iZ2 = iZ*iZ, iZ3 = iZ2*iZ, iW2 = iW*iW, iW3 = iW2*iW;
    const double Pre = (D*iW) * (Z*iW*iD - 1.0) * gamf;
    const double df0dZ = 2*(B2+Z)*(BS2*W2 + (W2-1)*Z3) * (iW2*iZ3);
    const double df0dW = 2*(B2+Z)*(B2+Z) * iW3;
    const double df1dZ = 1.0 + BS2*iZ3 - gamf*iW2;
    const double df1dW = B2*iW3 + (2*Z - D*W)*iW3 * gamf;
    const double f0 = -S2 + (Z+B2)*(Z+B2)*(W2-1)*iW2 - (2*Z+B2)*BS2*iZ2;
    const double f1 = -Tau + Z+B2 - Pre - 0.5*B2*iW2 - 0.5*BS2*iZ2 - D;
    const double idet = rcp(df0dZ*df1dW - df0dW*df1dZ);
    const double G0 = df1dW*idet, G1 = -df0dW*idet, G2 = -df1dZ*idet, G3 = df0dZ*idet;
    const double dZ = -(G0*f0 + G1*f1);
    const double dW = -(G2*f0 + G3*f1);
    double Z_new = Z + dZ;
    double W_new = W + dW;
    Z_new = (Z_new > 0.0) ? Z_new : -Z_new;
    Z_new = (Z_new < 1e20) ? Z_new : Z;
    W_new = (W_new > 1.0) ? W_new : 1.0;
    W_new = (W_new < 1e12) ? W_new : 1e12;
    Z = Z_new;
    W = W_new;
    iZ = rcp(Z);
    iW = rcp(W);
    error = fabs(dZ*iZ) + fabs(dW*iW);
    ++it;
    if (it == 250) { iters += it; return C2P_MAXITER; }
  }
  iters += it;
  if (zw) {
    zw[0] = Z;
    zw[1] = W;
  }
  return c2p_reconstruct(s, Z, W, P);
}

/* (Z, W) = (rho h W^2, W) of a primitive state: the Newton variables of anton2dzw at P,
 * as c2p_starting_prim forms them (rmhd-c2p.c:108-119), with reciprocals */
__device__ __forceinline__ void c2p_zw_of_prim(const double *P, double gamma, double &Z, double &W)
{
  const double V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
  const double W2 = rcp(1.0 - V2);
  const double h = 1.0 + P[pre] * rcp(P[rho]) * (gamma * rcp(gamma - 1.0));
  Z = P[rho] * h * W2;
  W = sqrt(W2);
}

/* anton2dzw in FP32 for the mixed form's HLL state: the same residuals, Jacobian and
 * bounds as c2p_anton2dzw, started from the primitive state Ps (the average of the two
 * reconstructed states), with a step tolerance and iteration cap for single precision.
 * Its result only seeds the FP32 p* Newton solve, whose own tolerance decides the flux.
 * P is written only on success; a failure sends the interface to the FP64 solver. */
constexpr float kHllC2pStep32 = 4e-6f;
constexpr int kHllC2pIter32 = 16;
__device__ __forceinline__ int hll_c2p_f32(const float *U, const float *Ps, float gamma, float pfloor, float *P,
                                           int &iters)
{
  const float D = U[ddd], Tau = U[tau];
  const float S2 = U[Sx]*U[Sx] + U[Sy]*U[Sy] + U[Sz]*U[Sz];
  const float B2 = U[Bx]*U[Bx] + U[By]*U[By] + U[Bz]*U[Bz];
  const float BS = U[Bx]*U[Sx] + U[By]*U[Sy] + U[Bz]*U[Sz];
  const float BS2 = BS*BS;
  const float gamf = (gamma - 1.0f) * rcp(gamma);
  // the start from Ps (rmhd-c2p.c:108-119)
  float Z, W;
  {
    const float V2 = Ps[vx]*Ps[vx] + Ps[vy]*Ps[vy] + Ps[vz]*Ps[vz];
    const float W2 = rcp(1.0f - V2);
    const float ir = rcp(Ps[rho]);
    const float h = 1.0f + Ps[pre] * ir * (rcp(gamma - 1.0f) + 1.0f);
    Z = Ps[rho] * h * W2;
    W = dsqrt(W2);
  }
  const float iD = rcp(D);
  float iZ = rcp(Z), iW = rcp(W);
  float error = 1.0f;
  int it = 0;
  while (error > kHllC2pStep32) {
    const float Z2 = Z*Z, Z3 = Z*Z2, W2 = W*W;
    const float iZ2 = iZ*iZ, iZ3 = iZ2*iZ, iW2 = iW*iW, iW3 = iW2*iW;
    const float Pre = (D*iW) * (Z*iW*iD - 1.0f) * gamf;
    const float df0dZ = 2.0f*(B2+Z)*(BS2*W2 + (W2-1.0f)*Z3) * (iW2*iZ3);
    const float df0dW = 2.0f*(B2+Z)*(B2+Z) * iW3;
    const float df1dZ = 1.0f + BS2*iZ3 - gamf*iW2;
    const float df1dW = B2*iW3 + (2.0f*Z - D*W)*iW3 * gamf;
    const float f0 = -S2 + (Z+B2)*(Z+B2)*(W2-1.0f)*iW2 - (2.0f*Z+B2)*BS2*iZ2;
    const float f1 = -Tau + Z+B2 - Pre - 0.5f*B2*iW2 - 0.5f*BS2*iZ2 - D;
    const float idet = rcp(df0dZ*df1dW - df0dW*df1dZ);
    const float dZ = -(df1dW*f0 - df0dW*f1) * idet;
    const float dW = -(-df1dZ*f0 + df0dZ*f1) * idet;
    float Z_new = Z + dZ, W_new = W + dW;
    Z_new = (Z_new > 0.0f) ? Z_new : -Z_new;
    Z_new = (Z_new < 1e20f) ? Z_new : Z;
    W_new = (W_new > 1.0f) ? W_new : 1.0f;
    W_new = (W_new < 1e12f) ? W_new : 1e12f;
    Z = Z_new;
    W = W_new;
    iZ = rcp(Z);
    iW = rcp(W);
    error = dfabs(dZ*iZ) + dfabs(dW*iW);
    ++it;
    if (it == kHllC2pIter32) { iters += it; return C2P_MAXITER; }
  }
  iters += it;
  // rmhd-c2p.c:121-150
  float Q[8];
  const float b0iW = BS * W * iZ * iW, iZB = rcp(Z + B2);
  Q[rho] = D*iW;
  Q[pre] = (D*iW) * (Z*iW*iD - 1.0f) * gamf;
  Q[vx] = (U[Sx] + b0iW*U[Bx]) * iZB;
  Q[vy] = (U[Sy] + b0iW*U[By]) * iZB;
  Q[vz] = (U[Sz] + b0iW*U[Bz]) * iZB;
  Q[Bx] = U[Bx];
  Q[By] = U[By];
  Q[Bz] = U[Bz];
  if (Q[pre] < pfloor && pfloor > 0.0f) Q[pre] = pfloor;
  const int err = check_prim(Q);
  if (err) return err;
#pragma unroll
  for (int i = 0; i < 8; ++i) P[i] = Q[i];
  return C2P_SUCCESS;
}

/* rmhd_c2p_solve_noble1dw, Mara/rmhd-c2p.c:261-326 (Noble et al. 2006, eq. 29) */
__device__ __forceinline__ int c2p_noble1dw(const C2PState &s, double Z, double *P, int &iters)
{
  double error = 1.0;
  int it = 0;
  const double D = s.D, B2 = s.B2, BS2 = s.BS2, S2 = s.S2, Tau = s.Tau, gamf = s.gamf;
  const double iD = rcp(D);
  double iZ = rcp(Z);
  while (error > 1e-12) {
    const double Z2 = Z*Z;
    const double iZ2 = iZ*iZ, iZ3 = iZ2*iZ;
    const double a = S2*Z2 + BS2*(B2 + 2*Z);
    const double b = (B2 + Z)*(B2 + Z)*Z2;
    const double ap = 2*(S2*Z + BS2);
    const double bp = 2*Z*(B2 + Z)*(B2 + 2*Z);
    const double ib = rcp(b);
    const double V2 = a * ib;
    const double W2 = rcp(1.0 - V2);
    const double W = sqrt(W2);
    const double W3 = W*W2;
    const double iW = rcp(W);
    const double Pre = (D*iW) * (Z*iW*iD - 1.0) * gamf;
    const double dv2dZ = (ap*b - bp*a) * (ib*ib);
    const double delPdelZ = gamf*(1.0 - V2);
    const double delPdelW = gamf*(D*(1.0 - V2) - 2*Z*(iW*iW*iW));
    const double dWdv2 = 0.5*W3;
    const double dPdZ = delPdelW * dWdv2 * dv2dZ + delPdelZ;
    const double f = Tau + D - 0.5*B2*(1+V2) + 0.5*BS2*iZ2 - Z + Pre;
    const double g = -0.5*B2*dv2dZ - BS2*iZ3 - 1.0 + dPdZ;
    const double dZ = -f*rcp(g);
    double Z_new = Z + dZ;
    Z_new = (Z_new > 0.0) ? Z_new : -Z_new;
    Z_new = (Z_new < 1e20) ? Z_new : Z;
    Z = Z_new;
    iZ = rcp(Z);
    error = fabs(dZ*iZ);
    ++it;
    if (it == 250) { iters += it; return C2P_MAXITER; }
  }
  iters += it;
  const double Z2 = Z*Z;
  const double a = S2*Z2 + BS2*(B2 + 2*Z);
  const double b = (B2 + Z)*(B2 + Z)*Z2;
  const double V2 = a * rcp(b);
  const double W2 = rcp(1.0 - V2);
  const double W = sqrt(W2);
  return c2p_reconstruct(s, Z, W, P);
}

/* Rmhd::ConsToPrim gamma-law cascade, Mara/rmhd.cpp:59-69,104-131. P holds the previous
 * primitive state (second guess) and is written only on success. `prev_valid` = false
 * means the previous state is the zero-filled PrimitiveArray (first RK stage,
 * Mara/mara.cpp:321), whose guess fails its first Newton step as NaN; those two
 * attempts are then skipped with the same outcome. `first` returns which attempt
 * succeeded (0..3) or 4.
 * `prev(P)` loads the previous state into P where the cascade first needs it: before the
 * first attempt from it, or on an early error (so that P ends as it would have with the
 * state loaded up front). Callers that load P beforehand pass nothing. The first attempt
 * starts from the estimate and nearly always succeeds, so a lazy caller reads the previous
 * state only for the rare zone that needs it. */
struct NoPrev {
  __device__ void operator()(double *) const {}
};
template <int UNROLL = 1, typename Prev = NoPrev>
__device__ __forceinline__ int cons_to_prim(const double *U, double *P, double gamma, double pfloor,
                                            bool prev_valid, int &iters, int &first, Prev prev = Prev())
{
  first = 4;
  int error = check_cons(U);
  if (error) {
    if (prev_valid) prev(P);
    return error;
  }
  C2PState s;
  c2p_new_state(s, U, gamma, pfloor);
  // attempts in the reference's order: anton2dzw from the estimate, then from the
  // previous state, noble1dw from the estimate, then from the previous state
#pragma unroll UNROLL
  for (int solver = 0; solver < 2; ++solver) {
#pragma unroll UNROLL
    for (int guess = 0; guess < 2; ++guess) {
      if (guess == 1 && !prev_valid) {
        // the zero-filled state gives Z = NaN: one step, then a NaN primitive state
        iters += 1;
        error = C2P_PRIM_CONTAINS_NAN;
        continue;
      }
      double Z, W;
      if (guess == 0) c2p_estimate_from_cons(s, Z, W);
      else {
        if (solver == 0) prev(P);
        c2p_starting_prim(s, P, Z, W);
      }
      error = solver == 0 ? c2p_anton2dzw(s, Z, W, P, iters) : c2p_noble1dw(s, Z, P, iters);
      if (!error) {
        first = 2 * solver + guess;
        return 0;
      }
    }
  }
  return error;
}

/* ---------------------------------------------------------------------------
 * Reconstruction, Mara/weno.c
 * ------------------------------------------------------------------------- */
template <typename R> __device__ __forceinline__ int sgn(R x) { return (x > R(0)) - (x < R(0)); }

/* __plm with __plm_minmod, weno.c:85-96: v points at the centre zone */
template <typename R> __device__ __forceinline__ R plm(R vm, R v0, R vp, R sg, R theta)
{
  const R a = theta * (v0 - vm);
  const R b = R(0.5) * (vp - vm);
  const R c = theta * (vp - v0);
  const R fa = dfabs(a), fb = dfabs(b), fc = dfabs(c);
  const R m01 = fa < fb ? fa : fb;
  const R mn = m01 < fc ? m01 : fc;
  const int s1 = sgn(a) + sgn(b);
  const int as1 = s1 < 0 ? -s1 : s1;
  const R lim = R(0.25)*R(as1)*R(sgn(a) + sgn(c))*mn;
  return v0 + sg*R(0.5)*lim;
}

/* __weno5, weno.c:97-154, with the FV C2R (right face of v2) or C2L (left face of v2)
 * coefficient sets of weno.c:24-31; vm2..vp2 are v[-2..2]. */
template <typename R, bool C2R, bool FD = false>
__device__ __forceinline__ R weno5(R vm2, R vm1, R v0, R vp1, R vp2, int IS, R sz10A)
{
  // c[k][j] and d[k]: finite-volume C2R / C2L, or (FD) the finite-difference flux
  // coefficients CeesC2R_FD / CeesC2L_FD, DeesC2R_FD / DeesC2L_FD of weno.c:30-38
  const R c00 = FD ? (C2R ? R(1./3.) : R(11./6.)) : (C2R ? R(3./8.) : R(15./8.));
  const R c01 = FD ? (C2R ? R(5./6.) : R(-7./6.)) : (C2R ? R(3./4.) : R(-5./4.));
  const R c02 = FD ? (C2R ? R(-1./6.) : R(1./3.)) : (C2R ? R(-1./8.) : R(3./8.));
  const R c10 = FD ? (C2R ? R(-1./6.) : R(1./3.)) : (C2R ? R(-1./8.) : R(3./8.));
  const R c11 = FD ? R(5./6.) : R(3./4.);
  const R c12 = FD ? (C2R ? R(1./3.) : R(-1./6.)) : (C2R ? R(3./8.) : R(-1./8.));
  const R c20 = FD ? (C2R ? R(1./3.) : R(-1./6.)) : (C2R ? R(3./8.) : R(-1./8.));
  const R c21 = FD ? (C2R ? R(-7./6.) : R(5./6.)) : (C2R ? R(-5./4.) : R(3./4.));
  const R c22 = FD ? (C2R ? R(11./6.) : R(1./3.)) : (C2R ? R(15./8.) : R(3./8.));
  const R d0 = FD ? (C2R ? R(0.3) : R(0.1)) : (C2R ? R(5./16.) : R(1./16.));
  const R d1 = FD ? R(0.6) : R(5./8.);
  const R d2 = FD ? (C2R ? R(0.1) : R(0.3)) : (C2R ? R(1./16.) : R(5./16.));

  const R vs0 = c00*v0 + c01*vp1 + c02*vp2;
  const R vs1 = c10*vm1 + c11*v0 + c12*vp1;
  const R vs2 = c20*vm2 + c21*vm1 + c22*v0;
  const R t0 = v0 - R(2)*vp1 + vp2, u0 = R(3)*v0 - R(4)*vp1 + vp2;
  const R t1 = vm1 - R(2)*v0 + vp1, u1 = vm1 - vp1;
  const R t2 = vm2 - R(2)*vm1 + v0, u2 = vm2 - R(4)*vm1 + R(3)*v0;
  R B0 = R(13./12.)*(t0*t0) + R(1./4.)*(u0*u0);
  R B1 = R(13./12.)*(t1*t1) + R(1./4.)*(u1*u1);
  R B2 = R(13./12.)*(t2*t2) + R(1./4.)*(u2*u2);
  R w0, w1, w2;
  if (IS == IS_B08) {
    const R eps = R(1e-14);
    const R tau5 = dfabs(B0 - B2);
    w0 = d0 * (R(1.0) + (tau5 / (B0 + eps)));
    w1 = d1 * (R(1.0) + (tau5 / (B1 + eps)));
    w2 = d2 * (R(1.0) + (tau5 / (B2 + eps)));
  }
  else if (IS == IS_SZ10) {
    const R epsp = R(1e-10);
    const R m01 = B0 < B1 ? B0 : B1, minB = m01 < B2 ? m01 : B2;
    const R x01 = B0 > B1 ? B0 : B1, maxB = x01 > B2 ? x01 : B2;
    const R R0 = minB / (maxB + epsp);
    B0 = R0*sz10A*minB + B0;
    B1 = R0*sz10A*minB + B1;
    B2 = R0*sz10A*minB + B2;
    w0 = d0 / ((epsp + B0)*(epsp + B0));
    w1 = d1 / ((epsp + B1)*(epsp + B1));
    w2 = d2 / ((epsp + B2)*(epsp + B2));
  }
  else {
    // Jiang & Shu (1996): w_k = d_k / (eps + B_k)^2. Multiplying all three by
    // prod_j (eps + B_j)^2 leaves the normalised result unchanged and needs no division.
    const R epsp = R(1e-6);
    R a0 = (epsp + B0)*(epsp + B0), a1 = (epsp + B1)*(epsp + B1), a2 = (epsp + B2)*(epsp + B2);
    if (sizeof(R) == 4) {
      const R m01 = a0 > a1 ? a0 : a1;
      const R im = rcp(m01 > a2 ? m01 : a2);
      a0 *= im; a1 *= im; a2 *= im;
    }
    w0 = d0 * (a1*a2);
    w1 = d1 * (a0*a2);
    w2 = d2 * (a0*a1);
  }
  const R wtot = w0 + w1 + w2;
  return (w0*vs0 + w1*vs1 + w2*vs2) * rcp(wtot);
}

/* ---------------------------------------------------------------------------
 * Riemann solvers. Inputs pl/pr are in the lab frame; DIM is the normal axis.
 * ------------------------------------------------------------------------- */
template <typename R> __device__ __forceinline__ void note_lambda(R &lam, R am, R ap)
{
  const R ml = (dfabs(am) < dfabs(ap)) ? dfabs(ap) : dfabs(am);
  if (lam < ml) lam = ml;
}

/* HllRiemannSolver, Mara/riemann_hll.cpp:20-64, s = 0 */
template <typename R, int DIM> __device__ __forceinline__ int riemann_hll(const R *Pl, const R *Pr, R *F,
                                                                         R gamma, R &lam)
{
  R Ul[8], Ur[8], Fl[8], Fr[8];
  R epl = 1, epr = 1, eml = 1, emr = 1;
  const R gm1 = gamma - R(1);
  if (prim_to_cons(Pl, Ul, gm1)) return 1;
  if (prim_to_cons(Pr, Ur, gm1)) return 1;
  flux_eig_pair<R, DIM>(Pl, Pr, Ul, Ur, Fl, Fr, epl, eml, epr, emr, gamma);
  const R ap = (epl > epr) ? epl : epr;
  const R am = (eml < emr) ? eml : emr;
  note_lambda(lam, am, ap);
  if (R(0) <= am) {
#pragma unroll
    for (int i = 0; i < 8; ++i) F[i] = Fl[i];
  }
  else if (R(0) <= ap) {
#pragma unroll
    for (int i = 0; i < 8; ++i) F[i] = (ap*Fl[i] - am*Fr[i] + ap*am*(Ur[i] - Ul[i])) * rcp(ap - am);
  }
  else {
#pragma unroll
    for (int i = 0; i < 8; ++i) F[i] = Fr[i];
  }
  return 0;
}

template <int DIM> struct Perm {
  // Mara/riemann_hlld-rmhd.cpp:164-182
  static constexpr int B1 = DIM == 1 ? Bx : DIM == 2 ? By : Bz;
  static constexpr int B2 = DIM == 1 ? By : DIM == 2 ? Bz : Bx;
  static constexpr int B3 = DIM == 1 ? Bz : DIM == 2 ? Bx : By;
  static constexpr int S1 = DIM == 1 ? Sx : DIM == 2 ? Sy : Sz;
  static constexpr int S2 = DIM == 1 ? Sy : DIM == 2 ? Sz : Sx;
  static constexpr int S3 = DIM == 1 ? Sz : DIM == 2 ? Sx : Sy;
  static constexpr int v1 = DIM == 1 ? vx : DIM == 2 ? vy : vz;
  static constexpr int v2 = DIM == 1 ? vy : DIM == 2 ? vz : vx;
  static constexpr int v3 = DIM == 1 ? vz : DIM == 2 ? vx : vy;
};

/* HllcRmhdRiemannSolver, Mara/riemann_hllc-rmhd.cpp:26-178 (Mignone & Bodo 2006), s = 0 */
template <typename R, int DIM> __device__ __forceinline__ int riemann_hllc(const R *pl, const R *pr, R *F,
                                                                          R gamma, R &lam)
{
  using T = Tol<R>;
  using Pm = Perm<DIM>;
  constexpr int B1 = Pm::B1, B2 = Pm::B2, B3 = Pm::B3, S1 = Pm::S1, S2 = Pm::S2, S3 = Pm::S3;
  constexpr int v1 = Pm::v1, v2 = Pm::v2, v3 = Pm::v3;
  R Pl[8], Pr[8], Ul[8], Ur[8], Fl[8], Fr[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) { Pl[i] = pl[i]; Pr[i] = pr[i]; }
  Pl[B1] = Pr[B1] = R(0.5)*(Pl[B1] + Pr[B1]);
  const R gm1 = gamma - R(1);
  if (prim_to_cons(Pl, Ul, gm1)) return 1;
  if (prim_to_cons(Pr, Ur, gm1)) return 1;
  R epl, epr, eml, emr;
  flux_eig_pair<R, DIM>(Pl, Pr, Ul, Ur, Fl, Fr, epl, eml, epr, emr, gamma);
  Ul[tau] += Ul[ddd]; Fl[tau] += Fl[ddd];
  Ur[tau] += Ur[ddd]; Fr[tau] += Fr[ddd];
  const R ap = (epl > epr) ? epl : epr;
  const R am = (eml < emr) ? eml : emr;
  note_lambda(lam, am, ap);

  if (R(0) <= am) {
#pragma unroll
    for (int i = 0; i < 8; ++i) F[i] = Fl[i];
    F[tau] -= F[ddd];
    return 0;
  }
  R U_hll[8], F_hll[8];
  const R ida = rcp(ap - am);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    U_hll[i] = (ap*Ur[i] - am*Ul[i] + (Fl[i] - Fr[i])) * ida;
    F_hll[i] = (ap*Fl[i] - am*Fr[i] + ap*am*(Ur[i] - Ul[i])) * ida;
  }
  R Ul_[8], Ur_[8], lc;
  if (dfabs(U_hll[B1]) > T::hllc_small_bx) {
    const R B_dot_FB = U_hll[B2]*F_hll[B2] + U_hll[B3]*F_hll[B3];
    const R a = F_hll[tau] - B_dot_FB;
    const R b = -F_hll[S1] - U_hll[tau] + (U_hll[B2]*U_hll[B2] + U_hll[B3]*U_hll[B3]) +
                (F_hll[B2]*F_hll[B2] + F_hll[B3]*F_hll[B3]);
    const R c = U_hll[S1] - B_dot_FB;
    const R pv1 = lc = (dfabs(a) < T::hllc_small_a) ? -c/b : (-b - dsqrt(b*b - R(4)*a*c)) / (R(2)*a);
    const R pv2 = (U_hll[B2]*pv1 - F_hll[B2]) / U_hll[B1];
    const R pv3 = (U_hll[B3]*pv1 - F_hll[B3]) / U_hll[B1];
    const R pB1 = U_hll[B1], pB2 = U_hll[B2], pB3 = U_hll[B3];
    const R v_dotB_ = pv1*pB1 + pv2*pB2 + pv3*pB3;
    const R v_dotv_ = pv1*pv1 + pv2*pv2 + pv3*pv3;
    const R gm2_ = R(1.0) / (R(1.0) - v_dotv_);
    const R ppre = F_hll[S1] + pB1*pB1/gm2_ - (F_hll[tau] - pB1*v_dotB_)*pv1;
    Ul_[ddd] = (am - Pl[v1]) / (am - pv1) * Ul[ddd];
    Ur_[ddd] = (ap - Pr[v1]) / (ap - pv1) * Ur[ddd];
    Ul_[tau] = (am*Ul[tau] - Ul[S1] + ppre*pv1 - v_dotB_*pB1) / (am - pv1);
    Ur_[tau] = (ap*Ur[tau] - Ur[S1] + ppre*pv1 - v_dotB_*pB1) / (ap - pv1);
    Ul_[S1] = (Ul_[tau] + ppre)*pv1 - v_dotB_*pB1;
    Ur_[S1] = (Ur_[tau] + ppre)*pv1 - v_dotB_*pB1;
    Ul_[S2] = (-pB1*(pB2/gm2_ + v_dotB_*pv2) + am*Ul[S2] - Fl[S2]) / (am - pv1);
    Ur_[S2] = (-pB1*(pB2/gm2_ + v_dotB_*pv2) + ap*Ur[S2] - Fr[S2]) / (ap - pv1);
    Ul_[S3] = (-pB1*(pB3/gm2_ + v_dotB_*pv3) + am*Ul[S3] - Fl[S3]) / (am - pv1);
    Ur_[S3] = (-pB1*(pB3/gm2_ + v_dotB_*pv3) + ap*Ur[S3] - Fr[S3]) / (ap - pv1);
    Ul_[B1] = Ur_[B1] = pB1;
    Ul_[B2] = Ur_[B2] = pB2;
    Ul_[B3] = Ur_[B3] = pB3;
  }
  else {
    const R a = F_hll[tau];
    const R b = -F_hll[S1] - U_hll[tau];
    const R c = U_hll[S1];
    const R v1_ = lc = (dfabs(a) < T::hllc_small_a) ? -c/b : (-b - dsqrt(b*b - R(4)*a*c)) / (R(2)*a);
    const R p_ = -F_hll[tau]*v1_ + F_hll[S1];
    Ul_[ddd] = (am - Pl[v1]) / (am - v1_) * Ul[ddd];
    Ur_[ddd] = (ap - Pr[v1]) / (ap - v1_) * Ur[ddd];
    Ul_[tau] = (am*Ul[tau] - Ul[S1] + p_*v1_) / (am - v1_);
    Ur_[tau] = (ap*Ur[tau] - Ur[S1] + p_*v1_) / (ap - v1_);
    Ul_[S1] = (Ul_[tau] + p_)*v1_;
    Ur_[S1] = (Ur_[tau] + p_)*v1_;
    Ul_[S2] = (am - Pl[v1]) / (am - v1_) * Ul[S2];
    Ur_[S2] = (ap - Pr[v1]) / (ap - v1_) * Ur[S2];
    Ul_[S3] = (am - Pl[v1]) / (am - v1_) * Ul[S3];
    Ur_[S3] = (ap - Pr[v1]) / (ap - v1_) * Ur[S3];
    Ul_[B1] = U_hll[B1];
    Ur_[B1] = U_hll[B1];
    Ul_[B2] = (am - Pl[v1]) / (am - v1_) * Ul[B2];
    Ur_[B2] = (ap - Pr[v1]) / (ap - v1_) * Ur[B2];
    Ul_[B3] = (am - Pl[v1]) / (am - v1_) * Ul[B3];
    Ur_[B3] = (ap - Pr[v1]) / (ap - v1_) * Ur[B3];
  }
  (void)v2; (void)v3;
  // the reference's region chain (:166-171) for s = 0 with s > am
  if (R(0) <= lc) {
#pragma unroll
    for (int i = 0; i < 8; ++i) F[i] = Fl[i] + am*(Ul_[i]-Ul[i]);
  }
  else if (R(0) <= ap) {
#pragma unroll
    for (int i = 0; i < 8; ++i) F[i] = Fr[i] + ap*(Ur_[i]-Ur[i]);
  }
  else {
#pragma unroll
    for (int i = 0; i < 8; ++i) F[i] = Fr[i];
  }
  F[tau] -= F[ddd];
  return 0;
}

/* HLLD (Mignone, Ugliano & Bodo 2008), Mara/riemann_hlld-rmhd.cpp:143-579, in the
 * x-normal frame. Holds the jump states across the fast waves (eqn 12). */
template <typename R> struct Hlld {
  R ap, am;
  R UlBx, UrBx;     // normal field (equal on both sides)
  R LE, LD, LB0, LB1, LB2, Lm0, Lm1, Lm2;
  R RE, RD, RB0, RB1, RB2, Rm0, Rm1, Rm2;

  struct Inner { R v0, v1, v2, B0, B1, B2, id; };

  /* FillInnerWaveStates, :321-384 */
  __device__ __forceinline__ void fill(R p, R Kl[3], R Kr[3], R Bc[3], R &etaL, R &etaR,
                                       R &wL, R &wR, Inner &aL, Inner &aR) const
  {
    using T = Tol<R>;
    {
      const R A = Lm0 - am*LE + p*(R(1) - am*am);
      const R G = LB1*LB1 + LB2*LB2;
      const R C = LB1*Lm1 + LB2*Lm2;
      const R Q = -A - G + (UlBx*UlBx)*(R(1) - am*am);
      R X = UlBx*(A*am*UlBx + C) - (A + G)*(am*p + LE);
      if (dfabs(X) < T::X_floor) X = T::X_floor;
      const R iX = rcp(X);
      const R CB = C + UlBx*(am*Lm0 - LE);
      aL.v0 = (UlBx*(A*UlBx + am*C) - (A + G)*(p + Lm0)) * iX;
      aL.v1 = (Q*Lm1 + LB1*CB) * iX;
      aL.v2 = (Q*Lm2 + LB2*CB) * iX;
      const R id = rcp(am - aL.v0);
      aL.B0 = (LB0 - UlBx*aL.v0) * id;
      aL.B1 = (LB1 - UlBx*aL.v1) * id;
      aL.B2 = (LB2 - UlBx*aL.v2) * id;
      aL.id = id;
    }
    {
      const R A = Rm0 - ap*RE + p*(R(1) - ap*ap);
      const R G = RB1*RB1 + RB2*RB2;
      const R C = RB1*Rm1 + RB2*Rm2;
      const R Q = -A - G + (UrBx*UrBx)*(R(1) - ap*ap);
      R X = UrBx*(A*ap*UrBx + C) - (A + G)*(ap*p + RE);
      if (dfabs(X) < T::X_floor) X = T::X_floor;
      const R iX = rcp(X);
      const R CB = C + UrBx*(ap*Rm0 - RE);
      aR.v0 = (UrBx*(A*UrBx + ap*C) - (A + G)*(p + Rm0)) * iX;
      aR.v1 = (Q*Rm1 + RB1*CB) * iX;
      aR.v2 = (Q*Rm2 + RB2*CB) * iX;
      const R id = rcp(ap - aR.v0);
      aR.B0 = (RB0 - UrBx*aR.v0) * id;
      aR.B1 = (RB1 - UrBx*aR.v1) * id;
      aR.B2 = (RB2 - UrBx*aR.v2) * id;
      aR.id = id;
    }
    wL = p + (LE - (aL.v0*Lm0 + aL.v1*Lm1 + aL.v2*Lm2)) * aL.id;
    wR = p + (RE - (aR.v0*Rm0 + aR.v1*Rm1 + aR.v2*Rm2)) * aR.id;
    if (wL < R(0)) wL *= R(-1);
    if (wR < R(0)) wR *= R(-1);
    etaL = R(-((UlBx > R(0)) - (UlBx < R(0)))) * dsqrt(wL);
    etaR = R(+((UrBx > R(0)) - (UrBx < R(0)))) * dsqrt(wR);
    const R idenL = rcp(am*p + LE + UlBx*etaL);
    const R idenR = rcp(ap*p + RE + UrBx*etaR);
    Kl[0] = (Lm0 + p + LB0*etaL) * idenL;
    Kl[1] = (Lm1 + LB1*etaL) * idenL;
    Kl[2] = (Lm2 + LB2*etaL) * idenL;
    Kr[0] = (Rm0 + p + RB0*etaR) * idenR;
    Kr[1] = (Rm1 + RB1*etaR) * idenR;
    Kr[2] = (Rm2 + RB2*etaR) * idenR;
    const R dK = Kr[0] - Kl[0];
    const R idK = rcp(dK);
    const bool zeroBc = dfabs(dK) < T::small_delK;
    const R aRv[3] = { aR.v0, aR.v1, aR.v2 }, aLv[3] = { aL.v0, aL.v1, aL.v2 };
    const R aRB[3] = { aR.B0, aR.B1, aR.B2 }, aLB[3] = { aL.B0, aL.B1, aL.B2 };
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Bc[k] = ((aRB[k]*(Kr[0] - aR.v0) + aR.B0*aRv[k]) - (aLB[k]*(Kl[0] - aL.v0) + aL.B0*aLv[k])) * idK;
      if (zeroBc) Bc[k] = R(0);
    }
  }

  /* HlldEquation48::Function, :386-415; returns false where the reference throws */
  __device__ __forceinline__ bool function(R p, R &y) const
  {
    using T = Tol<R>;
    if (is_inf(p)) return false;
    R Kl[3], Kr[3], Bc[3], etaL, etaR, wL, wR;
    Inner aL, aR;
    fill(p, Kl, Kr, Bc, etaL, etaR, wL, wR, aL, aR);
    const R delK = Kr[0] - Kl[0];
    const R Bch0 = Bc[0]*delK, Bch1 = Bc[1]*delK, Bch2 = Bc[2]*delK;
    const R Yl = (R(1) - (Kl[0]*Kl[0] + Kl[1]*Kl[1] + Kl[2]*Kl[2])) *
                 rcp(etaL*delK - (Kl[0]*Bch0 + Kl[1]*Bch1 + Kl[2]*Bch2));
    const R Yr = (R(1) - (Kr[0]*Kr[0] + Kr[1]*Kr[1] + Kr[2]*Kr[2])) *
                 rcp(etaR*delK - (Kr[0]*Bch0 + Kr[1]*Bch1 + Kr[2]*Bch2));
    if (dfabs(delK) < T::small_delK) y = R(0);
    else if (dfabs(Bc[0]) < T::small_Bc) y = delK;
    else y = delK*(R(1) - Bc[0]*(Yr - Yl));
    return true;
  }

  /* NewtonRaphesonSolver(15, 1e-12)::Solve with the centred-difference Jacobian of
   * HlldEquation48::Jacobian, Mara/nrsolver.cpp:20-83, riemann_hlld-rmhd.cpp:311-319.
   * FP32 (the mixed form) takes the derivative by a one-sided difference from the residual
   * already evaluated: two evaluations per step instead of three, the same L1 error; and it
   * takes a start whose residual is below newton_accept as the root without a step.
   * max_it: the reference's 15, fewer on the FP32 common path (Tol<float>). */
  __device__ __forceinline__ bool newton(R &x, int &iters, int max_it = 15) const
  {
    using T = Tol<R>;
    int worse = 0, it = 0;
    R eps = R(0), eps_last;
    do {
      eps_last = eps;
      R y = R(0), f0 = R(0), f1 = R(0);
      const R dp = T::fd_relative ? T::fd_dp * (dfabs(x) > R(1e-3) ? dfabs(x) : R(1e-3)) : T::fd_dp;
      bool ok = true;
      // FP32: the residual, then (one-sided) f(x + dp); FP64: the reference's f(x),
      // f(x - dp/2), f(x + dp/2)
      constexpr int nev = T::fd_one_sided ? 2 : 3;
#pragma unroll Shape<R>::unroll
      for (int k = 0; k < nev; ++k) {
        const R pk = k == 0 ? x : (T::fd_one_sided ? x + dp : (k == 1 ? x - R(0.5)*dp : x + R(0.5)*dp));
        R fk;
        if (!function(pk, fk)) { ok = false; break; }
        if (k == 0) {
          y = fk;
          if (T::fd_one_sided) f0 = fk;
          if constexpr (T::newton_accept > 0) {
            if (dfabs(fk) < R(T::newton_accept)) { iters += it; return true; }
          }
        }
        else if (k == 1 && !T::fd_one_sided) f0 = fk;
        else f1 = fk;
      }
      if (!ok) { iters += it; return false; }
      const R J = (f1 - f0)*rcp(dp);
      const R G = rcp(J);
      const R dx = R(0) + G*y;
      x -= dx;
      eps = (y > R(0)) ? y : -y;
      worse += (eps > eps_last);
      if (worse >= 10 || it == max_it) { iters += it; return false; }
      if (is_nan(eps)) { iters += it; return false; }
    } while (eps > T::newton_ftol && ++it <= max_it);
    iters += it;
    return true;
  }

  /* HlldEquation48::ReconstructSolution, :417-538, for s = 0: the flux only. The
   * outer states U and F of the side the flux is taken from are recomputed here from
   * the primitive states (bitwise the values of the first evaluation), which keeps them
   * out of the registers during the Newton iterations. */
  __device__ __forceinline__ bool flux(R p, const R *Pl, const R *Pr, R gamma, R *F) const
  {
    using T = Tol<R>;
    R Kl[3], Kr[3], Bc[3], etaL, etaR, wL, wR;
    Inner aL, aR;
    fill(p, Kl, Kr, Bc, etaL, etaR, wL, wR, aL, aR);
    const R dKl = Kl[0]*Kl[0] + Kl[1]*Kl[1] + Kl[2]*Kl[2];
    const R dKr = Kr[0]*Kr[0] + Kr[1]*Kr[1] + Kr[2]*Kr[2];
    const R denomL = etaL - (Kl[0]*Bc[0] + Kl[1]*Bc[1] + Kl[2]*Bc[2]);
    const R denomR = etaR - (Kr[0]*Bc[0] + Kr[1]*Bc[1] + Kr[2]*Bc[2]);
    R cLv[3], cRv[3];
    const R idL = rcp(denomL), idR = rcp(denomR);
    const R aLv[3] = { aL.v0, aL.v1, aL.v2 }, aRv[3] = { aR.v0, aR.v1, aR.v2 };
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      cLv[k] = Kl[k] - Bc[k]*(R(1) - dKl) * idL;
      cRv[k] = Kr[k] - Bc[k]*(R(1) - dKr) * idR;
      if (dfabs(denomL) < T::small_denom_c) cLv[k] = aLv[k];
      if (dfabs(denomR) < T::small_denom_c) cRv[k] = aRv[k];
    }
    const R lamC = cRv[0];
    const R lmaL = Kl[0];
    const R lmaR = Kr[0];
    const R TOL = T::tol54;
    if (p < R(0)) return false;
    if (lmaL < R(-1.0) || lmaR > R(1.0) || dfabs(lamC) > R(1.0)) return false;
    if (!((wL > p) && (aL.v0 > am - TOL) && (cLv[0] > lmaL - TOL))) return false;
    if (!((wR > p) && (aR.v0 < ap + TOL) && (cRv[0] < lmaR + TOL))) return false;
    if (dfabs(cRv[0] - cLv[0]) > T::contact) return false;

    // first region of the reference's chain (:518-524) whose upper speed bounds s = 0
    const int region = R(0) <= am ? 0 : R(0) <= lmaL ? 1 : R(0) <= lamC ? 2
                     : R(0) <= lmaR ? 3 : R(0) <= ap ? 4 : 5;
    const bool left = region <= 2;
    R Uo[8], Fo[8];
    {
      R Po[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) Po[i] = left ? Pl[i] : Pr[i];
      prim_to_cons(Po, Uo, gamma - R(1));
      R e1, e2;
      flux_and_eigenvalues<R, 1, false>(Uo, Po, Fo, e1, e2, gamma);
      Uo[tau] += Uo[ddd];
      Fo[tau] += Fo[ddd];
    }
    const R lam_o = left ? am : ap;
    if (region == 0 || region == 5) {
#pragma unroll
      for (int i = 0; i < 8; ++i) F[i] = Fo[i];
    }
    else {
      Inner a;
      a.v0 = left ? aL.v0 : aR.v0; a.v1 = left ? aL.v1 : aR.v1; a.v2 = left ? aL.v2 : aR.v2;
      a.B0 = left ? aL.B0 : aR.B0; a.B1 = left ? aL.B1 : aR.B1; a.B2 = left ? aL.B2 : aR.B2;
      a.id = left ? aL.id : aR.id;
      R aU[8];
      state_a(a, p, lam_o, left ? LD : RD, left ? LE : RE, Uo, aU);
      if (region == 1 || region == 4) {
#pragma unroll
        for (int i = 0; i < 8; ++i) F[i] = Fo[i] + lam_o*(aU[i] - Uo[i]);
      }
      else {
        R cU[8];
        const R lma = left ? lmaL : lmaR;
        const R cv[3] = { left ? cLv[0] : cRv[0], left ? cLv[1] : cRv[1], left ? cLv[2] : cRv[2] };
        state_c(aU, a.v0, cv, Bc, p, lma, lamC, cU);
#pragma unroll
        for (int i = 0; i < 8; ++i) F[i] = Fo[i] + lam_o*(aU[i] - Uo[i]) + lma*(cU[i] - aU[i]);
      }
    }
    F[tau] -= F[ddd];
    return true;
  }

  /* state between a fast and an Alfven wave, :455-484 */
  __device__ __forceinline__ void state_a(const Inner &a, R p, R lam, R Dj, R Ej, const R *Uo,
                                          R *Ua) const
  {
    using T = Tol<R>;
    const R denom = lam - a.v0;
    if (dfabs(denom) < T::small_denom) {
#pragma unroll
      for (int i = 0; i < 8; ++i) Ua[i] = Uo[i];
      return;
    }
    const R vB = a.v0*a.B0 + a.v1*a.B1 + a.v2*a.B2;
    const R idn = rcp(denom);
    Ua[ddd] = Dj * idn;
    Ua[tau] = (Ej + p*a.v0 - vB*a.B0) * idn;
    Ua[Sx] = (Ua[tau] + p)*a.v0 - vB*a.B0;
    Ua[Sy] = (Ua[tau] + p)*a.v1 - vB*a.B1;
    Ua[Sz] = (Ua[tau] + p)*a.v2 - vB*a.B2;
    Ua[Bx] = a.B0;
    Ua[By] = a.B1;
    Ua[Bz] = a.B2;
  }

  /* state between an Alfven wave and the contact, :486-516 */
  __device__ __forceinline__ void state_c(const R *aU, R av0, const R *cv, const R *Bc, R p,
                                          R lma, R lamC, R *cU) const
  {
    using T = Tol<R>;
    const R denom = lma - lamC;
    if (dfabs(denom) < T::small_denom) {
#pragma unroll
      for (int i = 0; i < 8; ++i) cU[i] = aU[i];
      return;
    }
    const R vBc = cv[0]*Bc[0] + cv[1]*Bc[1] + cv[2]*Bc[2];
    const R idn = rcp(denom);
    cU[ddd] = aU[ddd] * (lma - av0) * idn;
    cU[tau] = (lma*aU[tau] - aU[Sx] + p*lamC - vBc*Bc[0]) * idn;
    cU[Sx] = (cU[tau] + p)*cv[0] - vBc*Bc[0];
    cU[Sy] = (cU[tau] + p)*cv[1] - vBc*Bc[1];
    cU[Sz] = (cU[tau] + p)*cv[2] - vBc*Bc[2];
    cU[Bx] = Bc[0];
    cU[By] = Bc[1];
    cU[Bz] = Bc[2];
  }
};

struct FluxCounters {
  int hlld_ok;       // HLLD produced the flux
  int hlld_fail;     // HLLD fell back to HLLC
  int newton_iters;  // p* Newton iterations
  int c2p_iters;     // nested c2p Newton iterations
};

/* HlldRmhdRiemannSolver::IntercellFlux, Mara/riemann_hlld-rmhd.cpp:143-277, s = 0.
 * Live-state economy: the jump states across the fast waves are formed before the
 * nested c2p, after which the outer U and F are dead; only the primitive states are
 * kept (to rebuild the flux side at the end and for the HLLC fallback, which averages
 * the normal field exactly as here, riemann_hllc-rmhd.cpp:60).
 * FAST = true is the common path alone: the nested c2p from the average state and the
 * first Newton attempt from EstimateSolution(0), or with a small normal field the flux at
 * EstimateSolution(1). Anything else (an invalid state, the c2p cascade, attempts 2 and
 * 3, the HLLC fallback) returns kHlldDefer without writing f; the caller hands the
 * interface to the full solver, which then takes the same steps up to that point and
 * gives the same result. */
constexpr int kHlldDefer = 100;   // FAST mode: the interface needs the full solver

template <typename R, int DIM, bool FAST = false>
__device__ __forceinline__ int riemann_hlld(const R *pl, const R *pr, R *f, R gamma, double pfloor, R &lam,
                                            FluxCounters &cnt)
{
  using T = Tol<R>;
  using Pm = Perm<DIM>;
  R Pl[8], Pr[8];
  Pl[pre] = pl[pre]; Pr[pre] = pr[pre];
  Pl[rho] = pl[rho]; Pr[rho] = pr[rho];
  Pl[vx] = pl[Pm::v1]; Pr[vx] = pr[Pm::v1];
  Pl[vy] = pl[Pm::v2]; Pr[vy] = pr[Pm::v2];
  Pl[vz] = pl[Pm::v3]; Pr[vz] = pr[Pm::v3];
  Pl[By] = pl[Pm::B2]; Pr[By] = pr[Pm::B2];
  Pl[Bz] = pl[Pm::B3]; Pr[Bz] = pr[Pm::B3];
  Pl[Bx] = Pr[Bx] = R(0.5)*(pl[Pm::B1] + pr[Pm::B1]);

  Hlld<R> e;
  // the HLL state and its primitives, in the precision of the nested c2p's first attempt
  using RN = std::conditional_t<T::nested_c2p_fp32, float, double>;
  RN Uh[8], Ph[8];
  R Uhll_tau, Uhll_ddd, Uhll_Sx, Fhll_tau, Fhll_ddd, Fhll_Sx;
  {
    R Ul[8], Ur[8], Fl[8], Fr[8];
    const R gm1 = gamma - R(1);
    if (prim_to_cons(Pl, Ul, gm1)) return FAST ? kHlldDefer : 1;
    if (prim_to_cons(Pr, Ur, gm1)) return FAST ? kHlldDefer : 1;
    R epl, epr, eml, emr;
    flux_eig_pair<R, 1>(Pl, Pr, Ul, Ur, Fl, Fr, epl, eml, epr, emr, gamma);
    const R am = (eml < emr) ? eml : emr;
    const R ap = (epl > epr) ? epl : epr;
    note_lambda(lam, am, ap);

    // HLL state and the eqn (55) inputs
    const R ida = rcp(ap - am);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const R uh = (ap*Ur[i] - am*Ul[i] + (Fl[i] - Fr[i])) * ida;
      Uh[i] = (RN)uh;
      Ph[i] = (RN)(R(0.5)*(Pl[i] + Pr[i]));
    }
    Uhll_tau = (R)Uh[tau]; Uhll_ddd = (R)Uh[ddd]; Uhll_Sx = (R)Uh[Sx];
    Fhll_tau = (ap*Fl[tau] - am*Fr[tau] + ap*am*(Ur[tau] - Ul[tau])) * ida;
    Fhll_ddd = (ap*Fl[ddd] - am*Fr[ddd] + ap*am*(Ur[ddd] - Ul[ddd])) * ida;
    Fhll_Sx = (ap*Fl[Sx] - am*Fr[Sx] + ap*am*(Ur[Sx] - Ul[Sx])) * ida;

    // HlldEquation48 constructor, :292-317: jumps across the fast waves, eqn (12)
    Ul[tau] += Ul[ddd]; Fl[tau] += Fl[ddd];
    Ur[tau] += Ur[ddd]; Fr[tau] += Fr[ddd];
    e.am = am; e.ap = ap;
    e.UlBx = Ul[Bx]; e.UrBx = Ur[Bx];
    e.LE = am*Ul[tau] - Fl[tau]; e.RE = ap*Ur[tau] - Fr[tau];
    e.LD = am*Ul[ddd] - Fl[ddd]; e.RD = ap*Ur[ddd] - Fr[ddd];
    e.LB0 = am*Ul[Bx] - Fl[Bx]; e.RB0 = ap*Ur[Bx] - Fr[Bx];
    e.LB1 = am*Ul[By] - Fl[By]; e.RB1 = ap*Ur[By] - Fr[By];
    e.LB2 = am*Ul[Bz] - Fl[Bz]; e.RB2 = ap*Ur[Bz] - Fr[Bz];
    e.Lm0 = am*Ul[Sx] - Fl[Sx]; e.Rm0 = ap*Ur[Sx] - Fr[Sx];
    e.Lm1 = am*Ul[Sy] - Fl[Sy]; e.Rm1 = ap*Ur[Sy] - Fr[Sy];
    e.Lm2 = am*Ul[Sz] - Fl[Sz]; e.Rm2 = ap*Ur[Sz] - Fr[Sz];
  }
  int c2p_first;
  int hll_c2p_failed = 1;
  if constexpr (T::nested_c2p_from_average) {
    // mixed form: the HLL state lies between the two reconstructed states, so anton2dzw
    // starts from their average (the cascade's second guess) and the cascade runs only
    // if that fails; the result seeds the FP32 p* Newton solve
    if constexpr (sizeof(RN) == 4) {
      if (!check_cons(Uh)) hll_c2p_failed = hll_c2p_f32(Uh, Ph, (float)gamma, (float)pfloor, Ph, cnt.c2p_iters);
    }
    else if (!check_cons(Uh)) {
      C2PState s;
      c2p_new_state(s, Uh, (double)gamma, pfloor);
      double Z, W;
      c2p_starting_prim(s, Ph, Z, W);
      double Pg[8];
      hll_c2p_failed = c2p_anton2dzw(s, Z, W, Pg, cnt.c2p_iters);
      if (!hll_c2p_failed) {
#pragma unroll
        for (int i = 0; i < 8; ++i) Ph[i] = Pg[i];
      }
    }
  }
  if (hll_c2p_failed) {
    if constexpr (FAST) return kHlldDefer;
    if constexpr (sizeof(RN) == 4) {
      // the reference's cascade, in FP64, from the average state
      double Uc[8], Pc[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        Uc[i] = (double)Uh[i];
        Pc[i] = (double)(R(0.5)*(Pl[i] + Pr[i]));
      }
      hll_c2p_failed = cons_to_prim<Shape<R>::unroll>(Uc, Pc, (double)gamma, pfloor, true, cnt.c2p_iters, c2p_first);
#pragma unroll
      for (int i = 0; i < 8; ++i) Ph[i] = (RN)Pc[i];
    }
    else {
      hll_c2p_failed = cons_to_prim<Shape<R>::unroll>(Uh, Ph, (double)gamma, pfloor, true, cnt.c2p_iters, c2p_first);
    }
  }

  bool success = false;
  R Fx[8];
  if (!hll_c2p_failed) {
    // EstimateSolution(0): total pressure of the inverted HLL state, :543-555
    R est0;
    if constexpr (sizeof(R) == 8) {
      const double *P = Ph;
      const double V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
      const double B2 = P[Bx]*P[Bx] + P[By]*P[By] + P[Bz]*P[Bz];
      const double Bv = P[Bx]*P[vx] + P[By]*P[vy] + P[Bz]*P[vz];
      const double W2 = 1.0 / (1.0 - V2);
      const double W = sqrt(W2);
      const double b0 = W*Bv;
      const double b2 = (B2 + b0*b0) / W2;
      est0 = (R)(P[pre] + 0.5*b2);
    }
    else {
      // the same total pressure in the working precision: b2 = (B2 + W^2 Bv^2) / W^2
      // = B2 (1 - V2) + Bv^2, with no square root or division
      R P[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) P[i] = (R)Ph[i];
      const R V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
      const R B2 = P[Bx]*P[Bx] + P[By]*P[By] + P[Bz]*P[Bz];
      const R Bv = P[Bx]*P[vx] + P[By]*P[vy] + P[Bz]*P[vz];
      est0 = P[pre] + R(0.5)*(B2*(R(1) - V2) + Bv*Bv);
    }
    const bool small_bn = dfabs(Pl[Bx]) < T::small_bn || dfabs(Pr[Bx]) < T::small_bn;
    // EstimateSolution(1), eqn (55), :558-574 (the += / -= pair is kept: it rounds)
    auto estimate1 = [&]() {
      Fhll_tau += Fhll_ddd;
      Uhll_tau += Uhll_ddd;
      const R b = Uhll_tau - Fhll_Sx;
      const R c = Fhll_tau * Uhll_Sx - Uhll_tau * Fhll_Sx;
      const R p1 = dfabs((-b + dsqrt(b*b - R(4)*c))/R(2));
      const R p2 = dfabs((-b - dsqrt(b*b - R(4)*c))/R(2));
      Fhll_tau -= Fhll_ddd;
      Uhll_tau -= Uhll_ddd;
      return (p1 < p2) ? p1 : p2;
    };
    if constexpr (FAST) {
      // the small-normal-field branch (no Newton solve: the flux at EstimateSolution(1),
      // :226-231) is as cheap as the common path; its later attempts repeat it
      if (small_bn) success = e.flux(estimate1(), Pl, Pr, gamma, Fx);
      else {
        R p_star = est0;
        success = e.newton(p_star, cnt.newton_iters, T::newton_fast_max_it) && e.flux(p_star, Pl, Pr, gamma, Fx);
      }
      if (!success) return kHlldDefer;
    }
#pragma unroll Shape<R>::unroll
    for (int attempt = 0; attempt < (FAST ? 0 : 3) && !success; ++attempt) {
      R p_star;
      if (small_bn || attempt == 1) p_star = estimate1();
      else p_star = attempt == 0 ? est0 : R(0);
      if (small_bn) success = e.flux(p_star, Pl, Pr, gamma, Fx);
      else success = e.newton(p_star, cnt.newton_iters) && e.flux(p_star, Pl, Pr, gamma, Fx);
    }
    if (success) {
      ++cnt.hlld_ok;
      // rotate the x-normal flux back to the lab frame
      f[ddd] = Fx[ddd];
      f[tau] = Fx[tau];
      f[Pm::S1] = Fx[Sx];
      f[Pm::S2] = Fx[Sy];
      f[Pm::S3] = Fx[Sz];
      f[Pm::B1] = Fx[Bx];
      f[Pm::B2] = Fx[By];
      f[Pm::B3] = Fx[Bz];
      return 0;
    }
  }
  if constexpr (FAST) return kHlldDefer;
  ++cnt.hlld_fail;
  // HLLC of the same (normal-field-averaged) states in the lab frame
  R ql[8], qr[8];
  ql[rho] = Pl[rho]; qr[rho] = Pr[rho];
  ql[pre] = Pl[pre]; qr[pre] = Pr[pre];
  ql[Pm::v1] = Pl[vx]; qr[Pm::v1] = Pr[vx];
  ql[Pm::v2] = Pl[vy]; qr[Pm::v2] = Pr[vy];
  ql[Pm::v3] = Pl[vz]; qr[Pm::v3] = Pr[vz];
  ql[Pm::B1] = Pl[Bx]; qr[Pm::B1] = Pr[Bx];
  ql[Pm::B2] = Pl[By]; qr[Pm::B2] = Pr[By];
  ql[Pm::B3] = Pl[Bz]; qr[Pm::B3] = Pr[Bz];
  return riemann_hllc<R, DIM>(ql, qr, f, gamma, lam);
}

}  // namespace maragpu
