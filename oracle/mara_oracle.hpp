/* ORACLE — test infrastructure only (the checker, never the product).
 *
 * CPU restatement of Mara's plain FP64 relativistic-MHD update, per point.
 * Every function below follows one reference routine and keeps its order of
 * floating-point operations, so that with the reference's own compiler flags
 * (x86-64 SSE2, no FMA contraction) the results are bitwise identical to the
 * compiled reference. The scalar type is a template parameter so that the same
 * code, instantiated with oracle::Counted, yields operation counts.
 *
 * The reference keeps its solver state in file-scope statics (Mara/rmhd-c2p.c:47-54,
 * Mara/riemann_hlld-rmhd.cpp:23-25, Mara/weno.c:10-12, RiemannSolver::MaxLambda);
 * here the same state lives in per-call structs so the grid drivers can run
 * zones in parallel.
 */
#ifndef ORACLE_MARA_ORACLE_HPP
#define ORACLE_MARA_ORACLE_HPP

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdint>

namespace oracle {

using std::fabs;
using std::sqrt;
using std::pow;
using std::acos;
using std::cos;
using std::sin;
using std::isnan;
using std::isinf;

enum { ddd, tau, Sx, Sy, Sz, Bx, By, Bz };  // conserved, Mara/rmhd.hpp:19
enum { rho, pre, vx, vy, vz };              // primitive,  Mara/rmhd.hpp:20

// Mara/rmhd-c2p.h:10-20
enum C2PError {
  C2P_SUCCESS, C2P_CONS_CONTAINS_NAN, C2P_CONS_NEGATIVE_DENSITY,
  C2P_CONS_NEGATIVE_ENERGY, C2P_PRIM_CONTAINS_NAN, C2P_PRIM_NEGATIVE_PRESSURE,
  C2P_PRIM_NEGATIVE_RESTMASS, C2P_PRIM_SUPERLUMINAL, C2P_MAXITER
};

enum Recon { RECON_PCM = 0, RECON_PLM = 1, RECON_WENO5 = 2 };
enum Smooth { IS_JS96 = 0, IS_B08 = 1, IS_SZ10 = 2 };
enum Riemann { RIEMANN_HLL = 0, RIEMANN_HLLC = 1, RIEMANN_HLLD = 2 };

struct PhysParams {
  double gamma = 1.4;     // AdiabaticEos::Gamma
  double pfloor = -1.0;   // rmhd_c2p_set_pressure_floor, disabled if <= 0
  double theta = 2.0;     // weno.c plm_theta
  int IS = IS_JS96;       // weno.c IS_mode
  double sz10A = 50.0;    // weno.c shenzha10_A
  // Flop counting only: start the HLL state's c2p with anton2dzw from the average of the
  // two states, and run the cascade only if that fails, as the device's HLLD does (the
  // reference runs the cascade from the cons-based estimate). Parity runs keep false.
  bool nested_c2p_from_average = false;
};

// Accumulators the reference keeps in statics; merged by the grid driver.
template <class R> struct Tally {
  R max_lambda = R(0.0);               // RiemannSolver::MaxLambda
  std::int64_t hlld_success = 0;       // HlldRmhdRiemannSolver::NumberOfSuccess
  std::int64_t hlld_failure = 0;       // HlldRmhdRiemannSolver::NumberOfFailure
  std::int64_t c2p_iterations = 0;     // Newton iterations summed over solves
};

template <class R> inline void note_lambda(Tally<R> &t, R am, R ap)
{
  // Mara/riemann_hlld-rmhd.cpp:203-204 (same in hll, hllc)
  R ml = (fabs(am) < fabs(ap)) ? fabs(ap) : fabs(am);
  if (t.max_lambda < ml) t.max_lambda = ml;
}

/* -------------------------------------------------------------------------
 * Gamma-law EOS and the closed forms, Mara/eos.cpp:9-38, Mara/rmhd.cpp:133-163
 * ------------------------------------------------------------------------- */
template <class R> inline int c2p_check_cons(const R *U)
{
  // Mara/rmhd-c2p.c:153-164
  if (U[ddd] < 0.0) return C2P_CONS_NEGATIVE_DENSITY;
  if (U[tau] < 0.0) return C2P_CONS_NEGATIVE_ENERGY;
  for (int i = 0; i < 8; ++i) if (isnan(U[i])) return C2P_CONS_CONTAINS_NAN;
  return C2P_SUCCESS;
}
template <class R> inline int c2p_check_prim(const R *P)
{
  // Mara/rmhd-c2p.c:165-178
  const R v2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
  if (v2 >= 1.0) return C2P_PRIM_SUPERLUMINAL;
  if (P[pre] < 0.0) return C2P_PRIM_NEGATIVE_PRESSURE;
  if (P[rho] < 0.0) return C2P_PRIM_NEGATIVE_RESTMASS;
  for (int i = 0; i < 8; ++i) if (isnan(P[i])) return C2P_PRIM_CONTAINS_NAN;
  return C2P_SUCCESS;
}

template <class R> inline int prim_to_cons(const R *P, R *U, double gamma)
{
  // Mara/rmhd.cpp:133-163; e = Internal(rho, T=p)/rho with Internal = T/(Gamma-1)
  const R V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
  const R B2 = P[Bx]*P[Bx] + P[By]*P[By] + P[Bz]*P[Bz];
  const R Bv = P[Bx]*P[vx] + P[By]*P[vy] + P[Bz]*P[vz];
  const R W2 = 1.0 / (1.0 - V2);
  const R W = sqrt(W2);
  const R b0 = W*Bv;
  const R b2 = (B2 + b0*b0) / W2;
  const R bx = (P[Bx] + b0 * W*P[vx]) / W;
  const R by = (P[By] + b0 * W*P[vy]) / W;
  const R bz = (P[Bz] + b0 * W*P[vz]) / W;
  const R e = (P[pre] / (gamma - 1)) / P[rho];
  const R e_ = e + 0.5 * b2 / P[rho];
  const R p_ = P[pre] + 0.5 * b2;
  const R h_ = 1 + e_ + p_ / P[rho];

  U[ddd] = P[rho] * W;
  U[tau] = P[rho] * h_ * W2 - p_ - b0*b0 - U[ddd];
  U[Sx ] = P[rho] * h_ * W2 * P[vx] - b0*bx;
  U[Sy ] = P[rho] * h_ * W2 * P[vy] - b0*by;
  U[Sz ] = P[rho] * h_ * W2 * P[vz] - b0*bz;
  U[Bx ] = P[Bx];
  U[By ] = P[By];
  U[Bz ] = P[Bz];

  if (V2 >= 1.0) return C2P_PRIM_SUPERLUMINAL;
  return c2p_check_cons(U);
}

/* Mara/quartic.cpp:119-148, QuarticEquation::solve_cubic_equation */
template <class R> inline void solve_cubic(R c3, R c2, R c1, R c0, R &x1, R &x2, R &x3, int &nr)
{
  R a2 = c2/c3;
  R a1 = c1/c3;
  R a0 = c0/c3;

  R q = a1/3.0 - a2*a2/9.0;
  R r = (a1*a2 - 3.0*a0)/6.0 - a2*a2*a2 / 27.0;
  R delta = q*q*q + r*r;

  if (delta > 0.0) {
    R s1 = r + sqrt(delta);
    s1 = (s1 >= 0.0) ? pow(s1, R(1./3.)) : -pow(-s1, R(1./3.));
    R s2 = r - sqrt(delta);
    s2 = (s2 >= 0.0) ? pow(s2, R(1./3.)) : -pow(-s2, R(1./3.));
    x1 = (s1+s2) - a2/3.0;
    x2 = x3 = -0.5 * (s1+s2) - a2/3.0;
    nr = 1;
  }
  else if (delta < 0.0) {
    R theta = acos(r/sqrt(-q*q*q)) / 3.0;
    R costh = cos(theta);
    R sinth = sin(theta);
    R sq = sqrt(-q);
    x1 = 2.0*sq*costh - a2/3.0;
    x2 = -sq*costh - a2/3.0 - sqrt(R(3.)) * sq * sinth;
    x3 = -sq*costh - a2/3.0 + sqrt(R(3.)) * sq * sinth;
    nr = 3;
  }
  else {
    R s = (r >= 0.0) ? pow(r, R(1./3.)) : -pow(-r, R(1./3.));
    x1 = 2.0*s - a2/3.0;
    x2 = x3 = -s - a2/3.0;
    nr = 3;
  }
}

/* Mara/quartic.cpp:28-117, QuarticEquation::SolveOld (what Solve calls, :22-26) */
template <class R> inline int solve_quartic(R d4, R d3, R d2, R d1, R d0, R *roots)
{
  R r1, r2, r3, r4;
  int nr12, nr34;

  R a3 = d3/d4;
  R a2 = d2/d4;
  R a1 = d1/d4;
  R a0 = d0/d4;

  R au2 = -a2;
  R au1 = (a1*a3 - 4.0*a0);
  R au0 = 4.0*a0*a2 - a1*a1 - a0*a3*a3;

  R x1, x2, x3;
  int nr;
  solve_cubic(R(1.0), au2, au1, au0, x1, x2, x3, nr);

  R u1;
  if (nr == 1) u1 = x1;
  else u1 = (x1 > x3) ? x1 : x3;

  R R2 = 0.25*a3*a3 + u1 - a2;
  R Rr = (R2 > 0.0) ? sqrt(R2) : R(0.0);

  R D2, E2;
  if (Rr != 0.0) {
    R foo1 = 0.75*a3*a3 - R2 - 2.0*a2;
    R foo2 = 0.25*(4.0*a3*a2 - 8.0*a1 - a3*a3*a3) / Rr;
    D2 = foo1 + foo2;
    E2 = foo1 - foo2;
  }
  else {
    R foo1 = 0.75*a3*a3 - 2.0*a2;
    R foo2 = 2.0 * sqrt(u1*u1 - 4.0*a0);
    D2 = foo1 + foo2;
    E2 = foo1 - foo2;
  }

  if (D2 >= 0.0) {
    R D = sqrt(D2);
    r1 = -0.25*a3 + 0.5*Rr - 0.5*D;
    r2 = -0.25*a3 + 0.5*Rr + 0.5*D;
    nr12 = 2;
  }
  else {
    r1 = r2 = -0.25*a3 + 0.5*Rr;
    nr12 = 0;
  }
  if (E2 >= 0.0) {
    R E = sqrt(E2);
    r3 = -0.25*a3 - 0.5*Rr - 0.5*E;
    r4 = -0.25*a3 - 0.5*Rr + 0.5*E;
    nr34 = 2;
  }
  else {
    r3 = r4 = -0.25*a3 - 0.5*Rr;
    nr34 = 0;
  }
  int i = 0;
  if (nr12 != 0) { roots[i++] = r1; roots[i++] = r2; }
  if (nr34 != 0) { roots[i++] = r3; roots[i++] = r4; }
  (void)r1; (void)r3;
  return nr12 + nr34;
}

/* Mara/rmhd.cpp:164-265, Rmhd::FluxAndEigenvalues; dim is 1, 2 or 3 */
template <class R> inline void flux_and_eigenvalues(const R *U, const R *P, R *F, R *ap, R *am,
                                                    int dim, double gamma)
{
  const R V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
  const R B2 = P[Bx]*P[Bx] + P[By]*P[By] + P[Bz]*P[Bz];
  const R Bv = P[Bx]*P[vx] + P[By]*P[vy] + P[Bz]*P[vz];
  const R W2 = 1.0 / (1.0 - V2);
  const R W = sqrt(W2);
  const R b0 = W*Bv;
  const R b2 = (B2 + b0*b0) / W2;
  const R bx = (P[Bx] + b0 * W*P[vx]) / W;
  const R by = (P[By] + b0 * W*P[vy]) / W;
  const R bz = (P[Bz] + b0 * W*P[vz]) / W;
  const R e = (P[pre] / (gamma - 1)) / P[rho];
  const R p_ = P[pre] + 0.5 * b2;
  const R h = 1 + e + P[pre]/P[rho];

  switch (dim) {
  case 1:
    F[ddd] = U[ddd] * P[vx];
    F[tau] = U[tau] * P[vx] - b0*P[Bx] / W + p_*P[vx];
    F[Sx ] = U[Sx] * P[vx] - bx*P[Bx] / W + p_;
    F[Sy ] = U[Sy] * P[vx] - by*P[Bx] / W;
    F[Sz ] = U[Sz] * P[vx] - bz*P[Bx] / W;
    F[Bx ] = 0.0;
    F[By ] = P[vx]*P[By] - P[vy]*P[Bx];
    F[Bz ] = P[vx]*P[Bz] - P[vz]*P[Bx];
    break;
  case 2:
    F[ddd] = U[ddd] * P[vy];
    F[tau] = U[tau] * P[vy] - b0*P[By] / W + p_*P[vy];
    F[Sx ] = U[Sx] * P[vy] - bx*P[By] / W;
    F[Sy ] = U[Sy] * P[vy] - by*P[By] / W + p_;
    F[Sz ] = U[Sz] * P[vy] - bz*P[By] / W;
    F[Bx ] = P[vy]*P[Bx] - P[vx]*P[By];
    F[By ] = 0.0;
    F[Bz ] = P[vy]*P[Bz] - P[vz]*P[By];
    break;
  case 3:
    F[ddd] = U[ddd] * P[vz];
    F[tau] = U[tau] * P[vz] - b0*P[Bz] / W + p_*P[vz];
    F[Sx ] = U[Sx] * P[vz] - bx*P[Bz] / W;
    F[Sy ] = U[Sy] * P[vz] - by*P[Bz] / W;
    F[Sz ] = U[Sz] * P[vz] - bz*P[Bz] / W + p_;
    F[Bx ] = P[vz]*P[Bx] - P[vx]*P[Bz];
    F[By ] = P[vz]*P[By] - P[vy]*P[Bz];
    F[Bz ] = 0.0;
    break;
  }
  if (ap == nullptr || am == nullptr) return;

  R vi = 0, bi = 0;
  switch (dim) {
  case 1: vi = P[vx]; bi = bx; break;
  case 2: vi = P[vy]; bi = by; break;
  case 3: vi = P[vz]; bi = bz; break;
  }
  // SoundSpeed2Sr(rho, T=p): Gamma*T/(rho + T/(Gamma-1) + T), Mara/eos.cpp:23-27
  const R rhoh = P[rho] + P[pre] / (gamma - 1) + P[pre];
  const R cs2 = gamma*P[pre] / rhoh;
  const R W4 = W2*W2;
  const R v2 = vi*vi;
  const R v3 = vi*v2;
  const R v4 = vi*v3;

  const R K = P[rho]*h * (1.0/cs2-1) * W4;
  const R L = -(P[rho]*h + b2/cs2) * W2;

  const R A4 = K - L - b0*b0;
  const R A3 = -4*K*vi + L*vi*2 + 2*b0*bi;
  const R A2 = 6*K*v2 + L*(1.0-v2) + b0*b0 - bi*bi;
  const R A1 = -4*K*v3 - L*vi*2 - 2*b0*bi;
  const R A0 = K*v4 + L*v2 + bi*bi;

  R roots[4];
  int nr = solve_quartic(A4, A3, A2, A1, A0, roots);

  *ap = -1.0;
  *am = 1.0;
  for (int i = 0; i < nr; ++i) {
    if (roots[i] > *ap) *ap = roots[i];
    if (roots[i] < *am) *am = roots[i];
  }
  if (fabs(*ap) > 1.0 || fabs(*am) > 1.0) {
    *ap = 1.0;
    *am = -1.0;
  }
}

/* -------------------------------------------------------------------------
 * Conserved -> primitive, Mara/rmhd-c2p.c (gamma-law branch of Mara/rmhd.cpp:59-132)
 * ------------------------------------------------------------------------- */
template <class R> struct C2P {
  double gamma, gamf, pfloor;
  R D, Tau, S2, B2, BS, BS2;
  R Cons[8];
  R Z_start, W_start;
  int iterations;

  C2P(double gamma_, double pfloor_) : gamma(gamma_), gamf((gamma_ - 1.0) / gamma_),
    pfloor(pfloor_), iterations(0) {}

  void new_state(const R *U)  // :82-93
  {
    D = U[ddd];
    Tau = U[tau];
    S2 = U[Sx]*U[Sx] + U[Sy]*U[Sy] + U[Sz]*U[Sz];
    B2 = U[Bx]*U[Bx] + U[By]*U[By] + U[Bz]*U[Bz];
    BS = U[Bx]*U[Sx] + U[By]*U[Sy] + U[Bz]*U[Sz];
    BS2 = BS*BS;
    for (int i = 0; i < 8; ++i) Cons[i] = U[i];
  }
  void estimate_from_cons()  // :97-101
  {
    Z_start = sqrt(S2 + D*D);
    W_start = Z_start / D;
  }
  void set_starting_prim(const R *P)  // :108-119
  {
    const R V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
    const R W2 = 1.0 / (1.0 - V2);
    const R e = P[pre] / (P[rho] * (gamma - 1.0));
    const R h = 1.0 + e + P[pre] / P[rho];
    Z_start = P[rho] * h * W2;
    W_start = sqrt(W2);
  }
  int reconstruct_prim(R Z, R W, R *Pout) const  // :121-150
  {
    R P[8];
    const R b0 = BS * W / Z;
    P[rho] = D/W;
    P[pre] = (D/W) * (Z/(D*W) - 1.0) * gamf;
    P[vx ] = (Cons[Sx] + b0*Cons[Bx]/W) / (Z+B2);
    P[vy ] = (Cons[Sy] + b0*Cons[By]/W) / (Z+B2);
    P[vz ] = (Cons[Sz] + b0*Cons[Bz]/W) / (Z+B2);
    P[Bx ] = Cons[Bx];
    P[By ] = Cons[By];
    P[Bz ] = Cons[Bz];
    if (P[pre] < pfloor && pfloor > 0.0) P[pre] = pfloor;
    int err = c2p_check_prim(P);
    if (err) return err;
    for (int i = 0; i < 8; ++i) Pout[i] = P[i];
    return C2P_SUCCESS;
  }
  int solve_anton2dzw(R *P)  // :182-253, Anton et al. (2006) eqs. 84-85
  {
    int bad = c2p_check_cons(Cons);
    if (bad) return bad;
    iterations = 0;
    R error = 1.0;
    R W = W_start;
    R Z = Z_start;
    while (error > 1e-12) {
      const R Z2 = Z*Z;
      const R Z3 = Z*Z2;
      const R W2 = W*W;
      const R W3 = W*W2;
      const R Pre = (D/W) * (Z/(D*W) - 1.0) * gamf;

      const R df0dZ = 2*(B2+Z)*(BS2*W2 + (W2-1)*Z3) / (W2*Z3);
      const R df0dW = 2*(B2+Z)*(B2+Z) / W3;
      const R df1dZ = 1.0 + BS2/Z3 - gamf/W2;
      const R df1dW = B2/W3 + (2*Z - D*W)/W3 * gamf;

      R f0 = -S2 + (Z+B2)*(Z+B2)*(W2-1)/W2 - (2*Z+B2)*BS2/Z2;
      R f1 = -Tau + Z+B2 - Pre - 0.5*B2/W2 - 0.5*BS2/Z2 - D;

      const R J0 = df0dZ, J1 = df0dW, J2 = df1dZ, J3 = df1dW;
      const R det = J0*J3 - J1*J2;
      const R G0 = J3/det, G1 = -J1/det, G2 = -J2/det, G3 = J0/det;

      const R dZ = -(G0*f0 + G1*f1);
      const R dW = -(G2*f0 + G3*f1);

      R Z_new = Z + dZ;
      R W_new = W + dW;
      Z_new = (Z_new > 0.0) ? Z_new : -Z_new;
      Z_new = (Z_new < 1e20) ? Z_new : Z;
      W_new = (W_new > 1.0) ? W_new : R(1.0);
      W_new = (W_new < 1e12) ? W_new : R(1e12);
      Z = Z_new;
      W = W_new;

      error = fabs(dZ/Z) + fabs(dW/W);
      ++iterations;
      if (iterations == 250) return C2P_MAXITER;
    }
    return reconstruct_prim(Z, W, P);
  }
  int solve_noble1dw(R *P)  // :261-326, Noble et al. (2006) eq. 29
  {
    int bad = c2p_check_cons(Cons);
    if (bad) return bad;
    iterations = 0;
    R error = 1.0;
    R Z = Z_start;
    while (error > 1e-12) {
      const R Z2 = Z*Z;
      const R Z3 = Z*Z2;
      const R a = S2*Z2 + BS2*(B2 + 2*Z);
      const R b = (B2 + Z)*(B2 + Z)*Z2;
      const R ap = 2*(S2*Z + BS2);
      const R bp = 2*Z*(B2 + Z)*(B2 + 2*Z);
      const R V2 = a / b;
      const R W2 = 1.0 / (1.0 - V2);
      const R W = sqrt(W2);
      const R W3 = W*W2;
      const R Pre = (D/W) * (Z/(D*W) - 1.0) * gamf;

      const R dv2dZ = (ap*b - bp*a) / (b*b);
      const R delPdelZ = gamf/W2;
      const R delPdelW = gamf*(D/W2 - 2*Z/W3);
      const R dWdv2 = 0.5*W3;
      const R dPdZ = delPdelW * dWdv2 * dv2dZ + delPdelZ;

      const R f = Tau + D - 0.5*B2*(1+V2) + 0.5*BS2/Z2 - Z + Pre;
      const R g = -0.5*B2*dv2dZ - BS2/Z3 - 1.0 + dPdZ;
      const R dZ = -f/g;

      R Z_new = Z + dZ;
      Z_new = (Z_new > 0.0) ? Z_new : -Z_new;
      Z_new = (Z_new < 1e20) ? Z_new : Z;
      Z = Z_new;

      error = fabs(dZ/Z);
      ++iterations;
      if (iterations == 250) return C2P_MAXITER;
    }
    const R Z2 = Z*Z;
    const R a = S2*Z2 + BS2*(B2 + 2*Z);
    const R b = (B2 + Z)*(B2 + Z)*Z2;
    const R V2 = a / b;
    const R W2 = 1.0 / (1.0 - V2);
    const R W = sqrt(W2);
    return reconstruct_prim(Z, W, P);
  }
};

/* Rmhd::ConsToPrim gamma-law cascade, Mara/rmhd.cpp:59-69,104-131.
 * P holds the previous primitive state on entry (used as the second guess) and
 * is only overwritten on success. `iters` accumulates Newton iterations. */
template <class R> inline int cons_to_prim(const R *U, R *P, double gamma, double pfloor,
                                           std::int64_t *iters)
{
  int error = c2p_check_cons(U);
  if (error) return error;
  error = 1;
  C2P<R> s(gamma, pfloor);
  s.new_state(U);
  if (error) {
    s.estimate_from_cons();
    error = s.solve_anton2dzw(P);
    if (iters) *iters += s.iterations;
  }
  if (error) {
    s.set_starting_prim(P);
    error = s.solve_anton2dzw(P);
    if (iters) *iters += s.iterations;
  }
  if (error) {
    s.estimate_from_cons();
    error = s.solve_noble1dw(P);
    if (iters) *iters += s.iterations;
  }
  if (error) {
    s.set_starting_prim(P);
    error = s.solve_noble1dw(P);
    if (iters) *iters += s.iterations;
  }
  return error;
}

/* -------------------------------------------------------------------------
 * Reconstruction, Mara/weno.c
 * ------------------------------------------------------------------------- */
template <class R> inline R min3(const R *x)
{
  R x01 = x[0] < x[1] ? x[0] : x[1];
  return x01 < x[2] ? x01 : x[2];
}
template <class R> inline R max3(const R *x)
{
  R x01 = x[0] > x[1] ? x[0] : x[1];
  return x01 > x[2] ? x01 : x[2];
}
template <class R> inline int sgn(R x) { return (x > 0.0) - (x < 0.0); }

template <class R> inline R plm_minmod(R ul, R u0, R ur, double theta)  // weno.c:85-92
{
  const R a = theta * (u0 - ul);
  const R b = 0.5 * (ur - ul);
  const R c = theta * (ur - u0);
  const R fabc[3] = { fabs(a), fabs(b), fabs(c) };
  int s1 = sgn(a) + sgn(b);
  int aabs = s1 < 0 ? -s1 : s1;
  return 0.25*aabs*(sgn(a) + sgn(c))*min3(fabc);
}
template <class R> inline R plm(const R *v, double sg, double theta)  // weno.c:93-96
{
  return v[0] + sg*0.5*plm_minmod(v[-1], v[0], v[1], theta);
}

enum WenoOp { PLM_C2L, PLM_C2R, WENO5_FD_C2R, WENO5_FD_C2L, WENO5_FV_C2R, WENO5_FV_C2L,
              WENO5_FV_C2A, WENO5_FV_A2C };

struct WenoCoeffs { double c[3][3]; double d[3]; };
// weno.c:15-41
static const WenoCoeffs kC2L_FV = {{{15./8., -5./4., 3./8.}, {3./8., 3./4., -1./8.}, {-1./8., 3./4., 3./8.}},
                                   {1./16., 5./8., 5./16.}};
static const WenoCoeffs kC2R_FV = {{{3./8., 3./4., -1./8.}, {-1./8., 3./4., 3./8.}, {3./8., -5./4., 15./8.}},
                                   {5./16., 5./8., 1./16.}};
static const WenoCoeffs kA2C_FV = {{{23./24., 1./12., -1./24.}, {-1./24., 13./12., -1./24.}, {-1./24., 1./12., 23./24.}},
                                   {-9./80., 49./40., -9./80.}};
static const WenoCoeffs kC2A_FV = {{{25./24., -1./12., 1./24.}, {1./24., 11./12., 1./24.}, {1./24., -1./12., 25./24.}},
                                   {-17./240., 137./120., -17./240.}};
static const WenoCoeffs kC2R_FD = {{{1./3., 5./6., -1./6.}, {-1./6., 5./6., 1./3.}, {1./3., -7./6., 11./6.}},
                                   {0.3, 0.6, 0.1}};
static const WenoCoeffs kC2L_FD = {{{11./6., -7./6., 1./3.}, {1./3., 5./6., -1./6.}, {-1./6., 5./6., 1./3.}},
                                   {0.1, 0.6, 0.3}};

template <class R> inline R SQU(R x) { return x*x; }

template <class R> inline R weno5(const R *v, const WenoCoeffs &k, int IS, double sz10A)  // weno.c:97-154
{
  const double (*c)[3] = k.c;
  const double *d = k.d;
  double eps = 1e-6, eps_prime = 1e-6;
  const R vs[3] = {
    c[0][0]*v[+0] + c[0][1]*v[+1] + c[0][2]*v[2],
    c[1][0]*v[-1] + c[1][1]*v[+0] + c[1][2]*v[1],
    c[2][0]*v[-2] + c[2][1]*v[-1] + c[2][2]*v[0],
  };
  R B[3] = {
    (13./12.)*SQU(1*v[+0] - 2*v[+1] + 1*v[+2]) + (1./4.)*SQU(3*v[+0] - 4*v[+1] + 1*v[+2]),
    (13./12.)*SQU(1*v[-1] - 2*v[+0] + 1*v[+1]) + (1./4.)*SQU(1*v[-1] - 0*v[+0] - 1*v[+1]),
    (13./12.)*SQU(1*v[-2] - 2*v[-1] + 1*v[+0]) + (1./4.)*SQU(1*v[-2] - 4*v[-1] + 3*v[+0])
  };
  R w[3];
  if (IS == IS_B08) {
    eps = eps_prime = 1e-14;
    const R tau5 = fabs(B[0] - B[2]);
    w[0] = d[0] * (1.0 + (tau5 / (B[0] + eps)));
    w[1] = d[1] * (1.0 + (tau5 / (B[1] + eps)));
    w[2] = d[2] * (1.0 + (tau5 / (B[2] + eps)));
  }
  else if (IS == IS_SZ10) {
    eps = 1e-6;
    eps_prime = 1e-10;
    const double A = sz10A;
    const R minB = min3(B), maxB = max3(B);
    const R R0 = minB / (maxB + eps_prime);
    B[0] = R0*A*minB + B[0];
    B[1] = R0*A*minB + B[1];
    B[2] = R0*A*minB + B[2];
    w[0] = d[0] / SQU(eps_prime + B[0]);
    w[1] = d[1] / SQU(eps_prime + B[1]);
    w[2] = d[2] / SQU(eps_prime + B[2]);
  }
  else {
    eps = eps_prime = 1e-6;
    w[0] = d[0] / SQU(eps_prime + B[0]);
    w[1] = d[1] / SQU(eps_prime + B[1]);
    w[2] = d[2] / SQU(eps_prime + B[2]);
  }
  (void)eps;
  const R wtot = w[0] + w[1] + w[2];
  return (w[0]*vs[0] + w[1]*vs[1] + w[2]*vs[2])/wtot;
}

template <class R> inline R reconstruct(const R *v, int op, const PhysParams &pp)  // weno.c:58-71
{
  switch (op) {
  case WENO5_FD_C2L: return weno5(v, kC2L_FD, pp.IS, pp.sz10A);
  case WENO5_FD_C2R: return weno5(v, kC2R_FD, pp.IS, pp.sz10A);
  case WENO5_FV_C2L: return weno5(v, kC2L_FV, pp.IS, pp.sz10A);
  case WENO5_FV_C2R: return weno5(v, kC2R_FV, pp.IS, pp.sz10A);
  case WENO5_FV_A2C: return weno5(v, kA2C_FV, pp.IS, pp.sz10A);
  case WENO5_FV_C2A: return weno5(v, kC2A_FV, pp.IS, pp.sz10A);
  case PLM_C2L: return plm(v, -1.0, pp.theta);
  case PLM_C2R: return plm(v, +1.0, pp.theta);
  default: return R(0.0);
  }
}

/* -------------------------------------------------------------------------
 * Riemann solvers
 * ------------------------------------------------------------------------- */

/* Mara/riemann_hll.cpp:20-64 */
template <class R> inline int riemann_hll(const R *pl, const R *pr, R *F, int dim,
                                          const PhysParams &pp, Tally<R> &t)
{
  R epl = 1, epr = 1, eml = 1, emr = 1;
  R Ul[8], Ur[8], Pl[8], Pr[8], Fl[8], Fr[8];
  for (int i = 0; i < 8; ++i) { Pl[i] = pl[i]; Pr[i] = pr[i]; }
  if (prim_to_cons(Pl, Ul, pp.gamma)) return 1;
  if (prim_to_cons(Pr, Ur, pp.gamma)) return 1;
  flux_and_eigenvalues(Ul, Pl, Fl, &epl, &eml, dim, pp.gamma);
  flux_and_eigenvalues(Ur, Pr, Fr, &epr, &emr, dim, pp.gamma);
  R ap = (epl > epr) ? epl : epr;
  R am = (eml < emr) ? eml : emr;
  note_lambda(t, am, ap);
  R F_hll[8];
  for (int i = 0; i < 8; ++i)
    F_hll[i] = (ap*Fl[i] - am*Fr[i] + ap*am*(Ur[i] - Ul[i])) / (ap - am);
  const double s = 0.0;
  if (s <= am) for (int i = 0; i < 8; ++i) F[i] = Fl[i];
  else if (am < s && s <= ap) for (int i = 0; i < 8; ++i) F[i] = F_hll[i];
  else if (ap < s) for (int i = 0; i < 8; ++i) F[i] = Fr[i];
  return 0;
}

inline void axis_perm(int dim, int &B1, int &B2, int &B3, int &S1, int &S2, int &S3,
                      int &v1, int &v2, int &v3)
{
  // Mara/riemann_hlld-rmhd.cpp:164-182 and riemann_hllc-rmhd.cpp:40-58
  switch (dim) {
  case 1: B1 = Bx; B2 = By; B3 = Bz; S1 = Sx; S2 = Sy; S3 = Sz; v1 = vx; v2 = vy; v3 = vz; break;
  case 2: B1 = By; B2 = Bz; B3 = Bx; S1 = Sy; S2 = Sz; S3 = Sx; v1 = vy; v2 = vz; v3 = vx; break;
  default: B1 = Bz; B2 = Bx; B3 = By; S1 = Sz; S2 = Sx; S3 = Sy; v1 = vz; v2 = vx; v3 = vy; break;
  }
}

/* Mara/riemann_hllc-rmhd.cpp:26-178 (Mignone & Bodo 2006) */
template <class R> inline int riemann_hllc(const R *pl, const R *pr, R *F, int dim,
                                           const PhysParams &pp, Tally<R> &t)
{
  R epl, epr, eml, emr;
  R Ul[8], Ur[8], Pl[8], Pr[8], Fl[8], Fr[8];
  int B1, B2, B3, S1, S2, S3, v1, v2, v3;
  axis_perm(dim, B1, B2, B3, S1, S2, S3, v1, v2, v3);
  (void)v2; (void)v3;
  for (int i = 0; i < 8; ++i) { Pl[i] = pl[i]; Pr[i] = pr[i]; }
  Pl[B1] = Pr[B1] = 0.5*(Pl[B1] + Pr[B1]);
  if (prim_to_cons(Pl, Ul, pp.gamma)) return 1;
  if (prim_to_cons(Pr, Ur, pp.gamma)) return 1;
  flux_and_eigenvalues(Ul, Pl, Fl, &epl, &eml, dim, pp.gamma);
  flux_and_eigenvalues(Ur, Pr, Fr, &epr, &emr, dim, pp.gamma);
  Ul[tau] += Ul[ddd]; Fl[tau] += Fl[ddd];
  Ur[tau] += Ur[ddd]; Fr[tau] += Fr[ddd];
  R ap = (epl > epr) ? epl : epr;
  R am = (eml < emr) ? eml : emr;
  note_lambda(t, am, ap);

  R F_hll[8], U_hll[8];
  for (int i = 0; i < 8; ++i) {
    U_hll[i] = (ap*Ur[i] - am*Ul[i] + (Fl[i] - Fr[i])) / (ap - am);
    F_hll[i] = (ap*Fl[i] - am*Fr[i] + ap*am*(Ur[i] - Ul[i])) / (ap - am);
  }
  R Ul_[8], Ur_[8], P_[8], lc;
  if (fabs(U_hll[B1]) > 1e-12) {
    const R B_dot_FB = U_hll[B2]*F_hll[B2] + U_hll[B3]*F_hll[B3];
    const R a = F_hll[tau] - B_dot_FB;
    const R b = -F_hll[S1] - U_hll[tau] + (U_hll[B2]*U_hll[B2] + U_hll[B3]*U_hll[B3]) +
                (F_hll[B2]*F_hll[B2] + F_hll[B3]*F_hll[B3]);
    const R c = U_hll[S1] - B_dot_FB;

    P_[v1] = lc = (fabs(a) < 1e-10) ? -c/b : (-b - sqrt(b*b - 4*a*c)) / (2*a);
    P_[v2] = (U_hll[B2]*P_[v1] - F_hll[B2]) / U_hll[B1];
    P_[v3] = (U_hll[B3]*P_[v1] - F_hll[B3]) / U_hll[B1];
    P_[B1] = U_hll[B1];
    P_[B2] = U_hll[B2];
    P_[B3] = U_hll[B3];

    const R v_dotB_ = P_[v1]*P_[B1] + P_[v2]*P_[B2] + P_[v3]*P_[B3];
    const R v_dotv_ = P_[v1]*P_[v1] + P_[v2]*P_[v2] + P_[v3]*P_[v3];
    const R gm2_ = 1.0 / (1.0 - v_dotv_);

    P_[pre] = F_hll[S1] + P_[B1]*P_[B1]/gm2_ - (F_hll[tau] - P_[B1]*v_dotB_)*P_[v1];

    Ul_[ddd] = (am - Pl[v1]) / (am - P_[v1]) * Ul[ddd];
    Ur_[ddd] = (ap - Pr[v1]) / (ap - P_[v1]) * Ur[ddd];
    Ul_[tau] = (am*Ul[tau] - Ul[S1] + P_[pre]*P_[v1] - v_dotB_*P_[B1]) / (am - P_[v1]);
    Ur_[tau] = (ap*Ur[tau] - Ur[S1] + P_[pre]*P_[v1] - v_dotB_*P_[B1]) / (ap - P_[v1]);
    Ul_[S1] = (Ul_[tau] + P_[pre])*P_[v1] - v_dotB_*P_[B1];
    Ur_[S1] = (Ur_[tau] + P_[pre])*P_[v1] - v_dotB_*P_[B1];
    Ul_[S2] = (-P_[B1]*(P_[B2]/gm2_ + v_dotB_*P_[v2]) + am*Ul[S2] - Fl[S2]) / (am - P_[v1]);
    Ur_[S2] = (-P_[B1]*(P_[B2]/gm2_ + v_dotB_*P_[v2]) + ap*Ur[S2] - Fr[S2]) / (ap - P_[v1]);
    Ul_[S3] = (-P_[B1]*(P_[B3]/gm2_ + v_dotB_*P_[v3]) + am*Ul[S3] - Fl[S3]) / (am - P_[v1]);
    Ur_[S3] = (-P_[B1]*(P_[B3]/gm2_ + v_dotB_*P_[v3]) + ap*Ur[S3] - Fr[S3]) / (ap - P_[v1]);
    Ul_[B1] = Ur_[B1] = P_[B1];
    Ul_[B2] = Ur_[B2] = P_[B2];
    Ul_[B3] = Ur_[B3] = P_[B3];
  }
  else {
    const R a = F_hll[tau];
    const R b = -F_hll[S1] - U_hll[tau];
    const R c = U_hll[S1];
    const R v1_ = lc = (fabs(a) < 1e-10) ? -c/b : (-b - sqrt(b*b - 4*a*c)) / (2*a);
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
  const double s = 0.0;
  if (s <= am) for (int i = 0; i < 8; ++i) F[i] = Fl[i];
  else if (am < s && s <= lc) for (int i = 0; i < 8; ++i) F[i] = Fl[i] + am*(Ul_[i]-Ul[i]);
  else if (lc < s && s <= ap) for (int i = 0; i < 8; ++i) F[i] = Fr[i] + ap*(Ur_[i]-Ur[i]);
  else if (ap < s) for (int i = 0; i < 8; ++i) F[i] = Fr[i];
  F[tau] -= F[ddd];
  return 0;
}

/* HLLD, Mignone, Ugliano & Bodo (2008), Mara/riemann_hlld-rmhd.cpp:143-579.
 * Exceptions of the reference become `false` returns, with the same effect: the
 * attempt is abandoned and the next one tried. */
template <class R> struct HlldEqn {
  const R *Pl, *Pr;
  R *Ul, *Ur, *Fl, *Fr;
  R ap, am;
  R LE, LD, LB[3], Lm[3];
  R RE, RD, RB[3], Rm[3];
  R *P_hll, *U_hll, *F_hll;

  struct Inner { R U[8]; R B[3], v[3]; };

  static R dot(const R *v, const R *w) { return v[0]*w[0] + v[1]*w[1] + v[2]*w[2]; }
  static R dot(const R *v) { return v[0]*v[0] + v[1]*v[1] + v[2]*v[2]; }

  void init()  // constructor, :292-317
  {
    Ul[tau] += Ul[ddd];
    Fl[tau] += Fl[ddd];
    Ur[tau] += Ur[ddd];
    Fr[tau] += Fr[ddd];
    LE = am*Ul[tau] - Fl[tau];
    RE = ap*Ur[tau] - Fr[tau];
    LD = am*Ul[ddd] - Fl[ddd];
    RD = ap*Ur[ddd] - Fr[ddd];
    for (int k = 0; k < 3; ++k) {
      LB[k] = am*Ul[Bx+k] - Fl[Bx+k];
      RB[k] = ap*Ur[Bx+k] - Fr[Bx+k];
      Lm[k] = am*Ul[Sx+k] - Fl[Sx+k];
      Rm[k] = ap*Ur[Sx+k] - Fr[Sx+k];
    }
  }

  void fill_inner(R p, R *Kl, R *Kr, R *Bc, R &etaL, R &etaR, R &wL, R &wR,
                  Inner &aL, Inner &aR) const  // :321-384
  {
    {
      R A = Lm[0] - am*LE + p*(1 - am*am);
      R G = LB[1]*LB[1] + LB[2]*LB[2];
      R C = LB[1]*Lm[1] + LB[2]*Lm[2];
      R Q = -A - G + (Ul[Bx]*Ul[Bx])*(1 - am*am);
      R X = Ul[Bx]*(A*am*Ul[Bx] + C) - (A + G)*(am*p + LE);
      if (fabs(X) < 1e-8) X = 1e-8;
      aL.v[0] = (Ul[Bx]*(A*Ul[Bx] + am*C) - (A + G)*(p + Lm[0])) / X;
      aL.v[1] = (Q*Lm[1] + LB[1]*(C + Ul[Bx]*(am*Lm[0] - LE))) / X;
      aL.v[2] = (Q*Lm[2] + LB[2]*(C + Ul[Bx]*(am*Lm[0] - LE))) / X;
      aL.B[0] = (LB[0] - Ul[Bx]*aL.v[0]) / (am - aL.v[0]);
      aL.B[1] = (LB[1] - Ul[Bx]*aL.v[1]) / (am - aL.v[0]);
      aL.B[2] = (LB[2] - Ul[Bx]*aL.v[2]) / (am - aL.v[0]);
    }
    {
      R A = Rm[0] - ap*RE + p*(1 - ap*ap);
      R G = RB[1]*RB[1] + RB[2]*RB[2];
      R C = RB[1]*Rm[1] + RB[2]*Rm[2];
      R Q = -A - G + (Ur[Bx]*Ur[Bx])*(1 - ap*ap);
      R X = Ur[Bx]*(A*ap*Ur[Bx] + C) - (A + G)*(ap*p + RE);
      if (fabs(X) < 1e-8) X = 1e-8;
      aR.v[0] = (Ur[Bx]*(A*Ur[Bx] + ap*C) - (A + G)*(p + Rm[0])) / X;
      aR.v[1] = (Q*Rm[1] + RB[1]*(C + Ur[Bx]*(ap*Rm[0] - RE))) / X;
      aR.v[2] = (Q*Rm[2] + RB[2]*(C + Ur[Bx]*(ap*Rm[0] - RE))) / X;
      aR.B[0] = (RB[0] - Ur[Bx]*aR.v[0]) / (ap - aR.v[0]);
      aR.B[1] = (RB[1] - Ur[Bx]*aR.v[1]) / (ap - aR.v[0]);
      aR.B[2] = (RB[2] - Ur[Bx]*aR.v[2]) / (ap - aR.v[0]);
    }
    wL = p + (LE - dot(aL.v, Lm)) / (am - aL.v[0]);
    wR = p + (RE - dot(aR.v, Rm)) / (ap - aR.v[0]);
    if (wL < 0) wL *= -1;
    if (wR < 0) wR *= -1;
    etaL = -((Ul[Bx] > 0) - (Ul[Bx] < 0)) * sqrt(wL);
    etaR = +((Ur[Bx] > 0) - (Ur[Bx] < 0)) * sqrt(wR);
    for (int k = 0; k < 3; ++k) {
      Kl[k] = (Lm[k] + p*(k == 0) + LB[k]*etaL) / (am*p + LE + Ul[Bx]*etaL);
      Kr[k] = (Rm[k] + p*(k == 0) + RB[k]*etaR) / (ap*p + RE + Ur[Bx]*etaR);
      Bc[k] = ((aR.B[k]*(Kr[0] - aR.v[0]) + aR.B[0]*aR.v[k]) -
               (aL.B[k]*(Kl[0] - aL.v[0]) + aL.B[0]*aL.v[k])) / (Kr[0] - Kl[0]);
      if (fabs(Kr[0] - Kl[0]) < 1e-10) Bc[k] = 0.0;
    }
  }

  bool function(R p, R &y) const  // :386-415
  {
    if (fabs(Pl[Bx] - Pr[Bx]) > 1e-12) return false;
    if (isinf(p)) return false;
    R Kl[3], Kr[3], Bc[3], etaL, etaR, wL, wR;
    Inner aL, aR;
    fill_inner(p, Kl, Kr, Bc, etaL, etaR, wL, wR, aL, aR);
    const R delK = Kr[0] - Kl[0];
    const R Bch[3] = { Bc[0]*delK, Bc[1]*delK, Bc[2]*delK };
    const R Yl = (1 - dot(Kl)) / (etaL*delK - dot(Kl, Bch));
    const R Yr = (1 - dot(Kr)) / (etaR*delK - dot(Kr, Bch));
    if (fabs(delK) < 1e-10) y = 0.0;
    else if (fabs(Bc[0]) < 1e-10) y = delK;
    else y = delK*(1 - Bc[0]*(Yr - Yl));
    return true;
  }

  bool jacobian(R x, R &J) const  // :311-319, centred finite difference
  {
    const double dp = 1e-8;
    const R p0 = x - 0.5*dp;
    const R p1 = x + 0.5*dp;
    R f0, f1;
    if (!function(p0, f0)) return false;
    if (!function(p1, f1)) return false;
    J = (f1 - f0)/dp;
    return true;
  }

  /* NewtonRaphesonSolver(15, 1e-12)::Solve, Mara/nrsolver.cpp:20-83 with
   * EquationSystemBaseClass::GetErrorUpdateX, Mara/eqnbase.hpp:27-36 and the
   * rank-1 branch of matrix_inverse, Mara/matinv.c:42-46 */
  bool newton(R &x, int *iters) const
  {
    int worse = 0;
    R eps = 0.0, eps_last;
    int it = 0;
    do {
      eps_last = eps;
      R y, J;
      if (!function(x, y)) { *iters += it; return false; }
      if (!jacobian(x, J)) { *iters += it; return false; }
      const R G = 1.0 / J;
      R dx = 0.0;
      dx += G * y;
      x -= dx;
      eps = (y > 0 ? y : -y);
      worse += (eps > eps_last);
      if (worse >= 10 || it == 15) { *iters += it; return false; }
      if (isnan(eps)) { *iters += it; return false; }
    } while (eps > 1e-12 && ++it <= 15);
    *iters += it;
    return true;
  }

  bool reconstruct_solution(R p, R *U, R *F, double s) const  // :417-538
  {
    R Kl[3], Kr[3], Bc[3], etaL, etaR, wL, wR;
    Inner aL, aR, cL, cR;
    fill_inner(p, Kl, Kr, Bc, etaL, etaR, wL, wR, aL, aR);
    const R denomL = etaL - dot(Kl, Bc);
    const R denomR = etaR - dot(Kr, Bc);
    for (int k = 0; k < 3; ++k) {
      cL.v[k] = Kl[k] - Bc[k]*(1 - dot(Kl)) / denomL;
      cR.v[k] = Kr[k] - Bc[k]*(1 - dot(Kr)) / denomR;
      if (fabs(denomL) < 1e-10) cL.v[k] = aL.v[k];
      if (fabs(denomR) < 1e-10) cR.v[k] = aR.v[k];
    }
    const R lamC = cR.v[0];
    const R lmaL = Kl[0];
    const R lmaR = Kr[0];
    const double TOL = 1e-14;
    if (p < 0.0) return false;
    if (lmaL < -1.0 || lmaR > 1.0 || fabs(lamC) > 1.0) return false;
    if (!((wL > p) && (aL.v[0] > am-TOL) && (cL.v[0] > lmaL-TOL))) return false;
    if (!((wR > p) && (aR.v[0] < ap+TOL) && (cR.v[0] < lmaR+TOL))) return false;
    if (fabs(cR.v[0] - cL.v[0]) > 1e-8) return false;

    R denom;
    int i;
    if (fabs(denom = am - aL.v[0]) < 1e-8) {
      for (i = 0; i < 8; ++i) aL.U[i] = Ul[i];
    }
    else {
      const R vB = dot(aL.v, aL.B);
      aL.U[ddd] = LD / denom;
      aL.U[tau] = (LE + p*aL.v[0] - vB*aL.B[0]) / denom;
      aL.U[Sx] = (aL.U[tau] + p)*aL.v[0] - vB*aL.B[0];
      aL.U[Sy] = (aL.U[tau] + p)*aL.v[1] - vB*aL.B[1];
      aL.U[Sz] = (aL.U[tau] + p)*aL.v[2] - vB*aL.B[2];
      aL.U[Bx] = aL.B[0];
      aL.U[By] = aL.B[1];
      aL.U[Bz] = aL.B[2];
    }
    if (fabs(denom = ap - aR.v[0]) < 1e-8) {
      for (i = 0; i < 8; ++i) aR.U[i] = Ur[i];
    }
    else {
      const R vB = dot(aR.v, aR.B);
      aR.U[ddd] = RD / denom;
      aR.U[tau] = (RE + p*aR.v[0] - vB*aR.B[0]) / denom;
      aR.U[Sx] = (aR.U[tau] + p)*aR.v[0] - vB*aR.B[0];
      aR.U[Sy] = (aR.U[tau] + p)*aR.v[1] - vB*aR.B[1];
      aR.U[Sz] = (aR.U[tau] + p)*aR.v[2] - vB*aR.B[2];
      aR.U[Bx] = aR.B[0];
      aR.U[By] = aR.B[1];
      aR.U[Bz] = aR.B[2];
    }
    if (fabs(denom = lmaL - lamC) < 1e-8) {
      for (i = 0; i < 8; ++i) cL.U[i] = aL.U[i];
    }
    else {
      const R vBc = dot(cL.v, Bc);
      cL.U[ddd] = aL.U[ddd] * (lmaL - aL.v[0]) / denom;
      cL.U[tau] = (lmaL*aL.U[tau] - aL.U[Sx] + p*lamC - vBc*Bc[0]) / denom;
      cL.U[Sx] = (cL.U[tau] + p)*cL.v[0] - vBc*Bc[0];
      cL.U[Sy] = (cL.U[tau] + p)*cL.v[1] - vBc*Bc[1];
      cL.U[Sz] = (cL.U[tau] + p)*cL.v[2] - vBc*Bc[2];
      cL.U[Bx] = Bc[0];
      cL.U[By] = Bc[1];
      cL.U[Bz] = Bc[2];
    }
    if (fabs(denom = lmaR - lamC) < 1e-8) {
      for (i = 0; i < 8; ++i) cR.U[i] = aR.U[i];
    }
    else {
      const R vBc = dot(cR.v, Bc);
      cR.U[ddd] = aR.U[ddd] * (lmaR - aR.v[0]) / denom;
      cR.U[tau] = (lmaR*aR.U[tau] - aR.U[Sx] + p*lamC - vBc*Bc[0]) / denom;
      cR.U[Sx] = (cR.U[tau] + p)*cR.v[0] - vBc*Bc[0];
      cR.U[Sy] = (cR.U[tau] + p)*cR.v[1] - vBc*Bc[1];
      cR.U[Sz] = (cR.U[tau] + p)*cR.v[2] - vBc*Bc[2];
      cR.U[Bx] = Bc[0];
      cR.U[By] = Bc[1];
      cR.U[Bz] = Bc[2];
    }
    if (s <= am) for (i = 0; i < 8; ++i) F[i] = Fl[i];
    else if (am < s && s <= lmaL) for (i = 0; i < 8; ++i) F[i] = Fl[i] + am*(aL.U[i]-Ul[i]);
    else if (lmaL < s && s <= lamC) for (i = 0; i < 8; ++i) F[i] = Fl[i] + am*(aL.U[i]-Ul[i]) + lmaL*(cL.U[i]-aL.U[i]);
    else if (lamC < s && s <= lmaR) for (i = 0; i < 8; ++i) F[i] = Fr[i] + ap*(aR.U[i]-Ur[i]) + lmaR*(cR.U[i]-aR.U[i]);
    else if (lmaR < s && s <= ap) for (i = 0; i < 8; ++i) F[i] = Fr[i] + ap*(aR.U[i]-Ur[i]);
    else if (ap < s) for (i = 0; i < 8; ++i) F[i] = Fr[i];

    if (s <= am) for (i = 0; i < 8; ++i) U[i] = Ul[i];
    else if (am < s && s <= lmaL) for (i = 0; i < 8; ++i) U[i] = aL.U[i];
    else if (lmaL < s && s <= lamC) for (i = 0; i < 8; ++i) U[i] = cL.U[i];
    else if (lamC < s && s <= lmaR) for (i = 0; i < 8; ++i) U[i] = cR.U[i];
    else if (lmaR < s && s <= ap) for (i = 0; i < 8; ++i) U[i] = aR.U[i];
    else if (ap < s) for (i = 0; i < 8; ++i) U[i] = Ur[i];
    F[tau] -= F[ddd];
    U[tau] -= U[ddd];
    return true;
  }

  R estimate(int attempt)  // :540-579
  {
    if (attempt == 0) {
      const R *P = P_hll;
      const R V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
      const R B2 = P[Bx]*P[Bx] + P[By]*P[By] + P[Bz]*P[Bz];
      const R Bv = P[Bx]*P[vx] + P[By]*P[vy] + P[Bz]*P[vz];
      const R W2 = 1.0 / (1.0 - V2);
      const R W = sqrt(W2);
      const R b0 = W*Bv;
      const R b2 = (B2 + b0*b0) / W2;
      return P[pre] + 0.5*b2;
    }
    else if (attempt == 1) {
      F_hll[tau] += F_hll[ddd];
      U_hll[tau] += U_hll[ddd];
      const R a = 1.0;
      const R b = U_hll[tau] - F_hll[Sx];
      const R c = F_hll[tau] * U_hll[Sx] - U_hll[tau] * F_hll[Sx];
      const R p1 = fabs((-b + sqrt(b*b - 4*a*c))/(2*a));
      const R p2 = fabs((-b - sqrt(b*b - 4*a*c))/(2*a));
      F_hll[tau] -= F_hll[ddd];
      U_hll[tau] -= U_hll[ddd];
      return (p1 < p2) ? p1 : p2;
    }
    return R(0.0);
  }
};

struct HlldStats {
  std::int64_t newton_iters = 0;
};

template <class R> inline int riemann_hlld(const R *pl, const R *pr, R *f, int dim,
                                           const PhysParams &pp, Tally<R> &t)
{
  int B1, B2, B3, S1, S2, S3, v1, v2, v3;
  axis_perm(dim, B1, B2, B3, S1, S2, S3, v1, v2, v3);
  R epl, epr, eml, emr;
  R Ul[8], Ur[8], Pl[8], Pr[8], Fl[8], Fr[8];

  Pl[pre] = pl[pre]; Pr[pre] = pr[pre];
  Pl[rho] = pl[rho]; Pr[rho] = pr[rho];
  Pl[vx] = pl[v1]; Pr[vx] = pr[v1];
  Pl[vy] = pl[v2]; Pr[vy] = pr[v2];
  Pl[vz] = pl[v3]; Pr[vz] = pr[v3];
  Pl[Bx] = pl[B1]; Pr[Bx] = pr[B1];
  Pl[By] = pl[B2]; Pr[By] = pr[B2];
  Pl[Bz] = pl[B3]; Pr[Bz] = pr[B3];
  Pl[Bx] = Pr[Bx] = 0.5*(pl[B1] + pr[B1]);

  if (prim_to_cons(Pl, Ul, pp.gamma)) return 1;
  if (prim_to_cons(Pr, Ur, pp.gamma)) return 1;
  flux_and_eigenvalues(Ul, Pl, Fl, &epl, &eml, 1, pp.gamma);
  flux_and_eigenvalues(Ur, Pr, Fr, &epr, &emr, 1, pp.gamma);
  const R am = (eml < emr) ? eml : emr;
  const R ap = (epl > epr) ? epl : epr;
  note_lambda(t, am, ap);

  R P_hll[8], U_hll[8], F_hll[8];
  for (int i = 0; i < 8; ++i) {
    P_hll[i] = 0.5*(Pl[i] + Pr[i]);
    U_hll[i] = (ap*Ur[i] - am*Ul[i] + (Fl[i] - Fr[i])) / (ap - am);
    F_hll[i] = (ap*Fl[i] - am*Fr[i] + ap*am*(Ur[i] - Ul[i])) / (ap - am);
  }
  int hll_c2p_failed = 1;
  if (pp.nested_c2p_from_average && !c2p_check_cons(U_hll)) {
    C2P<R> s(pp.gamma, pp.pfloor);
    s.new_state(U_hll);
    s.set_starting_prim(P_hll);
    R Pg[8];
    hll_c2p_failed = s.solve_anton2dzw(Pg);
    t.c2p_iterations += s.iterations;
    if (!hll_c2p_failed)
      for (int i = 0; i < 8; ++i) P_hll[i] = Pg[i];
  }
  if (hll_c2p_failed) hll_c2p_failed = cons_to_prim(U_hll, P_hll, pp.gamma, pp.pfloor, &t.c2p_iterations);

  R U[8], F[8];
  HlldEqn<R> eqn;
  eqn.Pl = Pl; eqn.Pr = Pr; eqn.Ul = Ul; eqn.Ur = Ur; eqn.Fl = Fl; eqn.Fr = Fr;
  eqn.ap = ap; eqn.am = am;
  eqn.P_hll = P_hll; eqn.U_hll = U_hll; eqn.F_hll = F_hll;
  eqn.init();

  int attempt = 0, success = 0;
  int newton_iters = 0;
  R p_star;
  while (!hll_c2p_failed && attempt < 3 && !success) {
    if (fabs(Pl[Bx]) < 1e-8 || fabs(Pr[Bx]) < 1e-8) {
      p_star = eqn.estimate(1);
      success = eqn.reconstruct_solution(p_star, U, F, 0.0);
    }
    else {
      p_star = eqn.estimate(attempt);
      success = eqn.newton(p_star, &newton_iters) && eqn.reconstruct_solution(p_star, U, F, 0.0);
    }
    ++attempt;
  }
  if (success) {
    ++t.hlld_success;
    f[ddd] = F[ddd];
    f[tau] = F[tau];
    f[S1] = F[Sx];
    f[S2] = F[Sy];
    f[S3] = F[Sz];
    f[B1] = F[Bx];
    f[B2] = F[By];
    f[B3] = F[Bz];
    return 0;
  }
  ++t.hlld_failure;
  return riemann_hllc(pl, pr, f, dim, pp, t);
}

template <class R> inline int riemann_flux(int solver, const R *pl, const R *pr, R *F, int dim,
                                           const PhysParams &pp, Tally<R> &t)
{
  switch (solver) {
  case RIEMANN_HLL: return riemann_hll(pl, pr, F, dim, pp, t);
  case RIEMANN_HLLC: return riemann_hllc(pl, pr, F, dim, pp, t);
  default: return riemann_hlld(pl, pr, F, dim, pp, t);
  }
}

} // namespace oracle

#endif

namespace oracle {

/* ---------------------------------------------------------------------------------
 * RMHD characteristic eigensystem for the weno-split projection. No reference
 * counterpart: FluidEquations::Eigensystem (Mara/hydro.hpp:146-147) is empty for the RMHD
 * fluid. Pinned by invariants only (tests/test_eigensystem.py): A R = R diag(lam) for an
 * independent centred-difference Jacobian A = dF/dU, L R = I, the eigenvalues are the entropy speed,
 * the two Alfven speeds, the four roots of FluxAndEigenvalues' quartic and zero (the
 * normal field), and for a uniform state the projected split fluxes reproduce the
 * component-wise ones.
 *
 * Method: M = dU/dP and B = dF/dP by forward-mode differentiation of PrimToCons and the
 * flux formulas (exact to round-off); A_P = M^-1 B (similar to dF/dU);
 * eigenvalues in closed form where they have one and otherwise by Newton on the factored
 * quartic plus deflation; the eigenvectors of each eigenvalue cluster span the null space
 * of A_P - lam I (Gaussian elimination with complete pivoting); R = M R_P with unit
 * max-norm columns and L = R^-1. Returns 0, or nonzero when the eigenvectors are too
 * ill-conditioned to use (the caller falls back to the identity).
 * ------------------------------------------------------------------------------- */
namespace eig {

/* a value and its derivative along one primitive variable (forward-mode differentiation:
 * every operation applies its derivative rule, so the Jacobians below are exact to
 * round-off instead of carrying the cancellation of a difference quotient) */
struct D1 { double v, d; };
inline D1 operator+(D1 a, D1 b) { return {a.v + b.v, a.d + b.d}; }
inline D1 operator-(D1 a, D1 b) { return {a.v - b.v, a.d - b.d}; }
inline D1 operator*(D1 a, D1 b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline D1 operator/(D1 a, D1 b) { const double q = a.v / b.v; return {q, (a.d - q * b.d) / b.v}; }
inline D1 operator+(double s, D1 a) { return {s + a.v, a.d}; }
inline D1 operator-(double s, D1 a) { return {s - a.v, -a.d}; }
inline D1 operator*(double s, D1 a) { return {s * a.v, s * a.d}; }
inline D1 operator/(double s, D1 b) { const double q = s / b.v; return {q, -(q * b.d) / b.v}; }
inline D1 operator/(D1 a, double s) { return {a.v / s, a.d / s}; }
inline D1 sqrt(D1 a) { const double r = std::sqrt(a.v); return {r, a.d / (2.0 * r)}; }

/* U and the flux along dim of the primitive state P: PrimToCons (Mara/rmhd.cpp:133-163)
 * and the flux part of FluxAndEigenvalues (:164-199), in the reference's operation order;
 * T = double gives the same values as prim_to_cons and flux_and_eigenvalues above */
template <class T> inline void state_fluxes_t(const T *P, T *U, T *F, int dim, double gamma)
{
  using std::sqrt;
  const T V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
  const T B2 = P[Bx]*P[Bx] + P[By]*P[By] + P[Bz]*P[Bz];
  const T Bv = P[Bx]*P[vx] + P[By]*P[vy] + P[Bz]*P[vz];
  const T W2 = 1.0 / (1.0 - V2);
  const T W = sqrt(W2);
  const T b0 = W*Bv;
  const T b2 = (B2 + b0*b0) / W2;
  const T b[3] = {(P[Bx] + b0 * W*P[vx]) / W, (P[By] + b0 * W*P[vy]) / W, (P[Bz] + b0 * W*P[vz]) / W};
  const T e = (P[pre] / (gamma - 1)) / P[rho];
  const T e_ = e + 0.5 * b2 / P[rho];
  const T p_ = P[pre] + 0.5 * b2;
  const T h_ = 1.0 + e_ + p_ / P[rho];
  U[ddd] = P[rho] * W;
  U[tau] = P[rho] * h_ * W2 - p_ - b0*b0 - U[ddd];
  for (int a = 0; a < 3; ++a) U[Sx + a] = P[rho] * h_ * W2 * P[vx + a] - b0*b[a];
  for (int a = 0; a < 3; ++a) U[Bx + a] = P[Bx + a];
  const int vn = vx + dim - 1, bn = Bx + dim - 1;
  F[ddd] = U[ddd] * P[vn];
  F[tau] = U[tau] * P[vn] - b0*P[bn] / W + p_*P[vn];
  for (int a = 0; a < 3; ++a) {
    F[Sx + a] = U[Sx + a] * P[vn] - b[a]*P[bn] / W;
    if (a == dim - 1) F[Sx + a] = F[Sx + a] + p_;
    F[Bx + a] = (a == dim - 1) ? T{} : P[vn]*P[Bx + a] - P[vx + a]*P[bn];
  }
}

/* solves A X = Bm in place (A n x n, Bm n x m, row-major) by partial pivoting; 0 on success */
inline int solve(double *A, double *Bm, int n, int m)
{
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r) if (std::fabs(A[r*n + c]) > std::fabs(A[piv*n + c])) piv = r;
    if (!(std::fabs(A[piv*n + c]) > 0.0)) return 1;
    if (piv != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c*n + k], A[piv*n + k]);
      for (int k = 0; k < m; ++k) std::swap(Bm[c*m + k], Bm[piv*m + k]);
    }
    const double id = 1.0 / A[c*n + c];
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = A[r*n + c] * id;
      if (f == 0.0) continue;
      for (int k = c; k < n; ++k) A[r*n + k] -= f * A[c*n + k];
      for (int k = 0; k < m; ++k) Bm[r*m + k] -= f * Bm[c*m + k];
    }
  }
  for (int r = 0; r < n; ++r) {
    const double id = 1.0 / A[r*n + r];
    for (int k = 0; k < m; ++k) Bm[r*m + k] *= id;
  }
  return 0;
}

/* the `want` null vectors of the 8x8 matrix A (destroyed), whose null space has that
 * dimension: 8 - want steps of Gaussian elimination with complete pivoting, then back
 * substitution with each remaining column set to one in turn; written to the columns
 * col0.. of V (8x8, row-major). Returns 1 if a pivot vanishes. */
inline int null_space(double *A, double *V, int col0, int want)
{
  int rowp[8], colp[8];
  for (int i = 0; i < 8; ++i) { rowp[i] = i; colp[i] = i; }
  const int rank = 8 - want;
  for (int k = 0; k < rank; ++k) {
    int pr = k, pc = k;
    double best = -1.0;
    for (int r = k; r < 8; ++r)
      for (int c = k; c < 8; ++c) {
        const double a = std::fabs(A[rowp[r]*8 + colp[c]]);
        if (a > best) { best = a; pr = r; pc = c; }
      }
    if (!(best > 0.0)) return 1;
    std::swap(rowp[k], rowp[pr]);
    std::swap(colp[k], colp[pc]);
    const double piv = A[rowp[k]*8 + colp[k]];
    for (int r = k + 1; r < 8; ++r) {
      const double f = A[rowp[r]*8 + colp[k]] / piv;
      for (int c = k; c < 8; ++c) A[rowp[r]*8 + colp[c]] -= f * A[rowp[k]*8 + colp[c]];
    }
  }
  for (int fcount = 0; fcount < want; ++fcount) {
    double x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    x[colp[rank + fcount]] = 1.0;
    for (int k = rank - 1; k >= 0; --k) {
      double s = 0.0;
      for (int c = k + 1; c < 8; ++c) s += A[rowp[k]*8 + colp[c]] * x[colp[c]];
      x[colp[k]] = -s / A[rowp[k]*8 + colp[k]];
    }
    for (int i = 0; i < 8; ++i) V[i*8 + col0 + fcount] = x[i];
  }
  return 0;
}

/* root of K (x-v)^4 + (1-x^2)[L (x-v)^2 + (b0 x - bi)^2] by Newton from x (from outside the
 * roots it moves monotonically onto the extreme one; from a deflation estimate it polishes) */
inline double quartic_root_outside(double K, double L, double b0, double bi, double v, double x)
{
  for (int it = 0; it < 60; ++it) {
    const double y = x - v, y2 = y*y, c = b0*x - bi, q1 = L*y2 + c*c, om = (1.0 - x)*(1.0 + x);
    const double Q = K*y2*y2 + om*q1;
    const double dQ = 4.0*K*y2*y - 2.0*x*q1 + om*(2.0*L*y + 2.0*b0*c);
    const double step = Q / dQ;
    const double xn = x - step;
    if (!(std::fabs(xn) <= 1.0)) break;
    x = xn;
    if (!(std::fabs(step) > 1e-15)) break;
  }
  return x;
}

}  // namespace eig

inline int rmhd_eigensystem(const double *U, const double *P, double *L, double *R, double *lam, int dim,
                            double gamma)
{
  (void)U;
  using namespace eig;
  const int vn = dim == 1 ? vx : dim == 2 ? vy : vz;
  const int bn = dim == 1 ? Bx : dim == 2 ? By : Bz;
  // Jacobians in primitive variables, M = dU/dP and B = dF/dP, column k by
  // forward-mode differentiation along P[k]
  double M[64], B[64];
  for (int k = 0; k < 8; ++k) {
    D1 Pd[8], Ud[8], Fd[8];
    for (int i = 0; i < 8; ++i) Pd[i] = D1{P[i], i == k ? 1.0 : 0.0};
    state_fluxes_t(Pd, Ud, Fd, dim, gamma);
    for (int i = 0; i < 8; ++i) {
      M[i*8 + k] = Ud[i].d;
      B[i*8 + k] = Fd[i].d;
    }
  }
  double AP[64], Mc[64];
  for (int i = 0; i < 64; ++i) { AP[i] = B[i]; Mc[i] = M[i]; }
  if (solve(Mc, AP, 8, 8)) return 1;   // AP = M^-1 B
  // eigenvalues
  const double V2 = P[vx]*P[vx] + P[vy]*P[vy] + P[vz]*P[vz];
  const double Bv = P[Bx]*P[vx] + P[By]*P[vy] + P[Bz]*P[vz];
  const double B2 = P[Bx]*P[Bx] + P[By]*P[By] + P[Bz]*P[Bz];
  const double W = 1.0 / std::sqrt(1.0 - V2);
  const double b0 = W*Bv;
  const double b2 = (B2 + b0*b0) / (W*W);
  const double bi = (P[bn] + b0*W*P[vn]) / W;
  const double v = P[vn];
  const double rhoh = P[rho] + P[pre] / (gamma - 1.0) + P[pre];
  const double cs2 = gamma*P[pre] / rhoh;
  const double K = rhoh * (1.0/cs2 - 1.0) * W*W*W*W;
  const double Lq = -(rhoh + b2/cs2) * W*W;
  double ev[8];
  ev[0] = 0.0;                                   // normal field
  ev[1] = v;                                     // entropy
  {
    const double e = std::sqrt(rhoh + b2);       // Alfven: (b^n +- e u^n) / (b^0 +- e u^0)
    ev[2] = (bi + e*W*v) / (b0 + e*W);
    ev[3] = (bi - e*W*v) / (b0 - e*W);
  }
  {
    const double ca2 = b2 / (rhoh + b2);
    const double w2 = cs2 + ca2 - cs2*ca2;
    const double den = 1.0 / (1.0 - V2*w2);
    const double rt = std::sqrt(w2*(1.0 - V2)*(1.0 - V2*w2 - v*v*(1.0 - w2)));
    const double lp = std::min((v*(1.0 - w2) + rt)*den + 1e-12, 1.0);
    const double lm = std::max((v*(1.0 - w2) - rt)*den - 1e-12, -1.0);
    const double r1 = quartic_root_outside(K, Lq, b0, bi, v, lp);
    const double r2 = quartic_root_outside(K, Lq, b0, bi, v, lm);
    // deflate A4 x^4 + ... + A0 by (x - r1)(x - r2): the slow magnetosonic pair
    const double A4 = K - Lq - b0*b0;
    const double A3 = -4*K*v + Lq*v*2 + 2*b0*bi;
    const double A2 = 6*K*v*v + Lq*(1.0 - v*v) + b0*b0 - bi*bi;
    const double s = r1 + r2, pr = r1*r2;
    const double c1 = A3 + A4*s;
    const double c0 = A2 + c1*s - A4*pr;
    double disc = c1*c1 - 4.0*A4*c0;
    if (disc < 0.0) disc = 0.0;
    const double sq = std::sqrt(disc);
    ev[4] = r1;
    ev[5] = r2;
    // the deflated pair, polished by Newton on the factored quartic
    ev[6] = quartic_root_outside(K, Lq, b0, bi, v, (-c1 + sq) / (2.0*A4));
    ev[7] = quartic_root_outside(K, Lq, b0, bi, v, (-c1 - sq) / (2.0*A4));
  }
  std::sort(ev, ev + 8);
  // eigenvectors, cluster by cluster
  double RP[64];
  const double ctol = 1e-7;
  for (int i = 0; i < 8;) {
    int j = i + 1;
    while (j < 8 && ev[j] - ev[j - 1] < ctol) ++j;
    double mean = 0.0;
    for (int k = i; k < j; ++k) mean += ev[k];
    mean /= (j - i);
    double A[64];
    for (int r = 0; r < 8; ++r)
      for (int c = 0; c < 8; ++c) A[r*8 + c] = AP[r*8 + c] - (r == c ? mean : 0.0);
    if (null_space(A, RP, i, j - i)) return 2;
    for (int k = i; k < j; ++k) ev[k] = mean;
    i = j;
  }
  // conserved right eigenvectors, unit max-norm columns; L = R^-1
  for (int r = 0; r < 8; ++r)
    for (int c = 0; c < 8; ++c) {
      double s = 0.0;
      for (int k = 0; k < 8; ++k) s += M[r*8 + k] * RP[k*8 + c];
      R[r*8 + c] = s;
    }
  for (int c = 0; c < 8; ++c) {
    double m = 0.0;
    for (int r = 0; r < 8; ++r) m = std::max(m, std::fabs(R[r*8 + c]));
    if (!(m > 0.0)) return 1;
    for (int r = 0; r < 8; ++r) R[r*8 + c] /= m;
  }
  double Rc[64];
  for (int i = 0; i < 64; ++i) { Rc[i] = R[i]; L[i] = (i / 8 == i % 8) ? 1.0 : 0.0; }
  if (solve(Rc, L, 8, 8)) return 1;
  double lmax = 0.0, err = 0.0;
  for (int r = 0; r < 8; ++r)
    for (int c = 0; c < 8; ++c) {
      lmax = std::max(lmax, std::fabs(L[r*8 + c]));
      double s = 0.0;
      for (int k = 0; k < 8; ++k) s += L[r*8 + k] * R[k*8 + c];
      err = std::max(err, std::fabs(s - (r == c ? 1.0 : 0.0)));
    }
  if (!(lmax < 1e8) || !(err < 1e-8)) return 3;
  for (int k = 0; k < 8; ++k) lam[k] = ev[k];
  return 0;
}

}  // namespace oracle
