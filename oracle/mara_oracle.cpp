/* ORACLE — test infrastructure only (the checker, never the product).
 *
 * Grid-level CPU restatement of Mara's plain-method time step, following
 *   Mara/mara.cpp:297-341        luaC_Mara_advance (copy-in, P->U, RK, U->P, commit)
 *   Mara/hydro.cpp:80-202        prepare_integration, PrimToCons, ConsToPrim (+ emergency reset)
 *   Mara/plm-split.cpp:11-120    dUdt, intercell_flux_sweep, drive_sweeps_3d
 *   Mara/rmhd.cpp:293-337        ConstrainedTransport3d
 *   Mara/runge-kutta.hpp:7-55    single / rk2 / rk3 / rk4
 *   Mara/simple-cart.cpp:81-170  periodic Synchronize
 *   Mara/mara.cpp:374-381        get_timestep
 *   Mara/mara.cpp:2264-2324      measure.mean_max_divB
 * Arrays are the reference's own layout: AoS [i][j][k][q], q fastest, ghosts included.
 * Loops over zones run under OpenMP; each zone's arithmetic is the reference's, and
 * the only reductions are a max and integer sums, so results do not depend on the
 * thread count.
 *
 * Exposed through a C ABI (liboracle.so) for the Python tests and bench's CPU
 * baseline; built by oracle/Makefile.
 */
#include "mara_oracle.hpp"
#include "counted.hpp"

#include <algorithm>
#include <complex>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {
thread_local FlopTally g_tally;
}

using namespace oracle;

extern "C" {

typedef struct {
  int N[3];
  int Ng;
  double x0[3], x1[3];
  double gamma;
  int recon;      /* 0 pcm, 1 plm, 2 weno5 */
  double theta;
  int IS;         /* 0 js96, 1 b08, 2 sz10 */
  double sz10A;
  int riemann;    /* 0 hll, 1 hllc, 2 hlld */
  int rk;         /* 1 single, 2 rk2, 3 rk3, 4 rk4 */
  double pfloor;
  int ereset;
  double preset, dreset;
  int godunov;    /* 0 plm-split, 1 weno-split (local Lax-Friedrichs flux splitting) */
  int eigen;      /* weno-split projection: 0 component-wise, 1 RMHD characteristic */
} oracle_config;

typedef struct {
  double max_lambda;
  double hlld_success;
  double hlld_failure;
  double c2p_iterations;
  double c2p_solves;
  double riemann_retries;  /* interfaces that fell back to the PCM states */
  double riemann_failures; /* interfaces where both attempts failed */
  double reserved;
} oracle_stats;

}  // extern "C"

namespace {

struct Grid {
  oracle_config cfg;
  int n[3];        // augmented shape
  long T;          // total zones
  long sx, sy, sz; // strides in doubles (stride[1..3] of prepare_integration)
  double dx, dy, dz, min_dx;
  PhysParams pp;
  double max_lambda;
  std::vector<double> prim_array;   // MaraApplication::PrimitiveArray
  std::vector<int> failure_mask;    // MaraApplication::FailureMask
  oracle_stats last;
};

inline long wrap(long i, int N, int Ng)
{
  // index of the interior image of augmented index i (periodic)
  if (i < Ng) return i + N;
  if (i >= N + Ng) return i - N;
  return i;
}

void synchronize(const Grid &g, double *A)
{
  // Mara/simple-cart.cpp:81-170: all ghost zones end up holding their periodic image.
  const int Ng = g.cfg.Ng;
  const int *N = g.cfg.N;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < g.n[0]; ++i) {
    for (long j = 0; j < g.n[1]; ++j) {
      for (long k = 0; k < g.n[2]; ++k) {
        const bool ghost = i < Ng || i >= N[0] + Ng || j < Ng || j >= N[1] + Ng ||
                           k < Ng || k >= N[2] + Ng;
        if (!ghost) continue;
        const long src = ((wrap(i, N[0], Ng) * g.n[1] + wrap(j, N[1], Ng)) * g.n[2] +
                          wrap(k, N[2], Ng)) * 8;
        const long dst = ((i * g.n[1] + j) * g.n[2] + k) * 8;
        std::memcpy(A + dst, A + src, 8 * sizeof(double));
      }
    }
  }
}

void prepare(Grid &g)
{
  // Mara/simple-cart.cpp:8-72 and Mara/hydro.cpp:80-110
  const oracle_config &c = g.cfg;
  double gdx[3];
  double min_dx = c.x1[0] - c.x0[0];
  for (int d = 0; d < 3; ++d) {
    gdx[d] = (c.x1[d] - c.x0[d]) / c.N[d];
    if (gdx[d] < min_dx) min_dx = gdx[d];
  }
  g.min_dx = min_dx;
  double lx0[3], lx1[3];
  for (int d = 0; d < 3; ++d) {
    lx0[d] = c.x0[d] + gdx[d] * 0;
    lx1[d] = c.x0[d] + gdx[d] * (0 + c.N[d]);
    g.n[d] = c.N[d] + 2 * c.Ng;
  }
  g.T = (long)g.n[0] * g.n[1] * g.n[2];
  g.sz = 8;
  g.sy = (long)g.n[2] * 8;
  g.sx = (long)g.n[1] * g.n[2] * 8;
  g.dx = (lx1[0] - lx0[0]) / (g.n[0] - 2 * c.Ng);
  g.dy = (lx1[1] - lx0[1]) / (g.n[1] - 2 * c.Ng);
  g.dz = (lx1[2] - lx0[2]) / (g.n[2] - 2 * c.Ng);
  g.pp.gamma = c.gamma;
  g.pp.pfloor = c.pfloor;
  g.pp.theta = c.theta;
  g.pp.IS = c.IS;
  g.pp.sz10A = c.sz10A;
}

struct Acc {
  double max_lambda = 0.0;
  long hlld_success = 0, hlld_failure = 0, c2p_iter = 0, c2p_solves = 0;
  long retries = 0, failures = 0;
};

/* GodunovOperator::ConsToPrim, Mara/hydro.cpp:125-202. U is synchronized first (the
 * reference does so in place, through a const_cast). Returns the failed-zone count. */
int cons_to_prim_all(Grid &g, double *U, double *P, Acc &acc)
{
  synchronize(g, U);
  const long nz = g.T;
  long ttl = 0, iters = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : ttl, iters)
  for (long z = 0; z < nz; ++z) {
    std::int64_t it = 0;
    int err = (cons_to_prim(U + 8 * z, P + 8 * z, g.pp.gamma, g.pp.pfloor, &it) != 0);
    g.failure_mask[z] = err;
    ttl += err;
    iters += it;
  }
  acc.c2p_iter += iters;
  acc.c2p_solves += nz;
  if (ttl && g.cfg.ereset) {
#pragma omp parallel for schedule(static)
    for (long z = 0; z < nz; ++z) {
      if (g.failure_mask[z]) {
        double *p = P + 8 * z;
        p[0] = g.cfg.dreset;
        p[1] = g.cfg.preset;
        p[2] = p[3] = p[4] = 0.0;
        p[5] = U[8 * z + 5];
        p[6] = U[8 * z + 6];
        p[7] = U[8 * z + 7];
      }
    }
    // every zone's U is rebuilt from the repaired primitives (:186-193)
#pragma omp parallel for schedule(static)
    for (long z = 0; z < nz; ++z) prim_to_cons(P + 8 * z, U + 8 * z, g.pp.gamma);
    std::fill(g.failure_mask.begin(), g.failure_mask.end(), 0);
    ttl = 0;
  }
  return (int)ttl;
}

/* MethodOfLinesSplit::intercell_flux_sweep, Mara/plm-split.cpp:43-74 */
void flux_sweep(Grid &g, const double *P, double *F, int dim, Acc &acc)
{
  const long S = dim == 1 ? g.sx : dim == 2 ? g.sy : g.sz;
  const long end = 8 * g.T - 3 * S;
  const int recon = g.cfg.recon;
  const int solver = g.cfg.riemann;
  double lam = 0.0;
  long succ = 0, fail = 0, iters = 0, retries = 0, failures = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : lam) \
    reduction(+ : succ, fail, iters, retries, failures)
  for (long i = 2 * S; i < end; i += 8) {
    double Pl[8], Pr[8];
    for (int q = 0; q < 8; ++q) {
      const long m = i + q;
      double v[6] = { P[m-2*S], P[m-S], P[m+0], P[m+1*S], P[m+2*S], P[m+3*S] };
      switch (recon) {
      case RECON_PCM: Pl[q] = v[2]; Pr[q] = v[3]; break;
      case RECON_PLM:
        Pl[q] = reconstruct(&v[2], PLM_C2R, g.pp);
        Pr[q] = reconstruct(&v[3], PLM_C2L, g.pp);
        break;
      default:
        Pl[q] = reconstruct(&v[2], WENO5_FV_C2R, g.pp);
        Pr[q] = reconstruct(&v[3], WENO5_FV_C2L, g.pp);
        break;
      }
    }
    Tally<double> t;
    int error = riemann_flux(solver, Pl, Pr, &F[i], dim, g.pp, t);
    if (error) {
      ++retries;
      error = riemann_flux(solver, &P[i], &P[i + S], &F[i], dim, g.pp, t);
      if (error) {
        ++failures;
        for (int q = 0; q < 8; ++q) F[i + q] = 0.0;  // the reference leaves this memory unwritten
      }
    }
    if (t.max_lambda > lam) lam = t.max_lambda;
    succ += t.hlld_success;
    fail += t.hlld_failure;
    iters += t.c2p_iterations;
  }
  if (lam > acc.max_lambda) acc.max_lambda = lam;
  acc.hlld_success += succ;
  acc.hlld_failure += fail;
  acc.c2p_iter += iters;
  acc.retries += retries;
  acc.failures += failures;
}

/* The projection of the weno-split sweep: the identity (component-wise splitting), or
 * (eigen = 1) the RMHD characteristic eigenvectors. The reference's RMHD fluid has
 * no Eigensystem (FluidEquations::Eigensystem, Mara/hydro.hpp:146-147, is empty), so its
 * weno-split reads uninitialised matrices; these are the two working choices. */
void eigensystem(const Grid &g, const double *U, const double *P, double *L, double *R, double *lam,
                 int dim)
{
  if (g.cfg.eigen == 1 && rmhd_eigensystem(U, P, L, R, lam, dim, g.pp.gamma) == 0) return;
  // component-wise, or where the RMHD eigenvectors are too ill-conditioned to use
  for (int i = 0; i < 8; ++i) {
    lam[i] = 0.0;
    for (int j = 0; j < 8; ++j) L[i*8 + j] = R[i*8 + j] = (i == j) ? 1.0 : 0.0;
  }
}

/* matrix_vector_product, Mara/matmult.c:41-62: b_i = sum_j A_ij x_j, summed in j order */
inline void matvec8(const double *A, const double *x, double *b)
{
  for (int i = 0; i < 8; ++i) {
    b[i] = 0.0;
    for (int j = 0; j < 8; ++j) b[i] += A[i*8 + j] * x[j];
  }
}

/* WenoSplit::drive_single_sweep and intercell_flux_sweep with local Lax-Friedrichs flux
 * splitting, Mara/weno-split.cpp:163-207 and :210-245, over every line of the sweep
 * direction (ghost lines included, :296-314). Per line: the physical flux F and the
 * fastest speed A of every zone (MaxLambda); per interface i in [Ng-1, N+Ng-1]: the
 * eigensystem of the averaged primitive state, the split fluxes 0.5 (F +- ml U) of the
 * 6-zone stencil (ml = max A over it) projected with L, the finite-difference WENO5 of
 * each projected component and the projection back with R. The reference leaves the
 * other interfaces of a line unwritten (malloc'd); they only feed ghost zones, which the
 * next ConsToPrim overwrites; zero is used here. */
void weno_split_sweep(Grid &g, const double *U, const double *P, double *Fiph, int dim, Acc &acc)
{
  const int Ng = g.cfg.Ng;
  const int Nd = g.cfg.N[dim - 1];
  const int n = Nd + 2 * Ng;
  const long S = dim == 1 ? g.sx : dim == 2 ? g.sy : g.sz;
  // the two transverse axes and their strides (in doubles)
  const int a1 = dim == 1 ? 2 : 1, a2 = dim == 3 ? 2 : 3;
  const long s1 = a1 == 1 ? g.sx : g.sy, s2 = a2 == 2 ? g.sy : g.sz;
  const int n1 = g.n[a1 - 1], n2 = g.n[a2 - 1];
  double lam_max = 0.0;
#pragma omp parallel for schedule(dynamic, 4) reduction(max : lam_max)
  for (long line = 0; line < (long)n1 * n2; ++line) {
    const long base = (line / n2) * s1 + (line % n2) * s2;
    std::vector<double> F(8 * n), A(n);
    for (int i = 0; i < n; ++i) {
      double ap, am;
      flux_and_eigenvalues(U + base + i * S, P + base + i * S, &F[8 * i], &ap, &am, dim, g.pp.gamma);
      A[i] = (std::fabs(ap) > std::fabs(am)) ? std::fabs(ap) : std::fabs(am);
      if (lam_max < A[i]) lam_max = A[i];
    }
    for (int i = 0; i < n; ++i)
      for (int q = 0; q < 8; ++q) Fiph[base + i * S + q] = 0.0;
    for (int i = Ng - 1; i < Nd + Ng; ++i) {
      double Piph[8], Uiph[8], L[64], R[64], lam[8];
      for (int q = 0; q < 8; ++q) Piph[q] = 0.5 * (P[base + i * S + q] + P[base + (i + 1) * S + q]);
      prim_to_cons(Piph, Uiph, g.pp.gamma);
      eigensystem(g, Uiph, Piph, L, R, lam, dim);
      const double ml = *std::max_element(&A[i - 2], &A[i + 4]);
      double fpT[8 * 6], fmT[8 * 6];
      for (int j = 0; j < 6; ++j) {
        const double *Uj = U + base + (i + j - 2) * S;
        const double *Fj = &F[8 * (i + j - 2)];
        double Fp[8], Fm[8], fp[8], fm[8];
        for (int q = 0; q < 8; ++q) {
          Fp[q] = 0.5 * (Fj[q] + ml * Uj[q]);
          Fm[q] = 0.5 * (Fj[q] - ml * Uj[q]);
        }
        matvec8(L, Fp, fp);
        matvec8(L, Fm, fm);
        for (int q = 0; q < 8; ++q) {
          fpT[q * 6 + j] = fp[q];
          fmT[q * 6 + j] = fm[q];
        }
      }
      double f[8];
      for (int q = 0; q < 8; ++q)
        f[q] = reconstruct(&fpT[q * 6 + 2], WENO5_FD_C2R, g.pp) + reconstruct(&fmT[q * 6 + 3], WENO5_FD_C2L, g.pp);
      matvec8(R, f, &Fiph[base + i * S]);
    }
  }
  if (lam_max > acc.max_lambda) acc.max_lambda = lam_max;
}

/* Rmhd::ConstrainedTransport3d, Mara/rmhd.cpp:293-337. The reference's temporaries are
 * uninitialised outside [sx, stride0-sx); those zones are ghosts, zero is used here. */
void constrained_transport_3d(const Grid &g, double *Fx, double *Fy, double *Fz)
{
  const long nzone = g.T;
  std::vector<double> FxBy(nzone, 0.0), FxBz(nzone, 0.0), FyBz(nzone, 0.0), FyBx(nzone, 0.0),
      FzBx(nzone, 0.0), FzBy(nzone, 0.0);
  const long sx = g.sx, sy = g.sy, sz = g.sz, s0 = 8 * g.T;
#pragma omp parallel for schedule(static)
  for (long i = sx; i < s0 - sx; i += 8) {
    const double *F, *G, *H;
    F = &Fx[By + i];
    G = &Fy[Bx + i];
    FxBy[i/8] = (2*F[0]+F[sy]+F[-sy]-G[0]-G[sx]-G[-sy]-G[ sx-sy])*0.125;
    FyBx[i/8] = (2*G[0]+G[sx]+G[-sx]-F[0]-F[sy]-F[-sx]-F[-sx+sy])*0.125;
    G = &Fy[Bz + i];
    H = &Fz[By + i];
    FyBz[i/8] = (2*G[0]+G[sz]+G[-sz]-H[0]-H[sy]-H[-sz]-H[ sy-sz])*0.125;
    FzBy[i/8] = (2*H[0]+H[sy]+H[-sy]-G[0]-G[sz]-G[-sy]-G[-sy+sz])*0.125;
    H = &Fz[Bx + i];
    F = &Fx[Bz + i];
    FzBx[i/8] = (2*H[0]+H[sx]+H[-sx]-F[0]-F[sz]-F[-sx]-F[ sz-sx])*0.125;
    FxBz[i/8] = (2*F[0]+F[sz]+F[-sz]-H[0]-H[sx]-H[-sz]-H[-sz+sx])*0.125;
  }
#pragma omp parallel for schedule(static)
  for (long i = 0; i < s0; i += 8) {
    Fx[i+Bx] = 0.0;        Fx[i+By] = FxBy[i/8];  Fx[i+Bz] = FxBz[i/8];
    Fy[i+Bx] = FyBx[i/8];  Fy[i+By] = 0.0;        Fy[i+Bz] = FyBz[i/8];
    Fz[i+Bx] = FzBx[i/8];  Fz[i+By] = FzBy[i/8];  Fz[i+Bz] = 0.0;
  }
}

/* MethodOfLinesSplit::dUdt, Mara/plm-split.cpp:11-33 and drive_sweeps_3d :103-120.
 * Returns false on IntermediateFailure (c2p failed somewhere). */
/* PlmCtuHancockOperator::ctu_hancock (Mara/ctu-hancock.cpp:71-494) in 3-D: its own PLM
 * limiter (plm_minmod with the operator's theta, :22-31) on the stage primitives; Godunov
 * fluxes of the half-slope states (step 2); a Hancock predictor per axis with the
 * transverse Godunov fluxes (step 3); the predictors synchronised and inverted with P as
 * the guess; the corrector fluxes from them over [sx, stride0 - sx) (step 4). With
 * DoNotTolerateFailures (the reference never lowers it) every unphysical state or failed
 * inversion adds one to the failure mask, and the step ends at the next abort point.
 * Returns the mask's sum there, 0 when F, G, H hold the corrector fluxes. */
int ctu_hancock(Grid &g, const double *P, const double *U, double *F, double *G, double *H, double dt, Acc &acc)
{
  const long s0 = 8 * g.T, sx = g.sx, sy = g.sy, sz = g.sz;
  const double theta = g.cfg.theta;
  auto sgn = [](double x) { return (double)((x > 0) - (x < 0)); };
  auto minmod = [&](double ul, double u0, double ur) {
    const double a = theta * (u0 - ul);
    const double b =   0.5 * (ur - ul);
    const double c = theta * (ur - u0);
    const double ab = std::fabs(a) < std::fabs(b) ? std::fabs(a) : std::fabs(b);
    const double m3 = ab < std::fabs(c) ? ab : std::fabs(c);
    return 0.25*std::fabs(sgn(a) + sgn(b))*(sgn(a) + sgn(c))*m3;
  };
  std::vector<double> dPdx(s0, 0.0), dPdy(s0, 0.0), dPdz(s0, 0.0);
#pragma omp parallel for schedule(static)
  for (long i = sx; i < s0 - sx; ++i) {
    dPdx[i] = minmod(P[i-sx], P[i], P[i+sx]);
    dPdy[i] = minmod(P[i-sy], P[i], P[i+sy]);
    dPdz[i] = minmod(P[i-sz], P[i], P[i+sz]);
  }
  const double *dP[3] = {dPdx.data(), dPdy.data(), dPdz.data()};
  double *Fd[3] = {F, G, H};
  const long Sd[3] = {sx, sy, sz};
  std::vector<int> mask(g.T, 0);
  auto mask_sum = [&]() { long t = 0; for (int m : mask) t += m; return (int)t; };
  double lam = 0.0;
  long succ = 0, fail = 0, iters = 0;
  // Godunov fluxes of reconstructed states: step 2 from P over [0, s0 - sx), step 4 from the
  // predicted primitives over [sx, s0 - sx)
  auto faces = [&](const double *const Psrc[3], long begin) {
#pragma omp parallel for schedule(dynamic, 64) reduction(max : lam) reduction(+ : succ, fail, iters)
    for (long i = begin; i < s0 - sx; i += 8) {
      for (int d = 0; d < 3; ++d) {
        const long S = Sd[d];
        double Pl[8], Pr[8];
        for (int j = 0; j < 8; ++j) {
          Pr[j] = Psrc[d][i+S+j] - 0.5*dP[d][i+S+j];
          Pl[j] = Psrc[d][i+j] + 0.5*dP[d][i+j];
        }
        if (c2p_check_prim(Pl) || c2p_check_prim(Pr)) {
#pragma omp atomic
          mask[i/8] += 1;
          continue;
        }
        Tally<double> t;
        riemann_flux(g.cfg.riemann, Pl, Pr, Fd[d] + i, d + 1, g.pp, t);
        if (t.max_lambda > lam) lam = t.max_lambda;
        succ += t.hlld_success;
        fail += t.hlld_failure;
        iters += t.c2p_iterations;
      }
    }
  };
  auto collect = [&]() {
    if (lam > acc.max_lambda) acc.max_lambda = lam;
    acc.hlld_success += succ; acc.hlld_failure += fail; acc.c2p_iter += iters;
    lam = 0.0; succ = fail = iters = 0;
  };
  const double *Pthree[3] = {P, P, P};
  faces(Pthree, 0);
  collect();
  if (int e = mask_sum()) return e;
  // step 3: the Hancock predictor of each axis
  std::vector<double> Ux(s0, 0.0), Uy(s0, 0.0), Uz(s0, 0.0);
  double *Ud[3] = {Ux.data(), Uy.data(), Uz.data()};
  const double dx = g.dx, dy = g.dy, dz = g.dz;
#pragma omp parallel for schedule(static)
  for (long i = sx; i < s0; i += 8) {
    double Fo[3][2][8];                    // (FL, FR), (GL, GR), (HL, HR)
    bool bad = false;
    for (int d = 0; d < 3; ++d) {
      double PL[8], PR[8], UL[8], UR[8];
      for (int j = 0; j < 8; ++j) {
        PL[j] = P[i+j] - 0.5*dP[d][i+j];
        PR[j] = P[i+j] + 0.5*dP[d][i+j];
      }
      if (c2p_check_prim(PL) || c2p_check_prim(PR)) {
        mask[i/8] += 1;
        bad = true;
        continue;
      }
      prim_to_cons(PL, UL, g.pp.gamma);
      prim_to_cons(PR, UR, g.pp.gamma);
      flux_and_eigenvalues<double>(UL, PL, Fo[d][0], nullptr, nullptr, d + 1, g.pp.gamma);
      flux_and_eigenvalues<double>(UR, PR, Fo[d][1], nullptr, nullptr, d + 1, g.pp.gamma);
    }
    if (bad) continue;
    for (int j = 0; j < 8; ++j) {
      const double *FL = Fo[0][0], *FR = Fo[0][1], *GL = Fo[1][0], *GR = Fo[1][1], *HL = Fo[2][0], *HR = Fo[2][1];
      Ux[i+j] = U[i+j] - ((FR[j]-FL[j])/dx + (G[i+j]-G[i-sy+j])/dy + (H[i+j]-H[i-sz+j])/dz)*0.5*dt;
      Uy[i+j] = U[i+j] - ((GR[j]-GL[j])/dy + (H[i+j]-H[i-sz+j])/dz + (F[i+j]-F[i-sx+j])/dx)*0.5*dt;
      Uz[i+j] = U[i+j] - ((HR[j]-HL[j])/dz + (F[i+j]-F[i-sx+j])/dx + (G[i+j]-G[i-sy+j])/dy)*0.5*dt;
    }
  }
  if (int e = mask_sum()) return e;
  // the predictors, synchronised and inverted from P (Rmhd::ConsToPrim, no emergency reset)
  std::vector<double> Px(P, P + s0), Py(P, P + s0), Pz(P, P + s0);
  double *Pd[3] = {Px.data(), Py.data(), Pz.data()};
  for (int d = 0; d < 3; ++d) synchronize(g, Ud[d]);
  long c2p_it = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : c2p_it)
  for (long z = 0; z < g.T; ++z) {
    for (int d = 0; d < 3; ++d) {
      std::int64_t it = 0;
      if (cons_to_prim(Ud[d] + 8*z, Pd[d] + 8*z, g.pp.gamma, g.pp.pfloor, &it) != 0) mask[z] += 1;
      c2p_it += it;
    }
  }
  acc.c2p_iter += c2p_it;
  if (int e = mask_sum()) return e;
  const double *Ppred[3] = {Px.data(), Py.data(), Pz.data()};
  faces(Ppred, sx);
  collect();
  return mask_sum();
}

bool dUdt(Grid &g, const std::vector<double> &Uin, std::vector<double> &L, Acc &acc, int *err, double dt)
{
  std::vector<double> U = Uin;
  L.assign(U.size(), 0.0);
  std::vector<double> &P = g.prim_array;
  int e = cons_to_prim_all(g, U.data(), P.data(), acc);
  // PlmCtuHancockOperator::dUdt (Mara/ctu-hancock.cpp:52-69) ignores the count: failed zones
  // keep their previous primitives and the operator restarts the failure mask
  if (e != 0 && g.cfg.godunov != 2) { *err = e; return false; }
  const long s0 = 8 * g.T;
  std::vector<double> F(s0, 0.0), G(s0, 0.0), H(s0, 0.0);
  if (g.cfg.godunov == 2) {
    const int f = ctu_hancock(g, P.data(), U.data(), F.data(), G.data(), H.data(), dt, acc);
    if (f) { *err = f; return false; }
  }
  else if (g.cfg.godunov == 1) {
    // WenoSplit::drive_sweeps_3d, Mara/weno-split.cpp:291-335 (U is the synchronised copy)
    weno_split_sweep(g, U.data(), P.data(), F.data(), 1, acc);
    weno_split_sweep(g, U.data(), P.data(), G.data(), 2, acc);
    weno_split_sweep(g, U.data(), P.data(), H.data(), 3, acc);
  }
  else {
    flux_sweep(g, P.data(), F.data(), 1, acc);
    flux_sweep(g, P.data(), G.data(), 2, acc);
    flux_sweep(g, P.data(), H.data(), 3, acc);
  }
  constrained_transport_3d(g, F.data(), G.data(), H.data());
  const long sx = g.sx, sy = g.sy, sz = g.sz;
  const double dx = g.dx, dy = g.dy, dz = g.dz;
#pragma omp parallel for schedule(static)
  for (long i = sx; i < s0; ++i) {
    L[i] = -(F[i]-F[i-sx])/dx - (G[i]-G[i-sy])/dy - (H[i]-H[i-sz])/dz;
  }
  return true;
}

/* RungeKutta*::AdvanceState, Mara/runge-kutta.hpp:7-55 */
bool advance_state(Grid &g, std::vector<double> &U, double dt, Acc &acc, int *err)
{
  const size_t n = U.size();
  std::vector<double> L, U1(n);
  switch (g.cfg.rk) {
  case 1:
    if (!dUdt(g, U, L, acc, err, dt)) return false;
    for (size_t i = 0; i < n; ++i) U[i] += dt*L[i];
    return true;
  case 2:
    if (!dUdt(g, U, L, acc, err, dt)) return false;
    for (size_t i = 0; i < n; ++i) U1[i] = U[i] + dt*L[i];
    if (!dUdt(g, U1, L, acc, err, dt)) return false;
    for (size_t i = 0; i < n; ++i) U[i] = 0.5*(U[i] + U1[i] + dt*L[i]);
    return true;
  case 4: {
    std::vector<double> L1(n), L2(n), L3(n), T(n);
    if (!dUdt(g, U, L, acc, err, dt)) return false;
    for (size_t i = 0; i < n; ++i) L1[i] = dt*L[i];
    for (size_t i = 0; i < n; ++i) T[i] = U[i] + 0.5*L1[i];
    if (!dUdt(g, T, L, acc, err, dt)) return false;
    for (size_t i = 0; i < n; ++i) L2[i] = dt*L[i];
    for (size_t i = 0; i < n; ++i) T[i] = U[i] + 0.5*L2[i];
    if (!dUdt(g, T, L, acc, err, dt)) return false;
    for (size_t i = 0; i < n; ++i) L3[i] = dt*L[i];
    for (size_t i = 0; i < n; ++i) T[i] = U[i] + 1.0*L3[i];
    if (!dUdt(g, T, L, acc, err, dt)) return false;
    for (size_t i = 0; i < n; ++i) U[i] += (1.0/6.0)*(L1[i] + 2.0*L2[i] + 2.0*L3[i] + dt*L[i]);
    return true;
  }
  default:
    if (!dUdt(g, U, L, acc, err, dt)) return false;
    for (size_t i = 0; i < n; ++i) U1[i] = U[i] + dt*L[i];
    if (!dUdt(g, U1, L, acc, err, dt)) return false;
    for (size_t i = 0; i < n; ++i) U1[i] = 3./4*U[i] + 1./4*U1[i] + 1./4*dt*L[i];
    if (!dUdt(g, U1, L, acc, err, dt)) return false;
    for (size_t i = 0; i < n; ++i) U[i] = 1./3*U[i] + 2./3*U1[i] + 2./3*dt*L[i];
    return true;
  }
}

}  // namespace

extern "C" {

void oracle_set_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int oracle_get_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void *oracle_create(const oracle_config *cfg)
{
  Grid *g = new Grid;
  g->cfg = *cfg;
  prepare(*g);
  g->max_lambda = 0.0;
  g->prim_array.assign(8 * g->T, 0.0);
  g->failure_mask.assign(g->T, 0);
  std::memset(&g->last, 0, sizeof(g->last));
  return g;
}

void oracle_destroy(void *h) { delete static_cast<Grid *>(h); }

long oracle_num_zones(void *h) { return static_cast<Grid *>(h)->T; }

/* luaC_Mara_advance, Mara/mara.cpp:297-341: returns the error count; P_user is only
 * overwritten when it is zero. */
int oracle_advance(void *h, double *P_user, double dt, oracle_stats *stats)
{
  Grid &g = *static_cast<Grid *>(h);
  const long s0 = 8 * g.T;
  std::vector<double> P(P_user, P_user + s0), U(s0, 0.0);
  Acc acc;
  g.max_lambda = 0.0;
  std::fill(g.prim_array.begin(), g.prim_array.end(), 0.0);
  synchronize(g, P.data());
#pragma omp parallel for schedule(static)
  for (long z = 0; z < g.T; ++z) prim_to_cons(&P[8 * z], &U[8 * z], g.pp.gamma);

  int errors = 0;
  int ierr = 0;
  if (advance_state(g, U, dt, acc, &ierr)) {
    P = g.prim_array;
    errors = cons_to_prim_all(g, U.data(), P.data(), acc);
  }
  else {
    errors = ierr;
  }
  if (errors == 0) std::memcpy(P_user, P.data(), s0 * sizeof(double));
  g.max_lambda = acc.max_lambda;
  g.last.max_lambda = acc.max_lambda;
  g.last.hlld_success = (double)acc.hlld_success;
  g.last.hlld_failure = (double)acc.hlld_failure;
  g.last.c2p_iterations = (double)acc.c2p_iter;
  g.last.c2p_solves = (double)acc.c2p_solves;
  g.last.riemann_retries = (double)acc.retries;
  g.last.riemann_failures = (double)acc.failures;
  if (stats) *stats = g.last;
  return errors;
}

/* luaC_Mara_get_timestep, Mara/mara.cpp:374-381 (inf when no solve has run) */
double oracle_get_timestep(void *h, double cfl)
{
  Grid &g = *static_cast<Grid *>(h);
  return cfl * g.min_dx / g.max_lambda;
}

/* luaC_mean_max_divB, Mara/mara.cpp:2264-2324, on a primitive array (set_primitive).
 * Returns the reference's (mean, signed max) pair and, in absmax, max |div B| / dx. */
void oracle_mean_max_divB(void *h, const double *P, double *mean, double *max, double *absmax)
{
  Grid &g = *static_cast<Grid *>(h);
  const int Ng = g.cfg.Ng;
  const int *n = g.n;
  const long T = g.T;
  std::vector<double> div(T, 0.0);
  auto M = [&](long i, long j, long k) { return ((i * n[1] + j) * n[2] + k) * 8; };
  const int bx = 5, by = 6, bz = 7;
#pragma omp parallel for schedule(static)
  for (long i = Ng; i < n[0]-Ng-1; ++i) {
    for (long j = Ng; j < n[1]-Ng-1; ++j) {
      for (long k = Ng; k < n[2]-Ng-1; ++k) {
        const double *fx = P + bx, *fy = P + by, *fz = P + bz;
        div[M(i,j,k)/8] =
            ((fx[M(i+1,j  ,k  )] + fx[M(i+1,j+1,k  )] + fx[M(i+1,j  ,k+1)] + fx[M(i+1,j+1,k+1)]) -
             (fx[M(i  ,j  ,k  )] + fx[M(i  ,j+1,k  )] + fx[M(i  ,j  ,k+1)] + fx[M(i  ,j+1,k+1)])) / 4.0
          + ((fy[M(i  ,j+1,k  )] + fy[M(i  ,j+1,k+1)] + fy[M(i+1,j+1,k  )] + fy[M(i+1,j+1,k+1)]) -
             (fy[M(i  ,j  ,k  )] + fy[M(i  ,j  ,k+1)] + fy[M(i+1,j  ,k  )] + fy[M(i+1,j  ,k+1)])) / 4.0
          + ((fz[M(i  ,j  ,k+1)] + fz[M(i+1,j  ,k+1)] + fz[M(i  ,j+1,k+1)] + fz[M(i+1,j+1,k+1)]) -
             (fz[M(i  ,j  ,k  )] + fz[M(i+1,j  ,k  )] + fz[M(i  ,j+1,k  )] + fz[M(i+1,j+1,k  )])) / 4.0;
      }
    }
  }
  double sum = 0.0, mx = div[0], amx = 0.0;
  for (long z = 0; z < T; ++z) {   // std::valarray::sum / max: sequential order
    if (z > 0) sum += div[z]; else sum = div[0];
    if (div[z] > mx) mx = div[z];
    if (std::fabs(div[z]) > amx) amx = std::fabs(div[z]);
  }
  const double glb = (double)g.cfg.N[0] * g.cfg.N[1] * g.cfg.N[2];
  if (mean) *mean = sum / T / glb;
  if (max) *max = mx;
  if (absmax) *absmax = amx / g.min_dx;
}

/* The Mara.measure table on a primitive array (Mara/mara.cpp:1954-2262): sums over the
 * interior zones in memory order (absolute_index_is_ghost, :2326), divided by the global
 * interior count, and extremes, as luaC_mean_velocity (:1954), luaC_mean_cons (:1982),
 * luaC_mean_prim (:2015), luaC_mean_energies (:2043), luaC_mean_max_temperature (:2105),
 * luaC_mean_max_sonic_mach (:2128), luaC_mean_min_alfvenic_mach (:2167),
 * luaC_mean_max_magnetic_field (:2203) and luaC_max_lorentz_factor (:2232), for the
 * AdiabaticIdealRmhd fluid and the gamma-law AdiabaticEos (Mara/eos.cpp:5-47).
 * units = PhysicalUnits(Length, Mass, Time) (Mara/units.cpp:3-6, :92-141); out holds 33
 * values in the order of maragpu_measures. */
void oracle_measures(const int *N, int Ng, const double *P, double gamma, const double *units, double *out)
{
  // PhysicalUnits, Mara/units.cpp:92-141, and the constants of Mara/eos.hpp:12-15
  const double Length = units[0], Mass = units[1], Time = units[2];
  const double Gram = 1.0 / Mass, Centimeter = 1.0 / Length, Second = 1.0 / Time;
  const double Erg = Gram * std::pow(Centimeter / Second, 2.0);
  const double MeV = 1.60217648740e-06 * Erg;
  const double Gauss = std::sqrt(Erg / std::pow(Centimeter, 3.0) / (4 * M_PI));
  const double GramsPerCC = Gram / std::pow(Centimeter, 3.0);
  const double MeVPerFm3 = MeV / std::pow(1e-13 * Centimeter, 3);
  const double ErgPerCC = MeVPerFm3 * 1e-33;
  const double mp = 1.67262177774e-24 * Gram;
  const double kb = 1.38064881300e-16 * Erg;
  const double Kelvin_to_MeV = 8.621738e-11;          // Mara/eos.cpp:2

  const int n0 = N[0] + 2 * Ng, n1 = N[1] + 2 * Ng, n2 = N[2] + 2 * Ng;
  double Vrms[4] = {0, 0, 0, 0}, Uavg[8] = {0}, Pavg[8] = {0};
  double magnetic = 0, kinetic = 0, internal = 0, total = 0;
  double T_ttl = 0, T_max = 0, Ms_ttl = 0, Ms_max = 0, Ma_ttl = 0, Ma_min = 1e20;
  double B_ttl = 0, B_max = 0, W_max = 0;
  for (int i = Ng; i < n0 - Ng; ++i)
    for (int j = Ng; j < n1 - Ng; ++j)
      for (int k = Ng; k < n2 - Ng; ++k) {
        const double *P0 = P + (((long)i * n1 + j) * n2 + k) * 8;
        double U0[8];
        prim_to_cons(P0, U0, gamma);
        const double v2 = P0[vx]*P0[vx] + P0[vy]*P0[vy] + P0[vz]*P0[vz];
        const double B2 = P0[Bx]*P0[Bx] + P0[By]*P0[By] + P0[Bz]*P0[Bz];
        const double W0 = 1.0 / std::sqrt(1.0 - v2);
        // luaC_mean_velocity
        Vrms[0] += W0;
        Vrms[1] += std::fabs(P0[vx]);
        Vrms[2] += std::fabs(P0[vy]);
        Vrms[3] += std::fabs(P0[vz]);
        for (int q = 0; q < 8; ++q) { Uavg[q] += U0[q]; Pavg[q] += P0[q]; }
        // luaC_mean_energies: T = Temperature_p = p, u = Internal = T/(Gamma-1)
        const double T0 = P0[pre];
        const double u0 = T0 / (gamma - 1);
        const double rhoh = P0[rho] + u0 + P0[pre];
        const double e_f = rhoh * W0*W0 - P0[pre] - U0[ddd];
        const double e_m = U0[tau] - e_f;
        magnetic += e_m;
        kinetic += P0[rho] * W0 * (W0 - 1);
        internal += W0 * u0;
        total += U0[tau];
        // luaC_mean_max_temperature: AdiabaticEos::TemperatureMeV
        {
          double p = P0[pre], r = P0[rho];
          p /= ErgPerCC;
          r /= GramsPerCC;
          const double T_Kelvin = (p / r) * (mp / kb);
          const double TM = T_Kelvin * Kelvin_to_MeV;
          T_ttl += TM;
          if (TM > T_max) T_max = TM;
        }
        // luaC_mean_max_sonic_mach: SoundSpeed2Sr = Gamma T / rhoh
        {
          const double rh = P0[rho] + T0 / (gamma - 1) + T0;
          const double cs2 = gamma * T0 / rh;
          const double M0 = std::sqrt((v2 / (1 - v2)) / (cs2 / (1 - cs2)));
          Ms_ttl += M0;
          if (M0 > Ms_max) Ms_max = M0;
        }
        // luaC_mean_min_alfvenic_mach
        {
          const double va2 = B2 / (B2 + rhoh);
          const double M0 = std::sqrt((v2 / (1 - v2)) / (va2 / (1 - va2)));
          Ma_ttl += M0;
          if (M0 < Ma_min) Ma_min = M0;
        }
        // luaC_mean_max_magnetic_field
        {
          const double B0 = std::sqrt(B2);
          B_ttl += B0;
          if (B0 > B_max) B_max = B0;
        }
        // luaC_max_lorentz_factor
        if (W0 > W_max) W_max = W0;
      }
  const int ttl = N[0] * N[1] * N[2];                   // TTL_ZONES, an int in the reference
  for (int q = 0; q < 4; ++q) out[q] = Vrms[q] / ttl;
  for (int q = 0; q < 8; ++q) out[4 + q] = Uavg[q] / ttl;
  for (int q = 0; q < 8; ++q) out[12 + q] = Pavg[q] / ttl;
  out[20] = magnetic / ttl;
  out[21] = kinetic / ttl;
  out[22] = internal / ttl;
  out[23] = total / ttl;
  out[24] = T_ttl / ttl;
  out[25] = T_max;
  out[26] = Ms_ttl / ttl;
  out[27] = Ms_max;
  out[28] = Ma_ttl / ttl;
  out[29] = Ma_min;
  out[30] = (1.0 / Gauss) * B_ttl / ttl;
  out[31] = (1.0 / Gauss) * B_max;
  out[32] = W_max;
}

/* ---------------------------------------------------------------------------
 * Driving and cooling (Mara/ou-field.cpp, Mara/ornuhl.hpp, Mara/random.cpp,
 * Mara/driving.cpp, Mara/cooling.cpp), restated; fields travel in the reference's
 * serialized byte layout
 * ------------------------------------------------------------------------- */
}  // extern "C"

namespace ou {

/* RandomNumberStream, Mara/random.cpp:15-57 (Numerical Recipes' Ran). The reference
 * scales the 64-bit integer by a float literal, so the product is formed in single
 * precision before it meets (b - a) in double; kept. */
struct Ran {
  unsigned long long u, v, w;
  unsigned long long next()
  {
    u = u * 2862933555777941757ULL + 7046029254386353087ULL;
    v ^= v >> 17; v ^= v << 31; v ^= v >> 8;
    w = 4294957665U * (w & 0xffffffff) + (w >> 32);
    unsigned long long x = u ^ (u << 21);
    x ^= x >> 35; x ^= x << 4;
    return (x + v) ^ w;
  }
  void seed(int n)
  {
    v = 4101842887655102017ULL;
    w = 1;
    u = (unsigned long long)(long long)n ^ v; next();
    v = u; next();
    w = v; next();
  }
  double uniform(double a, double b)
  {
    const float f = (float)next() * 5.42101086242752217E-20f;
    return f * (b - a) + a;
  }
};

/* OrnsteinUhlenbeckProcess, Mara/ornuhl.hpp:20-68; the same 40-byte layout */
struct Proc {
  Ran r;
  double theta, sigma;
};
static_assert(sizeof(Proc) == 40, "layout of the serialized process");

inline double erfinv(double y)
{
  double x = 0.5*std::sqrt(M_PI)*(y + M_PI/12.0*std::pow(y, 3));
  double f, g;
  do {
    f = std::erf(x) - y;
    g = 2.0/std::sqrt(M_PI) * std::exp(-x*x);
    x -= f/g;
  } while (std::fabs(f/g) > 1e-12);
  return x;
}

inline double normal(Proc &p, double sig2)
{
  const double C = p.r.uniform(-1, 1);
  return std::sqrt(2*sig2) * erfinv(C) + 0.0;
}

/* Deviate(complex z, dt): the reference draws its two normals as the arguments of one
 * constructor call; compiled with g++ the imaginary part is drawn first */
inline std::complex<double> deviate(Proc &p, std::complex<double> z, double dt)
{
  const double im = normal(p, dt);
  const double re = normal(p, dt);
  const std::complex<double> dW(re, im);
  return -p.theta*z*dt + p.sigma*dW;
}

/* StochasticVectorField3d, Mara/ou-field.cpp */
struct Field3 {
  int k1 = 0, numk = 0;
  double P0 = 0, zeta = 0;
  std::vector<std::complex<double>> Kx, Ky, Kz, Fx, Fy, Fz;
  std::vector<Proc> procs;

  void build(double P0_, double zeta_, int k1_, int seed)     // :44-74, Stone spectrum
  {
    k1 = k1_; numk = (2*k1 + 1)*(2*k1 + 1)*(2*k1 + 1); P0 = P0_; zeta = zeta_;
    Kx.assign(numk, 0.0); Ky.assign(numk, 0.0); Kz.assign(numk, 0.0);
    Fx.assign(numk, 0.0); Fy.assign(numk, 0.0); Fz.assign(numk, 0.0);
    auto spectrum = [](double k) { return std::pow(k, 6) * std::exp(-8*k/(2*M_PI*2.0)); };
    const int si = (2*k1 + 1)*(2*k1 + 1), sj = 2*k1 + 1;
    double totpower = 0.0;
    for (int i = -k1; i <= k1; ++i)
      for (int j = -k1; j <= k1; ++j)
        for (int k = -k1; k <= k1; ++k) {
          const int m = (i + k1)*si + (j + k1)*sj + (k + k1);
          Kx[m] = 2*M_PI*i; Ky[m] = 2*M_PI*j; Kz[m] = 2*M_PI*k;
          const double K[3] = {Kx[m].real(), Ky[m].real(), Kz[m].real()};
          totpower += spectrum(std::sqrt(K[0]*K[0] + K[1]*K[1] + K[2]*K[2]));
        }
    procs.resize(numk);
    for (int m = 0; m < numk; ++m) {
      const double K[3] = {Kx[m].real(), Ky[m].real(), Kz[m].real()};
      const double Pk = spectrum(std::sqrt(K[0]*K[0] + K[1]*K[1] + K[2]*K[2]));
      procs[m].r.seed(seed + m);
      procs[m].theta = 1.0;
      procs[m].sigma = std::sqrt(2*P0*Pk/totpower);
    }
  }
  size_t bytes() const { return sizeof(int) + 2*sizeof(double) + 6*numk*sizeof(std::complex<double>) + numk*sizeof(Proc); }
  void save(char *b) const                                      // :261-278
  {
    auto put = [&](const void *p, size_t n) { std::memcpy(b, p, n); b += n; };
    const size_t cb = numk*sizeof(std::complex<double>);
    put(&k1, sizeof(int)); put(&P0, sizeof(double)); put(&zeta, sizeof(double));
    put(Kx.data(), cb); put(Ky.data(), cb); put(Kz.data(), cb);
    put(Fx.data(), cb); put(Fy.data(), cb); put(Fz.data(), cb);
    put(procs.data(), numk*sizeof(Proc));
  }
  void load(const char *b)                                      // :108-130
  {
    auto get = [&](void *p, size_t n) { std::memcpy(p, b, n); b += n; };
    get(&k1, sizeof(int));
    numk = (2*k1 + 1)*(2*k1 + 1)*(2*k1 + 1);
    for (auto *v : {&Kx, &Ky, &Kz, &Fx, &Fy, &Fz}) v->assign(numk, 0.0);
    procs.resize(numk);
    const size_t cb = numk*sizeof(std::complex<double>);
    get(&P0, sizeof(double)); get(&zeta, sizeof(double));
    get(Kx.data(), cb); get(Ky.data(), cb); get(Kz.data(), cb);
    get(Fx.data(), cb); get(Fy.data(), cb); get(Fz.data(), cb);
    get(procs.data(), numk*sizeof(Proc));
  }
  void advance(double dt)                                       // :186-210
  {
    for (int m = 0; m < numk; ++m) {
      std::complex<double> dW[3], dV[3];
      const double K[3] = {Kx[m].real(), Ky[m].real(), Kz[m].real()};
      const double kdotk = K[0]*K[0] + K[1]*K[1] + K[2]*K[2];
      if (kdotk < 1e-8) continue;
      dW[0] = deviate(procs[m], Fx[m], dt);
      dW[1] = deviate(procs[m], Fy[m], dt);
      dW[2] = deviate(procs[m], Fz[m], dt);
      for (int p = 0; p < 3; ++p) {
        dV[p] = 0.0;
        for (int q = 0; q < 3; ++q) {
          const double Pp = (1 - 2*zeta)*K[p]*K[q]/kdotk + zeta*(p == q);
          dV[p] += Pp * dW[q] / std::sqrt(1 - 1.25*zeta + 3*zeta*zeta);
        }
      }
      Fx[m] += dV[0]; Fy[m] += dV[1]; Fz[m] += dV[2];
    }
  }
  /* :226-239; the sum of a valarray expression (libstdc++ _Expr::sum) runs from the last
   * element to the first */
  void sample(double x, double y, double z, double *F) const
  {
    const std::complex<double> X(0, x), Y(0, y), Z(0, z);
    std::complex<double> s[3];
    for (int m = numk - 1; m >= 0; --m) {
      const std::complex<double> e = std::exp(Kx[m]*X + Ky[m]*Y + Kz[m]*Z);
      const std::complex<double> t[3] = {Fx[m]*e, Fy[m]*e, Fz[m]*e};
      for (int d = 0; d < 3; ++d) s[d] = m == numk - 1 ? t[d] : s[d] + t[d];
    }
    for (int d = 0; d < 3; ++d) F[d] = s[d].real();
  }
};

}  // namespace ou

extern "C" {

size_t oracle_ou_new(double P0, double zeta, int k1, int seed, char *buf, size_t cap)
{
  ou::Field3 f;
  f.build(P0, zeta, k1, seed);
  if (buf && f.bytes() <= cap) f.save(buf);
  return f.bytes();
}

void oracle_ou_advance(char *buf, size_t n, double dt, int times)
{
  ou::Field3 f;
  f.load(buf);
  if (f.bytes() != n) return;
  for (int t = 0; t < times; ++t) f.advance(dt);
  f.save(buf);
}

void oracle_ou_sample(const char *buf, size_t n, long npts, const double *xyz, double *F)
{
  ou::Field3 f;
  f.load(buf);
  if (f.bytes() != n) return;
  for (long i = 0; i < npts; ++i) f.sample(xyz[3*i], xyz[3*i + 1], xyz[3*i + 2], F + 3*i);
}

/* DrivingProcedure on an AoS primitive array with Ng ghosts over [x0, x1]: the field sampled
 * at every zone centre, ghosts included (ResampleField, Mara/driving.cpp:44-85, x_at of
 * Mara/simple-cart.hpp:42-44), then Drive_rmhd (:154-186); power = AveragePowerInFields */
void oracle_drive(const int *N, int Ng, const double *x0, const double *x1, const char *buf, size_t n,
                  double *P, double dt, double *power)
{
  ou::Field3 f;
  f.load(buf);
  if (f.bytes() != n) return;
  const int n0 = N[0] + 2*Ng, n1 = N[1] + 2*Ng, n2 = N[2] + 2*Ng;
  double dx[3];
  for (int d = 0; d < 3; ++d) dx[d] = (x1[d] - x0[d]) / N[d];
  const long T = (long)n0 * n1 * n2;
  std::vector<double> Fx(T), Fy(T), Fz(T);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n0; ++i)
    for (int j = 0; j < n1; ++j)
      for (int k = 0; k < n2; ++k) {
        const long z = ((long)i*n1 + j)*n2 + k;
        double F[3];
        f.sample(x0[0] + dx[0]*(i - Ng + 0.5), x0[1] + dx[1]*(j - Ng + 0.5), x0[2] + dx[2]*(k - Ng + 0.5), F);
        Fx[z] = F[0]; Fy[z] = F[1]; Fz[z] = F[2];
      }
  for (long z = 0; z < T; ++z) {
    double *P0 = P + 8*z;
    double uf[4] = {0.0, P0[vx], P0[vy], P0[vz]};
    const double v2 = uf[1]*uf[1] + uf[2]*uf[2] + uf[3]*uf[3];
    const double W0 = 1.0 / std::sqrt(1.0 - v2);
    uf[1] *= W0; uf[1] += Fx[z] * (dt/W0);
    uf[2] *= W0; uf[2] += Fy[z] * (dt/W0);
    uf[3] *= W0; uf[3] += Fz[z] * (dt/W0);
    const double u2 = uf[1]*uf[1] + uf[2]*uf[2] + uf[3]*uf[3];
    const double W1 = std::sqrt(1.0 + u2);
    P0[vx] = uf[1]/W1;
    P0[vy] = uf[2]/W1;
    P0[vz] = uf[3]/W1;
    P0[rho] *= W0/W1;
  }
  if (power) {
    // (Fx*Fx + Fy*Fy + Fz*Fz).sum(): a valarray expression, summed from the last zone
    double s = 0.0;
    for (long z = T - 1; z >= 0; --z) {
      const double e = Fx[z]*Fx[z] + Fy[z]*Fy[z] + Fz[z]*Fz[z];
      s = z == T - 1 ? e : s + e;
    }
    *power = s / T;
  }
}

/* CoolingModuleT4 (kind 1) or E4 (kind 2), Mara/cooling.cpp:13-103, over every zone of an AoS
 * primitive array (ghosts included, as there), with the gamma-law EOS in PhysicalUnits
 * units = {Length, Mass, Time}; removed = the energy taken out, summed in zone order */
void oracle_cool(const int *N, int Ng, const double *x0, const double *x1, double gamma, const double *units,
                 int kind, double ref, double t0, double *P, double dt, double *removed)
{
  const double Length = units[0], Mass = units[1], Time = units[2];
  const double Gram = 1.0 / Mass, Centimeter = 1.0 / Length, Second = 1.0 / Time;
  const double Erg = Gram * std::pow(Centimeter / Second, 2.0);
  const double MeV = 1.60217648740e-06 * Erg;
  const double GramsPerCC = Gram / std::pow(Centimeter, 3.0);
  const double ErgPerCC = MeV / std::pow(1e-13 * Centimeter, 3) * 1e-33;
  const double mp = 1.67262177774e-24 * Gram, kb = 1.38064881300e-16 * Erg;
  const double Kelvin_to_MeV = 8.621738e-11;
  const long T = (long)(N[0] + 2*Ng) * (N[1] + 2*Ng) * (N[2] + 2*Ng);
  double dxs[3];
  for (int d = 0; d < 3; ++d) dxs[d] = (x1[d] - x0[d]) / N[d];
  const double dV = dxs[0]*dxs[1]*dxs[2];
  double energy = 0.0;
  for (long z = 0; z < T; ++z) {
    double *P0 = P + 8*z;
    double U0[8], U1[8];
    double T1;
    if (kind == 1) {
      double p = P0[pre], r = P0[rho];                 // TemperatureMeV, Mara/eos.cpp:37-47
      p /= ErgPerCC;
      r /= GramsPerCC;
      const double T0 = (p/r) * (mp/kb) * Kelvin_to_MeV;
      const double T1_MeV = T0 - dt * (ref/t0) * std::pow(T0/ref, 4);
      double rr = P0[rho];                             // TemperatureArb, :48-58
      rr /= GramsPerCC;
      const double T_Kelvin = T1_MeV / Kelvin_to_MeV;
      T1 = ((kb*T_Kelvin) * (rr/mp)) * ErgPerCC;
    }
    else {
      const double T0 = P0[pre];                       // Temperature_p, Internal, Temperature_u
      const double e0 = (T0 / (gamma - 1)) / P0[rho];
      const double e1 = e0 - dt * (ref/t0) * std::pow(e0/ref, 4);
      T1 = (P0[rho]*e1) * (gamma - 1);
    }
    prim_to_cons(P0, U0, gamma);
    P0[pre] = T1;                                      // Pressure(rho, T) = T
    prim_to_cons(P0, U1, gamma);
    energy += (U0[tau] - U1[tau])*dV;
  }
  if (removed) *removed = energy;
}

/* ---------------------------------------------------------------------------
 * Point functions, batched
 * ------------------------------------------------------------------------- */
void oracle_prim_to_cons_batch(long n, const double *P, double *U, double gamma, int *err)
{
#pragma omp parallel for schedule(static)
  for (long z = 0; z < n; ++z) {
    int e = prim_to_cons(P + 8 * z, U + 8 * z, gamma);
    if (err) err[z] = e;
  }
}

/* Rmhd::ConsToPrim gamma-law cascade; P holds the previous prim state (guess) on entry */
void oracle_cons_to_prim_batch(long n, const double *U, double *P, double gamma, double pfloor,
                               int *err, int *iters)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (long z = 0; z < n; ++z) {
    std::int64_t it = 0;
    int e = cons_to_prim(U + 8 * z, P + 8 * z, gamma, pfloor, &it);
    if (err) err[z] = e;
    if (iters) iters[z] = (int)it;
  }
}

/* Single anton2dzw / noble1dw solves from the cons-based estimate (mara.cpp:1819-1879) */
void oracle_c2p_solver_batch(long n, int which, const double *U, double *P, double gamma,
                             double pfloor, int *err, int *iters)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (long z = 0; z < n; ++z) {
    C2P<double> s(gamma, pfloor);
    s.new_state(U + 8 * z);
    s.estimate_from_cons();
    int e = which == 0 ? s.solve_anton2dzw(P + 8 * z) : s.solve_noble1dw(P + 8 * z);
    if (err) err[z] = e;
    if (iters) iters[z] = s.iterations;
  }
}

/* anton2dzw from a given Newton start (Z, W) per state (start: 2n doubles); the
 * reference reaches the same solve through rmhd_c2p_set_starting_prim (Mara/rmhd-c2p.c:108-119),
 * which sets Z_start and W_start from a primitive state. Checker for the learned guess. */
void oracle_c2p_from_start_batch(long n, const double *U, const double *start, double *P, double gamma,
                                 double pfloor, int *err, int *iters)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (long z = 0; z < n; ++z) {
    C2P<double> s(gamma, pfloor);
    s.new_state(U + 8 * z);
    s.Z_start = start[2 * z + 0];
    s.W_start = start[2 * z + 1];
    int e = s.solve_anton2dzw(P + 8 * z);
    if (err) err[z] = e;
    if (iters) iters[z] = s.iterations;
  }
}

/* RMHD characteristic eigensystem of n primitive states (L, R: 64 doubles each, row-major;
 * lam: 8; ok: 1 where usable) */
void oracle_eigensystem_batch(long n, const double *P, double *L, double *R, double *lam, int *ok,
                              int dim, double gamma)
{
#pragma omp parallel for schedule(dynamic, 16)
  for (long z = 0; z < n; ++z) {
    double U[8];
    prim_to_cons(P + 8 * z, U, gamma);
    ok[z] = rmhd_eigensystem(U, P + 8 * z, L + 64 * z, R + 64 * z, lam + 8 * z, dim, gamma) == 0;
  }
}

void oracle_flux_eigen_batch(long n, const double *P, double *F, double *ap, double *am,
                             int dim, double gamma)
{
#pragma omp parallel for schedule(static)
  for (long z = 0; z < n; ++z) {
    double U[8];
    prim_to_cons(P + 8 * z, U, gamma);
    flux_and_eigenvalues(U, P + 8 * z, F + 8 * z, ap + z, am + z, dim, gamma);
  }
}

int oracle_quartic(double d4, double d3, double d2, double d1, double d0, double *roots)
{
  return solve_quartic(d4, d3, d2, d1, d0, roots);
}

void oracle_reconstruct_batch(long n, const double *V, double *out, int op, double theta,
                              int IS, double sz10A)
{
  PhysParams pp;
  pp.theta = theta;
  pp.IS = IS;
  pp.sz10A = sz10A;
  for (long z = 0; z < n; ++z) out[z] = reconstruct(V + 6 * z + 2, op, pp);
}

/* One Riemann solve per pair; lam receives each solve's max |lambda| (0 if none) */
void oracle_riemann_batch(long n, int solver, const double *PL, const double *PR, double *F,
                          int dim, double gamma, double pfloor, int *err, double *lam,
                          int *hlld_ok)
{
  PhysParams pp;
  pp.gamma = gamma;
  pp.pfloor = pfloor;
#pragma omp parallel for schedule(dynamic, 64)
  for (long z = 0; z < n; ++z) {
    Tally<double> t;
    for (int q = 0; q < 8; ++q) F[8 * z + q] = 0.0;
    int e = riemann_flux(solver, PL + 8 * z, PR + 8 * z, F + 8 * z, dim, pp, t);
    if (err) err[z] = e;
    if (lam) lam[z] = t.max_lambda;
    if (hlld_ok) hlld_ok[z] = (int)t.hlld_success;
  }
}

/* Flop counts on the oracle's own path with the Counted scalar (same roundings):
 * reconstruction + Riemann solve per interface over `nsample` interfaces of each axis
 * spread evenly through the reference's sweep range, and the c2p cascade (zero
 * previous-prim guess, as in the first RK stage) over `nsample` zones.
 * out[0..2] = mean flops per interface for x, y, z; out[3] = mean c2p flops per zone;
 * out[4..6] = mean transcendental calls per interface; out[7] = c2p transcendentals;
 * out[8..22] = per axis (add, mul, div, sqrt, transcendental) means; out[23..27] the same
 * for the c2p (out must hold 28 doubles). device_path != 0 counts the device's HLLD,
 * whose nested c2p starts from the average of the two states (PhysParams). */
void oracle_count_flops(void *h, const double *P_synced, long nsample, double *out, int device_path)
{
  Grid &g = *static_cast<Grid *>(h);
  PhysParams pp_saved = g.pp;
  g.pp.nested_c2p_from_average = device_path != 0;
  const long s0 = 8 * g.T;
  for (int dim = 1; dim <= 3; ++dim) {
    const long S = dim == 1 ? g.sx : dim == 2 ? g.sy : g.sz;
    const long first = 2 * S / 8, last = (s0 - 3 * S) / 8;  // zone range
    const long span = last - first;
    FlopTally total;
    for (long s = 0; s < nsample; ++s) {
      const long zone = first + (span * s) / nsample;
      const long i = 8 * zone;
      g_tally = FlopTally();
      Counted Pl[8], Pr[8], F[8];
      for (int q = 0; q < 8; ++q) {
        const long m = i + q;
        Counted v[6] = { P_synced[m-2*S], P_synced[m-S], P_synced[m], P_synced[m+S],
                         P_synced[m+2*S], P_synced[m+3*S] };
        switch (g.cfg.recon) {
        case RECON_PCM: Pl[q] = v[2]; Pr[q] = v[3]; break;
        case RECON_PLM:
          Pl[q] = reconstruct(&v[2], PLM_C2R, g.pp); Pr[q] = reconstruct(&v[3], PLM_C2L, g.pp); break;
        default:
          Pl[q] = reconstruct(&v[2], WENO5_FV_C2R, g.pp); Pr[q] = reconstruct(&v[3], WENO5_FV_C2L, g.pp); break;
        }
      }
      Tally<Counted> t;
      if (riemann_flux(g.cfg.riemann, Pl, Pr, F, dim, g.pp, t)) {
        Counted a[8], b[8];
        for (int q = 0; q < 8; ++q) { a[q] = P_synced[i + q]; b[q] = P_synced[i + S + q]; }
        riemann_flux(g.cfg.riemann, a, b, F, dim, g.pp, t);
      }
      total.add += g_tally.add; total.mul += g_tally.mul; total.div += g_tally.div;
      total.sqrt += g_tally.sqrt; total.trans += g_tally.trans;
    }
    out[dim - 1] = (double)total.flops() / nsample;
    out[3 + dim] = (double)total.trans / nsample;
    // per-category detail: out[8 + 5*(dim-1) + {0 add, 1 mul, 2 div, 3 sqrt, 4 trans}]
    out[8 + 5 * (dim - 1) + 0] = (double)total.add / nsample;
    out[8 + 5 * (dim - 1) + 1] = (double)total.mul / nsample;
    out[8 + 5 * (dim - 1) + 2] = (double)total.div / nsample;
    out[8 + 5 * (dim - 1) + 3] = (double)total.sqrt / nsample;
    out[8 + 5 * (dim - 1) + 4] = (double)total.trans / nsample;
  }
  FlopTally total;
  const long nzone = g.T;
  for (long s = 0; s < nsample; ++s) {
    const long z = (nzone * s) / nsample;
    Counted P[8], U[8], P0[8];
    for (int q = 0; q < 8; ++q) P[q] = P_synced[8 * z + q];
    prim_to_cons(P, U, g.pp.gamma);
    g_tally = FlopTally();
    std::int64_t it = 0;
    cons_to_prim(U, P0, g.pp.gamma, g.pp.pfloor, &it);
    total.add += g_tally.add; total.mul += g_tally.mul; total.div += g_tally.div;
    total.sqrt += g_tally.sqrt; total.trans += g_tally.trans;
  }
  out[3] = (double)total.flops() / nsample;
  out[7] = (double)total.trans / nsample;
  out[23] = (double)total.add / nsample;
  out[24] = (double)total.mul / nsample;
  out[25] = (double)total.div / nsample;
  out[26] = (double)total.sqrt / nsample;
  out[27] = (double)total.trans / nsample;
  g.pp = pp_saved;
}

}  // extern "C"
