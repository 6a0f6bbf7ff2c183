/* ORACLE — test infrastructure only.
 *
 * C-ABI harness over the compiled reference (Mara's own sources, built from where they
 * lie under the reference tree by oracle/Makefile into oracle/_ref/). It stands in for the
 * Lua host: the set-up sequence of ctf-core/run-mara.lua:113-148 and the body of
 * luaC_Mara_advance (Mara/mara.cpp:297-341) are restated here against the reference's
 * classes, so every number it returns is computed by the reference's code.
 * Function names and arguments match liboracle.so (prefix ref_ instead of oracle_).
 */
#include "mara.hpp"
#include "mara_oracle.hpp"   // rmhd_eigensystem: the RMHD eigensystem the reference lacks

#include <cmath>
#include <cstring>
#include <sstream>
#include <string>
#include <valarray>

extern "C" {

typedef struct {
  int N[3];
  int Ng;
  double x0[3], x1[3];
  double gamma;
  int recon;
  double theta;
  int IS;
  double sz10A;
  int riemann;
  int rk;
  double pfloor;
  int ereset;
  double preset, dreset;
  int godunov;    /* 0 plm-split, 1 weno-split */
  int eigen;      /* weno-split projection: 0 component-wise, 1 RMHD characteristic */
} oracle_config;

typedef struct {
  double max_lambda;
  double hlld_success;
  double hlld_failure;
  double c2p_iterations;
  double c2p_solves;
  double riemann_retries;
  double riemann_failures;
  double reserved;
} oracle_stats;

}  // extern "C"

static MaraApplication *App = nullptr;

/* The reference's RMHD fluid with an Eigensystem: FluidEquations::Eigensystem
 * (Mara/hydro.hpp:146-147) is empty for it, so WenoSplit would read uninitialised
 * matrices. Mode 0 supplies the identity (component-wise splitting), mode 1 the oracle's
 * RMHD eigensystem (identity where that is ill-conditioned), which lets the reference's own
 * weno-split sweep run on RMHD. */
class RmhdWithEigensystem : public AdiabaticIdealRmhd {
public:
  int mode = 0;
  double gamma = 4.0 / 3.0;
  void Eigensystem(const double *U, const double *P, double *L, double *R, double *lam,
                   int dim) const override
  {
    if (mode == 1 && oracle::rmhd_eigensystem(U, P, L, R, lam, dim, gamma) == 0) return;
    for (int i = 0; i < 8; ++i) {
      lam[i] = 0.0;
      for (int j = 0; j < 8; ++j) L[i*8 + j] = R[i*8 + j] = (i == j) ? 1.0 : 0.0;
    }
  }
};

static void ref_start(double gamma, bool with_eigensystem = false)
{
  if (App) delete App;
  App = new MaraApplication;
  HydroModule::Mara = App;
  delete App->fluid;
  // HllcRiemannSolver dispatches on typeid(fluid) == typeid(AdiabaticIdealRmhd)
  // (Mara/riemann_hllc.cpp:11-21), so the subclass is used only where no Riemann solver runs
  if (with_eigensystem) App->fluid = new RmhdWithEigensystem;
  else App->fluid = new AdiabaticIdealRmhd;
  delete App->eos;
  App->eos = new AdiabaticEos(gamma);
  rmhd_c2p_set_gamma(gamma);
}

static void ref_params(double theta, int IS, double sz10A, double pfloor)
{
  reconstruct_set_plm_theta(theta);
  reconstruct_set_smoothness_indicator(IS == 1 ? ImprovedBorges08
                                       : IS == 2 ? ImprovedShenZha10 : OriginalJiangShu96);
  reconstruct_set_shenzha10_A(sz10A);
  rmhd_c2p_set_pressure_floor(pfloor);
}

static RiemannSolver *make_riemann(int which)
{
  if (which == 0) return new HllRiemannSolver;
  if (which == 1) return new HllcRiemannSolver;
  return new HlldRmhdRiemannSolver;
}

extern "C" {

void *ref_create(const oracle_config *c)
{
  ref_start(c->gamma, c->godunov == 1);
  if (c->godunov == 2) {
    App->godunov = new PlmCtuHancockOperator;
    App->godunov->SetPlmTheta(c->theta);        // the operator's own limiter parameter
  }
  else if (c->godunov == 1) {
    App->godunov = new WenoSplit;
    GodunovOperator::fluxsplit_method = GodunovOperator::FLUXSPLIT_LOCAL_LAX_FRIEDRICHS;
    static_cast<RmhdWithEigensystem *>(App->fluid)->mode = c->eigen;
    static_cast<RmhdWithEigensystem *>(App->fluid)->gamma = c->gamma;
  }
  else App->godunov = new MethodOfLinesSplit;
  App->boundary = new PeriodicBoundary;
  App->riemann = make_riemann(c->riemann);
  switch (c->rk) {
  case 1: App->advance = new RungeKuttaSingleStep; break;
  case 2: App->advance = new RungeKuttaRk2Tvd; break;
  case 4: App->advance = new RungeKuttaClassicRk4; break;
  default: App->advance = new RungeKuttaShuOsherRk3; break;
  }
  GodunovOperator::reconstruct_method =
      c->recon == 0 ? GodunovOperator::RECONSTRUCT_PCM
      : c->recon == 1 ? GodunovOperator::RECONSTRUCT_PLM : GodunovOperator::RECONSTRUCT_WENO5;
  ref_params(c->theta, c->IS, c->sz10A, c->pfloor);
  if (c->ereset) App->godunov->enable_emergency_reset();
  else App->godunov->disable_emergency_reset();
  App->godunov->set_reset_pressure(c->preset);
  App->godunov->set_reset_density(c->dreset);
  App->domain = new SimpleCartesianDomain(c->x0, c->x1, c->N, 3, 8, c->Ng);
  RiemannSolver::ResetMaxLambda();
  return App;
}

void ref_destroy(void *) {}

long ref_num_zones(void *) { return App->domain->GetNumberOfZones(); }

int ref_advance(void *, double *UserPrim, double dt, oracle_stats *stats)
{
  const unsigned N = App->domain->GetNumberOfZones() * App->domain->get_Nq();
  std::valarray<double> P(N);
  std::valarray<double> U(N);
  int errors;
  const int s0 = HlldRmhdRiemannSolver::NumberOfSuccess;
  const int f0 = HlldRmhdRiemannSolver::NumberOfFailure;

  std::memcpy(&P[0], UserPrim, N * sizeof(double));
  RiemannSolver::ResetMaxLambda();
  App->FailureMask.resize(App->domain->GetNumberOfZones());
  App->PrimitiveArray.resize(P.size());
  App->godunov->PrimToCons(P, U);
  try {
    App->advance->AdvanceState(U, dt);
    errors = App->godunov->ConsToPrim(U, P);
  }
  catch (const GodunovOperator::IntermediateFailure &e) {
    errors = App->FailureMask.sum();
  }
  if (errors == 0) std::memcpy(UserPrim, &P[0], N * sizeof(double));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->max_lambda = RiemannSolver::GetMaxLambda();
    stats->hlld_success = HlldRmhdRiemannSolver::NumberOfSuccess - s0;
    stats->hlld_failure = HlldRmhdRiemannSolver::NumberOfFailure - f0;
  }
  return errors;
}

double ref_get_timestep(void *, double cfl)
{
  return cfl * App->domain->get_min_dx() / RiemannSolver::GetMaxLambda();
}

void ref_prim_to_cons_batch(long n, const double *P, double *U, double gamma, int *err)
{
  ref_start(gamma);
  for (long z = 0; z < n; ++z) {
    int e = App->fluid->PrimToCons(P + 8 * z, U + 8 * z);
    if (err) err[z] = e;
  }
}

void ref_cons_to_prim_batch(long n, const double *U, double *P, double gamma, double pfloor,
                            int *err, int *iters)
{
  ref_start(gamma);
  rmhd_c2p_set_pressure_floor(pfloor);
  for (long z = 0; z < n; ++z) {
    int e = App->fluid->ConsToPrim(U + 8 * z, P + 8 * z);
    if (err) err[z] = e;
    if (iters) iters[z] = rmhd_c2p_get_iterations();
  }
}

void ref_c2p_solver_batch(long n, int which, const double *U, double *P, double gamma,
                          double pfloor, int *err, int *iters)
{
  rmhd_c2p_set_gamma(gamma);
  rmhd_c2p_set_pressure_floor(pfloor);
  for (long z = 0; z < n; ++z) {
    rmhd_c2p_new_state(U + 8 * z);
    rmhd_c2p_estimate_from_cons();
    int e = which == 0 ? rmhd_c2p_solve_anton2dzw(P + 8 * z) : rmhd_c2p_solve_noble1dw(P + 8 * z);
    if (err) err[z] = e;
    if (iters) iters[z] = rmhd_c2p_get_iterations();
  }
}

void ref_flux_eigen_batch(long n, const double *P, double *F, double *ap, double *am, int dim,
                          double gamma)
{
  ref_start(gamma);
  for (long z = 0; z < n; ++z) {
    double U[8];
    App->fluid->PrimToCons(P + 8 * z, U);
    App->fluid->FluxAndEigenvalues(U, P + 8 * z, F + 8 * z, ap + z, am + z, dim);
  }
}

int ref_quartic(double d4, double d3, double d2, double d1, double d0, double *roots)
{
  QuarticEquation q(d4, d3, d2, d1, d0);
  return q.Solve(roots);
}

void ref_reconstruct_batch(long n, const double *V, double *out, int op, double theta, int IS,
                           double sz10A)
{
  ref_params(theta, IS, sz10A, -1.0);
  for (long z = 0; z < n; ++z) out[z] = reconstruct(V + 6 * z + 2, (ReconstructOperation)op);
}

void ref_riemann_batch(long n, int solver, const double *PL, const double *PR, double *F,
                       int dim, double gamma, double pfloor, int *err, double *lam, int *hlld_ok)
{
  ref_start(gamma);
  rmhd_c2p_set_pressure_floor(pfloor);
  RiemannSolver *rs = make_riemann(solver);
  for (long z = 0; z < n; ++z) {
    RiemannSolver::ResetMaxLambda();
    const int s0 = HlldRmhdRiemannSolver::NumberOfSuccess;
    for (int q = 0; q < 8; ++q) F[8 * z + q] = 0.0;
    int e = rs->IntercellFlux(PL + 8 * z, PR + 8 * z, 0, F + 8 * z, 0.0, dim);
    if (err) err[z] = e;
    if (lam) lam[z] = RiemannSolver::GetMaxLambda();
    if (hlld_ok) hlld_ok[z] = HlldRmhdRiemannSolver::NumberOfSuccess - s0;
  }
  delete rs;
}

/* ---------------------------------------------------------------------------
 * Driving and cooling: the reference's StochasticVectorField3d (Mara/ou-field.cpp),
 * DrivingProcedure (Mara/driving.cpp) and CoolingModuleT4/E4 (Mara/cooling.cpp) on the
 * domain and fluid of the last ref_create. Fields travel as the reference's own
 * serialized bytes.
 * ------------------------------------------------------------------------- */
}  // extern "C"

// the power spectrum luaC_new_ou_field passes (Mara/mara.cpp:1373-1377): k^6 e^(-8k/(2 pi k_peak)), k_peak = 2
static double harness_stone_spectrum(double k)
{
  const double k_peak = 2.0;
  return std::pow(k, 6) * std::exp(-8 * k / (2 * M_PI * k_peak));
}

static size_t put_bytes(const std::string &d, char *buf, size_t cap)
{
  if (buf && d.size() <= cap) std::memcpy(buf, d.data(), d.size());
  return d.size();
}

extern "C" {

size_t ref_ou_new(double P0, double zeta, int k1, int seed, char *buf, size_t cap)
{
  StochasticVectorField3d f(P0, zeta, k1, seed, harness_stone_spectrum);
  std::stringstream out;
  f.Serialize(out);
  return put_bytes(out.str(), buf, cap);
}

void ref_ou_advance(char *buf, size_t n, double dt, int times)
{
  std::stringstream in(std::string(buf, n));
  StochasticVectorField3d f(in);
  for (int t = 0; t < times; ++t) f.AdvanceField(dt);
  std::stringstream out;
  f.Serialize(out);
  put_bytes(out.str(), buf, n);
}

void ref_ou_sample(const char *buf, size_t n, long npts, const double *xyz, double *F)
{
  std::stringstream in(std::string(buf, n));
  StochasticVectorField3d f(in);
  for (long i = 0; i < npts; ++i) {
    const std::vector<double> v = f.SampleField(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    for (int d = 0; d < 3; ++d) F[3 * i + d] = v[d];
  }
}

/* DrivingProcedure built on the field (it samples every zone of the domain), then Drive on
 * the AoS primitives P; power = AveragePowerInFields */
void ref_drive(const char *buf, size_t n, double *P, double dt, double *power)
{
  std::stringstream in(std::string(buf, n));
  DrivingProcedure d(new StochasticVectorField3d(in));
  const size_t N = App->domain->GetNumberOfZones() * App->domain->get_Nq();
  std::valarray<double> V(P, N);
  d.Drive(V, dt);
  std::memcpy(P, &V[0], N * sizeof(double));
  if (power) *power = d.AveragePowerInFields();
}

/* CoolingModuleT4 (kind 1) or CoolingModuleE4 (kind 2) on the AoS primitives P; removed =
 * EnergyRemoved after the call */
void ref_cool(int kind, double ref, double t0, double *P, double dt, double *removed)
{
  CoolingModule *c = kind == 1 ? (CoolingModule *)new CoolingModuleT4(ref, t0)
                               : (CoolingModule *)new CoolingModuleE4(ref, t0);
  const size_t N = App->domain->GetNumberOfZones() * App->domain->get_Nq();
  std::valarray<double> V(P, N);
  c->Cool(V, dt);
  std::memcpy(P, &V[0], N * sizeof(double));
  if (removed) *removed = c->EnergyRemoved();
  delete c;
}

}  // extern "C"
