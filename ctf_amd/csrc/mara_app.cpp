/* Mara's Lua-facing application API (luaopen_Mara, Mara/mara.cpp:137-268) restated in
 * C++ over the maragpu device context. Mara keeps one MaraApplication per process and
 * every verb edits it; here the verbs edit a maragpu_config and the device context is
 * (re)built on first use after a change. Errors that the Lua binding reports with
 * luaL_error return MARA_E_* with a message in mara_last_error().
 */
#include <dlfcn.h>

#include <chrono>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "../../include/mara_host.h"
#include "../../include/maragpu.h"

namespace {

struct App {
  bool started = false;
  bool fluid_rmhd = false;   // MaraApplication's default fluid is Euler (hydro.cpp:22)
  bool have_domain = false;
  bool have_godunov = false;
  bool have_boundary = false;
  bool have_riemann = false;
  bool have_advance = false;
  maragpu_config cfg;
  maragpu_ctx *ctx = nullptr;
  maragpu_mlp *mlp = nullptr;      // learned c2p network, loaded on first use
  std::string mlp_path;            // config_solver key 'c2p_model'; default models/c2p_mlp.bin
  bool dirty = true;
  double last_max_lambda = 0.0;
  double units[3] = {1.0, 1.0, 1.0};   // PhysicalUnits(Length, Mass, Time): cgs, hydro.cpp:20
  // Mara->PrimitiveArray: the array of the last set_primitive, which the measures read
  // (advance does not change it, Mara/mara.cpp:297-341)
  std::vector<double> prim;
  bool prim_on_device = false;         // the context's committed state holds `prim`
  // driving: the field (advanced by driving_advance) and its bytes when last sampled
  // (set_driving, driving_resample), which is what the context drives with
  maragpu_ou_field *field = nullptr;
  std::vector<unsigned char> sampled;
  // cooling: 0 none, 1 T4, 2 E4, and the energy removed by contexts since destroyed
  int cool_kind = 0;
  double cool_ref = 0.0, cool_t0 = 0.0;
  double removed_pending = 0.0;
  std::string err;

  App()
  {
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.gamma = 1.4;                 // MaraApplication default eos, hydro.cpp:23
    cfg.recon = MARAGPU_PLM;         // GodunovOperator::reconstruct_method default, hydro.cpp:70
    cfg.theta = 2.0;                 // weno.c:10
    cfg.IS = MARAGPU_JS96;           // weno.c:12
    cfg.sz10A = 50.0;                // weno.c:11
    cfg.riemann = MARAGPU_HLLC;
    cfg.rk = 3;
    cfg.precision = MARAGPU_FP64;
    cfg.c2p = MARAGPU_C2P_PLAIN;
    cfg.pfloor = -1.0;               // rmhd-c2p.c:54
    cfg.ereset = 0;
    cfg.preset = 1.0;                // GodunovOperator(), hydro.cpp:75-78
    cfg.dreset = 1.0;
  }
};

App *g_app = nullptr;
// the CTU operator's limiter is a file-scope static of the reference (Mara/ctu-hancock.cpp:12),
// shared by every PlmCtuHancockOperator and kept across Mara.start() for the process
double g_ctu_theta = 2.0;
std::string g_err;

int fail(const std::string &msg, int code = MARAGPU_E_INVALID)
{
  g_err = msg;
  return code;
}

App *app()
{
  return g_app;
}

int need_app()
{
  if (!g_app) return fail("[Mara] call Mara.start() first");
  return 0;
}

void mark_dirty()
{
  if (g_app) g_app->dirty = true;
}

/* The network file: the 'c2p_model' path, else models/c2p_mlp.bin next to this library. */
int load_mlp(App *a)
{
  std::string path = a->mlp_path;
  if (path.empty()) {
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&mara_start), &info) && info.dli_fname) {
      std::string lib = info.dli_fname;
      const size_t slash = lib.rfind('/');
      path = (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/models/c2p_mlp.bin";
    }
  }
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail("[Mara] c2p='learned' needs a network file; none at '" + path + "'", MARAGPU_E_NOMODEL);
  std::vector<char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  const int rc = maragpu_mlp_load(&a->mlp, buf.data(), buf.size());
  if (rc) return fail("[Mara] cannot load the c2p network '" + path + "': " + maragpu_strerror(rc), rc);
  return 0;
}

/* The device context for the current configuration. An advance needs the whole solver
 * set up; a measure only the domain and the fluid (the solver fields keep their
 * defaults). */
int ensure_ctx(bool for_advance = true)
{
  App *a = app();
  if (!a->have_domain) return fail("[Mara] need a domain to run this, use set_domain");
  if (!a->fluid_rmhd) return fail("[Mara] this library runs the 'rmhd' fluid only; use set_fluid('rmhd')");
  if (for_advance) {
    if (!a->have_godunov) return fail("[Mara] need a godunov operator, use set_godunov");
    // WenoSplit splits the fluxes itself (Mara/weno-split.cpp:163-207); only plm-split calls
    // the Riemann solver
    if (!a->have_riemann && a->cfg.godunov != MARAGPU_WENO_SPLIT)
      return fail("[Mara] need a riemann solver, use set_riemann");
    if (!a->have_advance) return fail("[Mara] need an advance option, use set_advance");
    if (!a->have_boundary) return fail("[Mara] need boundary conditions, use set_boundary");
  }
  if (!a->dirty && a->ctx) return 0;
  if (a->ctx) {
    double e = 0.0;
    if (maragpu_cooling_energy_removed(a->ctx, &e) == MARAGPU_OK) a->removed_pending += e;
    maragpu_destroy(a->ctx);
    a->ctx = nullptr;
    a->prim_on_device = false;
  }
  maragpu_config cc = a->cfg;
  bool stand_in = false;
  if (cc.godunov == MARAGPU_PLM_MUSCL) {
    cc.theta = g_ctu_theta;
    if (cc.precision != MARAGPU_FP64 || cc.c2p != MARAGPU_C2P_PLAIN) {
      if (for_advance) return fail("[Mara] plm-muscl runs in precision 'fp64' with the 'plain' c2p");
      // a measure needs no operator: a plm-split context serves it, rebuilt for the next advance
      cc.godunov = MARAGPU_PLM_SPLIT;
      stand_in = true;
    }
  }
  int rc = maragpu_create(&cc, &a->ctx);
  if (rc) return fail(std::string("[Mara] device context: ") + maragpu_strerror(rc), rc);
  if (!a->sampled.empty()) {
    maragpu_ou_field *f = nullptr;
    rc = maragpu_ou_field_load(a->sampled.data(), a->sampled.size(), &f);
    if (!rc) rc = maragpu_set_driving(a->ctx, f);
    maragpu_ou_field_destroy(f);
    if (rc) return fail(std::string("[Mara] driving: ") + maragpu_strerror(rc), rc);
  }
  if (a->cfg.c2p == MARAGPU_C2P_LEARNED) {
    if (!a->mlp && (rc = load_mlp(a))) return rc;
    maragpu_attach_mlp(a->ctx, a->mlp);
  }
  a->dirty = stand_in;
  return 0;
}

size_t zone_doubles(const App *a)
{
  size_t n = 8;
  for (int d = 0; d < 3; ++d) n *= (size_t)(a->cfg.N[d] + 2 * a->cfg.Ng);
  return n;
}

}  // namespace

extern "C" {

const char *mara_last_error(void) { return g_err.c_str(); }

int mara_start(void)
{
  mara_close();
  g_app = new App;
  g_app->started = true;
  return 0;
}

int mara_close(void)
{
  if (g_app) {
    if (g_app->ctx) maragpu_destroy(g_app->ctx);
    if (g_app->mlp) maragpu_mlp_destroy(g_app->mlp);
    if (g_app->field) maragpu_ou_field_destroy(g_app->field);
    delete g_app;
    g_app = nullptr;
  }
  return 0;
}

int mara_set_fluid(const char *key)
{
  if (int rc = need_app()) return rc;
  if (!key) return fail("[Mara] no such fluid '(null)'");
  if (std::strcmp(key, "rmhd") == 0) {
    app()->fluid_rmhd = true;
    mark_dirty();
    return 0;
  }
  if (std::strcmp(key, "euler") == 0 || std::strcmp(key, "srhd") == 0)
    return fail(std::string("[Mara] fluid '") + key + "' is not supported by this library (rmhd only)");
  return fail(std::string("[Mara] no such fluid '") + key + "'");
}

int mara_set_eos(const char *key, double gamma)
{
  if (int rc = need_app()) return rc;
  if (key && std::strcmp(key, "gamma-law") == 0) {
    if (!(gamma > 1.0)) return fail("[Mara] gamma-law index must exceed 1");
    app()->cfg.gamma = gamma;
    mark_dirty();
    return 0;
  }
  if (key && (std::strcmp(key, "gamma-law + polytrope") == 0 || std::strcmp(key, "shen") == 0 ||
              std::strcmp(key, "tabulated") == 0))
    return fail(std::string("[Mara] eos '") + key + "' is not supported by this library (gamma-law only)");
  return fail(std::string("[Mara] no such eos '") + (key ? key : "(null)") + "'");
}

int mara_set_boundary(const char *key)
{
  if (int rc = need_app()) return rc;
  if (key && std::strcmp(key, "periodic") == 0) {
    app()->have_boundary = true;
    return 0;
  }
  if (key && (std::strcmp(key, "outflow") == 0 || std::strcmp(key, "perxouty") == 0 ||
              std::strcmp(key, "reflect2d") == 0 || std::strcmp(key, "magnetic-bubble") == 0 ||
              std::strcmp(key, "wind-inflow") == 0))
    return fail(std::string("[Mara] boundary '") + key + "' is not supported by this library (periodic only)");
  return fail(std::string("[Mara] no such boundary conditions '") + (key ? key : "(null)") + "'");
}

int mara_set_riemann(const char *key)
{
  if (int rc = need_app()) return rc;
  int r = -1;
  if (key && std::strcmp(key, "hll") == 0) r = MARAGPU_HLL;
  if (key && std::strcmp(key, "hllc") == 0) r = MARAGPU_HLLC;
  if (key && std::strcmp(key, "hlld") == 0) r = MARAGPU_HLLD;
  if (r < 0) return fail(std::string("[Mara] no such riemann solver '") + (key ? key : "(null)") + "'");
  app()->cfg.riemann = r;
  app()->have_riemann = true;
  mark_dirty();
  return 0;
}

int mara_set_godunov(const char *key)
{
  if (int rc = need_app()) return rc;
  if (key && std::strcmp(key, "plm-split") == 0) {
    app()->have_godunov = true;
    app()->cfg.godunov = MARAGPU_PLM_SPLIT;
    mark_dirty();
    return 0;
  }
  if (key && std::strcmp(key, "weno-split") == 0) {
    app()->have_godunov = true;
    app()->cfg.godunov = MARAGPU_WENO_SPLIT;
    mark_dirty();
    return 0;
  }
  if (key && std::strcmp(key, "plm-muscl") == 0) {
    app()->have_godunov = true;
    app()->cfg.godunov = MARAGPU_PLM_MUSCL;
    mark_dirty();
    return 0;
  }
  return fail(std::string("[Mara] no such integration scheme '") + (key ? key : "(null)") + "'");
}

int mara_set_advance(const char *key)
{
  if (int rc = need_app()) return rc;
  int r = 0;
  if (key && std::strcmp(key, "single") == 0) r = 1;
  if (key && std::strcmp(key, "rk2") == 0) r = 2;
  if (key && std::strcmp(key, "rk3") == 0) r = 3;
  if (key && std::strcmp(key, "rk4") == 0) r = 4;
  if (!r) return fail(std::string("[Mara] no such advance option '") + (key ? key : "(null)") + "'");
  app()->cfg.rk = r;
  app()->have_advance = true;
  mark_dirty();
  return 0;
}

/* luaC_config_solver, one table entry at a time; unknown keys are ignored as there */
int mara_config_solver_str(const char *key, const char *value)
{
  if (int rc = need_app()) return rc;
  if (!key || !value) return fail("[Mara] config_solver: null key or value");
  maragpu_config &c = app()->cfg;
  if (std::strcmp(key, "fsplit") == 0) {
    if (!std::strcmp(value, "llf")) return 0;   // GodunovOperator's default, hydro.cpp:72-73
    if (!std::strcmp(value, "marq"))
      // Mara/weno-split.cpp:73-75 hard-codes five fields for the Marquina branch
      return fail("[Mara] fsplit 'marq' is written for five fields and cannot run rmhd; use 'llf'");
    return fail(std::string("[Mara] no such fsplit '") + value + "'");
  }
  if (std::strcmp(key, "extrap") == 0) {
    if (!std::strcmp(value, "pcm")) c.recon = MARAGPU_PCM;
    else if (!std::strcmp(value, "plm")) c.recon = MARAGPU_PLM;
    else if (!std::strcmp(value, "weno5")) c.recon = MARAGPU_WENO5;
    else return fail(std::string("[Mara] no such extrap '") + value + "'");
    mark_dirty();
    return 0;
  }
  if (std::strcmp(key, "IS") == 0) {
    if (!std::strcmp(value, "js96")) c.IS = MARAGPU_JS96;
    else if (!std::strcmp(value, "b08")) c.IS = MARAGPU_B08;
    else if (!std::strcmp(value, "sz10")) c.IS = MARAGPU_SZ10;
    else return fail(std::string("no such IS: ") + value);
    mark_dirty();
    return 0;
  }
  if (std::strcmp(key, "precision") == 0) {
    if (!std::strcmp(value, "fp64")) c.precision = MARAGPU_FP64;
    else if (!std::strcmp(value, "mixed")) c.precision = MARAGPU_MIXED;
    else return fail(std::string("[Mara] no such precision '") + value + "'");
    mark_dirty();
    return 0;
  }
  if (std::strcmp(key, "eigensystem") == 0) {
    // the weno-split projection; the reference's RMHD fluid has none (Mara/hydro.hpp:146-147)
    if (!std::strcmp(value, "component")) c.eigen = MARAGPU_EIG_COMPONENT;
    else if (!std::strcmp(value, "rmhd")) c.eigen = MARAGPU_EIG_RMHD;
    else return fail(std::string("[Mara] no such eigensystem '") + value + "'");
    mark_dirty();
    return 0;
  }
  if (std::strcmp(key, "c2p_model") == 0) {
    if (app()->mlp) {
      maragpu_mlp_destroy(app()->mlp);
      app()->mlp = nullptr;
    }
    app()->mlp_path = value;
    mark_dirty();
    return 0;
  }
  if (std::strcmp(key, "c2p") == 0) {
    if (!std::strcmp(value, "plain")) c.c2p = MARAGPU_C2P_PLAIN;
    else if (!std::strcmp(value, "learned")) c.c2p = MARAGPU_C2P_LEARNED;
    else return fail(std::string("[Mara] no such c2p '") + value + "'");
    mark_dirty();
    return 0;
  }
  return 0;
}

int mara_config_solver_num(const char *key, double value)
{
  if (int rc = need_app()) return rc;
  if (!key) return fail("[Mara] config_solver: null key");
  maragpu_config &c = app()->cfg;
  if (!std::strcmp(key, "theta")) {
    c.theta = value;
    // luaC_config_solver also calls godunov->SetPlmTheta (Mara/mara.cpp:1068-1075), which
    // only the CTU operator implements: its limiter follows only while it is the operator
    if (app()->have_godunov && c.godunov == MARAGPU_PLM_MUSCL) g_ctu_theta = value;
  }
  else if (!std::strcmp(key, "sz10A")) c.sz10A = value;
  else if (!std::strcmp(key, "pfloor")) c.pfloor = value;
  else if (!std::strcmp(key, "preset")) c.preset = value;
  else if (!std::strcmp(key, "dreset")) c.dreset = value;
  else return 0;
  mark_dirty();
  return 0;
}

int mara_config_solver_bool(const char *key, int value)
{
  if (int rc = need_app()) return rc;
  if (!key) return fail("[Mara] config_solver: null key");
  if (!std::strcmp(key, "ereset")) {
    app()->cfg.ereset = value ? 1 : 0;
    mark_dirty();
  }
  return 0;
}

int mara_set_domain(const double *x0, const double *x1, const int *N, int Nd, int Nq, int Ng)
{
  if (int rc = need_app()) return rc;
  if (!x0 || !x1 || !N) return fail("[Mara] set_domain: null argument");
  if (Nd != 3) return fail("[Mara] this library runs 3D domains only");
  if (Nq != 8) return fail("[Mara] rmhd needs Nq = 8");
  maragpu_config &c = app()->cfg;
  for (int d = 0; d < 3; ++d) {
    if (N[d] < 1) return fail("[Mara] set_domain: N must be positive");
    c.N[d] = N[d];
    c.x0[d] = x0[d];
    c.x1[d] = x1[d];
  }
  c.Ng = Ng;
  app()->have_domain = true;
  app()->prim.clear();                 // the set_primitive array no longer fits
  app()->prim_on_device = false;
  mark_dirty();
  return 0;
}

int mara_advance(double *P, size_t n, double dt, double *kzps, int *nerr)
{
  if (int rc = need_app()) return rc;
  if (int rc = ensure_ctx()) return rc;
  App *a = app();
  if (!P) return fail("[Mara] advance: null buffer");
  if (n != zone_doubles(a)) return fail("[Mara] primitive has the wrong size for the domain", MARAGPU_E_SIZE);
  const auto t0 = std::chrono::steady_clock::now();
  int rc = maragpu_set_cooling(a->ctx, a->cool_kind, a->cool_ref, a->cool_t0, a->units);
  if (rc) return fail(std::string("[Mara] cooling: ") + maragpu_strerror(rc), rc);
  rc = maragpu_upload_prim(a->ctx, P, n);
  a->prim_on_device = false;
  if (rc) return fail(std::string("[Mara] upload: ") + maragpu_strerror(rc), rc);
  maragpu_stats st;
  const int errors = maragpu_advance(a->ctx, dt, &st);
  if (errors < 0) return fail(std::string("[Mara] advance: ") + maragpu_strerror(errors), errors);
  a->last_max_lambda = st.max_lambda;
  if (errors == 0) {
    rc = maragpu_download_prim(a->ctx, P, n);
    if (rc) return fail(std::string("[Mara] download: ") + maragpu_strerror(rc), rc);
  }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (kzps) *kzps = 1e-3 * (double)(n / 8) / sec;
  if (nerr) *nerr = errors;
  return 0;
}

double mara_get_timestep(double cfl)
{
  // CFL * min_dx / RiemannSolver::MaxLambda of the last step (inf before any solve)
  if (!g_app || !g_app->have_domain) return 0.0 / 0.0;
  double min_dx = g_app->cfg.x1[0] - g_app->cfg.x0[0];
  for (int d = 0; d < 3; ++d) {
    const double dx = (g_app->cfg.x1[d] - g_app->cfg.x0[d]) / g_app->cfg.N[d];
    if (dx < min_dx) min_dx = dx;
  }
  return cfl * min_dx / g_app->last_max_lambda;
}

int mara_set_primitive(const double *P, size_t n)
{
  if (int rc = need_app()) return rc;
  App *a = app();
  if (!a->have_domain) return fail("[Mara] need a domain to run this, use set_domain");
  if (!P) return fail("[Mara] set_primitive: null buffer");
  if (n != zone_doubles(a)) return fail("[Mara] primitive has the wrong size for the domain", MARAGPU_E_SIZE);
  a->prim.assign(P, P + n);
  a->prim_on_device = false;
  return 0;
}

namespace {
/* The context's committed state holds the set_primitive array (uploaded when it does not) */
int measure_ready()
{
  if (int rc = need_app()) return rc;
  App *a = app();
  if (int rc = ensure_ctx(false)) return rc;
  if (a->prim.size() != zone_doubles(a)) return fail("[Mara] no primitive state to measure, use set_primitive");
  if (!a->prim_on_device) {
    const int rc = maragpu_upload_prim(a->ctx, a->prim.data(), a->prim.size());
    if (rc) return fail(std::string("[Mara] upload: ") + maragpu_strerror(rc), rc);
    a->prim_on_device = true;
  }
  return 0;
}
}  // namespace

int mara_measure_mean_max_divB(double *mean, double *max)
{
  if (int rc = measure_ready()) return rc;
  const int rc = maragpu_measure_divB(app()->ctx, mean, max, nullptr);
  return rc ? fail(std::string("[Mara] measure: ") + maragpu_strerror(rc), rc) : 0;
}

int mara_set_plm_theta(double theta)
{
  if (int rc = need_app()) return rc;
  // luaC_set_plm_theta (Mara/mara.cpp:985-993): only the CTU operator takes it
  if (!app()->have_godunov) return fail("[Mara] need a godunov operator for this");
  if (app()->cfg.godunov == MARAGPU_PLM_MUSCL) {
    g_ctu_theta = theta;
    mark_dirty();
  }
  return 0;
}

int mara_set_units(double a, double b, double c)
{
  if (int rc = need_app()) return rc;
  if (!(a > 0.0) || !(b > 0.0) || !(c > 0.0)) return fail("[Mara] set_units: units must be positive");
  app()->units[0] = a;
  app()->units[1] = b;
  app()->units[2] = c;
  return 0;
}

namespace {
int measure_all(maragpu_measures *m)
{
  if (int rc = measure_ready()) return rc;
  const int rc = maragpu_measure(app()->ctx, app()->units, m);
  return rc ? fail(std::string("[Mara] measure: ") + maragpu_strerror(rc), rc) : 0;
}
}  // namespace

int mara_measure_mean_velocity(double out[4])
{
  maragpu_measures m;
  if (int rc = measure_all(&m)) return rc;
  for (int q = 0; q < 4; ++q) out[q] = m.mean_velocity[q];
  return 0;
}

int mara_measure_mean_cons(double out[8])
{
  maragpu_measures m;
  if (int rc = measure_all(&m)) return rc;
  for (int q = 0; q < 8; ++q) out[q] = m.mean_cons[q];
  return 0;
}

int mara_measure_mean_prim(double out[8])
{
  maragpu_measures m;
  if (int rc = measure_all(&m)) return rc;
  for (int q = 0; q < 8; ++q) out[q] = m.mean_prim[q];
  return 0;
}

int mara_measure_mean_energies(double out[4])
{
  maragpu_measures m;
  if (int rc = measure_all(&m)) return rc;
  for (int q = 0; q < 4; ++q) out[q] = m.mean_energies[q];
  return 0;
}

static int measure_pair(double (maragpu_measures::*field)[2], double *a, double *b)
{
  maragpu_measures m;
  if (int rc = measure_all(&m)) return rc;
  if (a) *a = (m.*field)[0];
  if (b) *b = (m.*field)[1];
  return 0;
}

int mara_measure_mean_max_temperature(double *mean, double *max)
{
  return measure_pair(&maragpu_measures::mean_max_temperature, mean, max);
}

int mara_measure_mean_max_sonic_mach(double *mean, double *max)
{
  return measure_pair(&maragpu_measures::mean_max_sonic_mach, mean, max);
}

int mara_measure_mean_min_alfvenic_mach(double *mean, double *min)
{
  return measure_pair(&maragpu_measures::mean_min_alfvenic_mach, mean, min);
}

int mara_measure_mean_max_magnetic_field(double *mean, double *max)
{
  return measure_pair(&maragpu_measures::mean_max_magnetic_field, mean, max);
}

int mara_measure_max_lorentz_factor(double *max)
{
  maragpu_measures m;
  if (int rc = measure_all(&m)) return rc;
  if (max) *max = m.max_lorentz_factor;
  return 0;
}

int mara_new_ou_field(int Nd, double P0, double zeta, int k1, int seed, void *buf, size_t cap, size_t *len)
{
  if (int rc = need_app()) return rc;
  if (Nd != 3) return fail("[Mara] new_ou_field: this library drives 3-d domains only");
  maragpu_ou_field *f = nullptr;
  int rc = maragpu_ou_field_create(P0, zeta, k1, seed, &f);
  if (rc) return fail("[Mara] new_ou_field: invalid parameters", rc);
  const size_t n = maragpu_ou_field_bytes(f);
  if (len) *len = n;
  if (buf && cap >= n) rc = maragpu_ou_field_save(f, buf, cap);
  maragpu_ou_field_destroy(f);
  return rc ? fail("[Mara] new_ou_field: " + std::string(maragpu_strerror(rc)), rc) : 0;
}

int mara_set_driving(const void *bytes, size_t n)
{
  if (int rc = need_app()) return rc;
  App *a = app();
  if (!a->have_domain) return fail("[Mara] need a domain to run this, use set_domain");
  maragpu_ou_field *f = nullptr;
  int rc = maragpu_ou_field_load(bytes, n, &f);
  if (rc) return fail("[Mara] set_driving: not a serialized 3-d driving field", rc);
  maragpu_ou_field_destroy(a->field);
  a->field = f;
  a->sampled.assign(static_cast<const unsigned char *>(bytes), static_cast<const unsigned char *>(bytes) + n);
  if (a->ctx && (rc = maragpu_set_driving(a->ctx, f))) return fail(std::string("[Mara] driving: ") + maragpu_strerror(rc), rc);
  return 0;
}

int mara_driving_advance(double dt)
{
  if (int rc = need_app()) return rc;
  if (app()->field) maragpu_ou_field_advance(app()->field, dt);
  return 0;
}

int mara_driving_resample(void)
{
  if (int rc = need_app()) return rc;
  App *a = app();
  if (!a->field) return 0;
  a->sampled.resize(maragpu_ou_field_bytes(a->field));
  maragpu_ou_field_save(a->field, a->sampled.data(), a->sampled.size());
  if (a->ctx) {
    const int rc = maragpu_set_driving(a->ctx, a->field);
    if (rc) return fail(std::string("[Mara] driving: ") + maragpu_strerror(rc), rc);
  }
  return 0;
}

int mara_driving_serialize(void *buf, size_t cap, size_t *len)
{
  if (int rc = need_app()) return rc;
  App *a = app();
  const size_t n = a->field ? maragpu_ou_field_bytes(a->field) : 0;
  if (len) *len = n;
  if (a->field && buf && cap >= n) maragpu_ou_field_save(a->field, buf, cap);
  return 0;
}

int mara_driving_power(double *power)
{
  if (int rc = need_app()) return rc;
  App *a = app();
  if (!power) return fail("[Mara] driving power: null output");
  *power = 0.0;
  if (!a->field) return 0;
  if (int rc = ensure_ctx(false)) return rc;
  const int rc = maragpu_driving_power(a->ctx, power);
  return rc ? fail(std::string("[Mara] driving power: ") + maragpu_strerror(rc), rc) : 0;
}

int mara_set_cooling(const char *key, double ref, double t0)
{
  if (int rc = need_app()) return rc;
  if (!key) return fail("[Mara] set_cooling: null key");
  App *a = app();
  // an unknown key leaves the cooling as it was (Mara/mara.cpp:1198-1217)
  const int kind = !std::strcmp(key, "T4") ? 1 : !std::strcmp(key, "E4") ? 2 : 0;
  if (!kind) return 0;
  if (!(ref > 0.0) || !(t0 > 0.0)) return fail("[Mara] set_cooling: reference value and time must be positive");
  a->cool_kind = kind;
  a->cool_ref = ref;
  a->cool_t0 = t0;
  return 0;
}

int mara_cooling_rate(double dt, double *rate)
{
  if (int rc = need_app()) return rc;
  App *a = app();
  if (!rate) return fail("[Mara] cooling_rate: null output");
  double e = a->removed_pending;
  a->removed_pending = 0.0;
  if (a->ctx) {
    double d = 0.0;
    const int rc = maragpu_cooling_energy_removed(a->ctx, &d);
    if (rc) return fail(std::string("[Mara] cooling_rate: ") + maragpu_strerror(rc), rc);
    e += d;
  }
  *rate = a->cool_kind ? e / dt : 0.0;
  return 0;
}

struct maragpu_ctx *mara_device_context(void)
{
  if (!g_app) return nullptr;
  if (ensure_ctx()) return nullptr;
  return g_app->ctx;
}

}  // extern "C"
