/* maragpu — MI355X (gfx950) relativistic-MHD update, C ABI.
 *
 * One context owns one device-resident grid and two HIP streams (the second runs the HLLD
 * sweeps' redo kernels beside the next sweep; maragpu_advance returns with both drained).
 * Nothing in this header uses HIP or torch types; all pointers are host pointers unless
 * stated otherwise.
 * Every entry point returns instead of throwing. Return codes: 0 success, > 0 a count
 * (failed zones, for maragpu_advance), < 0 an error (maragpu_strerror names it).
 *
 * Reference interfaces replaced (jzrake/ctf, Mara):
 *   maragpu_create / maragpu_destroy   MaraApplication set-up: Mara.set_domain, set_fluid('rmhd'),
 *                                      set_eos('gamma-law', G), set_godunov('plm-split'),
 *                                      set_riemann, set_advance, config_solver
 *                                      (Mara/mara.cpp:726-755, 813-867, 934-1167)
 *   maragpu_upload_prim / download     the memcpy in/out of luaC_Mara_advance (Mara/mara.cpp:317,333),
 *                                      host layout = Mara's AoS [i][j][k][q] (Mara/valman.hpp:25-36)
 *   maragpu_advance                    luaC_Mara_advance body (Mara/mara.cpp:297-341):
 *                                      GodunovOperator::PrimToCons, RungeKuttaIntegration::AdvanceState,
 *                                      ConsToPrim; state is not committed when zones fail. With
 *                                      godunov = MARAGPU_WENO_SPLIT the operator is WenoSplit with
 *                                      local Lax-Friedrichs splitting (set_godunov('weno-split'),
 *                                      fsplit = 'llf'; Mara/weno-split.cpp:13-245)
 *   maragpu_get_timestep               luaC_Mara_get_timestep (Mara/mara.cpp:374-381)
 *   maragpu_measure_divB               luaC_mean_max_divB (Mara/mara.cpp:2264-2324)
 *   maragpu_measure                    the other Mara.measure functions (Mara/mara.cpp:1954-2262)
 *   maragpu_ou_field_*, maragpu_set_driving, maragpu_set_cooling
 *                                      StochasticVectorField3d, DrivingProcedure, CoolingModuleT4/E4
 *   maragpu_batch_prim_to_cons         AdiabaticIdealRmhd::PrimToCons (Mara/rmhd.cpp:133-163)
 *   maragpu_batch_cons_to_prim         AdiabaticIdealRmhd::ConsToPrim cascade (Mara/rmhd.cpp:59-132)
 *                                      over the rmhd-c2p.h C API (Mara/rmhd-c2p.h:22-35)
 *   maragpu_batch_flux_eigen           AdiabaticIdealRmhd::FluxAndEigenvalues (Mara/rmhd.cpp:164-265)
 *   maragpu_batch_reconstruct          reconstruct(v, op) of weno.h (Mara/weno.h:17-20, weno.c:58-71)
 *   maragpu_batch_riemann              RiemannSolver::IntercellFlux (Mara/hydro.hpp:217-218) of
 *                                      HllRiemannSolver, HllcRmhdRiemannSolver, HlldRmhdRiemannSolver
 *   maragpu_batch_eigensystem          FluidEquations::Eigensystem (Mara/hydro.hpp:146-147), called
 *                                      by WenoSplit (Mara/weno-split.cpp:181); the reference's RMHD
 *                                      fluid leaves it empty, this is the RMHD characteristic one
 */
#ifndef MARAGPU_H
#define MARAGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MARAGPU_PCM = 0, MARAGPU_PLM = 1, MARAGPU_WENO5 = 2 };            /* config_solver extrap */
enum { MARAGPU_JS96 = 0, MARAGPU_B08 = 1, MARAGPU_SZ10 = 2 };            /* config_solver IS */
enum { MARAGPU_HLL = 0, MARAGPU_HLLC = 1, MARAGPU_HLLD = 2 };            /* set_riemann */
enum { MARAGPU_FP64 = 0, MARAGPU_MIXED = 1 };                            /* precision */
enum { MARAGPU_C2P_PLAIN = 0, MARAGPU_C2P_LEARNED = 1 };                 /* c2p initial guess */
enum { MARAGPU_PLM_SPLIT = 0, MARAGPU_WENO_SPLIT = 1, MARAGPU_PLM_MUSCL = 2 };  /* set_godunov */
enum { MARAGPU_EIG_COMPONENT = 0, MARAGPU_EIG_RMHD = 1 };                /* weno-split projection */

enum {
  MARAGPU_OK = 0,
  MARAGPU_E_INVALID = -1,   /* bad argument or configuration */
  MARAGPU_E_HIP = -2,       /* HIP runtime error */
  MARAGPU_E_NOMEM = -3,     /* device allocation failed */
  MARAGPU_E_SIZE = -4,      /* buffer size does not match the domain */
  MARAGPU_E_NOMODEL = -5,   /* learned c2p requested without loaded weights */
  MARAGPU_E_IO = -6         /* a checkpoint file could not be read or written */
};

typedef struct maragpu_config {
  int N[3];          /* interior zones per axis */
  int Ng;            /* ghost zones per side (3 for WENO5) */
  double x0[3], x1[3];
  double gamma;      /* gamma-law index */
  int recon;         /* MARAGPU_PCM / PLM / WENO5 */
  double theta;      /* PLM theta (2 = MC) */
  int IS;            /* WENO smoothness indicator */
  double sz10A;
  int riemann;       /* MARAGPU_HLL / HLLC / HLLD */
  int rk;            /* 1 single, 2 rk2, 3 rk3, 4 rk4 */
  int precision;     /* MARAGPU_FP64 or MARAGPU_MIXED */
  int c2p;           /* MARAGPU_C2P_PLAIN or MARAGPU_C2P_LEARNED */
  double pfloor;     /* pressure floor, disabled if <= 0 */
  int ereset;        /* emergency reset of failed zones */
  double preset, dreset;
  int godunov;       /* MARAGPU_PLM_SPLIT, MARAGPU_WENO_SPLIT (local Lax-Friedrichs splitting) or
                        MARAGPU_PLM_MUSCL (CTU-Hancock, FP64 and the plain c2p; theta is then the
                        operator's own limiter parameter, PlmCtuHancockOperator::SetPlmTheta) */
  int eigen;         /* weno-split projection: MARAGPU_EIG_COMPONENT or MARAGPU_EIG_RMHD */
} maragpu_config;

typedef struct maragpu_stats {
  double max_lambda;            /* RiemannSolver::MaxLambda over the step */
  int64_t hlld_success;         /* HLLD solves that produced the flux */
  int64_t hlld_failure;         /* HLLD solves that fell back to HLLC */
  int64_t riemann_retries;      /* interfaces retried with the piecewise-constant states */
  int64_t riemann_failures;     /* interfaces where both attempts failed (flux set to 0) */
  int64_t c2p_zones;            /* per-zone c2p solves (interior zones x (stages + 1)) */
  int64_t c2p_iterations;       /* Newton iterations of those solves */
  int64_t c2p_fallbacks;        /* per-zone solves not settled by their first attempt */
  int64_t c2p_failed;           /* failed zones of the failing stage, ghost images included */
  int64_t hlld_c2p_iterations;  /* Newton iterations of the c2p nested in HLLD */
  int64_t hlld_newton_iterations; /* p* Newton iterations in HLLD */
  int64_t mlp_guesses;          /* learned c2p: zones started from the network's guess */
  int64_t mlp_fallbacks;        /* learned c2p: zones that fell back to the plain cascade */
  int64_t fp64_redo;            /* mixed form: interfaces whose FP32 HLLD failed, redone in FP64 */
  int64_t fp64_redo_overflow;   /* ... beyond the redo queue (FP32 HLLC flux kept) */
  int64_t eig_identity;         /* weno-split, eigen = RMHD: interfaces projected with the
                                   identity because the eigensystem was rejected */
  double device_ms;             /* device time of the step */
} maragpu_stats;

typedef struct maragpu_ctx maragpu_ctx;

/* MARAGPU_E_INVALID for a malformed configuration, including an axis with fewer interior
 * zones than ghosts (every ghost zone is the periodic image of an interior one) */
int maragpu_create(const maragpu_config *cfg, maragpu_ctx **out);
int maragpu_destroy(maragpu_ctx *ctx);
/* the configuration the context was created with */
int maragpu_get_config(const maragpu_ctx *ctx, maragpu_config *cfg);
/* n = number of doubles = zones (ghosts included) x 8 */
int maragpu_upload_prim(maragpu_ctx *ctx, const double *P_aos, size_t n);
int maragpu_download_prim(maragpu_ctx *ctx, double *P_aos, size_t n);
/* One primitive field q (0..7: rho, pre, vx, vy, vz, Bx, By, Bz) of the committed state,
 * interior zones only, (i, j, k) order with k fastest; n = N0 N1 N2. This is the block that
 * DataManagerHDF5 writes to and reads from prim/<field> (modules/unigrid.lua:85-212,243-292).
 * The upload also writes the field's periodic ghost images. Uploading all eight fields
 * gives a committed state, as maragpu_upload_prim does. */
int maragpu_download_field(maragpu_ctx *ctx, int q, double *out, size_t n);
int maragpu_upload_field(maragpu_ctx *ctx, int q, const double *in, size_t n);
int maragpu_advance(maragpu_ctx *ctx, double dt, maragpu_stats *stats);
/* CFL * min dx / MaxLambda of the last step: +inf before any solve, as in Mara */
double maragpu_get_timestep(maragpu_ctx *ctx, double cfl);
/* mean and signed max in the reference's convention; absmax = max |corner div B| / dx */
int maragpu_measure_divB(maragpu_ctx *ctx, double *mean, double *max, double *absmax);
/* The Mara.measure table over the interior zones of the committed state (Mara/mara.cpp:1954-2262):
 * volume averages (sum / N^3) and extremes. units = {Length, Mass, Time} of PhysicalUnits
 * (Mara/units.cpp:3-6), NULL for Mara's default cgs (1, 1, 1). */
typedef struct maragpu_measures {
  double mean_velocity[4];           /* <W>, <|vx|>, <|vy|>, <|vz|>   luaC_mean_velocity (:1954-1980) */
  double mean_cons[8];               /* <U>                           luaC_mean_cons (:1982-2013) */
  double mean_prim[8];               /* <P>                           luaC_mean_prim (:2015-2041) */
  double mean_energies[4];           /* magnetic, kinetic, internal, total  luaC_mean_energies (:2043-2103) */
  double mean_max_temperature[2];    /* MeV                           luaC_mean_max_temperature (:2105-2126) */
  double mean_max_sonic_mach[2];     /*                               luaC_mean_max_sonic_mach (:2128-2165) */
  double mean_min_alfvenic_mach[2];  /* mean, min                     luaC_mean_min_alfvenic_mach (:2167-2201) */
  double mean_max_magnetic_field[2]; /* Gauss                         luaC_mean_max_magnetic_field (:2203-2230) */
  double max_lorentz_factor;         /*                               luaC_max_lorentz_factor (:2232-2258) */
} maragpu_measures;
int maragpu_measure(maragpu_ctx *ctx, const double *units, maragpu_measures *out);
/* maragpu_measure without any device allocation. No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). The per-block partials live in a fixed 270 KB device
 * array of the library, one per GPU, set up when the first context on that GPU is
 * created, so any grid size is accepted (the reduction runs over at most 1024 blocks).
 * All calls on one GPU share that array: calls that overlap in time on one GPU (from
 * different host threads or streams) race. Same result as maragpu_measure, bit for bit. */
int maragpu_measure_noalloc(maragpu_ctx *ctx, const double *units, maragpu_measures *out);
/* ---------------------------------------------------------------------------------
 * Driving and cooling: source steps applied to the committed state after every
 * successful advance, as luaC_Mara_advance does (Mara/mara.cpp:330-333).
 * ------------------------------------------------------------------------------- */
/* StochasticVectorField3d (Mara/ou-field.cpp:44-74): Fourier modes |k_i| <= k1 driven by
 * Ornstein-Uhlenbeck processes, with the spectrum of luaC_new_ou_field (Mara/mara.cpp:1373-1394).
 * Held on the host; bytes are the reference's Serialize layout (:261-278), so
 * Mara.new_ou_field output and Mara checkpoints load unchanged. */
typedef struct maragpu_ou_field maragpu_ou_field;
int maragpu_ou_field_create(double P0, double zeta, int k1, int seed, maragpu_ou_field **out);
int maragpu_ou_field_load(const void *bytes, size_t n, maragpu_ou_field **out);
size_t maragpu_ou_field_bytes(const maragpu_ou_field *field);
int maragpu_ou_field_save(const maragpu_ou_field *field, void *bytes, size_t n);
int maragpu_ou_field_advance(maragpu_ou_field *field, double dt);          /* AdvanceField (:186-210) */
int maragpu_ou_field_destroy(maragpu_ou_field *field);
/* SampleField (:226-239) at n points xyz (3 doubles each) on the device; F: 3 per point */
int maragpu_ou_field_sample(const maragpu_ou_field *field, size_t n, const double *xyz, double *F);
/* DrivingProcedure (Mara/driving.cpp): samples the field at every zone of the context
 * (ResampleField, :44-85; k1 <= 6) and drives the committed state after each successful
 * advance (Drive_rmhd, :154-186). NULL turns driving off. Call again to resample. */
int maragpu_set_driving(maragpu_ctx *ctx, const maragpu_ou_field *field);
int maragpu_driving_power(maragpu_ctx *ctx, double *power);               /* AveragePowerInFields */
/* CoolingModuleT4 (kind 1) / CoolingModuleE4 (kind 2) (Mara/cooling.cpp), 0 = none; units =
 * {Length, Mass, Time} (NULL: cgs) */
int maragpu_set_cooling(maragpu_ctx *ctx, int kind, double ref, double t0, const double *units);
/* EnergyRemoved: the energy taken out since the last call (then reset) */
int maragpu_cooling_energy_removed(maragpu_ctx *ctx, double *removed);

/* device time of the dominant kernel over the last advance (ms, per launch) and launches */
int maragpu_kernel_times(maragpu_ctx *ctx, double *flux_ms, int *flux_launches, double *c2p_ms,
                         int *c2p_launches, double *update_ms, int *update_launches);
const char *maragpu_strerror(int code);

/* Point functions over n states on the device (host arrays in and out). */
int maragpu_batch_prim_to_cons(size_t n, const double *P, double *U, double gamma, int *err);
/* P holds the previous primitive states (the cascade's second guess) on entry */
int maragpu_batch_cons_to_prim(size_t n, const double *U, double *P, double gamma, double pfloor,
                               int *err, int *iters);
int maragpu_batch_flux_eigen(size_t n, const double *P, double *F, double *ap, double *am, int dim,
                             double gamma);
/* V holds n stencils of 6 values v[-2..3]; op as weno.h ReconstructOperation (PLM_C2L=0,
 * PLM_C2R=1, WENO5_FD_C2R=2, WENO5_FD_C2L=3, WENO5_FV_C2R=4, WENO5_FV_C2L=5) applied at v[0] */
int maragpu_batch_reconstruct(size_t n, const double *V, double *out, int op, double theta, int IS,
                              double sz10A);
int maragpu_batch_riemann(size_t n, int solver, int precision, const double *PL, const double *PR,
                          double *F, int dim, double gamma, double pfloor, int *err, double *lam);
/* RMHD characteristic eigensystem of n primitive states for the flux along dim: L, R
 * (64 doubles each per state, row-major, L R = I), ascending eigenvalues lam (8 per state),
 * ok = 1 where the eigensystem is usable (else L = R = I and lam = 0, the projection the
 * weno-split sweep then uses) */
int maragpu_batch_eigensystem(size_t n, const double *P, double *L, double *R, double *lam, int *ok,
                              int dim, double gamma);

/* ---------------------------------------------------------------------------------
 * Learned c2p initial guess (no counterpart in the reference: the reference starts its
 * Newton solves from the cons-based estimate, Mara/rmhd-c2p.c:97-101, and from the
 * previous primitive state). A 6 -> 64 -> 64 -> 64 -> 2 network (hard-swish
 * activations), BF16 weights and FP32 accumulation on the matrix cores, maps scale-free
 * features of U to log W and log(p/D); the FP64 anton2dzw Newton solve polishes every
 * zone from that guess and the plain cascade is the fallback.
 * ------------------------------------------------------------------------------- */
typedef struct maragpu_mlp maragpu_mlp;

/* Training-pair sampling ranges (log-uniform): Lorentz factor 1 .. W_max,
 * magnetisation sigma = b^2/(rho h), p/rho; rho log-uniform on [0.1, 10], isotropic
 * velocity and field directions (random field-velocity angle). */
typedef struct maragpu_mlp_sampling {
  double W_max;
  double sigma_min, sigma_max;
  double ptr_min, ptr_max;
} maragpu_mlp_sampling;

typedef struct maragpu_mlp_train_config {
  uint64_t seed;           /* minibatch draws */
  int epochs;              /* passes of (pairs / batch) steps */
  int batch;               /* multiple of 256 */
  double lr, lr_final;     /* Adam, cosine decay */
} maragpu_mlp_train_config;

typedef struct maragpu_mlp_train_report {
  double loss_first, loss_last;  /* mean squared error of standardised targets */
  int64_t steps, samples;
  double seconds;
} maragpu_mlp_train_report;

int maragpu_mlp_create(maragpu_mlp **out, uint64_t seed);
int maragpu_mlp_destroy(maragpu_mlp *net);
size_t maragpu_mlp_param_count(void);
size_t maragpu_mlp_serialized_size(void);
int maragpu_mlp_save(const maragpu_mlp *net, void *buf, size_t nbytes);
int maragpu_mlp_load(maragpu_mlp **out, const void *buf, size_t nbytes);
/* theta: maragpu_mlp_param_count() floats; norms: 16 floats (x mean[6], x std[6],
 * y mean[2], y std[2]) */
int maragpu_mlp_get_params(const maragpu_mlp *net, float *theta, float *norms);
int maragpu_mlp_set_params(maragpu_mlp *net, const float *theta, const float *norms);
/* raw network outputs for standardised inputs (n x 6 -> n x 2) */
int maragpu_mlp_forward(const maragpu_mlp *net, size_t n, const float *Xs, float *Y);
/* gradient of the mean squared error over n (multiple of 256) standardised samples */
int maragpu_mlp_grad(maragpu_mlp *net, size_t n, const float *Xs, const float *Ts, float *grad,
                     float *loss);
/* Training pairs (conserved -> primitive) held on the device. */
typedef struct maragpu_mlp_data maragpu_mlp_data;
int maragpu_mlp_data_create(maragpu_mlp_data **out, int64_t capacity);
int maragpu_mlp_data_destroy(maragpu_mlp_data *data);
int64_t maragpu_mlp_data_size(const maragpu_mlp_data *data);
/* appends up to n pairs sampled over the ranges (states that fail prim_to_cons are dropped) */
int maragpu_mlp_data_add_synthetic(maragpu_mlp_data *data, int64_t n, uint64_t seed, double gamma,
                                   const maragpu_mlp_sampling *ranges);
/* appends up to n pairs drawn (with replacement) from the interior zones of ctx's
 * committed primitive state: a snapshot of the run */
int maragpu_mlp_data_add_state(maragpu_mlp_data *data, maragpu_ctx *ctx, int64_t n, uint64_t seed,
                               double gamma);
/* sets the standardisation from the data and trains (Adam) on the device */
int maragpu_mlp_train(maragpu_mlp *net, const maragpu_mlp_data *data, const maragpu_mlp_train_config *cfg,
                      maragpu_mlp_train_report *report);
/* the network a context uses for MARAGPU_C2P_LEARNED (not owned; NULL detaches) */
int maragpu_attach_mlp(maragpu_ctx *ctx, const maragpu_mlp *net);
/* learned-guess c2p on n states (host arrays): P holds the previous primitive states
 * (the fallback cascade's second guess) on entry and receives the primitives; iters the
 * Newton iterations, err the error code, used_guess 1 where the solve from the network's
 * guess converged (0 where the plain cascade took over); start (2n doubles, may be NULL)
 * the Newton start (Z, W) built from the network's outputs */
int maragpu_batch_cons_to_prim_learned(const maragpu_mlp *net, size_t n, const double *U, double *P,
                                       double gamma, double pfloor, int *err, int *iters,
                                       int *used_guess, double *start);

#ifdef __cplusplus
}
#endif

#endif /* MARAGPU_H */
