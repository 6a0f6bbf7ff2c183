/* mara_host — Mara's Lua-facing operator API over the maragpu device path, C ABI.
 *
 * The verbs keep the names, argument meaning and error behaviour of the Lua table that
 * luaopen_Mara registers (Mara/mara.cpp:137-268); where the Lua binding raises
 * luaL_error these return a negative code and mara_last_error() holds the message.
 * One application per process, as in Mara (the `Mara` global, Mara/mara.cpp).
 *
 *   mara_start / mara_close          luaC_Mara_start / close         (Mara/mara.cpp:270-295)
 *   mara_set_fluid                   luaC_set_fluid                  (:813-837)   'rmhd'
 *   mara_set_eos                     luaC_set_eos                    (:840-867)   'gamma-law', Gamma
 *   mara_set_boundary                luaC_set_boundary               (:887-932)   'periodic'
 *   mara_set_riemann                 luaC_set_riemann                (:934-958)   'hll' 'hllc' 'hlld'
 *   mara_set_godunov                 luaC_set_godunov                (:960-984)   'plm-split'
 *                                                                                 'weno-split'
 *                                                                                 'plm-muscl'
 *   mara_set_plm_theta               luaC_set_plm_theta              (:985-993)
 *   mara_set_advance                 luaC_set_advance                (:1140-1167) 'single' 'rk2' 'rk3' 'rk4'
 *   mara_config_solver_{str,num,bool} one key of luaC_config_solver  (:996-1138); extra keys
 *                                    'precision' = 'fp64' | 'mixed', 'c2p' = 'plain' | 'learned',
 *                                    'c2p_model' = path of the network file (default:
 *                                    models/c2p_mlp.bin next to the library) and
 *                                    'eigensystem' = 'component' | 'rmhd' (weno-split projection)
 *   mara_set_domain                  luaC_set_domain                 (:726-755)
 *   mara_advance                     luaC_Mara_advance               (:297-341) -> kz/s, errors
 *   mara_get_timestep                luaC_Mara_get_timestep          (:374-381)
 *   mara_set_primitive               luaC_set_primitive              (:1284-1295)
 *   mara_set_units                   luaC_set_units                  (:869-883)
 *   mara_measure_mean_max_divB       luaC_mean_max_divB              (:2264-2324)
 *   mara_measure_mean_velocity, _mean_cons, _mean_prim, _mean_energies, _mean_max_temperature,
 *   _mean_max_sonic_mach, _mean_min_alfvenic_mach, _mean_max_magnetic_field, _max_lorentz_factor
 *                                    the Mara.measure table          (:1954-2262)
 *   mara_new_ou_field                luaC_new_ou_field               (:1378-1400) -> bytes
 *   mara_set_driving                 luaC_set_driving                (:1169-1196)
 *   mara_driving_{advance,resample,serialize,power}  Mara.driving.*  (:1771-1808)
 *   mara_set_cooling                 luaC_set_cooling                (:1198-1217) 'T4' 'E4'
 *   mara_cooling_rate                luaC_cooling_rate               (:1297-1309)
 *   mara_init_turbulence             (synthetic data; the role of an init_prim callback, :531-632)
 */
#ifndef MARA_HOST_H
#define MARA_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int mara_start(void);
int mara_close(void);
int mara_set_fluid(const char *key);
int mara_set_eos(const char *key, double gamma);
int mara_set_boundary(const char *key);
int mara_set_riemann(const char *key);
int mara_set_godunov(const char *key);
int mara_set_advance(const char *key);
int mara_config_solver_str(const char *key, const char *value);
int mara_config_solver_num(const char *key, double value);
int mara_config_solver_bool(const char *key, int value);
int mara_set_domain(const double *x0, const double *x1, const int *N, int Nd, int Nq, int Ng);
/* P: AoS primitives, n doubles; on return kzps = 1e-3 zones (ghosts included) per wall
 * second and nerr = failed zones (P untouched when nerr > 0) */
int mara_advance(double *P, size_t n, double dt, double *kzps, int *nerr);
double mara_get_timestep(double cfl);
int mara_set_primitive(const double *P, size_t n);
int mara_measure_mean_max_divB(double *mean, double *max);
/* Mara.set_units(a, b, c) builds PhysicalUnits(a, b, c), whose parameters are (Length, Mass,
 * Time) although the Lua arguments are read as (Mass, Length, Time) (Mara/mara.cpp:871-877,
 * Mara/units.cpp:3-6): a sets the length unit, b the mass unit. Kept as is. */
int mara_set_units(double a, double b, double c);
/* the CTU operator's limiter parameter; like config_solver's theta it reaches the operator
 * only while plm-muscl is selected (the other operators ignore SetPlmTheta) */
int mara_set_plm_theta(double theta);
/* The measures, mean_max_divB included, read the array of the last set_primitive
 * (Mara->PrimitiveArray, which advance does not change), uploaded to the device on demand;
 * they need a domain and the fluid, not the rest of the solver. */
int mara_measure_mean_velocity(double out[4]);       /* <W>, <|vx|>, <|vy|>, <|vz|> */
int mara_measure_mean_cons(double out[8]);
int mara_measure_mean_prim(double out[8]);
int mara_measure_mean_energies(double out[4]);       /* magnetic, kinetic, internal, total */
int mara_measure_mean_max_temperature(double *mean, double *max);
int mara_measure_mean_max_sonic_mach(double *mean, double *max);
int mara_measure_mean_min_alfvenic_mach(double *mean, double *min);
int mara_measure_mean_max_magnetic_field(double *mean, double *max);
int mara_measure_max_lorentz_factor(double *max);
/* Driving and cooling. A serialized field is the reference's byte string: new_ou_field and
 * driving_serialize write it into buf when cap suffices and always report its length;
 * set_driving samples it over the domain (on the device), driving_advance steps the
 * field's OU processes, driving_resample samples the advanced field. Driving and then
 * cooling act on the state after every successful advance. */
int mara_new_ou_field(int Nd, double P0, double zeta, int k1, int seed, void *buf, size_t cap, size_t *len);
int mara_set_driving(const void *bytes, size_t n);
int mara_driving_advance(double dt);
int mara_driving_resample(void);
int mara_driving_serialize(void *buf, size_t cap, size_t *len);
int mara_driving_power(double *power);
int mara_set_cooling(const char *key, double ref, double t0);
int mara_cooling_rate(double dt, double *rate);   /* energy removed since the last call / dt */
const char *mara_last_error(void);
/* the device context behind the application (created on first use), for callers that
 * want to keep the state on the device between steps */
struct maragpu_ctx *mara_device_context(void);

typedef struct mara_ic_params {
  int N[3];
  int Ng;
  double rho, pre;
  double B0[3];
  double v_rms, v_kmin, v_kmax;
  int four_velocity;       /* v_rms is the rms of u = W v */
  double dB_rms, A_kmin, A_kmax;
  uint64_t seed;
} mara_ic_params;

/* fills an AoS primitive array of n = (N+2Ng)^3 x 8 doubles; 0 on success */
int mara_init_turbulence(const mara_ic_params *ic, double *P, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* MARA_HOST_H */
