/* The stochastic driving field on the host: StochasticVectorField3d (Mara/ou-field.cpp)
 * with its Ornstein-Uhlenbeck processes (Mara/ornuhl.hpp) and random streams
 * (Mara/random.cpp). The field is a few hundred Fourier modes; advancing it is serial
 * and cheap, so it stays on the host, and the device only evaluates it at every zone
 * (maragpu_set_driving). Serialization uses the reference's byte layout, so a field
 * made by Mara.new_ou_field or stored in a Mara checkpoint loads here unchanged.
 *
 * The arithmetic follows the reference step by step (complex products written out as
 * their real and imaginary parts) so that the serialized state after any number of
 * advances is the reference's, bit for bit.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/maragpu.h"
#include "ou_field.hpp"

namespace {

/* RandomNumberStream: Numerical Recipes' combined generator (3rd ed., section 7.1) */
struct Stream {
  uint64_t u, v, w;
  uint64_t draw()
  {
    u = u * 2862933555777941757ULL + 7046029254386353087ULL;
    v ^= v >> 17;
    v ^= v << 31;
    v ^= v >> 8;
    w = 4294957665U * (w & 0xffffffffULL) + (w >> 32);
    uint64_t x = u ^ (u << 21);
    x ^= x >> 35;
    x ^= x << 4;
    return (x + v) ^ w;
  }
  void init(int n)
  {
    v = 4101842887655102017ULL;
    w = 1;
    u = (uint64_t)(int64_t)n ^ v;
    draw();
    v = u;
    draw();
    w = v;
    draw();
  }
  // RandomDouble(a, b): the integer is scaled by a single-precision constant, so the
  // product is rounded to float before it is widened (Mara/random.cpp:31-34)
  double in(double a, double b)
  {
    const float s = (float)draw() * 5.42101086242752217E-20f;
    return s * (b - a) + a;
  }
};

struct Process {      // OrnsteinUhlenbeckProcess: 40 bytes, as serialized
  Stream s;
  double theta, sigma;
};
static_assert(sizeof(Process) == 40, "serialized process layout");

double inverse_erf(double y)
{
  // Newton on erf(x) = y from a two-term series start, to a 1e-12 step (Mara/ornuhl.hpp:55-66)
  double x = 0.5 * std::sqrt(M_PI) * (y + M_PI / 12.0 * std::pow(y, 3));
  double step;
  do {
    const double f = std::erf(x) - y;
    const double g = 2.0 / std::sqrt(M_PI) * std::exp(-x * x);
    step = f / g;
    x -= step;
  } while (std::fabs(step) > 1e-12);
  return x;
}

double gaussian(Process &p, double variance)
{
  const double c = p.s.in(-1, 1);
  return std::sqrt(2 * variance) * inverse_erf(c) + 0.0;
}

}  // namespace

namespace maragpu_host {

void OuField::build(double P0_, double zeta_, int k1_, int seed)
{
  k1 = k1_;
  P0 = P0_;
  zeta = zeta_;
  const int nk = 2 * k1 + 1;
  numk = nk * nk * nk;
  K.assign(3 * numk, 0.0);
  F.assign(6 * numk, 0.0);
  proc.assign((size_t)numk * sizeof(Process), 0);
  // the spectrum of luaC_new_ou_field (Mara/mara.cpp:1373-1377): k^6 exp(-8k / (2 pi k_peak)), k_peak = 2
  auto spectrum = [](double k) { return std::pow(k, 6) * std::exp(-8 * k / (2 * M_PI * 2.0)); };
  double total = 0.0;
  for (int i = -k1; i <= k1; ++i)
    for (int j = -k1; j <= k1; ++j)
      for (int k = -k1; k <= k1; ++k) {
        const int m = ((i + k1) * nk + (j + k1)) * nk + (k + k1);
        K[3 * m + 0] = 2 * M_PI * i;
        K[3 * m + 1] = 2 * M_PI * j;
        K[3 * m + 2] = 2 * M_PI * k;
        const double *q = &K[3 * m];
        total += spectrum(std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]));
      }
  Process *pr = reinterpret_cast<Process *>(proc.data());
  for (int m = 0; m < numk; ++m) {
    const double *q = &K[3 * m];
    const double Pk = spectrum(std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]));
    pr[m].s.init(seed + m);
    pr[m].theta = 1.0;
    pr[m].sigma = std::sqrt(2 * P0 * Pk / total);
  }
}

size_t OuField::bytes() const
{
  return sizeof(int) + 2 * sizeof(double) + (size_t)numk * (6 * 2 * sizeof(double) + sizeof(Process));
}

/* layout: k1, P0, zeta, then the complex arrays Kx, Ky, Kz, Fx, Fy, Fz (numk each, real
 * and imaginary parts interleaved), then the numk processes */
void OuField::save(unsigned char *b) const
{
  auto put = [&](const void *p, size_t n) { std::memcpy(b, p, n); b += n; };
  put(&k1, sizeof(int));
  put(&P0, sizeof(double));
  put(&zeta, sizeof(double));
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < numk; ++m) {
      const double c[2] = {K[3 * m + a], 0.0};
      put(c, sizeof(c));
    }
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < numk; ++m) put(&F[2 * (a * numk + m)], 2 * sizeof(double));
  put(proc.data(), proc.size());
}

bool OuField::load(const unsigned char *b, size_t n)
{
  if (n < sizeof(int) + 2 * sizeof(double)) return false;
  auto get = [&](void *p, size_t len) { std::memcpy(p, b, len); b += len; };
  int k;
  get(&k, sizeof(int));
  if (k < 0 || k > 64) return false;
  k1 = k;
  const int nk = 2 * k1 + 1;
  numk = nk * nk * nk;
  if (n != bytes()) return false;
  get(&P0, sizeof(double));
  get(&zeta, sizeof(double));
  K.assign(3 * numk, 0.0);
  F.assign(6 * numk, 0.0);
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < numk; ++m) {
      double c[2];
      get(c, sizeof(c));
      K[3 * m + a] = c[0];
    }
  for (int a = 0; a < 3; ++a)
    for (int m = 0; m < numk; ++m) get(&F[2 * (a * numk + m)], 2 * sizeof(double));
  proc.resize((size_t)numk * sizeof(Process));
  get(proc.data(), proc.size());
  return true;
}

/* AdvanceField (Mara/ou-field.cpp:186-210): each mode but k = 0 takes an OU increment per
 * component (the reference's complex constructor draws the imaginary part first, as its
 * compiler evaluates the two arguments), projected by (1 - 2 zeta) k k / k^2 + zeta I and
 * normalised */
void OuField::advance(double dt)
{
  Process *pr = reinterpret_cast<Process *>(proc.data());
  for (int m = 0; m < numk; ++m) {
    const double *k = &K[3 * m];
    const double kdotk = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    if (kdotk < 1e-8) continue;
    double dW[3][2];
    for (int a = 0; a < 3; ++a) {
      double *z = &F[2 * (a * numk + m)];
      const double im = gaussian(pr[m], dt);
      const double re = gaussian(pr[m], dt);
      // -theta z dt + sigma dW
      dW[a][0] = (-pr[m].theta * z[0]) * dt + pr[m].sigma * re;
      dW[a][1] = (-pr[m].theta * z[1]) * dt + pr[m].sigma * im;
    }
    const double norm = std::sqrt(1 - 1.25 * zeta + 3 * zeta * zeta);
    double dV[3][2];
    for (int p = 0; p < 3; ++p) {
      dV[p][0] = dV[p][1] = 0.0;
      for (int q = 0; q < 3; ++q) {
        const double Pr = (1 - 2 * zeta) * k[p] * k[q] / kdotk + zeta * (p == q);
        dV[p][0] += (Pr * dW[q][0]) / norm;
        dV[p][1] += (Pr * dW[q][1]) / norm;
      }
    }
    for (int a = 0; a < 3; ++a) {
      F[2 * (a * numk + m)] += dV[a][0];
      F[2 * (a * numk + m) + 1] += dV[a][1];
    }
  }
}

}  // namespace maragpu_host

struct maragpu_ou_field {
  maragpu_host::OuField f;
};

extern "C" {

int maragpu_ou_field_create(double P0, double zeta, int k1, int seed, maragpu_ou_field **out)
{
  if (!out || k1 < 0 || k1 > 64 || !(P0 >= 0.0)) return MARAGPU_E_INVALID;
  maragpu_ou_field *f = new (std::nothrow) maragpu_ou_field;
  if (!f) return MARAGPU_E_NOMEM;
  f->f.build(P0, zeta, k1, seed);
  *out = f;
  return MARAGPU_OK;
}

int maragpu_ou_field_load(const void *bytes, size_t n, maragpu_ou_field **out)
{
  if (!bytes || !out) return MARAGPU_E_INVALID;
  maragpu_ou_field *f = new (std::nothrow) maragpu_ou_field;
  if (!f) return MARAGPU_E_NOMEM;
  if (!f->f.load(static_cast<const unsigned char *>(bytes), n)) {
    delete f;
    return MARAGPU_E_SIZE;
  }
  *out = f;
  return MARAGPU_OK;
}

size_t maragpu_ou_field_bytes(const maragpu_ou_field *f) { return f ? f->f.bytes() : 0; }

int maragpu_ou_field_save(const maragpu_ou_field *f, void *bytes, size_t n)
{
  if (!f || !bytes) return MARAGPU_E_INVALID;
  if (n < f->f.bytes()) return MARAGPU_E_SIZE;
  f->f.save(static_cast<unsigned char *>(bytes));
  return MARAGPU_OK;
}

int maragpu_ou_field_advance(maragpu_ou_field *f, double dt)
{
  if (!f) return MARAGPU_E_INVALID;
  f->f.advance(dt);
  return MARAGPU_OK;
}

int maragpu_ou_field_destroy(maragpu_ou_field *f)
{
  delete f;
  return MARAGPU_OK;
}

const maragpu_host::OuField *maragpu_ou_field_modes(const maragpu_ou_field *f) { return f ? &f->f : nullptr; }

}  // extern "C"
