/* Synthetic initial data: uniform density, pressure and field plus seeded random-phase
 * Fourier fields, written as a Mara primitive array (AoS [i][j][k][q], ghosts filled
 * periodically). It plays the role of a Lua init_prim callback (Mara/mara.cpp:531-632)
 * for the turbulence set-ups; the velocity and field perturbations are defined here:
 *
 *  velocity:  v(x) = sum_n 2 Re[ a_n exp(2 pi i n.x) ] over integer wavevectors n with
 *             kmin <= |n| <= kmax (one of each +-n pair), a_n a complex Gaussian vector
 *             projected perpendicular to n, scaled so the rms of |v| over interior zones
 *             is v_rms. With four_velocity set the field is u = W v and v = u/sqrt(1+u^2).
 *  field:     dB = sum_n 2 Re[ (D_n x A_n) exp(i theta_n.(j + 1/2)) ] with A_n a complex
 *             Gaussian vector potential on kmin <= |n| <= kmax and D_n the Fourier symbol of
 *             the corner-centred divergence of luaC_mean_max_divB (Mara/mara.cpp:2298-2314):
 *             D_x = (e^{i tx} - 1)(1 + e^{i ty})(1 + e^{i tz})/4 and cyclic, t = 2 pi n/N.
 *             D_n . (D_n x A_n) = 0, so that stencil reads exactly zero divergence (to
 *             round-off); scaled to an rms of dB_rms.
 * The sums are evaluated as three partial 1-D transforms (pruned to the selected modes).
 */
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/mara_host.h"

namespace {

typedef std::complex<double> cplx;

struct Rng {
  // splitmix64
  std::uint64_t s;
  explicit Rng(std::uint64_t seed) : s(seed) {}
  std::uint64_t next()
  {
    std::uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal()
  {
    // Box-Muller, one value per call (the second is discarded for determinism)
    const double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

struct Mode {
  int n[3];
  cplx c[3];
};

/* f_c(i,j,k) = sum over modes m and their conjugates of c_m e^{i t.(idx + 1/2)} */
void synthesize(const std::vector<Mode> &modes, int K, const int *N, std::vector<double> *out)
{
  const int W = 2 * K + 1;
  std::vector<cplx> E[3];
  for (int d = 0; d < 3; ++d) {
    E[d].resize((size_t)W * N[d]);
    for (int m = -K; m <= K; ++m)
      for (int i = 0; i < N[d]; ++i) {
        const double ph = 2.0 * M_PI * m * (i + 0.5) / N[d];
        E[d][(size_t)(m + K) * N[d] + i] = cplx(std::cos(ph), std::sin(ph));
      }
  }
  for (int comp = 0; comp < 3; ++comp) {
    // dense coefficient cube with Hermitian symmetry
    std::vector<cplx> C((size_t)W * W * W, cplx(0.0, 0.0));
    for (const Mode &md : modes) {
      const size_t a = ((size_t)(md.n[0] + K) * W + (md.n[1] + K)) * W + (md.n[2] + K);
      const size_t b = ((size_t)(-md.n[0] + K) * W + (-md.n[1] + K)) * W + (-md.n[2] + K);
      C[a] += md.c[comp];
      C[b] += std::conj(md.c[comp]);
    }
    // T1[nx][ny][k]
    std::vector<cplx> T1((size_t)W * W * N[2]);
#pragma omp parallel for schedule(static)
    for (int ab = 0; ab < W * W; ++ab) {
      for (int k = 0; k < N[2]; ++k) {
        cplx s(0.0, 0.0);
        for (int c = 0; c < W; ++c) s += C[(size_t)ab * W + c] * E[2][(size_t)c * N[2] + k];
        T1[(size_t)ab * N[2] + k] = s;
      }
    }
    // T2[nx][j][k]
    std::vector<cplx> T2((size_t)W * N[1] * N[2]);
#pragma omp parallel for schedule(static)
    for (int a = 0; a < W; ++a) {
      for (int j = 0; j < N[1]; ++j) {
        for (int k = 0; k < N[2]; ++k) {
          cplx s(0.0, 0.0);
          for (int b = 0; b < W; ++b)
            s += T1[((size_t)a * W + b) * N[2] + k] * E[1][(size_t)b * N[1] + j];
          T2[((size_t)a * N[1] + j) * N[2] + k] = s;
        }
      }
    }
    std::vector<double> &f = out[comp];
    f.assign((size_t)N[0] * N[1] * N[2], 0.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N[0]; ++i) {
      for (int j = 0; j < N[1]; ++j) {
        for (int k = 0; k < N[2]; ++k) {
          double s = 0.0;
          for (int a = 0; a < W; ++a) {
            const cplx t = T2[((size_t)a * N[1] + j) * N[2] + k] * E[0][(size_t)a * N[0] + i];
            s += t.real();
          }
          f[((size_t)i * N[1] + j) * N[2] + k] = s;
        }
      }
    }
  }
}

bool in_half_space(int x, int y, int z)
{
  return x > 0 || (x == 0 && (y > 0 || (y == 0 && z > 0)));
}

std::vector<Mode> draw_modes(Rng &rng, double kmin, double kmax, int K, bool project)
{
  std::vector<Mode> modes;
  for (int x = -K; x <= K; ++x)
    for (int y = -K; y <= K; ++y)
      for (int z = -K; z <= K; ++z) {
        if (!in_half_space(x, y, z)) continue;
        const double k = std::sqrt((double)(x * x + y * y + z * z));
        if (k < kmin || k > kmax) continue;
        Mode m;
        m.n[0] = x; m.n[1] = y; m.n[2] = z;
        for (int c = 0; c < 3; ++c) {
          const double re = rng.normal(), im = rng.normal();
          m.c[c] = cplx(re, im);
        }
        if (project) {
          const double k2 = k * k;
          const cplx nd = (double)x * m.c[0] + (double)y * m.c[1] + (double)z * m.c[2];
          m.c[0] -= nd * ((double)x / k2);
          m.c[1] -= nd * ((double)y / k2);
          m.c[2] -= nd * ((double)z / k2);
        }
        modes.push_back(m);
      }
  return modes;
}

}  // namespace

extern "C" int mara_init_turbulence(const mara_ic_params *ic, double *P, size_t n)
{
  if (!ic || !P) return -1;
  const int *N = ic->N;
  const int Ng = ic->Ng;
  for (int d = 0; d < 3; ++d) if (N[d] < 1) return -1;
  const int n0 = N[0] + 2 * Ng, n1 = N[1] + 2 * Ng, n2 = N[2] + 2 * Ng;
  if (n != (size_t)n0 * n1 * n2 * 8) return -4;
  const size_t nint = (size_t)N[0] * N[1] * N[2];

  Rng rng(ic->seed);
  std::vector<double> v[3], b[3];
  for (int c = 0; c < 3; ++c) v[c].assign(nint, 0.0);
  if (ic->v_rms > 0.0 && ic->v_kmax >= ic->v_kmin && ic->v_kmax >= 1.0) {
    const int K = (int)std::floor(ic->v_kmax);
    std::vector<Mode> m = draw_modes(rng, ic->v_kmin, ic->v_kmax, K, true);
    synthesize(m, K, N, v);
    double s = 0.0;
    for (size_t z = 0; z < nint; ++z) s += v[0][z]*v[0][z] + v[1][z]*v[1][z] + v[2][z]*v[2][z];
    const double rms = std::sqrt(s / nint);
    const double f = rms > 0.0 ? ic->v_rms / rms : 0.0;
    for (int c = 0; c < 3; ++c) for (size_t z = 0; z < nint; ++z) v[c][z] *= f;
    if (ic->four_velocity) {
      for (size_t z = 0; z < nint; ++z) {
        const double u2 = v[0][z]*v[0][z] + v[1][z]*v[1][z] + v[2][z]*v[2][z];
        const double iw = 1.0 / std::sqrt(1.0 + u2);
        for (int c = 0; c < 3; ++c) v[c][z] *= iw;
      }
    }
  }
  for (int c = 0; c < 3; ++c) b[c].assign(nint, 0.0);
  if (ic->dB_rms > 0.0 && ic->A_kmax >= ic->A_kmin && ic->A_kmax >= 1.0) {
    const int K = (int)std::floor(ic->A_kmax);
    std::vector<Mode> m = draw_modes(rng, ic->A_kmin, ic->A_kmax, K, false);
    for (Mode &md : m) {
      cplx e[3], D[3];
      for (int d = 0; d < 3; ++d) {
        const double t = 2.0 * M_PI * md.n[d] / N[d];
        e[d] = cplx(std::cos(t), std::sin(t));
      }
      const cplx one(1.0, 0.0);
      D[0] = (e[0] - one) * (one + e[1]) * (one + e[2]) * 0.25;
      D[1] = (e[1] - one) * (one + e[0]) * (one + e[2]) * 0.25;
      D[2] = (e[2] - one) * (one + e[0]) * (one + e[1]) * 0.25;
      const cplx a0 = md.c[0], a1 = md.c[1], a2 = md.c[2];
      md.c[0] = D[1] * a2 - D[2] * a1;
      md.c[1] = D[2] * a0 - D[0] * a2;
      md.c[2] = D[0] * a1 - D[1] * a0;
    }
    synthesize(m, K, N, b);
    double s = 0.0;
    for (size_t z = 0; z < nint; ++z) s += b[0][z]*b[0][z] + b[1][z]*b[1][z] + b[2][z]*b[2][z];
    const double rms = std::sqrt(s / nint);
    const double f = rms > 0.0 ? ic->dB_rms / rms : 0.0;
    for (int c = 0; c < 3; ++c) for (size_t z = 0; z < nint; ++z) b[c][z] *= f;
  }
  // fill interior and ghosts (each ghost takes its periodic image)
  for (int i = 0; i < n0; ++i) {
    const int ii = ((i - Ng) % N[0] + N[0]) % N[0];
    for (int j = 0; j < n1; ++j) {
      const int jj = ((j - Ng) % N[1] + N[1]) % N[1];
      for (int k = 0; k < n2; ++k) {
        const int kk = ((k - Ng) % N[2] + N[2]) % N[2];
        const size_t s = ((size_t)ii * N[1] + jj) * N[2] + kk;
        double *p = P + (((size_t)i * n1 + j) * n2 + k) * 8;
        p[0] = ic->rho;
        p[1] = ic->pre;
        p[2] = v[0][s];
        p[3] = v[1][s];
        p[4] = v[2][s];
        p[5] = ic->B0[0] + b[0][s];
        p[6] = ic->B0[1] + b[1][s];
        p[7] = ic->B0[2] + b[2][s];
      }
    }
  }
  return 0;
}
