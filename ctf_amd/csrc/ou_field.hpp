/* Host-side stochastic driving field (see ou_field.cpp): modes, amplitudes and the state
 * of their Ornstein-Uhlenbeck processes. */
#pragma once

#include <cstddef>
#include <vector>

namespace maragpu_host {

struct OuField {
  int k1 = 0, numk = 0;
  double P0 = 0.0, zeta = 0.0;
  std::vector<double> K;            // wave vectors, 3 per mode (2 pi times integers)
  std::vector<double> F;            // complex amplitudes: axis-major, (re, im) per mode
  std::vector<unsigned char> proc;  // the processes, in the serialized layout

  void build(double P0, double zeta, int k1, int seed);
  size_t bytes() const;
  void save(unsigned char *b) const;
  bool load(const unsigned char *b, size_t n);
  void advance(double dt);
};

}  // namespace maragpu_host

struct maragpu_ou_field;
extern "C" const maragpu_host::OuField *maragpu_ou_field_modes(const maragpu_ou_field *f);
