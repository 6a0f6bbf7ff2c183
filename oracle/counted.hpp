/* ORACLE — test infrastructure only. Nothing in the product links this.
 *
 * A floating-point scalar that counts the arithmetic it performs. The oracle's
 * per-point physics is written against a scalar template parameter; instantiating
 * it with `Counted` instead of `double` yields the algorithmic operation count of
 * the same data-dependent path (Newton iterations included) without changing a
 * single rounding: every operation is carried out on the wrapped double.
 *
 * Counting convention: +, -, *, / and sqrt count one flop each; pow, acos, cos,
 * sin are counted separately as transcendentals; comparisons, fabs and copies
 * count nothing.
 */
#ifndef ORACLE_COUNTED_HPP
#define ORACLE_COUNTED_HPP

#include <cmath>
#include <cstdint>
#include <type_traits>

namespace oracle {

struct FlopTally {
  std::uint64_t add = 0, mul = 0, div = 0, sqrt = 0, trans = 0;
  std::uint64_t flops() const { return add + mul + div + sqrt; }
};

extern thread_local FlopTally g_tally;

struct Counted {
  double v;
  Counted() : v(0.0) {}
  Counted(double x) : v(x) {}
  template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
  Counted(T x) : v(static_cast<double>(x)) {}
  Counted &operator+=(Counted o) { ++g_tally.add; v += o.v; return *this; }
  Counted &operator-=(Counted o) { ++g_tally.add; v -= o.v; return *this; }
  Counted &operator*=(Counted o) { ++g_tally.mul; v *= o.v; return *this; }
  Counted &operator/=(Counted o) { ++g_tally.div; v /= o.v; return *this; }
  Counted operator-() const { return Counted(-v); }
};

inline Counted operator+(Counted a, Counted b) { ++g_tally.add; return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { ++g_tally.add; return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { ++g_tally.mul; return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { ++g_tally.div; return Counted(a.v / b.v); }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }

inline Counted sqrt(Counted a) { ++g_tally.sqrt; return Counted(std::sqrt(a.v)); }
inline Counted fabs(Counted a) { return Counted(std::fabs(a.v)); }
inline Counted pow(Counted a, Counted b) { ++g_tally.trans; return Counted(std::pow(a.v, b.v)); }
inline Counted acos(Counted a) { ++g_tally.trans; return Counted(std::acos(a.v)); }
inline Counted cos(Counted a) { ++g_tally.trans; return Counted(std::cos(a.v)); }
inline Counted sin(Counted a) { ++g_tally.trans; return Counted(std::sin(a.v)); }
inline bool isnan(Counted a) { return std::isnan(a.v); }
inline bool isinf(Counted a) { return std::isinf(a.v); }
inline double value(Counted a) { return a.v; }
inline double value(double a) { return a; }

} // namespace oracle

#endif
