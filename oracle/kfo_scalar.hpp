// kfo_scalar.hpp -- scalar type of the oracle.
//
// TEST INFRASTRUCTURE ONLY: part of the CPU oracle that checks the HIP
// fitter.  Nothing in the product links or calls it.
//
// With KFO_COUNTING defined the oracle is built on `Counted`, a double that
// counts its arithmetic.  An operation is counted only when it is not
// trivially exact: additions/subtractions with an exact-zero operand and
// multiplications/divisions by exactly 0, +1 or -1 are free.  This removes
// the structural zeros and identities of the reference's dense matrix
// products (D, jacToLocal, jacToGlobal, H) so the count is the algorithmic
// flop count of the fit.  Transcendentals (sin cos atan2 log pow hypot exp)
// are counted in a separate counter; sqrt counts as one flop.
//
// Flops inside an F32Region are also counted in g_flops32: the operations
// the HIP fitter's mixed-precision mode runs in float (transport Jacobian,
// bound-state covariance transport, material formulas).  Flops after
// kfo_mark_backward() (called where the smoother starts) are also counted in
// g_flops_bwd: the share of the HIP fitter's second kernel.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>

namespace kfo {

#ifdef KFO_COUNTING

extern thread_local uint64_t g_flops;
extern thread_local uint64_t g_flops32;
extern thread_local uint64_t g_trans;
extern thread_local int g_in32;
extern thread_local uint64_t g_flops_bwd;
extern thread_local uint64_t g_flops32_bwd;
extern thread_local int g_bwd;

inline void kfo_count() {
  ++g_flops;
  if (g_in32) ++g_flops32;
  if (g_bwd) ++g_flops_bwd;
  if (g_in32 && g_bwd) ++g_flops32_bwd;
}
inline void kfo_mark_backward() { g_bwd = 1; }

struct F32Region {
  F32Region() { ++g_in32; }  // NOLINT
  ~F32Region() { --g_in32; }
};

struct Counted {
  double v;
  Counted() : v(0.0) {}
  Counted(double x) : v(x) {}
  explicit operator double() const { return v; }
};

inline bool kfo_unit(double x) { return x == 1.0 || x == -1.0 || x == 0.0; }

inline Counted operator+(Counted a, Counted b) {
  if (a.v != 0.0 && b.v != 0.0) kfo_count();
  return Counted(a.v + b.v);
}
inline Counted operator-(Counted a, Counted b) {
  if (a.v != 0.0 && b.v != 0.0) kfo_count();
  return Counted(a.v - b.v);
}
inline Counted operator*(Counted a, Counted b) {
  if (!kfo_unit(a.v) && !kfo_unit(b.v)) kfo_count();
  return Counted(a.v * b.v);
}
inline Counted operator/(Counted a, Counted b) {
  if (a.v != 0.0 && b.v != 1.0 && b.v != -1.0) kfo_count();
  return Counted(a.v / b.v);
}
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted &operator+=(Counted &a, Counted b) { return a = a + b; }
inline Counted &operator-=(Counted &a, Counted b) { return a = a - b; }
inline Counted &operator*=(Counted &a, Counted b) { return a = a * b; }
inline Counted &operator/=(Counted &a, Counted b) { return a = a / b; }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }

inline double kval(Counted x) { return x.v; }
inline Counted ksqrt(Counted x) { kfo_count(); return Counted(std::sqrt(x.v)); }
inline Counted kabs(Counted x) { return Counted(std::fabs(x.v)); }
inline Counted kfloor(Counted x) { return Counted(std::floor(x.v)); }
inline Counted kcopysign(Counted a, Counted b) { return Counted(std::copysign(a.v, b.v)); }
inline Counted ksin(Counted x) { ++g_trans; return Counted(std::sin(x.v)); }
inline Counted kcos(Counted x) { ++g_trans; return Counted(std::cos(x.v)); }
inline Counted katan2(Counted y, Counted x) { ++g_trans; return Counted(std::atan2(y.v, x.v)); }
inline Counted klog(Counted x) { ++g_trans; return Counted(std::log(x.v)); }
inline Counted kexp(Counted x) { ++g_trans; return Counted(std::exp(x.v)); }
inline Counted kpow(Counted x, Counted y) { ++g_trans; return Counted(std::pow(x.v, y.v)); }
inline Counted khypot(Counted x, Counted y) { ++g_trans; return Counted(std::hypot(x.v, y.v)); }
inline bool kisnan(Counted x) { return std::isnan(x.v); }

using Real = Counted;

#else

inline double kval(double x) { return x; }
inline double ksqrt(double x) { return std::sqrt(x); }
inline double kabs(double x) { return std::fabs(x); }
inline double kfloor(double x) { return std::floor(x); }
inline double kcopysign(double a, double b) { return std::copysign(a, b); }
inline double ksin(double x) { return std::sin(x); }
inline double kcos(double x) { return std::cos(x); }
inline double katan2(double y, double x) { return std::atan2(y, x); }
inline double klog(double x) { return std::log(x); }
inline double kexp(double x) { return std::exp(x); }
inline double kpow(double x, double y) { return std::pow(x, y); }
inline double khypot(double x, double y) { return std::hypot(x, y); }
inline bool kisnan(double x) { return std::isnan(x); }

using Real = double;

struct F32Region {
  F32Region() {}  // user-provided: a scope marker, not an unused variable
};
inline void kfo_mark_backward() {}

#endif

inline Real kmax(Real a, Real b) { return (a < b) ? b : a; }
inline Real kmin(Real a, Real b) { return (b < a) ? b : a; }

}  // namespace kfo
