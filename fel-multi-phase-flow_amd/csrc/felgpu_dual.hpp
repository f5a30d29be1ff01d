// felgpu_dual.hpp -- the element type of the tangent mode: an FP64 value and
// its derivative along one direction of the inputs (forward-mode
// differentiation by dual numbers).
//
// Every kernel of the library is a template on the element type of the flow
// (double, float) and takes the volume-fraction class in PT<T>, so the
// tangent mode runs the same kernels with T = PT<T> = Dual.  The value part
// is computed by exactly the FP64 operations of the FP64 path (the library
// is built without contraction), so it is bitwise equal to the FP64 path's
// result; the derivative part is the chain rule applied to each operation.
// Branches (upwind direction, limiter and truncation tests, clamps, the
// Gauss-Seidel stop test) are taken on the value, which makes the derivative
// that of the branch the value takes: the exact derivative of the computed
// discrete map wherever no test sits on its threshold.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

namespace fel {

// The float and double overloads stay visible next to the Dual ones below,
// so an unqualified call inside this namespace resolves by argument type.
using ::acos;
using ::copysign;
using ::cos;
using ::cosh;
using ::exp;
using ::fabs;
using ::fmax;
using ::fmin;
using ::sin;
using ::sqrt;
using ::tanh;

struct Dual {
  double v, d;
  Dual() = default;  // trivial, so that arrays of Dual can live in LDS
  // constants (and every value that does not depend on the seeded inputs)
  __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}
  __host__ __device__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
};

// Arithmetic: value as in FP64, derivative by the chain rule.
__host__ __device__ __forceinline__ Dual operator+(Dual a) { return a; }
__host__ __device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__host__ __device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__host__ __device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__host__ __device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__host__ __device__ __forceinline__ Dual operator*(Dual a, Dual b) {
  return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
__host__ __device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__host__ __device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__host__ __device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ __forceinline__ Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__host__ __device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return Dual(q, -(q * b.d) / b.v);
}
__host__ __device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
__host__ __device__ __forceinline__ Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
__host__ __device__ __forceinline__ Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
__host__ __device__ __forceinline__ Dual& operator/=(Dual& a, Dual b) { return a = a / b; }

// Comparisons act on the value.
#define FEL_DUAL_CMP(op)                                                                       \
  __host__ __device__ __forceinline__ bool operator op(Dual a, Dual b) { return a.v op b.v; }  \
  __host__ __device__ __forceinline__ bool operator op(Dual a, double b) { return a.v op b; }  \
  __host__ __device__ __forceinline__ bool operator op(double a, Dual b) { return a op b.v; }
FEL_DUAL_CMP(==)
FEL_DUAL_CMP(!=)
FEL_DUAL_CMP(<)
FEL_DUAL_CMP(<=)
FEL_DUAL_CMP(>)
FEL_DUAL_CMP(>=)
#undef FEL_DUAL_CMP

// Elementary functions.  A zero derivative stays zero (no 0*inf where the
// function's derivative is unbounded, e.g. acos of a clamped +-1).
__host__ __device__ __forceinline__ Dual sqrt(Dual a) {
  const double s = ::sqrt(a.v);
  // at a.v == 0 (a vanishing gradient norm) the one-sided derivative of the
  // value is taken as 0
  return Dual(s, (a.d == 0.0 || s == 0.0) ? 0.0 : a.d / (2.0 * s));
}
__host__ __device__ __forceinline__ Dual exp(Dual a) {
  const double e = ::exp(a.v);
  return Dual(e, a.d == 0.0 ? 0.0 : e * a.d);
}
__host__ __device__ __forceinline__ Dual cosh(Dual a) {
  return Dual(::cosh(a.v), a.d == 0.0 ? 0.0 : ::sinh(a.v) * a.d);
}
__host__ __device__ __forceinline__ Dual tanh(Dual a) {
  const double t = ::tanh(a.v);
  return Dual(t, a.d == 0.0 ? 0.0 : (1.0 - t * t) * a.d);
}
__host__ __device__ __forceinline__ Dual cos(Dual a) {
  return Dual(::cos(a.v), a.d == 0.0 ? 0.0 : -::sin(a.v) * a.d);
}
__host__ __device__ __forceinline__ Dual sin(Dual a) {
  return Dual(::sin(a.v), a.d == 0.0 ? 0.0 : ::cos(a.v) * a.d);
}
__host__ __device__ __forceinline__ Dual acos(Dual a) {
  return Dual(::acos(a.v), a.d == 0.0 ? 0.0 : -a.d / ::sqrt(1.0 - a.v * a.v));
}
__host__ __device__ __forceinline__ Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
// fmin/fmax select one operand, as the FP64 functions do (ties: the first).
__host__ __device__ __forceinline__ Dual fmin(Dual a, Dual b) { return b.v < a.v ? b : a; }
__host__ __device__ __forceinline__ Dual fmax(Dual a, Dual b) { return b.v > a.v ? b : a; }
__host__ __device__ __forceinline__ Dual fmin(Dual a, double b) { return fmin(a, Dual(b)); }
__host__ __device__ __forceinline__ Dual fmin(double a, Dual b) { return fmin(Dual(a), b); }
__host__ __device__ __forceinline__ Dual fmax(Dual a, double b) { return fmax(a, Dual(b)); }
__host__ __device__ __forceinline__ Dual fmax(double a, Dual b) { return fmax(Dual(a), b); }
// |a| with the sign of b: the derivative follows a's sign flip only.
__host__ __device__ __forceinline__ Dual copysign(Dual a, Dual b) {
  const double r = ::copysign(a.v, b.v);
  const bool flip = (r < 0.0) != (a.v < 0.0);  // a.v != 0 wherever the library uses it
  return Dual(r, flip ? -a.d : a.d);
}
__host__ __device__ __forceinline__ Dual copysign(Dual a, double b) { return copysign(a, Dual(b)); }
__host__ __device__ __forceinline__ Dual copysign(double a, Dual b) { return copysign(Dual(a), b); }

// Two places where the reference's FP64 map is flat, or noise, at rounding
// level, and the derivative of the exact-arithmetic formula would not
// describe it:
//  * a - b with b non-zero but absorbed by rounding (a - b == a): the result
//    does not move with b for any perturbation the FP64 map resolves, so b's
//    derivative is dropped;
//  * the unit normal r/(|r| + 1e-30|dxinv|) of a gradient r that is flat,
//    |r|^2 < flat2 (no interface: the volume fraction is 0 or 1 to ~10
//    digits and r is rounding noise): its direction has derivatives of
//    order r'/|r| that grow without bound as |r| falls to rounding level,
//    while the FP64 map's response there is noise; the derivative is taken
//    as zero.
__host__ __device__ __forceinline__ double sub_rounded(double a, double b) { return a - b; }
__host__ __device__ __forceinline__ float sub_rounded(float a, float b) { return a - b; }
__host__ __device__ __forceinline__ Dual sub_rounded(Dual a, Dual b) {
  Dual r = a - b;
  if (b.v != 0.0 && r.v == a.v) r.d = a.d;
  return r;
}
__host__ __device__ __forceinline__ void flat_zero(double*, double) {}
__host__ __device__ __forceinline__ void flat_zero(float*, double) {}
__host__ __device__ __forceinline__ void flat_zero(Dual* r, double flat2) {
  if (r[0].v * r[0].v + r[1].v * r[1].v + r[2].v * r[2].v < flat2) r[0].d = r[1].d = r[2].d = 0.0;
}

// Value of an element of any of the library's types (scalar read-backs,
// stop tests and monitors use the value only).
__host__ __device__ __forceinline__ double val(double x) { return x; }
__host__ __device__ __forceinline__ double val(float x) { return (double)x; }
__host__ __device__ __forceinline__ double val(Dual x) { return x.v; }

// Element type of the volume-fraction class (volume fraction, its fluxes,
// the contact-line maps, divergence and pressure correction) for a given
// flow element type: FP64 with an FP64 or FP32 flow, Dual in the tangent mode.
template <class T>
struct PhaseOf {
  using type = double;
};
template <>
struct PhaseOf<Dual> {
  using type = Dual;
};
template <class T>
using PT = typename PhaseOf<T>::type;

template <class T>
inline constexpr bool is_f32 = std::is_same<T, float>::value;
template <class T>
inline constexpr bool is_dual = std::is_same<T, Dual>::value;

}  // namespace fel
