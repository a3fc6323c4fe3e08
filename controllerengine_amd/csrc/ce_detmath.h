// Deterministic double-precision elementary functions, single source for the host (g++/clang) and
// for gfx950 device code (hipcc).
//
// Every function below is built only from IEEE-754 operations that both compilers lower to
// correctly rounded instructions (add, sub, mul, div, fma, sqrt) plus exact operations (rint, floor,
// fabs, exponent-field manipulation).  Compiled with -ffp-contract=off on both sides, a call returns
// the same bits on the CPU and on the GPU, which is what lets the FP64 pipeline match its CPU
// counterpart bit for bit.  Accuracy is a few ulp (checked against mpmath in tests/test_detmath.py).
//
// Methods (standard, published):
//   exp    Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, degree-13 Taylor polynomial, 2^k scaling.
//   expm1  degree-17 Taylor polynomial on |x| < 1/2, exp(x) - 1 beyond.
//   log    x = 2^e m, m in [sqrt(1/2), sqrt 2); log m = 2 atanh(s), s = (m-1)/(m+1), odd series in s.
//   sin/cos  three-part Cody-Waite reduction by pi/2, Taylor polynomials on |r| <= pi/4, quadrant map.
//   atan2  ratio t in [0,1], nearest breakpoint c = j/8, atan t = atan c + atan((t-c)/(1+tc)),
//          odd Taylor series on |u| <= 1/16, octant map.
// All constants are mathematical constants (ln 2, pi/2, atan(j/8), 1/n!) rounded to double.
#pragma once

#if defined(__HIPCC__)
#define CE_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define CE_HD static inline
#endif

#include <stdint.h>

namespace cedm {

CE_HD double bits_to_d(uint64_t u) { return __builtin_bit_cast(double, u); }
CE_HD uint64_t d_to_bits(double d) { return __builtin_bit_cast(uint64_t, d); }

CE_HD double inf() { return bits_to_d(0x7ff0000000000000ull); }
CE_HD double qnan() { return bits_to_d(0x7ff8000000000000ull); }
CE_HD bool isnan_(double x) { return !(x == x); }
CE_HD bool isfinite_(double x) { return (d_to_bits(x) & 0x7ff0000000000000ull) != 0x7ff0000000000000ull; }
CE_HD double fabs_(double x) { return __builtin_fabs(x); }
CE_HD double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
CE_HD double sqrt_(double x) { return __builtin_sqrt(x); }
CE_HD double floor_(double x) { return __builtin_floor(x); }
CE_HD double rint_(double x) { return __builtin_rint(x); }
// C99 fmin/fmax semantics for quiet NaN: the other operand is returned.
CE_HD double fmin_(double a, double b) { return __builtin_fmin(a, b); }
CE_HD double fmax_(double a, double b) { return __builtin_fmax(a, b); }

// 2^k for -1022 <= k <= 1023, exact.
CE_HD double pow2i(int k) { return bits_to_d((uint64_t)(k + 1023) << 52); }

// x * 2^k without intermediate overflow/underflow for the k range exp() produces.
CE_HD double scale2(double x, int k) {
  if (k > 1023) return x * pow2i(1023) * pow2i(k - 1023);
  if (k < -1022) return x * pow2i(k + 600) * pow2i(-600);
  return x * pow2i(k);
}

constexpr double LN2_HI = 0x1.62e42fefa4000p-1;     // ln 2 rounded to 40 bits: k*LN2_HI exact for |k| < 2^13
constexpr double LN2_LO = -0x1.8432a1b0e2634p-43;   // ln 2 - LN2_HI
constexpr double INV_LN2 = 0x1.71547652b82fep+0;
constexpr double INV_LN10 = 0x1.bcb7b1526e50ep-2;
constexpr double SQRT2 = 0x1.6a09e667f3bcdp+0;
constexpr double PIO2_1 = 0x1.921fb54400000p+0;     // pi/2 split in three: 33 + 33 + 53 bits
constexpr double PIO2_2 = 0x1.0b4611a600000p-34;
constexpr double PIO2_3 = 0x1.3198a2e037073p-69;
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
constexpr double PIO2_HI = 0x1.921fb54442d18p+0;
constexpr double PIO2_LO = 0x1.1a62633145c07p-54;
constexpr double PI_HI = 0x1.921fb54442d18p+1;
constexpr double PI_LO = 0x1.1a62633145c07p-53;

// ---------------------------------------------------------------------------------------------
CE_HD double exp(double x) {
  if (isnan_(x)) return x;
  if (x > 0x1.62e42fefa39efp+9) return inf();      // > ln(DBL_MAX)
  if (x < -0x1.74910d52d3051p+9) return 0.0;       // < ln(min subnormal / 2)
  const double kd = rint_(x * INV_LN2);
  const int k = (int)kd;
  double r = fma_(-kd, LN2_HI, x);
  r = fma_(-kd, LN2_LO, r);
  // Taylor coefficients 1/n!, n = 13 .. 1
  double p = 0x1.6124613a86d09p-33;
  p = fma_(p, r, 0x1.1eed8eff8d898p-29);
  p = fma_(p, r, 0x1.ae64567f544e4p-26);
  p = fma_(p, r, 0x1.27e4fb7789f5cp-22);
  p = fma_(p, r, 0x1.71de3a556c734p-19);
  p = fma_(p, r, 0x1.a01a01a01a01ap-16);
  p = fma_(p, r, 0x1.a01a01a01a01ap-13);
  p = fma_(p, r, 0x1.6c16c16c16c17p-10);
  p = fma_(p, r, 0x1.1111111111111p-7);
  p = fma_(p, r, 0x1.5555555555555p-5);
  p = fma_(p, r, 0x1.5555555555555p-3);
  p = fma_(p, r, 0.5);
  p = fma_(p, r, 1.0);
  p = fma_(p, r, 1.0);
  return scale2(p, k);
}

// ---------------------------------------------------------------------------------------------
CE_HD double log(double x) {
  if (isnan_(x) || x < 0.0) return qnan();
  if (x == 0.0) return -inf();
  if (!isfinite_(x)) return x;
  int eadj = 0;
  if (x < 0x1p-1022) { x *= 0x1p54; eadj = -54; }  // subnormal input
  const uint64_t b = d_to_bits(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023 + eadj;
  double m = bits_to_d((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
  if (m > SQRT2) { m *= 0.5; e += 1; }
  const double f = m - 1.0;                 // exact (Sterbenz)
  const double s = f / (2.0 + f);
  const double z = s * s;
  // R = z/3 + z^2/5 + ... + z^10/21
  double R = 0x1.8618618618618p-5;          // 1/21
  R = fma_(R, z, 0x1.af286bca1af28p-5);     // 1/19
  R = fma_(R, z, 0x1.e1e1e1e1e1e1ep-5);     // 1/17
  R = fma_(R, z, 0x1.1111111111111p-4);     // 1/15
  R = fma_(R, z, 0x1.3b13b13b13b14p-4);     // 1/13
  R = fma_(R, z, 0x1.745d1745d1746p-4);     // 1/11
  R = fma_(R, z, 0x1.c71c71c71c71cp-4);     // 1/9
  R = fma_(R, z, 0x1.2492492492492p-3);     // 1/7
  R = fma_(R, z, 0x1.999999999999ap-3);     // 1/5
  R = fma_(R, z, 0x1.5555555555555p-2);     // 1/3
  R = R * z;
  const double two_s = s + s;
  const double ed = (double)e;
  const double lo = fma_(ed, LN2_LO, two_s * R);
  return fma_(ed, LN2_HI, two_s + lo);
}

CE_HD double log10(double x) { return log(x) * INV_LN10; }

// exp(x) - 1 without cancellation near 0: Taylor series (degree 17) on |x| < 1/2, exp(x) - 1 beyond.
CE_HD double expm1(double x) {
  if (!(fabs_(x) < 0.5)) return exp(x) - 1.0;
  double p = 0x1.6827863b97d97p-53;         // 1/18! ... evaluated as x + x^2 q(x)
  p = fma_(p, x, 0x1.952c77030ad4ap-49);    // 1/17!
  p = fma_(p, x, 0x1.ae7f3e733b81fp-45);    // 1/16!
  p = fma_(p, x, 0x1.ae7f3e733b81fp-41);    // 1/15!
  p = fma_(p, x, 0x1.93974a8c07c9dp-37);    // 1/14!
  p = fma_(p, x, 0x1.6124613a86d09p-33);    // 1/13!
  p = fma_(p, x, 0x1.1eed8eff8d898p-29);    // 1/12!
  p = fma_(p, x, 0x1.ae64567f544e4p-26);    // 1/11!
  p = fma_(p, x, 0x1.27e4fb7789f5cp-22);    // 1/10!
  p = fma_(p, x, 0x1.71de3a556c734p-19);    // 1/9!
  p = fma_(p, x, 0x1.a01a01a01a01ap-16);    // 1/8!
  p = fma_(p, x, 0x1.a01a01a01a01ap-13);    // 1/7!
  p = fma_(p, x, 0x1.6c16c16c16c17p-10);    // 1/6!
  p = fma_(p, x, 0x1.1111111111111p-7);     // 1/5!
  p = fma_(p, x, 0x1.5555555555555p-5);     // 1/4!
  p = fma_(p, x, 0x1.5555555555555p-3);     // 1/3!
  p = fma_(p, x, 0.5);                      // 1/2!
  return fma_(p * x, x, x);
}

// ---------------------------------------------------------------------------------------------
// sin and cos of r, |r| <= pi/4 (+ a little): Taylor series, truncation below 1e-19 relative.
CE_HD double sin_poly(double r) {
  // sin r = r + r z q(z), q = -1/3! + z/5! - z^2/7! + ... + z^7/17!
  const double z = r * r;
  double q = 0x1.952c77030ad4ap-49;         // +1/17!
  q = fma_(q, z, -0x1.ae7f3e733b81fp-41);   // -1/15!
  q = fma_(q, z, 0x1.6124613a86d09p-33);    // +1/13!
  q = fma_(q, z, -0x1.ae64567f544e4p-26);   // -1/11!
  q = fma_(q, z, 0x1.71de3a556c734p-19);    // +1/9!
  q = fma_(q, z, -0x1.a01a01a01a01ap-13);   // -1/7!
  q = fma_(q, z, 0x1.1111111111111p-7);     // +1/5!
  q = fma_(q, z, -0x1.5555555555555p-3);    // -1/3!
  return fma_(q * z, r, r);
}

CE_HD double cos_poly(double r) {
  // cos r = 1 - z p(z), p = 1/2! - z/4! + z^2/6! - ... + z^8/18!
  const double z = r * r;
  double p = 0x1.6827863b97d97p-53;         // +1/18!
  p = fma_(p, z, -0x1.ae7f3e733b81fp-45);   // -1/16!
  p = fma_(p, z, 0x1.93974a8c07c9dp-37);    // +1/14!
  p = fma_(p, z, -0x1.1eed8eff8d898p-29);   // -1/12!
  p = fma_(p, z, 0x1.27e4fb7789f5cp-22);    // +1/10!
  p = fma_(p, z, -0x1.a01a01a01a01ap-16);   // -1/8!
  p = fma_(p, z, 0x1.6c16c16c16c17p-10);    // +1/6!
  p = fma_(p, z, -0x1.5555555555555p-5);    // -1/4!
  p = fma_(p, z, 0.5);                      // +1/2!
  return fma_(-z, p, 1.0);
}

// Reduce x by pi/2: returns r and the quadrant (k mod 4).
CE_HD double reduce_pio2(double x, int& quadrant) {
  const double kd = rint_(x * TWO_OVER_PI);
  double r = fma_(-kd, PIO2_1, x);
  r = fma_(-kd, PIO2_2, r);
  r = fma_(-kd, PIO2_3, r);
  const double kq = kd - 4.0 * floor_(kd * 0.25);   // kd mod 4, exact for |kd| < 2^53
  quadrant = isfinite_(kq) ? (int)kq : 0;
  return r;
}

CE_HD void sincos(double x, double& s, double& c) {
  int q;
  const double r = reduce_pio2(x, q);
  const double sr = sin_poly(r), cr = cos_poly(r);
  // quadrant map without lane-divergent branches: swap on odd quadrants, then signs
  const bool odd = (q & 1) != 0;
  double ss = odd ? cr : sr;
  double cc = odd ? sr : cr;
  ss = (q & 2) ? -ss : ss;
  cc = ((q + 1) & 2) ? -cc : cc;
  const bool fin = isfinite_(x);
  s = fin ? ss : qnan();
  c = fin ? cc : qnan();
}

CE_HD double sin(double x) { double s, c; sincos(x, s, c); return s; }
CE_HD double cos(double x) { double s, c; sincos(x, s, c); return c; }

// ---------------------------------------------------------------------------------------------
// atan(j/8) rounded to double (hi) and the residual (lo), j = 0..8
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __constant__ static const double ATAN_HI[9] = {
#else
static const double ATAN_HI[9] = {
#endif
    0.0, 0x1.fd5ba9aac2f6ep-4, 0x1.f5b75f92c80ddp-3, 0x1.6f61941e4def1p-2, 0x1.dac670561bb4fp-2,
    0x1.1e00babdefeb4p-1, 0x1.4978fa3269ee1p-1, 0x1.700a7c5784634p-1, 0x1.921fb54442d18p-1};
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __constant__ static const double ATAN_LO[9] = {
#else
static const double ATAN_LO[9] = {
#endif
    0.0, -0x1.cd37686760c17p-59, 0x1.8ab6e3cf7afbdp-57, -0x1.c63aae6f6e918p-56, 0x1.a2b7f222f65e2p-56,
    -0x1.928df287a668fp-58, 0x1.2419a87f2a458p-56, -0x1.8c34d25aadef6p-56, 0x1.1a62633145c07p-55};

// atan(t) for t in [0, 1]
CE_HD double atan01(double t) {
  const double jd = rint_(t * 8.0);
  const int j = (int)jd;
  const double c = jd * 0.125;
  const double u = (t - c) / fma_(t, c, 1.0);
  const double z = u * u;
  // odd series u - u^3/3 + u^5/5 - ... - u^15/15, |u| <= 1/16
  double p = -0x1.1111111111111p-4;         // -1/15
  p = fma_(p, z, 0x1.3b13b13b13b14p-4);     // +1/13
  p = fma_(p, z, -0x1.745d1745d1746p-4);    // -1/11
  p = fma_(p, z, 0x1.c71c71c71c71cp-4);     // +1/9
  p = fma_(p, z, -0x1.2492492492492p-3);    // -1/7
  p = fma_(p, z, 0x1.999999999999ap-3);     // +1/5
  p = fma_(p, z, -0x1.5555555555555p-2);    // -1/3
  p = p * z;
  const double series = fma_(p, u, u);
  return ATAN_HI[j] + (ATAN_LO[j] + series);
}

// C99 atan2 semantics for zeros, infinities and NaN.  The finite nonzero path is evaluated for
// every lane (one atan01 call, selects for the octant), special cases are patched in afterwards.
CE_HD double atan2(double y, double x) {
  const bool yneg = (d_to_bits(y) >> 63) != 0;
  const bool xneg = (d_to_bits(x) >> 63) != 0;
  const double ax = fabs_(x), ay = fabs_(y);
  const bool swap = ay > ax;
  const double num = swap ? ax : ay, den = swap ? ay : ax;
  double t = num / den;
  t = (t == t && t <= 1.0) ? t : 0.0;  // keeps atan01's table index in range for 0/0, inf/inf
  double a = atan01(t);
  a = swap ? PIO2_HI - (a - PIO2_LO) : a;
  double r = xneg ? PI_HI - (a - PI_LO) : a;
  if (ay == 0.0) r = xneg ? PI_HI : 0.0;
  else if (ax == 0.0) r = PIO2_HI;
  else if (!isfinite_(ax) || !isfinite_(ay)) {
    if (!isfinite_(ax) && !isfinite_(ay)) r = xneg ? 3.0 * (PIO2_HI * 0.5) : PIO2_HI * 0.5;
    else if (!isfinite_(ax)) r = xneg ? PI_HI : 0.0;
    else r = PIO2_HI;
  }
  if (isnan_(x) || isnan_(y)) return qnan();
  return yneg ? -r : r;
}

CE_HD double hypot(double x, double y) {
  const double ax = fabs_(x), ay = fabs_(y);
  if (!isfinite_(ax) || !isfinite_(ay)) {
    if (ax == inf() || ay == inf()) return inf();
    return qnan();
  }
  const double big = fmax_(ax, ay);
  if (big > 0x1p500 || (big < 0x1p-500 && big > 0.0)) {
    // rescale by an exact power of two so that the squares neither overflow nor underflow
    const double s = big > 0x1p500 ? 0x1p-600 : 0x1p600;
    const double xs = ax * s, ys = ay * s;
    return sqrt_(fma_(xs, xs, ys * ys)) / s;
  }
  return sqrt_(fma_(ax, ax, ay * ay));
}

// Single-precision logf / expf / sinf / cosf / atan2f: the double functions above rounded to float.
// Their double results lie within a few double ulp of the exact value, so the rounded float is the
// correctly rounded one except when the exact value falls within that distance of a midpoint
// between two floats.
CE_HD float logf_(float x) { return (float)log((double)x); }
CE_HD float expf_(float x) { return (float)exp((double)x); }
CE_HD void sincosf_(float x, float& s, float& c) {
  double sd, cd;
  sincos((double)x, sd, cd);
  s = (float)sd;
  c = (float)cd;
}
CE_HD float atan2f_(float y, float x) { return (float)atan2((double)y, (double)x); }

}  // namespace cedm
