// synth_math.h -- counter-based random draws that give the same bits on the
// host (workloads.cpp, g++) and on the device (k_synth.hip, hipcc).
//
// u01: SplitMix64 of (seed, element, draw) -> f64 in [0, 1).
// gauss: Box-Muller, with ln, cos and sqrt written out in IEEE f64 add,
// subtract, multiply and divide only.  Each of those is correctly rounded
// on both sides, and with contraction off (-ffp-contract=off) both compilers
// evaluate them in the order written, so a host replay of a device-jittered
// mesh is bit-exact.  Library log/cos/sqrt are not: their last bits differ
// between the host's libm and the device's.  Accuracy is a few ulp, far
// below what the f32 vertices keep.
#pragma once

#include <cstdint>
#include <cstring>

#ifndef LBVH_HD
#define LBVH_HD
#endif

namespace lbvh_synth {

LBVH_HD inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Uniform f64 in [0, 1) for draw k of element i; hseed = splitmix64(seed).
LBVH_HD inline double u01(uint64_t hseed, uint64_t i, uint64_t k) {
  const uint64_t h = splitmix64(hseed ^ (i * 64 + k));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

LBVH_HD inline uint64_t bits_of(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return b;
}
LBVH_HD inline double from_bits(uint64_t b) {
  double x;
  memcpy(&x, &b, 8);
  return x;
}

// Natural log of a positive normal x: x = m 2^e with m in [1/sqrt2, sqrt2),
// ln m = 2 atanh(s), s = (m - 1)/(m + 1), |s| < 0.172, summed as the odd
// series to s^25 (the next term is below 2^-60 of the sum).
LBVH_HD inline double ln_pos(double x) {
  const uint64_t b = bits_of(x);
  int e = (int)((b >> 52) & 0x7FF) - 1023;
  double m = from_bits((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 1.4142135623730951) {
    m = m * 0.5;  // exact
    e += 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 25.0;
  for (int k = 23; k >= 1; k -= 2) p = p * s2 + 1.0 / (double)k;
  // ln 2 split so that e * ln2_hi is exact (21 significant bits)
  const double ln2_hi = 0.6931467056274414;       // 0x1.62e42p-1
  const double ln2_lo = 4.7493250390316726e-07;   // ln 2 - ln2_hi
  return (double)e * ln2_hi + ((double)e * ln2_lo + 2.0 * s * p);
}

// cos(2 pi y) for y in [0, 1): with r = |y - 1/2| (exact), cos(2 pi y) =
// -cos(2 pi r); for r > 1/4, -cos(2 pi r) = cos(2 pi (1/2 - r)) (exact).  The
// angle t <= pi/2 goes into the Taylor series to t^24 (the next term is
// below 2^-62), evaluated in t^2 by Horner with the coefficients
// (-1)^k / (2k)! as f64 constants: multiplies and adds only.
LBVH_HD inline double cos_2pi(double y) {
  double a = y - 0.5;
  if (a < 0.0) a = -a;
  double sign = -1.0;
  if (a > 0.25) {
    a = 0.5 - a;
    sign = 1.0;
  }
  const double t = a * 6.283185307179586;
  const double t2 = t * t;
  double p = 1.6117375710961184e-24;  //  1/24!
  p = p * t2 - 8.896791392450574e-22;  // -1/22!
  p = p * t2 + 4.110317623312165e-19;  // 1/20!
  p = p * t2 - 1.5619206968586225e-16;  // -1/18!
  p = p * t2 + 4.779477332387385e-14;  // 1/16!
  p = p * t2 - 1.1470745597729725e-11;  // -1/14!
  p = p * t2 + 2.08767569878681e-09;  // 1/12!
  p = p * t2 - 2.755731922398589e-07;  // -1/10!
  p = p * t2 + 2.48015873015873e-05;  // 1/8!
  p = p * t2 - 0.001388888888888889;  // -1/6!
  p = p * t2 + 0.041666666666666664;  // 1/4!
  p = p * t2 - 0.5;  // -1/2!
  p = p * t2 + 1.0;
  return sign * p;
}

// sqrt of a positive normal v = m 2^e (m in [1, 2)): Newton steps for
// 1/sqrt(v), multiplies only, from 2^-ceil(e/2), which is within a factor
// of sqrt 2 of it and not above sqrt 3 times it, so the steps converge
// (relative error 0.41 -> 1e-16 in seven; eight leave margin); then
// s = v * (1/sqrt(v)) and one correcting step s + (v - s^2) * r / 2.
LBVH_HD inline double sqrt_pos(double v) {
  const uint64_t b = bits_of(v);
  const int e = (int)((b >> 52) & 0x7FF) - 1023;
  double r = from_bits((uint64_t)(1023 - ((e + 1) >> 1)) << 52);
  const double h = 0.5 * v;
  for (int it = 0; it < 8; ++it) r = r * (1.5 - h * r * r);
  const double s = v * r;
  return s + (v - s * s) * (0.5 * r);
}

// Standard normal from draws k and k + 1 of element i (Box-Muller).
LBVH_HD inline double gauss(uint64_t hseed, uint64_t i, uint64_t k) {
  double a = u01(hseed, i, k);
  const double c = u01(hseed, i, k + 1);
  if (a < 1e-300) a = 1e-300;
  return sqrt_pos(-2.0 * ln_pos(a)) * cos_2pi(c);
}

}  // namespace lbvh_synth
