// vxo_math.h — oracle scalar math: IEEE binary32 arithmetic with no FMA
// contraction (compile with -ffp-contract=off), WGSL builtin semantics and the
// transcendental functions the passes use.
//
// TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP passes.
// It is never linked into, loaded by, or called from the product library.
//
// Transcendentals.  WGSL leaves sin/cos/exp/pow accuracy to the
// implementation (WGSL spec, "Floating Point Accuracy": absolute error 2^-11
// for sin/cos on [-pi, pi], 3 + 2|x| ULP for exp, inherited from exp2/log2
// for pow).  The reference's results therefore depend on its driver.  The
// oracle and the HIP kernels both use the functions below, written from
// textbook methods, evaluated with the same operation sequence, so that the
// two agree bit for bit:
//   * sin/cos: Cody & Waite argument reduction by pi/2 with a 3-part
//     constant, then the Taylor series of sin (degree 11) and cos (degree 12)
//     on [-pi/4, pi/4] in Horner form.
//   * exp: reduction x = k ln2 + r (2-part ln2), Taylor series of e^r
//     (degree 9), scaling by 2^k through the exponent field.
//   * log: frexp reduction to m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//     s = (m-1)/(m+1), odd series to s^13.
//   * pow(x, y) = exp(y log x) for x > 0.
// tests/test_oracle_math.py bounds their error against the C library.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace vxo {

constexpr float BIG_F32 = 1e30f;
constexpr float EPS_INV = 1e-8f;
constexpr uint32_t INVALID_U32 = 0xFFFFFFFFu;

struct v2 { float x, y; };
struct v3 { float x, y, z; };
struct v4 { float x, y, z, w; };

static inline v2 V2(float x, float y) { return v2{x, y}; }
static inline v3 V3(float x, float y, float z) { return v3{x, y, z}; }
static inline v3 V3s(float s) { return v3{s, s, s}; }

static inline v3 operator+(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 operator-(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 operator*(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 operator/(v3 a, v3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 operator*(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 operator*(float s, v3 a) { return V3(s * a.x, s * a.y, s * a.z); }
static inline v3 operator/(v3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline v3 operator+(v3 a, float s) { return V3(a.x + s, a.y + s, a.z + s); }
static inline v3 operator-(v3 a, float s) { return V3(a.x - s, a.y - s, a.z - s); }
static inline v3 operator-(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline v2 operator+(v2 a, v2 b) { return V2(a.x + b.x, a.y + b.y); }
static inline v2 operator-(v2 a, v2 b) { return V2(a.x - b.x, a.y - b.y); }
static inline v2 operator*(v2 a, v2 b) { return V2(a.x * b.x, a.y * b.y); }
static inline v2 operator/(v2 a, v2 b) { return V2(a.x / b.x, a.y / b.y); }
static inline v2 operator*(v2 a, float s) { return V2(a.x * s, a.y * s); }
static inline v2 operator*(float s, v2 a) { return V2(s * a.x, s * a.y); }
static inline v2 operator+(v2 a, float s) { return V2(a.x + s, a.y + s); }
static inline v2 operator-(v2 a, float s) { return V2(a.x - s, a.y - s); }

// WGSL min/max/clamp (IEEE minNum/maxNum for non-NaN operands).
static inline float mn(float a, float b) { return std::fmin(a, b); }
static inline float mx(float a, float b) { return std::fmax(a, b); }
static inline float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); }
static inline v3 mn3(v3 a, v3 b) { return V3(mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)); }
static inline v3 mx3(v3 a, v3 b) { return V3(mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)); }
static inline v3 clamp3(v3 v, v3 lo, v3 hi) { return mn3(mx3(v, lo), hi); }
static inline v3 clamp3s(v3 v, float lo, float hi) { return clamp3(v, V3s(lo), V3s(hi)); }
static inline v3 abs3(v3 a) { return V3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
static inline v3 floor3(v3 a) { return V3(std::floor(a.x), std::floor(a.y), std::floor(a.z)); }
static inline v2 floor2(v2 a) { return V2(std::floor(a.x), std::floor(a.y)); }
static inline float fractf(float x) { return x - std::floor(x); }
static inline v2 fract2(v2 a) { return V2(fractf(a.x), fractf(a.y)); }

static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float dot2(v2 a, v2 b) { return a.x * b.x + a.y * b.y; }
static inline float length3(v3 a) { return std::sqrt(dot3(a, a)); }
static inline float length2(v2 a) { return std::sqrt(dot2(a, a)); }
static inline v3 normalize3(v3 a) {
  float inv = 1.0f / std::sqrt(dot3(a, a));
  return a * inv;
}
static inline v2 normalize2(v2 a) {
  float inv = 1.0f / std::sqrt(dot2(a, a));
  return a * inv;
}
static inline float inverse_sqrt(float x) { return 1.0f / std::sqrt(x); }
static inline v3 cross3(v3 a, v3 b) {
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// WGSL mix(e1, e2, e3) = e1 * (1 - e3) + e2 * e3.
static inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static inline v3 mix3(v3 a, v3 b, float t) { return V3(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)); }
static inline v3 mix3v(v3 a, v3 b, v3 t) { return V3(mixf(a.x, b.x, t.x), mixf(a.y, b.y, t.y), mixf(a.z, b.z, t.z)); }
// WGSL smoothstep(low, high, x).
static inline float smoothstepf(float lo, float hi, float x) {
  float t = clampf((x - lo) / (hi - lo), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
static inline float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// WGSL float->int conversions saturate (NaN -> 0).
static inline int32_t f2i(float x) {
  if (!(x == x)) return 0;
  if (x >= 2147483520.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000;
  return (int32_t)x;
}
static inline uint32_t f2u(float x) {
  if (!(x == x) || x <= 0.0f) return 0u;
  if (x >= 4294967040.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}
static inline uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float from_bits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// ---- transcendentals (see header comment) --------------------------------
static inline float ldexp_i(float x, int k) {
  // x * 2^k for |k| <= 252 via two exponent-field multiplies (exact unless the
  // result is subnormal or overflows).
  int k1 = k / 2, k2 = k - k1;
  return x * from_bits((uint32_t)(127 + k1) << 23) * from_bits((uint32_t)(127 + k2) << 23);
}

static inline float vx_exp(float x) {
  if (!(x == x)) return x;
  if (x > 88.72f) return INFINITY;
  if (x < -103.9f) return 0.0f;
  const float ln2_hi = 0.693145751953125f;         // 11 significant bits
  const float ln2_lo = 1.428606765330187045e-06f;  // ln2 - ln2_hi
  float kf = std::floor(x * 1.44269504088896341f + 0.5f);
  float r = (x - kf * ln2_hi) - kf * ln2_lo;
  // Taylor e^r, degree 9, Horner.
  float p = 2.7557319224e-06f;                // 1/9!
  p = p * r + 2.4801587302e-05f;              // 1/8!
  p = p * r + 1.9841269841e-04f;              // 1/7!
  p = p * r + 1.3888888889e-03f;              // 1/6!
  p = p * r + 8.3333333333e-03f;              // 1/5!
  p = p * r + 4.1666666667e-02f;              // 1/4!
  p = p * r + 1.6666666667e-01f;              // 1/3!
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return ldexp_i(p, (int)kf);
}

static inline float vx_log(float x) {
  if (!(x == x) || x < 0.0f) return NAN;
  if (x == 0.0f) return -INFINITY;
  if (x == INFINITY) return INFINITY;
  int e = 0;
  uint32_t u = bits(x);
  if ((u >> 23) == 0) {  // subnormal: scale up by 2^24
    x = x * 16777216.0f;
    u = bits(x);
    e = -24;
  }
  e += (int)((u >> 23) & 0xFFu) - 127;
  float m = from_bits((u & 0x007FFFFFu) | 0x3F800000u);  // [1, 2)
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  // log m = 2 atanh(s) = 2 (s + s^3/3 + s^5/5 + ...), s = (m-1)/(m+1).
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float q = 2.0f / 13.0f;
  q = q * s2 + 2.0f / 11.0f;
  q = q * s2 + 2.0f / 9.0f;
  q = q * s2 + 2.0f / 7.0f;
  q = q * s2 + 2.0f / 5.0f;
  q = q * s2 + 2.0f / 3.0f;
  float logm = 2.0f * s + s * s2 * q;
  const float ln2_hi = 0.693145751953125f;
  const float ln2_lo = 1.428606765330187045e-06f;
  float ef = (float)e;
  return (ef * ln2_hi + logm) + ef * ln2_lo;
}

static inline float vx_pow(float x, float y) {
  if (y == 0.0f) return 1.0f;
  if (!(x > 0.0f)) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
    return x == x ? 0.0f : x;  // negative base: treated as 0 (WGSL: undefined)
  }
  if (y == 1.0f) return x;
  return vx_exp(y * vx_log(x));
}

// Reduces x to r in [-pi/4, pi/4], returns the quadrant.
static inline int vx_trig_reduce(float x, float* r) {
  const float two_over_pi = 0.636619772367581343f;
  const float pio2_1 = 1.5703125f;                 // 8 bits
  const float pio2_2 = 4.837512969970703125e-04f;  // next bits
  const float pio2_3 = 7.549789948768648e-08f;     // remainder
  float kf = std::floor(x * two_over_pi + 0.5f);
  *r = ((x - kf * pio2_1) - kf * pio2_2) - kf * pio2_3;
  return (int)(((int64_t)kf) & 3);
}
static inline float vx_sin_poly(float r) {
  float r2 = r * r;
  float p = -2.5052108385e-08f;      // -1/11!
  p = p * r2 + 2.7557319224e-06f;    // 1/9!
  p = p * r2 - 1.9841269841e-04f;    // -1/7!
  p = p * r2 + 8.3333333333e-03f;    // 1/5!
  p = p * r2 - 1.6666666667e-01f;    // -1/3!
  return r + r * r2 * p;
}
static inline float vx_cos_poly(float r) {
  float r2 = r * r;
  float p = 2.0876756988e-09f;       // 1/12!
  p = p * r2 - 2.7557319224e-07f;    // -1/10!
  p = p * r2 + 2.4801587302e-05f;    // 1/8!
  p = p * r2 - 1.3888888889e-03f;    // -1/6!
  p = p * r2 + 4.1666666667e-02f;    // 1/4!
  p = p * r2 - 0.5f;
  return 1.0f + r2 * p;
}
static inline float vx_sin(float x) {
  if (!(x == x) || std::fabs(x) == INFINITY) return NAN;
  float r;
  int q = vx_trig_reduce(x, &r);
  switch (q) {
    case 0: return vx_sin_poly(r);
    case 1: return vx_cos_poly(r);
    case 2: return -vx_sin_poly(r);
    default: return -vx_cos_poly(r);
  }
}
static inline float vx_cos(float x) {
  if (!(x == x) || std::fabs(x) == INFINITY) return NAN;
  float r;
  int q = vx_trig_reduce(x, &r);
  switch (q) {
    case 0: return vx_cos_poly(r);
    case 1: return -vx_sin_poly(r);
    case 2: return -vx_cos_poly(r);
    default: return vx_sin_poly(r);
  }
}

}  // namespace vxo
