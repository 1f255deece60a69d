// vxd_math.h — device scalar/vector math for the gfx950 passes.
//
// Floating point contract (all kernels are built with -ffp-contract=off and
// correctly rounded fp32 divide/sqrt): every WGSL expression is evaluated as
// IEEE binary32 operations in the order the reference writes them, so the
// passes are reproducible operation for operation.
//
// Transcendentals are this library's own, from textbook methods:
//   sin/cos  Cody-Waite reduction by pi/2 (3-part constant), Taylor series on
//            [-pi/4, pi/4] (sin to x^11, cos to x^12), Horner form;
//   exp      x = k ln2 + r (2-part ln2), Taylor series of e^r to r^9,
//            2^k applied through the exponent field;
//   log      frexp reduction to [sqrt(1/2), sqrt(2)), 2 atanh(s) series to s^13;
//   pow      exp(y log x), x > 0.
// WGSL leaves their accuracy to the implementation; these stay within a few
// ULP of the C library on the ranges the passes use.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vxd {

#define VXD __device__ __forceinline__

constexpr float BIG_F32 = 1e30f;
constexpr float EPS_INV = 1e-8f;
constexpr uint32_t INVALID_U32 = 0xFFFFFFFFu;

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };
struct i3 { int32_t x, y, z; };

VXD f2 F2(float x, float y) { return f2{x, y}; }
VXD f3 F3(float x, float y, float z) { return f3{x, y, z}; }
VXD f3 F3s(float s) { return f3{s, s, s}; }

VXD f3 operator+(f3 a, f3 b) { return F3(a.x + b.x, a.y + b.y, a.z + b.z); }
VXD f3 operator-(f3 a, f3 b) { return F3(a.x - b.x, a.y - b.y, a.z - b.z); }
VXD f3 operator*(f3 a, f3 b) { return F3(a.x * b.x, a.y * b.y, a.z * b.z); }
VXD f3 operator/(f3 a, f3 b) { return F3(a.x / b.x, a.y / b.y, a.z / b.z); }
VXD f3 operator*(f3 a, float s) { return F3(a.x * s, a.y * s, a.z * s); }
VXD f3 operator*(float s, f3 a) { return F3(s * a.x, s * a.y, s * a.z); }
VXD f3 operator/(f3 a, float s) { return F3(a.x / s, a.y / s, a.z / s); }
VXD f3 operator+(f3 a, float s) { return F3(a.x + s, a.y + s, a.z + s); }
VXD f3 operator-(f3 a, float s) { return F3(a.x - s, a.y - s, a.z - s); }
VXD f3 operator-(f3 a) { return F3(-a.x, -a.y, -a.z); }
VXD f2 operator+(f2 a, f2 b) { return F2(a.x + b.x, a.y + b.y); }
VXD f2 operator-(f2 a, f2 b) { return F2(a.x - b.x, a.y - b.y); }
VXD f2 operator*(f2 a, f2 b) { return F2(a.x * b.x, a.y * b.y); }
VXD f2 operator/(f2 a, f2 b) { return F2(a.x / b.x, a.y / b.y); }
VXD f2 operator*(f2 a, float s) { return F2(a.x * s, a.y * s); }
VXD f2 operator*(float s, f2 a) { return F2(s * a.x, s * a.y); }
VXD f2 operator+(f2 a, float s) { return F2(a.x + s, a.y + s); }
VXD f2 operator-(f2 a, float s) { return F2(a.x - s, a.y - s); }

VXD float mn(float a, float b) { return fminf(a, b); }
VXD float mx(float a, float b) { return fmaxf(a, b); }
VXD float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); }
VXD f3 mn3(f3 a, f3 b) { return F3(mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)); }
VXD f3 mx3(f3 a, f3 b) { return F3(mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)); }
VXD f3 clamp3s(f3 v, float lo, float hi) { return mn3(mx3(v, F3s(lo)), F3s(hi)); }
VXD f3 abs3(f3 a) { return F3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
VXD f3 floor3(f3 a) { return F3(floorf(a.x), floorf(a.y), floorf(a.z)); }
VXD f2 floor2(f2 a) { return F2(floorf(a.x), floorf(a.y)); }
VXD float fractf(float x) { return x - floorf(x); }
VXD f2 fract2(f2 a) { return F2(fractf(a.x), fractf(a.y)); }
VXD float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
VXD float dot2(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
VXD float length3(f3 a) { return sqrtf(dot3(a, a)); }
VXD float length2(f2 a) { return sqrtf(dot2(a, a)); }
VXD f3 normalize3(f3 a) { return a * (1.0f / sqrtf(dot3(a, a))); }
VXD f2 normalize2(f2 a) { return a * (1.0f / sqrtf(dot2(a, a))); }
VXD float inverse_sqrt(float x) { return 1.0f / sqrtf(x); }
VXD f3 cross3(f3 a, f3 b) { return F3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
VXD float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
VXD f3 mix3(f3 a, f3 b, float t) { return F3(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)); }
VXD float smoothstepf(float lo, float hi, float x) {
  float t = clampf((x - lo) / (hi - lo), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
VXD float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
VXD int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Saturating float->int (WGSL conversion semantics; NaN -> 0): values at or
// above 2147483520 (the largest float below 2^31) give INT32_MAX, values at
// or below -2^31 give INT32_MIN, others truncate.  Written as clamp, convert
// and selects: the early-return form compiles to nested branches.
VXD int32_t f2i(float x) {
  const float xc = fminf(fmaxf(x, -2147483648.0f), 2147483520.0f);  // in range (or NaN, replaced below)
  int32_t r = (int32_t)xc;
  r = x >= 2147483520.0f ? 2147483647 : r;
  return x == x ? r : 0;
}
// Saturating float->uint: NaN and values <= 0 give 0, values at or above
// 4294967040 (the largest float below 2^32) give UINT32_MAX, others truncate.
VXD uint32_t f2u(float x) {
  const float xc = fminf(fmaxf(x, 0.0f), 4294967040.0f);
  uint32_t r = (uint32_t)xc;
  r = x >= 4294967040.0f ? 0xFFFFFFFFu : r;
  return (x == x && x > 0.0f) ? r : 0u;
}
VXD uint32_t fbits(float f) { return __float_as_uint(f); }
VXD float from_bits(uint32_t u) { return __uint_as_float(u); }

// ---- transcendentals ------------------------------------------------------
VXD float ldexp_i(float x, int k) {
  int k1 = k / 2, k2 = k - k1;
  return x * from_bits((uint32_t)(127 + k1) << 23) * from_bits((uint32_t)(127 + k2) << 23);
}
// Special cases (NaN, overflow, underflow) are selected after the main path,
// which runs on x clamped into range: the same value for every x in range,
// without a branch around the polynomial.
VXD float vx_exp(float x_in) {
  const float x = fminf(fmaxf(x_in, -103.9f), 88.72f);
  const float ln2_hi = 0.693145751953125f;
  const float ln2_lo = 1.428606765330187045e-06f;
  float kf = floorf(x * 1.44269504088896341f + 0.5f);
  float r = (x - kf * ln2_hi) - kf * ln2_lo;
  float p = 2.7557319224e-06f;
  p = p * r + 2.4801587302e-05f;
  p = p * r + 1.9841269841e-04f;
  p = p * r + 1.3888888889e-03f;
  p = p * r + 8.3333333333e-03f;
  p = p * r + 4.1666666667e-02f;
  p = p * r + 1.6666666667e-01f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  float v = ldexp_i(p, (int)kf);
  v = x_in > 88.72f ? __builtin_inff() : v;
  v = x_in < -103.9f ? 0.0f : v;
  return x_in == x_in ? v : x_in;
}
// Special cases (NaN, negative, zero, infinity) are selected at the end; the
// subnormal rescale and the mantissa fold are selects too.
VXD float vx_log(float x) {
  const bool sub = (fbits(x) >> 23) == 0u;  // +0 or subnormal
  const uint32_t u = fbits(sub ? x * 16777216.0f : x);
  int e = (sub ? -24 : 0) + (int)((u >> 23) & 0xFFu) - 127;
  float m = from_bits((u & 0x007FFFFFu) | 0x3F800000u);
  const bool fold = m > 1.41421356f;
  m = fold ? m * 0.5f : m;
  e += fold ? 1 : 0;
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
  float v = (ef * ln2_hi + logm) + ef * ln2_lo;
  v = x == __builtin_inff() ? x : v;
  v = x == 0.0f ? -__builtin_inff() : v;
  return (!(x == x) || x < 0.0f) ? __builtin_nanf("") : v;
}
VXD float vx_pow(float x, float y) {
  if (y == 0.0f) return 1.0f;
  if (!(x > 0.0f)) {
    if (x == 0.0f) return y > 0.0f ? 0.0f : __builtin_inff();
    return x == x ? 0.0f : x;
  }
  if (y == 1.0f) return x;
  return vx_exp(y * vx_log(x));
}
VXD int vx_trig_reduce(float x, float* r) {
  const float two_over_pi = 0.636619772367581343f;
  const float pio2_1 = 1.5703125f;
  const float pio2_2 = 4.837512969970703125e-04f;
  const float pio2_3 = 7.549789948768648e-08f;
  float kf = floorf(x * two_over_pi + 0.5f);
  *r = ((x - kf * pio2_1) - kf * pio2_2) - kf * pio2_3;
  // kf mod 4 as two's complement.  |kf| >= 2^31 is a multiple of 256, so its
  // low bits are 0 (the value a 64-bit conversion gives too); below that the
  // 32-bit conversion is exact.  Avoids an emulated float -> int64 convert.
  return (fabsf(kf) < 2147483648.0f ? (int)kf : 0) & 3;
}
VXD float vx_sin_poly(float r) {
  float r2 = r * r;
  float p = -2.5052108385e-08f;
  p = p * r2 + 2.7557319224e-06f;
  p = p * r2 - 1.9841269841e-04f;
  p = p * r2 + 8.3333333333e-03f;
  p = p * r2 - 1.6666666667e-01f;
  return r + r * r2 * p;
}
VXD float vx_cos_poly(float r) {
  float r2 = r * r;
  float p = 2.0876756988e-09f;
  p = p * r2 - 2.7557319224e-07f;
  p = p * r2 + 2.4801587302e-05f;
  p = p * r2 - 1.3888888889e-03f;
  p = p * r2 + 4.1666666667e-02f;
  p = p * r2 - 0.5f;
  return 1.0f + r2 * p;
}
// NaN and infinities give NaN (selected after the main path).
VXD float vx_sin(float x) {
  float r;
  int q = vx_trig_reduce(x, &r);
  float s = vx_sin_poly(r), c = vx_cos_poly(r);
  float v = (q & 1) ? c : s;
  v = (q & 2) ? -v : v;
  return (x == x && fabsf(x) != __builtin_inff()) ? v : __builtin_nanf("");
}
VXD float vx_cos(float x) {
  float r;
  int q = vx_trig_reduce(x, &r);
  float s = vx_sin_poly(r), c = vx_cos_poly(r);
  float v = (q & 1) ? s : c;
  v = ((q + 1) & 2) ? -v : v;
  return (x == x && fabsf(x) != __builtin_inff()) ? v : __builtin_nanf("");
}

// ---- wavefront primitives ----
// The primary kernel's workgroup is one 64-lane wavefront.  Cross-lane reads
// take their source from active lanes only.
VXD uint64_t wave_exec() { return __builtin_amdgcn_read_exec(); }
VXD uint32_t wave_lane() { return __lane_id(); }
VXD uint64_t wave_ballot(bool p) { return __ballot(p ? 1 : 0); }
VXD bool wave_any(bool p) { return wave_ballot(p) != 0ull; }
VXD uint32_t wave_read(uint32_t v, uint32_t src) { return (uint32_t)__shfl((int)v, (int)src, 64); }
VXD float wave_read(float v, uint32_t src) { return __shfl(v, (int)src, 64); }
// Lane `dst` receives v (each active lane sends to one lane).
VXD uint32_t wave_send(uint32_t dst, uint32_t v) { return (uint32_t)__builtin_amdgcn_ds_permute((int)(dst << 2), (int)v); }
VXD uint32_t popc64(uint64_t x) { return (uint32_t)__popcll(x); }
// The largest v < 32 over the active lanes (ballots only, so any exec mask).
VXD uint32_t wave_max_small(uint32_t v) {
  uint32_t lo = 0u;
  for (uint32_t step = 16u; step > 0u; step >>= 1u)
    if (wave_any(v >= lo + step)) lo += step;
  return lo;
}

}  // namespace vxd
