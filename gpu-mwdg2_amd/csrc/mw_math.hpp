// Per-element arithmetic of the adaptive MWDG2/HWFV1 step, shared by the HIP
// kernels and the host set-up code. Every routine evaluates its expression in
// the order the reference writes it, so that with -ffp-contract=off each value
// is reproducible bit for bit on host and device.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MW_HD __host__ __device__ __forceinline__

namespace mw {

enum SolverType { HWFV1 = 1, MWDG2 = 2 };                                  // types/SolverTypes.h:3-7
enum InletType { CLOSED = 1, FREE = 2, HFIX = 4, HVAR = 8, QFIX = 16, QVAR = 32 };  // types/InletTypes.h:3-11
enum Direction { NORTH = 1, EAST = 2, SOUTH = 4, WEST = 8 };               // types/Directions.h:3-9

constexpr uint8_t LEAF_INTERNAL = 0xFF;   // node lies strictly inside the tree

// Start of level l in a hierarchy array, with the reference's 3-entry pad
// (utilities/get_lvl_idx.cuh:11-24); level starts are multiples of 4 for l >= 1.
MW_HD int lvl_idx(int l) { return ((1 << (2 * l)) - 1) / 3 + 3; }

// Level of a hierarchy index: 3(idx-3)+1 lies in [4^l, 4^(l+1)).
MW_HD int level_of(int idx) {
  const unsigned v = 3u * (unsigned)(idx - 3) + 1u;
  return (31 - __builtin_clz(v)) >> 1;
}

// Bit interleave and its inverse (zorder/dilate.cuh:25-41, compact.cuh:9-21).
MW_HD int dilate(int c) {
  c &= 0x0000ffff;
  c = (c ^ (c << 8)) & 0x00ff00ff;
  c = (c ^ (c << 4)) & 0x0f0f0f0f;
  c = (c ^ (c << 2)) & 0x33333333;
  c = (c ^ (c << 1)) & 0x55555555;
  return c;
}
MW_HD int compact(int code) {
  int c = code & 0x55555555;
  c = (c ^ (c >> 1)) & 0x33333333;
  c = (c ^ (c >> 2)) & 0x0f0f0f0f;
  c = (c ^ (c >> 4)) & 0x00ff00ff;
  c = (c ^ (c >> 8)) & 0x0000ffff;
  return c;
}
MW_HD int morton(int x, int y) { return dilate(x) | (dilate(y) << 1); }

template <typename real> MW_HD real rmax(real a, real b);
template <typename real> MW_HD real rmin(real a, real b);
template <> MW_HD float rmax<float>(float a, float b) { return fmaxf(a, b); }
template <> MW_HD double rmax<double>(double a, double b) { return fmax(a, b); }
template <> MW_HD float rmin<float>(float a, float b) { return fminf(a, b); }
template <> MW_HD double rmin<double>(double a, double b) { return fmin(a, b); }
template <typename real> MW_HD real rabs(real a) { return a < real(0) ? -a : (a == real(0) ? real(0) : a); }
template <> MW_HD float rabs<float>(float a) { return fabsf(a); }
template <> MW_HD double rabs<double>(double a) { return fabs(a); }
template <typename real> MW_HD real rsqrt_cr(real a);
template <> MW_HD float rsqrt_cr<float>(float a) { return sqrtf(a); }
template <> MW_HD double rsqrt_cr<double>(double a) { return sqrt(a); }

// Cube root in place of pow(x, 1/3) for the h^(1/3) of the Manning term
// (operators/apply_friction.cuh:41) and the FREE ghost depth
// (neighbours/add_ghost_cells.cu:135), without divisions: Newton steps
// r <- r (4 - x r^3) / 3 for r = x^(-1/3) from the seed whose bits are
// (4/3) * exponent bias - bits(x) / 3, then y = x r^2 and one step
// y <- y - (y^3 - x) r^2 / 3. Subnormal x is scaled by 2^54 (2^24) first.
// Host and device give identical bits; within 4 ulp of the C library's cbrt
// in FP64 and 2 ulp of the exact value in FP32.
template <typename real> MW_HD real cbrt_newton(real x);
template <> MW_HD double cbrt_newton<double>(double x) {
  if (!(x > 0.0)) return (x == 0.0) ? 0.0 : __builtin_nan("");
  if (__builtin_isinf(x)) return x;
  uint64_t b = __builtin_bit_cast(uint64_t, x);
  const bool sub = (b >> 52) == 0;
  const double xs = sub ? x * 18014398509481984.0 : x;
  b = __builtin_bit_cast(uint64_t, xs);
  double r = __builtin_bit_cast(double, 0x5540000000000000ull - b / 3);
  for (int it = 0; it < 4; ++it) r = r * (4.0 - xs * r * r * r) * (1.0 / 3.0);
  double y = xs * r * r;
  y = y - (y * y * y - xs) * (r * r) * (1.0 / 3.0);
  return sub ? y * (1.0 / 262144.0) : y;
}
template <> MW_HD float cbrt_newton<float>(float x) {
  if (!(x > 0.0f)) return (x == 0.0f) ? 0.0f : __builtin_nanf("");
  if (__builtin_isinf(x)) return x;
  uint32_t b = __builtin_bit_cast(uint32_t, x);
  const bool sub = (b >> 23) == 0;
  const float xs = sub ? x * 16777216.0f : x;
  b = __builtin_bit_cast(uint32_t, xs);
  float r = __builtin_bit_cast(float, 0x54AAAAABu - b / 3);
  for (int it = 0; it < 3; ++it) r = r * (4.0f - xs * r * r * r) * (1.0f / 3.0f);
  float y = xs * r * r;
  y = y - (y * y * y - xs) * (r * r) * (1.0f / 3.0f);
  return sub ? y * (1.0f / 256.0f) : y;
}

// e^x by Cody-Waite range reduction, x = n ln2 + r with |r| <= ln2 / 2 and ln2
// split into a head with trailing zero bits (so n * head is exact) and a tail
// (W. J. Cody and W. Waite, Software Manual for the Elementary Functions,
// 1980), then the Taylor series of e^r to degree 13 (FP64) or 8 (FP32) in
// nested form 1 + r (1 + r/2 (1 + r/3 (...))), scaled by 2^n through exponent
// bits. Used for the Oregon wavemaker pulse in place of exp, so host and device
// give identical bits; within 2 ulp of the C library's exp.
template <typename real> MW_HD real exp_cw(real x);
template <> MW_HD double exp_cw<double>(double x) {
  if (x != x) return x;
  if (x > 709.8) return __builtin_inf();
  if (x < -745.2) return 0.0;
  const double n = __builtin_rint(x * 1.4426950408889634);
  const double r = (x - n * 0.693147180369123816490) - n * 1.90821492927058770002e-10;
  double p = 1.0;
  for (int k = 13; k >= 1; --k) p = 1.0 + r * p / (double)k;
  const int e = (int)n;
  if (e > 1023) return (p * __builtin_bit_cast(double, (uint64_t)(e - 1 + 1023) << 52)) * 2.0;
  if (e >= -1022) return p * __builtin_bit_cast(double, (uint64_t)(e + 1023) << 52);
  return (p * __builtin_bit_cast(double, (uint64_t)(e + 54 + 1023) << 52)) * 5.551115123125783e-17;   // 2^-54
}
template <> MW_HD float exp_cw<float>(float x) {
  if (x != x) return x;
  if (x > 88.73f) return __builtin_inff();
  if (x < -103.98f) return 0.0f;
  const float n = __builtin_rintf(x * 1.44269504f);
  const float r = (x - n * 0.693145751953125f) - n * 1.428606765330187045e-06f;
  float p = 1.0f;
  for (int k = 8; k >= 1; --k) p = 1.0f + r * p / (float)k;
  const int e = (int)n;
  if (e > 127) return (p * __builtin_bit_cast(float, (uint32_t)(e - 1 + 127) << 23)) * 2.0f;
  if (e >= -126) return p * __builtin_bit_cast(float, (uint32_t)(e + 127) << 23);
  return (p * __builtin_bit_cast(float, (uint32_t)(e + 25 + 127) << 23)) * 2.98023223876953125e-08f;   // 2^-25
}

// Built-in synthetic cases: bed elevation and depth at a node (x, y)
// (synthetic/topo*.cuh, synthetic/h_init*.cuh); hl, hr are the case's 1D
// depths (classes/Depths1D.h). The c-property depth is the surface level
// minus the bed, negative above it, as the reference returns.
template <typename real>
MW_HD real c_prop_bed_1d(real a) {                              // topo_c_prop_1D.cuh
  real z = real(0.0);
  if (a >= real(22) && a < real(25)) z = real(0.05) * a - real(1.1);
  else if (a >= real(25) && a <= real(28)) z = real(-0.05) * a + real(1.4);
  else if (a > real(8) && a < real(12)) z = real(0.2) - real(0.05) * (a - real(10)) * (a - real(10));
  else if (a > real(39) && a < real(46.5)) z = real(0.3);
  return real(10.0) * z;
}
template <typename real>
MW_HD real triangle_bed(real a) {                               // topo_triangle.cuh
  const real slope = real(0.4) / real(3.0);
  if (a > real(25.5) && a <= real(28.5)) return (a - real(25.5)) * slope;
  if (a > real(28.5) && a <= real(31.5)) return real(0.4) - (a - real(28.5)) * slope;
  return real(0.0);
}
template <typename real>
MW_HD real node_dist(real x, real y, real x0, real y0) {
  return rsqrt_cr<real>((x - x0) * (x - x0) + (y - y0) * (y - y0));
}
template <typename real>
MW_HD real synthetic_bed(real x, real y, int tc) {
  switch (tc) {
    case 1: case 3: case 11: case 13: return c_prop_bed_1d<real>(x);
    case 2: case 4: case 12: case 14: return c_prop_bed_1d<real>(y);
    case 15: return triangle_bed<real>(x);
    case 16: return triangle_bed<real>(y);
    case 17: return real(0.01) * x * x;                         // topo_parabolic_bowl.cuh
    case 18: return real(0.01) * y * y;
    case 19: case 20: {                                         // topo_three_cones.cuh
      const real z1 = (real(8) - node_dist<real>(x, y, real(30), real(6))) / real(8);
      const real z2 = (real(8) - node_dist<real>(x, y, real(30), real(24))) / real(8);
      const real z3 = real(0.3) * (real(10) - node_dist<real>(x, y, real(47.5), real(15)));
      return rmax<real>(rmax<real>(rmax<real>(z1, z2), z3), real(0.0));
    }
    case 21: {                                                  // topo_diff.cuh
      real z = real(0);
      z = rmax<real>(z, real(1.0) - real(0.2) * node_dist<real>(x, y, real(20.0), real(15.0)));
      z = rmax<real>(z, real(2.0) - real(0.5) * node_dist<real>(x, y, real(40.0), real(15.0)));
      return rmax<real>(z, real(3.0) - real(0.3) * node_dist<real>(x, y, real(60.0), real(15.0)));
    }
    case 22: {                                                  // topo_non_diff.cuh
      const bool yb = real(11.0) <= y && y <= real(19.0);
      if (real(16.0) <= x && x <= real(24.0) && yb) return real(0.86);
      if (real(36.0) <= x && x <= real(44.0) && yb) return real(1.78);
      if (real(56.0) <= x && x <= real(64.0) && yb) return real(2.30);
      return real(0);
    }
    default: return real(0.0);
  }
}
template <typename real>
MW_HD real synthetic_depth(real x, real y, real z, real hl, real hr, int tc) {
  switch (tc) {
    case 1: case 2: case 3: case 4: return hl - z;
    case 5: case 7: case 9: case 11: case 13:                  // h_init_overtopping.cuh
      return (x <= real(25)) ? rmax<real>(real(0.0), hl - z) : rmax<real>(real(0.0), hr - z);
    case 6: case 8: case 10: case 12: case 14:
      return (y <= real(25)) ? rmax<real>(real(0.0), hl - z) : rmax<real>(real(0.0), hr - z);
    case 15: return (x < real(15.5)) ? real(0.75) : real(0.0);
    case 16: return (y < real(15.5)) ? real(0.75) : real(0.0);
    case 17: return rmax<real>(real(0.0), (real(0.2) * x + real(10.0)) - z);
    case 18: return rmax<real>(real(0.0), (real(0.2) * y + real(10.0)) - z);
    case 19: return real(1.0) - z;
    case 20: return (x < real(16.0)) ? rmax<real>(real(1.875) - z, real(0.0)) : real(0.0);
    case 21: return real(1.95) - z;
    case 22: return real(1.78) - z;
    case 23: return (rsqrt_cr<real>(x * x + y * y) < real(2.5)) ? real(2.5) : real(0.5);
    default: return real(0.0);
  }
}

// ---------------------------------------------------------------------------
// Multiwavelet filter bank. The analysis map of one variable is a 12x12
// orthogonal matrix acting on the 4 children's (0, 1x, 1y) coefficients; its
// entries are the decimal literals of the reference (mra/Filters.h:202-397).
// FB[family][k][i][j]: family 0..3 = HH (scale), GA, GB, GC (details alpha,
// beta, gamma); filter k acts on child PERM[k]; row i is the parent mode,
// column j the child mode. Both the double literal and the float literal are
// kept so the FP32 build uses exactly the float constants.
// ---------------------------------------------------------------------------
struct FLit { double d; float f; };
#define MW_FL(x) FLit{x, x##f}
template <typename real> MW_HD real flit(const FLit& l);
template <> MW_HD double flit<double>(const FLit& l) { return l.d; }
template <> MW_HD float flit<float>(const FLit& l) { return l.f; }

constexpr int PERM[4] = {0, 2, 1, 3};

#define Z0 MW_FL(0.0)
#define P43 MW_FL(0.4330127018922193)
#define N43 MW_FL(-0.4330127018922193)
#define P26 MW_FL(0.2672612419124244)
#define N26 MW_FL(-0.2672612419124244)
#define N46 MW_FL(-0.4629100498862757)
#define P46 MW_FL(0.46291004988627577)
#define P70 MW_FL(0.7071067811865476)
#define N70 MW_FL(-0.7071067811865476)
#define N16 MW_FL(-0.1636634176769942)
#define P16 MW_FL(0.16366341767699427)
#define N28 MW_FL(-0.2834733547569204)
#define P37 MW_FL(0.37796447300922725)
#define N37 MW_FL(-0.3779644730092272)
constexpr FLit FB[4][4][3][3] = {
  {  // HH
    {{MW_FL(0.5), Z0, Z0}, {N43, MW_FL(0.25), Z0}, {N43, Z0, MW_FL(0.25)}},
    {{MW_FL(0.5), Z0, Z0}, {N43, MW_FL(0.25), Z0}, {P43, Z0, MW_FL(0.25)}},
    {{MW_FL(0.5), Z0, Z0}, {P43, MW_FL(0.25), Z0}, {N43, Z0, MW_FL(0.25)}},
    {{MW_FL(0.5), Z0, Z0}, {P43, MW_FL(0.25), Z0}, {P43, Z0, MW_FL(0.25)}}},
  {  // GA
    {{N26, N46, N46}, {Z0, Z0, N70}, {Z0, N70, Z0}},
    {{Z0, Z0, Z0}, {Z0, Z0, Z0}, {Z0, Z0, Z0}},
    {{Z0, Z0, Z0}, {Z0, Z0, Z0}, {Z0, Z0, Z0}},
    {{P26, N46, N46}, {Z0, Z0, P70}, {Z0, P70, Z0}}},
  {  // GB
    {{Z0, Z0, Z0}, {Z0, Z0, Z0}, {Z0, Z0, Z0}},
    {{N26, N46, P46}, {Z0, Z0, N70}, {Z0, N70, Z0}},
    {{P26, N46, P46}, {Z0, Z0, P70}, {Z0, P70, Z0}},
    {{Z0, Z0, Z0}, {Z0, Z0, Z0}, {Z0, Z0, Z0}}},
  {  // GC
    {{MW_FL(0.5), Z0, Z0}, {N16, N28, P37}, {N16, P37, N28}},
    {{MW_FL(-0.5), Z0, Z0}, {N16, N28, N37}, {P16, N37, N28}},
    {{MW_FL(-0.5), Z0, Z0}, {P16, N28, N37}, {N16, N37, N28}},
    {{MW_FL(0.5), Z0, Z0}, {P16, N28, P37}, {P16, P37, N28}}}};
#undef Z0
#undef P43
#undef N43
#undef P26
#undef N26
#undef N46
#undef P46
#undef P70
#undef N70
#undef N16
#undef P16
#undef N28
#undef P37
#undef N37
constexpr FLit HAAR = MW_FL(0.7071067811865475);   // H0 = H1 = G0 = -G1

// One MW analysis row: sum over (k, j) of FB[fam][k][i][j] * u[PERM[k]][j],
// left to right, then / 2 (mra/encode_scale.cuh:13-38,
// mra/encode_details.cuh:159-239). Terms whose literal is 0 are skipped: they
// add a signed zero, which leaves every non-zero sum unchanged.
template <typename real, int FAM, int I>
MW_HD real enc_row(const real (&u)[4][3]) {
  real acc = real(0);
  bool have = false;
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (FB[FAM][k][I][j].d == 0.0) continue;
      const real t = flit<real>(FB[FAM][k][I][j]) * u[PERM[k]][j];
      acc = have ? acc + t : t;
      have = true;
    }
  return acc / real(2.0);
}

// u[c][m]: child c (Morton 0..3), mode m (0, 1x, 1y) -> s[m], d[family][m]
template <typename real>
MW_HD void encode_mw(const real (&u)[4][3], real (&s)[3], real (&d)[3][3]) {
  s[0] = enc_row<real, 0, 0>(u); s[1] = enc_row<real, 0, 1>(u); s[2] = enc_row<real, 0, 2>(u);
  d[0][0] = enc_row<real, 1, 0>(u); d[0][1] = enc_row<real, 1, 1>(u); d[0][2] = enc_row<real, 1, 2>(u);
  d[1][0] = enc_row<real, 2, 0>(u); d[1][1] = enc_row<real, 2, 1>(u); d[1][2] = enc_row<real, 2, 2>(u);
  d[2][0] = enc_row<real, 3, 0>(u); d[2][1] = enc_row<real, 3, 1>(u); d[2][2] = enc_row<real, 3, 2>(u);
}
template <typename real>
MW_HD void encode_mw_scale(const real (&u)[4][3], real (&s)[3]) {
  s[0] = enc_row<real, 0, 0>(u); s[1] = enc_row<real, 0, 1>(u); s[2] = enc_row<real, 0, 2>(u);
}

// Synthesis: child c, mode j = (sum_i HH*s_i + sum GA*a_i + GB*b_i + GC*g_i) * 2
// (mra/decode_scale_children.cuh:63-395).
template <typename real, int C, int J>
MW_HD real dec_entry(const real (&s)[3], const real (&d)[3][3]) {
  constexpr int K = PERM[C];
  real acc = real(0);
  bool have = false;
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (FB[f][K][i][J].d == 0.0) continue;
      const real v = (f == 0) ? s[i] : d[f - 1][i];
      const real t = flit<real>(FB[f][K][i][J]) * v;
      acc = have ? acc + t : t;
      have = true;
    }
  return acc * real(2.0);
}
template <typename real>
MW_HD void decode_mw(const real (&s)[3], const real (&d)[3][3], real (&u)[4][3]) {
#define MW_DEC_C(C) u[C][0] = dec_entry<real, C, 0>(s, d); u[C][1] = dec_entry<real, C, 1>(s, d); u[C][2] = dec_entry<real, C, 2>(s, d);
  MW_DEC_C(0) MW_DEC_C(1) MW_DEC_C(2) MW_DEC_C(3)
#undef MW_DEC_C
}

// Haar pair (mra/encode_scale.cuh:7-11, encode_details.cuh:7-24,
// decode_scale_children.cuh:38-60).
template <typename real>
MW_HD void encode_hw(const real (&c)[4], real& s, real& a, real& b, real& g) {
  const real H = flit<real>(HAAR), Gm = -flit<real>(HAAR);
  s = real(0.5) * (H * (H * c[0] + H * c[2]) + H * (H * c[1] + H * c[3]));
  a = real(0.5) * (H * (H * c[0] + Gm * c[2]) + H * (H * c[1] + Gm * c[3]));
  b = real(0.5) * (H * (H * c[0] + H * c[2]) + Gm * (H * c[1] + H * c[3]));
  g = real(0.5) * (H * (H * c[0] + Gm * c[2]) + Gm * (H * c[1] + Gm * c[3]));
}
template <typename real>
MW_HD void decode_hw(real u, real a, real b, real g, real (&c)[4]) {
  const real H = flit<real>(HAAR), Gm = -flit<real>(HAAR);
  c[0] = (H * (H * u + H * a) + H * (H * b + H * g)) * real(2.0);
  c[2] = (H * (H * u + Gm * a) + H * (H * b + Gm * g)) * real(2.0);
  c[1] = (H * (H * u + H * a) + Gm * (H * b + H * g)) * real(2.0);
  c[3] = (H * (H * u + Gm * a) + Gm * (H * b + Gm * g)) * real(2.0);
}

// max |detail| of one variable (classes/SubDetail.h:18-45)
template <typename real> MW_HD real submax(real a, real b, real g) {
  const real m = rmax<real>(rabs<real>(a), rabs<real>(b));
  return rmax<real>(rabs<real>(g), m);
}
template <typename real> MW_HD real detmax_mw(const real (&d)[3][3]) {
  const real m0 = submax<real>(d[0][0], d[1][0], d[2][0]);
  const real m1 = submax<real>(d[0][1], d[1][1], d[2][1]);
  const real m2 = submax<real>(d[0][2], d[1][2], d[2][2]);
  const real m = rmax<real>(m0, m1);
  return rmax<real>(m2, m);
}

// ---------------------------------------------------------------------------
// Shallow-water pieces (classes/FlowVector.h, operators/flux_HLL.cuh,
// operators/apply_friction.cuh, operators/get_bed_src.cuh).
// ---------------------------------------------------------------------------
template <typename real> struct FV { real h, qx, qy; };
template <typename real> MW_HD FV<real> fv_add(FV<real> a, FV<real> b) { return {a.h + b.h, a.qx + b.qx, a.qy + b.qy}; }
template <typename real> MW_HD FV<real> fv_sub(FV<real> a, FV<real> b) { return {a.h - b.h, a.qx - b.qx, a.qy - b.qy}; }
template <typename real> MW_HD FV<real> fv_scale(real s, FV<real> a) { return {s * a.h, s * a.qx, s * a.qy}; }
template <typename real> MW_HD FV<real> fv_div(FV<real> a, real s) { return {a.h / s, a.qx / s, a.qy / s}; }

template <typename real>
MW_HD FV<real> get_star(FV<real> U, real z, real zi, real tol_h) {        // FlowVector.h:23-48
  const real eta = U.h + z;
  const real hs = rmax<real>(real(0.0), eta - zi);
  const real ux = (U.h < tol_h) ? real(0) : U.qx / U.h;
  const real uy = (U.h < tol_h) ? real(0) : U.qy / U.h;
  return {hs, hs * ux, hs * uy};
}
template <typename real>
MW_HD FV<real> phys_flux_x(FV<real> U, real tol_h, real g) {             // FlowVector.h:50-65
  if (U.h < tol_h) return {real(0), real(0), real(0)};
  return {U.qx, U.qx * U.qx / U.h + g * U.h * U.h / real(2.0), U.qx * U.qy / U.h};
}
template <typename real>
MW_HD FV<real> phys_flux_y(FV<real> U, real tol_h, real g) {             // FlowVector.h:67-82
  if (U.h < tol_h) return {real(0), real(0), real(0)};
  return {U.qy, U.qx * U.qy / U.h, U.qy * U.qy / U.h + g * U.h * U.h / real(2.0)};
}
template <typename real>
MW_HD FV<real> flux_hll_x(FV<real> Un, FV<real> Up, real tol_h, real g) {  // flux_HLL.cuh:9-88
  if (Un.h < tol_h && Up.h < tol_h) return {real(0), real(0), real(0)};
  const real h_neg = (Un.h < tol_h) ? real(0.0) : Un.h;
  const real u_neg = (Un.h < tol_h) ? real(0.0) : Un.qx / Un.h;
  const real v_neg = (Un.h < tol_h) ? real(0.0) : Un.qy / Un.h;
  const real h_pos = (Up.h < tol_h) ? real(0.0) : Up.h;
  const real u_pos = (Up.h < tol_h) ? real(0.0) : Up.qx / Up.h;
  const real v_pos = (Up.h < tol_h) ? real(0.0) : Up.qy / Up.h;
  const real a_neg = rsqrt_cr<real>(g * h_neg);
  const real a_pos = rsqrt_cr<real>(g * h_pos);
  const real h_star = (real(0.5) * (a_neg + a_pos) + real(0.25) * (u_neg - u_pos)) *
                      (real(0.5) * (a_neg + a_pos) + real(0.25) * (u_neg - u_pos)) / g;
  const real u_star = real(0.5) * (u_neg + u_pos) + a_neg - a_pos;
  const real a_star = rsqrt_cr<real>(g * h_star);
  const real s_neg = (h_neg < tol_h) ? u_pos - real(2.0) * a_pos : rmin<real>(u_neg - a_neg, u_star - a_star);
  const real s_pos = (h_pos < tol_h) ? u_neg + real(2.0) * a_neg : rmax<real>(u_pos + a_pos, u_star + a_star);
  const FV<real> F_neg = {Un.qx, u_neg * Un.qx + real(0.5) * g * h_neg * h_neg, Un.qy * u_neg};
  const FV<real> F_pos = {Up.qx, u_pos * Up.qx + real(0.5) * g * h_pos * h_pos, Up.qy * u_pos};
  if (s_neg >= real(0.0)) return F_neg;
  if (s_neg < real(0.0) && s_pos >= real(0.0)) {
    FV<real> F;
    F.h = (s_pos * F_neg.h - s_neg * F_pos.h + s_neg * s_pos * (h_pos - h_neg)) / (s_pos - s_neg);
    F.qx = (s_pos * F_neg.qx - s_neg * F_pos.qx + s_neg * s_pos * (Up.qx - Un.qx)) / (s_pos - s_neg);
    const real s_mid = (s_neg * h_pos * (u_pos - s_pos) - s_pos * h_neg * (u_neg - s_neg)) /
                       (h_pos * (u_pos - s_pos) - h_neg * (u_neg - s_neg));
    F.qy = (s_neg < real(0.0) && s_mid >= real(0.0)) ? F.h * v_neg : F.h * v_pos;
    return F;
  }
  return F_pos;
}
template <typename real>
MW_HD FV<real> flux_hll_y(FV<real> Un, FV<real> Up, real tol_h, real g) {  // flux_HLL.cuh:90-111
  const FV<real> a = {Un.h, Un.qy, -Un.qx}, b = {Up.h, Up.qy, -Up.qx};
  const FV<real> F = flux_hll_x<real>(a, b, tol_h, g);
  return {F.h, -F.qy, F.qx};
}

// operators/apply_friction.cuh:7-53
template <typename real>
MW_HD void apply_friction(real h, real tol_h, real tol_s, real& qx, real& qy, real n, real g, real dt) {
  if (h < tol_h) { qx = real(0.0); qy = real(0.0); return; }
  const real ux = qx / h, uy = qy / h;
  if (rabs<real>(ux) < tol_s && rabs<real>(uy) < tol_s) { qx = real(0.0); qy = real(0.0); return; }
  const real Cf = g * n * n / cbrt_newton<real>(h);
  const real speed = rsqrt_cr<real>(ux * ux + uy * uy);
  const real Sf_x = -Cf * ux * speed;
  const real Sf_y = -Cf * uy * speed;
  const real D_x = real(1.0) + dt * Cf / h * (real(2.0) * ux * ux + uy * uy) / speed;
  const real D_y = real(1.0) + dt * Cf / h * (real(2.0) * uy * uy + ux * ux) / speed;
  qx += dt * Sf_x / D_x;
  qy += dt * Sf_y / D_y;
}

// Flow coefficients of one cell, h form: c[v][m], v = h,qx,qy; m = 0,1x,1y.
template <typename real> struct Cell { real c[3][3]; real z[3]; };

// Monai valley wavemaker (operators/non_reflective_wave.cuh:15-93): a Newton
// iteration (at most 101 steps, tolerance 1e-6 on the step and on the
// residual) locates the foot p of the outgoing characteristic inside the
// boundary cell, linearly between the still-water state (hb = 0.13535 m, at
// rest) outside and the cell state (hp, up); the ghost state then follows
// from the outgoing invariant u - 2c there and the incoming one built from
// the inlet surface elevation. pow(x, -0.5) is evaluated as 1 / sqrt(x).
template <typename real>
MW_HD void monai_wave(real h_inlet, real dt, real dxl, real hp, real up, real g, real& h_out, real& u_out) {
  const real hb = real(0.13535), ub = real(0.0);
  const real xa = -dxl / real(2.0), tol = real(1e-6);
  real p0 = dxl / real(2.0), p1 = p0 - dxl / real(2.0);
  real r = (p0 - xa) / dxl;
  real f0 = (p0 - xa) + dt * ((ub - r * (ub - up)) - rsqrt_cr<real>(g * (hb - r * (hb - hp))));
  int cond = 0;
  for (int k = 0; k <= 100 && cond == 0; ++k) {
    r = (p0 - xa) / dxl;
    const real df = real(1.0) + dt * ((up - ub) / dxl - real(0.5) * (real(1.0) / rsqrt_cr<real>(g * (hb - r * (hb - hp)))) *
                                                             (g * (hp - hb) / dxl));
    real dp;
    if (rabs<real>(df) <= real(0.0)) {
      cond = 1;
      dp = p1 - p0;
      p1 = p0;
    } else {
      dp = f0 / df;
      p1 = p0 - dp;
    }
    const real r1 = (p1 - xa) / dxl;
    const real f1 = (p1 - xa) + dt * ((ub - r1 * (ub - up)) - rsqrt_cr<real>(g * (hb - r1 * (hb - hp))));
    const real relerr = rabs<real>(dp) / (rabs<real>(p1) + real(0.0));
    if (relerr < tol) cond = 2;
    if (rabs<real>(f1) < tol) cond = 3;
    if (relerr <= tol && rabs<real>(f1) < tol) cond = 4;
    p0 = p1;
    f0 = f1;
  }
  const real hr = hb - (hb - hp) * (p1 - xa) / dxl;
  const real ur = ub - (ub - up) * (p1 - xa) / dxl;
  const real cr = rsqrt_cr<real>(g * hr);
  const real bm = ur - real(2.0) * cr;
  const real bp = real(2.0) * rsqrt_cr<real>(g * (h_inlet + real(0.13535))) + h_inlet * rsqrt_cr<real>(g / real(0.13535));
  h_out = real(1.0) / real(16.0) / g * (bp - bm) * (bp - bm);
  u_out = real(0.5) * (bp + bm);
}

// MWDG2 implicit Manning on one cell's mean and its 4 Gauss-like points
// (operators/friction_implicit.cu:191-315).
template <typename real>
MW_HD void friction_mw(Cell<real>& s, real tol_h, real tol_s, real n, real g, real dt) {
  const real h = s.c[0][0], h1x = s.c[0][1], h1y = s.c[0][2];
  real hlx = h - h1x, hux = h + h1x, hly = h - h1y, huy = h + h1y;
  real qxlx = s.c[1][0] - s.c[1][1], qxux = s.c[1][0] + s.c[1][1];
  real qxly = s.c[1][0] - s.c[1][2], qxuy = s.c[1][0] + s.c[1][2];
  real qylx = s.c[2][0] - s.c[2][1], qyux = s.c[2][0] + s.c[2][1];
  real qyly = s.c[2][0] - s.c[2][2], qyuy = s.c[2][0] + s.c[2][2];
  apply_friction<real>(h, tol_h, tol_s, s.c[1][0], s.c[2][0], n, g, dt);
  apply_friction<real>(hlx, tol_h, tol_s, qxlx, qylx, n, g, dt);
  apply_friction<real>(hux, tol_h, tol_s, qxux, qyux, n, g, dt);
  apply_friction<real>(hly, tol_h, tol_s, qxly, qyly, n, g, dt);
  apply_friction<real>(huy, tol_h, tol_s, qxuy, qyuy, n, g, dt);
  s.c[1][1] = real(0.5) * (qxux - qxlx);
  s.c[1][2] = real(0.5) * (qxuy - qxly);
  s.c[2][1] = real(0.5) * (qyux - qylx);
  s.c[2][2] = real(0.5) * (qyuy - qyly);
}

}  // namespace mw
