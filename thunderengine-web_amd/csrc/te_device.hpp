// Device-side building blocks of the CM1 large step on gfx950.
//
// Arithmetic contract (shared by every kernel of this library, and the reason
// the build flags are what they are):
//   * one IEEE binary32 operation per source operation: -ffp-contract=off, and
//     __builtin_fmaf exactly where the reference step fuses (its WGSL writes
//     fma(), e.g. data/kernels/acoustic_w_solve.wgsl:129-133);
//   * subnormal operands and results flush to zero with the sign kept
//     (-fgpu-flush-denormals-to-zero): the reference recording was made that way;
//   * division and sqrt are correctly rounded (hipcc's default for f32);
//   * min/max are explicit selects with the reference's tie rule.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace te {

// Per-dispatch argument block: thread grid and the word offset of each
// parameter inside the flat memory M (parameter n at p[n]).
struct KArg {
  uint32_t n[3];
  uint32_t p[48];  // parameters 0..31 as in the reference's slots; 32..47 for fused launches
  // launch shape, filled by the runtime: block dims (3), grid blocks in x and
  // y (2), m = ceil(2^40 / d) as (low, high) words for d = grid blocks in x
  // and in y (4; floor(l / d) = (l * m) >> 40 exactly for l < 2^24 and
  // d <= 2^16: Granlund & Montgomery, PLDI 1994, thm. 4.2), grid blocks (1),
  // flat launch (1: see te_flat), plane launch (1: see TE_TID3) with the
  // division constant of the plane's row length n0 (2)
  uint32_t lg[14];
  // the parameters as device pointers (the memory's base + p[n] words),
  // filled by the runtime so that kernels do no per-wave address arithmetic
  uint32_t* q[48];
};

#define TE_DEV __device__ __forceinline__

TE_DEV float ldf(const uint32_t* M, uint32_t o) { return __uint_as_float(M[o]); }
TE_DEV int32_t ldi(const uint32_t* M, uint32_t o) { return (int32_t)M[o]; }
TE_DEV float* fp(uint32_t* M, uint32_t o) { return reinterpret_cast<float*>(M + o); }
TE_DEV float* fq(uint32_t* p) { return reinterpret_cast<float*>(p); }

// A field addressed by 32-bit byte offsets from a wave-uniform base: loads and
// stores take the scalar-base form (one VGPR per offset, shared by every field
// of the same layout, instead of a 64-bit address per field).  The field must
// span less than 2^32 bytes from its base.
struct FieldB {
  char* b;
  TE_DEV explicit FieldB(const float* p) : b(reinterpret_cast<char*>(const_cast<float*>(p))) {}
  TE_DEV float operator()(uint32_t o) const { return *reinterpret_cast<const float*>(b + o); }
  TE_DEV void put(uint32_t o, float v) const { *reinterpret_cast<float*>(b + o) = v; }
  // the same field from a wave-uniform byte displacement (a level)
  TE_DEV FieldB operator+(uint64_t d) const { return FieldB(reinterpret_cast<const float*>(b + d)); }
};

// Materialise x here: the value is computed at this point of the program
// (its operands die) rather than sunk towards a later use.
TE_DEV void te_pin(float& x) { asm volatile("" : "+v"(x)); }

struct Dims {
  int nx, ny, nz, sJ, sK;
  float rdx, rdy, rdz;
  TE_DEV int64_t sk() const { return (int64_t)sJ * sK; }
  TE_DEV int64_t idx(int i, int j, int k) const { return (int64_t)(i + 3) + (int64_t)sJ * ((j + 3) + (int64_t)sK * k); }
};
// The periodic halo cells whose image is interior cell (i, j) of a level:
// hx in {i, i - nx (i >= nx - 3), i + nx (i <= 3)}, hy likewise with ny, all
// pairs but (i, j) itself -- the cells bc_periodic_x then bc_periodic_y
// (bc_periodic_x.wgsl:111-133, bc_periodic_y.wgsl:111-133) copy from it,
// for nx, ny >= 4.  put(o) gets each one's word offset from (i, j).
template <class F>
TE_DEV void periodic_images(const Dims& g, int i, int j, F&& put) {
  int64_t dx[3] = {0, 0, 0}, dy[3] = {0, 0, 0};
  int nxo = 1, nyo = 1;
  if (i >= g.nx - 3) dx[nxo++] = -g.nx;
  if (i <= 3) dx[nxo++] = g.nx;
  if (j >= g.ny - 3) dy[nyo++] = -(int64_t)g.ny * g.sJ;
  if (j <= 3) dy[nyo++] = (int64_t)g.ny * g.sJ;
  for (int b = 0; b < nyo; ++b)
    for (int a = 0; a < nxo; ++a)
      if (a != 0 || b != 0) put(dx[a] + dy[b]);
}
// A pointwise update of levels [k0, k1) of a cell-centred field on a
// periodic domain (nx, ny >= 4) followed by the field's halo fill
// (bc_periodic_x, bc_periodic_y): each cell writes its new value and its
// periodic images, the cells next to levels 0 and nz also the images of
// those levels (copied as they are: the update does not write them, and no
// thread of the launch reads the field's halo).
template <class F>
TE_DEV void update_with_images(const Dims& g, float* f, int i, int j, int k, int k0, int k1, F value) {
  const int64_t c = g.idx(i, j, k), sk = g.sk();
  const bool lo = k == k0 && k0 == 1, hi = k == k1 - 1 && k1 == g.nz;
  const float flo = lo ? f[c - sk] : 0.0f, fhi = hi ? f[c + sk] : 0.0f;
  const float v = value(c);
  f[c] = v;
  periodic_images(g, i, j, [&](int64_t o) {
    f[c + o] = v;
    if (lo) f[c - sk + o] = flo;
    if (hi) f[c + sk + o] = fhi;
  });
}
// GridDims bytes: nx ny nz strideJ strideK 1/dx 1/dy 1/dz (timestep.js:88-90).
TE_DEV Dims dims(const uint32_t* M, uint32_t o) {
  Dims g;
  g.nx = ldi(M, o); g.ny = ldi(M, o + 1); g.nz = ldi(M, o + 2); g.sJ = ldi(M, o + 3); g.sK = ldi(M, o + 4);
  g.rdx = ldf(M, o + 5); g.rdy = ldf(M, o + 6); g.rdz = ldf(M, o + 7);
  return g;
}
struct Acoustic { float dts, alph, beta, kdiv, smeps; };
// AcousticParamsGPU bytes: dts alph beta kdiv smeps (timestep.js:46-52).
TE_DEV Acoustic acoustic(const uint32_t* M, uint32_t o) {
  return {ldf(M, o), ldf(M, o + 1), ldf(M, o + 2), ldf(M, o + 3), ldf(M, o + 4)};
}

TE_DEV float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
TE_DEV float mx(float a, float b) { return (a < b || a != a) ? b : a; }
TE_DEV float mn(float a, float b) { return (b < a || a != a) ? b : a; }
TE_DEV float ab(float a) { return __builtin_fabsf(a); }
TE_DEV float copysign_mag(float mag, float sgn) {
  return __uint_as_float((__float_as_uint(mag) & 0x7fffffffu) | (__float_as_uint(sgn) & 0x80000000u));
}

// Metric buffer (grid.js:200-214): uh | uf | vh | vf ...
TE_DEV float hm_uh(const float* hm, const Dims& g, int i) { return hm[i + 3]; }
TE_DEV float hm_uf(const float* hm, const Dims& g, int i) { return hm[g.sJ + i + 3]; }
TE_DEV float hm_vh(const float* hm, const Dims& g, int j) { return hm[2 * g.sJ + j + 3]; }
TE_DEV float hm_vf(const float* hm, const Dims& g, int j) { return hm[2 * g.sJ + g.sK + j + 3]; }

constexpr float RD = 287.04f, RV = 461.5f, CP = 1005.7f, CV = 718.66003f, CPV = 1.87e+03f, CVV = 1408.5f,
                ROVCP = 0.28541315f, CPL = 4.19e+03f, CPI = 2106.0f, P00 = 1e+05f, GRAV = 9.81f;

// ---- advection stencils
// 5th-order upwind face value, weights (2,-13,47,27,-3)/60 (adv_flux_x.wgsl:106-113)
TE_DEV float flx5(float s1, float s2, float s3, float s4, float s5) {
  return fma_(-3.0f / 60.0f, s5, fma_(27.0f / 60.0f, s4, fma_(47.0f / 60.0f, s3, fma_(2.0f / 60.0f, s1, (-13.0f / 60.0f) * s2))));
}
// 3rd-order face value, weights (-1,5,2)/6 (adv_flux_z.wgsl:106-111)
TE_DEV float flx3(float s1, float s2, float s3) {
  return fma_(2.0f / 6.0f, s3, fma_(-1.0f / 6.0f, s1, (5.0f / 6.0f) * s2));
}
// WENO5 with WENO-Z style weights (adv_flux_x_weno.wgsl:121-151)
TE_DEV float weno5(float s1, float s2, float s3, float s4, float s5, float weps) {
  const float c1312 = 13.0f / 12.0f;
  float a1 = fma_(-2.0f, s2, s1) + s3, q1 = fma_(3.0f, s3, fma_(-4.0f, s2, s1));
  float b1 = fma_(c1312 * a1, a1, (0.25f * q1) * q1);
  float a2 = fma_(-2.0f, s3, s2) + s4, q2 = s2 - s4;
  float b2 = fma_(c1312 * a2, a2, (0.25f * q2) * q2);
  float a3 = fma_(-2.0f, s4, s3) + s5, q3 = fma_(3.0f, s3, -(4.0f * s4)) + s5;
  float b3 = fma_(c1312 * a3, a3, (0.25f * q3) * q3);
  float tau = ab(b1 - b3);
  float r1 = mn(1e+30f, tau / (b1 + weps));
  float r2 = mn(1e+30f, tau / (b2 + weps));
  float r3 = mn(1e+30f, tau / (b3 + weps));
  float m = mx(r1, mx(r2, r3));
  float one = 1.0f;
  if (m > 1e+15f) {
    float im = 1.0f / m;
    r1 = r1 * im; r2 = r2 * im; r3 = r3 * im;
    one = im * im;
  }
  float w1 = 0.1f * fma_(r1, r1, one);
  float w2 = 0.6f * fma_(r2, r2, one);
  float w3 = 0.3f * fma_(r3, r3, one);
  float p1 = fma_(11.0f / 6.0f, s3, fma_(2.0f / 6.0f, s1, (-7.0f / 6.0f) * s2));
  float p2 = fma_(2.0f / 6.0f, s4, fma_(-1.0f / 6.0f, s2, (5.0f / 6.0f) * s3));
  float p3 = fma_(-1.0f / 6.0f, s5, fma_(2.0f / 6.0f, s3, (5.0f / 6.0f) * s4));
  return fma_(w3, p3, fma_(w1, p1, w2 * p2)) / ((w1 + w2) + w3);
}
TE_DEV bool weno_pdef_ok(float s1, float s2, float s3, float s4, float s5, float weps, bool pdef) {
  if (!pdef) return true;
  float bmax = mx(mx(mx(ab(s1), ab(s2)), mx(ab(s3), ab(s4))), ab(s5));
  return bmax > weps;
}
// flux-limited 2nd-order upstream value (adv_flux_z_weno.wgsl:111-120)
TE_DEV float upstrpd(float s1, float s2, float s3, float weps) {
  float dd = s2 - s1;
  float rr = (s3 - s2) / (copysign_mag(weps, dd) + dd);
  float phi = mx(0.0f, mn(rr, mn(fma_(2.0f / 6.0f, rr, 1.0f / 6.0f), 1.0f)));
  return fma_(phi, s2 - s1, s2);
}

// ---- double-float log/exp/pow with a final rounding fix
// (pressure_tendency.wgsl:114-227: the reference's own algorithm, kept so the
// thermodynamic powers match its recording bit for bit).
struct FF { float x, y; };
TE_DEV FF ff_two_sum(float a, float b) {
  float s = a + b, bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
TE_DEV FF ff_quick_two_sum(float a, float b) {
  float s = a + b;
  return {s, b - (s - a)};
}
TE_DEV FF ff_mul(FF a, FF b) {
  float p = a.x * b.x;
  float e = fma_(a.x, b.x, -p);
  e = fma_(a.x, b.y, e);
  e = fma_(a.y, b.x, e);
  return ff_quick_two_sum(p, e);
}
TE_DEV FF ff_add(FF a, FF b) {
  FF s = ff_two_sum(a.x, b.x), t = ff_two_sum(a.y, b.y);
  FF u = ff_quick_two_sum(s.x, s.y + t.x);
  return ff_quick_two_sum(u.x, u.y + t.y);
}
TE_DEV FF ff_log(float x) {
  int e;
  float m = frexpf(x, &e);
  if (m < 0.70710677f) { m = m * 2.0f; e = e - 1; }
  float f = m - 1.0f;
  FF d = ff_two_sum(2.0f, f);
  float sh = f / d.x;
  float r = fma_(-sh, d.x, f);
  r = fma_(-sh, d.y, r);
  FF sf = ff_quick_two_sum(sh, r / d.x);
  FF s2 = ff_mul(sf, sf), s3 = ff_mul(s2, sf);
  FF P = {0.11764706f, -4.3826945e-10f};
  P = ff_add(ff_mul(P, s2), FF{0.13333334f, -6.9538753e-09f});
  P = ff_add(ff_mul(P, s2), FF{0.15384616f, -5.731216e-09f});
  P = ff_add(ff_mul(P, s2), FF{0.18181819f, -5.4186042e-09f});
  P = ff_add(ff_mul(P, s2), FF{0.22222222f, -1.6556846e-09f});
  P = ff_add(ff_mul(P, s2), FF{0.2857143f, -1.2772424e-08f});
  P = ff_add(ff_mul(P, s2), FF{0.4f, -5.9604646e-09f});
  P = ff_add(ff_mul(P, s2), FF{0.6666667f, -1.9868216e-08f});
  FF l1p = ff_add(FF{2.0f * sf.x, 2.0f * sf.y}, ff_mul(s3, P));
  const float ln2h = 0.69314575f, ln2m = 1.4285906e-06f, ln2l = 1.6198487e-11f;
  float fe = (float)e;
  FF a = ff_add(FF{fe * ln2h, fe * ln2m}, l1p);
  return ff_add(a, FF{fe * ln2l, 0.0f});
}
TE_DEV float ff_exp(FF y) {
  const float ln2h = 0.69314575f, ln2m = 1.4285906e-06f, ln2l = 1.6198487e-11f;
  float kq = rintf(y.x / 0.6931472f);
  FF r = ff_add(y, FF{-kq * ln2h, -kq * ln2m});
  r = ff_add(r, FF{-kq * ln2l, 0.0f});
  FF P = {2.5052108e-08f, 4.417623e-16f};
  P = ff_add(ff_mul(P, r), FF{2.755732e-07f, -7.575112e-15f});
  P = ff_add(ff_mul(P, r), FF{2.7557319e-06f, 3.7935712e-14f});
  P = ff_add(ff_mul(P, r), FF{2.4801588e-05f, -3.406996e-13f});
  P = ff_add(ff_mul(P, r), FF{0.0001984127f, -2.7255969e-12f});
  P = ff_add(ff_mul(P, r), FF{0.0013888889f, -3.3631094e-11f});
  P = ff_add(ff_mul(P, r), FF{0.008333334f, -4.346172e-10f});
  P = ff_add(ff_mul(P, r), FF{0.041666668f, -1.2417635e-09f});
  P = ff_add(ff_mul(P, r), FF{0.16666667f, -4.967054e-09f});
  P = ff_add(ff_mul(P, r), FF{0.5f, 0.0f});
  P = ff_add(ff_mul(P, r), FF{1.0f, 0.0f});
  FF Q = ff_mul(P, r);
  float s = 1.0f + Q.x;
  float t = (1.0f - s) + Q.x;
  FF ur = ff_two_sum(t, Q.y);
  float res = s + ur.x;
  float bb = res - s;
  float e3 = (s - (res - bb)) + (ur.x - bb);
  if (e3 != 0.0f && ur.y != 0.0f && ((e3 > 0.0f) == (ur.y > 0.0f))) {
    float nb = __uint_as_float(__float_as_uint(res) + ((e3 > 0.0f) ? 1u : 0xFFFFFFFFu));
    if (e3 + e3 == nb - res) res = nb;
  }
  return ldexpf(res, (int)kq);
}
TE_DEV float rr_pow(float x, float p) {
  if (!(x > 0.0f) || (__float_as_uint(x) & 0x7fffffffu) == 0x7f800000u) return powf(x, p);
  FF L = ff_log(x);
  float yh = L.x * p;
  return ff_exp(ff_quick_two_sum(yh, fma_(L.y, p, fma_(L.x, p, -yh))));
}
TE_DEV float rr_log(float x) {
  if (!(x > 0.0f) || (__float_as_uint(x) & 0x7fffffffu) == 0x7f800000u) return logf(x);
  FF a = ff_log(x);
  return a.x + a.y;
}
TE_DEV float rr_exp(float x) {
  if (!(ab(x) < 87.0f)) return expf(x);
  return ff_exp(FF{x, 0.0f});
}

// The same results faster: the routines above round a double-float
// approximation of log / exp / pow to binary32.  A binary64 evaluation
// (device libm, within a few ulp of binary64) rounds to the same float
// whenever the value is not within 2^-40 (pow: 2^-32, its double-float
// argument carries the error of log x times p) of a rounding boundary between
// two floats (both approximations are closer than that); near a boundary
// the double-float routine decides.  The library's self-test compares each
// fast_* with its rr_* on every binary32 input (pow: at the exponents the
// kernels use), bit for bit.
TE_DEV bool f32_clear(double d, float& f, double margin = 0x1p-40) {
  f = (float)d;
  const double df = (double)f;
  const uint32_t b = __float_as_uint(f);
  // the neighbour of f on d's side: one step up in magnitude when d lies
  // further from zero than f (signed zeros included)
  const float nb = __uint_as_float(b + (((d > df) != ((b >> 31) != 0u)) ? 1u : 0xFFFFFFFFu));
  return __builtin_fabs(d - 0.5 * (df + (double)nb)) > margin * __builtin_fabs(d);
}
TE_DEV float fast_log(float x) {
  if (!(x > 0.0f) || (__float_as_uint(x) & 0x7fffffffu) == 0x7f800000u) return logf(x);
  float f;
  return f32_clear(log((double)x), f) ? f : rr_log(x);
}
TE_DEV float fast_exp(float x) {
  if (!(ab(x) < 87.0f)) return expf(x);
  float f;
  return f32_clear(exp((double)x), f) ? f : rr_exp(x);
}
TE_DEV float fast_pow(float x, float p) {
  if (!(x > 0.0f) || (__float_as_uint(x) & 0x7fffffffu) == 0x7f800000u) return powf(x, p);
  float f;
  return f32_clear(exp((double)p * log((double)x)), f, 0x1p-32) ? f : rr_pow(x, p);
}

// 1 / x rounded to nearest even, for the tridiagonal recurrences' pivots:
// the hardware reciprocal estimate (within 1 ulp) refined by one
// Newton-Raphson step with an FMA residual (Markstein's reciprocal
// iteration, "IA-64 and Elementary Functions", 2000), for |x| in
// [2^-64, 2^64]; other x (zero, infinities, NaN, extremes) take the IEEE
// division.  te_selftest_math compares it with 1.0f / x on every binary32
// input (the estimate alone differs on about a tenth of them; after the
// step, on none).
TE_DEV float recip(float x) {
  const float a = __builtin_fabsf(x);
  if (!(a >= 0x1p-64f && a <= 0x1p64f)) return 1.0f / x;
  const float y = __builtin_amdgcn_rcpf(x);
  return fma_(fma_(-x, y, 1.0f), y, y);
}

// global thread ids of a 3-D launch
// Elimination coefficients of a column sweep: level k of this lane's column,
// either in the field at p (cA / cB of the reference) or, for the fused
// launches, in dynamic LDS (slot k + off of nl lanes; they are dead scratch).
extern __shared__ float te_lds[];
template <bool LDS>
struct ColStore {
  float* p;
  int64_t base, sk;
  int lane, nl, off;
  __device__ __forceinline__ float& at(int k) const {
    if (LDS) return te_lds[(k + off) * nl + lane];
    return p[base + k * sk];
  }
};
#define TE_LANE ((int)(threadIdx.x + a.lg[0] * threadIdx.y))
#define TE_NLANES ((int)(a.lg[0] * a.lg[1]))

// Workgroup coordinates, renumbered for the 8 XCDs.  The dispatcher hands
// workgroups to the XCDs round-robin in dispatch order (workgroup w, counted
// i-fastest over the grid, runs on XCD w mod 8), and each XCD has its own L2.
// Neighbouring workgroups share cache lines (row ends of 199-word rows,
// stencil halos), so numbering the grid such that XCD x runs the x-th
// contiguous eighth of it keeps those lines in one L2 instead of fetching
// (and partially writing) them from two.  A bijection of the workgroups:
// XCD x holds q + (x < r) of the n = 8q + r of them.
#ifndef TE_XCD_REMAP
#define TE_XCD_REMAP 1
#endif
struct Blk {
  int x, y, z;
};
TE_DEV uint32_t te_div(uint32_t l, uint32_t mlo, uint32_t mhi) {
  return (uint32_t)((((uint64_t)mhi << 32 | mlo) * (uint64_t)l) >> 40);
}
__device__ __forceinline__ Blk te_block(const KArg& a) {
  const uint32_t nbx = a.lg[3], nby = a.lg[4];
#if TE_XCD_REMAP
  const uint32_t n = a.lg[9];
  const uint32_t w = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
  const uint32_t q = n >> 3, r = n & 7u, x = w & 7u;
  const uint32_t l = x * q + (x < r ? x : r) + (w >> 3);
  const uint32_t yz = te_div(l, a.lg[5], a.lg[6]), z = te_div(yz, a.lg[7], a.lg[8]);
  return Blk{(int)(l - yz * nbx), (int)(yz - z * nby), (int)z};
#else
  (void)nbx;
  (void)nby;
  return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
#endif
}
// block dims as the runtime launched them (kernel arguments, not the
// dispatch packet)
#define TE_BX ((int)a.lg[0])
#define TE_BY ((int)a.lg[1])
#define TE_BZ ((int)a.lg[2])

// Flat launches.  A pointwise op over a whole-field grid [strideJ, strideK,
// nz+1] (or over a word range) touches every word of its fields once, in
// memory order, with no dependence on the position: the runtime then
// launches it over the n0*n1*n2 words in runs of 4 per thread (16-byte
// vector accesses; it checks every field is 16-byte aligned).  te_flat calls
// op(word, width) for this thread's run: width 4 when whole, else 1 per word
// of the tail, and reports whether the launch was flat.
template <class F>
TE_DEV bool te_flat(const KArg& a, F op) {
  if (a.lg[10] == 0) return false;
  const uint32_t W = a.n[0] * a.n[1] * a.n[2];
  const uint32_t f = 4u * ((uint32_t)te_block(a).x * (uint32_t)a.lg[0] + threadIdx.x);
  if (f + 4 <= W) {
    op(f, 4);
  } else {
    for (uint32_t w = f; w < W; ++w) op(w, 1);
  }
  return true;
}
// te_flat for ops whose per-word work depends on the level only: op(word,
// width, k) with k the word's level (L words per level); a run of 4 that
// straddles two levels is done word by word.
template <class F>
TE_DEV bool te_flat_levels(const KArg& a, uint32_t L, F op) {
  if (a.lg[10] == 0) return false;
  const uint32_t W = a.n[0] * a.n[1] * a.n[2];
  const uint32_t f = 4u * ((uint32_t)te_block(a).x * (uint32_t)a.lg[0] + threadIdx.x);
  if (f >= W) return true;
  const uint32_t k0 = f / L, next = (k0 + 1) * L;
  if (f + 4 <= W && f + 4 <= next) {
    op(f, 4, (int)k0);
  } else {
    for (uint32_t w = f; w < W && w < f + 4; ++w) op(w, 1, (int)(w < next ? k0 : k0 + 1));
  }
  return true;
}
// 4 or 1 floats at word w of p
struct F4 {
  float v[4];
};
TE_DEV F4 ld4(const uint32_t* p, uint32_t w, int width) {
  F4 r;
  if (width == 4) {
    const float4 t = *reinterpret_cast<const float4*>(p + w);
    r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
  } else {
    r.v[0] = __uint_as_float(p[w]);
  }
  return r;
}
TE_DEV void st4(uint32_t* p, uint32_t w, int width, const F4& r) {
  if (width == 4) *reinterpret_cast<float4*>(p + w) = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
  else p[w] = __float_as_uint(r.v[0]);
}

// Thread coordinates of a 3-D launch.  Plane launches (rows whose length is
// not a multiple of 64: whole-field rows of nx+7, staggered rows of nx+1)
// number the (i, j) plane's points consecutively, 256 per workgroup, with
// the level on the grid's y axis, so no wave idles at a row's end.
TE_DEV void te_tid3(const KArg& a, int& gx, int& gy, int& gz) {
  const Blk b = te_block(a);
  if (a.lg[11] != 0) {
    const uint32_t t = (uint32_t)b.x * 256u + threadIdx.x;
    const uint32_t row = te_div(t, a.lg[12], a.lg[13]);
    gx = (int)(t - row * a.n[0]);
    gy = (int)row;
    gz = b.y;
    return;
  }
  gx = b.x * TE_BX + (int)threadIdx.x;
  gy = b.y * TE_BY + (int)threadIdx.y;
  gz = b.z * TE_BZ + (int)threadIdx.z;
}
#define TE_TID3                                                          \
  int gx, gy, gz;                                                        \
  te_tid3(a, gx, gy, gz);                                                \
  if (gx >= (int)a.n[0] || gy >= (int)a.n[1] || gz >= (int)a.n[2]) return;

}  // namespace te
