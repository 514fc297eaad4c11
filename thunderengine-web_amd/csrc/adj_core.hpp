// Reverse-mode building blocks for the adjoint of the large step.
//
// The adjoint of a reference kernel is built from a per-point body that
// restates the kernel's arithmetic on dual numbers D1 (value, derivative).
// For one point of the kernel's thread grid, adj_point() runs the body
//   * once to find the point's outputs (their adjoints are read and then
//     cleared: the kernel overwrites them) and to count its active loads, then
//   * once per active load s, with that load seeded (derivative 1): each
//     output y then carries dy/dx_s, and sum_o ybar_o * dy_o/dx_s is added to
//     the adjoint of the loaded word x_s.
// This is the transposed local Jacobian of the point, so a launch over the
// kernel's grid adds J^T ybar to the adjoint memory.  Loads of profiles, the
// metric buffer and parameter words are passive (no derivative).  Branch
// decisions are taken on the body's own intermediates, formed from the forward
// state's words, so the derivative is that of the branch the forward step took
// (up to the rounding of those intermediates right at a switch).
//
// The header compiles for gfx950 (atomic adds into the adjoint memory) and for
// the host (plain adds), so the same bodies can be exercised on a CPU.
#pragma once
#include <stdint.h>
#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define TE_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define TE_HD inline
#endif

namespace te {
namespace adj {

// ---- dual numbers.  Values and derivatives are carried in binary64: the
// local Jacobians of the WENO weights divide by smoothness indicators formed
// from differences of nearly equal values, where binary32 intermediates would
// put the linearisation point off by ~1e-3 relative.  The inputs are the
// binary32 words of the forward state, so branch decisions follow the values
// the forward step computed from, up to rounding of the intermediates.
struct D1 {
  double v, d;
  TE_HD D1() : v(0.0), d(0.0) {}
  TE_HD D1(double x) : v(x), d(0.0) {}
  TE_HD D1(double x, double dx) : v(x), d(dx) {}
};
TE_HD double fmad_(double a, double b, double c) { return __builtin_fma(a, b, c); }
TE_HD float fmaf_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
TE_HD D1 operator+(D1 a, D1 b) { return D1(a.v + b.v, a.d + b.d); }
TE_HD D1 operator-(D1 a, D1 b) { return D1(a.v - b.v, a.d - b.d); }
TE_HD D1 operator-(D1 a) { return D1(-a.v, -a.d); }
TE_HD D1 operator*(D1 a, D1 b) { return D1(a.v * b.v, fmad_(a.d, b.v, a.v * b.d)); }
TE_HD D1 operator/(D1 a, D1 b) {
  const double q = a.v / b.v;
  return D1(q, (a.d - q * b.d) / b.v);
}
TE_HD bool operator<(D1 a, D1 b) { return a.v < b.v; }
TE_HD bool operator>(D1 a, D1 b) { return a.v > b.v; }
TE_HD bool operator<=(D1 a, D1 b) { return a.v <= b.v; }
TE_HD bool operator>=(D1 a, D1 b) { return a.v >= b.v; }
TE_HD bool operator!=(D1 a, D1 b) { return a.v != b.v; }
TE_HD bool operator==(D1 a, D1 b) { return a.v == b.v; }
// fma(a, b, c): product rule in the derivative
TE_HD D1 fma_(D1 a, D1 b, D1 c) { return D1(fmad_(a.v, b.v, c.v), fmad_(a.d, b.v, fmad_(a.v, b.d, c.d))); }
// WGSL max / min as selects (ties keep the first operand), as the forward kernels
TE_HD D1 mx(D1 a, D1 b) { return (a.v < b.v || a.v != a.v) ? b : a; }
TE_HD D1 mn(D1 a, D1 b) { return (b.v < a.v || a.v != a.v) ? b : a; }
TE_HD D1 ab(D1 a) { return a.v < 0.0 ? D1(-a.v, -a.d) : a; }
TE_HD bool sign_bit(double x) {
  union { double f; uint64_t u; } q{x};
  return (q.u >> 63) != 0u;
}
// |mag| with the sign of sgn
TE_HD D1 copysign_mag(D1 mag, D1 sgn) {
  const bool neg = sign_bit(sgn.v);
  const D1 m = ab(mag);
  return neg ? D1(-m.v, -m.d) : m;
}
// x^p for a constant p
TE_HD D1 powc(D1 x, double p) {
#ifdef __HIP_DEVICE_COMPILE__
  const double y = pow(x.v, p);
#else
  const double y = std::pow(x.v, p);
#endif
  return D1(y, x.d * ((p * y) / x.v));
}

// ---- dispatch argument: thread grid, word offsets of parameters 0..31 (the
// reference's 64-word argument slot: A[0..2] grid, A[4 + n] parameter n) and
// where each parameter's values are read: the program memory, or the reverse
// sweep's write log when a later dispatch of the step has overwritten them.
// Adjoints always go to word p[n] + c of the adjoint memory.
struct AdjArg {
  uint32_t n[3];
  uint32_t p[32];
  const float* v[32];
};

// A parameter as seen by a body: its values and its adjoint offset.
struct P {
  const float* v;
  uint32_t off;
  TE_HD operator const float*() const { return v; }
  TE_HD float operator[](int64_t i) const { return v[i]; }
};

// GridDims bytes (timestep.js:88-90): nx ny nz strideJ strideK 1/dx 1/dy 1/dz.
struct ADims {
  int nx, ny, nz, sJ, sK;
  float rdx, rdy, rdz;
  TE_HD int64_t sk() const { return (int64_t)sJ * sK; }
  TE_HD int64_t idx(int i, int j, int k) const {
    return (int64_t)(i + 3) + (int64_t)sJ * ((j + 3) + (int64_t)sK * k);
  }
  // inverse of idx (floor division: words before the field's first word
  // decode to k = -1)
  TE_HD void coords(int64_t q, int& i, int& j, int& k) const {
    const uint32_t plane = (uint32_t)(sJ * sK);
    uint32_t u;  // 32-bit divisions: a field has fewer than 2^31 words
    if (q >= 0) {
      u = (uint32_t)q;
      k = (int)(u / plane);
    } else {
      const int64_t up = q + (int64_t)plane;  // words before the field: level -1
      u = (uint32_t)up;
      k = -1;
    }
    const uint32_t rem = q >= 0 ? u - (uint32_t)k * plane : u;
    const uint32_t jj = rem / (uint32_t)sJ;
    j = (int)jj - 3;
    i = (int)(rem - jj * (uint32_t)sJ) - 3;
  }
  // (i, j, k) moved by d along i (axis 0) or j (axis 1), wrapping into the
  // neighbouring row / level as the linear index does (|d| < sJ, sK)
  TE_HD void step(int axis, int d, int& i, int& j, int& k) const {
    if (axis == 0) {
      i += d;
      if (i < -3) { i += sJ; --j; }
      else if (i > sJ - 4) { i -= sJ; ++j; }
      if (j < -3) { j += sK; --k; }
      else if (j > sK - 4) { j -= sK; ++k; }
    } else {
      j += d;
      if (j < -3) { j += sK; --k; }
      else if (j > sK - 4) { j -= sK; ++k; }
    }
  }
};

constexpr int MAXO = 8;   // outputs per point
constexpr int MAXL = 32;  // active loads whose values the discovery pass keeps for the seeded passes

// Accessor handed to the bodies: memory, adjoint memory and the pass state.
struct Acc {
  const float* M;  // values (word 0 of the program memory)
  float* Mb;       // adjoints (word 0: word p[n] + c is the adjoint of parameter n's value c)
  const AdjArg* a;
  int mode;  // 0: discovery pass, 1: seeded pass
  bool cache = true;  // seeded passes read the kept load values (lv)
  int cnt, seed;
  int64_t rec;
  double acc;
  double yb[MAXO];
  int64_t oa[MAXO];
  bool has[MAXO];
  bool clr[MAXO];  // the output overwrites its word (out), or adds to it (keep)
  float lv[MAXL];  // values of the first MAXL active loads, in load order

  TE_HD Acc(const float* M_, float* Mb_, const AdjArg* a_)
      : M(M_), Mb(Mb_), a(a_), mode(0), cnt(0), seed(-1), rec(-1), acc(0.0) {
    for (int o = 0; o < MAXO; ++o) { yb[o] = 0.0; oa[o] = 0; has[o] = false; clr[o] = false; }
  }
  // parameter n: its values and adjoint offset / its words
  TE_HD P f(int n) const { return P{a->v[n], a->p[n]}; }
  TE_HD uint32_t word(int n, int k = 0) const {
    union { float f; uint32_t u; } q{a->v[n][k]};
    return q.u;
  }
  TE_HD int32_t I(int n, int k = 0) const { return (int32_t)word(n, k); }
  TE_HD float F(int n, int k = 0) const { return a->v[n][k]; }
  TE_HD ADims dims() const {
    ADims g;
    g.nx = I(0, 0); g.ny = I(0, 1); g.nz = I(0, 2); g.sJ = I(0, 3); g.sK = I(0, 4);
    g.rdx = F(0, 5); g.rdy = F(0, 6); g.rdz = F(0, 7);
    return g;
  }
  // active load of p[c].  The seeded passes take the same loads in the same
  // order (their branches are decided on the same values), so they read the
  // values the discovery pass kept instead of the memory.
  TE_HD D1 ld(P p, int64_t c) {
    float x;
    if (mode == 0 || !cache) {
      x = p.v[c];
      if (cache && cnt < MAXL) lv[cnt] = x;
    } else {
      x = cnt < MAXL ? lv[cnt] : p.v[c];
    }
    D1 r((double)x);
    if (mode != 0 && cnt == seed) {
      r.d = 1.0;
      rec = (int64_t)p.off + c;
    }
    ++cnt;
    return r;
  }
  // output y to p[c]
  template <int O>
  TE_HD void out(P p, int64_t c, D1 y) {
    static_assert(O < MAXO, "too many outputs");
    if (mode == 0) {
      const int64_t w = (int64_t)p.off + c;
      oa[O] = w;
      yb[O] = (double)Mb[w];
      has[O] = true;
      clr[O] = true;
    } else {
      acc = fmad_(yb[O], y.d, acc);
    }
  }
  // y added to p[c] (p[c] += y, the old value not among the body's loads):
  // the word's adjoint stays as it is, the loads of y receive ybar * dy/dx
  template <int O>
  TE_HD void keep(P p, int64_t c, D1 y) {
    out<O>(p, c, y);
    if (mode == 0) clr[O] = false;
  }
};

TE_HD void adj_add(float* p, float v) {
#ifdef __HIP_DEVICE_COMPILE__
  unsafeAtomicAdd(p, v);
#else
  *p += v;
#endif
}

// Bodies declaring kReload re-read their loads in every seeded pass: for
// them the kept values cost more registers (or scratch) than the reloads.
template <class B, class = void>
struct KeepLoads {
  static constexpr bool value = true;
};
template <class B>
struct KeepLoads<B, decltype((void)B::kReload)> {
  static constexpr bool value = false;
};

// Adjoint of one point of a kernel whose per-point body is B::run(Acc&, x, y, z).
template <class B>
TE_HD void adj_point(const float* M, float* Mb, const AdjArg& a, int x, int y, int z) {
  Acc r(M, Mb, &a);
  r.cache = KeepLoads<B>::value;
  B::run(r, x, y, z);
  bool any = false;
  for (int o = 0; o < MAXO; ++o) any = any || r.has[o];
  if (!any) return;
  const int n = r.cnt;
  for (int o = 0; o < MAXO; ++o)
    if (r.has[o] && r.clr[o]) Mb[r.oa[o]] = 0.0f;
  r.mode = 1;
  for (int s = 0; s < n; ++s) {
    r.cnt = 0;
    r.seed = s;
    r.acc = 0.0;
    r.rec = -1;
    B::run(r, x, y, z);
    if (r.acc != 0.0 && r.rec >= 0) adj_add(Mb + r.rec, (float)r.acc);
  }
}

// A kernel whose per-point update is a sum of independent parts, each a body
// that adds its part with keep<> (the seeded passes of one body grow with the
// square of its active loads; parts split them).
template <class... Bs>
struct Parts {};
template <class B>
struct PointAdj {
  static TE_HD void run(const float* M, float* Mb, const AdjArg& a, int x, int y, int z) {
    adj_point<B>(M, Mb, a, x, y, z);
  }
};
template <class... Bs>
struct PointAdj<Parts<Bs...>> {
  static TE_HD void run(const float* M, float* Mb, const AdjArg& a, int x, int y, int z) {
    (adj_point<Bs>(M, Mb, a, x, y, z), ...);
  }
};

// Metric buffer (grid.js:200-214): uh | uf | vh | vf ...
TE_HD float hm_uh(const float* hm, const ADims& g, int i) { return hm[i + 3]; }
TE_HD float hm_uf(const float* hm, const ADims& g, int i) { return hm[g.sJ + i + 3]; }
TE_HD float hm_vh(const float* hm, const ADims& g, int j) { return hm[2 * g.sJ + j + 3]; }
TE_HD float hm_vf(const float* hm, const ADims& g, int j) { return hm[2 * g.sJ + g.sK + j + 3]; }

constexpr float RD = 287.04f, RV = 461.5f, CP = 1005.7f, CV = 718.66003f, CPV = 1.87e+03f, CVV = 1408.5f,
                ROVCP = 0.28541315f, CPL = 4.19e+03f, CPI = 2106.0f, P00 = 1e+05f, GRAV = 9.81f;

}  // namespace adj
}  // namespace te
