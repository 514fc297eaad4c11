// CPU oracle for the CM1 large step — TEST INFRASTRUCTURE ONLY.
//
// This is a strict binary32 restatement of the reference kernels
// (data/kernels/*.wgsl of ThunderEngine), one C++ function per kernel, written
// to be checked against the reference's recorded hashes and against the HIP
// product path.  Nothing in the product links this code: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Arithmetic rules (reference: adv_flux_x_weno.wgsl:13-71, f32.js:13-41):
//  * every `a op b` is one IEEE binary32 operation, round-to-nearest-even;
//    the build uses -ffp-contract=off so nothing is fused implicitly;
//  * fma() appears exactly where the reference writes fma(), same operand order;
//  * div_cr / sqrt_cr are the correctly rounded IEEE operations;
//  * WGSL min/max are restated as explicit selects (ties keep the first operand),
//    the same selects the HIP kernels use, so both sides agree on signed zeros.
//
// The kernels are templates over the arithmetic type T: T=float is the oracle,
// T=double is the FP64 build of the same step used for error reports.  In the
// FP64 build the constants keep their binary32 values (same discrete model,
// wider arithmetic).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <xmmintrin.h>

namespace orc {

// One dispatch: word memory `w` (u32 view, used for byte parameters and
// hashes), value memory `v` (T view of the same word indices), and the 64-word
// argument slot A: A[0..2] thread grid, A[4+n] word offset of parameter n
// (replay.js:7-11 layout).
template <class T>
struct K {
  uint32_t* w;
  T* v;
  const uint32_t* A;
  uint32_t off(int n) const { return A[4 + n]; }
  T* p(int n) const { return v + off(n); }
  int32_t I(int n, int i = 0) const { return (int32_t)w[off(n) + i]; }
  float F(int n, int i = 0) const {
    float f;
    std::memcpy(&f, &w[off(n) + i], 4);
    return f;
  }
};

// GridDims bytes (timestep.js:88-90): nx ny nz strideJ strideK rdx rdy rdz.
template <class T>
struct Dims {
  int nx, ny, nz, sJ, sK;
  T rdx, rdy, rdz;
  int64_t sk() const { return (int64_t)sJ * sK; }
  int64_t idx(int i, int j, int k) const { return (int64_t)(i + 3) + (int64_t)sJ * ((j + 3) + (int64_t)sK * k); }
};
template <class T>
inline Dims<T> dims(const K<T>& k, int n = 0) {
  Dims<T> g;
  g.nx = k.I(n, 0); g.ny = k.I(n, 1); g.nz = k.I(n, 2); g.sJ = k.I(n, 3); g.sK = k.I(n, 4);
  g.rdx = (T)k.F(n, 5); g.rdy = (T)k.F(n, 6); g.rdz = (T)k.F(n, 7);
  return g;
}

// AcousticParamsGPU bytes (timestep.js:46-52): dts alph beta kdiv smeps.
template <class T>
struct Acoustic { T dts, alph, beta, kdiv, smeps; };
template <class T>
inline Acoustic<T> acoustic(const K<T>& k, int n = 1) {
  return {(T)k.F(n, 0), (T)k.F(n, 1), (T)k.F(n, 2), (T)k.F(n, 3), (T)k.F(n, 4)};
}

// ---- scalar helpers shared by every kernel
inline float fma_(float a, float b, float c) { return std::fmaf(a, b, c); }
inline double fma_(double a, double b, double c) { return std::fma(a, b, c); }
template <class T> inline T mx(T a, T b) { return (a < b || a != a) ? b : a; }   // WGSL max
template <class T> inline T mn(T a, T b) { return (b < a || a != a) ? b : a; }   // WGSL min
template <class T> inline T ab(T a) { return std::fabs(a); }

// ---- FP64 tangent-linear build: T = TL carries (value, tangent) through
// every operation of a kernel; branch decisions use the double values.  The
// tangent of an output is the kernel's Jacobian applied to the input tangents
// (the reference for the adjoint's transposed Jacobians).
struct TL {
  double v, t;
  TL() : v(0.0), t(0.0) {}
  TL(double x) : v(x), t(0.0) {}
  TL(double x, double dx) : v(x), t(dx) {}
};
inline TL operator+(TL a, TL b) { return TL(a.v + b.v, a.t + b.t); }
inline TL operator-(TL a, TL b) { return TL(a.v - b.v, a.t - b.t); }
inline TL operator-(TL a) { return TL(-a.v, -a.t); }
inline TL operator*(TL a, TL b) { return TL(a.v * b.v, a.t * b.v + a.v * b.t); }
inline TL operator/(TL a, TL b) {
  const double q = a.v / b.v;
  return TL(q, (a.t - q * b.t) / b.v);
}
inline bool operator<(TL a, TL b) { return a.v < b.v; }
inline bool operator>(TL a, TL b) { return a.v > b.v; }
inline bool operator<=(TL a, TL b) { return a.v <= b.v; }
inline bool operator>=(TL a, TL b) { return a.v >= b.v; }
inline bool operator==(TL a, TL b) { return a.v == b.v; }
inline bool operator!=(TL a, TL b) { return a.v != b.v; }
inline TL fma_(TL a, TL b, TL c) { return TL(std::fma(a.v, b.v, c.v), a.t * b.v + a.v * b.t + c.t); }
inline TL ab(TL a) { return a.v < 0.0 ? TL(-a.v, -a.t) : a; }
inline TL copysign_mag(TL mag, TL sgn) {
  const TL m = ab(mag);
  return std::signbit(sgn.v) ? TL(-m.v, -m.t) : m;
}
inline TL rr_pow(TL x, TL p) {
  const double y = std::pow(x.v, p.v);
  return TL(y, x.t * (p.v * y / x.v));
}

// Metric buffer hmetric (grid.js:200-214): uh | uf | vh | vf | xf | yf | axi...
template <class T> inline T hm_uh(const T* hm, const Dims<T>& g, int i) { return hm[i + 3]; }
template <class T> inline T hm_uf(const T* hm, const Dims<T>& g, int i) { return hm[g.sJ + i + 3]; }
template <class T> inline T hm_vh(const T* hm, const Dims<T>& g, int j) { return hm[2 * g.sJ + j + 3]; }
template <class T> inline T hm_vf(const T* hm, const Dims<T>& g, int j) { return hm[2 * g.sJ + g.sK + j + 3]; }

// Subnormal handling.  The reference recording flushes subnormal operands and
// results to zero with the sign kept (its division helper div_cr exists because
// "the GPU flushes subnormals", adv_flux_x_weno.wgsl:23-28; the recorded
// acoustic_w_solve hashes only reproduce with per-operation flushing).  The
// oracle therefore runs every kernel with MXCSR FTZ|DAZ set on each thread.
inline void ftz_daz() { _mm_setcsr(_mm_getcsr() | 0x8040u); }

// Thread-grid loops.  The body receives the global invocation id and plays the
// role of the WGSL main() after its ARG[0..2] bounds check.
template <class F>
inline void for3(const uint32_t* A, F&& body) {
  const int64_t X = A[0], Y = A[1], Z = A[2];
#pragma omp parallel
  {
    ftz_daz();
#pragma omp for collapse(2) schedule(static)
    for (int64_t z = 0; z < Z; ++z)
      for (int64_t y = 0; y < Y; ++y)
        for (int64_t x = 0; x < X; ++x) body((int)x, (int)y, (int)z);
  }
}
// Column kernels (rank 2): body(i, j) per thread of the [X, Y] grid.
template <class F>
inline void for2(const uint32_t* A, F&& body) {
  const int64_t X = A[0], Y = A[1];
#pragma omp parallel
  {
    ftz_daz();
#pragma omp for collapse(2) schedule(static)
    for (int64_t y = 0; y < Y; ++y)
      for (int64_t x = 0; x < X; ++x) body((int)x, (int)y);
  }
}

// Physical constants as the kernels hold them (update_diagnostics_moist.wgsl:85-101).
constexpr float RD = 287.04f, RV = 461.5f, CP = 1005.7f, CV = 718.66003f, CPV = 1.87e+03f, CVV = 1408.5f,
                ROVCP = 0.28541315f, CPL = 4.19e+03f, CPI = 2106.0f, P00 = 1e+05f, GRAV = 9.81f;

// ---- registry
template <class T>
using KernelFn = void (*)(const K<T>&);
struct KernelEntry {
  const char* name;
  KernelFn<float> f32;
  KernelFn<double> f64;
  KernelFn<TL> ftl;  // null: no tangent-linear build of this kernel
};
const KernelEntry* find_kernel(const std::string& name);
int kernel_count();
const KernelEntry* kernel_at(int i);

}  // namespace orc
