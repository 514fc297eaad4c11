// Advection kernels, one per reference kernel: 5th-order and WENO5 fluxes of
// scalars and momentum, and their assembly into tendencies.
#include "te_device.hpp"
#include "te_kernels.hpp"

namespace te {

// Every flux below is a value function fv_*(..., i, j, k, out): it returns
// false where the reference kernel writes nothing, else the face value the
// kernel stores.  The flux kernels store it; the fused advection kernel
// (k_fused.hip) evaluates the two faces of a cell directly.
//
// Each function first loads its whole stencil window unconditionally and
// then picks the upwind operands with selects: the arithmetic is the
// reference's, on the same values, but no load waits behind a lane-divergent
// branch, so a wave issues all its loads together.  A window reaches at most
// 3 levels (or 3 halo-padded rows) past the field; the words there belong to
// a neighbouring field or the runtime's guard zones, and no result uses them.
// the window w[0..6] = s at offsets -3..3 along stride st
struct Win7 {
  float v[7];
  TE_DEV float operator()(int n) const { return v[n + 3]; }
};
TE_DEV Win7 hwin(const float* s, int64_t c, int64_t st) {
  Win7 w;
#pragma unroll
  for (int n = -3; n <= 3; ++n) w.v[n + 3] = s[c + n * st];
  return w;
}
TE_DEV Win7 vwin(const float* s, int64_t c, int64_t sk) { return hwin(s, c, sk); }
TE_DEV float sel(bool up, float a, float b) { return up ? a : b; }
// upwind-biased 5-point operands of the face between offsets SH-1 and SH:
// (-3, -2, -1, 0, 1) + SH when up, else (2, 1, 0, -1, -2) + SH
struct Up5 {
  float s1, s2, s3, s4, s5;
};
template <int SH>
TE_DEV Up5 up5(const Win7& W, bool up) {
  return Up5{sel(up, W(-3 + SH), W(2 + SH)), sel(up, W(-2 + SH), W(1 + SH)), sel(up, W(-1 + SH), W(SH)),
             sel(up, W(SH), W(-1 + SH)), sel(up, W(1 + SH), W(-2 + SH))};
}
TE_DEV float flx5(const Up5& o) { return flx5(o.s1, o.s2, o.s3, o.s4, o.s5); }
TE_DEV float weno5(const Up5& o, float weps) { return weno5(o.s1, o.s2, o.s3, o.s4, o.s5, weps); }

// A face value in two parts: fp_* issues every load the face may need (its
// window and velocity or mass-flux operands; per-level profiles at the level
// clamped into [0, nz]) and fe_* decides and computes from them.  fv_* is
// fe_*(fp_*(...)); the fused kernel issues all six faces' loads before any
// of their branches, so a wave waits for memory once.
struct FacePre {
  Win7 W;
  float a, b, c;
};
TE_DEV int lev(const Dims& g, int k) { return k < 0 ? 0 : (k > g.nz ? g.nz : k); }

// adv_flux_x.wgsl:120-143 / adv_flux_y.wgsl:120-144
template <int DIR>
TE_DEV FacePre fp_h5(const Dims& g, const float* s, const float* rr, int64_t c) {
  return {hwin(s, c, DIR == 0 ? 1 : (int64_t)g.sJ), rr[c], 0.0f, 0.0f};
}
template <int DIR>
TE_DEV bool fe_h5(const Dims& g, const FacePre& p, int i, int j, int k, float& out) {
  if (DIR == 0 ? (i > g.nx || j >= g.ny || k >= g.nz) : (i >= g.nx || j > g.ny || k >= g.nz)) return false;
  const float u = p.a;
  out = u * flx5(up5<0>(p.W, u >= 0.0f));
  return true;
}
template <int DIR>
TE_DEV bool fv_h5(const Dims& g, const float* s, const float* rr, int i, int j, int k, int64_t c, float& out) {
  return fe_h5<DIR>(g, fp_h5<DIR>(g, s, rr, c), i, j, k, out);
}
template <int DIR>
__device__ void flux_h5(uint32_t* M, const KArg& a, int i, int j, int k) {
  Dims g = dims(a.q[0], 0);
  float f;
  if (fv_h5<DIR>(g, fq(a.q[1]), fq(a.q[2]), i, j, k, g.idx(i, j, k), f)) fq(a.q[3])[g.idx(i, j, k)] = f;
}
__global__ __launch_bounds__(256) void k_adv_flux_x(uint32_t* M, KArg a) { TE_TID3 flux_h5<0>(M, a, gx, gy, gz); }
__global__ __launch_bounds__(256) void k_adv_flux_y(uint32_t* M, KArg a) { TE_TID3 flux_h5<1>(M, a, gx, gy, gz); }

// adv_flux_z.wgsl:127-200
TE_DEV FacePre fp_z5(const Dims& g, const float* s, const float* rrw, const float* c1, const float* c2, int k, int64_t c) {
  const int l = lev(g, k);
  return {vwin(s, c, g.sk()), rrw[c], c1[l], c2[l]};
}
TE_DEV bool fe_z5(const Dims& g, const FacePre& p, int i, int j, int k, float& out) {
  if (i >= g.nx || j >= g.ny || k > g.nz) return false;
  if (k == 0 || k == g.nz) { out = 0.0f; return true; }
  const float w = p.a, cc1 = p.b, cc2 = p.c;
  const Win7& W = p.W;
  const bool up = w >= 0.0f;
  float f;
  if (k == 1) f = up ? fma_(cc1, W(-1), cc2 * W(0)) : flx3(W(1), W(0), W(-1));
  else if (k == 2) f = up ? flx3(W(-2), W(-1), W(0)) : flx5(W(2), W(1), W(0), W(-1), W(-2));
  else if (k == g.nz - 2) f = up ? flx5(W(-3), W(-2), W(-1), W(0), W(1)) : flx3(W(1), W(0), W(-1));
  else if (k == g.nz - 1) f = up ? flx3(W(-2), W(-1), W(0)) : fma_(cc1, W(-1), cc2 * W(0));
  else f = flx5(up5<0>(W, up));
  out = w * f;
  return true;
}
TE_DEV bool fv_z5(const Dims& g, const float* s, const float* rrw, const float* c1, const float* c2, int i, int j, int k, int64_t c,
                  float& out) {
  return fe_z5(g, fp_z5(g, s, rrw, c1, c2, k, c), i, j, k, out);
}
__global__ __launch_bounds__(256) void k_adv_flux_z(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  float f;
  if (fv_z5(g, fq(a.q[1]), fq(a.q[2]), fq(a.q[3]), fq(a.q[4]), gx, gy, gz, g.idx(gx, gy, gz), f))
    fq(a.q[5])[g.idx(gx, gy, gz)] = f;
}

// adv_flux_x_weno.wgsl:158-200 / adv_flux_y_weno.wgsl:158-201
template <int DIR>
TE_DEV FacePre fp_h_weno(const Dims& g, const float* s, const float* rr, int64_t c) { return fp_h5<DIR>(g, s, rr, c); }
template <int DIR>
TE_DEV bool fe_h_weno(const Dims& g, const FacePre& p, float weps, bool pdef, int i, int j, int k, float& out) {
  if (i > g.nx || j > g.ny || k >= g.nz) return false;
  const float u = p.a;
  const Up5 o = up5<0>(p.W, u >= 0.0f);
  out = weno_pdef_ok(o.s1, o.s2, o.s3, o.s4, o.s5, weps, pdef) ? u * weno5(o, weps) : 0.0f;
  return true;
}
template <int DIR>
TE_DEV bool fv_h_weno(const Dims& g, const float* s, const float* rr, float weps, bool pdef, int i, int j, int k, int64_t c,
                      float& out) {
  return fe_h_weno<DIR>(g, fp_h_weno<DIR>(g, s, rr, c), weps, pdef, i, j, k, out);
}
template <int DIR>
__device__ void flux_h_weno(uint32_t* M, const KArg& a, int i, int j, int k) {
  Dims g = dims(a.q[0], 0);
  float f;
  if (fv_h_weno<DIR>(g, fq(a.q[1]), fq(a.q[2]), ldf(a.q[3], 0), ldi(a.q[4], 0) == 1, i, j, k, g.idx(i, j, k), f))
    fq(a.q[5])[g.idx(i, j, k)] = f;
}
__global__ __launch_bounds__(256) void k_adv_flux_x_weno(uint32_t* M, KArg a) { TE_TID3 flux_h_weno<0>(M, a, gx, gy, gz); }
__global__ __launch_bounds__(256) void k_adv_flux_y_weno(uint32_t* M, KArg a) { TE_TID3 flux_h_weno<1>(M, a, gx, gy, gz); }

// adv_flux_z_weno.wgsl:174-279 (face level k = kf - 1)
TE_DEV FacePre fp_z_weno(const Dims& g, const float* s, const float* rrw, const float* c1, const float* c2, int k,
                         int64_t c) {
  return fp_z5(g, s, rrw, c1, c2, k, c);
}
TE_DEV bool fe_z_weno(const Dims& g, const FacePre& p, float weps, int pdefweno, int i, int j, int k, float& out) {
  const int kf = k + 1;
  if (i >= g.nx || j >= g.ny || k < 0 || kf > g.nz + 1) return false;
  if (kf == 1 || kf == g.nz + 1) { out = 0.0f; return true; }
  const Win7& W = p.W;
#define L(n) W(n)
  const float w = p.a, ck1 = p.b, ck2 = p.c;
  const int nk = g.nz;
  if (kf == 2) {
    out = (w >= 0.0f) ? w * fma_(ck1, L(-1), ck2 * L(0)) : w * upstrpd(L(1), L(0), L(-1), weps);
    return true;
  }
  if (kf == nk) {
    out = (w > 0.0f) ? w * upstrpd(L(-2), L(-1), L(0), weps) : w * fma_(ck1, L(-1), ck2 * L(0));
    return true;
  }
  const Up5 o = up5<0>(W, w >= 0.0f);
  const float s1 = o.s1, s2 = o.s2, s3 = o.s3, s4 = o.s4, s5 = o.s5;
  if (pdefweno == 1) {
    bool doit = weno_pdef_ok(s1, s2, s3, s4, s5, weps, true);
    if (kf == 3 && w > 0.0f) doit = false;
    if (kf == nk - 1 && w < 0.0f) doit = false;
    if (doit) { out = w * weno5(s1, s2, s3, s4, s5, weps); return true; }
    if (kf != 3 && kf != nk - 1) { out = 0.0f; return true; }
  }
  if (kf == 3) {
    out = (w >= 0.0f) ? w * upstrpd(L(-2), L(-1), L(0), weps) : w * weno5(L(2), L(1), L(0), L(-1), L(-2), weps);
    return true;
  }
  if (kf == nk - 1) {
    out = (w >= 0.0f) ? w * weno5(L(-3), L(-2), L(-1), L(0), L(1), weps) : w * upstrpd(L(1), L(0), L(-1), weps);
    return true;
  }
  out = w * weno5(s1, s2, s3, s4, s5, weps);
  return true;
#undef L
}
TE_DEV bool fv_z_weno(const Dims& g, const float* s, const float* rrw, const float* c1, const float* c2, float weps,
                      int pdefweno, int i, int j, int k, int64_t c, float& out) {
  return fe_z_weno(g, fp_z_weno(g, s, rrw, c1, c2, k, c), weps, pdefweno, i, j, k, out);
}
__global__ __launch_bounds__(256) void k_adv_flux_z_weno(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  float f;
  if (fv_z_weno(g, fq(a.q[1]), fq(a.q[2]), fq(a.q[3]), fq(a.q[4]), ldf(a.q[5], 0), ldi(a.q[7], 0), gx, gy,
                gz, g.idx(gx, gy, gz), f))
    fq(a.q[8])[g.idx(gx, gy, gz)] = f;
}

TE_DEV float flx5_upwind(const float* s, int64_t c, int64_t st, bool up, bool shifted) {
  const Win7 W = hwin(s, c, st);
  return shifted ? flx5(up5<1>(W, up)) : flx5(up5<0>(W, up));
}
TE_DEV float weno5_upwind(const float* s, int64_t c, int64_t st, bool up, bool shifted, float weps) {
  const Win7 W = hwin(s, c, st);
  return shifted ? weno5(up5<1>(W, up), weps) : weno5(up5<0>(W, up), weps);
}

// adv_flux_x_stag.wgsl:144-186 / adv_flux_y_stag.wgsl (rr = rru for x, rrv for y)
// p.a, p.b: the mass flux at c and at the other point of the stagger's
// average (stagger 4: c - sk, with p.c = c2[k]); stagger 1: p.a only
template <int DIR>
TE_DEV FacePre fp_h_stag(const Dims& g, int stag, const float* s, const float* rr, const float* c2, int k, int64_t c) {
  const int64_t sj = g.sJ, sk = g.sk(), st = DIR == 0 ? 1 : sj;
  const int64_t o = stag == 2 ? (DIR == 0 ? 1 : -1) : stag == 3 ? (DIR == 0 ? -sj : sj) : stag == 4 ? -sk : 0;
  return {hwin(s, c, st), rr[c], rr[c + o], stag == 4 ? c2[lev(g, k)] : 0.0f};
}
template <int DIR>
TE_DEV bool fe_h_stag(const Dims& g, int stag, const FacePre& p, int i, int j, int k, float& out) {
  if (i > g.nx || j > g.ny || k >= g.nz) return false;
  if (stag == 4 && k == 0) { out = 0.0f; return true; }
  float m;
  if (stag == 2 || stag == 3) m = 0.5f * (p.a + p.b);
  else if (stag == 4) { const float cc2 = p.c; m = fma_(cc2, p.a, (1.0f - cc2) * p.b); }
  else m = p.a;
  const bool up = m >= 0.0f;
  out = m * (stag == (DIR == 0 ? 2 : 3) ? flx5(up5<1>(p.W, up)) : flx5(up5<0>(p.W, up)));
  return true;
}
template <int DIR>
TE_DEV bool fv_h_stag(const Dims& g, int stag, const float* s, const float* rr, const float* c2, int i, int j, int k, int64_t c,
                      float& out) {
  return fe_h_stag<DIR>(g, stag, fp_h_stag<DIR>(g, stag, s, rr, c2, k, c), i, j, k, out);
}
template <int DIR>
__device__ void flux_h_stag(uint32_t* M, const KArg& a, int i, int j, int k) {
  Dims g = dims(a.q[0], 0);
  float f;
  if (fv_h_stag<DIR>(g, ldi(a.q[1], 0), fq(a.q[2]), fq(a.q[DIR == 0 ? 3 : 4]), fq(a.q[7]), i, j, k, g.idx(i, j, k), f))
    fq(a.q[8])[g.idx(i, j, k)] = f;
}
__global__ __launch_bounds__(256) void k_adv_flux_x_stag(uint32_t* M, KArg a) { TE_TID3 flux_h_stag<0>(M, a, gx - 1, gy - 1, gz); }
__global__ __launch_bounds__(256) void k_adv_flux_y_stag(uint32_t* M, KArg a) { TE_TID3 flux_h_stag<1>(M, a, gx - 1, gy - 1, gz); }

// adv_flux_z_stag.wgsl:127-268
// p.a = m (the mass flux average), p.b, p.c = c1[k], c2[k]
TE_DEV FacePre fp_z_stag(const Dims& g, int stag, const float* s, const float* rrw, const float* c1, const float* c2, int k,
                         int64_t c) {
  const int64_t sj = g.sJ, sk = g.sk();
  const int64_t o = stag == 2 ? -1 : stag == 3 ? -sj : sk;
  const int l = lev(g, k);
  return {vwin(s, c, sk), 0.5f * (rrw[c] + rrw[c + o]), c1[l], c2[l]};
}
TE_DEV bool fe_z_stag(const Dims& g, int stag, const FacePre& p, int i, int j, int k, float& out) {
  if (i > g.nx || j > g.ny || k > g.nz) return false;
  const float m = p.a;
  const bool up = m >= 0.0f;
  const float cc1 = p.b, cc2 = p.c;
  const Win7& W = p.W;
#define A_(n) W(n)
  float f;
  if (stag == 4) {
    if (k > g.nz - 1) { out = 0.0f; return true; }
    if (k == 0) f = up ? 0.5f * (A_(0) + A_(1)) : flx3(A_(2), A_(1), A_(0));
    else if (k == 1) f = up ? flx3(A_(-1), A_(0), A_(1)) : flx5(A_(3), A_(2), A_(1), A_(0), A_(-1));
    else if (k == g.nz - 2) f = up ? flx5(A_(-2), A_(-1), A_(0), A_(1), A_(2)) : flx3(A_(2), A_(1), A_(0));
    else if (k == g.nz - 1) f = up ? flx3(A_(-1), A_(0), A_(1)) : 0.5f * (A_(0) + A_(1));
    else f = up ? flx5(A_(-2), A_(-1), A_(0), A_(1), A_(2)) : flx5(A_(3), A_(2), A_(1), A_(0), A_(-1));
  } else {
    if (k == 0 || k == g.nz) { out = 0.0f; return true; }
    if (k == 1) f = up ? fma_(cc1, A_(-1), cc2 * A_(0)) : flx3(A_(1), A_(0), A_(-1));
    else if (k == 2) f = up ? flx3(A_(-2), A_(-1), A_(0)) : flx5(A_(2), A_(1), A_(0), A_(-1), A_(-2));
    else if (k == g.nz - 2) f = up ? flx5(A_(-3), A_(-2), A_(-1), A_(0), A_(1)) : flx3(A_(1), A_(0), A_(-1));
    else if (k == g.nz - 1) f = up ? flx3(A_(-2), A_(-1), A_(0)) : fma_(cc1, A_(-1), cc2 * A_(0));
    else f = up ? flx5(A_(-3), A_(-2), A_(-1), A_(0), A_(1)) : flx5(A_(2), A_(1), A_(0), A_(-1), A_(-2));
  }
#undef A_
  out = m * f;
  return true;
}
TE_DEV bool fv_z_stag(const Dims& g, int stag, const float* s, const float* rrw, const float* c1, const float* c2, int i,
                      int j, int k, int64_t c, float& out) {
  return fe_z_stag(g, stag, fp_z_stag(g, stag, s, rrw, c1, c2, k, c), i, j, k, out);
}
__global__ __launch_bounds__(256) void k_adv_flux_z_stag(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  float f;
  if (fv_z_stag(g, ldi(a.q[1], 0), fq(a.q[2]), fq(a.q[5]), fq(a.q[6]), fq(a.q[7]), gx, gy, gz, g.idx(gx, gy, gz), f))
    fq(a.q[8])[g.idx(gx, gy, gz)] = f;
}

// adv_flux_x_weno_w.wgsl:143-175 / adv_flux_y_weno_w.wgsl
// p.a, p.b = rr at c and c - sk, p.c = c2[k]
template <int DIR>
TE_DEV FacePre fp_h_weno_w(const Dims& g, const float* s, const float* rr, const float* c2, int k, int64_t c) {
  return {hwin(s, c, DIR == 0 ? 1 : (int64_t)g.sJ), rr[c], rr[c - g.sk()], c2[lev(g, k)]};
}
template <int DIR>
TE_DEV bool fe_h_weno_w(const Dims& g, const FacePre& p, float weps, int i, int j, int k, float& out) {
  if (i < 0 || j < 0 || i > g.nx || j > g.ny || k >= g.nz) return false;
  if (k == 0) { out = 0.0f; return true; }
  float cc2 = p.c, cc1 = 1.0f - cc2;
  float ub = fma_(cc2, p.a, cc1 * p.b);
  out = ub * weno5(up5<0>(p.W, ub >= 0.0f), weps);
  return true;
}
template <int DIR>
TE_DEV bool fv_h_weno_w(const Dims& g, const float* s, const float* rr, const float* c2, float weps, int i, int j, int k, int64_t c,
                        float& out) {
  return fe_h_weno_w<DIR>(g, fp_h_weno_w<DIR>(g, s, rr, c2, k, c), weps, i, j, k, out);
}
template <int DIR>
__device__ void flux_h_weno_w(uint32_t* M, const KArg& a, int i, int j, int k) {
  Dims g = dims(a.q[0], 0);
  float f;
  if (fv_h_weno_w<DIR>(g, fq(a.q[1]), fq(a.q[2]), fq(a.q[3]), ldf(a.q[4], 0), i, j, k, g.idx(i, j, k), f))
    fq(a.q[5])[g.idx(i, j, k)] = f;
}
__global__ __launch_bounds__(256) void k_adv_flux_x_weno_w(uint32_t* M, KArg a) { TE_TID3 flux_h_weno_w<0>(M, a, gx - 1, gy - 1, gz); }
__global__ __launch_bounds__(256) void k_adv_flux_y_weno_w(uint32_t* M, KArg a) { TE_TID3 flux_h_weno_w<1>(M, a, gx - 1, gy - 1, gz); }

// adv_flux_z_weno_w.wgsl:159-231
TE_DEV FacePre fp_z_weno_w(const Dims& g, const float* s, const float* rrw, int64_t c) {
  const int64_t sk = g.sk();
  return {vwin(s, c, sk), 0.5f * (rrw[c] + rrw[c + sk]), 0.0f, 0.0f};
}
TE_DEV bool fe_z_weno_w(const Dims& g, const FacePre& p, float weps, int i, int j, int k, float& out) {
  if (i > g.nx || j > g.ny || k > g.nz) return false;
  if (k > g.nz - 1) { out = 0.0f; return true; }
  const Win7& W = p.W;
#define L(n) W(n)
  float wbar = p.a;
  const int nz = g.nz;
  float f;
  if (k == 0) f = (wbar >= 0.0f) ? 0.5f * (L(0) + L(1)) : upstrpd(L(2), L(1), L(0), weps);
  else if (k == nz - 1) f = (wbar > 0.0f) ? upstrpd(L(-1), L(0), L(1), weps) : 0.5f * (L(0) + L(1));
  else if (k == 1) f = (wbar >= 0.0f) ? upstrpd(L(-1), L(0), L(1), weps) : weno5(L(3), L(2), L(1), L(0), L(-1), weps);
  else if (k == nz - 2) f = (wbar > 0.0f) ? weno5(L(-2), L(-1), L(0), L(1), L(2), weps) : upstrpd(L(2), L(1), L(0), weps);
  else f = (wbar >= 0.0f) ? weno5(L(-2), L(-1), L(0), L(1), L(2), weps) : weno5(L(3), L(2), L(1), L(0), L(-1), weps);
#undef L
  out = wbar * f;
  return true;
}
TE_DEV bool fv_z_weno_w(const Dims& g, const float* s, const float* rrw, float weps, int i, int j, int k, int64_t c, float& out) {
  return fe_z_weno_w(g, fp_z_weno_w(g, s, rrw, c), weps, i, j, k, out);
}
__global__ __launch_bounds__(256) void k_adv_flux_z_weno_w(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  float f;
  if (fv_z_weno_w(g, fq(a.q[1]), fq(a.q[2]), ldf(a.q[3], 0), gx, gy, gz, g.idx(gx, gy, gz), f)) fq(a.q[4])[g.idx(gx, gy, gz)] = f;
}

// adv_flux_x_weno_uv.wgsl:168-188 / adv_flux_y_weno_uv.wgsl
template <int DIR>
__device__ void flux_h_weno_uv(uint32_t* M, const KArg& a, int i, int j, int k) {
  Dims g = dims(a.q[0], 0);
  if (i > g.nx || j > g.ny || k >= g.nz) return;
  int stag = ldi(a.q[1], 0);
  const float *s = fq(a.q[2]), *rr = fq(a.q[3]);
  float weps = ldf(a.q[4], 0);
  const int64_t sj = g.sJ, st = DIR == 0 ? 1 : sj, c = g.idx(i, j, k);
  float m;
  if (DIR == 0) m = (stag == 2) ? 0.5f * (rr[c] + rr[c + 1]) : 0.5f * (rr[c] + rr[c - sj]);
  else m = (stag == 2) ? 0.5f * (rr[c] + rr[c - 1]) : 0.5f * (rr[c] + rr[c + sj]);
  fq(a.q[5])[c] = m * weno5_upwind(s, c, st, m >= 0.0f, stag == (DIR == 0 ? 2 : 3), weps);
}
__global__ __launch_bounds__(256) void k_adv_flux_x_weno_uv(uint32_t* M, KArg a) { TE_TID3 flux_h_weno_uv<0>(M, a, gx - 1, gy - 1, gz); }
__global__ __launch_bounds__(256) void k_adv_flux_y_weno_uv(uint32_t* M, KArg a) { TE_TID3 flux_h_weno_uv<1>(M, a, gx - 1, gy - 1, gz); }

// adv_flux_z_weno_uv.wgsl:159-236
__global__ __launch_bounds__(256) void k_adv_flux_z_weno_uv(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz;
  if (i > g.nx || j > g.ny || k > g.nz) return;
  int stag = ldi(a.q[1], 0);
  const float *s = fq(a.q[2]), *rrw = fq(a.q[3]), *c1 = fq(a.q[4]), *c2 = fq(a.q[5]);
  float weps = ldf(a.q[6], 0);
  float* fz = fq(a.q[7]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  if (k == 0 || k == g.nz) { fz[c] = 0.0f; return; }
#define L(n) s[c + (n) * sk]
  float m = (stag == 2) ? 0.5f * (rrw[c] + rrw[c - 1]) : 0.5f * (rrw[c] + rrw[c - sj]);
  bool up = m >= 0.0f;
  float f;
  if (k == 1) f = up ? fma_(c1[k], L(-1), c2[k] * L(0)) : upstrpd(L(1), L(0), L(-1), weps);
  else if (k == 2) f = up ? upstrpd(L(-2), L(-1), L(0), weps) : weno5(L(2), L(1), L(0), L(-1), L(-2), weps);
  else if (k == g.nz - 2) f = up ? weno5(L(-3), L(-2), L(-1), L(0), L(1), weps) : upstrpd(L(1), L(0), L(-1), weps);
  else if (k == g.nz - 1) f = up ? upstrpd(L(-2), L(-1), L(0), weps) : fma_(c1[k], L(-1), c2[k] * L(0));
  else f = up ? weno5(L(-3), L(-2), L(-1), L(0), L(1), weps) : weno5(L(2), L(1), L(0), L(-1), L(-2), weps);
#undef L
  fz[c] = m * f;
}

// adv_assemble.wgsl:139-279: flux divergence + mass-divergence term, divided
// by the reference density; open-boundary inflow corrections included.
// STC != 0: the stagger is known at compile time (specialised launches).
// c: index of (i, j, k).
// The field at the cell and the two mass-divergence values the assembly
// averages (divx at c and at the stagger's other point), loaded ahead of the
// assembly's branches.
struct AsmPre {
  float s, d0, d1;
};
template <int STC = 0>
TE_DEV AsmPre asm_pre(const KArg& a, const Dims& g, int64_t c) {
  const int stag = STC != 0 ? STC : ldi(a.q[1], 0), st = stag < 0 ? -stag : stag;
  const float *s = fq(a.q[2]), *divx = fq(a.q[6]);
  const int64_t nb = st == 2 ? c - 1 : st == 3 ? c - g.sJ : st == 4 ? c - g.sk() : c;
  return {s[c], divx[c], divx[nb]};
}
template <int STC = 0, class FLUX, class STORE>
TE_DEV void assemble_at(uint32_t* M, const KArg& a, const Dims& g, int i, int j, int k, int64_t c, const AsmPre& P,
                        FLUX F, STORE store) {
  const int stag = STC != 0 ? STC : ldi(a.q[1], 0);
  int st = stag < 0 ? -stag : stag;
  int iMax = st == 2 ? g.nx : g.nx - 1, jMax = st == 3 ? g.ny : g.ny - 1, kLo = st == 4 ? 1 : 0, kMax = g.nz - 1;
  if (i > iMax || j > jMax || k < kLo || k > kMax) return;
  int openx = ldi(a.q[17], 0), openy = ldi(a.q[17], 1);
  if (openx != 0 && st == 2 && (i == 0 || i == g.nx)) return;
  if (openy != 0 && st == 3 && (j == 0 || j == g.ny)) return;
  const float *rho0 = fq(a.q[7]), *rf0 = fq(a.q[8]), *mh = fq(a.q[9]), *mf = fq(a.q[10]),
              *rru = fq(a.q[13]), *rrv = fq(a.q[14]), *c1 = fq(a.q[15]), *c2 = fq(a.q[16]),
              *hm = fq(a.q[30]);
  float* tend = fq(a.q[11]);
  int accumulate = ldi(a.q[12], 0);
  const int64_t sj = g.sJ, sk = g.sk();
  float dFx = st == 2 ? F(0, i, j, k) - F(0, i - 1, j, k) : F(0, i + 1, j, k) - F(0, i, j, k);
  float dFy = st == 3 ? F(1, i, j, k) - F(1, i, j - 1, k) : F(1, i, j + 1, k) - F(1, i, j, k);
  float dFz = st == 4 ? F(2, i, j, k) - F(2, i, j, k - 1) : F(2, i, j, k + 1) - F(2, i, j, k);
  float mz = stag < 0 ? 1.0f : (stag == 4 ? mf[k] : mh[k]);
  float dv;  // divx at c and at c - 1, c - sj or c - sk
  if (st == 2 || st == 3) dv = 0.5f * (P.d0 + P.d1);
  else if (st == 4) dv = fma_(c2[k], P.d0, c1[k] * P.d1);
  else dv = P.d0;
  if (openx != 0 && st != 2) {
    bool west = i == 0, east = i == g.nx - 1;
    if (west || east) {
      int64_t cf = west ? c : c + 1;
      float un, piece;
      if (st == 3) {
        un = rru[cf] + rru[cf - sj];
        piece = ((0.5f * ((rru[c + 1 - sj] - rru[c - sj]) + (rru[c + 1] - rru[c]))) * g.rdx) * hm_uh(hm, g, i);
      } else if (st == 4) {
        float cc2 = 0.5f * (c2[k - 1] + c2[k]), cc1 = 1.0f - cc2;
        un = fma_(cc2, rru[cf], cc1 * rru[cf - sk]);
        piece = (fma_(c1[k], rru[c + 1 - sk] - rru[c - sk], c2[k] * (rru[c + 1] - rru[c])) * g.rdx) * hm_uh(hm, g, i);
      } else {
        un = rru[cf];
        piece = ((rru[c + 1] - rru[c]) * g.rdx) * hm_uh(hm, g, i);
      }
      bool inflow = west ? un >= 0.0f : un <= 0.0f;
      if (inflow) { dFx = 0.0f; dv = dv - piece; }
    }
  }
  if (openy != 0 && st != 3) {
    bool south = j == 0, north = j == g.ny - 1;
    if (south || north) {
      int64_t cf = south ? c : c + sj;
      float vn, piece;
      if (st == 2) {
        vn = rrv[cf] + rrv[cf - 1];
        piece = ((0.5f * ((rrv[c + sj - 1] - rrv[c - 1]) + (rrv[c + sj] - rrv[c]))) * g.rdy) * hm_vh(hm, g, j);
      } else if (st == 4) {
        float cc2 = 0.5f * (c2[k - 1] + c2[k]), cc1 = 1.0f - cc2;
        vn = fma_(cc2, rrv[cf], cc1 * rrv[cf - sk]);
        piece = (fma_(c1[k], rrv[c + sj - sk] - rrv[c - sk], c2[k] * (rrv[c + sj] - rrv[c])) * g.rdy) * hm_vh(hm, g, j);
      } else {
        vn = rrv[cf];
        piece = ((rrv[c + sj] - rrv[c]) * g.rdy) * hm_vh(hm, g, j);
      }
      bool inflow = south ? vn >= 0.0f : vn <= 0.0f;
      if (inflow) { dFy = 0.0f; dv = dv - piece; }
    }
  }
  float mxm = st == 2 ? hm_uf(hm, g, i) : hm_uh(hm, g, i);
  float mym = st == 3 ? hm_vf(hm, g, j) : hm_vh(hm, g, j);
  float divF = fma_(dFz * g.rdz, mz, fma_(dFx * g.rdx, mxm, (dFy * g.rdy) * mym));
  float rho = st == 4 ? rf0[k] : rho0[k];
  float t = fma_(P.s, dv, -divF) / rho;
  store(c, t, accumulate);
}

__global__ __launch_bounds__(256) void k_adv_assemble(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  const float *fx = fq(a.q[3]), *fy = fq(a.q[4]), *fz = fq(a.q[5]);
  float* tend = fq(a.q[11]);
  const int64_t c0 = g.idx(gx, gy, gz);
  assemble_at(
      M, a, g, gx, gy, gz, c0, asm_pre(a, g, c0),
      [&](int d, int i, int j, int k) {
        const int64_t c = g.idx(i, j, k);
        return d == 0 ? fx[c] : (d == 1 ? fy[c] : fz[c]);
      },
      [&](int64_t c, float t, int acc) { tend[c] = acc != 0 ? tend[c] + t : t; });
}

// The three flux dispatches of one advected field and adv_assemble in one
// launch: each cell evaluates its six face fluxes with the flux kernels' own
// value functions instead of reading them back from fx/fy/fz (which are then
// left unwritten; they are dead scratch).  A face is the same value whichever
// cell evaluates it; the two faces of a direction share one stencil window,
// so evaluating both costs arithmetic, not loads (measured faster than
// publishing one face per cell through LDS, which needs a barrier between
// the loads of the two halves).
// p: adv_assemble's parameters, with 3 -> flux kind (0 flx5, 1 WENO5,
// 2 staggered flx5, 3 WENO5 of w-staggered fields), 4 rrw, 5 -> weps,
// 20 -> pdef, 21 -> pdefweno (22-24: fx, fy, fz of the unfused dispatches).
template <int KIND>
struct FusedFlux {
  Dims g;
  int stag, pdef, pdw;
  const float *s, *rru, *rrv, *rrw, *c1, *c2;
  float weps;
  // loads of the face of direction D at index c (level k: the face's)
  template <int D>
  __device__ __forceinline__ FacePre pre(int k, int64_t c) const {
    if (KIND == 0)
      return D == 0 ? fp_h5<0>(g, s, rru, c) : D == 1 ? fp_h5<1>(g, s, rrv, c) : fp_z5(g, s, rrw, c1, c2, k, c);
    if (KIND == 1)
      return D == 0 ? fp_h_weno<0>(g, s, rru, c)
                    : D == 1 ? fp_h_weno<1>(g, s, rrv, c) : fp_z_weno(g, s, rrw, c1, c2, k, c);
    if (KIND == 2)
      return D == 0 ? fp_h_stag<0>(g, stag, s, rru, c2, k, c)
                    : D == 1 ? fp_h_stag<1>(g, stag, s, rrv, c2, k, c) : fp_z_stag(g, stag, s, rrw, c1, c2, k, c);
    return D == 0 ? fp_h_weno_w<0>(g, s, rru, c2, k, c)
                  : D == 1 ? fp_h_weno_w<1>(g, s, rrv, c2, k, c) : fp_z_weno_w(g, s, rrw, c);
  }
  // face flux of direction D at (i, j, k) from its loads; NaN where no flux
  // dispatch writes
  template <int D>
  __device__ __forceinline__ float eval(const FacePre& p, int i, int j, int k) const {
    float f = __builtin_nanf("");
    bool ok;
    if (KIND == 0)
      ok = D == 0 ? fe_h5<0>(g, p, i, j, k, f) : D == 1 ? fe_h5<1>(g, p, i, j, k, f) : fe_z5(g, p, i, j, k, f);
    else if (KIND == 1)
      ok = D == 0 ? fe_h_weno<0>(g, p, weps, pdef == 1, i, j, k, f)
                  : D == 1 ? fe_h_weno<1>(g, p, weps, pdef == 1, i, j, k, f) : fe_z_weno(g, p, weps, pdw, i, j, k, f);
    else if (KIND == 2)
      ok = D == 0 ? fe_h_stag<0>(g, stag, p, i, j, k, f)
                  : D == 1 ? fe_h_stag<1>(g, stag, p, i, j, k, f) : fe_z_stag(g, stag, p, i, j, k, f);
    else
      ok = D == 0 ? fe_h_weno_w<0>(g, p, weps, i, j, k, f)
                  : D == 1 ? fe_h_weno_w<1>(g, p, weps, i, j, k, f) : fe_z_weno_w(g, p, weps, i, j, k, f);
    return ok ? f : __builtin_nanf("");  // a face no flux dispatch writes: poisons the result visibly
  }
};

template <int KIND, int STC>
__device__ __forceinline__ void adv_fused_body(uint32_t* M, const KArg& a) {
  // launched as 64 x Y x Z blocks (Y * Z <= 4; runtime launch(), te_autotune):
  // a wave is one row of cells, so j and k are wave-uniform and branches on
  // them are scalar
  const Blk b_ = te_block(a);
  const int i = b_.x * 64 + (int)threadIdx.x,
            j = b_.y * TE_BY + (int)__builtin_amdgcn_readfirstlane(threadIdx.y),
            k = b_.z * TE_BZ + (int)__builtin_amdgcn_readfirstlane(threadIdx.z);
  if (j >= (int)a.n[1] || k >= (int)a.n[2] || i >= (int)a.n[0]) return;
  Dims g = dims(a.q[0], 0);
  const int stag = STC != 0 ? STC : ldi(a.q[1], 0), st = stag < 0 ? -stag : stag;
  const FusedFlux<KIND> F{g, stag, ldi(a.q[20], 0), ldi(a.q[21], 0), fq(a.q[2]), fq(a.q[13]), fq(a.q[14]),
                          fq(a.q[4]), fq(a.q[15]), fq(a.q[16]), ldf(a.q[5], 0)};
  // faces of this cell's divergence: low and high in x, y and z, indexed
  // from the cell's own index (one index computation per cell)
  const int64_t c = g.idx(i, j, k), sj = g.sJ, sk = g.sk();
  const int xlo = st == 2 ? i - 1 : i, ylo = st == 3 ? j - 1 : j, zlo = st == 4 ? k - 1 : k;
  const int64_t cx = st == 2 ? c - 1 : c, cy = st == 3 ? c - sj : c, cz = st == 4 ? c - sk : c;
  // every face's loads first, with the assembly's and the chain's operands
  // (below), then the faces (one wait for memory)
  const FacePre pxl = F.template pre<0>(k, cx), pxh = F.template pre<0>(k, cx + 1);
  const FacePre pyl = F.template pre<1>(k, cy), pyh = F.template pre<1>(k, cy + sj);
  const FacePre pzl = F.template pre<2>(zlo, cz), pzh = F.template pre<2>(zlo + 1, cz + sk);
  const AsmPre P = asm_pre<STC>(a, g, c);
  const int chain = ldi(a.q[32], 0), acc = ldi(a.q[12], 0);
  float* tend = fq(a.q[11]);
  const float t_old = (chain & 16) == 0 && acc != 0 ? tend[c] : 0.0f;
  const float d3 = (chain & 1) ? fq(a.q[33])[c] : 0.0f, th3 = (chain & 1) ? fq(a.q[34])[c] : 0.0f;
  const float add1 = (chain & 2) ? fq(a.q[35])[c] : 0.0f, add2 = (chain & 4) ? fq(a.q[37])[c] : 0.0f;
  const float phi_a = (chain & 8) ? fq(a.q[39])[c] : 0.0f;
  const bool buoy = (chain & 64) && k >= 1 && k <= g.nz - 1;
  const float bm = buoy ? fq(a.q[42])[c - sk] : 0.0f, b0 = buoy ? fq(a.q[42])[c] : 0.0f;
  const float bc1 = buoy ? fq(a.q[15])[k] : 0.0f, bc2 = buoy ? fq(a.q[16])[k] : 0.0f;
  const float fxl = F.template eval<0>(pxl, xlo, j, k), fxh = F.template eval<0>(pxh, xlo + 1, j, k);
  const float fyl = F.template eval<1>(pyl, i, ylo, k), fyh = F.template eval<1>(pyh, i, ylo + 1, k);
  const float fzl = F.template eval<2>(pzl, i, j, zlo), fzh = F.template eval<2>(pzh, i, j, zlo + 1);
  // p[32] -> chain flags: the pointwise dispatches around a field's assembly
  // in the reference, applied to the tendency in registers:
  // 16 turb_zero of the accumulating sink before it (tend = 0 + t),
  // 1 apply_thterm (tend = fma(-div3d, thterm, tend); p33 div3d, p34 thterm),
  // 2 / 4 turb_add_tendency (tend = fma(scale, src, tend); p35/p36, p37/p38),
  // 8 rk_update (phi = fma(dttmp, tend, phi_a); p39 phi_a, p40 phi, p41 dttmp;
  // phi must not be the advected field, which other workgroups still read),
  // 32 with 8 on a periodic domain (nx, ny >= 4, cell-centred field): the
  // halo fill of phi after it, each cell's periodic images written by the
  // cell, level nz's by the cell below it (no thread of the launch reads phi),
  // 64 buoyancy_apply after the assembly of w (tend = tend + fma(c1[k],
  // buoy[k-1], c2[k] * buoy[k]) at levels 1 .. nz-1; p42 buoy, p15 c1, p16 c2).
  // With an update in the chain the tendency itself is not stored.
  assemble_at<STC>(
      M, a, g, i, j, k, c, P,
      [&](int d, int fi, int fj, int fk) {
        if (d == 0) return fi == xlo ? fxl : fxh;
        if (d == 1) return fj == ylo ? fyl : fyh;
        return fk == zlo ? fzl : fzh;
      },
      [&](int64_t c, float t, int acc) {
        float v = (chain & 16) ? 0.0f + t : (acc != 0 ? t_old + t : t);
        if (chain & 1) v = fma_(-d3, th3, v);
        if (chain & 2) v = fma_(ldf(a.q[36], 0), add1, v);
        if (chain & 4) v = fma_(ldf(a.q[38], 0), add2, v);
        if (buoy) v = v + fma_(bc1, bm, bc2 * b0);
        if (chain & 8) {
          float* phi = fq(a.q[40]);
          const float nv = fma_(ldf(a.q[41], 0), v, phi_a);
          phi[c] = nv;
          if (chain & 32) {
            const bool top = k == g.nz - 1;
            const float ptop = top ? phi[c + sk] : 0.0f;
            periodic_images(g, i, j, [&](int64_t o) {
              phi[c + o] = nv;
              if (top) phi[c + sk + o] = ptop;
            });
          }
        } else {
          tend[c] = v;
        }
      });
}

// One kernel per specialised body (adv_fused_kernel), so that each carries
// only its own registers; k_adv_fused selects the body at run time.
template <int KIND, int STC>
__global__ __launch_bounds__(256) void k_adv_fused_v(uint32_t* M, KArg a) {
  adv_fused_body<KIND, STC>(M, a);
}
const void* adv_fused_kernel(int kind, int stag) {
  if (kind == 0 && stag == 1) return (const void*)&k_adv_fused_v<0, 1>;
  if (kind == 1 && stag == 1) return (const void*)&k_adv_fused_v<1, 1>;
  if (kind == 2 && stag == 2) return (const void*)&k_adv_fused_v<2, 2>;
  if (kind == 2 && stag == 3) return (const void*)&k_adv_fused_v<2, 3>;
  if (kind == 2 && stag == 4) return (const void*)&k_adv_fused_v<2, 4>;
  if (kind == 3 && stag == 4) return (const void*)&k_adv_fused_v<3, 4>;
  return nullptr;
}

// The three flux dispatches of one advected field and adv_assemble in one
// launch (see above); the flux kind selects one specialised body.
__global__ __launch_bounds__(256) void k_adv_fused(uint32_t* M, KArg a) {
  const int kind = ldi(a.q[3], 0), stag = ldi(a.q[1], 0);
  // the (kind, stagger) pairs the encoder emits get bodies specialised on both
  if (kind == 0 && stag == 1) adv_fused_body<0, 1>(M, a);
  else if (kind == 1 && stag == 1) adv_fused_body<1, 1>(M, a);
  else if (kind == 2 && stag == 2) adv_fused_body<2, 2>(M, a);
  else if (kind == 2 && stag == 3) adv_fused_body<2, 3>(M, a);
  else if (kind == 2 && stag == 4) adv_fused_body<2, 4>(M, a);
  else if (kind == 3 && stag == 4) adv_fused_body<3, 4>(M, a);
  else if (kind == 0) adv_fused_body<0, 0>(M, a);
  else if (kind == 1) adv_fused_body<1, 0>(M, a);
  else if (kind == 2) adv_fused_body<2, 0>(M, a);
  else adv_fused_body<3, 0>(M, a);
}

}  // namespace te
