// Acoustic substep kernels (one per reference kernel) and the
// positive-definite flux correction.
#include "te_device.hpp"
#include "te_kernels.hpp"

namespace te {

// acoustic_prepare.wgsl:130-184 (FMA operand order switches with
// vec_v = nx>3 && i<(nx/4)*4, as in the reference)
__global__ __launch_bounds__(256) void k_acoustic_prepare(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  Acoustic ac = acoustic(a.q[1], 0);
  const float *ppterm = fq(a.q[2]), *piadv = fq(a.q[3]), *thv = fq(a.q[4]), *rr0 = fq(a.q[5]),
              *rf0 = fq(a.q[6]), *mh = fq(a.q[7]), *mf = fq(a.q[8]), *c1 = fq(a.q[9]), *c2 = fq(a.q[10]);
  float *qk = fq(a.q[11]), *pk1 = fq(a.q[12]), *pk2 = fq(a.q[13]), *mm = fq(a.q[14]);
  const int64_t sk = g.sk(), c = g.idx(i, j, k);
  float tem = (ac.dts * g.rdz) * mh[k];
  float r2 = (tem * rr0[k]) * rf0[k + 1];
  float r1 = (tem * rr0[k]) * rf0[k];
  float tem1 = ((ac.dts * g.rdz) * CP) * mf[k];
  float c1a = c1[k], c1b = c1[k + 1], c2a = c2[k], c2b = c2[k + 1];
  bool vec_v = g.nx > 3 && i < (g.nx / 4) * 4;
  qk[c] = tem * ppterm[c];
  float p2 = 0.0f, p1 = 0.0f;
  if (k != g.nz - 1)
    p2 = vec_v ? fma_(-c2b, piadv[c + sk], (1.0f - c1b) * piadv[c]) : fma_(1.0f - c1b, piadv[c], -(c2b * piadv[c + sk]));
  if (k != 0)
    p1 = vec_v ? fma_(c2a - 1.0f, piadv[c], c1a * piadv[c - sk]) : fma_(c1a, piadv[c - sk], (c2a - 1.0f) * piadv[c]);
  pk2[c] = k == g.nz - 1 ? 0.0f : r2 * p2;
  pk1[c] = k == 0 ? 0.0f : r1 * p1;
  mm[c] = tem1 * fma_(c2a, thv[c], c1a * (k == 0 ? thv[c] : thv[c - sk]));
}

// acoustic_begin.wgsl:128-149
__global__ __launch_bounds__(256) void k_acoustic_begin(uint32_t* M, KArg a) {
  if (te_flat(a, [&](uint32_t w, int width) {
        const float beta = acoustic(a.q[1], 0).beta;
        F4 z, p = ld4(a.q[2], w, width);
        for (int e = 0; e < 4; ++e) z.v[e] = 0.0f;
        for (int e = 0; e < width; ++e) p.v[e] = p.v[e] * beta;
        st4(a.q[3], w, width, z);
        st4(a.q[4], w, width, z);
        st4(a.q[5], w, width, p);
      }))
    return;
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  Acoustic ac = acoustic(a.q[1], 0);
  const int64_t c = g.idx(i, j, k);
  fq(a.q[3])[c] = 0.0f;
  fq(a.q[4])[c] = 0.0f;
  fq(a.q[5])[c] = fq(a.q[2])[c] * ac.beta;
}

// acoustic_uv.wgsl:144-181
__global__ __launch_bounds__(256) void k_acoustic_uv(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz;
  if (i > g.nx || j > g.ny || k >= g.nz) return;
  Acoustic ac = acoustic(a.q[1], 0);
  const float *ppd = fq(a.q[2]), *thv = fq(a.q[3]), *uten = fq(a.q[4]), *vten = fq(a.q[5]),
              *hm = fq(a.q[30]);
  float *u = fq(a.q[6]), *v = fq(a.q[7]);
  int openx = ldi(a.q[8], 0), openy = ldi(a.q[8], 1);
  const int64_t sj = g.sJ, c = g.idx(i, j, k);
  float tem1 = (g.rdx * CP) * 0.5f;
  float tem2 = (g.rdy * CP) * 0.5f;
  bool uEdge = openx != 0 && (i == 0 || i == g.nx);
  bool vEdge = openy != 0 && (j == 0 || j == g.ny);
  if (!uEdge) u[c] = fma_(ac.dts, fma_(-((tem1 * (ppd[c] - ppd[c - 1])) * hm_uf(hm, g, i)), thv[c] + thv[c - 1], uten[c]), u[c]);
  if (!vEdge) v[c] = fma_(ac.dts, fma_(-((tem2 * (ppd[c] - ppd[c - sj])) * hm_vf(hm, g, j)), thv[c] + thv[c - sj], vten[c]), v[c]);
}

// acoustic_fpk.wgsl:167-207
__global__ __launch_bounds__(256) void k_acoustic_fpk(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  Acoustic ac = acoustic(a.q[1], 0);
  const float *u = fq(a.q[2]), *v = fq(a.q[3]), *w = fq(a.q[4]), *piadv = fq(a.q[5]),
              *ppten = fq(a.q[6]), *ppterm = fq(a.q[7]), *pk1 = fq(a.q[8]), *pk2 = fq(a.q[9]),
              *qk = fq(a.q[10]), *hm = fq(a.q[30]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  float temx = (ac.dts * 0.5f) * g.rdx;
  float temy = (ac.dts * 0.5f) * g.rdy;
  float div = fma_((u[c + 1] - u[c]) * g.rdx, hm_uh(hm, g, i), ((v[c + sj] - v[c]) * g.rdy) * hm_vh(hm, g, j));
  if (ab(div) < ac.smeps) div = 0.0f;
  float Au = fma_(u[c], piadv[c] - piadv[c - 1], u[c + 1] * (piadv[c + 1] - piadv[c]));
  float Av = fma_(v[c], piadv[c] - piadv[c - sj], v[c + sj] * (piadv[c + sj] - piadv[c]));
  float advPi = fma_((-Au) * temx, hm_uh(hm, g, i), -((Av * temy) * hm_vh(hm, g, j)));
  fq(a.q[11])[c] = fma_(ac.beta, fma_(w[c], pk1[c] + qk[c], w[c + sk] * (pk2[c] - qk[c])),
                           fma_(ac.dts, fma_(-ppterm[c], div, ppten[c]), advPi));
}

// acoustic_w_solve.wgsl:124-212: one lane per (i,j) column, Thomas sweep over
// k with the elimination coefficients in cA/cB; lanes stay coalesced along i.
// The sweeps run in chunks of WS_CHUNK levels: every load of a chunk is
// issued before its serial recurrence, so one wave keeps ~8 levels of loads
// in flight instead of waiting on each level.  Sums of level k-1 (pk1+qk,
// pk2-qk, pp3d, pp3d+fpk) carry in registers: they are the same operations on
// the same values as the reference's re-evaluation at level k.
#ifndef WS_CHUNK
#define WS_CHUNK 8
#endif
template <bool LDS>
__device__ void w_solve(uint32_t* M, const KArg& a) {
  const Blk b_ = te_block(a);
  const int i = b_.x * TE_BX + (int)threadIdx.x, j = b_.y * TE_BY + (int)threadIdx.y;
  if (i >= (int)a.n[0] || j >= (int)a.n[1]) return;
  Dims g = dims(a.q[0], 0);
  if (i >= g.nx || j >= g.ny) return;
  Acoustic ac = acoustic(a.q[1], 0);
  const float *pk1 = fq(a.q[2]), *pk2 = fq(a.q[3]), *qk = fq(a.q[4]), *mm = fq(a.q[5]),
              *fpk = fq(a.q[6]), *wten = fq(a.q[7]), *pp3d = fq(a.q[8]);
  float *w = fq(a.q[9]), *cA = fq(a.q[10]), *cB = fq(a.q[11]);
  const int nz = g.nz;
  const int64_t sk = g.sk(), base = g.idx(i, j, 0);
  const float asq = ac.alph * ac.alph;
  const ColStore<LDS> CA{cA, base, sk, TE_LANE, TE_NLANES, 0}, CB{cB, base, sk, TE_LANE, TE_NLANES, nz};
  // level 0 values (read before w[0] is cleared: w is not among them)
  float pq_m = pk1[base] + qk[base], p2q_m = pk2[base] - qk[base], pp_m = pp3d[base];
  float ppf_m = pp_m + fpk[base];
  w[base] = 0.0f;
  w[base + nz * sk] = 0.0f;
  float pa = 0.0f, pb = 0.0f;
  for (int k0 = 1; k0 <= nz - 1; k0 += WS_CHUNK) {
    float m[WS_CHUNK], a1[WS_CHUNK], q[WS_CHUNK], a2[WS_CHUNK], wt[WS_CHUNK], p[WS_CHUNK], f[WS_CHUNK], wc[WS_CHUNK];
#pragma unroll
    for (int t = 0; t < WS_CHUNK; ++t) {
      const int k = k0 + t;
      if (k <= nz - 1) {
        const int64_t c = base + k * sk;
        m[t] = mm[c]; a1[t] = pk1[c]; q[t] = qk[c]; a2[t] = pk2[c];
        wt[t] = wten[c]; p[t] = pp3d[c]; f[t] = fpk[c]; wc[t] = w[c];
      }
    }
#pragma unroll
    for (int t = 0; t < WS_CHUNK; ++t) {
      const int k = k0 + t;
      if (k <= nz - 1) {
        const int64_t c = base + k * sk;
        const float pq = a1[t] + q[t], p2q = a2[t] - q[t], ppf = p[t] + f[t];
        const float aa = (-asq) * (m[t] * pq_m);
        const float cc = (-asq) * ((-m[t]) * p2q);
        const float bb = fma_(-asq, m[t] * ((-pq) + p2q_m), 1.0f);
        const float fw = fma_(ac.dts, wt[t], ac.beta * ((-m[t]) * (p[t] - pp_m)));
        const float dd = fma_(ac.alph, (-m[t]) * (ppf - ppf_m), wc[t] + fw);
        if (k == 1) {
          const float r = 1.0f / bb;
          pa = (-cc) * r;
          pb = dd * r;
          CA.at(k) = pa;
          CB.at(k) = pb;
        } else if (k < nz - 1) {
          const float r = 1.0f / fma_(aa, pa, bb);
          pa = (-cc) * r;
          pb = fma_(-aa, pb, dd) * r;
          CA.at(k) = pa;
          CB.at(k) = pb;
        } else {
          const float r = 1.0f / fma_(aa, pa, bb);
          pb = fma_(-aa, pb, dd) * r;  // w at the top interior level
          w[c] = pb;
        }
        pq_m = pq; p2q_m = p2q; pp_m = p[t]; ppf_m = ppf;
      }
    }
  }
  float wn = pb;
  for (int k0 = nz - 2; k0 >= 1; k0 -= WS_CHUNK) {
    float ca[WS_CHUNK], cb[WS_CHUNK];
#pragma unroll
    for (int t = 0; t < WS_CHUNK; ++t) {
      const int k = k0 - t;
      if (k >= 1) { ca[t] = CA.at(k); cb[t] = CB.at(k); }
    }
#pragma unroll
    for (int t = 0; t < WS_CHUNK; ++t) {
      const int k = k0 - t;
      if (k >= 1) {
        wn = fma_(ca[t], wn, cb[t]);
        w[base + k * sk] = wn;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_acoustic_w_solve(uint32_t* M, KArg a) { w_solve<false>(M, a); }
// Same sweep with cA / cB in LDS (2 nz floats per lane), for fused step programs.
__global__ __launch_bounds__(256) void k_acoustic_w_solve_lds(uint32_t* M, KArg a) { w_solve<true>(M, a); }

// acoustic_pp.wgsl:131-167
__global__ __launch_bounds__(256) void k_acoustic_pp(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  Acoustic ac = acoustic(a.q[1], 0);
  const float *w = fq(a.q[2]), *pk1 = fq(a.q[3]), *pk2 = fq(a.q[4]), *qk = fq(a.q[5]), *fpk = fq(a.q[6]);
  float *pp3d = fq(a.q[7]), *ppd = fq(a.q[8]), *ppx = fq(a.q[9]);
  int saveX = ldi(a.q[10], 0);
  const int64_t sk = g.sk(), c = g.idx(i, j, k);
  float prev = pp3d[c];
  float p = fma_(ac.alph, fma_(w[c], pk1[c] + qk[c], w[c + sk] * (pk2[c] - qk[c])), prev + fpk[c]);
  if (ab(p) < ac.smeps) p = 0.0f;
  pp3d[c] = p;
  float dum1 = ac.kdiv * (p - prev);
  ppd[c] = p + dum1;
  if (saveX != 0) ppx[c] = dum1;
}

// acoustic_accumulate.wgsl:129-158
__global__ __launch_bounds__(256) void k_acoustic_accumulate(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  Acoustic ac = acoustic(a.q[1], 0);
  const float *u = fq(a.q[2]), *v = fq(a.q[3]), *w = fq(a.q[4]);
  float *rru = fq(a.q[5]), *rrv = fq(a.q[6]), *rrw = fq(a.q[7]);
  float tavg = ldf(a.q[8], 0);
  const int64_t c = g.idx(i, j, k);
  if (tavg > 0.0f) {
    rru[c] = (rru[c] + u[c]) * tavg;
    rrv[c] = (rrv[c] + v[c]) * tavg;
    rrw[c] = fma_(w[c], ac.alph, rrw[c]) * tavg;
  } else {
    rru[c] = rru[c] + u[c];
    rrv[c] = rrv[c] + v[c];
    rrw[c] = rrw[c] + w[c];
  }
}

// ---- positive-definite flux correction
TE_DEV bool pdef_inflow(const Dims& g, int dir, int i, int j, int64_t c, int64_t step, const float* rr, int openx,
                        int openy) {
  if (dir == 0 && openx != 0) {
    if (i == 0 && rr[c] >= 0.0f) return true;
    if (i == g.nx - 1 && rr[c + step] <= 0.0f) return true;
  }
  if (dir == 1 && openy != 0) {
    if (j == 0 && rr[c] >= 0.0f) return true;
    if (j == g.ny - 1 && rr[c + step] <= 0.0f) return true;
  }
  return false;
}

// pdef_dum.wgsl:177-231: the value at cell (i, j, k), index c
TE_DEV float pdef_dum_at(uint32_t* M, const KArg& a, const Dims g, int i, int j, int k, int64_t c) {
  int dir = ldi(a.q[1], 0);
  const float *s = fq(a.q[2]), *s0 = fq(a.q[3]), *f = fq(a.q[4]), *rr = fq(a.q[5]), *rho0 = fq(a.q[6]),
              *mh = fq(a.q[7]), *hm = fq(a.q[30]);
  float dt = ldf(a.q[8], 0);
  int openx = ldi(a.q[9], 0), openy = ldi(a.q[9], 1);
  const int64_t step = dir == 0 ? 1 : (dir == 1 ? (int64_t)g.sJ : g.sk());
  float rd = dir == 0 ? g.rdx : (dir == 1 ? g.rdy : g.rdz);
  float base = rho0[k] * s0[c];
  if (pdef_inflow(g, dir, i, j, c, step, rr, openx, openy)) return base;
  float adv = (-(f[c + step] - f[c])) * rd;
  float sdiv = (s[c] * (rr[c + step] - rr[c])) * rd;
  float m = dir == 0 ? hm_uh(hm, g, i) : (dir == 1 ? hm_vh(hm, g, j) : mh[k]);
  adv = adv * m;
  sdiv = sdiv * m;
  return fma_(dt, adv + sdiv, base);
}
__global__ __launch_bounds__(256) void k_pdef_dum(uint32_t* M, KArg a) {
  TE_TID3
  const Dims g = dims(a.q[0], 0);
  const int i = gx, j = gy, k = gz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int64_t c = g.idx(i, j, k);
  fq(a.q[10])[c] = pdef_dum_at(M, a, g, i, j, k, c);
}
// pdef_dum, then the halo fill of its output on a periodic domain (nx, ny >=
// 4; timestep.js:1593 haloFill): each cell writes its periodic images.
// Arguments as pdef_dum's.
__global__ __launch_bounds__(256) void k_pdef_dum_xy(uint32_t* M, KArg a) {
  TE_TID3
  const Dims g = dims(a.q[0], 0);
  const int i = gx, j = gy, k = gz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  update_with_images(g, fq(a.q[10]), i, j, k, 0, g.nz, [&](int64_t c) { return pdef_dum_at(M, a, g, i, j, k, c); });
}

struct PdefCell { bool neg; float f1, f2, foo3; };
TE_DEV PdefCell pdef_cell(float d, float dm, float dp) {
  PdefCell r{false, 0.0f, 0.0f, 0.0f};
  if (d < 0.0f) {
    float f1 = mx(0.0f, dm), f2 = mx(0.0f, dp);
    if (f1 + f2 > 1e-30f) {
      r.neg = true; r.f1 = f1; r.f2 = f2;
      r.foo3 = mx(d, -(f1 + f2)) / (f1 + f2);
    }
  }
  return r;
}
TE_DEV float pdef_mass(float dm2, float dm1, float d0, float dp1, float dp2, bool& flag) {
  PdefCell l = pdef_cell(dm1, dm2, d0), c = pdef_cell(d0, dm1, dp1), r = pdef_cell(dp1, d0, dp2);
  float mass = 0.0f;
  if (l.neg) mass = fma_(l.f2, l.foo3, mass);
  if (c.neg) mass = fma_(-(c.f1 + c.f2), c.foo3, mass);
  if (r.neg) mass = fma_(r.f1, r.foo3, mass);
  flag = c.neg || (l.neg && d0 > 1e-30f) || (r.neg && d0 > 1e-30f);
  return mass;
}

// pdef_apply.wgsl:207-282
__global__ __launch_bounds__(256) void k_pdef_apply(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const float *s = fq(a.q[1]), *s0 = fq(a.q[2]), *fx = fq(a.q[3]), *fy = fq(a.q[4]), *fz = fq(a.q[5]),
              *rru = fq(a.q[6]), *rrv = fq(a.q[7]), *rrw = fq(a.q[8]), *dx = fq(a.q[9]),
              *dy = fq(a.q[10]), *dz = fq(a.q[11]), *rho0 = fq(a.q[12]), *mh = fq(a.q[13]),
              *hm = fq(a.q[30]);
  float dt = ldf(a.q[14], 0);
  int openx = ldi(a.q[15], 0), openy = ldi(a.q[15], 1);
  float* tend = fq(a.q[16]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  float rdt = 1.0f / dt;
  float base = rho0[k] * s0[c];
  float delta = 0.0f;
  bool flag;
  // the three directions' dum windows and the tendency loaded together
  // (what a flagged direction reads besides comes after)
  const float x0 = dx[c - 2], x1 = dx[c - 1], x2 = dx[c], x3 = dx[c + 1], x4 = dx[c + 2];
  const float y0 = dy[c - 2 * sj], y1 = dy[c - sj], y2 = dy[c], y3 = dy[c + sj], y4 = dy[c + 2 * sj];
  const float zm2 = k >= 2 ? dz[c - 2 * sk] : 0.0f, zm1 = k >= 1 ? dz[c - sk] : 0.0f, z0 = dz[c];
  const float zp1 = k + 1 < g.nz ? dz[c + sk] : 0.0f, zp2 = k + 2 < g.nz ? dz[c + 2 * sk] : 0.0f;
  const float t0 = tend[c];
  {
    float ms = pdef_mass(x0, x1, x2, x3, x4, flag);
    if (flag) {
      float sdiv = ((s[c] * (rru[c + 1] - rru[c])) * g.rdx) * hm_uh(hm, g, i);
      float advOld = pdef_inflow(g, 0, i, j, c, 1, rru, openx, openy) ? -sdiv : ((-(fx[c + 1] - fx[c])) * g.rdx) * hm_uh(hm, g, i);
      float advNew = fma_((x2 + ms) - base, rdt, -sdiv);
      delta = delta + (advNew - advOld);
    }
  }
  {
    float ms = pdef_mass(y0, y1, y2, y3, y4, flag);
    if (flag) {
      float sdiv = ((s[c] * (rrv[c + sj] - rrv[c])) * g.rdy) * hm_vh(hm, g, j);
      float advOld = pdef_inflow(g, 1, i, j, c, sj, rrv, openx, openy) ? -sdiv : ((-(fy[c + sj] - fy[c])) * g.rdy) * hm_vh(hm, g, j);
      float advNew = fma_((y2 + ms) - base, rdt, -sdiv);
      delta = delta + (advNew - advOld);
    }
  }
  {
    float ms = pdef_mass(zm2, zm1, z0, zp1, zp2, flag);
    if (flag) {
      float sdiv = ((s[c] * (rrw[c + sk] - rrw[c])) * g.rdz) * mh[k];
      float advOld = ((-(fz[c + sk] - fz[c])) * g.rdz) * mh[k];
      float advNew = fma_((z0 + ms) - base, rdt, -sdiv);
      delta = delta + (advNew - advOld);
    }
  }
  tend[c] = t0 + delta / rho0[k];
}

}  // namespace te
