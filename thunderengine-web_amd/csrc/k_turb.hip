// Rayleigh damping and 1.5-order TKE (LES) turbulence kernels, one per
// reference kernel.  Column kernels run one lane per (i,j) column with the
// tridiagonal sweep carried in registers.
#include "te_device.hpp"
#include "te_kernels.hpp"

// Column sweeps load TS_CHUNK levels ahead of their serial arithmetic.
#ifndef TS_CHUNK
#define TS_CHUNK 8
#endif

namespace te {

// rayleigh_damping.wgsl:111-164
__global__ __launch_bounds__(256) void k_rayleigh_damping(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz;
  if (i > g.nx || j > g.ny || k > g.nz) return;
  const float *u = fq(a.q[1]), *v = fq(a.q[2]), *w = fq(a.q[3]), *th = fq(a.q[4]), *u0 = fq(a.q[5]),
              *v0 = fq(a.q[6]), *tauh = fq(a.q[7]), *tauf = fq(a.q[8]), *hdx = fq(a.q[14]),
              *hdy = fq(a.q[15]), *avg = fq(a.q[18]), *th0 = fq(a.q[19]);
  float *uten = fq(a.q[10]), *vten = fq(a.q[11]), *wten = fq(a.q[12]), *thten = fq(a.q[13]);
  float rdalpha = ldf(a.q[9], 0);
  int lateral = ldi(a.q[16], 0), toAverage = ldi(a.q[17], 0);
  const int64_t c = g.idx(i, j, k);
  // the four fields loaded before the first store (the loads go out together)
  const float uc = u[c], vc = v[c], thc = th[c], wc = w[c];
  if (k < g.nz) {
    float t = tauh[k];
    float ur = toAverage != 0 ? avg[k] : u0[k];
    float vr = toAverage != 0 ? avg[g.nz + k] : v0[k];
    if (j < g.ny) uten[c] = ((-rdalpha) * t) * (uc - ur);
    if (i < g.nx) vten[c] = ((-rdalpha) * t) * (vc - vr);
    if (i < g.nx && j < g.ny) {
      float d = toAverage != 0 ? (th0[k] - avg[2 * g.nz + k]) + thc : thc;
      thten[c] = ((-rdalpha) * t) * d;
    }
  }
  if (i < g.nx && j < g.ny) {
    float tf = tauf[k];
    if (lateral != 0 && k < g.nz) tf = mx(mx(tf, hdx[i]), hdy[j]);
    wten[c] = ((-rdalpha) * tf) * wc;
  }
}

// turb_rho_faces.wgsl:111-151
__global__ __launch_bounds__(256) void k_turb_rho_faces(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 2, j = gy - 2, k = gz;
  if (i > g.nx + 1 || j > g.ny + 1 || k > g.nz) return;
  const float *rho = fq(a.q[1]), *c1 = fq(a.q[2]), *c2 = fq(a.q[3]);
  float *rr = fq(a.q[4]), *rf = fq(a.q[5]);
  const uint32_t *os = a.q[6], *ot = a.q[7];
  const int64_t sk = g.sk(), c = g.idx(i, j, k);
  // rf formed before either store (the loads go out together)
  float r;
  if (k == 0) {
    r = fma_(ldf(os, 0), rho[c], ldf(os, 1) * rho[c + sk]);
    if (g.nz >= 3) r = fma_(ldf(os, 2), rho[c + 2 * sk], r);
  } else if (k == g.nz) {
    r = fma_(ldf(ot, 0), rho[c - sk], ldf(ot, 1) * rho[c - 2 * sk]);
    if (g.nz >= 3) r = fma_(ldf(ot, 2), rho[c - 3 * sk], r);
  } else {
    r = fma_(c1[k], rho[c - sk], c2[k] * rho[c]);
  }
  if (k < g.nz) rr[c] = 1.0f / rho[c];
  rf[c] = r;
}

// turb_def_centers.wgsl:139-168
__global__ __launch_bounds__(256) void k_turb_def_centers(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 1, j = gy - 1, k = gz;
  if (i > g.nx + 1 || j > g.ny + 1 || k >= g.nz) return;
  const float *u = fq(a.q[1]), *v = fq(a.q[2]), *w = fq(a.q[3]), *rho = fq(a.q[4]), *mh = fq(a.q[5]),
              *hm = fq(a.q[30]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  // every value formed before the first store (the loads go out together)
  const float r0 = rho[c];
  const float s11 = ((r0 * (u[c + 1] - u[c])) * g.rdx) * hm_uh(hm, g, i);
  const float s22 = ((r0 * (v[c + sj] - v[c])) * g.rdy) * hm_vh(hm, g, j);
  const float s33 = ((r0 * (w[c + sk] - w[c])) * g.rdz) * mh[k];
  const float rcorner = 0.25f * ((rho[c - 1 - sj] + r0) + (rho[c - 1] + rho[c - sj]));
  const float s12 =
      (0.5f * fma_((u[c] - u[c - sj]) * g.rdy, hm_vf(hm, g, j), ((v[c] - v[c - 1]) * g.rdx) * hm_uf(hm, g, i))) * rcorner;
  fq(a.q[6])[c] = s11;
  fq(a.q[7])[c] = s22;
  fq(a.q[8])[c] = s33;
  fq(a.q[9])[c] = s12;
}

// turb_def_faces.wgsl:125-151
__global__ __launch_bounds__(256) void k_turb_def_faces(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz + 1;
  if (i > g.nx + 1 || j > g.ny + 1 || k > g.nz - 1) return;
  const float *u = fq(a.q[1]), *v = fq(a.q[2]), *w = fq(a.q[3]), *rf = fq(a.q[4]), *mf = fq(a.q[5]),
              *hm = fq(a.q[30]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  float mfu = 0.5f * (mf[k] + mf[k]);
  float mfv = 0.5f * (mf[k] + mf[k]);
  // both values formed before the first store
  const float s13 =
      ((0.5f * fma_((w[c] - w[c - 1]) * g.rdx, hm_uf(hm, g, i), ((u[c] - u[c - sk]) * g.rdz) * mfu)) * 0.5f) * (rf[c - 1] + rf[c]);
  const float s23 =
      ((0.5f * fma_((w[c] - w[c - sj]) * g.rdy, hm_vf(hm, g, j), ((v[c] - v[c - sk]) * g.rdz) * mfv)) * 0.5f) * (rf[c - sj] + rf[c]);
  fq(a.q[6])[c] = s13;
  fq(a.q[7])[c] = s23;
}

// turb_def_assemble.wgsl:111-149
__global__ __launch_bounds__(256) void k_turb_def_assemble(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz + 1;
  if (i >= g.nx || j >= g.ny || k > g.nz - 1) return;
  const float *s11 = fq(a.q[1]), *s22 = fq(a.q[2]), *s33 = fq(a.q[3]), *s12 = fq(a.q[4]),
              *s13 = fq(a.q[5]), *s23 = fq(a.q[6]), *c1 = fq(a.q[7]), *c2 = fq(a.q[8]), *rf = fq(a.q[9]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  float A = c1[k], B = c2[k];
  float t11 = fma_(A * s11[c - sk], s11[c - sk], (B * s11[c]) * s11[c]);
  float t22 = fma_(A * s22[c - sk], s22[c - sk], (B * s22[c]) * s22[c]);
  float t33 = fma_(A * s33[c - sk], s33[c - sk], (B * s33[c]) * s33[c]);
  auto sq2 = [&](int64_t p, int64_t q) { return fma_(s12[p], s12[p], s12[q] * s12[q]); };
  float lo = sq2(c - sk, c + 1 + sj - sk) + sq2(c + sj - sk, c + 1 - sk);
  float hi = sq2(c, c + 1 + sj) + sq2(c + sj, c + 1);
  float t12 = 0.25f * fma_(A, lo, B * hi);
  float t13 = 0.5f * fma_(s13[c], s13[c], s13[c + 1] * s13[c + 1]);
  float t23 = 0.5f * fma_(s23[c], s23[c], s23[c + sj] * s23[c + sj]);
  float rrf = 1.0f / (rf[c] * rf[c]);
  fq(a.q[10])[c] = (4.0f * (t13 + t23)) * rrf;
  fq(a.q[11])[c] = fma_(2.0f, (t11 + t22) + t33, 4.0f * t12) * rrf;
}

// saturation mixing ratios over liquid / ice (turb_nm.wgsl:204-230)
TE_DEV float mp_rslf(float p, float t) {
  float esl = 611.2f * fast_exp((17.67f * (t - 273.15f)) / (t - 29.65f));
  esl = mn(esl, p * 0.5f);
  return ((RD / RV) * esl) / (p - esl);
}
TE_DEV float mp_rsif(float p, float t) {
  float esi = 611.2f * fast_exp((21.874557f * (t - 273.15f)) / (t - 7.66f));
  esi = mn(esi, p * 0.5f);
  return ((RD / RV) * esi) / (p - esi);
}

// turb_nm.wgsl:236-355 (moist squared buoyancy frequency)
__global__ __launch_bounds__(256) void k_turb_nm(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz + 1;
  if (i >= g.nx || j >= g.ny || k > g.nz - 1) return;
  const float *thP = fq(a.q[1]), *ppP = fq(a.q[2]), *th0 = fq(a.q[3]), *pi0 = fq(a.q[4]),
              *prs = fq(a.q[5]), *qv = fq(a.q[6]), *qc = fq(a.q[7]), *qr = fq(a.q[8]), *qi = fq(a.q[9]),
              *qs = fq(a.q[10]), *qg = fq(a.q[11]), *c1 = fq(a.q[12]), *c2 = fq(a.q[13]), *mf = fq(a.q[14]);
  float *nm = fq(a.q[15]), *sat = fq(a.q[17]);
  int moist = ldi(a.q[16], 0), wantSat = ldi(a.q[18], 0);
  const int64_t sk = g.sk(), c = g.idx(i, j, k);
  const float nmsat = 1e-10f, eps = RD / RV, reps = RV / RD, gAcc = 9.81f;
  const float LV1 = 3134708.0f, LV2 = 2.32e+03f, LS1 = 2898463.5f, LS2 = 236.0f;
  float thA = th0[k - 1] + thP[c - sk];
  float thB = th0[k] + thP[c];
  if (moist == 0) {
    if (wantSat != 0) sat[c] = 0.0f;
    nm[c] = ((fast_log(thB / thA) * gAcc) * g.rdz) * mf[k];
    return;
  }
  float piA = pi0[k - 1] + ppP[c - sk], piB = pi0[k] + ppP[c];
  float tA = thA * piA, tB = thB * piB;
  float qtA = ((((qv[c - sk] + qc[c - sk]) + qr[c - sk]) + qi[c - sk]) + qs[c - sk]) + qg[c - sk];
  float qtB = ((((qv[c] + qc[c]) + qr[c]) + qi[c]) + qs[c]) + qg[c];
  float qciA = (qv[c - sk] + qc[c - sk]) + qi[c - sk];
  float qciB = (qv[c] + qc[c]) + qi[c];
  float thvA = (thA * fma_(reps, qv[c - sk], 1.0f)) / (1.0f + qtA);
  float thvB = (thB * fma_(reps, qv[c], 1.0f)) / (1.0f + qtB);
  float rmf = g.rdz * mf[k];
  float A = c1[k], B = c2[k];
  float pavg = fma_(A, prs[c - sk], B * prs[c]);
  float tavg = fma_(A, tA, B * tB);
  float qtavg = fma_(A, qtA, B * qtB);
  float qciavg = fma_(A, qciA, B * qciB);
  // The three temperature ranges only choose lam's operands (liquid, ice, or
  // both blended by ff); lam and the logarithms it needs are evaluated once,
  // below, for the saturated points alone.
  bool satd = false, blend = false;
  float qvs1 = 0.0f, lh1 = 0.0f, drdt1 = 0.0f, qvs2 = 0.0f, lh2 = 0.0f, drdt2 = 0.0f, cpml = 0.0f, ff = 0.0f;
  if (tavg >= 273.15f) {
    float qvs = mp_rslf(pavg, tavg);
    float ql = mx(qciavg - qvs, 0.0f);
    if (ql > nmsat) {
      satd = true;
      float qvl = mx(qciavg - ql, 0.0f);
      cpml = fma_(CPL, ql, fma_(CPV, qvl, CP));
      qvs1 = qvs;
      lh1 = fma_(-LV2, tavg, LV1);
      drdt1 = ((17.67f * (273.15f - 29.65f)) * qvs) / ((tavg - 29.65f) * (tavg - 29.65f));
    }
  } else if (tavg <= 233.15f) {
    float qvs = mp_rsif(pavg, tavg);
    float qice = mx(qciavg - qvs, 0.0f);
    if (qice > nmsat) {
      satd = true;
      float qvl = mx(qciavg - qice, 0.0f);
      cpml = fma_(CPI, qice, fma_(CPV, qvl, CP));
      qvs1 = qvs;
      lh1 = fma_(-LS2, tavg, LS1);
      drdt1 = ((21.874557f * (273.15f - 7.66f)) * qvs) / ((tavg - 7.66f) * (tavg - 7.66f));
    }
  } else {
    ff = (tavg - 233.15f) / (273.15f - 233.15f);
    float qvsl = mp_rslf(pavg, tavg), qvsi = mp_rsif(pavg, tavg);
    float qvs = fma_(ff, qvsl, (1.0f - ff) * qvsi);
    float qice = mx(0.0f, (1.0f - ff) * (qciavg - qvs));
    float ql = mx(0.0f, ff * (qciavg - qvs));
    if (ql + qice > nmsat) {
      satd = blend = true;
      float qvl = mx(0.0f, (qciavg - ql) - qice);
      cpml = fma_(CPI, qice, fma_(CPL, ql, fma_(CPV, qvl, CP)));
      qvs1 = qvsl;
      lh1 = fma_(-LV2, tavg, LV1);
      drdt1 = ((17.67f * (273.15f - 29.65f)) * qvsl) / ((tavg - 29.65f) * (tavg - 29.65f));
      qvs2 = qvsi;
      lh2 = fma_(-LS2, tavg, LS1);
      drdt2 = ((21.874557f * (273.15f - 7.66f)) * qvsi) / ((tavg - 7.66f) * (tavg - 7.66f));
    }
  }
  if (wantSat != 0) sat[c] = satd ? 1.0f : 0.0f;
  float out;
  if (satd) {
    const float dlnT = fast_log(tB / tA) * rmf;
    const float dlnQt = fast_log((1.0f + qtB) / (1.0f + qtA)) * rmf;
    auto lam = [&](float qvs, float lh, float drdt) {
      float gamma = ((gAcc * (1.0f + qtavg)) * (1.0f + (lh * qvs) / (RD * tavg))) / fma_(lh, drdt, cpml);
      return gAcc * fma_(dlnT + gamma / tavg, 1.0f + (tavg * drdt) / (eps + qvs), -dlnQt);
    };
    const float n1 = lam(qvs1, lh1, drdt1);
    out = blend ? fma_(ff, n1, (1.0f - ff) * lam(qvs2, lh2, drdt2)) : n1;
  } else {
    out = (gAcc * fast_log(thvB / thvA)) * rmf;
  }
  nm[c] = out;
}

// turb_km_kh.wgsl:125-187; IMG: then the halo fill of kmh, kmv, khh, khv on
// a periodic domain (nx, ny >= 4; timestep.js:1593 haloFill): each cell
// writes its periodic images, the cells of levels 1 and nz-1 also the images
// of levels 0 and nz (not written here; no thread of the launch reads K).
template <bool IMG>
__device__ __forceinline__ void turb_km_kh_body(uint32_t* M, const KArg& a) {
  TE_TID3
  const Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz + 1;
  if (i >= g.nx || j >= g.ny || k > g.nz - 1) return;
  const float *tkea = fq(a.q[1]), *nm = fq(a.q[2]), *mf = fq(a.q[3]), *hm = fq(a.q[30]);
  float *kmh = fq(a.q[4]), *kmv = fq(a.q[5]), *khh = fq(a.q[6]), *khv = fq(a.q[7]), *lenscl = fq(a.q[8]),
        *dissten = fq(a.q[9]);
  float kmScale = ldf(a.q[10], 0), kmFixed = ldf(a.q[11], 0), turbfac = ldf(a.q[12], 0);
  const int64_t c = g.idx(i, j, k);
  const float c_m = 0.1f, c_l = 0.8165f, ri_c = 0.25f, tke_min = 1e-10f, nm_min = 1e-06f, small_len_frac = 0.001f;
  float cmz = mx(0.001f, turbfac * c_m);
  float ce1z = ((cmz * c_l) * c_l) * (1.0f / ri_c - 1.0f);
  float ce2z = mx(0.0f, fma_(cmz * 3.1415927f, 3.1415927f, -ce1z));
  float dzk = (1.0f / g.rdz) / mf[k];
  float grdscl = fast_pow((((1.0f / g.rdx) / hm_uh(hm, g, i)) * ((1.0f / g.rdy) / hm_vh(hm, g, j))) * dzk, 0.33333334f);
  float rgrdscl = 1.0f / grdscl;
  // the cell's operands, loaded once before the stores below
  const float tkc = tkea[c], nmc = nm[c], dissc = dissten[c];
  float tk, len;
  if (tkc <= tke_min) {
    tk = tke_min;
    len = small_len_frac * grdscl;
  } else {
    tk = tkc;
    len = grdscl;
    if (nmc > nm_min) {
      len = c_l * sqrtf(tk / nmc);
      len = mn(len, grdscl);
      len = mx(len, small_len_frac * grdscl);
    }
  }
  lenscl[c] = len;
  dissten[c] = dissc + ((fma_(ce2z * len, rgrdscl, ce1z) * tk) * sqrtf(tk)) / len;
  float km = kmFixed > 0.0f ? kmFixed : ((cmz * sqrtf(tk)) * len) * kmScale;
  kmh[c] = km;
  kmv[c] = km;
  float prinv = 3.0f;
  if (nmc > nm_min) prinv = mn(fma_(2.0f * len, rgrdscl, 1.0f), 3.0f);
  khh[c] = km * prinv;
  khv[c] = km * prinv;
  if (IMG) {
    float* F[4] = {kmh, kmv, khh, khv};
    const float v[4] = {km, km, km * prinv, km * prinv};
    const int64_t sk = g.sk(), e = k == 1 ? -sk : (k == g.nz - 1 ? sk : 0);
    const int64_t e2 = k == 1 && k == g.nz - 1 ? sk : 0;  // nz = 2: level 1 is next to both
    float ev[4], ev2[4];
    for (int q = 0; q < 4; ++q) {
      ev[q] = e != 0 ? F[q][c + e] : 0.0f;
      ev2[q] = e2 != 0 ? F[q][c + e2] : 0.0f;
    }
    periodic_images(g, i, j, [&](int64_t o) {
      for (int q = 0; q < 4; ++q) {
        F[q][c + o] = v[q];
        if (e != 0) F[q][c + e + o] = ev[q];
        if (e2 != 0) F[q][c + e2 + o] = ev2[q];
      }
    });
  }
}
__global__ __launch_bounds__(256) void k_turb_km_kh(uint32_t* M, KArg a) { turb_km_kh_body<false>(M, a); }
__global__ __launch_bounds__(256) void k_turb_km_kh_xy(uint32_t* M, KArg a) { turb_km_kh_body<true>(M, a); }

// turb_dissipative_heating.wgsl:139-194
__global__ __launch_bounds__(256) void k_turb_dissipative_heating(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const float *dissten = fq(a.q[1]), *pi0 = fq(a.q[2]), *ppi = fq(a.q[3]), *qv = fq(a.q[4]),
              *qc = fq(a.q[5]), *qr = fq(a.q[6]), *qi = fq(a.q[7]), *qs = fq(a.q[8]), *qg = fq(a.q[9]),
              *sfc = fq(a.q[12]);
  float* thten = fq(a.q[10]);
  const uint32_t* op = a.q[11];
  int moist = ldi(op, 0), bbc3 = ldi(op, 1), zntField = ldi(op, 4);
  float znt = ldf(op, 2), dz1 = ldf(op, 3);
  const int64_t sk = g.sk(), c = g.idx(i, j, k);
  const float cv = CP - RD, cvv = CPV - RV, cpdcv = CP / (CP - RD);
  float epsd;
  if (k == 0 && bbc3 != 0) {
    float ust = sfc[3 * sk + g.idx(i, j, 0)];
    float z0 = zntField != 0 ? sfc[4 * sk + g.idx(i, j, 0)] : znt;
    epsd = (((ust * ust) * ust) * fast_log((dz1 + z0) / z0)) / (0.4f * dz1);
  } else {
    epsd = 0.5f * (dissten[c] + dissten[c + sk]);
  }
  float pi = pi0[k] + ppi[c];
  float dheat;
  if (moist != 0) {
    float cvm = fma_(2106.0f, (qi[c] + qs[c]) + qg[c], fma_(4.19e+03f, qc[c] + qr[c], fma_(cvv, qv[c], cv)));
    dheat = epsd / ((cpdcv * cvm) * pi);
  } else {
    dheat = epsd / (CP * pi);
  }
  thten[c] = thten[c] + dheat;
}

// turb_gettau.wgsl:111-143
__global__ __launch_bounds__(256) void k_turb_gettau(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 1, j = gy - 1, k = gz;
  if (i > g.nx || j > g.ny || k >= g.nz) return;
  const float *kmh = fq(a.q[1]), *kmv = fq(a.q[2]);
  float *t11 = fq(a.q[3]), *t22 = fq(a.q[4]), *t33 = fq(a.q[5]), *t12 = fq(a.q[6]), *t13 = fq(a.q[7]),
        *t23 = fq(a.q[8]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  // every operand loaded before the first store (the loads go out together;
  // t13 / t23 and kmv at the column's first level are read but not used)
  const float tem = kmh[c] + kmh[c + sk];
  const float q = ((kmh[c - 1 - sj] + kmh[c]) + (kmh[c - 1] + kmh[c - sj])) +
                  ((kmh[c - 1 - sj + sk] + kmh[c + sk]) + (kmh[c - 1 + sk] + kmh[c - sj + sk]));
  const float a11 = t11[c], a22 = t22[c], a33 = t33[c], a12 = t12[c], a13 = t13[c], a23 = t23[c];
  const float kv = kmv[c], kvx = kmv[c - 1], kvy = kmv[c - sj];
  t11[c] = a11 * tem;
  t22[c] = a22 * tem;
  t33[c] = a33 * tem;
  t12[c] = a12 * (0.25f * q);
  if (k >= 1 && k <= g.nz - 1 && i >= 0 && j >= 0) {
    t13[c] = a13 * (kvx + kv);
    t23[c] = a23 * (kvy + kv);
  }
}

// turb_sfcstress.wgsl:132-207 (free-slip and no-slip bottom; surface-flux
// modes need the surface-layer planes and are not on this path)
__global__ __launch_bounds__(256) void k_turb_sfcstress(uint32_t* M, KArg a) {
  const Blk b_ = te_block(a);
  const int i = b_.x * TE_BX + (int)threadIdx.x, j = b_.y * TE_BY + (int)threadIdx.y;
  if (i >= (int)a.n[0] || j >= (int)a.n[1]) return;
  Dims g = dims(a.q[0], 0);
  if (i > g.nx || j > g.ny) return;
  const float *rf = fq(a.q[2]), *u = fq(a.q[6]), *v = fq(a.q[7]), *kmv = fq(a.q[9]), *mf = fq(a.q[11]);
  float *t13 = fq(a.q[3]), *t23 = fq(a.q[4]);
  int mode = ldi(a.q[5], 0), topNoSlip = ldi(a.q[10], 0);
  float movex = ldf(a.q[8], 0), movey = ldf(a.q[8], 1);
  const int64_t plane = g.sk(), sjj = g.sJ;
  {
    const int64_t ct = g.idx(i, j, g.nz);
    if (topNoSlip != 0) {
      const int64_t cu = ct - plane;
      float mftu = 0.5f * (mf[g.nz] + mf[g.nz]), mftv = 0.5f * (mf[g.nz] + mf[g.nz]);
      t13[ct] = ((((((-2.0f * (u[cu] + movex)) * g.rdz) * mftu) * 0.5f) * (rf[ct - 1] + rf[ct])) * 0.5f) * (kmv[cu - 1] + kmv[cu]);
      t23[ct] = ((((((-2.0f * (v[cu] + movey)) * g.rdz) * mftv) * 0.5f) * (rf[ct - sjj] + rf[ct])) * 0.5f) * (kmv[cu - sjj] + kmv[cu]);
    } else {
      t13[ct] = t13[ct - plane];
      t23[ct] = t23[ct - plane];
    }
  }
  const int64_t c = g.idx(i, j, 0);
  if (mode == 4) {
    float mfbu = 0.5f * (mf[0] + mf[0]), mfbv = 0.5f * (mf[0] + mf[0]);
    t13[c] = ((((((2.0f * (u[c] + movex)) * g.rdz) * mfbu) * 0.5f) * (rf[c - 1] + rf[c])) * 0.5f) * (kmv[c - 1 + plane] + kmv[c + plane]);
    t23[c] = ((((((2.0f * (v[c] + movey)) * g.rdz) * mfbv) * 0.5f) * (rf[c - sjj] + rf[c])) * 0.5f) * (kmv[c - sjj + plane] + kmv[c + plane]);
    return;
  }
  t13[c] = t13[c + plane];
  t23[c] = t23[c + plane];
}

// tke_epst.wgsl:170-206 (shear production).  Two levels per thread (the
// runtime halves the launch's z extent): k0 = 2 gz + 1 and k0 + 1.  The
// terms of the face between them (L1, the four-point L2 sum, L5 at k0) enter
// level k0 as its upper terms and level k0 + 1 as its lower ones, formed once.
__global__ __launch_bounds__(256) void k_tke_epst(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  const int i = gx, j = gy, k0 = 2 * gz + 1;
  if (i >= g.nx || j >= g.ny || k0 > g.nz - 1) return;
  const float *u = fq(a.q[1]), *v = fq(a.q[2]), *w = fq(a.q[3]), *t11 = fq(a.q[4]), *t12 = fq(a.q[5]),
              *t13 = fq(a.q[6]), *t22 = fq(a.q[7]), *t23 = fq(a.q[8]), *t33 = fq(a.q[9]), *rf = fq(a.q[10]),
              *mh = fq(a.q[11]), *mf = fq(a.q[12]), *c1 = fq(a.q[13]), *c2 = fq(a.q[14]), *hm = fq(a.q[30]);
  const int64_t sj = g.sJ, sk = g.sk(), c0 = g.idx(i, j, k0);
  auto L1 = [&](int64_t cc) {
    return fma_(((u[cc + 1] - u[cc]) * g.rdx) * hm_uh(hm, g, i), t11[cc], (((v[cc + sj] - v[cc]) * g.rdy) * hm_vh(hm, g, j)) * t22[cc]);
  };
  auto L2 = [&](int64_t cc, int ii, int jj) {
    return t12[cc] * fma_((u[cc] - u[cc - sj]) * g.rdy, hm_vf(hm, g, jj), ((v[cc] - v[cc - 1]) * g.rdx) * hm_uf(hm, g, ii));
  };
  auto L3 = [&](int64_t cc, int ii, int k) {
    return fma_((w[cc] - w[cc - 1]) * g.rdx, hm_uf(hm, g, ii), ((u[cc] - u[cc - sk]) * g.rdz) * mf[k]) * t13[cc];
  };
  auto L4 = [&](int64_t cc, int jj, int k) {
    return fma_((w[cc] - w[cc - sj]) * g.rdy, hm_vf(hm, g, jj), ((v[cc] - v[cc - sk]) * g.rdz) * mf[k]) * t23[cc];
  };
  auto L5 = [&](int64_t cc, int lev) { return (((w[cc + sk] - w[cc]) * g.rdz) * mh[lev]) * t33[cc]; };
  auto L2s = [&](int64_t cc) {
    return (L2(cc, i, j) + L2(cc + 1 + sj, i + 1, j + 1)) + (L2(cc + sj, i, j + 1) + L2(cc + 1, i + 1, j));
  };
  // level k at c from its lower (l1d, dn, l5d) and upper (l1c, up, l5c) terms
  auto level = [&](int k, int64_t c, float l1d, float l1c, float dn, float up, float l5d, float l5c) {
    float e = fma_(c1[k], l1d, c2[k] * l1c);
    e = fma_(0.25f, fma_(c2[k], up, c1[k] * dn), e);
    e = fma_(0.5f, (L3(c, i, k) + L3(c + 1, i + 1, k)) + (L4(c, j, k) + L4(c + sj, j + 1, k)), e);
    e = e + fma_(c1[k], l5d, c2[k] * l5c);
    return e / rf[c];
  };
  const int64_t cd = c0 - sk, cu = c0 + sk;
  const float l1m = L1(cd), l10 = L1(c0), s2m = L2s(cd), s20 = L2s(c0), l5m = L5(cd, k0 - 1), l50 = L5(c0, k0);
  const float e0 = level(k0, c0, l1m, l10, s2m, s20, l5m, l50);
  float* epst = fq(a.q[15]);
  if (k0 + 1 <= g.nz - 1) {
    const float e1 = level(k0 + 1, cu, l10, L1(cu), s20, L2s(cu), l50, L5(cu, k0 + 1));
    epst[c0] = e0;
    epst[cu] = e1;
  } else {
    epst[c0] = e0;
  }
}

// tke_sources.wgsl:111-129
__global__ __launch_bounds__(256) void k_tke_sources(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz + 1;
  if (i >= g.nx || j >= g.ny || k > g.nz - 1) return;
  const int64_t c = g.idx(i, j, k);
  const float *khv = fq(a.q[1]), *nm = fq(a.q[2]), *dissten = fq(a.q[3]), *epst = fq(a.q[4]);
  fq(a.q[5])[c] = fma_(-khv[c], nm[c], -dissten[c]) + epst[c];
}

// turb_u_tendency.wgsl:146-294 / turb_v_tendency.wgsl
template <int DIR, bool LDS>
__device__ void uv_tendency(uint32_t* M, const KArg& a, int i, int j) {
  Dims g = dims(a.q[0], 0);
  if (DIR == 0 ? (i > g.nx || j >= g.ny) : (i >= g.nx || j > g.ny)) return;
  const float *q = fq(a.q[1]), *w = fq(a.q[2]), *rho = fq(a.q[3]), *rf = fq(a.q[4]), *kmv = fq(a.q[5]),
              *tA = fq(a.q[6]), *tB = fq(a.q[7]), *tC = fq(a.q[8]), *mh = fq(a.q[9]), *mf = fq(a.q[10]),
              *hm = fq(a.q[30]);
  float dt = ldf(a.q[11], 0);
  float *cA = fq(a.q[12]), *cB = fq(a.q[13]), *dwd = fq(a.q[14]), *ten = fq(a.q[15]);
  int openx = ldi(a.q[16], 0), openy = ldi(a.q[16], 1), explicitZ = ldi(a.q[17], 0);
  const int64_t sj = g.sJ, sk = g.sk(), st = DIR == 0 ? 1 : sj, base = g.idx(i, j, 0);
  bool zeroX = DIR == 0 ? (openx != 0 && (i == 0 || i == g.nx)) : (openx != 0 && (i == 0 || i == g.nx - 1));
  bool zeroY = DIR == 0 ? (openy != 0 && (j == 0 || j == g.ny - 1)) : (openy != 0 && (j == 0 || j == g.ny));
  auto turbxy = [&](int64_t c, float& tx, float& ty) {
    if (DIR == 0) {
      tx = zeroX ? 0.0f : ((tA[c] - tA[c - 1]) * g.rdx) * hm_uf(hm, g, i);
      ty = zeroY ? 0.0f : ((tB[c + sj] - tB[c]) * g.rdy) * hm_vh(hm, g, j);
    } else {
      tx = zeroX ? 0.0f : ((tA[c + 1] - tA[c]) * g.rdx) * hm_uh(hm, g, i);
      ty = zeroY ? 0.0f : ((tB[c] - tB[c - sj]) * g.rdy) * hm_vf(hm, g, j);
    }
  };
  auto rinv = [&](int k) { const int64_t c = base + k * sk; return 1.0f / (0.5f * (rho[c - st] + rho[c])); };
  if (explicitZ != 0) {
    for (int k = 0; k < g.nz; ++k) {
      const int64_t c = base + k * sk;
      float tx, ty;
      turbxy(c, tx, ty);
      float tz = (((tC[c + sk] - tC[c]) * g.rdz) * 0.5f) * (mh[k] + mh[k]);
      ten[c] = ten[c] + ((tx + ty) + tz) / (0.5f * (rho[c - st] + rho[c]));
    }
    return;
  }
  float tema = ((-0.0625f * dt) * g.rdz) * g.rdz;
  float temc = (0.5f * dt) * g.rdz;
  float rdt = 0.5f / dt;
  const float hmf = DIR == 0 ? hm_uf(hm, g, i) : hm_vf(hm, g, j), rdh = DIR == 0 ? g.rdx : g.rdy;
  const int nz = g.nz;
  const ColStore<LDS> CA{cA, base, sk, TE_LANE, TE_NLANES, 0}, CB{cB, base, sk, TE_LANE, TE_NLANES, nz};
  // Every pass runs in chunks of TS_CHUNK levels whose loads are issued before
  // the chunk's arithmetic.  Y(l) = ((kmv_l + kmv_l') * (mf+mf)) * (rf_l + rf_l')
  // is the level-l factor of L1(k, l) = Y(l) * tem(k); Y(k+1) carries to the
  // next level.  The back substitution and the tendency pass share one sweep.
  for (int k0 = 0; k0 <= nz; k0 += TS_CHUNK) {
    float w0[TS_CHUNK], w1[TS_CHUNK], k0v[TS_CHUNK], k1v[TS_CHUNK], r0[TS_CHUNK], r1[TS_CHUNK];
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 + t;
      if (k <= nz) {
        const int64_t c = base + k * sk;
        w0[t] = w[c]; w1[t] = w[c - st]; k0v[t] = kmv[c - st]; k1v[t] = kmv[c]; r0[t] = rf[c - st]; r1[t] = rf[c];
      }
    }
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 + t;
      if (k <= nz)
        dwd[base + k * sk] = (((((w0[t] - w1[t]) * rdh) * hmf) * 0.25f) * (k0v[t] + k1v[t])) * (r0[t] + r1[t]);
    }
  }
  auto Y = [&](int lev) {
    const int64_t cl = base + lev * sk;
    return ((kmv[cl - st] + kmv[cl]) * (mf[lev] + mf[lev])) * (rf[cl - st] + rf[cl]);
  };
  float pa, pb;
  float Yn = Y(1);
  {
    float tem = (mh[0] + mh[0]) * rinv(0);
    float cfc = tema * (Yn * tem);
    float cfb = 1.0f - cfc;
    float cfd = fma_(-((temc * tC[base]) * (mh[0] + mh[0])), rinv(0), q[base]);
    float t = 1.0f / cfb;
    pa = (-cfc) * t;
    pb = cfd * t;
    CA.at(0) = pa;
    CB.at(0) = pb;
  }
  for (int k0 = 1; k0 < nz - 1; k0 += TS_CHUNK) {
    float ro0[TS_CHUNK], ro1[TS_CHUNK], mhk[TS_CHUNK], ka[TS_CHUNK], kb[TS_CHUNK], mfn[TS_CHUNK], ra[TS_CHUNK],
        rb[TS_CHUNK], qv[TS_CHUNK];
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 + t;
      if (k < nz - 1) {
        const int64_t c = base + k * sk, cn = c + sk;
        ro0[t] = rho[c - st]; ro1[t] = rho[c]; mhk[t] = mh[k];
        ka[t] = kmv[cn - st]; kb[t] = kmv[cn]; mfn[t] = mf[k + 1]; ra[t] = rf[cn - st]; rb[t] = rf[cn];
        qv[t] = q[c];
      }
    }
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 + t;
      if (k < nz - 1) {
        const int64_t c = base + k * sk;
        const float rinvk = 1.0f / (0.5f * (ro0[t] + ro1[t]));
        const float tem = (mhk[t] + mhk[t]) * rinvk;
        const float Yk = Yn;
        Yn = ((ka[t] + kb[t]) * (mfn[t] + mfn[t])) * (ra[t] + rb[t]);
        float cfa = tema * (Yk * tem), cfc = tema * (Yn * tem);
        float cfb = (1.0f - cfa) - cfc;
        float tt = 1.0f / fma_(cfa, pa, cfb);
        pa = (-cfc) * tt;
        pb = fma_(-cfa, pb, qv[t]) * tt;
        CA.at(k) = pa;
        CB.at(k) = pb;
      }
    }
  }
  const int kt = nz - 1;
  const int64_t ct = base + kt * sk;
  {
    float tem = (mh[kt] + mh[kt]) * rinv(kt);
    float cfa = tema * (Yn * tem);
    float cfb = 1.0f - cfa;
    float cfd = fma_((temc * tC[base + nz * sk]) * (mh[kt] + mh[kt]), rinv(kt), q[ct]);
    float t = 1.0f / fma_(cfa, pa, cfb);
    pb = fma_(-cfa, pb, cfd) * t;
    CB.at(kt) = pb;
  }
  // level kt's tendency, then back substitution + tendency for k = kt-1 .. 0
  auto tendency = [&](int k, int64_t c, float cb, float qk, float ro0, float ro1, float d0, float d1, float tx, float ty,
                      float mhk, float tv) {
    float tz = ((ro0 + ro1) * (cb - qk)) * rdt;
    tz = fma_(((d1 - d0) * g.rdz) * 0.5f, mhk + mhk, tz);
    ten[c] = fma_((tx + ty) + tz, 1.0f / (0.5f * (ro0 + ro1)), tv);
  };
  {
    float tx, ty;
    turbxy(ct, tx, ty);
    tendency(kt, ct, pb, q[ct], rho[ct - st], rho[ct], dwd[ct], dwd[ct + sk], tx, ty, mh[kt], ten[ct]);
  }
  for (int k0 = nz - 2; k0 >= 0; k0 -= TS_CHUNK) {
    float ca[TS_CHUNK], cb[TS_CHUNK], qv[TS_CHUNK], ro0[TS_CHUNK], ro1[TS_CHUNK], d0[TS_CHUNK], d1[TS_CHUNK],
        txv[TS_CHUNK], tyv[TS_CHUNK], mhk[TS_CHUNK], tv[TS_CHUNK];
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 - t;
      if (k >= 0) {
        const int64_t c = base + k * sk;
        ca[t] = CA.at(k); cb[t] = CB.at(k); qv[t] = q[c]; ro0[t] = rho[c - st]; ro1[t] = rho[c];
        d0[t] = dwd[c]; d1[t] = dwd[c + sk]; mhk[t] = mh[k]; tv[t] = ten[c];
        turbxy(c, txv[t], tyv[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 - t;
      if (k >= 0) {
        const int64_t c = base + k * sk;
        pb = fma_(ca[t], pb, cb[t]);
        CB.at(k) = pb;
        tendency(k, c, pb, qv[t], ro0[t], ro1[t], d0[t], d1[t], txv[t], tyv[t], mhk[t], tv[t]);
      }
    }
  }
}
// one lane per column (columns of fewer than 4 levels; deeper ones run
// uv_waves)
template <int DIR>
__device__ void uv_tendency_lds(uint32_t* M, const KArg& a) {
  const Blk b_ = te_block(a);
  const int i = b_.x * TE_BX + (int)threadIdx.x, j = b_.y * TE_BY + (int)threadIdx.y;
  if (i < (int)a.n[0] && j < (int)a.n[1]) uv_tendency<DIR, true>(M, a, i, j);
}

// ---------------------------------------------------------------------------
// Column sweeps as 64 columns of a row x AC_WAVES waves: each wave holds a
// contiguous range of levels, its rows' coefficients in registers, and the
// Thomas recurrence passes between the waves through LDS -- the forward
// elimination upwards ((pa, pb) of the wave's last level), the back
// substitution downwards (the new value of the wave's first level).  Wave w
// waits until fwd_done == w (back_next == w), which the wave below (above)
// sets once its share is done; the waits poll at most 2^22 times.  Level by
// level the operations are the sequential sweep's, the first row's missing
// sub-diagonal and the last row's missing super-diagonal being zeros:
// fma(0, 0, b) = b, fma(-0, p, d) = d, (1 - 0) - c = 1 - c, x - 0 = x.
struct WaveRelay {
  float xb[2][64];  // (pa, pb) of the wave below's last level
  float xw[64];     // the new value at the wave above's first level
  int fwd_done, back_next;
};
TE_DEV void relay_init(WaveRelay& r) {
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    r.fwd_done = 0;
    r.back_next = AC_WAVES - 1;
  }
  __syncthreads();
}
TE_DEV void relay_wait(int* flag, int want) {
  for (int n = 0; n < (1 << 22); ++n) {
    const int v = __builtin_amdgcn_readfirstlane(__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
    if (v == want) break;
    __builtin_amdgcn_s_sleep(1);
  }
}
TE_DEV void relay_pass(int* flag, int v) { __hip_atomic_store(flag, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP); }
// The column of a lane: a launch numbers the op's n0 x n1 columns row by
// row, 64 per workgroup (the workgroup's number l = y * grid width + x), so
// that no workgroup is left with a few columns of a row end.
struct WaveColumn {
  int i, j;
  bool in;  // within the op's n0 x n1
};
TE_DEV WaveColumn wave_column(const KArg& a, int lane) {
  const Blk b = te_block(a);
  const uint32_t l = (uint32_t)b.y * a.lg[3] + (uint32_t)b.x, n0 = a.n[0];
  const uint32_t c = l * 64u + (uint32_t)lane, j = c / n0;
  return {(int)(c - j * n0), (int)j, c < n0 * a.n[1]};
}
// forward elimination of rows (A, C, R) over the wave's Lq levels: A becomes
// pa, R becomes pb
template <int LMAX>
TE_DEV void relay_forward(WaveRelay& r, int wave, int lane, int Lq, float (&A)[LMAX], const float (&C)[LMAX],
                          float (&R)[LMAX]) {
  relay_wait(&r.fwd_done, wave);
  if (Lq > 0) {
    float pa = 0.0f, pb = 0.0f;
    if (wave > 0) {
      pa = r.xb[0][lane];
      pb = r.xb[1][lane];
    }
#pragma unroll
    for (int t = 0; t < LMAX; ++t) {
      if (t < Lq) {
        const float cfa = A[t], cfc = C[t];
        const float cfb = (1.0f - cfa) - cfc;
        const float tt = recip(fma_(cfa, pa, cfb));
        pa = (-cfc) * tt;
        pb = fma_(-cfa, pb, R[t]) * tt;
        A[t] = pa;
        R[t] = pb;
      }
    }
    r.xb[0][lane] = pa;
    r.xb[1][lane] = pb;
  }
  relay_pass(&r.fwd_done, wave + 1);
}
// back substitution: R becomes the solution (top: the wave holds the last row)
template <int LMAX>
TE_DEV void relay_back(WaveRelay& r, int wave, int lane, int Lq, bool top, const float (&A)[LMAX], float (&R)[LMAX]) {
  relay_wait(&r.back_next, wave);
  if (Lq > 0) {
    float nb = top ? R[0] : r.xw[lane];
#pragma unroll
    for (int t = LMAX - 1; t >= 0; --t) {
      if (t < Lq) {
        nb = top && t == Lq - 1 ? R[t] : fma_(A[t], nb, R[t]);
        R[t] = nb;
      }
    }
    r.xw[lane] = nb;
  }
  relay_pass(&r.back_next, wave - 1);
}

// uv_tendency in that form (turb_u_tendency_lds / turb_v_tendency_lds): each
// wave forms dwd over its levels and one past (level nz for the top wave),
// its rows, the solution, and its levels' tendencies; all loads of a phase
// are issued before its stores.
template <int DIR, int LMAX>
__device__ __forceinline__ void uv_waves(uint32_t* M, const KArg& a) {
  __shared__ WaveRelay rl;
  const int lane = (int)threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const Dims g = dims(a.q[0], 0);
  const WaveColumn col = wave_column(a, lane);
  const int i = col.i, j = col.j, nz = g.nz;
  const int Lw = (nz + AC_WAVES - 1) / AC_WAVES, k0 = wave * Lw, k1 = min(nz, k0 + Lw), Lq = max(0, k1 - k0);
  // a wave past the top (Lq = 0, when AC_WAVES does not divide nz) touches no level
  const bool live = col.in && (DIR == 0 ? (i <= g.nx && j < g.ny) : (i < g.nx && j <= g.ny)) && Lq > 0;
  relay_init(rl);
  auto F = [&](int n) { return FieldB(reinterpret_cast<float*>(M) + a.p[n]); };
  const FieldB q = F(1), w = F(2), rho = F(3), rf = F(4), kmv = F(5), tA = F(6), tB = F(7), tC = F(8), dwd = F(14),
               ten = F(15);
  const float *mh = reinterpret_cast<float*>(M) + a.p[9], *mf = reinterpret_cast<float*>(M) + a.p[10],
              *hm = reinterpret_cast<float*>(M) + a.p[30];
  const float dt = ldf(a.q[11], 0);
  const int openx = ldi(a.q[16], 0), openy = ldi(a.q[16], 1), explicitZ = ldi(a.q[17], 0);
  // Addresses: the level as a wave-uniform displacement of the field's base
  // (L(k), scalar registers), the column as a 32-bit byte offset (o, with
  // the staggered neighbour at o - ost and the next / previous row at
  // o +- osj) shared by every level and field.
  const int ii = live ? i : 0, jj = live ? j : 0;
  const uint32_t o = 4u * (uint32_t)g.idx(ii, jj, 0), osj = 4u * (uint32_t)g.sJ, ost = DIR == 0 ? 4u : osj;
  const uint64_t osk = 4u * (uint64_t)g.sk();
  auto L = [&](int k) { return (uint64_t)k * osk; };
  const bool zeroX = DIR == 0 ? (openx != 0 && (i == 0 || i == g.nx)) : (openx != 0 && (i == 0 || i == g.nx - 1));
  const bool zeroY = DIR == 0 ? (openy != 0 && (j == 0 || j == g.ny - 1)) : (openy != 0 && (j == 0 || j == g.ny));
  const float hx = DIR == 0 ? hm_uf(hm, g, ii) : hm_uh(hm, g, ii), hy = DIR == 0 ? hm_vh(hm, g, jj) : hm_vf(hm, g, jj);
  auto txy = [&](uint64_t d) {
    const FieldB A = tA + d, B = tB + d;
    float tx, ty;
    if (DIR == 0) {
      tx = zeroX ? 0.0f : ((A(o) - A(o - 4u)) * g.rdx) * hx;
      ty = zeroY ? 0.0f : ((B(o + osj) - B(o)) * g.rdy) * hy;
    } else {
      tx = zeroX ? 0.0f : ((A(o + 4u) - A(o)) * g.rdx) * hx;
      ty = zeroY ? 0.0f : ((B(o) - B(o - osj)) * g.rdy) * hy;
    }
    return tx + ty;
  };
  if (explicitZ != 0) {
    if (live)
      for (int k = k0; k < k1; ++k) {
        const uint64_t d = L(k);
        const float tz = ((((tC + L(k + 1))(o) - (tC + d)(o)) * g.rdz) * 0.5f) * (mh[k] + mh[k]);
        const FieldB rk = rho + d;
        (ten + d).put(o, (ten + d)(o) + (txy(d) + tz) / (0.5f * (rk(o - ost) + rk(o))));
      }
    return;
  }
  const float tema = ((-0.0625f * dt) * g.rdz) * g.rdz, temc = (0.5f * dt) * g.rdz, rdt = 0.5f / dt;
  const float rdh = DIR == 0 ? g.rdx : g.rdy;
  const bool top = k1 == nz && Lq > 0;
  float CA[LMAX], CC[LMAX], R[LMAX], DW[LMAX + 1];
  if (live) {
    const float hmf = DIR == 0 ? hx : hy;
    // Level t's dwd and Y = ((kmv + kmv') (mf + mf)) (rf + rf'), then row
    // t - 1 (Y of its level and the next; level nz: dwd only), in chunks of
    // UV_CH levels: the chunk's loads (unconditional: slots past the wave's
    // range re-read its last level) all issued, then its arithmetic, each
    // value materialised where it is formed (te_pin) so that the loaded
    // operands die there.
    constexpr int UV_CH = 4;
    const float tCb = tC(o), tCt = (tC + L(nz))(o);
    float Yp = 0.0f;
#pragma unroll
    for (int t0 = 0; t0 <= LMAX; t0 += UV_CH) {
      float KA[UV_CH], KB[UV_CH], RA[UV_CH], RB[UV_CH], W0[UV_CH], W1[UV_CH], MF[UV_CH], P0[UV_CH], P1[UV_CH],
          MH[UV_CH], QK[UV_CH];
#pragma unroll
      for (int e = 0; e < UV_CH; ++e) {
        const int t = t0 + e;
        if (t > LMAX) break;
        const int k = min(k0 + t, nz);
        const uint64_t d = L(k);
        const FieldB km = kmv + d, rk = rf + d, wk = w + d;
        KA[e] = km(o - ost);
        KB[e] = km(o);
        RA[e] = rk(o - ost);
        RB[e] = rk(o);
        W0[e] = wk(o);
        W1[e] = wk(o - ost);
        MF[e] = mf[k];
        if (t > 0) {
          const int kr = min(k0 + t - 1, nz - 1);
          const uint64_t dr = L(kr);
          const FieldB ro = rho + dr;
          P0[e] = ro(o - ost);
          P1[e] = ro(o);
          MH[e] = mh[kr];
          QK[e] = (q + dr)(o);
        }
      }
#pragma unroll
      for (int e = 0; e < UV_CH; ++e) {
        const int t = t0 + e;
        if (t > LMAX) break;
        DW[t] = (((((W0[e] - W1[e]) * rdh) * hmf) * 0.25f) * (KA[e] + KB[e])) * (RA[e] + RB[e]);
        te_pin(DW[t]);
        float Yt = ((KA[e] + KB[e]) * (MF[e] + MF[e])) * (RA[e] + RB[e]);
        te_pin(Yt);
        if (t > 0) {
          const int r = t - 1, kr = min(k0 + r, nz - 1);
          const float rinvk = 1.0f / (0.5f * (P0[e] + P1[e]));
          const float mhk = MH[e], qk = QK[e];
          const float tem = (mhk + mhk) * rinvk;
          CA[r] = kr == 0 ? 0.0f : tema * (Yp * tem);
          CC[r] = kr < nz - 1 ? tema * (Yt * tem) : 0.0f;
          const float tb = kr == 0 ? -((temc * tCb) * (mhk + mhk)) : (temc * tCt) * (mhk + mhk);
          R[r] = kr == 0 || kr == nz - 1 ? fma_(tb, rinvk, qk) : qk;
          te_pin(CA[r]);
          te_pin(CC[r]);
          te_pin(R[r]);
        }
        Yp = Yt;
      }
    }
#pragma unroll
    for (int t = 0; t <= LMAX; ++t)
      if (t < Lq || (t == Lq && top)) (dwd + L(k0 + t)).put(o, DW[t]);
  }
  relay_forward<LMAX>(rl, wave, lane, Lq, CA, CC, R);
  // the horizontal divergences and old tendencies, loaded while the back
  // substitution comes down (into CC, free after the elimination); the
  // vertical difference of dwd in place of DW
  // (the top wave, first in the back substitution, loads them after it)
  float T0[LMAX];
  auto prefetch = [&]() {
    constexpr int PF_CH = 5;
#pragma unroll
    for (int t0 = 0; t0 < LMAX; t0 += PF_CH) {
      float A0[PF_CH], A1[PF_CH], B0[PF_CH], B1[PF_CH];
#pragma unroll
      for (int e = 0; e < PF_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const uint64_t d = L(min(k0 + t, nz - 1));
        const FieldB A = tA + d, B = tB + d;
        A0[e] = DIR == 0 ? A(o) : A(o + 4u);
        A1[e] = DIR == 0 ? A(o - 4u) : A(o);
        B0[e] = DIR == 0 ? B(o + osj) : B(o);
        B1[e] = DIR == 0 ? B(o) : B(o - osj);
        CC[t] = (ten + d)(o);
      }
#pragma unroll
      for (int e = 0; e < PF_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const float tx = zeroX ? 0.0f : ((A0[e] - A1[e]) * g.rdx) * hx;
        const float ty = zeroY ? 0.0f : ((B0[e] - B1[e]) * g.rdy) * hy;
        T0[t] = tx + ty;
        te_pin(T0[t]);
        DW[t] = ((DW[t + 1] - DW[t]) * g.rdz) * 0.5f;
      }
    }
  };
  if (live && !top) prefetch();
  relay_back<LMAX>(rl, wave, lane, Lq, top, CA, R);
  if (live && top) prefetch();
  if (live) {
    // every level's new tendency into R, then the stores (no load waits
    // behind a store); loads in chunks as above
    constexpr int FN_CH = 5;
#pragma unroll
    for (int t0 = 0; t0 < LMAX; t0 += FN_CH) {
      float P0[FN_CH], P1[FN_CH], QK[FN_CH], MH[FN_CH];
#pragma unroll
      for (int e = 0; e < FN_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const int k = min(k0 + t, nz - 1);
        const uint64_t d = L(k);
        const FieldB rk = rho + d;
        P0[e] = rk(o - ost);
        P1[e] = rk(o);
        QK[e] = (q + d)(o);
        MH[e] = mh[k];
      }
#pragma unroll
      for (int e = 0; e < FN_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const float ro = P0[e] + P1[e];
        float tz = (ro * (R[t] - QK[e])) * rdt;
        tz = fma_(DW[t], MH[e] + MH[e], tz);
        R[t] = fma_(T0[t] + tz, 1.0f / (0.5f * ro), CC[t]);
        te_pin(R[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < LMAX; ++t)
      if (t < Lq) (ten + L(k0 + t)).put(o, R[t]);
  }
}
#define UV_INST(L)                                                                                           \
  __global__ __launch_bounds__(64 * AC_WAVES, (L) <= 10 ? 6 : 1) void k_u_waves##L(uint32_t* M, KArg a) { uv_waves<0, L>(M, a); } \
  __global__ __launch_bounds__(64 * AC_WAVES, (L) <= 10 ? 6 : 1) void k_v_waves##L(uint32_t* M, KArg a) { uv_waves<1, L>(M, a); }
UV_INST(1) UV_INST(2) UV_INST(3) UV_INST(4) UV_INST(5) UV_INST(6) UV_INST(7) UV_INST(8) UV_INST(9) UV_INST(10)
UV_INST(11) UV_INST(12) UV_INST(13) UV_INST(14) UV_INST(15) UV_INST(16)
#undef UV_INST
const void* uv_waves_kernel(int dir, int levels_per_wave) {
#define E(L) {(const void*)&k_u_waves##L, (const void*)&k_v_waves##L}
  static const void* const fns[][2] = {{nullptr, nullptr}, E(1), E(2), E(3), E(4), E(5), E(6), E(7), E(8),
                                       E(9), E(10), E(11), E(12), E(13), E(14), E(15), E(16)};
#undef E
  return levels_per_wave >= 1 && levels_per_wave <= 16 && (dir == 0 || dir == 1) ? fns[levels_per_wave][dir] : nullptr;
}

__global__ __launch_bounds__(256) void k_turb_u_tendency(uint32_t* M, KArg a) {
  const Blk b_ = te_block(a);
  const int i = b_.x * TE_BX + (int)threadIdx.x, j = b_.y * TE_BY + (int)threadIdx.y;
  if (i < (int)a.n[0] && j < (int)a.n[1]) uv_tendency<0, false>(M, a, i, j);
}
__global__ __launch_bounds__(256) void k_turb_u_tendency_lds(uint32_t* M, KArg a) { uv_tendency_lds<0>(M, a); }
__global__ __launch_bounds__(256) void k_turb_v_tendency(uint32_t* M, KArg a) {
  const Blk b_ = te_block(a);
  const int i = b_.x * TE_BX + (int)threadIdx.x, j = b_.y * TE_BY + (int)threadIdx.y;
  if (i < (int)a.n[0] && j < (int)a.n[1]) uv_tendency<1, false>(M, a, i, j);
}
__global__ __launch_bounds__(256) void k_turb_v_tendency_lds(uint32_t* M, KArg a) { uv_tendency_lds<1>(M, a); }

// Implicit vertical solve shared by turb_w_tendency and tke_transport (both
// solve for a face field x on levels 1..nz-1 with the same matrix):
//   Z(k) = ((kmh[k] + kmh[k+1]) * mh[k]) * rho[k],   mfrf(k) = mf[k] / rf[k],
//   upper(k) = tema * (Z(k) * mfrf(k)),  lower(k) = tema * (Z(k-1) * mfrf(k)),
// first / last right-hand sides cfd1 / cfdT given by the caller; the
// back-substituted solution is handed to out(k, c, x_new, x_old, rf) in one
// descending sweep (levels nz-1 .. 1).  Loads run TS_CHUNK levels ahead.
template <class OUT, bool LDS>
__device__ __forceinline__ void face_implicit(const float* kmh, const float* mh, const float* rho, const float* mf,
                                              const float* rf, const float* x, const ColStore<LDS>& CA,
                                              const ColStore<LDS>& CB, int64_t base, int64_t sk, int nz, float tema,
                                              float cfd1, float cfdT, OUT out) {
  float Zp = ((kmh[base] + kmh[base + sk]) * mh[0]) * rho[base];  // Z(0)
  float Zk = ((kmh[base + sk] + kmh[base + 2 * sk]) * mh[1]) * rho[base + sk];
  float pa, pb;
  {
    const int64_t c = base + sk;
    const float mfrf1 = mf[1] / rf[c];
    float cfc = tema * (Zk * mfrf1);
    float cfb = 1.0f - cfc;
    float inv = 1.0f / cfb;
    pa = (-cfc) * inv;
    pb = cfd1 * inv;
    CA.at(1) = pa;
    CB.at(1) = pb;
  }
  Zp = Zk;  // Z(1)
  for (int k0 = 2; k0 < nz - 1; k0 += TS_CHUNK) {
    float ka[TS_CHUNK], kb[TS_CHUNK], mhk[TS_CHUNK], ro[TS_CHUNK], mfk[TS_CHUNK], rfk[TS_CHUNK], xv[TS_CHUNK];
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 + t;
      if (k < nz - 1) {
        const int64_t c = base + k * sk;
        ka[t] = kmh[c]; kb[t] = kmh[c + sk]; mhk[t] = mh[k]; ro[t] = rho[c]; mfk[t] = mf[k]; rfk[t] = rf[c];
        xv[t] = x[c];
      }
    }
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 + t;
      if (k < nz - 1) {
        const int64_t c = base + k * sk;
        const float mfrf = mfk[t] / rfk[t];
        Zk = ((ka[t] + kb[t]) * mhk[t]) * ro[t];
        float cfa = tema * (Zp * mfrf), cfc = tema * (Zk * mfrf);
        float cfb = (1.0f - cfa) - cfc;
        float inv = 1.0f / fma_(cfa, pa, cfb);
        pa = (-cfc) * inv;
        pb = fma_(-cfa, pb, xv[t]) * inv;
        CA.at(k) = pa;
        CB.at(k) = pb;
        Zp = Zk;
      }
    }
  }
  const int kt = nz - 1;
  const int64_t ct = base + kt * sk;
  {
    // Z(kt-1): Zp holds it when the middle sweep ran (nz > 3); else Z(1) = Z(kt-1)
    const float mfrfT = mf[kt] / rf[ct];
    float cfa = tema * (Zp * mfrfT);
    float cfb = 1.0f - cfa;
    float inv = 1.0f / fma_(cfa, pa, cfb);
    pb = fma_(-cfa, pb, cfdT) * inv;
    CB.at(kt) = pb;
    out(kt, ct, pb, x[ct], rf[ct]);
  }
  for (int k0 = nz - 2; k0 >= 1; k0 -= TS_CHUNK) {
    float ca[TS_CHUNK], cb[TS_CHUNK], xv[TS_CHUNK], rfk[TS_CHUNK];
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 - t;
      if (k >= 1) {
        const int64_t c = base + k * sk;
        ca[t] = CA.at(k); cb[t] = CB.at(k); xv[t] = x[c]; rfk[t] = rf[c];
        out.prefetch(t, c);
      }
    }
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 - t;
      if (k >= 1) {
        const int64_t c = base + k * sk;
        pb = fma_(ca[t], pb, cb[t]);
        CB.at(k) = pb;
        out.at(t, c, pb, xv[t], rfk[t]);
      }
    }
  }
}

// Tendency of a face field from the implicit solution plus horizontal flux
// divergences: ten += ((tx + ty) + (rf * (x_new - x_old)) * rdt) / rf.
struct FaceTendency {
  const float *hx, *hy, *hm;
  float* ten;
  Dims g;
  int i, j;
  bool zeroX, zeroY, negate;
  float rdt;
  float txv[TS_CHUNK], tyv[TS_CHUNK], tv[TS_CHUNK];
  __device__ void fluxes(int64_t c, float& tx, float& ty) const {
    if (negate) {  // tke_transport: -(f[i+1] - f[i])
      tx = zeroX ? 0.0f : ((-(hx[c + 1] - hx[c])) * g.rdx) * hm_uh(hm, g, i);
      ty = zeroY ? 0.0f : ((-(hy[c + g.sJ] - hy[c])) * g.rdy) * hm_vh(hm, g, j);
    } else {       // turb_w_tendency: (t13[i+1] - t13[i])
      tx = zeroX ? 0.0f : ((hx[c + 1] - hx[c]) * g.rdx) * hm_uh(hm, g, i);
      ty = zeroY ? 0.0f : ((hy[c + g.sJ] - hy[c]) * g.rdy) * hm_vh(hm, g, j);
    }
  }
  __device__ void apply(int64_t c, float xn, float xo, float rfc, float tx, float ty, float t0) const {
    float tz = (rfc * (xn - xo)) * rdt;
    ten[c] = t0 + ((tx + ty) + tz) / rfc;
  }
  __device__ void operator()(int, int64_t c, float xn, float xo, float rfc) {
    float tx, ty;
    fluxes(c, tx, ty);
    apply(c, xn, xo, rfc, tx, ty, ten[c]);
  }
  __device__ void prefetch(int t, int64_t c) {
    fluxes(c, txv[t], tyv[t]);
    tv[t] = ten[c];
  }
  __device__ void at(int t, int64_t c, float xn, float xo, float rfc) { apply(c, xn, xo, rfc, txv[t], tyv[t], tv[t]); }
};

// face_implicit in the AC_WAVES-wave form (relay_forward / relay_back):
// levels 1 .. nz-1 split into contiguous ranges, one per wave; each level's
// row is face_implicit's (cfa = 0 at level 1, cfc = 0 at level nz-1 give its
// first and last forms), then the tendency of each level from the solution.
template <int LMAX>
__device__ __forceinline__ void face_waves(WaveRelay& rl, bool live, int wave, int lane, const float* kmh,
                                           const float* mh, const float* rho, const float* mf, const float* rf,
                                           const float* x, int64_t base, int64_t sk, int nz, float tema, float cfd1,
                                           float cfdT, const FaceTendency& out) {
  const int Lw = (nz - 1 + AC_WAVES - 1) / AC_WAVES, k0 = 1 + wave * Lw, k1 = min(nz, k0 + Lw),
            Lq = max(0, k1 - k0);
  live = live && Lq > 0;
  const bool top = k1 == nz && Lq > 0;
  // addresses: the level as a wave-uniform displacement of each field's base,
  // the column as a 32-bit byte offset (o; o + 4 / o + osj the next column /
  // row) shared by every level and field
  const uint32_t o = 4u * (uint32_t)base, osj = 4u * (uint32_t)out.g.sJ;
  const uint64_t osk = 4u * (uint64_t)sk;
  auto L = [&](int k) { return (uint64_t)k * osk; };
  const FieldB KMH(kmh), RHO(rho), RF_(rf), X(x), HX(out.hx), HY(out.hy), TEN(out.ten);
  const float hxm = out.zeroX ? 0.0f : hm_uh(out.hm, out.g, live ? out.i : 0),
              hym = out.zeroY ? 0.0f : hm_vh(out.hm, out.g, live ? out.j : 0);
  float CA[LMAX], CC[LMAX], R[LMAX];
  if (live) {
    // Z(l) = ((kmh_l + kmh_l+1) mh_l) rho_l at l = k0-1 .. k1-1, then the
    // rows, in chunks of FW_CH levels: the chunk's loads (unconditional: slots
    // past the wave's range re-read its last level) all issued, then its
    // arithmetic, each value materialised where it is formed (te_pin).
    constexpr int FW_CH = 4;
    float Zp = 0.0f;
#pragma unroll
    for (int t0 = 0; t0 <= LMAX; t0 += FW_CH) {
      float K0[FW_CH], K1[FW_CH], MHl[FW_CH], RH[FW_CH], XV[FW_CH], RV[FW_CH], MF[FW_CH];
#pragma unroll
      for (int e = 0; e < FW_CH; ++e) {
        const int t = t0 + e;
        if (t > LMAX) break;
        const int l = min(k0 - 1 + t, nz - 1);
        const uint64_t d = L(l);
        K0[e] = (KMH + d)(o);
        K1[e] = (KMH + d + osk)(o);
        MHl[e] = mh[l];
        RH[e] = (RHO + d)(o);
        if (t > 0) {
          const int k = min(k0 + t - 1, nz - 1);
          const uint64_t dk = L(k);
          XV[e] = (X + dk)(o);
          RV[e] = (RF_ + dk)(o);
          MF[e] = mf[k];
        }
      }
#pragma unroll
      for (int e = 0; e < FW_CH; ++e) {
        const int t = t0 + e;
        if (t > LMAX) break;
        float Zt = ((K0[e] + K1[e]) * MHl[e]) * RH[e];
        te_pin(Zt);
        if (t > 0) {
          const int r = t - 1, k = min(k0 + r, nz - 1);
          const float mfrf = MF[e] / RV[e];
          CA[r] = k == 1 ? 0.0f : tema * (Zp * mfrf);
          CC[r] = k == nz - 1 ? 0.0f : tema * (Zt * mfrf);
          R[r] = k == 1 ? cfd1 : (k == nz - 1 ? cfdT : XV[e]);
          te_pin(CA[r]);
          te_pin(CC[r]);
        }
        Zp = Zt;
      }
    }
  }
  relay_forward<LMAX>(rl, wave, lane, Lq, CA, CC, R);
  // the horizontal divergences (tx + ty, into CC: free after the
  // elimination) and the old tendencies, loaded while the back substitution
  // comes down (FaceTendency::fluxes; the top wave, first in the back
  // substitution, loads them after it)
  float TV[LMAX];
  auto prefetch = [&]() {
    constexpr int PF_CH = 5;
#pragma unroll
    for (int t0 = 0; t0 < LMAX; t0 += PF_CH) {
      float A0[PF_CH], A1[PF_CH], B0[PF_CH], B1[PF_CH];
#pragma unroll
      for (int e = 0; e < PF_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const uint64_t d = L(min(k0 + t, nz - 1));
        const FieldB A = HX + d, B = HY + d;
        A0[e] = A(o + 4u);
        A1[e] = A(o);
        B0[e] = B(o + osj);
        B1[e] = B(o);
        TV[t] = (TEN + d)(o);
      }
#pragma unroll
      for (int e = 0; e < PF_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const float dx = A0[e] - A1[e], dy = B0[e] - B1[e];
        const float tx = out.zeroX ? 0.0f : ((out.negate ? -dx : dx) * out.g.rdx) * hxm;
        const float ty = out.zeroY ? 0.0f : ((out.negate ? -dy : dy) * out.g.rdy) * hym;
        CC[t] = tx + ty;
        te_pin(CC[t]);
      }
    }
  };
  if (live && !top) prefetch();
  relay_back<LMAX>(rl, wave, lane, Lq, top, CA, R);
  if (live && top) prefetch();
  if (live) {
    // FaceTendency::apply, x and rf loaded again (chunks as above)
    constexpr int FN_CH = 5;
#pragma unroll
    for (int t0 = 0; t0 < LMAX; t0 += FN_CH) {
      float XO[FN_CH], RF[FN_CH];
#pragma unroll
      for (int e = 0; e < FN_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const uint64_t d = L(min(k0 + t, nz - 1));
        XO[e] = (X + d)(o);
        RF[e] = (RF_ + d)(o);
      }
#pragma unroll
      for (int e = 0; e < FN_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const float tz = (RF[e] * (R[t] - XO[e])) * out.rdt;
        R[t] = TV[t] + (CC[t] + tz) / RF[e];
        te_pin(R[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < LMAX; ++t)
      if (t < Lq) (TEN + L(k0 + t)).put(o, R[t]);
  }
}

// turb_w_tendency.wgsl:152-287
// LMAX > 0: 64 x AC_WAVES workgroups, grid (columns / 64, rows): face_waves;
// else one lane per column
template <bool LDS, int LMAX = 0>
__device__ void w_tendency(uint32_t* M, const KArg& a) {
  const Blk b_ = te_block(a);
  const bool wv = LMAX > 0;
  const WaveColumn col = wv ? wave_column(a, (int)threadIdx.x) : WaveColumn{0, 0, true};
  const int i = wv ? col.i : b_.x * TE_BX + (int)threadIdx.x;
  const int j = wv ? col.j : b_.y * TE_BY + (int)threadIdx.y;
  Dims g = dims(a.q[0], 0);
  const bool live = col.in && i < (int)a.n[0] && j < (int)a.n[1] && i < g.nx && j < g.ny;
  if (!live && !wv) return;
  const float *w = fq(a.q[1]), *rho = fq(a.q[2]), *rf = fq(a.q[3]), *kmh = fq(a.q[4]), *t13 = fq(a.q[5]),
              *t23 = fq(a.q[6]), *t33 = fq(a.q[7]), *mh = fq(a.q[8]), *mf = fq(a.q[9]), *hm = fq(a.q[30]);
  float dt = ldf(a.q[10], 0);
  float *cA = fq(a.q[11]), *cB = fq(a.q[12]), *wten = fq(a.q[13]);
  int openx = ldi(a.q[14], 0), openy = ldi(a.q[14], 1), explicitZ = ldi(a.q[15], 0);
  const int64_t sj = g.sJ, sk = g.sk(), base = g.idx(i, j, 0);
  bool zeroX = openx != 0 && (i == 0 || i == g.nx - 1);
  bool zeroY = openy != 0 && (j == 0 || j == g.ny - 1);
  auto turbxy = [&](int64_t c, float& tx, float& ty) {
    tx = zeroX ? 0.0f : ((t13[c + 1] - t13[c]) * g.rdx) * hm_uh(hm, g, i);
    ty = zeroY ? 0.0f : ((t23[c + sj] - t23[c]) * g.rdy) * hm_vh(hm, g, j);
  };
  if (explicitZ != 0) {  // uniform: no wait is reached; in AC_WAVES-wave blocks wave 0 does the columns
    if (!live || (wv && threadIdx.y != 0)) return;
    for (int k = 1; k < g.nz; ++k) {
      const int64_t c = base + k * sk;
      float tx, ty;
      turbxy(c, tx, ty);
      float tz = ((t33[c] - t33[c - sk]) * g.rdz) * mf[k];
      wten[c] = wten[c] + ((tx + ty) + tz) / rf[c];
    }
    return;
  }
  float tema = ((-1.0f * dt) * g.rdz) * g.rdz;
  float temc = dt * g.rdz;
  float rdt = 1.0f / dt;
  const int nz = g.nz;
  const int64_t c1 = base + sk, ct = base + (nz - 1) * sk;
  const float cfd1 = live ? fma_(-(temc * t33[c1 - sk]), mf[1] / rf[c1], w[c1]) : 0.0f;
  const float cfdT = live ? fma_(temc * t33[ct], mf[nz - 1] / rf[ct], w[ct]) : 0.0f;
  FaceTendency out{t13, t23, hm, wten, g, i, j, zeroX, zeroY, false, rdt, {}, {}, {}};
  if constexpr (LMAX > 0) {
    __shared__ WaveRelay rl;
    relay_init(rl);
    face_waves<LMAX>(rl, live, __builtin_amdgcn_readfirstlane((int)threadIdx.y), (int)threadIdx.x, kmh, mh, rho, mf,
                     rf, w, base, sk, nz, tema, cfd1, cfdT, out);
    return;
  }
  const ColStore<LDS> CA{cA, base, sk, TE_LANE, TE_NLANES, 0}, CB{cB, base, sk, TE_LANE, TE_NLANES, nz};
  face_implicit(kmh, mh, rho, mf, rf, w, CA, CB, base, sk, nz, tema, cfd1, cfdT, out);
}
__global__ __launch_bounds__(256) void k_turb_w_tendency(uint32_t* M, KArg a) { w_tendency<false>(M, a); }
__global__ __launch_bounds__(256) void k_turb_w_tendency_lds(uint32_t* M, KArg a) { w_tendency<true>(M, a); }

// turb_scalar_flux_h.wgsl:125-148
__global__ __launch_bounds__(256) void k_turb_scalar_flux_h(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz;
  if (i > g.nx || j > g.ny || k >= g.nz) return;
  const float *s = fq(a.q[1]), *rho = fq(a.q[2]), *khh = fq(a.q[3]), *hm = fq(a.q[30]);
  float leNum = ldf(a.q[4], 0);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  // both fluxes formed before the first store (the loads go out together)
  const float fx = (((((-0.125f * (rho[c] + rho[c - 1])) * leNum) * ((khh[c] + khh[c - 1]) + (khh[c + sk] + khh[c - 1 + sk]))) *
                     (s[c] - s[c - 1])) * g.rdx) * hm_uf(hm, g, i);
  const float fy = (((((-0.125f * (rho[c] + rho[c - sj])) * leNum) * ((khh[c] + khh[c - sj]) + (khh[c + sk] + khh[c - sj + sk]))) *
                     (s[c] - s[c - sj])) * g.rdy) * hm_vf(hm, g, j);
  fq(a.q[5])[c] = fx;
  fq(a.q[6])[c] = fy;
}

// turb_scalar_vert_implicit.wgsl:210-327
template <bool LDS>
__device__ void scalar_vert_implicit(uint32_t* M, const KArg& a) {
  const Blk b_ = te_block(a);
  const int i = b_.x * TE_BX + (int)threadIdx.x, j = b_.y * TE_BY + (int)threadIdx.y;
  if (i >= (int)a.n[0] || j >= (int)a.n[1]) return;
  Dims g = dims(a.q[0], 0);
  if (i >= g.nx || j >= g.ny) return;
  const float *s = fq(a.q[1]), *rho = fq(a.q[2]), *rr = fq(a.q[3]), *rf = fq(a.q[4]), *khv = fq(a.q[5]),
              *mh = fq(a.q[6]), *mf = fq(a.q[7]), *fx = fq(a.q[8]), *fy = fq(a.q[9]), *sflux = fq(a.q[14]),
              *hm = fq(a.q[30]);
  float dt = ldf(a.q[10], 0);
  float *cA = fq(a.q[11]), *cB = fq(a.q[12]), *sten = fq(a.q[13]);
  int useFlux = ldi(a.q[15], 0), explicitZ = ldi(a.q[16], 0), bcturbs = ldi(a.q[18], 0);
  const uint32_t* ow = a.q[17];
  float wall0 = ldf(ow, 0), wall1 = ldf(ow, 1), wall2 = ldf(ow, 2), wall3 = ldf(ow, 3);
  const int64_t sj = g.sJ, sk = g.sk(), base = g.idx(i, j, 0);
  auto L2 = [&](int k) {
    const int64_t cc = base + k * sk;
    return ((((-khv[cc]) * (s[cc] - s[cc - sk])) * g.rdz) * mf[k]) * rf[cc];
  };
  auto hdiv = [&](int64_t c) {
    return ((-(fx[c + 1] - fx[c])) * g.rdx) * hm_uh(hm, g, i) + ((-(fy[c + sj] - fy[c])) * g.rdy) * hm_vh(hm, g, j);
  };
  if (explicitZ != 0) {
    auto L1 = [&](int kf) -> float {
      const int64_t c = base + kf * sk;
      if (kf == 0) {
        if (wall3 == 0.0f) return useFlux != 0 ? sflux[base] * rf[c] : (bcturbs == 2 ? L2(1) : 0.0f);
        if (wall3 == 2.0f) return (wall0 * wall1) * rf[c];
        return (((((-wall0) * 2.0f) * (s[c] - wall1)) * g.rdz) * mf[0]) * rf[c];
      }
      if (kf == g.nz) {
        if (wall3 == 0.0f) return bcturbs == 2 ? L2(g.nz - 1) : 0.0f;
        const int64_t cb = base + (g.nz - 1) * sk;
        if (wall3 == 2.0f) return (wall0 * wall2) * rf[c];
        return (((((-wall0) * 2.0f) * (wall2 - s[cb])) * g.rdz) * mf[g.nz]) * rf[c];
      }
      return L2(kf);
    };
    float below = L1(0);
    for (int k = 0; k < g.nz; ++k) {
      const int64_t c = base + k * sk;
      float above = L1(k + 1);
      float tz = ((-(above - below)) * g.rdz) * mh[k];
      sten[c] = fma_(hdiv(c) + tz, rr[c], sten[c]);
      below = above;
    }
    return;
  }
  float tema = ((-dt) * g.rdz) * g.rdz;
  float rdt = 1.0f / dt;
  float temc = dt * g.rdz;
  // L4(k) = ((X(k) * mh[k]) * rr[k]) and L3(k) = ((X(k+1) * mh[k]) * rr[k]) with
  // X(k) = (khv[k] * mf[k]) * rf[k]: X(k+1) is evaluated once and carried.
  const int nz = g.nz;
  const ColStore<LDS> CA{cA, base, sk, TE_LANE, TE_NLANES, 0}, CB{cB, base, sk, TE_LANE, TE_NLANES, nz};
  auto X = [&](int k) { const int64_t c = base + k * sk; return (khv[c] * mf[k]) * rf[c]; };
  float pa, pb;
  float Xk1 = X(1);  // X(k+1) for k = 0
  {
    float cfc = tema * ((Xk1 * mh[0]) * rr[base]);
    float cfb = 1.0f - cfc;
    float t = 1.0f / cfb;
    pa = (-cfc) * t;
    CA.at(0) = pa;
    float cfd = s[base];
    if (useFlux != 0 || bcturbs == 2) {
      float dumz;
      if (useFlux != 0) dumz = sflux[base] * rf[base];
      else {
        const int64_t c1 = base + sk;
        dumz = ((((-khv[c1]) * (s[c1] - s[base])) * g.rdz) * mf[1]) * rf[c1];
      }
      cfd = fma_((temc * dumz) * mh[0], rr[base], s[base]);
    }
    pb = cfd * t;
    CB.at(0) = pb;
  }
  for (int k0 = 1; k0 < nz - 1; k0 += TS_CHUNK) {
    float kh[TS_CHUNK], rfu[TS_CHUNK], mfu[TS_CHUNK], mhk[TS_CHUNK], rrk[TS_CHUNK], sk_[TS_CHUNK];
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 + t;
      if (k < nz - 1) {
        const int64_t c = base + k * sk;
        kh[t] = khv[c + sk]; rfu[t] = rf[c + sk]; mfu[t] = mf[k + 1]; mhk[t] = mh[k]; rrk[t] = rr[c]; sk_[t] = s[c];
      }
    }
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 + t;
      if (k < nz - 1) {
        const int64_t c = base + k * sk;
        const float Xk = Xk1;
        Xk1 = (kh[t] * mfu[t]) * rfu[t];
        float cfa = tema * ((Xk * mhk[t]) * rrk[t]), cfc = tema * ((Xk1 * mhk[t]) * rrk[t]);
        float cfb = (1.0f - cfa) - cfc;
        float tt = 1.0f / fma_(cfa, pa, cfb);
        pa = (-cfc) * tt;
        pb = fma_(-cfa, pb, sk_[t]) * tt;
        CA.at(k) = pa;
        CB.at(k) = pb;
      }
    }
  }
  const int kt = nz - 1;
  const int64_t ct = base + kt * sk;
  {
    float cfa = tema * ((Xk1 * mh[kt]) * rr[ct]);
    float cfb = 1.0f - cfa;
    float t = 1.0f / fma_(cfa, pa, cfb);
    float cfd = s[ct];
    if (bcturbs == 2) {
      float dumz = ((((-khv[ct]) * (s[ct] - s[ct - sk])) * g.rdz) * mf[kt]) * rf[ct];
      cfd = fma_(-((temc * dumz) * mh[kt]), rr[ct], s[ct]);
    }
    pb = fma_(-cfa, pb, cfd) * t;
    CB.at(kt) = pb;
    sten[ct] = fma_(hdiv(ct) + (rho[ct] * (pb - s[ct])) * rdt, rr[ct], sten[ct]);
  }
  for (int k0 = nz - 2; k0 >= 0; k0 -= TS_CHUNK) {
    float ca[TS_CHUNK], cb[TS_CHUNK], hd[TS_CHUNK], ro[TS_CHUNK], sv[TS_CHUNK], rv[TS_CHUNK], tv[TS_CHUNK];
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 - t;
      if (k >= 0) {
        const int64_t c = base + k * sk;
        ca[t] = CA.at(k); cb[t] = CB.at(k); hd[t] = hdiv(c); ro[t] = rho[c]; sv[t] = s[c]; rv[t] = rr[c]; tv[t] = sten[c];
      }
    }
#pragma unroll
    for (int t = 0; t < TS_CHUNK; ++t) {
      const int k = k0 - t;
      if (k >= 0) {
        const int64_t c = base + k * sk;
        pb = fma_(ca[t], pb, cb[t]);
        CB.at(k) = pb;
        sten[c] = fma_(hd[t] + (ro[t] * (pb - sv[t])) * rdt, rv[t], tv[t]);
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_turb_scalar_vert_implicit(uint32_t* M, KArg a) { scalar_vert_implicit<false>(M, a); }
__global__ __launch_bounds__(256) void k_turb_scalar_vert_implicit_lds(uint32_t* M, KArg a) {
  scalar_vert_implicit<true>(M, a);
}

// turb_scalar_vert_implicit as a launch of 64 columns of a row x AC_WAVES
// waves per workgroup (the fused program's form): each wave takes a
// contiguous range of LMAX levels, forms its rows' coefficients and
// right-hand sides and the explicit parts of the tendency in registers, the
// Thomas recurrence passes from wave to wave through LDS (forward upwards,
// back substitution downwards; wave w waits for fwd_done == w and for
// back_next == w, set by the wave below / above once its share is done, with
// the waits bounded all the same), and each wave updates its levels' tendency
// as soon as its back substitution is done.  Values are the sequential
// sweep's, level by level: the first row's missing sub-diagonal and the last
// row's missing super-diagonal are zeros, which give the reference's special
// rows' bits.  cA / cB are not written (dead scratch of the fused program).
template <int LMAX>
__device__ __forceinline__ void svi_waves(uint32_t* M, const KArg& a) {
  __shared__ WaveRelay rl;
  const int lane = (int)threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const Dims g = dims(a.q[0], 0);
  const WaveColumn col = wave_column(a, lane);
  const int i = col.i, j = col.j, nz = g.nz;
  const int Lw = (nz + AC_WAVES - 1) / AC_WAVES, k0 = wave * Lw, k1 = min(nz, k0 + Lw), Lq = max(0, k1 - k0);
  // a wave past the top (Lq = 0, when AC_WAVES does not divide nz) touches no level
  const bool active = col.in && i < g.nx && j < g.ny && Lq > 0;
  relay_init(rl);
  auto F = [&](int q) { return reinterpret_cast<float*>(M) + a.p[q]; };
  const float *s = F(1), *rho = F(2), *rr = F(3), *rf = F(4), *khv = F(5), *mh = F(6), *mf = F(7), *fx = F(8),
              *fy = F(9), *sflux = F(14), *hm = F(30);
  float* sten = F(13);
  const float dt = ldf(a.q[10], 0);
  const int useFlux = ldi(a.q[15], 0), explicitZ = ldi(a.q[16], 0), bcturbs = ldi(a.q[18], 0);
  const uint32_t* ow = a.q[17];
  const float wall0 = ldf(ow, 0), wall1 = ldf(ow, 1), wall2 = ldf(ow, 2), wall3 = ldf(ow, 3);
  const int64_t sj = g.sJ, sk = g.sk(), base = g.idx(active ? i : 0, active ? j : 0, 0);
  const float uh = hm_uh(hm, g, active ? i : 0), vh = hm_vh(hm, g, active ? j : 0);
  auto hdiv = [&](int64_t c) {
    return ((-(fx[c + 1] - fx[c])) * g.rdx) * uh + ((-(fy[c + sj] - fy[c])) * g.rdy) * vh;
  };
  auto L2 = [&](int k) {
    const int64_t cc = base + k * sk;
    return ((((-khv[cc]) * (s[cc] - s[cc - sk])) * g.rdz) * mf[k]) * rf[cc];
  };
  if (explicitZ != 0) {  // explicit vertical diffusion: every level on its own
    if (active) {
      auto L1 = [&](int kf) -> float {
        const int64_t c = base + kf * sk;
        if (kf == 0) {
          if (wall3 == 0.0f) return useFlux != 0 ? sflux[base] * rf[c] : (bcturbs == 2 ? L2(1) : 0.0f);
          if (wall3 == 2.0f) return (wall0 * wall1) * rf[c];
          return (((((-wall0) * 2.0f) * (s[c] - wall1)) * g.rdz) * mf[0]) * rf[c];
        }
        if (kf == nz) {
          if (wall3 == 0.0f) return bcturbs == 2 ? L2(nz - 1) : 0.0f;
          const int64_t cb = base + (nz - 1) * sk;
          if (wall3 == 2.0f) return (wall0 * wall2) * rf[c];
          return (((((-wall0) * 2.0f) * (wall2 - s[cb])) * g.rdz) * mf[nz]) * rf[c];
        }
        return L2(kf);
      };
      for (int k = k0; k < k1; ++k) {
        const int64_t c = base + k * sk;
        const float tz = ((-(L1(k + 1) - L1(k))) * g.rdz) * mh[k];
        sten[c] = fma_(hdiv(c) + tz, rr[c], sten[c]);
      }
    }
    return;
  }
  const float tema = ((-dt) * g.rdz) * g.rdz, rdt = 1.0f / dt, temc = dt * g.rdz;
  // addresses: the level as a wave-uniform displacement of each field's base,
  // the column as a 32-bit byte offset (o; o + 4 / o + osj the next column /
  // row) shared by every level and field
  const uint32_t o = 4u * (uint32_t)base, osj = 4u * (uint32_t)sj;
  const uint64_t osk = 4u * (uint64_t)sk;
  auto L = [&](int k) { return (uint64_t)k * osk; };
  const FieldB S_(s), RHO(rho), RR_(rr), RF(rf), KHV(khv), FX(fx), FY(fy), STEN(sten);
  // rows of the wave's levels: sub-diagonal CA, super-diagonal CC, right-hand
  // side D.  Every level's loads are unconditional (slots past the wave's
  // range re-read its last level), four levels at a time (sched_barrier),
  // each value materialised where it is formed (te_pin); only the stores are
  // guarded.
  const bool top = k1 == nz && Lq > 0;
  float CA[LMAX], CC[LMAX], D[LMAX];
  if (active) {
    auto X = [&](int k) { const uint64_t d = L(k); return ((KHV + d)(o) * mf[k]) * (RF + d)(o); };
    // the first / last rows' right-hand sides where a boundary flux enters
    const bool flux0 = k0 == 0 && (useFlux != 0 || bcturbs == 2), fluxT = top && bcturbs == 2;
    float d0 = 0.0f, dT = 0.0f;
    if (flux0) {
      float dumz;
      if (useFlux != 0) dumz = sflux[base] * rf[base];
      else {
        const int64_t c1 = base + sk;
        dumz = ((((-khv[c1]) * (s[c1] - s[base])) * g.rdz) * mf[1]) * rf[c1];
      }
      d0 = fma_((temc * dumz) * mh[0], rr[base], s[base]);
    }
    if (fluxT) {
      const int64_t c = base + (nz - 1) * sk;
      const float dumz = ((((-khv[c]) * (s[c] - s[c - sk])) * g.rdz) * mf[nz - 1]) * rf[c];
      dT = fma_(-((temc * dumz) * mh[nz - 1]), rr[c], s[c]);
    }
    float Xk = X(k0 == 0 ? 1 : k0);  // X(k) (k0 = 0: not used)
    constexpr int SV_CH = 4;
#pragma unroll
    for (int t0 = 0; t0 < LMAX; t0 += SV_CH) {
      float RV[SV_CH], MH[SV_CH], KN[SV_CH], MFN[SV_CH], RFN[SV_CH], SV[SV_CH];
#pragma unroll
      for (int e = 0; e < SV_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const int k = min(k0 + t, nz - 1), kn = min(k + 1, nz - 1);
        const uint64_t d = L(k), dn = L(kn);
        RV[e] = (RR_ + d)(o);
        MH[e] = mh[k];
        KN[e] = (KHV + dn)(o);
        MFN[e] = mf[kn];
        RFN[e] = (RF + dn)(o);
        SV[e] = (S_ + d)(o);
      }
#pragma unroll
      for (int e = 0; e < SV_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const int k = min(k0 + t, nz - 1);
        const float r = RV[e], m = MH[e];
        float Xn = (KN[e] * MFN[e]) * RFN[e];
        Xn = k + 1 < nz ? Xn : 0.0f;
        te_pin(Xn);
        CA[t] = k == 0 ? 0.0f : tema * ((Xk * m) * r);
        CC[t] = k == nz - 1 ? 0.0f : tema * (((k == 0 ? Xk : Xn) * m) * r);
        D[t] = k == 0 && flux0 ? d0 : (k == nz - 1 && fluxT ? dT : SV[e]);
        te_pin(CA[t]);
        te_pin(CC[t]);
        te_pin(D[t]);
        Xk = k == 0 ? Xk : Xn;
      }
    }
  }
  relay_forward<LMAX>(rl, wave, lane, Lq, CA, CC, D);
  // the horizontal divergence (into CC, free after the elimination) and the
  // old tendency, loaded while the back substitution comes down (the top
  // wave, first in the back substitution, loads them after it)
  float TV[LMAX];
  auto prefetch = [&]() {
    constexpr int PF_CH = 5;
#pragma unroll
    for (int t0 = 0; t0 < LMAX; t0 += PF_CH) {
      float A0[PF_CH], A1[PF_CH], B0[PF_CH], B1[PF_CH];
#pragma unroll
      for (int e = 0; e < PF_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const uint64_t d = L(min(k0 + t, nz - 1));
        const FieldB A = FX + d, B = FY + d;
        A0[e] = A(o + 4u);
        A1[e] = A(o);
        B0[e] = B(o + osj);
        B1[e] = B(o);
        TV[t] = (STEN + d)(o);
      }
#pragma unroll
      for (int e = 0; e < PF_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        CC[t] = ((-(A0[e] - A1[e])) * g.rdx) * uh + ((-(B0[e] - B1[e])) * g.rdy) * vh;
        te_pin(CC[t]);
      }
    }
  };
  if (active && !top) prefetch();
  relay_back<LMAX>(rl, wave, lane, Lq, top, CA, D);
  if (active && top) prefetch();
  if (active) {
    // every level's new tendency into D, then the stores
    constexpr int FN_CH = 5;
#pragma unroll
    for (int t0 = 0; t0 < LMAX; t0 += FN_CH) {
      float RH[FN_CH], SO[FN_CH], RV[FN_CH];
#pragma unroll
      for (int e = 0; e < FN_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        const uint64_t d = L(min(k0 + t, nz - 1));
        RH[e] = (RHO + d)(o);
        SO[e] = (S_ + d)(o);
        RV[e] = (RR_ + d)(o);
      }
#pragma unroll
      for (int e = 0; e < FN_CH; ++e) {
        const int t = t0 + e;
        if (t >= LMAX) break;
        D[t] = fma_(CC[t] + (RH[e] * (D[t] - SO[e])) * rdt, RV[e], TV[t]);
        te_pin(D[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < LMAX; ++t)
      if (t < Lq) (STEN + L(k0 + t)).put(o, D[t]);
  }
}
#define SVI_INST(L) \
  __global__ __launch_bounds__(64 * AC_WAVES, (L) <= 10 ? 6 : 1) void k_svi_waves##L(uint32_t* M, KArg a) { svi_waves<L>(M, a); }
SVI_INST(1) SVI_INST(2) SVI_INST(3) SVI_INST(4) SVI_INST(5) SVI_INST(6) SVI_INST(7) SVI_INST(8) SVI_INST(9)
SVI_INST(10) SVI_INST(11) SVI_INST(12) SVI_INST(13) SVI_INST(14) SVI_INST(15) SVI_INST(16)
#undef SVI_INST
const void* svi_waves_kernel(int levels_per_wave) {
  static const void* const fns[] = {nullptr, (const void*)&k_svi_waves1, (const void*)&k_svi_waves2,
                                    (const void*)&k_svi_waves3, (const void*)&k_svi_waves4,
                                    (const void*)&k_svi_waves5, (const void*)&k_svi_waves6,
                                    (const void*)&k_svi_waves7, (const void*)&k_svi_waves8,
                                    (const void*)&k_svi_waves9, (const void*)&k_svi_waves10,
                                    (const void*)&k_svi_waves11, (const void*)&k_svi_waves12,
                                    (const void*)&k_svi_waves13, (const void*)&k_svi_waves14,
                                    (const void*)&k_svi_waves15, (const void*)&k_svi_waves16};
  return levels_per_wave >= 1 && levels_per_wave <= 16 ? fns[levels_per_wave] : nullptr;
}

// tke_flux_h.wgsl:125-146
__global__ __launch_bounds__(256) void k_tke_flux_h(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz + 1;
  if (i > g.nx || j > g.ny || k > g.nz - 1) return;
  const float *t = fq(a.q[1]), *rf = fq(a.q[2]), *kmh = fq(a.q[3]), *hm = fq(a.q[30]);
  const int64_t sj = g.sJ, c = g.idx(i, j, k);
  // both fluxes formed before the first store
  const float fx =
      (((((-0.25f * (rf[c] + rf[c - 1])) * 2.0f) * (kmh[c] + kmh[c - 1])) * (t[c] - t[c - 1])) * g.rdx) * hm_uf(hm, g, i);
  const float fy =
      (((((-0.25f * (rf[c] + rf[c - sj])) * 2.0f) * (kmh[c] + kmh[c - sj])) * (t[c] - t[c - sj])) * g.rdy) * hm_vf(hm, g, j);
  fq(a.q[4])[c] = fx;
  fq(a.q[5])[c] = fy;
}

// tke_transport.wgsl:158-301
// LMAX > 0: 64 x AC_WAVES workgroups, grid (columns / 64, rows): face_waves;
// else one lane per column
template <bool LDS, int LMAX = 0>
__device__ void tke_transport(uint32_t* M, const KArg& a) {
  const Blk b_ = te_block(a);
  const bool wv = LMAX > 0;
  const WaveColumn col = wv ? wave_column(a, (int)threadIdx.x) : WaveColumn{0, 0, true};
  const int i = wv ? col.i : b_.x * TE_BX + (int)threadIdx.x;
  const int j = wv ? col.j : b_.y * TE_BY + (int)threadIdx.y;
  Dims g = dims(a.q[0], 0);
  const bool live = col.in && i < (int)a.n[0] && j < (int)a.n[1] && i < g.nx && j < g.ny;
  if (!live && !wv) return;
  const float *t = fq(a.q[1]), *rho = fq(a.q[2]), *rf = fq(a.q[3]), *kmh = fq(a.q[4]), *kmv = fq(a.q[5]),
              *mh = fq(a.q[6]), *mf = fq(a.q[7]), *fx = fq(a.q[11]), *fy = fq(a.q[12]), *hm = fq(a.q[30]);
  float dt = ldf(a.q[8], 0);
  float *cA = fq(a.q[9]), *cB = fq(a.q[10]), *tketen = fq(a.q[13]);
  int openx = ldi(a.q[14], 0), openy = ldi(a.q[14], 1), explicitZ = ldi(a.q[15], 0);
  const int64_t sj = g.sJ, sk = g.sk(), base = g.idx(i, j, 0);
  bool zeroX = openx != 0 && (i == 0 || i == g.nx - 1);
  bool zeroY = openy != 0 && (j == 0 || j == g.ny - 1);
  auto turbxy = [&](int64_t c, float& tx, float& ty) {
    tx = zeroX ? 0.0f : ((-(fx[c + 1] - fx[c])) * g.rdx) * hm_uh(hm, g, i);
    ty = zeroY ? 0.0f : ((-(fy[c + sj] - fy[c])) * g.rdy) * hm_vh(hm, g, j);
  };
  auto mfrf = [&](int k) { return mf[k] / rf[base + k * sk]; };
  if (explicitZ != 0) {  // uniform: no wait is reached; in AC_WAVES-wave blocks wave 0 does the columns
    if (!live || (wv && threadIdx.y != 0)) return;
    auto L1 = [&](int k) {
      const int64_t c = base + k * sk;
      return ((((-(kmv[c] + kmv[c + sk])) * (t[c + sk] - t[c])) * g.rdz) * mh[k]) * rho[c];
    };
    float below = L1(0);
    for (int k = 1; k < g.nz; ++k) {
      const int64_t c = base + k * sk;
      float here = L1(k);
      float tx, ty;
      turbxy(c, tx, ty);
      float tz = ((-(here - below)) * g.rdz) * mf[k];
      tketen[c] = tketen[c] + ((tx + ty) + tz) / rf[c];
      below = here;
    }
    return;
  }
  float tema = ((-1.0f * dt) * g.rdz) * g.rdz;
  float temc = dt * g.rdz;
  float rdt = 1.0f / dt;
  const int nz = g.nz;
  const int64_t c1 = base + sk, ct = base + (nz - 1) * sk;
  float cfd1 = 0.0f, cfdT = 0.0f;
  if (live) {
    const float tem1 = ((((-(kmv[c1 - sk] + kmv[c1])) * (t[c1] - t[c1 - sk])) * g.rdz) * mh[0]) * rho[c1 - sk];
    cfd1 = fma_(temc * tem1, mf[1] / rf[c1], t[c1]);
    const float temT = ((((-(kmv[ct] + kmv[ct + sk])) * (t[ct + sk] - t[ct])) * g.rdz) * mh[nz - 1]) * rho[ct];
    cfdT = fma_(-(temc * temT), mf[nz - 1] / rf[ct], t[ct]);
  }
  FaceTendency out{fx, fy, hm, tketen, g, i, j, zeroX, zeroY, true, rdt, {}, {}, {}};
  if constexpr (LMAX > 0) {
    __shared__ WaveRelay rl;
    relay_init(rl);
    face_waves<LMAX>(rl, live, __builtin_amdgcn_readfirstlane((int)threadIdx.y), (int)threadIdx.x, kmh, mh, rho, mf,
                     rf, t, base, sk, nz, tema, cfd1, cfdT, out);
    return;
  }
  const ColStore<LDS> CA{cA, base, sk, TE_LANE, TE_NLANES, 0}, CB{cB, base, sk, TE_LANE, TE_NLANES, nz};
  face_implicit(kmh, mh, rho, mf, rf, t, CA, CB, base, sk, nz, tema, cfd1, cfdT, out);
}
__global__ __launch_bounds__(256) void k_tke_transport(uint32_t* M, KArg a) { tke_transport<false>(M, a); }
__global__ __launch_bounds__(256) void k_tke_transport_lds(uint32_t* M, KArg a) { tke_transport<true>(M, a); }
#define FACE_INST(L)                                                                                      \
  __global__ __launch_bounds__(64 * AC_WAVES, (L) <= 10 ? 6 : 1) void k_w_waves##L(uint32_t* M, KArg a) { w_tendency<true, L>(M, a); } \
  __global__ __launch_bounds__(64 * AC_WAVES, (L) <= 10 ? 6 : 1) void k_tke_waves##L(uint32_t* M, KArg a) { tke_transport<true, L>(M, a); }
FACE_INST(1) FACE_INST(2) FACE_INST(3) FACE_INST(4) FACE_INST(5) FACE_INST(6) FACE_INST(7) FACE_INST(8)
FACE_INST(9) FACE_INST(10) FACE_INST(11) FACE_INST(12) FACE_INST(13) FACE_INST(14) FACE_INST(15) FACE_INST(16)
#undef FACE_INST
const void* face_waves_kernel(int which, int levels_per_wave) {
#define E(L) {(const void*)&k_w_waves##L, (const void*)&k_tke_waves##L}
  static const void* const fns[][2] = {{nullptr, nullptr}, E(1), E(2), E(3), E(4), E(5), E(6), E(7), E(8),
                                       E(9), E(10), E(11), E(12), E(13), E(14), E(15), E(16)};
#undef E
  return levels_per_wave >= 1 && levels_per_wave <= 16 && (which == 0 || which == 1) ? fns[levels_per_wave][which]
                                                                                      : nullptr;
}

}  // namespace te
