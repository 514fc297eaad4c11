// Step kernels, one per reference kernel: diagnostics, resets, mass fluxes,
// divergence, RK update, buoyancy, pressure tendency, end-of-step updates.
// Every kernel has the signature (M, KArg) so a recorded or encoded program can
// launch it by name; the thread mapping follows the reference dispatch grid.
#include "te_device.hpp"
#include "te_kernels.hpp"

namespace te {

// update_diagnostics_moist.wgsl:128-184
__global__ __launch_bounds__(256) void k_update_diagnostics_moist(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k >= g.nz) return;
  const float *thP = fp(M, a.p[1]), *ppP = fp(M, a.p[2]), *th0 = fp(M, a.p[3]), *pi0 = fp(M, a.p[4]),
              *qv0 = fp(M, a.p[5]), *qc0 = fp(M, a.p[6]), *qi0 = fp(M, a.p[7]), *rth0 = fp(M, a.p[8]),
              *qv = fp(M, a.p[9]), *qc = fp(M, a.p[10]), *qr = fp(M, a.p[11]), *qi = fp(M, a.p[12]),
              *qs = fp(M, a.p[13]), *qg = fp(M, a.p[14]);
  float *thv = fp(M, a.p[15]), *ppterm = fp(M, a.p[16]), *piadv = fp(M, a.p[17]), *thterm = fp(M, a.p[18]),
        *buoy = fp(M, a.p[19]);
  int64_t c = g.idx(i, j, k);
  const float reps = RV / RD, repsm1 = reps - 1.0f;
  float qvv = mx(qv[c], 0.0f);
  float liq = qc[c] + qr[c];
  float ice = (qi[c] + qs[c]) + qg[c];
  float qli = mx(0.0f, liq + ice);
  float cpli = fma_(CPL, mx(0.0f, liq), CPI * mx(0.0f, ice));
  float th = th0[k] + thP[c];
  float pi = pi0[k] + ppP[c];
  float cpm = fma_(CPV, qvv, CP) + cpli;
  float cvm = 1.0f / (fma_(CVV, qvv, CV) + cpli);
  thv[c] = (th * fma_(reps, qvv, 1.0f)) / ((1.0f + qvv) + qli);
  piadv[c] = pi;
  ppterm[c] = ((pi * ROVCP) * cpm) * cvm;
  thterm[c] = (th * fma_(-ROVCP, cpm, fma_(RV, qvv, RD))) * cvm;
  buoy[c] = GRAV * (fma_(thP[c], rth0[k], repsm1 * (qvv - qv0[k])) - ((qli - qc0[k]) - qi0[k]));
}

// update_diagnostics.wgsl:114-144 (dry)
__global__ __launch_bounds__(256) void k_update_diagnostics(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k >= g.nz) return;
  const float *thP = fp(M, a.p[1]), *ppP = fp(M, a.p[2]), *th0 = fp(M, a.p[3]), *pi0 = fp(M, a.p[4]),
              *rth0 = fp(M, a.p[8]);
  float *thv = fp(M, a.p[5]), *ppterm = fp(M, a.p[6]), *piadv = fp(M, a.p[7]), *buoy = fp(M, a.p[9]);
  const float rddcv = RD / (CP - RD);
  int64_t c = g.idx(i, j, k);
  float pi = pi0[k] + ppP[c];
  thv[c] = th0[k] + thP[c];
  piadv[c] = pi;
  ppterm[c] = pi * rddcv;
  buoy[c] = (GRAV * thP[c]) * rth0[k];
}

// stage_tendency_reset.wgsl:110-129
__global__ __launch_bounds__(256) void k_stage_tendency_reset(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  int64_t c = g.idx(i, j, k);
  fp(M, a.p[1])[c] = 0.0f;
  fp(M, a.p[2])[c] = 0.0f;
  fp(M, a.p[3])[c] = 0.0f;
}

// stage_state_reset.wgsl:110-138
__global__ __launch_bounds__(256) void k_stage_state_reset(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  int64_t c = g.idx(i, j, k);
  // u <- u_a, v <- v_a, w <- w_a, pp3d <- ppi, ppd <- ppx
  const int src[5] = {1, 2, 3, 4, 10};
  for (int q = 0; q < 5; ++q) fp(M, a.p[5 + q])[c] = fp(M, a.p[src[q]])[c];
}

// massflux_from_state.wgsl:111-137
__global__ __launch_bounds__(256) void k_massflux_from_state(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  const float *u = fp(M, a.p[1]), *v = fp(M, a.p[2]), *w = fp(M, a.p[3]), *rho0 = fp(M, a.p[4]), *rf0 = fp(M, a.p[5]);
  float *rru = fp(M, a.p[6]), *rrv = fp(M, a.p[7]), *rrw = fp(M, a.p[8]);
  int64_t c = g.idx(i, j, k);
  if (k < g.nz) {
    rru[c] = rho0[k] * u[c];
    rrv[c] = rho0[k] * v[c];
  }
  rrw[c] = (k == 0 || k == g.nz) ? 0.0f : rf0[k] * w[c];
}

// velocity_to_massflux.wgsl:111-134
__global__ __launch_bounds__(256) void k_velocity_to_massflux(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  const float *rho0 = fp(M, a.p[1]), *rf0 = fp(M, a.p[2]);
  float *rru = fp(M, a.p[3]), *rrv = fp(M, a.p[4]), *rrw = fp(M, a.p[5]);
  int64_t c = g.idx(i, j, k);
  if (k < g.nz) {
    rru[c] = rru[c] * rho0[k];
    rrv[c] = rrv[c] * rho0[k];
  }
  rrw[c] = (k == 0 || k == g.nz) ? 0.0f : rrw[c] * rf0[k];
}

// div_massflux.wgsl:125-147
__global__ __launch_bounds__(256) void k_div_massflux(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 2 || j > g.ny + 2 || k >= g.nz) return;
  const float *rru = fp(M, a.p[1]), *rrv = fp(M, a.p[2]), *rrw = fp(M, a.p[3]), *mh = fp(M, a.p[4]),
              *hm = fp(M, a.p[30]);
  float* divx = fp(M, a.p[5]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  divx[c] = fma_((rrw[c + sk] - rrw[c]) * g.rdz, mh[k],
                 fma_((rru[c + 1] - rru[c]) * g.rdx, hm_uh(hm, g, i), ((rrv[c + sj] - rrv[c]) * g.rdy) * hm_vh(hm, g, j)));
}

// divergence3d.wgsl:125-147
__global__ __launch_bounds__(256) void k_divergence3d(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx, j = gy, k = gz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const float *u = fp(M, a.p[1]), *v = fp(M, a.p[2]), *w = fp(M, a.p[3]), *mh = fp(M, a.p[4]), *hm = fp(M, a.p[30]);
  float* dv = fp(M, a.p[5]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  float d = fma_((w[c + sk] - w[c]) * g.rdz, mh[k],
                 fma_((u[c + 1] - u[c]) * g.rdx, hm_uh(hm, g, i), ((v[c + sj] - v[c]) * g.rdy) * hm_vh(hm, g, j)));
  dv[c] = (ab(d) < 1e-30f) ? 0.0f : d;
}

// buoyancy_apply.wgsl:111-129
__global__ __launch_bounds__(256) void k_buoyancy_apply(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx, j = gy, k = gz + 1;
  if (i >= g.nx || j >= g.ny || k > g.nz - 1) return;
  const float *buoy = fp(M, a.p[1]), *c1 = fp(M, a.p[2]), *c2 = fp(M, a.p[3]);
  float* wten = fp(M, a.p[4]);
  const int64_t sk = g.sk(), c = g.idx(i, j, k);
  wten[c] = wten[c] + fma_(c1[k], buoy[c - sk], c2[k] * buoy[c]);
}

// rk_update.wgsl:111-132
__global__ __launch_bounds__(256) void k_rk_update(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int stag = ldi(M, a.p[1]);
  int i = gx, j = gy, k = gz;
  int iMax = stag == 2 ? g.nx : g.nx - 1, jMax = stag == 3 ? g.ny : g.ny - 1, kLo = stag == 4 ? 1 : 0;
  if (i > iMax || j > jMax || k < kLo || k >= g.nz) return;
  const float *phiA = fp(M, a.p[2]), *tend = fp(M, a.p[3]);
  float* phi = fp(M, a.p[4]);
  float dttmp = ldf(M, a.p[5]);
  int64_t c = g.idx(i, j, k);
  phi[c] = fma_(dttmp, tend[c], phiA[c]);
}

// build_sadv.wgsl:111-127
__global__ __launch_bounds__(256) void k_build_sadv(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k >= g.nz) return;
  const float *th0 = fp(M, a.p[1]), *thP = fp(M, a.p[2]);
  int64_t c = g.idx(i, j, k);
  fp(M, a.p[3])[c] = (th0[k] - 3e+02f) + thP[c];
}

// apply_thterm.wgsl:111-127
__global__ __launch_bounds__(256) void k_apply_thterm(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  const float *dv = fp(M, a.p[1]), *thterm = fp(M, a.p[2]);
  float* tend = fp(M, a.p[3]);
  int64_t c = g.idx(gx, gy, gz);
  tend[c] = fma_(-dv[c], thterm[c], tend[c]);
}

// turb_add_tendency.wgsl:111-127
__global__ __launch_bounds__(256) void k_turb_add_tendency(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  const float* src = fp(M, a.p[1]);
  float* dst = fp(M, a.p[2]);
  float scale = ldf(M, a.p[3]);
  int64_t c = g.idx(i, j, k);
  dst[c] = fma_(scale, src[c], dst[c]);
}

// turb_zero.wgsl:103-116
__global__ __launch_bounds__(256) void k_turb_zero(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  if (gx >= g.sJ || gy >= g.sK || gz > g.nz + 1) return;
  fp(M, a.p[1])[gx + (int64_t)g.sJ * (gy + (int64_t)g.sK * gz)] = 0.0f;
}

// tke_advance.wgsl:111-129
__global__ __launch_bounds__(256) void k_tke_advance(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx, j = gy, k = gz + 1;
  if (i >= g.nx || j >= g.ny || k > g.nz - 1) return;
  const float *tkea = fp(M, a.p[1]), *tketen = fp(M, a.p[2]);
  float dttmp = ldf(M, a.p[3]);
  int64_t c = g.idx(i, j, k);
  float v = fma_(dttmp, tketen[c], tkea[c]);
  fp(M, a.p[4])[c] = (v < 1e-10f) ? 0.0f : v;
}

// ppx_finish.wgsl:111-126
__global__ __launch_bounds__(256) void k_ppx_finish(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  int64_t c = g.idx(gx, gy, gz);
  float* ppx = fp(M, a.p[2]);
  ppx[c] = fp(M, a.p[1])[c] + ppx[c];
}

// diabatic_unapply.wgsl:111-140
__global__ __launch_bounds__(256) void k_diabatic_unapply(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  float dt = ldf(M, a.p[1]);
  int64_t c = g.idx(gx, gy, gz);
  for (int q = 0; q < 4; ++q) {
    float* s = fp(M, a.p[2 + q]);
    float* r = fp(M, a.p[6 + q]);
    s[c] = fma_(-dt, r[c], s[c]);
    r[c] = s[c];
  }
}

// diabatic_rates.wgsl:111-137
__global__ __launch_bounds__(256) void k_diabatic_rates(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  float dt = ldf(M, a.p[1]);
  float rdt = 1.0f / dt;
  int64_t c = g.idx(gx, gy, gz);
  for (int q = 0; q < 4; ++q) {
    const float* s = fp(M, a.p[2 + q]);
    float* r = fp(M, a.p[6 + q]);
    r[c] = (s[c] - r[c]) * rdt;
  }
}

// pressure_tendency.wgsl:234-291
__global__ __launch_bounds__(256) void k_pressure_tendency(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  const float *th = fp(M, a.p[1]), *qv = fp(M, a.p[2]), *rho = fp(M, a.p[3]), *ppi = fp(M, a.p[4]),
              *thten = fp(M, a.p[5]), *qpten = fp(M, a.p[6]), *th0 = fp(M, a.p[7]), *pi0 = fp(M, a.p[8]),
              *qvt = fp(M, a.p[13]);
  float* ppten = fp(M, a.p[11]);
  float dt = ldf(M, a.p[9]);
  int hasDiabatic = ldi(M, a.p[10]), hasStepTend = ldi(M, a.p[12]), moist = ldi(M, a.p[14]);
  const int k = gz;
  int64_t c = g.idx(gx, gy, gz);
  const float rp00 = 1.0f / P00, rddcv = RD / (CP - RD);
  float tem = 0.0001f * 0.0001f;
  float tht = hasStepTend != 0 ? thten[c] : 0.0f;
  if (moist == 0) {
    float p = 0.0f;
    if (ab(dt * tht) > tem) {
      float thnew = fma_(dt, tht, th[c]);
      float pinew = rr_pow(((rho[c] * (th0[k] + thnew)) * RD) * rp00, rddcv) - pi0[k];
      p = (pinew - ppi[c]) / dt;
    }
    ppten[c] = p;
    return;
  }
  float qvten = qvt[c];
  float p = 0.0f;
  if (ab(dt * tht) > tem || ab(dt * qvten) > 1e-14f) {
    float thnew = fma_(dt, tht, th[c]);
    float qvnew = fma_(dt, qvten, qv[c]);
    float pinew = rr_pow(((rho[c] * (th0[k] + thnew)) * fma_(RV, qvnew, RD)) * rp00, rddcv) - pi0[k];
    p = (pinew - ppi[c]) / dt;
  }
  if (hasDiabatic != 0) p = p + qpten[c];
  ppten[c] = p;
}

// update_rho_prs.wgsl:234-267
__global__ __launch_bounds__(256) void k_update_rho_prs(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  int moist = ldi(M, a.p[1]);
  const float *thP = fp(M, a.p[2]), *ppP = fp(M, a.p[3]), *th0 = fp(M, a.p[4]), *pi0 = fp(M, a.p[5]),
              *qv = fp(M, a.p[6]);
  float *prs = fp(M, a.p[7]), *rho = fp(M, a.p[8]);
  const int k = gz;
  int64_t c = g.idx(gx, gy, gz);
  const float cpdrd = CP / RD;
  float pi = pi0[k] + ppP[c];
  float th = th0[k] + thP[c];
  float p = P00 * rr_pow(pi, cpdrd);
  prs[c] = p;
  float rmix = moist != 0 ? fma_(mx(0.0f, qv[c]), RV, RD) : RD;
  rho[c] = p / ((th * pi) * rmix);
}

}  // namespace te
