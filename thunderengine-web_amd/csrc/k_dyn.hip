// Step kernels, one per reference kernel: diagnostics, resets, mass fluxes,
// divergence, RK update, buoyancy, pressure tendency, end-of-step updates.
// Every kernel has the signature (M, KArg) so a recorded or encoded program can
// launch it by name; the thread mapping follows the reference dispatch grid.
#include "te_device.hpp"
#include "te_kernels.hpp"

namespace te {

// update_diagnostics_moist.wgsl:128-184
__global__ __launch_bounds__(256) void k_update_diagnostics_moist(uint32_t* M, KArg a) {
  {
    const Dims g = dims(a.q[0], 0);
    if (te_flat_levels(a, (uint32_t)(g.sJ * g.sK), [&](uint32_t w, int width, int k) {
          if (k >= g.nz) return;
          const float reps = RV / RD, repsm1 = reps - 1.0f;
          const float t0 = fq(a.q[3])[k], p0 = fq(a.q[4])[k], qv0k = fq(a.q[5])[k], qc0k = fq(a.q[6])[k],
                      qi0k = fq(a.q[7])[k], rt0 = fq(a.q[8])[k];
          const F4 thP = ld4(a.q[1], w, width), ppP = ld4(a.q[2], w, width), qv = ld4(a.q[9], w, width),
                   qc = ld4(a.q[10], w, width), qr = ld4(a.q[11], w, width), qi = ld4(a.q[12], w, width),
                   qs = ld4(a.q[13], w, width), qg = ld4(a.q[14], w, width);
          F4 thv, ppterm, piadv, thterm, buoy;
          for (int e = 0; e < width; ++e) {
            const float qvv = mx(qv.v[e], 0.0f);
            const float liq = qc.v[e] + qr.v[e];
            const float ice = (qi.v[e] + qs.v[e]) + qg.v[e];
            const float qli = mx(0.0f, liq + ice);
            const float cpli = fma_(CPL, mx(0.0f, liq), CPI * mx(0.0f, ice));
            const float th = t0 + thP.v[e];
            const float pi = p0 + ppP.v[e];
            const float cpm = fma_(CPV, qvv, CP) + cpli;
            const float cvm = 1.0f / (fma_(CVV, qvv, CV) + cpli);
            thv.v[e] = (th * fma_(reps, qvv, 1.0f)) / ((1.0f + qvv) + qli);
            piadv.v[e] = pi;
            ppterm.v[e] = ((pi * ROVCP) * cpm) * cvm;
            thterm.v[e] = (th * fma_(-ROVCP, cpm, fma_(RV, qvv, RD))) * cvm;
            buoy.v[e] = GRAV * (fma_(thP.v[e], rt0, repsm1 * (qvv - qv0k)) - ((qli - qc0k) - qi0k));
          }
          st4(a.q[15], w, width, thv);
          st4(a.q[16], w, width, ppterm);
          st4(a.q[17], w, width, piadv);
          st4(a.q[18], w, width, thterm);
          st4(a.q[19], w, width, buoy);
        }))
      return;
  }
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k >= g.nz) return;
  const float *thP = fq(a.q[1]), *ppP = fq(a.q[2]), *th0 = fq(a.q[3]), *pi0 = fq(a.q[4]),
              *qv0 = fq(a.q[5]), *qc0 = fq(a.q[6]), *qi0 = fq(a.q[7]), *rth0 = fq(a.q[8]),
              *qv = fq(a.q[9]), *qc = fq(a.q[10]), *qr = fq(a.q[11]), *qi = fq(a.q[12]),
              *qs = fq(a.q[13]), *qg = fq(a.q[14]);
  float *thv = fq(a.q[15]), *ppterm = fq(a.q[16]), *piadv = fq(a.q[17]), *thterm = fq(a.q[18]),
        *buoy = fq(a.q[19]);
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
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k >= g.nz) return;
  const float *thP = fq(a.q[1]), *ppP = fq(a.q[2]), *th0 = fq(a.q[3]), *pi0 = fq(a.q[4]),
              *rth0 = fq(a.q[8]);
  float *thv = fq(a.q[5]), *ppterm = fq(a.q[6]), *piadv = fq(a.q[7]), *buoy = fq(a.q[9]);
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
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  int64_t c = g.idx(i, j, k);
  fq(a.q[1])[c] = 0.0f;
  fq(a.q[2])[c] = 0.0f;
  fq(a.q[3])[c] = 0.0f;
}

// stage_state_reset.wgsl:110-138
__global__ __launch_bounds__(256) void k_stage_state_reset(uint32_t* M, KArg a) {
  if (te_flat(a, [&](uint32_t w, int width) {
        const int src[5] = {1, 2, 3, 4, 10};
        for (int q = 0; q < 5; ++q) st4(a.q[5 + q], w, width, ld4(a.q[src[q]], w, width));
      }))
    return;
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  int64_t c = g.idx(i, j, k);
  // u <- u_a, v <- v_a, w <- w_a, pp3d <- ppi, ppd <- ppx
  const int src[5] = {1, 2, 3, 4, 10};
  for (int q = 0; q < 5; ++q) fq(a.q[5 + q])[c] = fq(a.q[src[q]])[c];
}

// massflux_from_state.wgsl:111-137
__global__ __launch_bounds__(256) void k_massflux_from_state(uint32_t* M, KArg a) {
  {
    const Dims g = dims(a.q[0], 0);
    if (te_flat_levels(a, (uint32_t)(g.sJ * g.sK), [&](uint32_t w, int width, int k) {
          if (k < g.nz) {
            const float r = fq(a.q[4])[k];
            F4 u = ld4(a.q[1], w, width), v = ld4(a.q[2], w, width);
            for (int e = 0; e < width; ++e) {
              u.v[e] = r * u.v[e];
              v.v[e] = r * v.v[e];
            }
            st4(a.q[6], w, width, u);
            st4(a.q[7], w, width, v);
          }
          F4 m = ld4(a.q[3], w, width);
          const bool edge = k == 0 || k == g.nz;
          const float rf = edge ? 0.0f : fq(a.q[5])[k];
          for (int e = 0; e < width; ++e) m.v[e] = edge ? 0.0f : rf * m.v[e];
          st4(a.q[8], w, width, m);
        }))
      return;
  }
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  const float *u = fq(a.q[1]), *v = fq(a.q[2]), *w = fq(a.q[3]), *rho0 = fq(a.q[4]), *rf0 = fq(a.q[5]);
  float *rru = fq(a.q[6]), *rrv = fq(a.q[7]), *rrw = fq(a.q[8]);
  int64_t c = g.idx(i, j, k);
  if (k < g.nz) {
    rru[c] = rho0[k] * u[c];
    rrv[c] = rho0[k] * v[c];
  }
  rrw[c] = (k == 0 || k == g.nz) ? 0.0f : rf0[k] * w[c];
}

// massflux_from_state then div_massflux (div_massflux.wgsl:125-147) in one
// launch: each cell stores its mass fluxes and forms the divergence from its
// own and its i+1, j+1, k+1 neighbours' fluxes, each the product
// massflux_from_state stores there (so the same bits), instead of reading them
// back.  p: 1-8 as massflux_from_state's, 9 mh, 10 divx.  Grid: whole.
__global__ __launch_bounds__(256) void k_massflux_div(uint32_t* M, KArg a) {
  TE_TID3
  const Dims g = dims(a.q[0], 0);
  const int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  const float *u = fq(a.q[1]), *v = fq(a.q[2]), *w = fq(a.q[3]), *rho0 = fq(a.q[4]), *rf0 = fq(a.q[5]),
              *mh = fq(a.q[9]), *hm = fq(a.q[30]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  auto fw = [&](int kk, int64_t cc) { return (kk == 0 || kk == g.nz) ? 0.0f : rf0[kk] * w[cc]; };
  const bool div = i <= g.nx + 2 && j <= g.ny + 2 && k < g.nz;
  const float r = k < g.nz ? rho0[k] : 0.0f;
  const float uc = k < g.nz ? u[c] : 0.0f, vc = k < g.nz ? v[c] : 0.0f;
  const float up = div ? u[c + 1] : 0.0f, vp = div ? v[c + sj] : 0.0f;
  const float mw = fw(k, c), mwp = div ? fw(k + 1, c + sk) : 0.0f;
  const float mu = r * uc, mv = r * vc;
  // the divergence formed before the first store (its profile and metric
  // loads go out with the others)
  const float d = div ? fma_((mwp - mw) * g.rdz, mh[k],
                             fma_((r * up - mu) * g.rdx, hm_uh(hm, g, i), ((r * vp - mv) * g.rdy) * hm_vh(hm, g, j)))
                      : 0.0f;
  if (k < g.nz) {
    fq(a.q[6])[c] = mu;
    fq(a.q[7])[c] = mv;
  }
  fq(a.q[8])[c] = mw;
  if (div) fq(a.q[10])[c] = d;
}

// velocity_to_massflux.wgsl:111-134
__global__ __launch_bounds__(256) void k_velocity_to_massflux(uint32_t* M, KArg a) {
  {
    const Dims g = dims(a.q[0], 0);
    if (te_flat_levels(a, (uint32_t)(g.sJ * g.sK), [&](uint32_t w, int width, int k) {
          if (k < g.nz) {
            const float r = fq(a.q[1])[k];
            F4 u = ld4(a.q[3], w, width), v = ld4(a.q[4], w, width);
            for (int e = 0; e < width; ++e) {
              u.v[e] = u.v[e] * r;
              v.v[e] = v.v[e] * r;
            }
            st4(a.q[3], w, width, u);
            st4(a.q[4], w, width, v);
          }
          const bool edge = k == 0 || k == g.nz;
          F4 m = ld4(a.q[5], w, width);
          const float rf = edge ? 0.0f : fq(a.q[2])[k];
          for (int e = 0; e < width; ++e) m.v[e] = edge ? 0.0f : m.v[e] * rf;
          st4(a.q[5], w, width, m);
        }))
      return;
  }
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  const float *rho0 = fq(a.q[1]), *rf0 = fq(a.q[2]);
  float *rru = fq(a.q[3]), *rrv = fq(a.q[4]), *rrw = fq(a.q[5]);
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
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 2 || j > g.ny + 2 || k >= g.nz) return;
  const float *rru = fq(a.q[1]), *rrv = fq(a.q[2]), *rrw = fq(a.q[3]), *mh = fq(a.q[4]),
              *hm = fq(a.q[30]);
  float* divx = fq(a.q[5]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  divx[c] = fma_((rrw[c + sk] - rrw[c]) * g.rdz, mh[k],
                 fma_((rru[c + 1] - rru[c]) * g.rdx, hm_uh(hm, g, i), ((rrv[c + sj] - rrv[c]) * g.rdy) * hm_vh(hm, g, j)));
}

// divergence3d.wgsl:125-147
__global__ __launch_bounds__(256) void k_divergence3d(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const float *u = fq(a.q[1]), *v = fq(a.q[2]), *w = fq(a.q[3]), *mh = fq(a.q[4]), *hm = fq(a.q[30]);
  float* dv = fq(a.q[5]);
  const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, k);
  float d = fma_((w[c + sk] - w[c]) * g.rdz, mh[k],
                 fma_((u[c + 1] - u[c]) * g.rdx, hm_uh(hm, g, i), ((v[c + sj] - v[c]) * g.rdy) * hm_vh(hm, g, j)));
  dv[c] = (ab(d) < 1e-30f) ? 0.0f : d;
}

// buoyancy_apply.wgsl:111-129
__global__ __launch_bounds__(256) void k_buoyancy_apply(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz + 1;
  if (i >= g.nx || j >= g.ny || k > g.nz - 1) return;
  const float *buoy = fq(a.q[1]), *c1 = fq(a.q[2]), *c2 = fq(a.q[3]);
  float* wten = fq(a.q[4]);
  const int64_t sk = g.sk(), c = g.idx(i, j, k);
  wten[c] = wten[c] + fma_(c1[k], buoy[c - sk], c2[k] * buoy[c]);
}

// rk_update.wgsl:111-132
__global__ __launch_bounds__(256) void k_rk_update(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int stag = ldi(a.q[1], 0);
  int i = gx, j = gy, k = gz;
  int iMax = stag == 2 ? g.nx : g.nx - 1, jMax = stag == 3 ? g.ny : g.ny - 1, kLo = stag == 4 ? 1 : 0;
  if (i > iMax || j > jMax || k < kLo || k >= g.nz) return;
  const float *phiA = fq(a.q[2]), *tend = fq(a.q[3]);
  float* phi = fq(a.q[4]);
  float dttmp = ldf(a.q[5], 0);
  int64_t c = g.idx(i, j, k);
  phi[c] = fma_(dttmp, tend[c], phiA[c]);
}

// build_sadv.wgsl:111-127
__global__ __launch_bounds__(256) void k_build_sadv(uint32_t* M, KArg a) {
  {
    const Dims g = dims(a.q[0], 0);
    if (te_flat_levels(a, (uint32_t)(g.sJ * g.sK), [&](uint32_t w, int width, int k) {
          if (k >= g.nz) return;
          const float t0 = fq(a.q[1])[k] - 3e+02f;
          F4 s = ld4(a.q[2], w, width);
          for (int e = 0; e < width; ++e) s.v[e] = t0 + s.v[e];
          st4(a.q[3], w, width, s);
        }))
      return;
  }
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k >= g.nz) return;
  const float *th0 = fq(a.q[1]), *thP = fq(a.q[2]);
  int64_t c = g.idx(i, j, k);
  fq(a.q[3])[c] = (th0[k] - 3e+02f) + thP[c];
}

// apply_thterm.wgsl:111-127
__global__ __launch_bounds__(256) void k_apply_thterm(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  const float *dv = fq(a.q[1]), *thterm = fq(a.q[2]);
  float* tend = fq(a.q[3]);
  int64_t c = g.idx(gx, gy, gz);
  tend[c] = fma_(-dv[c], thterm[c], tend[c]);
}

// turb_add_tendency.wgsl:111-127
__global__ __launch_bounds__(256) void k_turb_add_tendency(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  const float* src = fq(a.q[1]);
  float* dst = fq(a.q[2]);
  float scale = ldf(a.q[3], 0);
  int64_t c = g.idx(i, j, k);
  dst[c] = fma_(scale, src[c], dst[c]);
}

// turb_zero.wgsl:103-116
__global__ __launch_bounds__(256) void k_turb_zero(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  if (gx >= g.sJ || gy >= g.sK || gz > g.nz + 1) return;
  fq(a.q[1])[gx + (int64_t)g.sJ * (gy + (int64_t)g.sK * gz)] = 0.0f;
}

// tke_advance.wgsl:111-129
__global__ __launch_bounds__(256) void k_tke_advance(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx, j = gy, k = gz + 1;
  if (i >= g.nx || j >= g.ny || k > g.nz - 1) return;
  const float *tkea = fq(a.q[1]), *tketen = fq(a.q[2]);
  float dttmp = ldf(a.q[3], 0);
  int64_t c = g.idx(i, j, k);
  float v = fma_(dttmp, tketen[c], tkea[c]);
  fq(a.q[4])[c] = (v < 1e-10f) ? 0.0f : v;
}

// ppx_finish.wgsl:111-126
__global__ __launch_bounds__(256) void k_ppx_finish(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  int64_t c = g.idx(gx, gy, gz);
  float* ppx = fq(a.q[2]);
  ppx[c] = fq(a.q[1])[c] + ppx[c];
}
// ppx_finish, then the halo fill of ppx on a periodic domain (nx, ny >= 4;
// timestep.js:1593 haloFill): each cell writes its periodic images.
__global__ __launch_bounds__(256) void k_ppx_finish_xy(uint32_t* M, KArg a) {
  TE_TID3
  const Dims g = dims(a.q[0], 0);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  const float* pp = fq(a.q[1]);
  float* ppx = fq(a.q[2]);
  update_with_images(g, ppx, gx, gy, gz, 0, g.nz, [&](int64_t c) { return pp[c] + ppx[c]; });
}

// diabatic_unapply.wgsl:111-140
__global__ __launch_bounds__(256) void k_diabatic_unapply(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  float dt = ldf(a.q[1], 0);
  int64_t c = g.idx(gx, gy, gz);
  for (int q = 0; q < 4; ++q) {
    float* s = fq(a.q[2 + q]);
    float* r = fq(a.q[6 + q]);
    s[c] = fma_(-dt, r[c], s[c]);
    r[c] = s[c];
  }
}

// diabatic_rates.wgsl:111-137
__global__ __launch_bounds__(256) void k_diabatic_rates(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  float dt = ldf(a.q[1], 0);
  float rdt = 1.0f / dt;
  int64_t c = g.idx(gx, gy, gz);
  for (int q = 0; q < 4; ++q) {
    const float* s = fq(a.q[2 + q]);
    float* r = fq(a.q[6 + q]);
    r[c] = (s[c] - r[c]) * rdt;
  }
}

// pressure_tendency.wgsl:234-291
__global__ __launch_bounds__(256) void k_pressure_tendency(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  const float *th = fq(a.q[1]), *qv = fq(a.q[2]), *rho = fq(a.q[3]), *ppi = fq(a.q[4]),
              *thten = fq(a.q[5]), *qpten = fq(a.q[6]), *th0 = fq(a.q[7]), *pi0 = fq(a.q[8]),
              *qvt = fq(a.q[13]);
  float* ppten = fq(a.q[11]);
  float dt = ldf(a.q[9], 0);
  int hasDiabatic = ldi(a.q[10], 0), hasStepTend = ldi(a.q[12], 0), moist = ldi(a.q[14], 0);
  const int k = gz;
  int64_t c = g.idx(gx, gy, gz);
  const float rp00 = 1.0f / P00, rddcv = RD / (CP - RD);
  float tem = 0.0001f * 0.0001f;
  float tht = hasStepTend != 0 ? thten[c] : 0.0f;
  if (moist == 0) {
    float p = 0.0f;
    if (ab(dt * tht) > tem) {
      float thnew = fma_(dt, tht, th[c]);
      float pinew = fast_pow(((rho[c] * (th0[k] + thnew)) * RD) * rp00, rddcv) - pi0[k];
      p = (pinew - ppi[c]) / dt;
    }
    ppten[c] = p;
    return;
  }
  float qvten = qvt[c];
  // every operand loaded with the tendencies (one wait for memory), whether
  // or not the cell takes the update
  const float thc = th[c], qvc = qv[c], rhoc = rho[c], ppic = ppi[c], qp = hasDiabatic != 0 ? qpten[c] : 0.0f;
  float p = 0.0f;
  if (ab(dt * tht) > tem || ab(dt * qvten) > 1e-14f) {
    float thnew = fma_(dt, tht, thc);
    float qvnew = fma_(dt, qvten, qvc);
    float pinew = fast_pow(((rhoc * (th0[k] + thnew)) * fma_(RV, qvnew, RD)) * rp00, rddcv) - pi0[k];
    p = (pinew - ppic) / dt;
  }
  if (hasDiabatic != 0) p = p + qp;
  ppten[c] = p;
}

// update_rho_prs.wgsl:234-267
__global__ __launch_bounds__(256) void k_update_rho_prs(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  if (gx >= g.nx || gy >= g.ny || gz >= g.nz) return;
  int moist = ldi(a.q[1], 0);
  const float *thP = fq(a.q[2]), *ppP = fq(a.q[3]), *th0 = fq(a.q[4]), *pi0 = fq(a.q[5]),
              *qv = fq(a.q[6]);
  float *prs = fq(a.q[7]), *rho = fq(a.q[8]);
  const int k = gz;
  int64_t c = g.idx(gx, gy, gz);
  const float cpdrd = CP / RD;
  float pi = pi0[k] + ppP[c];
  float th = th0[k] + thP[c];
  float p = P00 * fast_pow(pi, cpdrd);
  prs[c] = p;
  float rmix = moist != 0 ? fma_(mx(0.0f, qv[c]), RV, RD) : RD;
  rho[c] = p / ((th * pi) * rmix);
}

// stability_max.wgsl:120-178 (adapt_dt = 1): max over the grid of the
// squared advective Courant number c2 = x^2 + y^2 + z^2 (x = (u_i + u_i+1)
// dt / (2 dx) uh, and alike) and of the squared diffusion number of kmh,
// khh, kmv at levels k >= 1, into two words: the bits of non-negative floats
// order like the floats, so an unsigned max of the bits is the float max
// and the result does not depend on the order of the workgroups.  Each
// workgroup reduces in LDS and adds one global atomic per word.
__global__ __launch_bounds__(256) void k_stability_max(uint32_t* M, KArg a) {
  __shared__ uint32_t red[2];
  const uint32_t t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  if (t == 0) {
    red[0] = 0u;
    red[1] = 0u;
  }
  __syncthreads();
  int gx, gy, gz;
  te_tid3(a, gx, gy, gz);
  const Dims g = dims(a.q[0], 0);
  float c2 = 0.0f, m = 0.0f;
  if (gx < (int)a.n[0] && gy < (int)a.n[1] && gz < (int)a.n[2] && gx < g.nx && gy < g.ny && gz < g.nz) {
    const uint32_t* P = a.q[1];
    const float dtdx = ldf(P, 0), dtdy = ldf(P, 1), dtdz = ldf(P, 2), ksdx = ldf(P, 3), ksdy = ldf(P, 4);
    const int twoD = ldi(P, 5), withKs = ldi(P, 6);
    const float *u = fq(a.q[2]), *v = fq(a.q[3]), *w = fq(a.q[4]), *mh = fq(a.q[5]), *hm = fq(a.q[30]);
    const int64_t c = g.idx(gx, gy, gz);
    const float uh = hm_uh(hm, g, gx), vh = hm_vh(hm, g, gy);
    const float x = ((u[c] + u[c + 1]) * dtdx) * uh;
    const float z = ((w[c] + w[c + g.sk()]) * dtdz) * mh[gz];
    if (twoD != 0) {
      c2 = fma_(x, x, z * z);
    } else {
      const float y = ((v[c] + v[c + g.sJ]) * dtdy) * vh;
      c2 = fma_(z, z, fma_(x, x, y * y));
    }
    if (withKs != 0 && gz >= 1) {
      const float K[3] = {fq(a.q[6])[c], fq(a.q[8])[c], fq(a.q[7])[c]};  // kmh, khh, kmv
      for (int n = 0; n < 3; ++n) {
        const float aa = ((K[n] * ksdx) * uh) * uh;
        float tt = aa;
        if (twoD == 0) {
          const float bb = ((K[n] * ksdy) * vh) * vh;
          tt = fma_(aa, aa, bb * bb);
        }
        m = mx(m, tt);
      }
    }
  }
  atomicMax(&red[0], __float_as_uint(c2));
  atomicMax(&red[1], __float_as_uint(m));
  __syncthreads();
  if (t == 0) {
    atomicMax(a.q[9], red[0]);
    atomicMax(a.q[9] + 1, red[1]);
  }
}

}  // namespace te
