// Launches that replace a run of reference dispatches (encoder option
// `fuse`).  Each produces the same bits as the run it replaces on every
// field the step carries.
#include "te_device.hpp"
#include "te_kernels.hpp"

namespace te {

// Periodic halos of up to 8 fields in one launch.  Replaces, per field, the
// pair bc_periodic_x + bc_periodic_y (bc_periodic_x.wgsl:111-133,
// bc_periodic_y.wgsl:111-133; timestep.js haloFill): the x pass copies
// interior columns into the i-halo of every row, the y pass then copies whole
// rows (including the fresh i-halo) into the j-halo, so every halo cell ends
// as the image f[i mod nx, j mod ny, k] of an interior cell; this kernel
// writes exactly those images.
// Grid: x = halo cells of one level (7 rows of strideJ, then ny rows of 7),
// y = k in [0, nz].  p[1..8] fields, p[9] -> field count.
__global__ __launch_bounds__(256) void k_bc_periodic_xy(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  const int k = gy;
  if (k > g.nz) return;
  const int rowCells = 7 * g.sJ;
  int i, j;
  if (gx < rowCells) {
    const int r = gx / g.sJ;           // 0..6 -> j = -3..-1, ny..ny+3
    i = gx - r * g.sJ - 3;
    j = r < 3 ? r - 3 : g.ny + (r - 3);
  } else {
    const int t = gx - rowCells;       // ny rows x 7 halo columns, row by row
    j = t / 7;
    const int cidx = t - j * 7;
    if (j >= g.ny) return;
    i = cidx < 3 ? cidx - 3 : g.nx + (cidx - 3);
  }
  const int si = ((i % g.nx) + g.nx) % g.nx, sj = ((j % g.ny) + g.ny) % g.ny;
  const int64_t dst = g.idx(i, j, k), src = g.idx(si, sj, k);
  const int n = ldi(M, a.p[9]);
  for (int q = 0; q < 8; ++q) {
    if (q >= n) break;
    float* f = fp(M, a.p[1 + q]);
    f[dst] = f[src];
  }
}

// turb_zero (turb_zero.wgsl, timestep.js:1070-1072) of up to 8 fields in one
// launch over the whole grid [strideJ, strideK, nz+1].  p[1..8] fields,
// p[9] -> field count.
__global__ __launch_bounds__(256) void k_zero_fields(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  if (gx >= g.sJ || gy >= g.sK || gz > g.nz) return;
  const int64_t c = gx + (int64_t)g.sJ * (gy + (int64_t)g.sK * gz);
  const int n = ldi(M, a.p[9]);
  for (int q = 0; q < 8; ++q) {
    if (q >= n) break;
    fp(M, a.p[1 + q])[c] = 0.0f;
  }
}

// turb_add_tendency (dst = fma(scale, src, dst) over the whole grid) for up
// to 4 independent triples: p[1..4] sources, p[5..8] destinations,
// p[9..12] -> scales, p[13] -> count.
__global__ __launch_bounds__(256) void k_add_fields(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  const int64_t c = g.idx(i, j, k);
  const int n = ldi(M, a.p[13]);
  for (int q = 0; q < 4; ++q) {
    if (q >= n) break;
    float* dst = fp(M, a.p[5 + q]);
    dst[c] = fma_(ldf(M, a.p[9 + q]), fp(M, a.p[1 + q])[c], dst[c]);
  }
}

// Up to 4 buffer copies of a.n[0] words each (the state swap of
// timestep.js:766-774): p[1..4] sources, p[5..8] destinations, p[9] -> count.
__global__ __launch_bounds__(256) void k_copy_fields(uint32_t* M, KArg a) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n[0]) return;
  const int n = ldi(M, a.p[9]);
  for (int f = 0; f < 4; ++f) {
    if (f >= n) break;
    M[a.p[5 + f] + q] = M[a.p[1 + f] + q];
  }
}

// ---- acoustic substep
// acoustic_fpk (acoustic_fpk.wgsl:167-207) and acoustic_w_solve
// (acoustic_w_solve.wgsl:124-212) of one column in one sweep: fpk(k) is formed
// on the way up, right before the elimination step that needs it, and
// written out for acoustic_pp.  w at k = 0 and nz is read before the solve
// clears it, as the separate fpk dispatch would have.  Loads run AC_CHUNK
// levels ahead of the serial elimination.
// p: 0 dims, 1 acoustic, 2 u, 3 v, 4 w, 5 piadv, 6 ppten, 7 ppterm, 8 pk1,
//    9 pk2, 10 qk, 11 fpk, 12 mm, 13 wten, 14 pp3d, 15 cA, 16 cB, 30 metric.
#define AC_CHUNK 4
__global__ __launch_bounds__(256) void k_ac_fpk_solve(uint32_t* M, KArg a) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i >= (int)a.n[0] || j >= (int)a.n[1]) return;
  Dims g = dims(M, a.p[0]);
  if (i >= g.nx || j >= g.ny) return;
  Acoustic ac = acoustic(M, a.p[1]);
  const float *u = fp(M, a.p[2]), *v = fp(M, a.p[3]), *piadv = fp(M, a.p[5]), *ppten = fp(M, a.p[6]),
              *ppterm = fp(M, a.p[7]), *pk1 = fp(M, a.p[8]), *pk2 = fp(M, a.p[9]), *qk = fp(M, a.p[10]),
              *mm = fp(M, a.p[12]), *wten = fp(M, a.p[13]), *pp3d = fp(M, a.p[14]), *hm = fp(M, a.p[30]);
  float *w = fp(M, a.p[4]), *fpk = fp(M, a.p[11]), *cA = fp(M, a.p[15]), *cB = fp(M, a.p[16]);
  const int nz = g.nz;
  const int64_t sj = g.sJ, sk = g.sk(), base = g.idx(i, j, 0);
  const float temx = (ac.dts * 0.5f) * g.rdx, temy = (ac.dts * 0.5f) * g.rdy;
  const float huh = hm_uh(hm, g, i), hvh = hm_vh(hm, g, j);
  const float asq = ac.alph * ac.alph;
  // fpk of one level from its loads (acoustic_fpk.wgsl:185-205)
  auto fpk_of = [&](float u0, float u1, float v0, float v1, float pc, float pw, float pe, float ps, float pn, float wk,
                    float wk1, float a1, float a2, float q, float pterm, float pten) {
    float div = fma_((u1 - u0) * g.rdx, huh, ((v1 - v0) * g.rdy) * hvh);
    if (ab(div) < ac.smeps) div = 0.0f;
    float Au = fma_(u0, pc - pw, u1 * (pe - pc));
    float Av = fma_(v0, pc - ps, v1 * (pn - pc));
    float advPi = fma_((-Au) * temx, huh, -((Av * temy) * hvh));
    return fma_(ac.beta, fma_(wk, a1 + q, wk1 * (a2 - q)), fma_(ac.dts, fma_(-pterm, div, pten), advPi));
  };
  const float wTop = w[base + nz * sk];
  float pq_m, p2q_m, pp_m, ppf_m, wk = w[base];
  {  // level 0: fpk only
    const int64_t c = base;
    const float a1 = pk1[c], a2 = pk2[c], q = qk[c], wk1 = nz > 1 ? w[c + sk] : wTop;
    const float f = fpk_of(u[c], u[c + 1], v[c], v[c + sj], piadv[c], piadv[c - 1], piadv[c + 1], piadv[c - sj],
                           piadv[c + sj], wk, wk1, a1, a2, q, ppterm[c], ppten[c]);
    fpk[c] = f;
    pq_m = a1 + q; p2q_m = a2 - q; pp_m = pp3d[c]; ppf_m = pp_m + f;
    wk = wk1;
  }
  w[base] = 0.0f;
  w[base + nz * sk] = 0.0f;
  float pa = 0.0f, pb = 0.0f;
  for (int k0 = 1; k0 <= nz - 1; k0 += AC_CHUNK) {
    float u0[AC_CHUNK], u1[AC_CHUNK], v0[AC_CHUNK], v1[AC_CHUNK], pc[AC_CHUNK], pw[AC_CHUNK], pe[AC_CHUNK],
        ps[AC_CHUNK], pn[AC_CHUNK], wn1[AC_CHUNK], a1[AC_CHUNK], a2[AC_CHUNK], q[AC_CHUNK], pt[AC_CHUNK],
        pten[AC_CHUNK], m[AC_CHUNK], wt[AC_CHUNK], p[AC_CHUNK];
#pragma unroll
    for (int t = 0; t < AC_CHUNK; ++t) {
      const int k = k0 + t;
      if (k <= nz - 1) {
        const int64_t c = base + k * sk;
        u0[t] = u[c]; u1[t] = u[c + 1]; v0[t] = v[c]; v1[t] = v[c + sj];
        pc[t] = piadv[c]; pw[t] = piadv[c - 1]; pe[t] = piadv[c + 1]; ps[t] = piadv[c - sj]; pn[t] = piadv[c + sj];
        wn1[t] = k + 1 < nz ? w[c + sk] : wTop;
        a1[t] = pk1[c]; a2[t] = pk2[c]; q[t] = qk[c]; pt[t] = ppterm[c]; pten[t] = ppten[c];
        m[t] = mm[c]; wt[t] = wten[c]; p[t] = pp3d[c];
      }
    }
#pragma unroll
    for (int t = 0; t < AC_CHUNK; ++t) {
      const int k = k0 + t;
      if (k <= nz - 1) {
        const int64_t c = base + k * sk;
        const float f = fpk_of(u0[t], u1[t], v0[t], v1[t], pc[t], pw[t], pe[t], ps[t], pn[t], wk, wn1[t], a1[t], a2[t],
                               q[t], pt[t], pten[t]);
        fpk[c] = f;
        const float pq = a1[t] + q[t], p2q = a2[t] - q[t], ppf = p[t] + f;
        const float aa = (-asq) * (m[t] * pq_m);
        const float cc = (-asq) * ((-m[t]) * p2q);
        const float bb = fma_(-asq, m[t] * ((-pq) + p2q_m), 1.0f);
        const float fw = fma_(ac.dts, wt[t], ac.beta * ((-m[t]) * (p[t] - pp_m)));
        const float dd = fma_(ac.alph, (-m[t]) * (ppf - ppf_m), wk + fw);
        if (k == 1) {
          const float r = 1.0f / bb;
          pa = (-cc) * r;
          pb = dd * r;
          cA[c] = pa;
          cB[c] = pb;
        } else if (k < nz - 1) {
          const float r = 1.0f / fma_(aa, pa, bb);
          pa = (-cc) * r;
          pb = fma_(-aa, pb, dd) * r;
          cA[c] = pa;
          cB[c] = pb;
        } else {
          const float r = 1.0f / fma_(aa, pa, bb);
          pb = fma_(-aa, pb, dd) * r;
          w[c] = pb;
        }
        pq_m = pq; p2q_m = p2q; pp_m = p[t]; ppf_m = ppf;
        wk = wn1[t];
      }
    }
  }
  float wn = pb;
  for (int k0 = nz - 2; k0 >= 1; k0 -= 8) {
    float ca[8], cb[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = k0 - t;
      if (k >= 1) { ca[t] = cA[base + k * sk]; cb[t] = cB[base + k * sk]; }
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = k0 - t;
      if (k >= 1) {
        wn = fma_(ca[t], wn, cb[t]);
        w[base + k * sk] = wn;
      }
    }
  }
}

// acoustic_pp (acoustic_pp.wgsl:131-167) at interior points, then
// acoustic_accumulate (acoustic_accumulate.wgsl:129-158) at every point of
// the whole grid.  Neither reads what the other writes, so one pass does
// both.  p: 0 dims, 1 acoustic, 2 w, 3 pk1, 4 pk2, 5 qk, 6 fpk, 7 pp3d, 8 ppd,
// 9 ppx, 10 -> saveX, 11 u, 12 v, 13 rru, 14 rrv, 15 rrw, 16 -> tavg.
__global__ __launch_bounds__(256) void k_ac_pp_accumulate(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(M, a.p[0]);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  Acoustic ac = acoustic(M, a.p[1]);
  const float* w = fp(M, a.p[2]);
  const int64_t sk = g.sk(), c = g.idx(i, j, k);
  const float wc = w[c];
  if (i >= 0 && i < g.nx && j >= 0 && j < g.ny && k < g.nz) {
    const float *pk1 = fp(M, a.p[3]), *pk2 = fp(M, a.p[4]), *qk = fp(M, a.p[5]), *fpk = fp(M, a.p[6]);
    float *pp3d = fp(M, a.p[7]), *ppd = fp(M, a.p[8]), *ppx = fp(M, a.p[9]);
    const float prev = pp3d[c];
    float p = fma_(ac.alph, fma_(wc, pk1[c] + qk[c], w[c + sk] * (pk2[c] - qk[c])), prev + fpk[c]);
    if (ab(p) < ac.smeps) p = 0.0f;
    pp3d[c] = p;
    const float dum1 = ac.kdiv * (p - prev);
    ppd[c] = p + dum1;
    if (ldi(M, a.p[10]) != 0) ppx[c] = dum1;
  }
  const float *u = fp(M, a.p[11]), *v = fp(M, a.p[12]);
  float *rru = fp(M, a.p[13]), *rrv = fp(M, a.p[14]), *rrw = fp(M, a.p[15]);
  const float tavg = ldf(M, a.p[16]);
  if (tavg > 0.0f) {
    rru[c] = (rru[c] + u[c]) * tavg;
    rrv[c] = (rrv[c] + v[c]) * tavg;
    rrw[c] = fma_(wc, ac.alph, rrw[c]) * tavg;
  } else {
    rru[c] = rru[c] + u[c];
    rrv[c] = rrv[c] + v[c];
    rrw[c] = rrw[c] + wc;
  }
}

}  // namespace te
