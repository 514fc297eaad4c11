// apmasscon dry-mass conservation at the end of the step: row, level and
// total sums in double-float (float pairs) in a fixed sequential order, so
// every launch reproduces the reference's bits; then the uniform pi'
// correction with prs and rho from the corrected state.
#include "te_device.hpp"
#include "te_kernels.hpp"

namespace te {

struct DFs {
  float x, y;
};
// f_df_add (mass_row_sums.wgsl:120-127)
TE_DEV void df_add(DFs& acc, float x) {
  float s = acc.x + x;
  float bb = s - acc.x;
  float err = (acc.x - (s - bb)) + (x - bb);
  acc.x = s;
  acc.y = acc.y + err;
}
// f_df_add_pair (mass_level_sums.wgsl:98-102)
TE_DEV void df_add_pair(DFs& acc, float vx, float vy) {
  df_add(acc, vx);
  acc.y = acc.y + vy;
}

// mass_row_sums.wgsl:134-170, one thread per (j, k) row.  Grid [ny, nz].
__global__ __launch_bounds__(256) void k_mass_row_sums(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  const int j = gx, k = gy;
  if (j >= g.ny || k >= g.nz) return;
  const float *rho = fq(a.q[1]), *pi0 = fq(a.q[2]), *pp3d = fq(a.q[3]), *mh = fq(a.q[4]),
              *hm = fq(a.q[30]);
  float *rows1 = fq(a.q[5]), *rows2 = fq(a.q[6]);
  DFs s1{0.0f, 0.0f}, s2{0.0f, 0.0f};
  const float rvh = 1.0f / hm_vh(hm, g, j);
  for (int i = 0; i < g.nx; ++i) {
    const int64_t c = g.idx(i, j, k);
    const float rmh = 1.0f / mh[k];
    const float w = ((rho[c] * (1.0f / hm_uh(hm, g, i))) * rvh) * rmh;
    df_add(s1, w);
    df_add(s2, pi0[k] + pp3d[c]);
  }
  const int64_t r = ((int64_t)k * g.sK + j) * 2;
  rows1[r] = s1.x; rows1[r + 1] = s1.y;
  rows2[r] = s2.x; rows2[r + 1] = s2.y;
}

// mass_level_sums.wgsl:118-146, one thread per level.  Grid [nz].
__global__ __launch_bounds__(256) void k_mass_level_sums(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  const int k = gx;
  if (k >= g.nz) return;
  const float *rows1 = fq(a.q[1]), *rows2 = fq(a.q[2]);
  float *lev1 = fq(a.q[3]), *lev2 = fq(a.q[4]);
  DFs s1{0.0f, 0.0f}, s2{0.0f, 0.0f};
  for (int j = 0; j < g.ny; ++j) {
    const int64_t r = ((int64_t)k * g.sK + j) * 2;
    df_add_pair(s1, rows1[r], rows1[r + 1]);
    df_add_pair(s2, rows2[r], rows2[r + 1]);
  }
  lev1[k * 2] = s1.x; lev1[k * 2 + 1] = s1.y;
  lev2[k * 2] = s2.x; lev2[k * 2 + 1] = s2.y;
}

// mass_reference_init.wgsl:118-142: the reference dry mass, once (ref[2]).
__global__ __launch_bounds__(256) void k_mass_reference_init(uint32_t* M, KArg a) {
  TE_TID3
  if (gx != 0) return;
  Dims g = dims(a.q[0], 0);
  const float* lev1 = fq(a.q[1]);
  float* ref = fq(a.q[2]);
  if (ref[2] != 0.0f) return;
  DFs s{0.0f, 0.0f};
  for (int k = 0; k < g.nz; ++k) df_add_pair(s, lev1[k * 2], lev1[k * 2 + 1]);
  ref[0] = s.x;
  ref[1] = s.y;
  ref[2] = 1.0f;
}

// f_powm1 (mass_correction_apply.wgsl:121-128).  The |r| > 0.05 branch calls
// the WGSL pow builtin; the correctly rounded pow stands in for it.
TE_DEV float powm1(float r, float a) {
  if (ab(r) > 0.05f) return fast_pow(1.0f + r, a) - 1.0f;
  return r * fma_(r, fma_(a * (a - 1.0f), 0.5f, (((r * a) * (a - 1.0f)) * (a - 2.0f)) / 6.0f), a);
}

// mass_correction_apply.wgsl:258-321.  Grid [nx, ny, nz].
__global__ __launch_bounds__(256) void k_mass_correction_apply(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  const int i = gx, j = gy, k = gz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const int moist = ldi(a.q[1], 0);
  const float *lev1 = fq(a.q[2]), *lev2 = fq(a.q[3]), *ref = fq(a.q[4]), *thP = fq(a.q[5]),
              *th0 = fq(a.q[7]), *pi0 = fq(a.q[8]), *qv = fq(a.q[9]);
  float *pp3d = fq(a.q[6]), *prs = fq(a.q[10]), *rho = fq(a.q[11]), *out = fq(a.q[12]);
  DFs m2{0.0f, 0.0f}, pp2{0.0f, 0.0f};
  for (int kk = 0; kk < g.nz; ++kk) {
    df_add_pair(m2, lev1[kk * 2], lev1[kk * 2 + 1]);
    df_add_pair(pp2, lev2[kk * 2], lev2[kk * 2 + 1]);
  }
  const float rddcv = RD / (CP - RD);
  const float r = ((ref[0] - m2.x) + (ref[1] - m2.y)) / m2.x;
  const float p2 = pp2.x + pp2.y;
  const float tem = (powm1(r, rddcv) * p2) / (float)((g.nx * g.ny) * g.nz);
  if (i == 0 && j == 0 && k == 0) {
    out[0] = tem; out[1] = m2.x; out[2] = p2; out[3] = r; out[4] = m2.y;
  }
  const int64_t c = g.idx(i, j, k);
  const float pp = pp3d[c] + tem;
  pp3d[c] = pp;
  const float cpdrd = CP / RD;
  const float pi = pi0[k] + pp;
  const float th = th0[k] + thP[c];
  const float p = P00 * fast_pow(pi, cpdrd);
  prs[c] = p;
  const float rmix = moist != 0 ? fma_(mx(0.0f, qv[c]), RV, RD) : RD;
  rho[c] = p / ((th * pi) * rmix);
}

}  // namespace te
