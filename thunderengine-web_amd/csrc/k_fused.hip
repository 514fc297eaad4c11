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
  Dims g = dims(a.q[0], 0);
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
  const int n = ldi(a.q[9], 0);
  for (int q = 0; q < 8; ++q) {
    if (q >= n) break;
    float* f = fq(a.q[1 + q]);
    f[dst] = f[src];
  }
}

// turb_zero (turb_zero.wgsl, timestep.js:1070-1072) of up to 8 fields in one
// launch over the whole grid [strideJ, strideK, nz+1].  p[1..8] fields,
// p[9] -> field count.
__global__ __launch_bounds__(256) void k_zero_fields(uint32_t* M, KArg a) {
  if (te_flat(a, [&](uint32_t w, int width) {
        const int n = ldi(a.q[9], 0);
        F4 z;
        z.v[0] = z.v[1] = z.v[2] = z.v[3] = 0.0f;
        for (int q = 0; q < 8; ++q) {
          if (q >= n) break;
          st4(a.q[1 + q], w, width, z);
        }
      }))
    return;
  TE_TID3
  Dims g = dims(a.q[0], 0);
  if (gx >= g.sJ || gy >= g.sK || gz > g.nz) return;
  const int64_t c = gx + (int64_t)g.sJ * (gy + (int64_t)g.sK * gz);
  const int n = ldi(a.q[9], 0);
  for (int q = 0; q < 8; ++q) {
    if (q >= n) break;
    fq(a.q[1 + q])[c] = 0.0f;
  }
}

// turb_add_tendency (dst = fma(scale, src, dst) over the whole grid) for up
// to 4 independent triples: p[1..4] sources, p[5..8] destinations,
// p[9..12] -> scales, p[13] -> count.
__global__ __launch_bounds__(256) void k_add_fields(uint32_t* M, KArg a) {
  if (te_flat(a, [&](uint32_t w, int width) {
        const int n = ldi(a.q[13], 0);
        for (int q = 0; q < 4; ++q) {
          if (q >= n) break;
          const float sc = ldf(a.q[9 + q], 0);
          const F4 s = ld4(a.q[1 + q], w, width);
          F4 d = ld4(a.q[5 + q], w, width);
          for (int e = 0; e < width; ++e) d.v[e] = fma_(sc, s.v[e], d.v[e]);
          st4(a.q[5 + q], w, width, d);
        }
      }))
    return;
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  const int64_t c = g.idx(i, j, k);
  const int n = ldi(a.q[13], 0);
  for (int q = 0; q < 4; ++q) {
    if (q >= n) break;
    float* dst = fq(a.q[5 + q]);
    dst[c] = fma_(ldf(a.q[9 + q], 0), fq(a.q[1 + q])[c], dst[c]);
  }
}

// stage_tendency_reset (stage_tendency_reset.wgsl:111-129) followed by
// turb_add_tendency of the three SGS / damping momentum tendencies:
// dst = fma(scale, src, 0), the bits of adding src*scale to the cleared field.
// p: 1-3 uten vten wten, 4-6 their sources, 7 -> scale.  Grid: whole.
__global__ __launch_bounds__(256) void k_tendency_reset_add(uint32_t* M, KArg a) {
  if (te_flat(a, [&](uint32_t w, int width) {
        const float sc = ldf(a.q[7], 0);
        for (int q = 0; q < 3; ++q) {
          F4 d = ld4(a.q[4 + q], w, width);
          for (int e = 0; e < width; ++e) d.v[e] = fma_(sc, d.v[e], 0.0f);
          st4(a.q[1 + q], w, width, d);
        }
      }))
    return;
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  const int64_t c = g.idx(i, j, k);
  const float scale = ldf(a.q[7], 0);
  for (int q = 0; q < 3; ++q) fq(a.q[1 + q])[c] = fma_(scale, fq(a.q[4 + q])[c], 0.0f);
}

// Up to 4 buffer copies of a.n[0] words each (the state swap of
// timestep.js:766-774): p[1..4] sources, p[5..8] destinations, p[9] -> count.
__global__ __launch_bounds__(256) void k_copy_fields(uint32_t* M, KArg a) {
  if (te_flat(a, [&](uint32_t w, int width) {
        const int n = ldi(a.q[9], 0);
        for (int f = 0; f < 4; ++f) {
          if (f >= n) break;
          st4(a.q[5 + f], w, width, ld4(a.q[1 + f], w, width));
        }
      }))
    return;
  const uint32_t q = (uint32_t)(te_block(a).x * TE_BX) + threadIdx.x;
  if (q >= a.n[0]) return;
  const int n = ldi(a.q[9], 0);
  for (int f = 0; f < 4; ++f) {
    if (f >= n) break;
    a.q[5 + f][q] = a.q[1 + f][q];
  }
}

// ---- acoustic substep
// acoustic_pp (acoustic_pp.wgsl:131-167) at interior points, then
// acoustic_accumulate (acoustic_accumulate.wgsl:129-158) at every point of
// the whole grid.  Neither reads what the other writes, so one pass does
// both.  p: 0 dims, 1 acoustic, 2 w, 3 pk1, 4 pk2, 5 qk, 6 fpk, 7 pp3d, 8 ppd,
// 9 ppx, 10 -> saveX, 11 u, 12 v, 13 rru, 14 rrv, 15 rrw, 16 -> tavg.
__global__ __launch_bounds__(256) void k_ac_pp_accumulate(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  int i = gx - 3, j = gy - 3, k = gz;
  if (i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
  Acoustic ac = acoustic(a.q[1], 0);
  const float* w = fq(a.q[2]);
  const int64_t sk = g.sk(), c = g.idx(i, j, k);
  const float wc = w[c];
  if (i >= 0 && i < g.nx && j >= 0 && j < g.ny && k < g.nz) {
    const float *pk1 = fq(a.q[3]), *pk2 = fq(a.q[4]), *qk = fq(a.q[5]), *fpk = fq(a.q[6]);
    float *pp3d = fq(a.q[7]), *ppd = fq(a.q[8]), *ppx = fq(a.q[9]);
    const float prev = pp3d[c];
    float p = fma_(ac.alph, fma_(wc, pk1[c] + qk[c], w[c + sk] * (pk2[c] - qk[c])), prev + fpk[c]);
    if (ab(p) < ac.smeps) p = 0.0f;
    pp3d[c] = p;
    const float dum1 = ac.kdiv * (p - prev);
    ppd[c] = p + dum1;
    if (ldi(a.q[10], 0) != 0) ppx[c] = dum1;
  }
  const float *u = fq(a.q[11]), *v = fq(a.q[12]);
  float *rru = fq(a.q[13]), *rrv = fq(a.q[14]), *rrw = fq(a.q[15]);
  const float tavg = ldf(a.q[16], 0);
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
