// Launches that replace a run of reference dispatches (encoder option
// `fuse`).  Each produces the same bits as the run it replaces on every
// field the step carries.
#include "te_device.hpp"
#include <type_traits>

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

// rk_update of a scalar (stagger 1) then its halo fill (rk_update.wgsl:111-132,
// timestep.js:1593 haloFill); arguments as rk_update's.
__global__ __launch_bounds__(256) void k_rk_update_xy(uint32_t* M, KArg a) {
  TE_TID3
  const Dims g = dims(a.q[0], 0);
  const int i = gx, j = gy, k = gz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;
  const float *phiA = fq(a.q[2]), *tend = fq(a.q[3]);
  const float dttmp = ldf(a.q[5], 0);
  update_with_images(g, fq(a.q[4]), i, j, k, 0, g.nz, [&](int64_t c) { return fma_(dttmp, tend[c], phiA[c]); });
}

// tke_advance (levels 1 .. nz-1) then the halo fill of tke
// (tke_advance.wgsl:111-129, timestep.js:1593 haloFill); arguments as
// tke_advance's.
__global__ __launch_bounds__(256) void k_tke_advance_xy(uint32_t* M, KArg a) {
  TE_TID3
  const Dims g = dims(a.q[0], 0);
  const int i = gx, j = gy, k = gz + 1;
  if (i >= g.nx || j >= g.ny || k > g.nz - 1) return;
  const float *tkea = fq(a.q[1]), *tketen = fq(a.q[2]);
  const float dttmp = ldf(a.q[3], 0);
  update_with_images(g, fq(a.q[4]), i, j, k, 1, g.nz, [&](int64_t c) {
    const float v = fma_(dttmp, tketen[c], tkea[c]);
    return (v < 1e-10f) ? 0.0f : v;
  });
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

namespace te {

// One acoustic substep after acoustic_uv, on a periodic domain, in one launch:
// acoustic_fpk, acoustic_w_solve, acoustic_pp and acoustic_accumulate, then
// the periodic halo fill of ppd, u, v and w, and on the stage's last substep
// the one of pp3d that follows the substeps (timestep.js:2023-2054,
// acoustic_fpk.wgsl:167-207, acoustic_w_solve.wgsl:124-212,
// acoustic_pp.wgsl:131-167, acoustic_accumulate.wgsl:129-158); on the first
// substep of a stage also acoustic_begin (acoustic_begin.wgsl:129-149).
// acoustic_prepare's qk, pk1, pk2 and mm (acoustic_prepare.wgsl:131-184) are
// formed per level from ppterm, piadv and thv with its operations, instead of
// being stored once per stage and read on every substep.
//
// A workgroup owns 64 columns of one row (lanes along i) and all levels:
// AC_WAVES waves each take a contiguous range of levels for the pointwise
// parts, and the tridiagonal sweep is passed from wave to wave (forward
// elimination upwards, back substitution downwards), each wave keeping its
// levels' coefficients in registers.  Values of the sweep are the reference's
// sequential Thomas recurrence, level by level.
//
// Halo cells (every halo cell of ppd, u, v, w is the image of interior cell
// (i mod nx, j mod ny), as the halo fill leaves it), each updated by the
// workgroup of its interior cell, at each level by the wave that owns the
// level, after that wave's own stores:
//  * w and ppd: rrw accumulates the image's old w (the accumulate runs
//    before the fill), then the new values are copied in.
//  * u and v: rru, rrv accumulate the value there (acoustic_uv's at i = nx or
//    j = ny, else the old image), then the image is copied in.  u at
//    (nx, j) and v at (i, ny) are updated by the workgroup of column nx-1 /
//    row ny-1, the only one that reads them (acoustic_fpk), after its reads.
// The row's own images (hx = i) go column by column, lane = column; the
// x-images and corners through a short per-workgroup list.
// p: 0 dims, 1 acoustic, 2 u, 3 v, 4 w, 5 piadv, 6 ppten, 7 ppterm, 8 thv,
//    9 wten, 10 pp3d, 11 ppd, 12 ppx, 13 rru, 14 rrv, 15 rrw, 16 -> save ppx,
//    17 -> tavg, 18 rr0, 19 rf0, 20 mh, 21 mf, 22 c1, 23 c2, 30 metric,
//    33 -> first substep of the stage, 34 -> (if bound) fill pp3d's halo
//    too (with ppd's, group 4 below); the fields the replaced dispatches
//    write as scratch (for expansions of this launch into them): 24 qk,
//    25 pk1, 26 pk2, 27 mm, 28 fpk, 29 cA, 31 cB, and 32 rho0.
// Grid: 64 x AC_WAVES threads per workgroup, ceil(nx / 64) x ny workgroups.
// Needs nx, ny >= 4 and 4 <= nz <= AC_WAVES * LMAX.
#define AC_HALO_CELLS 40  // per workgroup: at most 7 x-images, 1 (nx, j) cell, 14 corners
template <int LMAX>
__device__ __forceinline__ void ac_column(uint32_t* M, const KArg& a) {
  constexpr int KMAX = AC_WAVES * 16 + 1;
  __shared__ float xa[AC_WAVES][3][64];  // pk1+qk, pk2-qk, pp3d+fpk at each wave's last level
  __shared__ float xb[2][64];            // elimination carry between waves
  __shared__ float xw[AC_WAVES][64];     // new w at each wave's first level
  __shared__ float lv[6][KMAX];          // acoustic_prepare's per-level factors: tem, r1, r2, tem1, c1, c2
  __shared__ int hcnt;
  __shared__ int fwd_done, back_next;  // the sweep's hand-offs between waves
  __shared__ short hcell[AC_HALO_CELLS][3];
  // dynamic: per level and lane mm (then cA, then the new w) and the w
  // row's right-hand side part (then cB): 2 * nz * 64 floats
  float* me0 = te_lds;
  const Blk blk = te_block(a);
  const int lane = (int)threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one wave per threadIdx.y
  const Dims g = dims(a.q[0], 0);
  const int i = blk.x * 64 + lane, j = blk.y, nx = g.nx, ny = g.ny, nz = g.nz;
  float* me1 = te_lds + nz * 64;
  const bool active = i < nx && j < ny;
  const int Lw = (nz + AC_WAVES - 1) / AC_WAVES, k0 = wave * Lw, k1 = min(nz, k0 + Lw), Lq = max(0, k1 - k0);
  const bool top = k1 == nz && Lq > 0;
  const Acoustic ac = acoustic(a.q[1], 0);
  // fields as 32-bit word offsets into M (fewer scalar registers than pointers)
  auto F = [&](int q) { return reinterpret_cast<float*>(M) + a.p[q]; };
  const bool saveX = ldi(a.q[16], 0) != 0, first = ldi(a.q[33], 0) != 0;
  const bool imgP = a.p[34] != 0 && ldi(a.q[34], 0) != 0;
  const float tavg = ldf(a.q[17], 0);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    hcnt = 0;
    fwd_done = 0;
    back_next = AC_WAVES - 1;
  }
  // Halo cells this workgroup updates, (hx, hy, groups):
  // group 1 = u, rru; 2 = v, rrv; 4 = w, rrw, ppd.  A halo cell's groups are
  // handled by the workgroup of its image (i mod nx, j mod ny), except u at
  // (nx, j) and v at (i, ny): acoustic_fpk of column nx-1 / row ny-1 reads
  // them, so that column's / row's workgroup updates them after its reads.
  // The images of the workgroup's own row (hx = i: hy = j -+ ny, and v at
  // (i, ny) for row ny-1) are handled column by column (row_images); the list
  // holds the x-images, the corners and u at (nx, j).
  const int ys[2] = {j >= ny - 3 ? j - ny : j, j <= 3 ? j + ny : j};
  if (active && wave == 0) {
    const int xs[2] = {i >= nx - 3 ? i - nx : i, i <= 3 ? i + nx : i};
    const int nxi = (i >= nx - 3) + (i <= 3), nyi = (j >= ny - 3) + (j <= 3);
    const int n = nxi * (1 + nyi) + (i == nx - 1);
    if (n > 0) {
      int q = atomicAdd(&hcnt, n);
      auto put = [&](int hx, int hy, int m) {
        hcell[q][0] = (short)hx; hcell[q][1] = (short)hy; hcell[q][2] = (short)m;
        ++q;
      };
      for (int by = 0; by < 3; ++by)
        for (int bx = 1; bx < 3; ++bx) {
          const int hx = xs[bx - 1], hy = by == 0 ? j : ys[by - 1];
          if (hx == i || (by > 0 && hy == j)) continue;
          put(hx, hy, 7 & ~(hx == nx && hy == j ? 1 : 0));
        }
      if (i == nx - 1) put(nx, j, 1);
    }
  }

  {
    const float *rr0 = F(18), *rf0 = F(19), *mh = F(20), *mf = F(21), *c1 = F(22), *c2 = F(23);
    for (int k = (int)(threadIdx.y * 64 + threadIdx.x); k <= nz; k += 64 * AC_WAVES) {
      if (k < nz) {
        const float tem = (ac.dts * g.rdz) * mh[k];
        lv[0][k] = tem;
        lv[1][k] = (tem * rr0[k]) * rf0[k];
        lv[2][k] = (tem * rr0[k]) * rf0[k + 1];
      }
      lv[3][k] = ((ac.dts * g.rdz) * CP) * mf[k];
      lv[4][k] = c1[k];
      lv[5][k] = c2[k];
    }
  }
  __syncthreads();
  const int64_t sj = g.sJ, sk = g.sk();
  const float asq = ac.alph * ac.alph;
  const float temx = (ac.dts * 0.5f) * g.rdx, temy = (ac.dts * 0.5f) * g.rdy;
  const bool vec_v = nx > 3 && i < (nx / 4) * 4;
  auto acc = [&](float base, float x) { return tavg > 0.0f ? (base + x) * tavg : base + x; };
  auto accw = [&](float base, float x) { return tavg > 0.0f ? fma_(x, ac.alph, base) * tavg : base + x; };
  // Halo cells, each wave at its own levels (and level nz for the top wave)
  // once its stores above are done.  u, v (groups 1, 2): rru, rrv accumulate
  // the value there (acoustic_uv's at i = nx or j = ny, else the old image),
  // then the image is copied in; the interior values are acoustic_uv's, and
  // the wave's own acoustic_fpk reads of u at i = nx and v at j = ny came
  // first.  w, ppd (group 4): rrw accumulates the old w there, then the new w
  // and ppd of the interior cell are copied in; the interior cell is in this
  // workgroup's row at the same level, written by this wave.  pp3d (with
  // group 4 when imgP) likewise: no thread of the launch reads its halo.
  // halo(): the cells of the list, 2 (cell, level) items per lane per round
  // with their loads before their stores.
  // The row's own images, lane = column, 2 levels of loads before their
  // stores: r = 0, 1 the images at hy = j -+ ny (v excluded at hy = ny: row
  // ny-1's workgroup handles v there), r = 2 v at (i, ny) for row ny-1.
  auto row_images = [&](auto G) {
    constexpr int groups = decltype(G)::value;
    if (!active) return;
    float *u = F(2), *v = F(3), *w = F(4), *pp3d = F(10), *ppd = F(11), *rru = F(13), *rrv = F(14), *rrw = F(15);
    const int nl = Lq + (top ? 1 : 0);
    for (int r = 0; r < 3; ++r) {
      int hy, m;
      if (r < 2) {
        hy = ys[r];
        if (hy == j || (r == 1 && ys[0] == ys[1])) continue;
        m = 7 & ~(hy == ny ? 2 : 0);
      } else {
        if (j != ny - 1) continue;
        hy = ny;
        m = 2;
      }
      m &= groups;
      if (m == 0) continue;
      const int sy = hy < 0 ? hy + ny : (hy >= ny ? hy - ny : hy);
      for (int t0 = 0; t0 < nl; t0 += 2) {
        float x[2][11];
#pragma unroll
        for (int d = 0; d < 2; ++d)
          if (t0 + d < nl) {
            const int t = t0 + d, k = t < Lq ? k0 + t : nz;
            const int64_t ch = g.idx(i, hy, k), cs = g.idx(i, sy, k);
            if (m & 1) { x[d][0] = u[ch]; x[d][1] = u[cs]; x[d][2] = rru[ch]; }
            if (m & 2) { x[d][3] = v[ch]; x[d][4] = v[cs]; x[d][5] = rrv[ch]; }
            if (m & 4) {
              x[d][6] = w[ch]; x[d][7] = rrw[ch]; x[d][8] = w[cs]; x[d][9] = ppd[cs];
              if (imgP) x[d][10] = pp3d[cs];
            }
          }
#pragma unroll
        for (int d = 0; d < 2; ++d)
          if (t0 + d < nl) {
            const int t = t0 + d, k = t < Lq ? k0 + t : nz;
            const int64_t ch = g.idx(i, hy, k);
            if (m & 1) {
              rru[ch] = acc(first ? 0.0f : x[d][2], x[d][0]);
              u[ch] = x[d][1];
            }
            if (m & 2) {
              rrv[ch] = acc(first ? 0.0f : x[d][5], x[d][3]);
              v[ch] = x[d][4];
            }
            if (m & 4) {
              rrw[ch] = accw(first ? x[d][6] * ac.beta : x[d][7], x[d][6]);
              w[ch] = x[d][8];
              ppd[ch] = x[d][9];
              if (imgP) pp3d[ch] = x[d][10];
            }
          }
      }
    }
  };
  auto halo = [&](auto G) {
    constexpr int groups = decltype(G)::value;
    float *u = F(2), *v = F(3), *w = F(4), *pp3d = F(10), *ppd = F(11), *rru = F(13), *rrv = F(14), *rrw = F(15);
    const int nc = hcnt, nl = Lq + (top ? 1 : 0), items = nc * nl;
    constexpr int B = 2;
    for (int it0 = lane; it0 < items; it0 += B * 64) {
      float x[B][11];
      int64_t hc[B];
      int mask[B];
#pragma unroll
      for (int d = 0; d < B; ++d) {
        const int it = it0 + d * 64;
        mask[d] = 0;
        if (it < items) {
          const int kk = it / nc, q = it - kk * nc, k = kk < Lq ? k0 + kk : nz;
          const int hx = hcell[q][0], hy = hcell[q][1];
          const int sx = hx < 0 ? hx + nx : (hx >= nx ? hx - nx : hx), sy = hy < 0 ? hy + ny : (hy >= ny ? hy - ny : hy);
          const int64_t ch = g.idx(hx, hy, k), cs = g.idx(sx, sy, k);
          hc[d] = ch;
          mask[d] = hcell[q][2] & groups;
          if (mask[d] & 1) { x[d][0] = u[ch]; x[d][1] = u[cs]; x[d][2] = rru[ch]; }
          if (mask[d] & 2) { x[d][3] = v[ch]; x[d][4] = v[cs]; x[d][5] = rrv[ch]; }
          if (mask[d] & 4) {
            x[d][6] = w[ch]; x[d][7] = rrw[ch]; x[d][8] = w[cs]; x[d][9] = ppd[cs];
            if (imgP) x[d][10] = pp3d[cs];
          }
        }
      }
#pragma unroll
      for (int d = 0; d < B; ++d) {
        const int64_t ch = hc[d];
        if (mask[d] & 1) {
          rru[ch] = acc(first ? 0.0f : x[d][2], x[d][0]);
          u[ch] = x[d][1];
        }
        if (mask[d] & 2) {
          rrv[ch] = acc(first ? 0.0f : x[d][5], x[d][3]);
          v[ch] = x[d][4];
        }
        if (mask[d] & 4) {
          rrw[ch] = accw(first ? x[d][6] * ac.beta : x[d][7], x[d][6]);
          w[ch] = x[d][8];
          ppd[ch] = x[d][9];
          if (imgP) pp3d[ch] = x[d][10];
        }
      }
    }
  };
  float PQ[LMAX], P2Q[LMAX], PF[LMAX];
  // ---- pointwise: acoustic_prepare's terms, acoustic_fpk, the w row's
  // right-hand side (levels -1 and nz, read at the column's ends and not
  // used, lie in the field slot or the memory's front guard)
  if (active) {
    const float *u = F(2), *v = F(3), *w = F(4), *piadv = F(5), *ppten = F(6), *ppterm = F(7), *thv = F(8),
                *wten = F(9), *pp3d = F(10), *hm = F(30);
    const float uh = hm_uh(hm, g, i), vh = hm_vh(hm, g, j);
    float pp_prev = pp3d[g.idx(i, j, min(k0, nz) - 1)];
#pragma unroll
    for (int t = 0; t < LMAX; ++t) {
      if (t < Lq) {
        const int k = k0 + t;
        const int64_t c = g.idx(i, j, k);
        const float uc = u[c], up = u[c + 1], vc = v[c], vp = v[c + sj], wk = w[c], wk1 = w[c + sk];
        const float pa0 = piadv[c], paxm = piadv[c - 1], paxp = piadv[c + 1], paym = piadv[c - sj],
                    payp = piadv[c + sj], pzp0 = piadv[c + sk], pzm0 = piadv[c - sk];
        const float ppt = ppterm[c], ppn = ppten[c], pp = pp3d[c], th = thv[c], thm0 = thv[c - sk], wt = wten[c];
        const bool bot = k == 0, ktop = k == nz - 1;
        const float pzp = ktop ? 0.0f : pzp0, pzm = bot ? 0.0f : pzm0, thm = bot ? th : thm0;
        const float tem = lv[0][k], r1 = lv[1][k], r2 = lv[2][k], tem1 = lv[3][k];
        const float c1a = lv[4][k], c2a = lv[5][k], c1b = lv[4][k + 1], c2b = lv[5][k + 1];
        const float qk = tem * ppt;
        const float p2 = ktop ? 0.0f : (vec_v ? fma_(-c2b, pzp, (1.0f - c1b) * pa0) : fma_(1.0f - c1b, pa0, -(c2b * pzp)));
        const float p1 = bot ? 0.0f : (vec_v ? fma_(c2a - 1.0f, pa0, c1a * pzm) : fma_(c1a, pzm, (c2a - 1.0f) * pa0));
        const float pk2 = ktop ? 0.0f : r2 * p2;
        const float pk1 = bot ? 0.0f : r1 * p1;
        const float mm = tem1 * fma_(c2a, th, c1a * thm);
        float div = fma_((up - uc) * g.rdx, uh, ((vp - vc) * g.rdy) * vh);
        if (ab(div) < ac.smeps) div = 0.0f;
        const float Au = fma_(uc, pa0 - paxm, up * (paxp - pa0));
        const float Av = fma_(vc, pa0 - paym, vp * (payp - pa0));
        const float advPi = fma_((-Au) * temx, uh, -((Av * temy) * vh));
        const float fpk = fma_(ac.beta, fma_(wk, pk1 + qk, wk1 * (pk2 - qk)),
                               fma_(ac.dts, fma_(-ppt, div, ppn), advPi));
        PQ[t] = pk1 + qk;
        P2Q[t] = pk2 - qk;
        PF[t] = pp + fpk;
        me0[k * 64 + lane] = mm;
        me1[k * 64 + lane] = bot ? 0.0f : wk + fma_(ac.dts, wt, ac.beta * ((-mm) * (pp - pp_prev)));
        pp_prev = pp;
      }
    }
#pragma unroll
    for (int t = 0; t < LMAX; ++t)
      if (t == Lq - 1) {
        xa[wave][0][lane] = PQ[t];
        xa[wave][1][lane] = P2Q[t];
        xa[wave][2][lane] = PF[t];
      }
  }
  // The sweep passes from wave to wave through LDS counters instead of
  // workgroup barriers, so a wave's share of the serial recurrence overlaps
  // the loads of the waves above it (forward) and the stores of the waves
  // that finished (backward).  Every wave passes every hand-off, in wave
  // order: wave w waits for fwd_done == w (set by wave w-1 after its
  // elimination) and for back_next == w (set by wave w+1 after its back
  // substitution; the top wave is the last of the forward chain), so each
  // wait ends once the waves below / above it have run.  The waits are
  // bounded (2^22 polls) all the same.
  auto wait_for = [&](int* flag, int want) {
    for (int s = 0; s < (1 << 22); ++s) {
      const int v = __builtin_amdgcn_readfirstlane(
          __hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
      if (v == want) break;
      __builtin_amdgcn_s_sleep(1);
    }
  };
  auto pass = [&](int* flag, int v) { __hip_atomic_store(flag, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP); };
  // ---- forward elimination, wave after wave (cA over mm, cB over the
  // right-hand side part; the top row leaves w at nz-1 there).  Rows k = 1
  // and k = nz-1 are the reference's special cases, taken by selects.  Lanes
  // past nx compute on unset values and touch no global memory here.
  // The rows' operands from LDS into registers before the wait (the
  // recurrence then waits on nothing but itself).
  float ME[LMAX], EE[LMAX];
#pragma unroll
  for (int t = 0; t < LMAX; ++t)
    if (t < Lq) {
      ME[t] = me0[(k0 + t) * 64 + lane];
      EE[t] = me1[(k0 + t) * 64 + lane];
    }
  wait_for(&fwd_done, wave);
  {
    if (Lq > 0) {
      float pa = 0.0f, pb = 0.0f, pq_m = 0.0f, p2q_m = 0.0f, ppf_m = 0.0f;
      if (wave > 0) {
        pa = xb[0][lane];
        pb = xb[1][lane];
        pq_m = xa[wave - 1][0][lane];
        p2q_m = xa[wave - 1][1][lane];
        ppf_m = xa[wave - 1][2][lane];
      }
#pragma unroll
      for (int t = 0; t < LMAX; ++t) {
        const int k = k0 + t;
        if (t < Lq) {
          if (k >= 1) {
            const float m = ME[t], e = EE[t];
            const float aa = (-asq) * (m * pq_m);
            const float cc = (-asq) * ((-m) * P2Q[t]);
            const float bb = fma_(-asq, m * ((-PQ[t]) + p2q_m), 1.0f);
            const float dd = fma_(ac.alph, (-m) * (PF[t] - ppf_m), e);
            const bool first_row = k == 1;
            const float r = recip(first_row ? bb : fma_(aa, pa, bb));
            pa = (-cc) * r;
            pb = (first_row ? dd : fma_(-aa, pb, dd)) * r;
            ME[t] = pa;  // not read for the top row (w there is pb)
            EE[t] = pb;
          }
          pq_m = PQ[t];
          p2q_m = P2Q[t];
          ppf_m = PF[t];
        }
      }
      xb[0][lane] = pa;
      xb[1][lane] = pb;
    }
    pass(&fwd_done, wave + 1);
  }
  // ---- back substitution, wave after wave downwards (new w over cA); the
  // elimination coefficients of the wave's levels stayed in registers (ME, EE)
  wait_for(&back_next, wave);
  {
    if (Lq > 0) {
      float wn = top ? 0.0f : xw[wave + 1][lane];
#pragma unroll
      for (int t = LMAX - 1; t >= 0; --t) {
        const int k = k0 + t;
        if (t < Lq) {
          wn = k == nz - 1 ? EE[t] : (k >= 1 ? fma_(ME[t], wn, EE[t]) : 0.0f);
          me0[k * 64 + lane] = wn;
        }
      }
      xw[wave][lane] = wn;
    }
    pass(&back_next, wave - 1);
  }
  // ---- acoustic_pp, the rru / rrv / rrw accumulation (acoustic_begin's
  // values first on the stage's first substep), new w; the loads of 4 levels
  // before their stores
  if (active && Lq > 0) {
    float *u = F(2), *v = F(3), *w = F(4), *pp3d = F(10), *ppd = F(11), *ppx = F(12), *rru = F(13), *rrv = F(14),
          *rrw = F(15);
    const float wnext = top ? 0.0f : xw[wave + 1][lane];
#pragma unroll
    for (int t0 = 0; t0 < LMAX; t0 += 4) {
      float RU[4], RV[4], RW[4], UU[4], VV[4], PV[4];
#pragma unroll
      for (int d = 0; d < 4; ++d)
        if (t0 + d < Lq) {
          const int64_t c = g.idx(i, j, k0 + t0 + d);
          PV[d] = pp3d[c];
          UU[d] = u[c];
          VV[d] = v[c];
          RU[d] = first ? 0.0f : rru[c];
          RV[d] = first ? 0.0f : rrv[c];
          RW[d] = first ? w[c] * ac.beta : rrw[c];
        }
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const int t = t0 + d;
        if (t < Lq) {
          const int k = k0 + t;
          const int64_t c = g.idx(i, j, k);
          const float w0 = me0[k * 64 + lane];
          const float w1 = t + 1 < Lq ? me0[(k + 1) * 64 + lane] : wnext;
          float p = fma_(ac.alph, fma_(w0, PQ[t], w1 * P2Q[t]), PF[t]);
          if (ab(p) < ac.smeps) p = 0.0f;
          pp3d[c] = p;
          const float dum1 = ac.kdiv * (p - PV[d]);
          ppd[c] = p + dum1;
          if (saveX) ppx[c] = dum1;
          w[c] = w0;
          rru[c] = acc(RU[d], UU[d]);
          rrv[c] = acc(RV[d], VV[d]);
          rrw[c] = accw(RW[d], w0);
        }
      }
    }
    if (top) {  // level nz: w stays 0 (acoustic_w_solve clears it); ppd there is not updated
      const int64_t c = g.idx(i, j, nz);
      const float uo = u[c], vo = v[c], ru = first ? 0.0f : rru[c], rv = first ? 0.0f : rrv[c];
      const float rw = first ? w[c] * ac.beta : rrw[c];
      w[c] = 0.0f;
      rru[c] = acc(ru, uo);
      rrv[c] = acc(rv, vo);
      rrw[c] = accw(rw, 0.0f);
    }
  }
  // ---- halo cells of the wave's levels, after its own stores above
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  row_images(std::integral_constant<int, 7>());
  halo(std::integral_constant<int, 7>());
}
// instantiations by levels per wave: the runtime picks the one with
// ceil(nz / AC_WAVES) levels (k_ac_column: 16)
__global__ __launch_bounds__(64 * AC_WAVES) void k_ac_column(uint32_t* M, KArg a) { ac_column<16>(M, a); }
#define AC_INST(L) \
  __global__ __launch_bounds__(64 * AC_WAVES) void k_ac_column##L(uint32_t* M, KArg a) { ac_column<L>(M, a); }
AC_INST(1) AC_INST(2) AC_INST(3) AC_INST(4) AC_INST(5) AC_INST(6) AC_INST(7) AC_INST(8) AC_INST(9) AC_INST(10)
AC_INST(11) AC_INST(12) AC_INST(13) AC_INST(14) AC_INST(15)
#undef AC_INST

const void* ac_column_kernel(int levels_per_wave) {
  static const void* const fns[17] = {nullptr,
                                      (const void*)&k_ac_column1,  (const void*)&k_ac_column2,
                                      (const void*)&k_ac_column3,  (const void*)&k_ac_column4,
                                      (const void*)&k_ac_column5,  (const void*)&k_ac_column6,
                                      (const void*)&k_ac_column7,  (const void*)&k_ac_column8,
                                      (const void*)&k_ac_column9,  (const void*)&k_ac_column10,
                                      (const void*)&k_ac_column11, (const void*)&k_ac_column12,
                                      (const void*)&k_ac_column13, (const void*)&k_ac_column14,
                                      (const void*)&k_ac_column15, (const void*)&k_ac_column};
  return levels_per_wave >= 1 && levels_per_wave <= 16 ? fns[levels_per_wave] : nullptr;
}

}  // namespace te
