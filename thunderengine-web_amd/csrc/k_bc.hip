// Lateral boundary kernels: periodic and open halo fills, open-boundary
// radiative speeds and normal-velocity tendencies, turbulence edge copies.
#include "te_device.hpp"
#include "te_kernels.hpp"

namespace te {

// bc_periodic_x.wgsl:111-133 / bc_periodic_y.wgsl:111-133
template <int DIR>
__device__ void periodic(uint32_t* M, const KArg& a, int h, int t, int k) {
  Dims g = dims(a.q[0], 0);
  const int n = DIR == 0 ? g.nx : g.ny, other = DIR == 0 ? g.ny : g.nx;
  if (h > 3 || t > other + 3 || k > g.nz) return;
  float* f = fq(a.q[1]);
  int lo = (((-1 - h) % n) + n) % n;
  int hi = (((n + h) % n) + n) % n;
  auto at = [&](int q) { return DIR == 0 ? g.idx(q, t, k) : g.idx(t, q, k); };
  if (h < 3) f[at(-1 - h)] = f[at(lo)];
  f[at(n + h)] = f[at(hi)];
}
__global__ __launch_bounds__(256) void k_bc_periodic_x(uint32_t* M, KArg a) { TE_TID3 periodic<0>(M, a, gx, gy - 3, gz); }
__global__ __launch_bounds__(256) void k_bc_periodic_y(uint32_t* M, KArg a) { TE_TID3 periodic<1>(M, a, gx, gy - 3, gz); }

// bc_open_x.wgsl:111-131 / bc_open_y.wgsl:111-131 (zero-gradient halo)
template <int DIR>
__device__ void open_halo(uint32_t* M, const KArg& a, int h, int t, int k) {
  Dims g = dims(a.q[0], 0);
  const int n = DIR == 0 ? g.nx : g.ny, other = DIR == 0 ? g.ny : g.nx;
  if (h >= 3 || t > other + 3 || k > g.nz) return;
  float* f = fq(a.q[1]);
  int last = n + ldi(a.q[2], 0) - 1;
  auto at = [&](int q) { return DIR == 0 ? g.idx(q, t, k) : g.idx(t, q, k); };
  f[at(-1 - h)] = f[at(0)];
  f[at(last + 1 + h)] = f[at(last)];
}
__global__ __launch_bounds__(256) void k_bc_open_x(uint32_t* M, KArg a) { TE_TID3 open_halo<0>(M, a, gx, gy - 3, gz); }
__global__ __launch_bounds__(256) void k_bc_open_y(uint32_t* M, KArg a) { TE_TID3 open_halo<1>(M, a, gx, gy - 3, gz); }

// openbc_uv_tendency.wgsl:167-208
__global__ __launch_bounds__(256) void k_openbc_uv_tendency(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  const int n = gx, k = gy;
  if (k >= g.nz) return;
  Acoustic ac = acoustic(a.q[1], 0);
  const float *rw = fq(a.q[2]), *re = fq(a.q[3]), *rs = fq(a.q[4]), *rn = fq(a.q[5]), *uten = fq(a.q[6]),
              *vten = fq(a.q[7]), *hm = fq(a.q[30]);
  float *u = fq(a.q[8]), *v = fq(a.q[9]);
  int openx = ldi(a.q[10], 0), openy = ldi(a.q[10], 1);
  const int64_t sj = g.sJ;
  if (openx != 0 && n < g.ny) {
    int64_t cw = g.idx(0, n, k);
    u[cw] = fma_(ac.dts, fma_(((-rw[n]) * (u[cw + 1] - u[cw])) * g.rdx, hm_uh(hm, g, 0), uten[cw]), u[cw]);
    int64_t ce = g.idx(g.nx, n, k);
    u[ce] = fma_(ac.dts, fma_(((-re[n]) * (u[ce] - u[ce - 1])) * g.rdx, hm_uh(hm, g, g.nx - 1), uten[ce]), u[ce]);
  }
  if (openy != 0 && n < g.nx) {
    int64_t cs = g.idx(n, 0, k);
    v[cs] = fma_(ac.dts, fma_(((-rs[n]) * (v[cs + sj] - v[cs])) * g.rdy, hm_vh(hm, g, 0), vten[cs]), v[cs]);
    int64_t cn = g.idx(n, g.ny, k);
    v[cn] = fma_(ac.dts, fma_(((-rn[n]) * (v[cn] - v[cn - sj])) * g.rdy, hm_vh(hm, g, g.ny - 1), vten[cn]), v[cn]);
  }
}

// radbc_ew.wgsl:118-161 / radbc_ns.wgsl:118-161
template <int DIR>
__device__ void radbc(uint32_t* M, const KArg& a, int t) {
  Dims g = dims(a.q[0], 0);
  if (t >= (DIR == 0 ? g.ny : g.nx)) return;
  const float *ua = fq(a.q[1]), *u = fq(a.q[2]), *hm = fq(a.q[30]);
  float dt = ldf(a.q[3], 0);
  float *lo = fq(a.q[4]), *hi = fq(a.q[5]);
  const float umax = 3.5e+02f;
  const int nq = DIR == 0 ? g.nx : g.ny;
  float d1 = DIR == 0 ? (1.0f / g.rdx) / hm_uf(hm, g, 1) : (1.0f / g.rdy) / hm_vf(hm, g, 1);
  float d2 = DIR == 0 ? (1.0f / g.rdx) / hm_uf(hm, g, g.nx - 1) : (1.0f / g.rdy) / hm_vf(hm, g, g.ny - 1);
  auto at = [&](int q, int k) { return DIR == 0 ? g.idx(q, t, k) : g.idx(t, q, k); };
  float avg1 = 0.0f, avg2 = 0.0f;
  for (int k = 0; k < g.nz; ++k) {
    int64_t c1 = at(1, k), c2 = at(2, k);
    float foo = ua[c2] - ua[c1];
    float cb1 = (d1 * (ua[c1] - u[c1])) / (dt * (copysign_mag(1e-10f, foo) + foo));
    avg1 = avg1 + mx(mn(cb1, 0.0f), -umax);
    int64_t ce = at(nq - 1, k), ce2 = at(nq - 2, k);
    foo = ua[ce] - ua[ce2];
    float cb2 = (d2 * (ua[ce] - u[ce])) / (dt * (copysign_mag(1e-10f, foo) + foo));
    avg2 = avg2 + mn(mx(cb2, 0.0f), umax);
  }
  lo[t] = avg1 / (float)g.nz;
  hi[t] = avg2 / (float)g.nz;
}
__global__ __launch_bounds__(256) void k_radbc_ew(uint32_t* M, KArg a) { TE_TID3 radbc<0>(M, a, gx); }
__global__ __launch_bounds__(256) void k_radbc_ns(uint32_t* M, KArg a) { TE_TID3 radbc<1>(M, a, gx); }

// turb_openbc_flux.wgsl:111-140
__global__ __launch_bounds__(256) void k_turb_openbc_flux(uint32_t* M, KArg a) {
  TE_TID3
  Dims g = dims(a.q[0], 0);
  const int n = gx, k = gy;
  if (k > g.nz) return;
  float *fx = fq(a.q[1]), *fy = fq(a.q[2]);
  int openx = ldi(a.q[3], 0), openy = ldi(a.q[3], 1);
  if (openx != 0 && n <= g.ny) {
    fx[g.idx(0, n, k)] = fx[g.idx(1, n, k)];
    fx[g.idx(g.nx, n, k)] = fx[g.idx(g.nx - 1, n, k)];
  }
  if (openy != 0 && n <= g.nx) {
    fy[g.idx(n, 0, k)] = fy[g.idx(n, 1, k)];
    fy[g.idx(n, g.ny, k)] = fy[g.idx(n, g.ny - 1, k)];
  }
}

// turb_openbc_stress_x.wgsl:111-137 / turb_openbc_stress_y.wgsl
template <int DIR>
__device__ void openbc_stress(uint32_t* M, const KArg& a, int t, int k) {
  Dims g = dims(a.q[0], 0);
  if (t > (DIR == 0 ? g.ny : g.nx) || k >= g.nz) return;
  float *sd = fq(a.q[1]), *s12 = fq(a.q[2]);
  const int nq = DIR == 0 ? g.nx : g.ny;
  auto at = [&](int q) { return DIR == 0 ? g.idx(q, t, k) : g.idx(t, q, k); };
  sd[at(-1)] = sd[at(0)];
  sd[at(nq)] = sd[at(nq - 1)];
  if (t >= 0) {
    s12[at(0)] = s12[at(1)];
    s12[at(nq)] = s12[at(nq - 1)];
  }
}
__global__ __launch_bounds__(256) void k_turb_openbc_stress_x(uint32_t* M, KArg a) { TE_TID3 openbc_stress<0>(M, a, gx - 1, gy); }
__global__ __launch_bounds__(256) void k_turb_openbc_stress_y(uint32_t* M, KArg a) { TE_TID3 openbc_stress<1>(M, a, gx - 1, gy); }

// copy_words (replay.js:403-406): end-of-step state swap copies.
__global__ __launch_bounds__(256) void k__copy(uint32_t* M, KArg a) {
  if (te_flat(a, [&](uint32_t w, int width) { st4(a.q[1], w, width, ld4(a.q[0], w, width)); })) return;
  const uint32_t q = (uint32_t)(te_block(a).x * TE_BX) + threadIdx.x;
  if (q >= a.n[0]) return;
  a.q[1][q] = a.q[0][q];
}

}  // namespace te
