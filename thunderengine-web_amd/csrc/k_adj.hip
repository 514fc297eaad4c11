// Adjoint kernels of the large step for gfx950.
//
// Each reference kernel on the path has an adjoint launch over the same
// thread grid: k_adj<B> runs adj_point<B> (adj_core.hpp) for every point of
// the grid, adding the transposed local Jacobian times the output adjoints to
// the adjoint memory with float atomics; acoustic_w_solve has a hand-written
// column adjoint (one thread per column).  The registry lists, for every
// kernel, the parameters it writes: the reverse sweep saves those regions
// before the forward dispatch and restores them before the adjoint one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "adj_bodies.hpp"
#include "te_adjoint.hpp"

namespace te {
namespace adj {

template <class B>
__global__ __launch_bounds__(256) void k_adj(const float* M, float* Mb, AdjArg a) {
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  const uint64_t n0 = a.n[0], n01 = n0 * a.n[1];
  if (t >= n01 * a.n[2]) return;
  const uint64_t z = t / n01, rem = t - z * n01, y = rem / n0, x = rem - y * n0;
  PointAdj<B>::run(M, Mb, a, (int)x, (int)y, (int)z);
}

// acoustic_w_solve: the rows' coefficients, column sweeps, the levels'
// coefficient adjoints, then the scratch words cleared
__global__ __launch_bounds__(256) void k_adj_w_coef(const float* M, AdjArg a, float* scr, uint64_t fw, uint32_t nz) {
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  const uint64_t n0 = a.n[0], n01 = n0 * a.n[1];
  if (t >= n01 * nz) return;
  const uint64_t z = t / n01, rem = t - z * n01, y = rem / n0, x = rem - y * n0;
  w_solve_coef(M, a, scr, fw, (int)x, (int)y, (int)z);
}
__global__ __launch_bounds__(64) void k_adj_w_sweep(const float* M, float* Mb, AdjArg a, float* scr, uint64_t fw,
                                                     uint32_t nz1) {
  extern __shared__ float te_col[];  // cA and cB / z: 2 (nz + 1) floats per lane, lane-interleaved
  const uint64_t t = (uint64_t)blockIdx.x * 64u + threadIdx.x;
  const uint64_t n0 = a.n[0];
  if (t >= n0 * a.n[1]) return;
  const uint64_t y = t / n0, x = t - y * n0;
  w_solve_sweep(M, Mb, a, scr, fw, te_col + threadIdx.x, te_col + 64 * nz1 + threadIdx.x, 64, (int)x, (int)y);
}
// The same with the column's coefficients and incoming adjoint preloaded into
// LDS by independent loads (32 columns per block: 7 arrays of nz + 1 floats
// per lane), so the serial phases wait on LDS, not on memory.
__global__ __launch_bounds__(32) void k_adj_w_sweep_pre(const float* M, float* Mb, AdjArg a, float* scr, uint64_t fw,
                                                        uint32_t nz1) {
  extern __shared__ float te_col[];
  const uint64_t t = (uint64_t)blockIdx.x * 32u + threadIdx.x;
  const uint64_t n0 = a.n[0];
  if (t >= n0 * a.n[1]) return;
  const uint64_t y = t / n0, x = t - y * n0;
  float* tA = te_col + threadIdx.x;
  float* tB = te_col + 32 * nz1 + threadIdx.x;
  float* pre = te_col + 2 * 32 * nz1 + threadIdx.x;
  Acc r(M, Mb, &a);
  const ADims g = r.dims();
  if ((int)x < g.nx && (int)y < g.ny) {
    const int64_t sk = g.sk(), c0 = g.idx((int)x, (int)y, 0);
    const float* wb = Mb + a.p[9] + c0;
#pragma unroll 4
    for (uint32_t k = 0; k < nz1; ++k) {
      pre[(0 * nz1 + k) * 32] = scr[c0 + k * sk];
      pre[(1 * nz1 + k) * 32] = scr[fw + c0 + k * sk];
      pre[(2 * nz1 + k) * 32] = scr[2 * fw + c0 + k * sk];
      pre[(3 * nz1 + k) * 32] = scr[3 * fw + c0 + k * sk];
      pre[(4 * nz1 + k) * 32] = wb[k * sk];
    }
  }
  w_solve_sweep(M, Mb, a, scr, fw, tA, tB, 32, (int)x, (int)y, pre, 32);
}
__global__ __launch_bounds__(256) void k_adj_w_level(const float* M, float* Mb, AdjArg a, uint32_t nz) {
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  const uint64_t n0 = a.n[0], n01 = n0 * a.n[1];
  if (t >= n01 * nz) return;
  const uint64_t z = t / n01, rem = t - z * n01, y = rem / n0, x = rem - y * n0;
  w_solve_level(M, Mb, a, (int)x, (int)y, (int)z);
}
__global__ __launch_bounds__(256) void k_adj_w_clear(const float* M, float* Mb, AdjArg a) {
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  const uint64_t n0 = a.n[0];
  if (t >= n0 * a.n[1]) return;
  const uint64_t y = t / n0, x = t - y * n0;
  Acc r(M, Mb, &a);
  w_solve_clear(Mb, a, r.dims(), (int)x, (int)y);
}

// gather-form adjoints: prep over the forward grid, gather over every word of
// the field grid
template <class G>
__global__ __launch_bounds__(256) void k_adj_prep(const float* M, float* Mb, AdjArg a, float* scr) {
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  const uint64_t n0 = a.n[0], n01 = n0 * a.n[1];
  if (t >= n01 * a.n[2]) return;
  const uint64_t z = t / n01, rem = t - z * n01, y = rem / n0, x = rem - y * n0;
  G::prep(M, Mb, a, scr, (int)x, (int)y, (int)z);
}
template <class G>
__global__ __launch_bounds__(256) void k_adj_gather(const float* M, float* Mb, AdjArg a, const float* scr, uint64_t lead,
                                                    uint64_t words) {
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (t >= lead + words) return;
  G::gather(M, Mb, a, scr, (int64_t)t - (int64_t)lead);
}
template <class G>
uint64_t gather_lead(int nx, int ny, int nz, int sJ, int sK) {
  ADims g;
  g.nx = nx; g.ny = ny; g.nz = nz; g.sJ = sJ; g.sK = sK;
  g.rdx = g.rdy = g.rdz = 0.0f;
  return (uint64_t)G::lead(g);
}

// bc_periodic_x then bc_periodic_y on one field leave every halo word
// (i in -3..nx+3, j in -3..ny+3, levels 0..nz) equal to its interior image
// (i mod nx, j mod ny): the pair's adjoint moves each halo word's adjoint onto
// its image.  One thread per halo word (7 halo rows of strideJ words and
// 7 halo columns of ny words per level); several halo words share an image,
// so the adds are atomic.
__global__ __launch_bounds__(256) void k_adj_halo_xy(float* Mb, HaloArg fields, int nx, int ny, int nz1, int sJ, int sK) {
  const uint64_t off = fields.off[blockIdx.y];  // one field per grid row
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  const uint64_t ring = (uint64_t)7 * sJ + (uint64_t)7 * ny;
  if (t >= ring * nz1) return;
  const int k = (int)(t / ring);
  const int u = (int)(t - (uint64_t)k * ring);
  int i, j;
  if (u < 7 * sJ) {
    const int rr = u / sJ;
    j = rr < 3 ? rr - 3 : ny + (rr - 3);
    i = u - rr * sJ - 3;
  } else {
    const int v = u - 7 * sJ, rr = v / ny;
    i = rr < 3 ? rr - 3 : nx + (rr - 3);
    j = v - rr * ny;
  }
  float* f = Mb + off + (uint64_t)sJ * sK * k;
  float* h = f + (i + 3) + (int64_t)sJ * (j + 3);
  const float hb = *h;
  if (hb == 0.0f) return;
  *h = 0.0f;
  const int ii = ((i % nx) + nx) % nx, jj = ((j % ny) + ny) % ny;
  adj_add(f + (ii + 3) + (int64_t)sJ * (jj + 3), hb);
}

hipError_t halo_xy(float* Mb, const HaloArg& fields, int nf, int nx, int ny, int nz, int sJ, int sK,
                   hipStream_t stream) {
  if (nf <= 0) return hipSuccess;
  if (nf > kHaloFields) return hipErrorInvalidValue;
  int nz1 = nz + 1;
  const uint64_t n = ((uint64_t)7 * sJ + (uint64_t)7 * ny) * nz1;
  HaloArg h = fields;
  void* args[] = {(void*)&Mb, (void*)&h, (void*)&nx, (void*)&ny, (void*)&nz1, (void*)&sJ, (void*)&sK};
  return hipLaunchKernel(reinterpret_cast<const void*>(&k_adj_halo_xy), dim3((unsigned)((n + 255) / 256), (unsigned)nf),
                         dim3(256), args, 0, stream);
}

struct GatherEntry {
  const char* name;
  const void* prep;
  const void* gather;
  uint64_t (*lead)(int, int, int, int, int);
};
#define GK(n, G) {n, G::kPrep ? reinterpret_cast<const void*>(&k_adj_prep<G>) : nullptr, \
                  reinterpret_cast<const void*>(&k_adj_gather<G>), &gather_lead<G>},
static const GatherEntry kGather[] = {TE_ADJ_GATHER_KERNELS(GK)};
#undef GK

struct Entry {
  const char* name;
  const void* fn;
  int nw;
  int writes[8];
  bool column;
};

#define AK(n, B, nw, ...) {n, reinterpret_cast<const void*>(&k_adj<B>), nw, __VA_ARGS__, false},
static const Entry kAdj[] = {
    TE_ADJ_POINT_KERNELS(AK)
    {"acoustic_w_solve", reinterpret_cast<const void*>(&k_adj_w_sweep), 3, {9, 10, 11}, true},
};
#undef AK

// Up to 8 regions of n words each (blockIdx.y = region): the reverse sweep's
// write log and restores, one launch per dispatch; 16-byte accesses when
// every region is 16-byte aligned and n is a multiple of 4.
template <bool VEC>
__global__ __launch_bounds__(256) void k_adj_copy(CopyArg c) {
  const uint32_t q = blockIdx.y;
  const uint64_t stride = (uint64_t)gridDim.x * 256u;
  if (VEC) {
    const uint4* __restrict__ s = reinterpret_cast<const uint4*>(c.src[q]);
    uint4* __restrict__ d = reinterpret_cast<uint4*>(c.dst[q]);
    for (uint64_t w = (uint64_t)blockIdx.x * 256u + threadIdx.x; w < c.n / 4; w += stride) d[w] = s[w];
  } else {
    const uint32_t* __restrict__ s = c.src[q];
    uint32_t* __restrict__ d = c.dst[q];
    for (uint64_t w = (uint64_t)blockIdx.x * 256u + threadIdx.x; w < c.n; w += stride) d[w] = s[w];
  }
}

hipError_t copy_regions(const CopyArg& c, int nreg, hipStream_t stream) {
  if (nreg <= 0 || c.n == 0) return hipSuccess;
  bool vec = c.n % 4 == 0;
  for (int q = 0; q < nreg; ++q)
    vec = vec && ((reinterpret_cast<uintptr_t>(c.src[q]) | reinterpret_cast<uintptr_t>(c.dst[q])) & 15u) == 0;
  const uint64_t items = vec ? c.n / 4 : c.n;
  const uint64_t blocks = std::min<uint64_t>((items + 255) / 256, 2048);
  void* args[] = {(void*)&c};
  const void* fn = vec ? reinterpret_cast<const void*>(&k_adj_copy<true>) : reinterpret_cast<const void*>(&k_adj_copy<false>);
  return hipLaunchKernel(fn, dim3((unsigned)blocks, (unsigned)nreg), dim3(256), args, 0, stream);
}

int find(const char* name) {
  for (int i = 0; i < (int)(sizeof(kAdj) / sizeof(kAdj[0])); ++i)
    if (std::strcmp(kAdj[i].name, name) == 0) return i;
  return -1;
}

int writes(int id, int out[8]) {
  for (int q = 0; q < kAdj[id].nw; ++q) out[q] = kAdj[id].writes[q];
  return kAdj[id].nw;
}

hipError_t launch(int id, const uint32_t* A, const float* const* vals, const float* M, float* Mb, int nz, int sJ, int sK,
                  float* scr, uint64_t fw, hipStream_t stream) {
  const Entry& e = kAdj[id];
  AdjArg a;
  a.n[0] = A[0]; a.n[1] = A[1]; a.n[2] = A[2];
  for (int q = 0; q < 32; ++q) {
    a.p[q] = A[4 + q];
    a.v[q] = vals ? vals[q] : M + a.p[q];
  }
  void* args[] = {(void*)&M, (void*)&Mb, (void*)&a};
  for (const GatherEntry& ge : kGather) {
    if (std::strcmp(ge.name, e.name) != 0) continue;
    if (!scr || fw < (uint64_t)sJ * sK * (nz + 1)) return hipErrorInvalidValue;
    if (ge.prep) {
      const uint64_t n = (uint64_t)a.n[0] * a.n[1] * a.n[2];
      void* pargs[] = {(void*)&M, (void*)&Mb, (void*)&a, (void*)&scr};
      if (n > 0) {
        hipError_t r = hipLaunchKernel(ge.prep, dim3((unsigned)((n + 255) / 256)), dim3(256), pargs, 0, stream);
        if (r != hipSuccess) return r;
      }
    }
    const float* cscr = scr;
    uint64_t words = (uint64_t)sJ * sK * (nz + 1), lead = ge.lead(0, 0, nz, sJ, sK);
    void* gargs[] = {(void*)&M, (void*)&Mb, (void*)&a, (void*)&cscr, (void*)&lead, (void*)&words};
    return hipLaunchKernel(ge.gather, dim3((unsigned)((lead + words + 255) / 256)), dim3(256), gargs, 0, stream);
  }
  if (e.column) {  // acoustic_w_solve: coefficients, sweeps, levels, clear
    const uint64_t cols = (uint64_t)a.n[0] * a.n[1];
    if (cols == 0) return hipSuccess;
    if (!scr) return hipErrorInvalidValue;
    uint32_t nzu = (uint32_t)nz;
    const uint64_t pts = cols * nzu;
    void* cargs[] = {(void*)&M, (void*)&a, (void*)&scr, (void*)&fw, (void*)&nzu};
    hipError_t r = hipLaunchKernel(reinterpret_cast<const void*>(&k_adj_w_coef), dim3((unsigned)((pts + 255) / 256)),
                                   dim3(256), cargs, 0, stream);
    if (r != hipSuccess) return r;
    uint32_t nz1 = nzu + 1;
    void* sargs[] = {(void*)&M, (void*)&Mb, (void*)&a, (void*)&scr, (void*)&fw, (void*)&nz1};
    const size_t lds_pre = (size_t)7 * 32 * nz1 * sizeof(float), lds = (size_t)2 * 64 * nz1 * sizeof(float);
    if (lds_pre <= 65536) {
      r = hipLaunchKernel(reinterpret_cast<const void*>(&k_adj_w_sweep_pre), dim3((unsigned)((cols + 31) / 32)), dim3(32),
                          sargs, lds_pre, stream);
    } else {
      if (lds > 65536) return hipErrorInvalidConfiguration;  // nz <= 127
      r = hipLaunchKernel(e.fn, dim3((unsigned)((cols + 63) / 64)), dim3(64), sargs, lds, stream);
    }
    if (r != hipSuccess) return r;
    void* largs[] = {(void*)&M, (void*)&Mb, (void*)&a, (void*)&nzu};
    r = hipLaunchKernel(reinterpret_cast<const void*>(&k_adj_w_level), dim3((unsigned)((pts + 255) / 256)), dim3(256),
                        largs, 0, stream);
    if (r != hipSuccess) return r;
    return hipLaunchKernel(reinterpret_cast<const void*>(&k_adj_w_clear), dim3((unsigned)((cols + 255) / 256)),
                           dim3(256), args, 0, stream);
  }
  const uint64_t n = (uint64_t)a.n[0] * a.n[1] * a.n[2];
  if (n == 0) return hipSuccess;
  if ((n + 255) / 256 > 0x7fffffffull) return hipErrorInvalidConfiguration;
  return hipLaunchKernel(e.fn, dim3((unsigned)((n + 255) / 256)), dim3(256), args, 0, stream);
}

}  // namespace adj
}  // namespace te
