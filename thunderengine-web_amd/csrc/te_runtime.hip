// Program executor and the C ABI (include/te.h).
//
// The executor owns one flat device memory M (every buffer of a program,
// word addressed, one hipMalloc) and launches the program's kernels in order on
// one HIP stream.  A full step is captured once into a HIP graph and replayed.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <array>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <set>

#include "te_internal.hpp"
#include "te_kernels.hpp"

using te::KArg;
using te::Op;

namespace te {

struct KernelInfo {
  const char* name;
  const void* fn;
};
#define TE_ROW(n) {#n, (const void*)&k_##n},
static const KernelInfo kKernels[] = {TE_KERNEL_LIST(TE_ROW) TE_WIDE_LIST(TE_ROW)};
#undef TE_ROW
static const int kNumKernels = (int)(sizeof(kKernels) / sizeof(kKernels[0]));

int kernel_id(const char* name) {
  for (int i = 0; i < kNumKernels; ++i)
    if (std::strcmp(kKernels[i].name, name) == 0) return i;
  return -1;
}
const char* kernel_name(int kid) { return (kid >= 0 && kid < kNumKernels) ? kKernels[kid].name : "?"; }

uint64_t Layout::field(const std::string& n) const {
  for (size_t i = 0; i < fields.size(); ++i)
    if (fields[i] == n) return arena + i * field_words;
  return UINT64_MAX;
}

// Block shape for a thread grid: 64 lanes along i (one wavefront, coalesced
// i-fastest accesses), then j, then k, up to 256 threads.
static dim3 block_for(const uint32_t* n) {
  auto p2 = [](uint32_t v) { uint32_t r = 1; while (r < v) r <<= 1; return r; };
  uint32_t bx = std::min<uint32_t>(64, p2(std::max<uint32_t>(n[0], 1)));
  uint32_t by = std::min<uint32_t>(256 / bx, p2(std::max<uint32_t>(n[1], 1)));
  uint32_t bz = std::min<uint32_t>(256 / (bx * by), p2(std::max<uint32_t>(n[2], 1)));
  return dim3(bx, by, bz);
}

// Block dims, grid blocks and the division constants te_block() uses to
// renumber workgroups for the XCDs (te_device.hpp).
static bool launch_shape(KArg& a, const dim3& b, const dim3& g) {
  if ((uint64_t)g.x * g.y * g.z >= (1u << 24) || g.x > (1u << 16) || g.y > (1u << 16)) return false;
  auto magic = [](uint32_t d, uint32_t* w) {
    const uint64_t m = ((uint64_t)1 << 40) / d + (((uint64_t)1 << 40) % d != 0);
    w[0] = (uint32_t)m;
    w[1] = (uint32_t)(m >> 32);
  };
  a.lg[0] = b.x; a.lg[1] = b.y; a.lg[2] = b.z;
  a.lg[3] = g.x; a.lg[4] = g.y;
  magic(g.x, &a.lg[5]);
  magic(g.y, &a.lg[7]);
  a.lg[9] = g.x * g.y * g.z;
  return true;
}

static KArg karg_of(const Op& op) {
  KArg a{};
  a.n[0] = op.A[0]; a.n[1] = op.A[1]; a.n[2] = op.A[2];
  for (int q = 0; q < 48; ++q) a.p[q] = op.A[4 + q];
  return a;
}

// Column kernels: one lane per (i, j) column sweeping k serially, a wave of
// TE_COLUMN_LANES columns along i (64 by default: measured fastest at
// 192 x 192 x 80 against 32, 16 and 8, which split rows into shorter
// segments without adding resident waves at these kernels' register use).
static bool is_column_kernel(int kid) {
  static const char* const cols[] = {"acoustic_w_solve", "turb_scalar_vert_implicit",
                                     "turb_u_tendency", "turb_v_tendency", "turb_w_tendency", "tke_transport"};
  const char* nm = kKernels[kid].name;
  for (const char* n : cols) {
    const size_t l = std::strlen(n);
    if (std::strncmp(nm, n, l) == 0 && (nm[l] == 0 || std::strcmp(nm + l, "_lds") == 0)) return true;
  }
  return false;
}
// Column kernels whose elimination coefficients live in dynamic LDS:
// 2 * nz floats per lane.
static bool uses_column_lds(int kid) {
  const char* nm = kKernels[kid].name;
  const size_t l = std::strlen(nm);
  return l > 4 && std::strcmp(nm + l - 4, "_lds") == 0;
}
// The turbulence sweeps launched as 64 columns x AC_WAVES waves (k_turb.hip:
// svi_waves_kernel, uv_waves_kernel, face_waves_kernel).
static bool wave_sweep(int kid) {
  static const int ids[] = {kernel_id("turb_scalar_vert_implicit_lds"), kernel_id("turb_u_tendency_lds"),
                            kernel_id("turb_v_tendency_lds"), kernel_id("turb_w_tendency_lds"),
                            kernel_id("tke_transport_lds")};
  for (int id : ids)
    if (kid == id) return true;
  return false;
}
static uint32_t column_lanes() {
  static uint32_t lanes = [] {
    const char* e = std::getenv("TE_COLUMN_LANES");
    int v = e ? std::atoi(e) : 64;
    return (uint32_t)((v == 8 || v == 16 || v == 32 || v == 64) ? v : 64);
  }();
  return lanes;
}

static hipError_t launch(te_ctx* c, const Op& op) {
  KArg a = karg_of(op);
  for (int q = 0; q < 48; ++q) a.q[q] = c->dM + a.p[q];
  if (a.n[0] == 0 || a.n[1] == 0 || a.n[2] == 0) return hipSuccess;
  dim3 b = block_for(a.n);
  static const int kAdvFused = kernel_id("adv_fused"), kTkeEpst = kernel_id("tke_epst");
  if (op.kid == kAdvFused) b = dim3(64, 4, 1);  // one row of cells per wave (k_adv_fused)
  if (op.kid == kTkeEpst) a.n[2] = (a.n[2] + 1) / 2;  // two levels per thread (k_tke_epst)
  static const int kAcColumn = kernel_id("ac_column");
  if (op.kid == kAcColumn) {  // 64 columns of a row per workgroup, AC_WAVES waves over the levels
    const uint32_t nz = op.nz, lw = (nz + AC_WAVES - 1) / AC_WAVES;
    if (nz < 4 || lw > 16) return hipErrorInvalidConfiguration;
    b = dim3(64, AC_WAVES, 1);
    const dim3 g((a.n[0] + 63) / 64, a.n[1], 1);
    if (!launch_shape(a, b, g)) return hipErrorInvalidConfiguration;
    const void* fn = ac_column_kernel((int)lw);
    const size_t dyn = (size_t)2 * nz * 64 * sizeof(float);  // per level and lane: mm / cA / w and cB
    if (dyn > 65536) {
      const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
      if (e != hipSuccess) return e;
    }
    void* args[] = {(void*)&c->dM, (void*)&a};
    return hipLaunchKernel(fn, g, b, args, dyn, c->stream);
  }
  if (a.n[2] == 1 && op.nz >= 4 && wave_sweep(op.kid)) {
    // 64 columns of a row per workgroup, AC_WAVES waves over the levels
    // (levels 0 .. nz-1, or 1 .. nz-1 for the face-level sweeps); shallower
    // columns take the one-lane-per-column form below
    static const int kSviLds = kernel_id("turb_scalar_vert_implicit_lds"), kULds = kernel_id("turb_u_tendency_lds"),
                     kVLds = kernel_id("turb_v_tendency_lds"), kWLds = kernel_id("turb_w_tendency_lds");
    const bool face = op.kid != kSviLds && op.kid != kULds && op.kid != kVLds;
    const uint32_t lw = (op.nz - (face ? 1 : 0) + AC_WAVES - 1) / AC_WAVES;
    if (lw > 16) return hipErrorInvalidConfiguration;
    b = dim3(64, AC_WAVES, 1);
    // the n0 x n1 columns numbered row by row, 64 per workgroup (wave_column)
    const uint64_t nb = ((uint64_t)a.n[0] * a.n[1] + 63) / 64;
    const dim3 g = nb <= 65535 ? dim3((uint32_t)nb, 1, 1) : dim3(4096, (uint32_t)((nb + 4095) / 4096), 1);
    if (!launch_shape(a, b, g)) return hipErrorInvalidConfiguration;
    void* args[] = {(void*)&c->dM, (void*)&a};
    const void* fn = op.kid == kSviLds ? svi_waves_kernel((int)lw)
                     : !face       ? uv_waves_kernel(op.kid == kULds ? 0 : 1, (int)lw)
                                   : face_waves_kernel(op.kid == kWLds ? 0 : 1, (int)lw);
    return hipLaunchKernel(fn, g, b, args, 0, c->stream);
  }
  size_t lds = 0;
  if (op.flat) {  // runs of 4 words per thread over the op's n0*n1*n2 words (te_flat)
    const uint64_t words = (uint64_t)a.n[0] * a.n[1] * a.n[2], runs = (words + 3) / 4;
    b = dim3(256, 1, 1);
    const dim3 g((uint32_t)((runs + 255) / 256), 1, 1);
    if (!launch_shape(a, b, g)) return hipErrorInvalidConfiguration;
    a.lg[10] = 1;
    void* args[] = {(void*)&c->dM, (void*)&a};
    return hipLaunchKernel(kKernels[op.kid].fn, g, b, args, 0, c->stream);
  }
  if (a.n[2] == 1 && is_column_kernel(op.kid)) {
    b = dim3(column_lanes(), 1, 1);
    if (uses_column_lds(op.kid)) lds = (size_t)2 * op.nz * b.x * sizeof(float);
  }
  if (op.tuned && !op.tplane) b = dim3(op.tb[0], op.tb[1], op.tb[2]);
  if (op.tuned ? op.tplane : op.plane) {  // (i, j) points numbered consecutively, level on grid y (te_tid3)
    const uint64_t pts = (uint64_t)a.n[0] * a.n[1];
    b = dim3(256, 1, 1);
    const dim3 g((uint32_t)((pts + 255) / 256), a.n[2], 1);
    if (pts + 256 > (1u << 24) || a.n[0] > (1u << 16) || !launch_shape(a, b, g)) return hipErrorInvalidConfiguration;
    const uint64_t m = ((uint64_t)1 << 40) / a.n[0] + (((uint64_t)1 << 40) % a.n[0] != 0);
    a.lg[11] = 1;
    a.lg[12] = (uint32_t)m;
    a.lg[13] = (uint32_t)(m >> 32);
    void* args[] = {(void*)&c->dM, (void*)&a};
    return hipLaunchKernel(kKernels[op.kid].fn, g, b, args, 0, c->stream);
  }
  dim3 g((a.n[0] + b.x - 1) / b.x, (a.n[1] + b.y - 1) / b.y, (a.n[2] + b.z - 1) / b.z);
  if (!launch_shape(a, b, g)) return hipErrorInvalidConfiguration;  // beyond te_block()'s exact range
  void* args[] = {(void*)&c->dM, (void*)&a};
  return hipLaunchKernel(op.fn ? op.fn : kKernels[op.kid].fn, g, b, args, lds, c->stream);
}

// region_hash (replay.js:408-418): per-lane sums of a 32-bit mix, reduced
// with one atomic add per block.
__global__ void k_region_hash(const uint32_t* M, uint64_t off, uint64_t n, uint32_t* out) {
  uint32_t s = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t h = (M[off + i] ^ ((uint32_t)i * 0x9E3779B1u)) * 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    s += h;
  }
  for (int d = 32; d > 0; d >>= 1) s += __shfl_down(s, d, 64);
  __shared__ uint32_t part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, part[0] + part[1] + part[2] + part[3]);
}
__global__ void k_zero_word(uint32_t* p) { *p = 0u; }
__global__ void k_zero_words(uint32_t* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace te

// Hash accumulators of te_region_hashes_noalloc: one word per region, in a
// device array of the library (each device holds its own copy until the
// process ends), so that entry point allocates nothing.
static __device__ __align__(256) uint32_t scratch[TE_NOALLOC_MAX_REGIONS];

// Every binary32 bit pattern through fast_* and rr_*, and through recip and
// the IEEE division (te_selftest_math): out[0..5] mismatch counts, [6..11]
// deferrals (64-bit), first[0..5].
__global__ __launch_bounds__(256) void k_selftest_math(unsigned long long* out, uint32_t* first) {
  const float cpdrd = te::CP / te::RD, rddcv = te::RD / (te::CP - te::RD), third = 0.33333334f;
  uint32_t bad[6] = {0, 0, 0, 0, 0, 0}, slow[6] = {0, 0, 0, 0, 0, 0};
  for (uint64_t n = blockIdx.x * 256ull + threadIdx.x; n < (1ull << 32); n += (uint64_t)gridDim.x * 256ull) {
    const uint32_t bits = (uint32_t)n;
    const float x = __uint_as_float(bits);
    float r[5], q[5];
    r[0] = te::fast_log(x); q[0] = te::rr_log(x);
    r[1] = te::fast_exp(x); q[1] = te::rr_exp(x);
    r[2] = te::fast_pow(x, cpdrd); q[2] = te::rr_pow(x, cpdrd);
    r[3] = te::fast_pow(x, rddcv); q[3] = te::rr_pow(x, rddcv);
    r[4] = te::fast_pow(x, third); q[4] = te::rr_pow(x, third);
    double d[5] = {log((double)x), exp((double)x), exp((double)cpdrd * log((double)x)),
                   exp((double)rddcv * log((double)x)), exp((double)third * log((double)x))};
    const bool special[5] = {!(x > 0.0f) || (bits & 0x7fffffffu) == 0x7f800000u, !(te::ab(x) < 87.0f),
                             !(x > 0.0f) || (bits & 0x7fffffffu) == 0x7f800000u,
                             !(x > 0.0f) || (bits & 0x7fffffffu) == 0x7f800000u,
                             !(x > 0.0f) || (bits & 0x7fffffffu) == 0x7f800000u};
    for (int f = 0; f < 5; ++f) {
      if (__float_as_uint(r[f]) != __float_as_uint(q[f])) {
        ++bad[f];
        atomicMin(&first[f], bits);
      }
      float t;
      if (!special[f] && !te::f32_clear(d[f], t, f < 2 ? 0x1p-40 : 0x1p-32)) ++slow[f];
    }
    // recip against the IEEE division (deferrals: x outside [2^-64, 2^64])
    const float rc = te::recip(x), qc = 1.0f / x;
    if (__float_as_uint(rc) != __float_as_uint(qc) && !(rc != rc && qc != qc)) {
      ++bad[5];
      atomicMin(&first[5], bits);
    }
    const float ax = te::ab(x);
    if (!(ax >= 0x1p-64f && ax <= 0x1p64f)) ++slow[5];
  }
  for (int f = 0; f < 6; ++f) {
    if (bad[f]) atomicAdd(&out[f], (unsigned long long)bad[f]);
    if (slow[f]) atomicAdd(&out[6 + f], (unsigned long long)slow[f]);
  }
}

// ------------------------------------------------------------------ helpers
static int fail(te_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}
static int hip_fail(te_ctx* c, hipError_t e, const char* what) {
  return fail(c, TE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define TE_HIP(call, what)                       \
  do {                                           \
    hipError_t e_ = (call);                      \
    if (e_ != hipSuccess) return hip_fail(c, e_, what); \
  } while (0)

static int validate_ops(te_ctx* c) {
  for (size_t i = 0; i < c->ops.size(); ++i) {
    const Op& op = c->ops[i];
    const char* nm = te::kernel_name(op.kid);
    if (std::strcmp(nm, "_copy") == 0) {
      if ((uint64_t)op.A[4] + op.A[0] > c->words || (uint64_t)op.A[5] + op.A[0] > c->words)
        return fail(c, TE_ERR_ARG, "copy op " + std::to_string(i) + " leaves the memory");
      continue;
    }
    if (std::strcmp(nm, "copy_fields") == 0)
      for (int q = 1; q <= 8; ++q)
        if ((uint64_t)op.A[4 + q] + op.A[0] > c->words)
          return fail(c, TE_ERR_ARG, "copy_fields op " + std::to_string(i) + " leaves the memory");
    for (int q = 0; q < 48; ++q)
      if (op.A[4 + q] >= c->words) return fail(c, TE_ERR_ARG, "op " + std::to_string(i) + " (" + nm + ") parameter offset outside memory");
    if (std::strcmp(nm, "adv_fused") == 0) {
      // the fields a fused advection writes (tendency, chained update) must
      // not be ones its workgroups read at neighbouring points
      for (int w : {11, 40})
        for (int r : {2, 4, 13, 14})
          if (op.A[4 + w] != 0 && op.A[4 + w] == op.A[4 + r])
            return fail(c, TE_ERR_ARG, "adv_fused op " + std::to_string(i) + " writes a field its stencil reads");
    }
  }
  return TE_OK;
}

// Words of zeroed memory kept in front of M (at least; the guard grows to the
// back guard's size once the program's grids are known): kernels read
// stencil windows up to 3 levels beyond a field's ends, so both guards cover
// a whole field plus 2 levels and every access of a validated program stays
// inside the allocation.
static const uint64_t kFrontGuard = 1u << 20;

static int alloc_memory(te_ctx* c, uint64_t words, uint64_t front = kFrontGuard) {
  front = std::max(front, std::max<uint64_t>(c->front, kFrontGuard));
  front = (front + 63) / 64 * 64;  // M starts 256-byte aligned
  uint32_t* base = nullptr;
  TE_HIP(hipMalloc((void**)&base, (front + words) * 4), "hipMalloc(M)");
  TE_HIP(hipMemsetAsync(base, 0, (front + words) * 4, c->stream), "hipMemset(M)");
  if (c->dBase) {
    TE_HIP(hipMemcpyAsync(base + front, c->dM, c->prog_words * 4, hipMemcpyDeviceToDevice, c->stream), "copy M");
    TE_HIP(hipStreamSynchronize(c->stream), "copy M sync");
    (void)hipFree(c->dBase);
  }
  c->dBase = base;
  c->dM = base + front;
  c->front = front;
  c->words = words;
  return TE_OK;
}

static int ctx_alloc(te_ctx* c, uint64_t words) {
  c->prog_words = words;
  TE_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  TE_HIP(hipEventCreate(&c->ev0), "hipEventCreate");
  TE_HIP(hipEventCreate(&c->ev1), "hipEventCreate");
  int rc = alloc_memory(c, words);
  if (rc != TE_OK) return rc;
  TE_HIP(hipMalloc((void**)&c->dHash, 256), "hipMalloc(hash)");
  return TE_OK;
}

// Guard region: every kernel indexes at most one field footprint
// (strideJ*strideK*(nz+1) words, plus one plane of stencil reach) past a
// parameter's offset, and the offsets were checked to lie inside the program's
// memory.  Before the first run the grid of every op is read back and the
// allocation grown so that footprint past the end of the program's memory is
// still allocated: no launch of a validated program can leave M.
// kernels that map threads themselves (columns, fused advection, copies)
static const std::set<std::string>& own_mapping() {
  static const std::set<std::string> s = {
      "acoustic_w_solve", "acoustic_w_solve_lds", "adv_fused", "_copy", "copy_fields", "turb_sfcstress",
      "turb_u_tendency", "turb_u_tendency_lds", "turb_v_tendency", "turb_v_tendency_lds", "turb_w_tendency",
      "turb_w_tendency_lds", "turb_scalar_vert_implicit", "turb_scalar_vert_implicit_lds", "tke_transport",
      "tke_transport_lds", "ac_column"};
  return s;
}

static int grow_guards(te_ctx* c) {
  if (c->guard_ok) return TE_OK;
  uint64_t need = 0;
  std::map<uint32_t, std::array<uint32_t, 5>> seen;  // grid parameter blocks already read
  // pointwise ops that may run flat (te_flat): kernel -> the parameters it
  // accesses in 16-byte runs (fields; profiles are read per level)
  static const std::map<std::string, std::vector<int>> flatOps = {
      {"zero_fields", {1, 2, 3, 4, 5, 6, 7, 8}},
      {"add_fields", {1, 2, 3, 4, 5, 6, 7, 8}},
      {"tendency_reset_add", {1, 2, 3, 4, 5, 6}},
      {"stage_state_reset", {1, 2, 3, 4, 5, 6, 7, 8, 9, 10}},
      {"copy_fields", {1, 2, 3, 4, 5, 6, 7, 8}},
      {"acoustic_begin", {2, 3, 4, 5}},
      {"massflux_from_state", {1, 2, 3, 6, 7, 8}},
      {"velocity_to_massflux", {3, 4, 5}},
      {"build_sadv", {2, 3}},
      {"update_diagnostics_moist", {1, 2, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19}},
      {"_copy", {0, 1}}};
  auto aligned = [](const Op& op, const std::vector<int>& ps) {
    for (int q : ps)
      if (op.A[4 + q] % 4 != 0) return false;
    return true;
  };
  for (Op& op : c->ops) {
    const char* kn = te::kernel_name(op.kid);
    const auto fl = flatOps.find(kn);
    if (std::strcmp(kn, "_copy") == 0) {
      op.flat = aligned(op, fl->second);
      continue;
    }
    uint32_t d[5];
    auto it = seen.find(op.A[4]);
    if (it != seen.end()) {
      std::memcpy(d, it->second.data(), sizeof(d));
    } else {
      TE_HIP(hipMemcpy(d, c->dM + op.A[4], sizeof(d), hipMemcpyDeviceToHost), "read grid");
      seen[op.A[4]] = {d[0], d[1], d[2], d[3], d[4]};
    }
    const int32_t nz = (int32_t)d[2], sJ = (int32_t)d[3], sK = (int32_t)d[4];
    if (nz < 1 || sJ < 1 || sK < 1 || nz > 100000 || sJ > 100000 || sK > 100000)
      return fail(c, TE_ERR_ARG, std::string("op ") + te::kernel_name(op.kid) + ": invalid grid parameters");
    op.nz = nz;
    if (fl != flatOps.end()) {
      const bool wholeGrid = op.A[0] == (uint32_t)sJ && op.A[1] == (uint32_t)sK && op.A[2] == (uint32_t)(nz + 1);
      const bool words1d = std::strcmp(kn, "copy_fields") == 0 && op.A[1] == 1 && op.A[2] == 1;
      op.flat = (wholeGrid || words1d) && aligned(op, fl->second);
    }
    const std::set<std::string>& ownMapping = own_mapping();
    // (te_div is exact for numerators below 2^24, divisors up to 2^16)
    op.plane = !op.flat && !ownMapping.count(kn) && op.A[0] % 64 != 0 && op.A[1] > 1 && op.A[0] <= (1u << 16) &&
               (uint64_t)op.A[0] * op.A[1] + 256 <= (1u << 24);
    if (te::wave_sweep(op.kid) ? nz > AC_WAVES * 16
                          : te::uses_column_lds(op.kid) && (size_t)2 * nz * te::column_lanes() * sizeof(float) > 65536)
      return fail(c, TE_ERR_UNSUPPORTED, std::string(te::kernel_name(op.kid)) + ": nz too large for the column sweep");
    // the wave sweeps address a column within a level by a 32-bit byte offset
    if (te::wave_sweep(op.kid) && (uint64_t)sJ * sK >= (1ull << 30))
      return fail(c, TE_ERR_UNSUPPORTED, std::string(te::kernel_name(op.kid)) + ": a level spans 2^32 bytes or more");
    need = std::max<uint64_t>(need, (uint64_t)sJ * sK * (uint64_t)(nz + 3));
    if (std::strcmp(kn, "adv_fused") == 0) {
      int32_t stag = 0, kind = 0;
      TE_HIP(hipMemcpy(&stag, c->dM + op.A[4 + 1], 4, hipMemcpyDeviceToHost), "read stagger");
      TE_HIP(hipMemcpy(&kind, c->dM + op.A[4 + 3], 4, hipMemcpyDeviceToHost), "read flux kind");
      op.fn = te::adv_fused_kernel(kind, stag);
    }
    if (std::strcmp(te::kernel_name(op.kid), "turb_sfcstress") == 0) {
      int32_t mode = 0;
      TE_HIP(hipMemcpy(&mode, c->dM + op.A[4 + 5], 4, hipMemcpyDeviceToHost), "read mode");
      if (mode != 0 && mode != 4)
        return fail(c, TE_ERR_UNSUPPORTED, "turb_sfcstress: surface-flux stress modes (bbc=3) are not on this path");
    }
  }
  const uint64_t want = c->prog_words + need;
  if (want > c->words || need > c->front) {
    int rc = alloc_memory(c, std::max(want, c->words), need);
    if (rc != TE_OK) return rc;
  }
  c->guard_ok = true;
  return TE_OK;
}

int te::ensure_guard(te_ctx* c) { return grow_guards(c); }

static int run_ops(te_ctx* c, int first, int last) {
  for (int i = first; i <= last; ++i) {
    hipError_t e = te::launch(c, c->ops[i]);
    if (e != hipSuccess) return hip_fail(c, e, (std::string("launch ") + te::kernel_name(c->ops[i].kid)).c_str());
  }
  return TE_OK;
}

static int ensure_graph(te_ctx* c) {
  if (c->graph || c->graph_failed) return TE_OK;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    c->graph_failed = true;
    return TE_OK;
  }
  int rc = run_ops(c, 0, (int)c->ops.size() - 1);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc != TE_OK || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    c->graph_failed = true;
    (void)hipGetLastError();
    return TE_OK;
  }
  e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->graph = nullptr;
    c->graph_failed = true;
    (void)hipGetLastError();
  }
  return TE_OK;
}

int te::run_ops_plain(te_ctx* c, int first, int last) {
  if (!c || !c->dM) return fail(c, TE_ERR_STATE, "context has no memory");
  if (first < 0 || last >= (int)c->ops.size() || first > last) return fail(c, TE_ERR_ARG, "bad op range");
  int rc = grow_guards(c);
  if (rc != TE_OK) return rc;
  return run_ops(c, first, last);
}

extern "C" {

int te_program_create(uint64_t mem_words, int32_t nops, const char* const* names, const uint32_t* args, te_ctx** out) {
  if (!out || (nops > 0 && (!names || !args)) || mem_words == 0) return TE_ERR_ARG;
  *out = nullptr;
  te_ctx* c = new te_ctx();
  c->ops.resize(nops);
  for (int i = 0; i < nops; ++i) {
    int kid = te::kernel_id(names[i]);
    if (kid < 0) {
      int rc = fail(c, TE_ERR_UNSUPPORTED, std::string("no kernel '") + names[i] + "' on this path");
      *out = c;
      return rc;
    }
    c->ops[i].kid = kid;
    c->op_names.push_back(names[i]);
    std::memcpy(c->ops[i].A, args + (size_t)i * 64, 64 * 4);
    if (std::strcmp(names[i], "_copy") == 0) {  // copy slot: A[0]=n, A[4]=src, A[5]=dst
      c->ops[i].A[1] = c->ops[i].A[2] = 1;
    }
  }
  *out = c;
  c->words = c->prog_words = mem_words;
  int rc = validate_ops(c);
  if (rc != TE_OK) return rc;
  return ctx_alloc(c, mem_words);
}

int te_upload_words(te_ctx* c, uint64_t off, const uint32_t* src, uint64_t n) {
  if (!c || !src || n > c->prog_words || off > c->prog_words - n) return fail(c, TE_ERR_ARG, "upload outside memory");
  TE_HIP(hipMemcpyAsync(c->dM + off, src, n * 4, hipMemcpyHostToDevice, c->stream), "upload");
  TE_HIP(hipStreamSynchronize(c->stream), "upload sync");
  return TE_OK;
}

int te_download_words(te_ctx* c, uint64_t off, uint32_t* dst, uint64_t n) {
  if (!c || !dst || n > c->prog_words || off > c->prog_words - n) return fail(c, TE_ERR_ARG, "download outside memory");
  if (!c->dM) {  // encoded program without device memory: read the host image
    if (off + n > c->image.size()) return fail(c, TE_ERR_STATE, "no memory image");
    std::memcpy(dst, c->image.data() + off, n * 4);
    return TE_OK;
  }
  TE_HIP(hipMemcpyAsync(dst, c->dM + off, n * 4, hipMemcpyDeviceToHost, c->stream), "download");
  TE_HIP(hipStreamSynchronize(c->stream), "download sync");
  return TE_OK;
}

int te_run(te_ctx* c, int32_t nsteps, int32_t first, int32_t last) {
  if (!c || !c->dM) return fail(c, TE_ERR_STATE, "context has no memory");
  const int n = (int)c->ops.size();
  if (last < 0 || last >= n) last = n - 1;
  if (first < 0 || first > last || nsteps < 0) return fail(c, TE_ERR_ARG, "bad op range");
  int rc0 = grow_guards(c);
  if (rc0 != TE_OK) return rc0;
  const bool full = first == 0 && last == n - 1;
  if (full && nsteps > 1) ensure_graph(c);
  TE_HIP(hipEventRecord(c->ev0, c->stream), "event");
  for (int s = 0; s < nsteps; ++s) {
    if (full && c->graph) {
      TE_HIP(hipGraphLaunch(c->graph, c->stream), "hipGraphLaunch");
    } else {
      int rc = run_ops(c, first, last);
      if (rc != TE_OK) return rc;
    }
  }
  TE_HIP(hipEventRecord(c->ev1, c->stream), "event");
  return TE_OK;
}

int te_profile_ops(te_ctx* c, int32_t nsteps, float* op_ms) {
  if (!c || !c->dM || !op_ms || nsteps < 1) return fail(c, TE_ERR_ARG, "te_profile_ops: bad arguments");
  const int n = (int)c->ops.size();
  for (int i = 0; i < n; ++i)
    if (c->ops[i].kid < 0) return fail(c, TE_ERR_UNSUPPORTED, "kernel '" + c->op_names[i] + "' is not on this path");
  int rc0 = grow_guards(c);
  if (rc0 != TE_OK) return rc0;
  std::vector<hipEvent_t> ev(n + 1, nullptr);
  int rc = TE_OK;
  for (int i = 0; i <= n && rc == TE_OK; ++i)
    if (hipEventCreate(&ev[i]) != hipSuccess) rc = fail(c, TE_ERR_HIP, "hipEventCreate");
  for (int i = 0; i < n; ++i) op_ms[i] = 0.0f;
  for (int s = 0; s < nsteps && rc == TE_OK; ++s) {
    if (hipEventRecord(ev[0], c->stream) != hipSuccess) { rc = fail(c, TE_ERR_HIP, "event"); break; }
    for (int i = 0; i < n && rc == TE_OK; ++i) {
      hipError_t e = te::launch(c, c->ops[i]);
      if (e == hipSuccess) e = hipEventRecord(ev[i + 1], c->stream);
      if (e != hipSuccess) rc = hip_fail(c, e, "profiled launch");
    }
    if (rc == TE_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(c, TE_ERR_HIP, "sync");
    for (int i = 0; i < n && rc == TE_OK; ++i) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) rc = fail(c, TE_ERR_HIP, "elapsed");
      op_ms[i] += ms;
    }
  }
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

int te_sync(te_ctx* c) {
  if (!c) return TE_ERR_ARG;
  TE_HIP(hipStreamSynchronize(c->stream), "sync");
  TE_HIP(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1), "elapsed");
  return TE_OK;
}

int te_last_run_ms(te_ctx* c, float* ms) {
  if (!c || !ms) return TE_ERR_ARG;
  TE_HIP(hipEventSynchronize(c->ev1), "event sync");
  TE_HIP(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1), "elapsed");
  *ms = c->last_ms;
  return TE_OK;
}

int te_autotune(te_ctx* c, int32_t reps, int32_t force) {
  if (!c || !c->dM) return fail(c, TE_ERR_STATE, "context has no memory");
  if (reps < 1) return fail(c, TE_ERR_ARG, "reps must be at least 1");
  const int n = (int)c->ops.size();
  for (int i = 0; i < n; ++i)
    if (c->ops[i].kid < 0) return fail(c, TE_ERR_UNSUPPORTED, "kernel '" + c->op_names[i] + "' is not on this path");
  int rc = grow_guards(c);
  if (rc != TE_OK) return rc;
  // the trial steps change the memory: every trial starts from a copy of it,
  // and the copy is put back at the end
  struct Hold {
    uint32_t* p = nullptr;
    std::vector<hipEvent_t> ev;
    ~Hold() {
      if (p) (void)hipFree(p);
      for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    }
  } hold;
  const uint64_t bytes = c->prog_words * 4;
  TE_HIP(hipMalloc((void**)&hold.p, bytes), "hipMalloc(autotune copy)");
  TE_HIP(hipMemcpyAsync(hold.p, c->dM, bytes, hipMemcpyDeviceToDevice, c->stream), "autotune copy");
  hold.ev.assign(n + 1, nullptr);
  for (hipEvent_t& e : hold.ev) TE_HIP(hipEventCreate(&e), "hipEventCreate");
  auto restore = [&]() { return hipMemcpyAsync(c->dM, hold.p, bytes, hipMemcpyDeviceToDevice, c->stream); };

  // dispatches on the generic 3-D path: not the column sweeps, the fused
  // advection or the flat copies, which fix their own thread mapping
  // (adv_fused too, among the shapes that keep one row of cells per wave)
  static const int kAdvFused = te::kernel_id("adv_fused");
  std::vector<char> eligible(n, 0), planeOk(n, 0);
  for (int i = 0; i < n; ++i) {
    const te::Op& op = c->ops[i];
    const char* kn = te::kernel_name(op.kid);
    if (op.flat || (own_mapping().count(kn) && op.kid != kAdvFused) ||
        (op.A[2] == 1 && te::is_column_kernel(op.kid)))
      continue;
    if (op.A[0] == 0 || op.A[1] == 0 || op.A[2] == 0) continue;
    eligible[i] = 1;
    planeOk[i] = op.A[1] > 1 && op.A[0] <= (1u << 16) && (uint64_t)op.A[0] * op.A[1] + 256 <= (1u << 24);
  }
  struct Shape { bool plane; uint16_t x, y, z; };
  static const Shape kShapes[] = {{false, 64, 4, 1}, {false, 64, 2, 2}, {false, 64, 1, 4}, {false, 64, 2, 1},
                                  {false, 64, 1, 2}, {false, 64, 1, 1}, {false, 128, 2, 1}, {false, 128, 1, 2},
                                  {false, 128, 1, 1}, {false, 256, 1, 1}, {false, 32, 8, 1}, {false, 32, 4, 2},
                                  {true, 256, 1, 1}};
  constexpr int kCand = (int)(sizeof(kShapes) / sizeof(kShapes[0]));
  auto p2 = [](uint32_t v) { uint32_t r = 1; while (r < v) r <<= 1; return r; };
  // the grid a shape gives must stay within launch()'s limits (te_block's exact range)
  auto grid_ok = [](uint64_t gx, uint64_t gy, uint64_t gz) {
    return gx * gy * gz < (1u << 24) && gx <= (1u << 16) && gy <= (1u << 16);
  };
  auto fits = [&](int i, const Shape& sh) {
    const te::Op& op = c->ops[i];
    if (sh.plane) {
      const uint64_t pts = (uint64_t)op.A[0] * op.A[1];
      return op.kid != kAdvFused && planeOk[i] != 0 && grid_ok((pts + 255) / 256, op.A[2], 1);
    }
    if (!grid_ok((op.A[0] + sh.x - 1) / sh.x, (op.A[1] + sh.y - 1) / sh.y, (op.A[2] + sh.z - 1) / sh.z)) return false;
    if (op.kid == kAdvFused) return sh.x == 64 && sh.y * sh.z <= 4;
    return sh.x <= std::max<uint32_t>(64, p2(op.A[0])) && sh.y <= p2(op.A[1]) && sh.z <= p2(op.A[2]);
  };
  const std::vector<te::Op> base = c->ops;  // shapes as the context had them
  // every return before the end (a failed launch, event or copy) leaves the
  // context as it was: its own shapes, and the memory from before the trials
  bool done = false;
  auto undo = [&]() {
    if (done) return;
    c->ops = base;
    (void)restore();
    (void)hipStreamSynchronize(c->stream);
  };
  struct OnExit {
    decltype(undo)& f;
    ~OnExit() { f(); }
  } onExit{undo};
  std::vector<te::Op> best = base;
  std::vector<float> bestMs(n, 1e30f), ms(n);

  // each candidate in context: `reps` eager steps from the same memory with
  // an event after every launch, so each dispatch is timed with the caches
  // as the step leaves them (candidate -1: the untuned shapes)
  for (int cand = -1; cand < kCand; ++cand) {
    for (int i = 0; i < n; ++i) {
      if (!eligible[i]) continue;
      te::Op& op = c->ops[i];
      op = base[i];
      op.tuned = false;
      if (cand >= 0 && fits(i, kShapes[cand])) {
        op.tuned = true;
        op.tplane = kShapes[cand].plane;
        op.tb[0] = kShapes[cand].x; op.tb[1] = kShapes[cand].y; op.tb[2] = kShapes[cand].z;
      }
    }
    TE_HIP(restore(), "autotune restore");
    std::fill(ms.begin(), ms.end(), 0.0f);
    for (int s = 0; s < reps; ++s) {
      TE_HIP(hipEventRecord(hold.ev[0], c->stream), "event");
      for (int i = 0; i < n; ++i) {
        TE_HIP(te::launch(c, c->ops[i]), "autotune launch");
        TE_HIP(hipEventRecord(hold.ev[i + 1], c->stream), "event");
      }
      TE_HIP(hipStreamSynchronize(c->stream), "autotune sync");
      for (int i = 0; i < n; ++i) {
        float t = 0.0f;
        TE_HIP(hipEventElapsedTime(&t, hold.ev[i], hold.ev[i + 1]), "elapsed");
        ms[i] += t;
      }
    }
    for (int i = 0; i < n; ++i)
      if (eligible[i] && (cand < 0 || c->ops[i].tuned) && ms[i] < bestMs[i]) {
        bestMs[i] = ms[i];
        best[i] = c->ops[i];
      }
  }

  // keep the per-dispatch choices only if whole steps are faster with them
  auto steps_ms = [&](const std::vector<te::Op>& ops, float* out) -> int {
    c->ops = ops;
    TE_HIP(restore(), "autotune restore");
    TE_HIP(hipEventRecord(hold.ev[0], c->stream), "event");
    for (int s = 0; s < reps; ++s) {
      int r = run_ops(c, 0, n - 1);
      if (r != TE_OK) return r;
    }
    TE_HIP(hipEventRecord(hold.ev[1], c->stream), "event");
    TE_HIP(hipEventSynchronize(hold.ev[1]), "event sync");
    TE_HIP(hipEventElapsedTime(out, hold.ev[0], hold.ev[1]), "elapsed");
    return TE_OK;
  };
  float tBase = 0.0f, tBest = 0.0f;
  for (int rep = 0; rep < 3 && rc == TE_OK; ++rep) {  // interleaved A/B
    float t = 0.0f;
    rc = steps_ms(base, &t);
    tBase += t;
    if (rc == TE_OK) rc = steps_ms(best, &t);
    tBest += t;
  }
  if (rc != TE_OK) {
    c->ops = base;
    return rc;
  }
  const bool keepTuned = force != 0 || tBest < 0.995f * tBase;
  c->ops = keepTuned ? best : base;
  int tuned = 0;
  for (int i = 0; i < n; ++i) tuned += c->ops[i].tuned ? 1 : 0;
  TE_HIP(restore(), "autotune restore");
  TE_HIP(hipStreamSynchronize(c->stream), "autotune sync");
  done = true;
  c->tune_base_ms = tBase / (3.0f * reps);
  c->tune_best_ms = tBest / (3.0f * reps);
  const bool hadGraph = c->graph != nullptr;
  if (hadGraph) {  // the step graph holds the old shapes: capture it again
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  c->graph_failed = false;
  c->tuned_ops = tuned;
  return hadGraph ? ensure_graph(c) : TE_OK;
}

int te_autotune_ms(const te_ctx* c, float* untuned, float* tuned) {
  if (!c || !untuned || !tuned) return TE_ERR_ARG;
  *untuned = c->tune_base_ms;
  *tuned = c->tune_best_ms;
  return TE_OK;
}

int32_t te_tuned_ops(const te_ctx* c) { return c ? c->tuned_ops : 0; }

int te_region_hash(te_ctx* c, uint64_t off, uint64_t n, uint32_t* out) {
  if (!c || !out || n > c->words || off > c->words - n) return fail(c, TE_ERR_ARG, "hash outside memory");
  hipLaunchKernelGGL(te::k_zero_word, dim3(1), dim3(1), 0, c->stream, c->dHash);
  hipLaunchKernelGGL(te::k_region_hash, dim3(256), dim3(256), 0, c->stream, (const uint32_t*)c->dM, off, n, c->dHash);
  TE_HIP(hipGetLastError(), "hash launch");
  TE_HIP(hipMemcpyAsync(out, c->dHash, 4, hipMemcpyDeviceToHost, c->stream), "hash copy");
  TE_HIP(hipStreamSynchronize(c->stream), "hash sync");
  return TE_OK;
}

// Device address of `scratch` per device, looked up once per device with that
// device current (the array's host-side name is not a device address).
static uint32_t* noalloc_scratch(int dev) {
  static std::mutex mu;
  static std::map<int, uint32_t*> ptrs;
  std::lock_guard<std::mutex> lock(mu);
  auto it = ptrs.find(dev);
  if (it != ptrs.end()) return it->second;
  uint32_t* p = nullptr;
  if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
  ptrs[dev] = p;
  return p;
}

int te_region_hashes_noalloc(const uint32_t* d_mem, const uint64_t* offs, const uint64_t* counts, int32_t n,
                             uint32_t* out) {
  if (n < 0 || n > TE_NOALLOC_MAX_REGIONS || (n > 0 && (!d_mem || !offs || !counts || !out))) return TE_ERR_ARG;
  for (int32_t r = 0; r < n; ++r)  // byte addresses 4 * (off + count) stay representable
    if (counts[r] > UINT64_MAX / 8 || offs[r] > UINT64_MAX / 8 - counts[r]) return TE_ERR_ARG;
  if (n == 0) return TE_OK;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return TE_ERR_HIP;
  uint32_t* acc = noalloc_scratch(dev);
  if (!acc) return TE_ERR_HIP;
  hipStream_t s = 0;  // the device's null stream: no stream is created
  hipLaunchKernelGGL(te::k_zero_words, dim3((n + 255) / 256), dim3(256), 0, s, acc, (int)n);
  for (int32_t r = 0; r < n; ++r)
    if (counts[r] > 0)
      hipLaunchKernelGGL(te::k_region_hash, dim3(256), dim3(256), 0, s, d_mem, offs[r], counts[r], acc + r);
  if (hipGetLastError() != hipSuccess) return TE_ERR_HIP;
  if (hipMemcpyAsync(out, acc, 4 * (size_t)n, hipMemcpyDeviceToHost, s) != hipSuccess) return TE_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return TE_ERR_HIP;
  return TE_OK;
}

int te_fnv1a(te_ctx* c, uint64_t off, uint64_t n, uint32_t* out) {
  if (!c || !out || n > c->prog_words || off > c->prog_words - n) return fail(c, TE_ERR_ARG, "fnv outside memory");
  std::vector<uint32_t> h(n);
  int rc = te_download_words(c, off, h.data(), n);
  if (rc != TE_OK) return rc;
  uint32_t x = 0x811c9dc5u;
  for (uint64_t i = 0; i < n; ++i) x = (x ^ h[i]) * 0x01000193u;
  *out = x;
  return TE_OK;
}

uint64_t te_memory_words(const te_ctx* c) { return c ? c->prog_words : 0; }
int32_t te_op_count(const te_ctx* c) { return c ? (int32_t)c->ops.size() : 0; }
int te_op_get(const te_ctx* c, int32_t i, const char** name, uint32_t* args64) {
  if (!c || i < 0 || i >= (int32_t)c->ops.size()) return TE_ERR_ARG;
  if (name) *name = i < (int32_t)c->op_names.size() ? c->op_names[i].c_str() : te::kernel_name(c->ops[i].kid);
  if (args64) std::memcpy(args64, c->ops[i].A, 64 * 4);
  return TE_OK;
}

int te_selftest_math(uint64_t mismatch[6], uint32_t first[6], uint64_t slow[6]) {
  if (!mismatch || !first || !slow) return TE_ERR_ARG;
  unsigned long long* out = nullptr;
  uint32_t* fb = nullptr;
  const uint32_t ones[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
  unsigned long long h[12];
  hipError_t e = hipMalloc((void**)&out, sizeof(h));
  if (e == hipSuccess) e = hipMalloc((void**)&fb, sizeof(ones));
  if (e == hipSuccess) e = hipMemset(out, 0, sizeof(h));
  if (e == hipSuccess) e = hipMemcpy(fb, ones, sizeof(ones), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_selftest_math, dim3(1 << 16), dim3(256), 0, 0, out, fb);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(h, out, sizeof(h), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(first, fb, sizeof(ones), hipMemcpyDeviceToHost);
  if (out) (void)hipFree(out);
  if (fb) (void)hipFree(fb);
  if (e != hipSuccess) return TE_ERR_HIP;
  for (int f = 0; f < 6; ++f) {
    mismatch[f] = h[f];
    slow[f] = h[6 + f];
  }
  return TE_OK;
}

int32_t te_kernel_count(void) { return te::kNumKernels; }
const char* te_kernel_name(int32_t i) { return (i >= 0 && i < te::kNumKernels) ? te::kKernels[i].name : nullptr; }

const char* te_last_error(const te_ctx* c) { return c ? c->err.c_str() : "null context"; }

void te_destroy(te_ctx* c) {
  if (!c) return;
  for (te_adjoint* a : c->adjoints) te::adjoint_detach(a);
  te_clock_destroy(c->clock);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  if (c->dBase) (void)hipFree(c->dBase);
  if (c->dHash) (void)hipFree(c->dHash);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// ------------------------------------------------------------------ config
int te_encode(const te_config* cfg, te_ctx** out) {
  if (!cfg || !out) return TE_ERR_ARG;
  *out = nullptr;
  te_ctx* c = new te_ctx();
  *out = c;
  c->has_config = true;
  c->cfg = *cfg;
  if (cfg->zf) {
    c->zf.assign(cfg->zf, cfg->zf + cfg->nz + 1);
    c->cfg.zf = c->zf.data();
  }
  c->cfg.arena_fields = nullptr;  // the caller's list is consumed by the encoder only
  if (cfg->input_sounding) {
    c->sounding = cfg->input_sounding;
    c->cfg.input_sounding = c->sounding.c_str();
  }
  te_config first = *cfg;
  if (cfg->adaptive_dt) first.nsound = 4;  // setTimeStep(dtl, 4) before the clock starts (driver.js:41)
  c->cfg.nsound = first.nsound;
  try {
    te::Encoded e = te::encode_step(first);
    c->layout = std::move(e.L);
    c->image = std::move(e.image);
    c->ops = std::move(e.ops);
    c->op_names = std::move(e.names);
  } catch (const std::exception& ex) {
    return fail(c, TE_ERR_UNSUPPORTED, ex.what());
  }
  c->words = c->prog_words = c->layout.words;
  if (cfg->adaptive_dt) {
    // parameter words for any later (dt, nsound): the most substeps, and slack
    // for payloads that coincide for one dt and not for another
    c->dtl = cfg->dt;
    c->stab_off = c->layout.region.at("stability");
    te_config most = first;
    most.nsound = 192;
    uint64_t words = c->layout.words;
    try {
      words = std::max<uint64_t>(words, te::encode_step(most, false).L.words);
    } catch (const std::exception& ex) {
      return fail(c, TE_ERR_UNSUPPORTED, ex.what());
    }
    words += 4096;
    c->image.resize(words, 0u);
    c->words = c->prog_words = words;
  }
  return TE_OK;
}

int te_create(const te_config* cfg, te_ctx** out) {
  int rc = te_encode(cfg, out);
  if (rc != TE_OK) return rc;
  te_ctx* c = *out;
  for (size_t i = 0; i < c->ops.size(); ++i)
    if (c->ops[i].kid < 0)
      return fail(c, TE_ERR_UNSUPPORTED, "kernel '" + c->op_names[i] + "' is not on this path (microphysics)");
  rc = validate_ops(c);
  if (rc != TE_OK) return rc;
  rc = ctx_alloc(c, c->prog_words);
  if (rc != TE_OK) return rc;
  return te_upload_words(c, 0, c->image.data(), c->image.size());
}

// TimeStep.setTimeStep and the next encode (timestep.js:711-716): the step
// encoded for (dt, nsound) without its arena; the parameter words above the
// arena replace the old ones in the image and on the device, and the
// dispatch list replaces the old one when it differs (then the graph, the
// guard sizes and the tuned shapes are rebuilt).
static int retime(te_ctx* c, float dt, int32_t nsound) {
  if (!c || !c->has_config) return fail(c, TE_ERR_STATE, "not a config context");
  if (!c->adjoints.empty()) return fail(c, TE_ERR_STATE, "the time step of a context with adjoints is fixed");
  if (!(dt > 0.0f) || nsound < 4 || nsound > 192 || nsound % 2 != 0)
    return fail(c, TE_ERR_ARG, "dt must be positive and nsound even in [4, 192]");
  te_config cfg = c->cfg;
  cfg.dt = dt;
  cfg.nsound = nsound;
  te::Encoded e;
  try {
    e = te::encode_step(cfg, false);
  } catch (const std::exception& ex) {
    return fail(c, TE_ERR_ARG, ex.what());
  }
  const te::Layout& L = c->layout;
  if (e.L.arena != L.arena || e.L.field_words != L.field_words || e.L.fields != L.fields)
    return fail(c, TE_ERR_STATE, "the re-encoded step has another memory layout");
  if (e.L.words > c->prog_words || c->image.size() < c->prog_words)
    return fail(c, TE_ERR_STATE, "the re-encoded step's parameter words exceed the context's memory");
  for (const te::Op& op : e.ops)
    if (op.kid < 0) return fail(c, TE_ERR_UNSUPPORTED, "the re-encoded step has kernels off this path");
  const uint64_t pay0 = L.arena + e.skipped;
  std::copy(e.image.begin() + L.arena, e.image.end(), c->image.begin() + pay0);
  std::fill(c->image.begin() + e.L.words, c->image.begin() + c->prog_words, 0u);
  bool same = e.ops.size() == c->ops.size();
  for (size_t i = 0; same && i < e.ops.size(); ++i)
    same = e.ops[i].kid == c->ops[i].kid && std::memcmp(e.ops[i].A, c->ops[i].A, sizeof(e.ops[i].A)) == 0;
  if (!same) {
    std::vector<te::Op> keep = std::move(c->ops);
    c->ops = std::move(e.ops);
    int rc = validate_ops(c);
    if (rc != TE_OK) {
      c->ops = std::move(keep);
      return rc;
    }
    c->op_names = std::move(e.names);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
    c->graph_failed = false;
    c->guard_ok = false;
    c->tuned_ops = 0;
  }
  c->cfg.dt = dt;
  c->cfg.nsound = nsound;
  if (!c->dM) return TE_OK;  // encoded program: the host image is the memory
  TE_HIP(hipMemcpyAsync(c->dM + pay0, c->image.data() + pay0, (c->prog_words - pay0) * 4, hipMemcpyHostToDevice,
                        c->stream),
         "upload step parameters");
  TE_HIP(hipStreamSynchronize(c->stream), "upload sync");
  return same ? TE_OK : grow_guards(c);
}

int te_set_time_step(te_ctx* c, float dt, int32_t nsound) { return retime(c, dt, nsound); }

// stabilityFromWords (adaptive.js:111-116): square roots of the two maxima,
// read and cleared for the next step (view.js:546-549).
static int read_stability(te_ctx* c, float* cfl, float* ks) {
  uint32_t w[2] = {0u, 0u};
  TE_HIP(hipMemcpyAsync(w, c->dM + c->stab_off, 8, hipMemcpyDeviceToHost, c->stream), "read stability");
  TE_HIP(hipMemsetAsync(c->dM + c->stab_off, 0, 8, c->stream), "clear stability");
  TE_HIP(hipStreamSynchronize(c->stream), "stability sync");
  float v[2];
  std::memcpy(v, w, 8);
  *cfl = (float)std::sqrt((double)v[0]);
  *ks = (float)std::sqrt((double)v[1]);
  return TE_OK;
}

static bool adaptive_ctx(const te_ctx* c) { return c && c->dM && c->has_config && c->cfg.adaptive_dt; }

int te_adaptive_start(te_ctx* c) {
  if (!adaptive_ctx(c)) return fail(c, TE_ERR_STATE, "not a context created with adaptive_dt = 1");
  int stab = -1;
  for (int i = (int)c->ops.size() - 1; i >= 0 && stab < 0; --i)
    if (std::strcmp(te::kernel_name(c->ops[i].kid), "stability_max") == 0) stab = i;
  if (stab < 0) return fail(c, TE_ERR_STATE, "the step has no stability_max dispatch");
  TE_HIP(hipMemsetAsync(c->dM + c->stab_off, 0, 8, c->stream), "clear stability");
  int rc = te_run(c, 1, stab, stab);  // encodeInitial (driver.js:46): the t = 0 maxima only
  if (rc != TE_OK) return rc;
  float cfl = 0.0f, ks = 0.0f;
  rc = read_stability(c, &cfl, &ks);
  if (rc != TE_OK) return rc;
  const float nan = std::numeric_limits<float>::quiet_NaN();
  const float freq[6] = {c->cfg.tapfrq, c->cfg.rstfrq, c->cfg.statfrq, nan, nan, nan};
  te_clock_destroy(c->clock);
  c->clock = te_clock_create(c->dtl, c->layout.min_spacing, 0.0, 0, 1, freq);
  if (!c->clock) return fail(c, TE_ERR_STATE, "clock allocation failed");
  te_clock_start(c->clock, cfl, ks);
  te_clock_state st;
  te_clock_get(c->clock, &st);
  return retime(c, st.dt, st.nsound);
}

int te_adaptive_step(te_ctx* c, int32_t nsteps) {
  if (!adaptive_ctx(c) || !c->clock) return fail(c, TE_ERR_STATE, "te_adaptive_start has not run on this context");
  if (nsteps < 0) return fail(c, TE_ERR_ARG, "nsteps must be non-negative");
  for (int s = 0; s < nsteps; ++s) {
    te_clock_begin_step(c->clock);
    int rc = grow_guards(c);
    if (rc != TE_OK) return rc;
    ensure_graph(c);
    if (c->graph) {
      TE_HIP(hipGraphLaunch(c->graph, c->stream), "hipGraphLaunch");
    } else {
      rc = run_ops(c, 0, (int)c->ops.size() - 1);
      if (rc != TE_OK) return rc;
    }
    float cfl = 0.0f, ks = 0.0f;
    rc = read_stability(c, &cfl, &ks);
    if (rc != TE_OK) return rc;
    te_clock_end_step(c->clock, cfl, ks);
    te_clock_state st;
    te_clock_get(c->clock, &st);
    rc = retime(c, st.dt, st.nsound);
    if (rc != TE_OK) return rc;
  }
  return TE_OK;
}

int te_adaptive_clock(const te_ctx* c, te_clock_state* out) {
  if (!c || !out || !c->clock) return TE_ERR_STATE;
  te_clock_get(c->clock, out);
  return TE_OK;
}

int te_init_state(te_ctx* c) {
  if (!c || !c->has_config) return fail(c, TE_ERR_STATE, "not a config context");
  try {
    te::fill_initial_state(c->cfg, c->layout, c->image);
  } catch (const std::exception& ex) {
    return fail(c, TE_ERR_ARG, ex.what());
  }
  if (!c->dM) return TE_OK;  // encoded-only context: the host image holds the state
  return te_upload_words(c, 0, c->image.data(), c->image.size());
}

int te_step(te_ctx* c, int32_t n) { return te_run(c, n, 0, -1); }

int te_field_offset(const te_ctx* c, const char* name, uint64_t* off) {
  if (!c || !name || !off || !c->has_config) return TE_ERR_ARG;
  uint64_t o = c->layout.field(name);
  if (o == UINT64_MAX) return TE_ERR_ARG;
  *off = o;
  return TE_OK;
}
int te_region_offset(const te_ctx* c, const char* name, uint64_t* off) {
  if (!c || !name || !off || !c->has_config) return TE_ERR_ARG;
  auto it = c->layout.region.find(name);
  if (it == c->layout.region.end()) return TE_ERR_ARG;
  *off = it->second;
  return TE_OK;
}
int32_t te_region_count(const te_ctx* c) { return c && c->has_config ? (int32_t)c->layout.region.size() : 0; }
const char* te_region_name(const te_ctx* c, int32_t i) {
  if (!c || !c->has_config || i < 0 || i >= (int32_t)c->layout.region.size()) return nullptr;
  std::vector<std::pair<uint64_t, const std::string*>> order;
  for (const auto& kv : c->layout.region) order.push_back({kv.second, &kv.first});
  std::sort(order.begin(), order.end(),
            [](const auto& a, const auto& b) { return a.first != b.first ? a.first < b.first : *a.second < *b.second; });
  return order[(size_t)i].second->c_str();
}
uint64_t te_field_stride_words(const te_ctx* c) { return c ? c->layout.field_words : 0; }
int32_t te_field_count(const te_ctx* c) { return c ? (int32_t)c->layout.fields.size() : 0; }
const char* te_field_name(const te_ctx* c, int32_t i) {
  return (c && i >= 0 && i < (int32_t)c->layout.fields.size()) ? c->layout.fields[i].c_str() : nullptr;
}
int te_grid(const te_ctx* c, int32_t o[5]) {
  if (!c || !o || !c->has_config) return TE_ERR_ARG;
  o[0] = c->layout.nx; o[1] = c->layout.ny; o[2] = c->layout.nz; o[3] = c->layout.sJ; o[4] = c->layout.sK;
  return TE_OK;
}

// ------------------------------------------------------------------ ensembles
struct te_ensemble {
  std::vector<te_ctx*> members;
  std::string err;
  uint64_t points = 0;
};

int te_ensemble_create(const te_config* cfgs, int32_t n, te_ensemble** out) {
  if (!cfgs || n < 1 || !out) return TE_ERR_ARG;
  te_ensemble* e = new te_ensemble();
  *out = e;
  for (int32_t m = 0; m < n; ++m) {
    te_ctx* c = nullptr;
    int rc = te_create(&cfgs[m], &c);
    if (rc == TE_OK) rc = te_init_state(c);
    if (rc == TE_OK) rc = grow_guards(c);
    if (rc == TE_OK) rc = ensure_graph(c);
    if (rc != TE_OK) {
      e->err = "member " + std::to_string(m) + ": " + (c ? c->err : std::string("create failed"));
      te_destroy(c);
      return rc;
    }
    e->members.push_back(c);
    e->points += (uint64_t)cfgs[m].nx * cfgs[m].ny * cfgs[m].nz;
  }
  return TE_OK;
}

int te_ensemble_step(te_ensemble* e, int32_t nsteps) {
  if (!e || nsteps < 0) return TE_ERR_ARG;
  // step s of every member before step s+1 of any: members advance together,
  // each on its own stream, so the device runs them concurrently
  for (int32_t s = 0; s < nsteps; ++s)
    for (te_ctx* c : e->members) {
      hipError_t r = c->graph ? hipGraphLaunch(c->graph, c->stream) : hipSuccess;
      int rc = r == hipSuccess ? (c->graph ? TE_OK : run_ops(c, 0, (int)c->ops.size() - 1)) : hip_fail(c, r, "hipGraphLaunch");
      if (rc != TE_OK) {
        e->err = c->err;
        return rc;
      }
    }
  return TE_OK;
}

int te_ensemble_sync(te_ensemble* e) {
  if (!e) return TE_ERR_ARG;
  for (te_ctx* c : e->members)
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
      e->err = "ensemble sync failed";
      return TE_ERR_HIP;
    }
  return TE_OK;
}

int32_t te_ensemble_size(const te_ensemble* e) { return e ? (int32_t)e->members.size() : 0; }
te_ctx* te_ensemble_member(te_ensemble* e, int32_t m) {
  return (e && m >= 0 && m < (int32_t)e->members.size()) ? e->members[m] : nullptr;
}
uint64_t te_ensemble_points(const te_ensemble* e) { return e ? e->points : 0; }
const char* te_ensemble_last_error(const te_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }
void te_ensemble_destroy(te_ensemble* e) {
  if (!e) return;
  for (te_ctx* c : e->members) te_destroy(c);
  delete e;
}

}  // extern "C"
