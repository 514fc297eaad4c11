// Host build of the library's adjoint bodies — TEST INFRASTRUCTURE ONLY.
//
// Compiles thunderengine-web_amd/csrc/adj_bodies.hpp (the same per-point and
// per-column adjoint bodies the gfx950 kernels of k_adj.hip run) for the CPU
// and runs one dispatch's adjoint serially over its thread grid, so the CPU
// suite can check every adjoint against the oracle's FP64 tangent-linear
// build without a GPU.  The GPU tests check the device launches the same way.
#include <cstring>
#include <vector>
#include <xmmintrin.h>

#include "../thunderengine-web_amd/csrc/adj_bodies.hpp"

using namespace te::adj;

namespace {

template <class B>
void grid(const float* M, float* Mb, const AdjArg& a) {
  for (uint32_t z = 0; z < a.n[2]; ++z)
    for (uint32_t y = 0; y < a.n[1]; ++y)
      for (uint32_t x = 0; x < a.n[0]; ++x) PointAdj<B>::run(M, Mb, a, (int)x, (int)y, (int)z);
}

void columns(const float* M, float* Mb, const AdjArg& a) {
  Acc r(M, Mb, &a);
  const ADims g = r.dims();
  const uint64_t fw = (uint64_t)g.sJ * g.sK * (g.nz + 1);
  std::vector<float> scr(4 * fw, 0.0f);
  for (int z = 0; z < g.nz; ++z)
    for (uint32_t y = 0; y < a.n[1]; ++y)
      for (uint32_t x = 0; x < a.n[0]; ++x) w_solve_coef(M, a, scr.data(), fw, (int)x, (int)y, z);
  std::vector<float> tA(g.nz + 1), tB(g.nz + 1);
  for (uint32_t y = 0; y < a.n[1]; ++y)
    for (uint32_t x = 0; x < a.n[0]; ++x) w_solve_sweep(M, Mb, a, scr.data(), fw, tA.data(), tB.data(), 1, (int)x, (int)y);
  for (int z = 0; z < g.nz; ++z)
    for (uint32_t y = 0; y < a.n[1]; ++y)
      for (uint32_t x = 0; x < a.n[0]; ++x) w_solve_level(M, Mb, a, (int)x, (int)y, z);
  for (uint32_t y = 0; y < a.n[1]; ++y)
    for (uint32_t x = 0; x < a.n[0]; ++x) w_solve_clear(Mb, a, g, (int)x, (int)y);
}

template <class G>
void gathered(const float* M, float* Mb, const AdjArg& a) {
  Acc r(M, Mb, &a);
  const ADims g = r.dims();
  const uint64_t fw = (uint64_t)g.sJ * g.sK * (g.nz + 1);
  std::vector<float> scr(fw, 0.0f);
  if (G::kPrep)
    for (uint32_t z = 0; z < a.n[2]; ++z)
      for (uint32_t y = 0; y < a.n[1]; ++y)
        for (uint32_t x = 0; x < a.n[0]; ++x) G::prep(M, Mb, a, scr.data(), (int)x, (int)y, (int)z);
  const int64_t lead = G::lead(g), words = (int64_t)fw;
  for (int64_t q = -lead; q < words; ++q) G::gather(M, Mb, a, scr.data(), q);
}

struct HostEntry {
  const char* name;
  void (*fn)(const float*, float*, const AdjArg&);
  int nw;
  int writes[8];
};

#define HK(n, B, nw, ...) {n, &grid<B>, nw, __VA_ARGS__},
const HostEntry kHost[] = {
    TE_ADJ_POINT_KERNELS(HK)
    {"acoustic_w_solve", &columns, 3, {9, 10, 11}},
};
#undef HK

}  // namespace

extern "C" {

// Adjoint of one dispatch (kernel `name`, 64-word slot A) on value memory M
// and adjoint memory Mb (word 0 of the program memory), in gather form where
// the library has one unless scatter_only.  Returns -1 for a kernel without
// an adjoint.
int adjhost_op(const char* name, const uint32_t* A, const float* M, float* Mb, int scatter_only) {
  _mm_setcsr(_mm_getcsr() | 0x8040u);  // flush subnormals, as the GPU build
#define HG(kname, G) \
  if (std::strcmp(kname, name) == 0 && !scatter_only) {                      \
    AdjArg a;                                                                \
    a.n[0] = A[0]; a.n[1] = A[1]; a.n[2] = A[2];                             \
    for (int q = 0; q < 32; ++q) { a.p[q] = A[4 + q]; a.v[q] = M + a.p[q]; }                          \
    gathered<G>(M, Mb, a);                                                   \
    return 0;                                                                \
  }
  TE_ADJ_GATHER_KERNELS(HG)
#undef HG
  for (const HostEntry& e : kHost)
    if (std::strcmp(e.name, name) == 0) {
      AdjArg a;
      a.n[0] = A[0]; a.n[1] = A[1]; a.n[2] = A[2];
      for (int q = 0; q < 32; ++q) { a.p[q] = A[4 + q]; a.v[q] = M + a.p[q]; }
      if (std::strcmp(name, "_copy") == 0) a.n[1] = a.n[2] = 1;
      e.fn(M, Mb, a);
      return 0;
    }
  return -1;
}

// Parameters kernel `name` writes (at most 8), or -1.
int adjhost_writes(const char* name, int* out) {
  for (const HostEntry& e : kHost)
    if (std::strcmp(e.name, name) == 0) {
      for (int q = 0; q < e.nw; ++q) out[q] = e.writes[q];
      return e.nw;
    }
  return -1;
}

int adjhost_count() { return (int)(sizeof(kHost) / sizeof(kHost[0])); }
const char* adjhost_name(int i) { return kHost[i].name; }
}
