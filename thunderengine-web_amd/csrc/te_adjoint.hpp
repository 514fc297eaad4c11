// Adjoint kernel registry (k_adj.hip), used by the reverse sweep (te_adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace te {
namespace adj {

// Index of the adjoint of reference kernel `name`, or -1.
int find(const char* name);
// Parameters the forward kernel writes (at most 8); returns their count.
int writes(int id, int out[8]);
// Adjoint launch of one dispatch (64-word argument slot A) on value memory M
// and adjoint memory Mb (both word 0 of the program memory); nz is the
// grid's level count and sJ, sK its strides (the dims block's words 2..4);
// scr is 4 * fw floats of scratch (fw >= one field's words) for the column
// solve's coefficients and the gather-form adjoints' output copies.
// vals: where each of the 32 parameters' values are read (null: M + offset).
hipError_t launch(int id, const uint32_t* A, const float* const* vals, const float* M, float* Mb, int nz, int sJ, int sK,
                  float* scr, uint64_t fw, hipStream_t stream);
// Adjoint of bc_periodic_x followed by bc_periodic_y on each of the nf
// (at most kHaloFields) fields at words fields.off[0..nf) of the adjoint
// memory, in one launch.
constexpr int kHaloFields = 8;
struct HaloArg {
  uint64_t off[kHaloFields];
};
hipError_t halo_xy(float* Mb, const HaloArg& fields, int nf, int nx, int ny, int nz, int sJ, int sK,
                   hipStream_t stream);
// Copies of up to kCopyRegions regions of n words each, in one launch.
constexpr int kCopyRegions = 16;
struct CopyArg {
  const uint32_t* src[kCopyRegions];
  uint32_t* dst[kCopyRegions];
  uint64_t n;
};
hipError_t copy_regions(const CopyArg& c, int nreg, hipStream_t stream);

}  // namespace adj
}  // namespace te
