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
// grid's level count (the dims block's third word).
hipError_t launch(int id, const uint32_t* A, const float* M, float* Mb, int nz, hipStream_t stream);

}  // namespace adj
}  // namespace te
