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

}  // namespace te
