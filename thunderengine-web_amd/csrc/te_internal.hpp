// Host-side state of a context (program, device memory, stream, encoder
// layout).  Internal to the library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/te.h"

namespace te {

// One program entry: kernel id (index into the registry) and its 64-word
// argument slot (A[0..2] thread grid, A[4+n] parameter n).
struct Op {
  int32_t kid;
  uint32_t A[64];
  int32_t nz = 0;  // levels of the op's grid (set before the first launch)
  bool flat = false;  // launched flat over its words (set with nz)
  bool plane = false;  // launched over its (i, j) plane's points (set with nz)
  // launch shape chosen by te_autotune (generic 3-D kernels only)
  bool tuned = false, tplane = false;
  uint16_t tb[3] = {0, 0, 0};
  // the kernel launched in place of the op's table entry: a body specialised
  // on parameters the op fixes (set with nz; null: the table's kernel)
  const void* fn = nullptr;
};

int kernel_id(const char* name);
const char* kernel_name(int kid);

// Word layout of a program built by the encoder: named regions of M.
struct Layout {
  uint64_t words = 0;                       // total words of M
  uint64_t arena = 0, field_words = 0;      // arena base and words per field slot
  std::vector<std::string> fields;          // arena field names, sorted
  std::map<std::string, uint64_t> region;   // profile/parameter regions by name
  int nx = 0, ny = 0, nz = 0, sJ = 0, sK = 0;
  double min_spacing = 0;  // finest horizontal face spacing (m; the acoustic schedule's)
  uint64_t field(const std::string& n) const;
};

// The encoder's product: the layout, the initial memory image (profiles,
// parameter bytes, metric buffer; the arena part is zero), and the op list.
struct Encoded {
  Layout L;
  std::vector<uint32_t> image;
  // words of the arena left out of `image` (encode_step without the arena):
  // word w >= L.arena + skipped of M is image[w - skipped]
  uint64_t skipped = 0;
  std::vector<Op> ops;
  std::vector<std::string> names;  // kernel name of each op (also for kernels outside this path)
};

Encoded encode_step(const te_config& cfg, bool with_arena = true);
// Fills the arena part of `image` with the t=0 state of the scenario.
void fill_initial_state(const te_config& cfg, const Layout& L, std::vector<uint32_t>& image);

}  // namespace te

struct te_ctx;
struct te_adjoint;
namespace te {
// Launches ops [first, last] of a context in order on its stream (no events,
// no graph); TE_OK or a status with the context's error set.
int run_ops_plain(te_ctx* c, int first, int last);
// Reads every op's grid and grows the context's guard regions so that no
// launch of its program leaves the allocation; idempotent.  Afterwards
// c->front and c->words are final.
int ensure_guard(te_ctx* c);
// Called by te_destroy for each adjoint still built on the context: the
// adjoint forgets the context, and its calls that need it fail.
void adjoint_detach(te_adjoint* a);

}  // namespace te

struct te_ctx {
  uint32_t* dBase = nullptr;  // allocation (front guard + M + back guard)
  uint64_t front = 0;         // words of the front guard (dM = dBase + front)
  uint32_t* dM = nullptr;     // word 0 of the program's memory M
  uint64_t words = 0;       // allocated words (program memory + guard region)
  uint64_t prog_words = 0;  // words of the program's memory M
  bool guard_ok = false;
  uint32_t* dHash = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.0f;
  std::vector<te::Op> ops;
  std::vector<std::string> op_names;
  hipGraphExec_t graph = nullptr;
  bool graph_failed = false;
  int32_t tuned_ops = 0;  // ops whose launch shape te_autotune chose
  float tune_base_ms = 0.0f, tune_best_ms = 0.0f;  // its eager step times, untuned / tuned
  std::string err;
  // config contexts
  bool has_config = false;
  te_config cfg{};
  std::vector<float> zf;  // owned copy of cfg.zf
  std::string sounding;   // owned copy of cfg.input_sounding
  te::Layout layout;
  std::vector<uint32_t> image;
  // adjoints built on this context (te_adjoint_create); te_destroy detaches
  // them so that their later calls fail instead of using freed state
  std::vector<te_adjoint*> adjoints;
  // adapt_dt = 1: the namelist dtl, the clock (te_adaptive_start) and the
  // stability words' offset
  float dtl = 0.0f;
  te_clock* clock = nullptr;
  uint64_t stab_off = 0;
};
