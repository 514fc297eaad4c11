// Adjoint of the large step: checkpointed forward, recomputing reverse sweep.
//
// Memory (one MI355X holds all of it for the grids of this path):
//   * checkpoints: the state words [0, arena end) before every `every`-th step;
//   * the write log: every region each dispatch writes, saved before the
//     dispatch (a whole FieldArena slot per written field), for a whole
//     checkpoint segment when that fits (each step is then recomputed once),
//     else for one step, with the segment's step-start states recomputed
//     first;
//   * the adjoint memory: one float per word of M, with M's guards.
// The reverse of a step runs the dispatches' adjoints from the last, each
// reading its inputs where they are at that point: in the write log slot of
// the next dispatch that overwrote them, or in M.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/te.h"
#include "te_adjoint.hpp"
#include "te_internal.hpp"

struct te_adjoint {
  te_ctx* fwd = nullptr;  // the caller's context
  te_ctx* rec = nullptr;  // unfused dispatch list on memory of its own (owned)
  int every = 1;
  uint64_t state_words = 0;  // [0, arena end)
  float* dMbBase = nullptr;
  float* dMb = nullptr;
  float* dScr = nullptr;  // 4 planes of one field: the column solve's coefficients
  std::vector<uint32_t*> ckpt;
  int nsteps = 0;
  uint32_t* dSteps = nullptr;  // `every` step-start states
  uint32_t* dLog = nullptr;
  uint64_t log_words = 0;     // one step's write log
  bool segment_log = false;   // the log holds a whole checkpoint segment
  std::vector<int> adj_id;                                        // per dispatch
  std::vector<std::vector<std::pair<uint64_t, uint64_t>>> wreg;  // (offset, words) per dispatch
  std::vector<uint64_t> stride;  // log words per region of each dispatch (its length rounded up to 64)
  std::vector<char> halo_pair;    // dispatch i is bc_periodic_y right after bc_periodic_x of the same field
  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
  float fwd_ms = 0.0f, bwd_ms = 0.0f;
  std::string err;
};

static int afail(te_adjoint* a, int code, const std::string& msg) {
  if (a) a->err = msg;
  return code;
}
#define TA_HIP(call, what)                                                                     \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return afail(a, TE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

static int run_rec_ops(te_adjoint* a, int first, int last) {
  int rc = te_run(a->rec, 1, first, last);
  if (rc != TE_OK) return afail(a, rc, std::string("forward dispatches: ") + te_last_error(a->rec));
  return TE_OK;
}

// Logged forward of `steps` steps on rec's memory (the write log of each
// dispatch appended at log position `pos`), then their reverse.
static int reverse_steps(te_adjoint* a, int steps) {
  te_ctx* r = a->rec;
  const int n = (int)r->ops.size();
  std::vector<uint64_t> at((size_t)n * steps, 0);
  // save (restore) the regions of dispatch i at log position p: one launch
  auto copy = [&](int i, uint64_t p, bool save) -> hipError_t {
    const auto& w = a->wreg[i];
    te::adj::CopyArg c{};
    c.n = w.empty() ? 0 : w[0].second;
    for (size_t q = 0; q < w.size(); ++q) {
      uint32_t* m = r->dM + w[q].first;
      uint32_t* l = a->dLog + p + q * a->stride[i];
      c.src[q] = save ? m : l;
      c.dst[q] = save ? l : m;
    }
    return te::adj::copy_regions(c, (int)w.size(), r->stream);
  };
  uint64_t pos = 0;
  for (int s = 0; s < steps; ++s)
    for (int i = 0; i < n; ++i) {
      at[(size_t)s * n + i] = pos;
      TA_HIP(copy(i, pos, true), "log");
      pos += a->wreg[i].size() * a->stride[i];
      int rc = run_rec_ops(a, i, i);
      if (rc != TE_OK) return rc;
    }
  // reverse: a dispatch's inputs are the values before it.  Walking back,
  // each field's value before the dispatch is in the log slot of the first
  // write at or after it, or still in M when the segment never wrote it after
  // the dispatch; the adjoint reads its parameters there (no restores).
  std::map<uint64_t, const float*> cur;  // field offset -> its values before the dispatch being reversed
  const float* M = reinterpret_cast<const float*>(r->dM);
  const float* vals[32];
  for (int s = steps - 1; s >= 0; --s)
    for (int i = n - 1; i >= 0; --i) {
      const uint64_t p = at[(size_t)s * n + i];
      for (size_t q = 0; q < a->wreg[i].size(); ++q)
        cur[a->wreg[i][q].first] = reinterpret_cast<const float*>(a->dLog + p + q * a->stride[i]);
      const uint32_t* A = r->ops[i].A;
      if (a->halo_pair[i]) {  // with the bc_periodic_x before it, as one map
        TA_HIP(te::adj::halo_xy(a->dMb, A[4 + 1], r->layout.nx, r->layout.ny, r->layout.nz, r->layout.sJ,
                                r->layout.sK, r->stream),
               "adjoint of the periodic halos");
        continue;
      }
      if (i + 1 < n && a->halo_pair[i + 1]) continue;
      for (int q = 0; q < 32; ++q) {
        auto it = cur.find(A[4 + q]);
        vals[q] = it != cur.end() ? it->second : M + A[4 + q];
      }
      TA_HIP(te::adj::launch(a->adj_id[i], A, vals, M, a->dMb, r->layout.nz, r->layout.sJ, r->layout.sK, a->dScr,
                             r->layout.field_words, r->stream),
             std::string("adjoint of ") + r->op_names[i]);
    }
  return TE_OK;
}

extern "C" {

int te_adjoint_create(te_ctx* ctx, int32_t every, te_adjoint** out) {
  if (!out) return TE_ERR_ARG;
  *out = nullptr;
  te_adjoint* a = new te_adjoint();
  *out = a;
  if (!ctx || !ctx->has_config || !ctx->dM) return afail(a, TE_ERR_ARG, "needs a config context with device memory");
  if (every < 1) return afail(a, TE_ERR_ARG, "checkpoint interval must be at least 1");
  a->fwd = ctx;
  a->every = every;
  te_config cfg = ctx->cfg;
  cfg.fuse = 0;
  int rc = te_create(&cfg, &a->rec);
  if (rc != TE_OK) return afail(a, rc, std::string("unfused program: ") + (a->rec ? te_last_error(a->rec) : ""));
  te_ctx* r = a->rec;
  const te::Layout &L = r->layout, &F = ctx->layout;
  if (L.arena != F.arena || L.field_words != F.field_words || L.fields != F.fields)
    return afail(a, TE_ERR_STATE, "fused and unfused layouts differ");
  a->state_words = L.arena + L.field_words * L.fields.size();
  const uint64_t arena_end = a->state_words;
  // per dispatch: its adjoint and the regions it writes
  const int n = (int)r->ops.size();
  a->adj_id.resize(n);
  a->wreg.resize(n);
  for (int i = 0; i < n; ++i) {
    const std::string& nm = r->op_names[i];
    const int id = te::adj::find(nm.c_str());
    if (id < 0) return afail(a, TE_ERR_UNSUPPORTED, "no adjoint for kernel '" + nm + "'");
    a->adj_id[i] = id;
    const uint32_t* A = r->ops[i].A;
    if (nm == "rayleigh_damping" && (r->image[A[4 + 17]] != 0u || r->image[A[4 + 16]] != 0u))
      return afail(a, TE_ERR_UNSUPPORTED, "no adjoint for Rayleigh damping toward horizontal means or at lateral edges");
    if (nm == "_copy") {
      a->wreg[i].push_back({A[5], A[0]});
      continue;
    }
    int wp[8];
    const int nw = te::adj::writes(id, wp);
    for (int q = 0; q < nw; ++q) {
      const uint64_t off = A[4 + wp[q]];
      if (nm == "acoustic_pp" && wp[q] == 9 && r->image[A[4 + 10]] == 0u) continue;  // ppx unused
      if (off < L.arena || off >= arena_end || (off - L.arena) % L.field_words != 0)
        return afail(a, TE_ERR_UNSUPPORTED, "dispatch " + std::to_string(i) + " (" + nm + ") writes outside the FieldArena");
      bool dup = false;
      for (auto& w : a->wreg[i]) dup = dup || w.first == off;
      if (!dup) a->wreg[i].push_back({off, L.field_words});
    }
  }
  a->halo_pair.assign(n, 0);
  for (int i = 1; i < n; ++i)
    if (r->op_names[i] == "bc_periodic_y" && r->op_names[i - 1] == "bc_periodic_x" &&
        r->ops[i].A[4 + 1] == r->ops[i - 1].A[4 + 1] && r->ops[i].A[4] == r->ops[i - 1].A[4])
      a->halo_pair[i] = 1;
  a->stride.resize(n);
  for (int i = 0; i < n; ++i) {
    a->stride[i] = a->wreg[i].empty() ? 0 : (a->wreg[i][0].second + 63) / 64 * 64;
    a->log_words += a->stride[i] * a->wreg[i].size();
  }
  hipStream_t s = r->stream;
  const uint64_t mbw = r->front + r->words;
  TA_HIP(hipMalloc((void**)&a->dMbBase, mbw * 4), "hipMalloc(adjoint memory)");
  TA_HIP(hipMemsetAsync(a->dMbBase, 0, mbw * 4, s), "clear adjoint memory");
  a->dMb = a->dMbBase + r->front;
  TA_HIP(hipMalloc((void**)&a->dScr, 4 * L.field_words * 4), "hipMalloc(column scratch)");
  // a whole segment's write log when it fits the budget (one recompute per
  // step); otherwise one step's log and the segment's step-start states
  // (budget: TE_ADJOINT_LOG_GIB, default 96 GiB of the 288 GiB HBM)
  uint64_t budget = 96ull << 30;
  if (const char* env = std::getenv("TE_ADJOINT_LOG_GIB")) budget = (uint64_t)std::strtoull(env, nullptr, 10) << 30;
  if (every > 1 && (uint64_t)every * a->log_words * 4 <= budget &&
      hipMalloc((void**)&a->dLog, (uint64_t)every * a->log_words * 4) == hipSuccess) {
    a->segment_log = true;
  } else {
    (void)hipGetLastError();
    TA_HIP(hipMalloc((void**)&a->dLog, std::max<uint64_t>(a->log_words, 1) * 4), "hipMalloc(write log)");
    TA_HIP(hipMalloc((void**)&a->dSteps, (uint64_t)every * a->state_words * 4), "hipMalloc(step states)");
  }
  TA_HIP(hipEventCreate(&a->e0), "event");
  TA_HIP(hipEventCreate(&a->e1), "event");
  TA_HIP(hipEventCreate(&a->e2), "event");
  TA_HIP(hipEventCreate(&a->e3), "event");
  TA_HIP(hipStreamSynchronize(s), "sync");
  return TE_OK;
}

te_ctx* te_adjoint_program(te_adjoint* a) { return a ? a->rec : nullptr; }

int te_fwd_ckpt(te_adjoint* a, int32_t nsteps) {
  if (!a || !a->rec || nsteps < 0) return afail(a, TE_ERR_ARG, "bad arguments");
  te_ctx* f = a->fwd;
  const int nseg = (nsteps + a->every - 1) / a->every;
  while ((int)a->ckpt.size() < nseg) {
    uint32_t* p = nullptr;
    TA_HIP(hipMalloc((void**)&p, a->state_words * 4), "hipMalloc(checkpoint)");
    a->ckpt.push_back(p);
  }
  TA_HIP(hipEventRecord(a->e0, f->stream), "event");
  for (int sg = 0; sg < nseg; ++sg) {
    TA_HIP(hipMemcpyAsync(a->ckpt[sg], f->dM, a->state_words * 4, hipMemcpyDeviceToDevice, f->stream), "checkpoint");
    const int k = std::min(a->every, nsteps - sg * a->every);
    int rc = te_run(f, k, 0, -1);
    if (rc != TE_OK) return afail(a, rc, std::string("forward: ") + te_last_error(f));
  }
  TA_HIP(hipEventRecord(a->e1, f->stream), "event");
  TA_HIP(hipEventSynchronize(a->e1), "sync");
  TA_HIP(hipEventElapsedTime(&a->fwd_ms, a->e0, a->e1), "elapsed");
  a->nsteps = nsteps;
  return TE_OK;
}

int te_bwd(te_adjoint* a) {
  if (!a || !a->rec) return afail(a, TE_ERR_ARG, "bad arguments");
  if (a->nsteps <= 0) return afail(a, TE_ERR_STATE, "no recorded steps (te_fwd_ckpt)");
  te_ctx* r = a->rec;
  TA_HIP(hipStreamSynchronize(a->fwd->stream), "sync forward");
  TA_HIP(hipEventRecord(a->e2, r->stream), "event");
  const int nseg = (a->nsteps + a->every - 1) / a->every;
  const uint64_t sw = a->state_words;
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int s0 = sg * a->every, s1 = std::min(a->nsteps, s0 + a->every);
    TA_HIP(hipMemcpyAsync(r->dM, a->ckpt[sg], sw * 4, hipMemcpyDeviceToDevice, r->stream), "load checkpoint");
    if (a->segment_log) {
      int rc = reverse_steps(a, s1 - s0);
      if (rc != TE_OK) return rc;
      continue;
    }
    for (int s = s0; s < s1; ++s) {
      TA_HIP(hipMemcpyAsync(a->dSteps + (uint64_t)(s - s0) * sw, r->dM, sw * 4, hipMemcpyDeviceToDevice, r->stream),
             "step state");
      if (s < s1 - 1) {
        int rc = run_rec_ops(a, 0, (int)r->ops.size() - 1);
        if (rc != TE_OK) return rc;
      }
    }
    for (int s = s1 - 1; s >= s0; --s) {
      TA_HIP(hipMemcpyAsync(r->dM, a->dSteps + (uint64_t)(s - s0) * sw, sw * 4, hipMemcpyDeviceToDevice, r->stream),
             "load step state");
      int rc = reverse_steps(a, 1);
      if (rc != TE_OK) return rc;
    }
  }
  TA_HIP(hipEventRecord(a->e3, r->stream), "event");
  TA_HIP(hipEventSynchronize(a->e3), "sync");
  TA_HIP(hipEventElapsedTime(&a->bwd_ms, a->e2, a->e3), "elapsed");
  return TE_OK;
}

int te_adj_ops(te_adjoint* a, int32_t first, int32_t last) {
  if (!a || !a->rec) return afail(a, TE_ERR_ARG, "bad arguments");
  te_ctx* r = a->rec;
  const int n = (int)r->ops.size();
  if (first < 0 || last >= n || first > last) return afail(a, TE_ERR_ARG, "bad dispatch range");
  for (int i = last; i >= first; --i)
    TA_HIP(te::adj::launch(a->adj_id[i], r->ops[i].A, nullptr, reinterpret_cast<const float*>(r->dM), a->dMb,
                           r->layout.nz, r->layout.sJ, r->layout.sK,
                           a->dScr, r->layout.field_words, r->stream),
           std::string("adjoint of ") + r->op_names[i]);
  TA_HIP(hipStreamSynchronize(r->stream), "sync");
  return TE_OK;
}

int te_adj_upload(te_adjoint* a, uint64_t off, const float* src, uint64_t n) {
  if (!a || !a->rec || !src || off + n > a->rec->prog_words) return afail(a, TE_ERR_ARG, "upload outside memory");
  TA_HIP(hipMemcpyAsync(a->dMb + off, src, n * 4, hipMemcpyHostToDevice, a->rec->stream), "upload");
  TA_HIP(hipStreamSynchronize(a->rec->stream), "sync");
  return TE_OK;
}

int te_adj_download(te_adjoint* a, uint64_t off, float* dst, uint64_t n) {
  if (!a || !a->rec || !dst || off + n > a->rec->prog_words) return afail(a, TE_ERR_ARG, "download outside memory");
  TA_HIP(hipMemcpyAsync(dst, a->dMb + off, n * 4, hipMemcpyDeviceToHost, a->rec->stream), "download");
  TA_HIP(hipStreamSynchronize(a->rec->stream), "sync");
  return TE_OK;
}

int te_adj_zero(te_adjoint* a) {
  if (!a || !a->rec) return afail(a, TE_ERR_ARG, "bad arguments");
  const uint64_t mbw = a->rec->front + a->rec->words;
  TA_HIP(hipMemsetAsync(a->dMbBase, 0, mbw * 4, a->rec->stream), "clear");
  TA_HIP(hipStreamSynchronize(a->rec->stream), "sync");
  return TE_OK;
}

int te_adj_last_ms(te_adjoint* a, float* fwd_ms, float* bwd_ms) {
  if (!a) return TE_ERR_ARG;
  if (fwd_ms) *fwd_ms = a->fwd_ms;
  if (bwd_ms) *bwd_ms = a->bwd_ms;
  return TE_OK;
}

const char* te_adjoint_last_error(const te_adjoint* a) { return a ? a->err.c_str() : "null adjoint"; }

void te_adjoint_destroy(te_adjoint* a) {
  if (!a) return;
  if (a->rec && a->rec->stream) (void)hipStreamSynchronize(a->rec->stream);
  for (uint32_t* p : a->ckpt) (void)hipFree(p);
  if (a->dMbBase) (void)hipFree(a->dMbBase);
  if (a->dScr) (void)hipFree(a->dScr);
  if (a->dLog) (void)hipFree(a->dLog);
  if (a->dSteps) (void)hipFree(a->dSteps);
  for (hipEvent_t e : {a->e0, a->e1, a->e2, a->e3})
    if (e) (void)hipEventDestroy(e);
  if (a->rec) te_destroy(a->rec);
  delete a;
}

}  // extern "C"
