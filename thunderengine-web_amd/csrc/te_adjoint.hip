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

// One reference dispatch of a program's dispatch: the adjoint of a fused
// launch is the adjoints of the dispatches it stands for, in reverse order,
// all reading the values from before the launch (none of them reads a value
// another of them writes, other than linearly).
struct Unit {
  int adj = -1;         // adjoint registry index, or -1: periodic halos of `field` (x then y)
  uint32_t A[64] = {};  // the reference dispatch's argument slot
  uint64_t field = 0;
};

struct te_adjoint {
  te_ctx* fwd = nullptr;  // the caller's context
  te_ctx* rec = nullptr;  // the same configuration's program on memory of its own (owned)
  int every = 1;
  uint64_t state_words = 0;  // [0, arena end)
  float* dMbBase = nullptr;
  float* dMb = nullptr;
  uint64_t mb_words = 0;  // words of the adjoint allocation (rec's front guard + memory + back guard)
  float* dScr = nullptr;  // 4 planes of one field: the column solve's coefficients
  std::vector<uint32_t*> ckpt;
  int nsteps = 0;
  uint32_t* dSteps = nullptr;  // `every` step-start states
  uint32_t* dLog = nullptr;
  uint64_t log_words = 0;     // one step's write log
  bool segment_log = false;   // the log holds a whole checkpoint segment
  std::vector<std::vector<Unit>> units;                          // per dispatch, in forward order
  std::vector<std::vector<std::pair<uint64_t, uint64_t>>> wreg;  // (offset, words) per dispatch
  std::vector<uint64_t> stride;  // log words per region of each dispatch (its length rounded up to 64)
  std::vector<char> halo_pair;    // dispatch i is bc_periodic_y right after bc_periodic_x of the same field
  std::vector<std::vector<char>> live;  // per dispatch and region: some adjoint reads its saved values
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
  int rc = te::run_ops_plain(a->rec, first, last);
  if (rc != TE_OK) return afail(a, rc, std::string("forward dispatches: ") + te_last_error(a->rec));
  return TE_OK;
}

// Adjoint of the periodic halos (x then y) of the fields of a run of halo
// units, up to kHaloFields fields per launch (a field listed twice is done
// once: the second forward fill left the halos as the first did).
static int halo_units(te_adjoint* a, const Unit* us, int nu) {
  te_ctx* r = a->rec;
  te::adj::HaloArg h{};
  int nf = 0;
  for (int u = 0; u <= nu; ++u) {
    if (u == nu || nf == te::adj::kHaloFields) {
      TA_HIP(te::adj::halo_xy(a->dMb, h, nf, r->layout.nx, r->layout.ny, r->layout.nz, r->layout.sJ, r->layout.sK,
                              r->stream),
             "adjoint of the periodic halos");
      nf = 0;
      if (u == nu) break;
    }
    bool dup = false;
    for (int t = 0; t < nf; ++t) dup = dup || h.off[t] == us[u].field;
    if (!dup) h.off[nf++] = us[u].field;
  }
  return TE_OK;
}

// Logged forward of `steps` steps on rec's memory (the write log of each
// dispatch appended at log position `pos`), then their reverse.
static int reverse_steps(te_adjoint* a, int steps) {
  te_ctx* r = a->rec;
  const int n = (int)r->ops.size();
  std::vector<uint64_t> at((size_t)n * steps, 0);
  // Logged forward.  The regions a dispatch writes are saved to its log slot
  // before it runs; a run of dispatches none of which writes a region another
  // of them writes saves all their regions in one launch before the run (a
  // region's values before the run are its values before its own writer).
  uint64_t pos = 0;
  for (int s = 0; s < steps; ++s)
    for (int i = 0; i < n;) {
      te::adj::CopyArg c{};
      int nreg = 0, e = i;
      std::vector<uint64_t> wr;  // every region the run writes (saved or not)
      for (; e < n; ++e) {
        const auto& w = a->wreg[e];
        int nlive = 0;
        for (size_t q = 0; q < w.size(); ++q) nlive += a->live[e][q];
        bool fits = nreg + nlive <= te::adj::kCopyRegions && (nlive == 0 || nreg == 0 || w[0].second == c.n);
        for (size_t q = 0; q < w.size() && fits; ++q)
          for (uint64_t o : wr) fits = fits && o != w[q].first;
        if (!fits) break;
        at[(size_t)s * n + e] = pos;
        for (size_t q = 0; q < w.size(); ++q) {
          wr.push_back(w[q].first);
          if (!a->live[e][q]) continue;  // no adjoint reads these values: not saved
          c.src[nreg] = r->dM + w[q].first;
          c.dst[nreg] = a->dLog + pos + q * a->stride[e];
          c.n = w[q].second;
          ++nreg;
        }
        pos += w.size() * a->stride[e];
      }
      TA_HIP(te::adj::copy_regions(c, nreg, r->stream), "log");
      int rc = run_rec_ops(a, i, e - 1);
      if (rc != TE_OK) return rc;
      i = e;
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
        if (a->live[i][q]) cur[a->wreg[i][q].first] = reinterpret_cast<const float*>(a->dLog + p + q * a->stride[i]);
      const uint32_t* A = r->ops[i].A;
      if (a->halo_pair[i]) {  // with the bc_periodic_x before it, as one map
        Unit h;
        h.field = A[4 + 1];
        int rc = halo_units(a, &h, 1);
        if (rc != TE_OK) return rc;
        continue;
      }
      if (i + 1 < n && a->halo_pair[i + 1]) continue;
      const std::vector<Unit>& us = a->units[i];
      if (!us.empty() && us[0].adj < 0) {  // bc_periodic_xy: halo units only, fields independent
        int rc = halo_units(a, us.data(), (int)us.size());
        if (rc != TE_OK) return rc;
        continue;
      }
      for (int u = (int)us.size() - 1; u >= 0; --u) {
        for (int q = 0; q < 32; ++q) {
          auto it = cur.find(us[u].A[4 + q]);
          vals[q] = it != cur.end() ? it->second : M + us[u].A[4 + q];
        }
        TA_HIP(te::adj::launch(us[u].adj, us[u].A, vals, M, a->dMb, r->layout.nz, r->layout.sJ, r->layout.sK, a->dScr,
                               r->layout.field_words, r->stream),
               std::string("adjoint of ") + r->op_names[i]);
      }
    }
  return TE_OK;
}

// The reference dispatches a program dispatch stands for, with the
// parameters it writes.  Fused launches follow their definitions in
// k_fused.hip / k_adv.hip (the oracle restates the same expansion).
static int decompose(te_adjoint* a, int i, std::vector<Unit>& units, std::vector<uint64_t>& writes) {
  te_ctx* r = a->rec;
  const std::string& nm = r->op_names[i];
  const uint32_t* A = r->ops[i].A;
  const std::vector<uint32_t>& W = r->image;
  auto unit = [&](const char* kname, const uint32_t* slot) -> int {
    Unit u;
    u.adj = te::adj::find(kname);
    if (u.adj < 0) return afail(a, TE_ERR_UNSUPPORTED, std::string("no adjoint for kernel '") + kname + "'");
    std::memcpy(u.A, slot, sizeof(u.A));
    units.push_back(u);
    return TE_OK;
  };
  auto slot = [&](uint32_t g0, uint32_t g1, uint32_t g2, std::initializer_list<std::pair<int, uint32_t>> ps,
                  uint32_t* X) {
    std::memset(X, 0, 64 * 4);
    X[0] = g0; X[1] = g1; X[2] = g2;
    X[4] = A[4];     // dims
    X[34] = A[34];   // metric
    for (auto& pr : ps) X[4 + pr.first] = pr.second;
  };
  const uint32_t nx = W[A[4]], ny = W[A[4] + 1], nz = W[A[4] + 2];
  uint32_t X[64];
  if (nm == "bc_periodic_xy") {
    const uint32_t nf = std::min<uint32_t>(W[A[4 + 9]], 8);
    for (uint32_t q = 0; q < nf; ++q) {
      Unit u;
      u.field = A[5 + q];
      units.push_back(u);
      writes.push_back(A[5 + q]);
    }
    return TE_OK;
  }
  if (nm == "adv_fused") {
    const uint32_t kind = W[A[4 + 3]], chain = W[A[4 + 32]];
    const uint32_t S = A[4 + 2], RU = A[4 + 13], RV = A[4 + 14], RW = A[4 + 4], C1 = A[4 + 15], C2 = A[4 + 16],
                   ST = A[4 + 1], WE = A[4 + 5], PD = A[4 + 20], PW = A[4 + 21], FX = A[4 + 22], FY = A[4 + 23],
                   FZ = A[4 + 24], TEND = A[4 + 11];
    if (chain & 16u) {
      slot(nx + 7, ny + 7, nz + 1, {{1, TEND}}, X);
      if (unit("turb_zero", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    }
    if (kind == 0) {
      slot(nx + 1, ny, nz, {{1, S}, {2, RU}, {3, FX}}, X); if (unit("adv_flux_x", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      slot(nx, ny + 1, nz, {{1, S}, {2, RV}, {3, FY}}, X); if (unit("adv_flux_y", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      slot(nx, ny, nz + 1, {{1, S}, {2, RW}, {3, C1}, {4, C2}, {5, FZ}}, X);
      if (unit("adv_flux_z", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    } else if (kind == 1) {
      slot(nx + 1, ny, nz, {{1, S}, {2, RU}, {3, WE}, {4, PD}, {5, FX}}, X);
      if (unit("adv_flux_x_weno", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      slot(nx, ny + 1, nz, {{1, S}, {2, RV}, {3, WE}, {4, PD}, {5, FY}}, X);
      if (unit("adv_flux_y_weno", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      slot(nx, ny, nz + 1, {{1, S}, {2, RW}, {3, C1}, {4, C2}, {5, WE}, {6, PD}, {7, PW}, {8, FZ}}, X);
      if (unit("adv_flux_z_weno", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    } else if (kind == 2) {
      slot(nx + 2, ny + 2, nz, {{1, ST}, {2, S}, {3, RU}, {4, RV}, {5, RW}, {6, C1}, {7, C2}, {8, FX}}, X);
      if (unit("adv_flux_x_stag", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      slot(nx + 2, ny + 2, nz, {{1, ST}, {2, S}, {3, RU}, {4, RV}, {5, RW}, {6, C1}, {7, C2}, {8, FY}}, X);
      if (unit("adv_flux_y_stag", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      slot(nx + 1, ny + 1, nz + 1, {{1, ST}, {2, S}, {3, RU}, {4, RV}, {5, RW}, {6, C1}, {7, C2}, {8, FZ}}, X);
      if (unit("adv_flux_z_stag", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    } else {
      slot(nx + 2, ny + 2, nz, {{1, S}, {2, RU}, {3, C2}, {4, WE}, {5, FX}}, X);
      if (unit("adv_flux_x_weno_w", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      slot(nx + 2, ny + 2, nz, {{1, S}, {2, RV}, {3, C2}, {4, WE}, {5, FY}}, X);
      if (unit("adv_flux_y_weno_w", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      slot(nx + 1, ny + 1, nz + 1, {{1, S}, {2, RW}, {3, WE}, {4, FZ}}, X);
      if (unit("adv_flux_z_weno_w", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    }
    std::memcpy(X, A, sizeof(X));
    X[4 + 3] = FX; X[4 + 4] = FY; X[4 + 5] = FZ;
    for (int q = 20; q < 32; ++q) X[4 + q] = A[4 + q] * 0u;  // only parameters 0..19 and 30 belong to adv_assemble
    X[4 + 30] = A[4 + 30];
    if (unit("adv_assemble", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    const uint32_t full[3] = {nx, ny, nz}, whole[3] = {nx + 7, ny + 7, nz + 1};
    if (chain & 1u) {
      slot(full[0], full[1], full[2], {{1, A[4 + 33]}, {2, A[4 + 34]}, {3, TEND}}, X);
      if (unit("apply_thterm", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    }
    if (chain & 2u) {
      slot(whole[0], whole[1], whole[2], {{1, A[4 + 35]}, {2, TEND}, {3, A[4 + 36]}}, X);
      if (unit("turb_add_tendency", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    }
    if (chain & 4u) {
      slot(whole[0], whole[1], whole[2], {{1, A[4 + 37]}, {2, TEND}, {3, A[4 + 38]}}, X);
      if (unit("turb_add_tendency", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    }
    writes.push_back(TEND);
    if (chain & 8u) {
      slot(A[0], A[1], A[2], {{1, A[4 + 1]}, {2, A[4 + 39]}, {3, TEND}, {4, A[4 + 40]}, {5, A[4 + 41]}}, X);
      if (unit("rk_update", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      writes.push_back(A[4 + 40]);
    }
    return TE_OK;
  }
  if (nm == "ac_pp_accumulate") {
    std::memset(X, 0, sizeof(X));
    X[0] = nx; X[1] = ny; X[2] = nz;
    X[4] = A[4]; X[5] = A[5]; X[34] = A[34];
    for (int q = 2; q <= 10; ++q) X[4 + q] = A[4 + q];  // w pk1 pk2 qk fpk pp3d ppd ppx saveX
    if (unit("acoustic_pp", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    std::memset(X, 0, sizeof(X));
    X[0] = A[0]; X[1] = A[1]; X[2] = A[2];
    X[4] = A[4]; X[5] = A[5]; X[34] = A[34];
    const int as[7] = {11, 12, 2, 13, 14, 15, 16};  // u v w rru rrv rrw tavg
    for (int q = 0; q < 7; ++q) X[4 + 2 + q] = A[4 + as[q]];
    if (unit("acoustic_accumulate", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    writes.insert(writes.end(), {A[4 + 7], A[4 + 8], A[4 + 13], A[4 + 14], A[4 + 15]});
    if (W[A[4 + 10]] != 0u) writes.push_back(A[4 + 9]);
    return TE_OK;
  }
  if (nm == "tendency_reset_add") {
    std::memset(X, 0, sizeof(X));
    X[0] = A[0]; X[1] = A[1]; X[2] = A[2];
    X[4] = A[4]; X[5] = A[5]; X[6] = A[6]; X[7] = A[7]; X[34] = A[34];
    if (unit("stage_tendency_reset", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    for (int q = 0; q < 3; ++q) {
      slot(A[0], A[1], A[2], {{1, A[4 + 4 + q]}, {2, A[4 + 1 + q]}, {3, A[4 + 7]}}, X);
      if (unit("turb_add_tendency", X) != TE_OK) return TE_ERR_UNSUPPORTED;
    }
    writes.insert(writes.end(), {A[5], A[6], A[7]});
    return TE_OK;
  }
  if (nm == "add_fields") {
    const uint32_t nf = std::min<uint32_t>(W[A[4 + 13]], 4);
    for (uint32_t q = 0; q < nf; ++q) {
      slot(A[0], A[1], A[2], {{1, A[5 + q]}, {2, A[9 + q]}, {3, A[13 + q]}}, X);
      if (unit("turb_add_tendency", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      writes.push_back(A[9 + q]);
    }
    return TE_OK;
  }
  if (nm == "zero_fields") {
    const uint32_t nf = std::min<uint32_t>(W[A[4 + 9]], 8);
    for (uint32_t q = 0; q < nf; ++q) {
      slot(A[0], A[1], A[2], {{1, A[5 + q]}}, X);
      if (unit("turb_zero", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      writes.push_back(A[5 + q]);
    }
    return TE_OK;
  }
  if (nm == "copy_fields") {
    const uint32_t nf = std::min<uint32_t>(W[A[4 + 9]], 4);
    for (uint32_t q = 0; q < nf; ++q) {
      std::memset(X, 0, sizeof(X));
      X[0] = A[0]; X[1] = 1; X[2] = 1;
      X[4] = A[5 + q]; X[5] = A[9 + q];
      if (unit("_copy", X) != TE_OK) return TE_ERR_UNSUPPORTED;
      writes.push_back(A[9 + q]);
    }
    return TE_OK;
  }
  std::string ref = nm;
  const bool lds = ref.size() > 4 && ref.compare(ref.size() - 4, 4, "_lds") == 0;
  if (lds) ref.resize(ref.size() - 4);  // the same sweep with its elimination coefficients in LDS
  if (unit(ref.c_str(), A) != TE_OK) return TE_ERR_UNSUPPORTED;
  if (ref == "rayleigh_damping" && (W[A[4 + 17]] != 0u || W[A[4 + 16]] != 0u))
    return afail(a, TE_ERR_UNSUPPORTED, "no adjoint for Rayleigh damping toward horizontal means or at lateral edges");
  if (ref == "_copy") {
    writes.push_back(A[5]);
    return TE_OK;
  }
  int wp[8];
  const int nw = te::adj::writes(units.back().adj, wp);
  for (int q = 0; q < nw; ++q) {
    if (ref == "acoustic_pp" && wp[q] == 9 && W[A[4 + 10]] == 0u) continue;  // ppx unused
    if (lds && ref == "acoustic_w_solve" && wp[q] != 9) continue;            // cA, cB stay in LDS
    writes.push_back(A[4 + wp[q]]);
  }
  return TE_OK;
}

void te::adjoint_detach(te_adjoint* a) {
  if (a) a->fwd = nullptr;
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
  // the same program (fused or the reference's list) on memory of its own,
  // with the acoustic substep as its separate launches (fuse = 2): their
  // adjoints read the per-stage terms and fpk the one-launch substep keeps
  // in registers
  te_config cfg = ctx->cfg;
  if (cfg.fuse == 1) cfg.fuse = 2;
  int rc = te_create(&cfg, &a->rec);
  if (rc != TE_OK) return afail(a, rc, std::string("program: ") + (a->rec ? te_last_error(a->rec) : ""));
  te_ctx* r = a->rec;
  const te::Layout &L = r->layout, &F = ctx->layout;
  if (L.arena != F.arena || L.field_words != F.field_words || L.fields != F.fields)
    return afail(a, TE_ERR_STATE, "layouts differ");
  a->state_words = L.arena + L.field_words * L.fields.size();
  const uint64_t arena_end = a->state_words;
  // per dispatch: the reference dispatches it stands for and the regions it writes
  const int n = (int)r->ops.size();
  a->units.resize(n);
  a->wreg.resize(n);
  for (int i = 0; i < n; ++i) {
    std::vector<uint64_t> writes;
    rc = decompose(a, i, a->units[i], writes);
    if (rc != TE_OK) return rc;
    const std::string& nm = r->op_names[i];
    for (uint64_t off : writes) {
      if (off < L.arena || off >= arena_end || (off - L.arena) % L.field_words != 0)
        return afail(a, TE_ERR_UNSUPPORTED, "dispatch " + std::to_string(i) + " (" + nm + ") writes outside the FieldArena");
      bool dup = false;
      for (auto& w : a->wreg[i]) dup = dup || w.first == off;
      if (!dup) a->wreg[i].push_back({off, (nm == "_copy" || nm == "copy_fields") ? (uint64_t)r->ops[i].A[0] : L.field_words});
    }
    if (a->wreg[i].size() > (size_t)te::adj::kCopyRegions)
      return afail(a, TE_ERR_UNSUPPORTED, "dispatch " + std::to_string(i) + " (" + nm + ") writes too many fields");
  }
  // A region's saved values (written before dispatch i overwrites field f)
  // are what the adjoints of dispatch i and of the dispatches since f's
  // previous writer read as f (walking back, cur[f] points at them until
  // that writer).  When none of those has f among its parameters the
  // region is not saved.  The walk wraps around the step (the previous
  // writer may be in the step before); at a segment's first step it covers
  // dispatches before the segment too, which only keeps more regions.
  {
    std::vector<std::vector<uint64_t>> reads(n);
    for (int j = 0; j < n; ++j)
      for (const Unit& u : a->units[j])
        if (u.adj >= 0)
          for (int q = 0; q < 32; ++q) reads[j].push_back(u.A[4 + q]);
    auto has = [](const std::vector<uint64_t>& v, uint64_t x) { return std::find(v.begin(), v.end(), x) != v.end(); };
    auto writes = [&](int j, uint64_t f) {
      for (auto& w : a->wreg[j])
        if (w.first == f) return true;
      return false;
    };
    a->live.resize(n);
    for (int i = 0; i < n; ++i) {
      a->live[i].assign(a->wreg[i].size(), 0);
      for (size_t q = 0; q < a->wreg[i].size(); ++q) {
        const uint64_t f = a->wreg[i][q].first;
        for (int t = 0; t < n; ++t) {
          const int j = (i - t + n) % n;
          if (t > 0 && writes(j, f)) break;
          if (has(reads[j], f)) {
            a->live[i][q] = 1;
            break;
          }
        }
      }
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
  // the adjoint launches index the adjoint memory as the forward's index M:
  // give it the same front and back guards, grown to their final size first
  rc = te::ensure_guard(r);
  if (rc != TE_OK) return afail(a, rc, std::string("program: ") + te_last_error(r));
  const uint64_t mbw = r->front + r->words;
  TA_HIP(hipMalloc((void**)&a->dMbBase, mbw * 4), "hipMalloc(adjoint memory)");
  TA_HIP(hipMemsetAsync(a->dMbBase, 0, mbw * 4, s), "clear adjoint memory");
  a->dMb = a->dMbBase + r->front;
  a->mb_words = mbw;
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
  ctx->adjoints.push_back(a);
  return TE_OK;
}

te_ctx* te_adjoint_program(te_adjoint* a) { return a ? a->rec : nullptr; }

int te_fwd_ckpt(te_adjoint* a, int32_t nsteps) {
  if (!a || !a->rec || nsteps < 0) return afail(a, TE_ERR_ARG, "bad arguments");
  if (!a->fwd) return afail(a, TE_ERR_STATE, "the forward context was destroyed");
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
  // the checkpoints were written on the forward context's stream; a
  // destroyed context synchronised its stream before it went
  if (a->fwd) TA_HIP(hipStreamSynchronize(a->fwd->stream), "sync forward");
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
    for (int u = (int)a->units[i].size() - 1; u >= 0; --u) {
      const Unit& un = a->units[i][u];
      if (un.adj < 0) {
        int rc = halo_units(a, &un, 1);
        if (rc != TE_OK) return rc;
      } else
        TA_HIP(te::adj::launch(un.adj, un.A, nullptr, reinterpret_cast<const float*>(r->dM), a->dMb, r->layout.nz,
                               r->layout.sJ, r->layout.sK, a->dScr, r->layout.field_words, r->stream),
               std::string("adjoint of ") + r->op_names[i]);
    }
  TA_HIP(hipStreamSynchronize(r->stream), "sync");
  return TE_OK;
}

int te_adj_upload(te_adjoint* a, uint64_t off, const float* src, uint64_t n) {
  if (!a || !a->rec || !src || n > a->rec->prog_words || off > a->rec->prog_words - n)
    return afail(a, TE_ERR_ARG, "upload outside memory");
  TA_HIP(hipMemcpyAsync(a->dMb + off, src, n * 4, hipMemcpyHostToDevice, a->rec->stream), "upload");
  TA_HIP(hipStreamSynchronize(a->rec->stream), "sync");
  return TE_OK;
}

int te_adj_download(te_adjoint* a, uint64_t off, float* dst, uint64_t n) {
  if (!a || !a->rec || !dst || n > a->rec->prog_words || off > a->rec->prog_words - n)
    return afail(a, TE_ERR_ARG, "download outside memory");
  TA_HIP(hipMemcpyAsync(dst, a->dMb + off, n * 4, hipMemcpyDeviceToHost, a->rec->stream), "download");
  TA_HIP(hipStreamSynchronize(a->rec->stream), "sync");
  return TE_OK;
}

int te_adj_zero(te_adjoint* a) {
  if (!a || !a->rec) return afail(a, TE_ERR_ARG, "bad arguments");
  TA_HIP(hipMemsetAsync(a->dMbBase, 0, a->mb_words * 4, a->rec->stream), "clear");
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
  if (a->fwd) {
    auto& v = a->fwd->adjoints;
    v.erase(std::remove(v.begin(), v.end(), a), v.end());
  }
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
