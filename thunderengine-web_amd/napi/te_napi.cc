// Node N-API binding of libte (include/te.h) for the JavaScript host.
//
// Each export is a thin wrapper: JS numbers / typed arrays in, te_* call,
// nonzero status -> JS Error carrying te_last_error().  Contexts are
// napi_external values whose finalizer calls te_destroy.
#include <node_api.h>

#include <cstddef>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/te.h"

namespace {

#define NAPI_OK(call)                                                   \
  do {                                                                  \
    if ((call) != napi_ok) {                                            \
      napi_throw_error(env, nullptr, "N-API call failed: " #call);      \
      return nullptr;                                                   \
    }                                                                   \
  } while (0)

struct Holder {
  te_ctx* ctx = nullptr;
  std::vector<std::string> names;  // keeps te_config.arena_fields alive during te_encode
};

void finalize(napi_env, void* data, void*) {
  Holder* h = static_cast<Holder*>(data);
  if (h->ctx) te_destroy(h->ctx);
  delete h;
}

napi_value throw_status(napi_env env, const te_ctx* ctx, int rc, const char* what) {
  std::string msg = std::string(what) + " failed (" + std::to_string(rc) + "): " + (ctx ? te_last_error(ctx) : "");
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < want) {
    napi_throw_type_error(env, nullptr, "missing arguments");
    return false;
  }
  return true;
}

te_ctx* ctx_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a te context");
    return nullptr;
  }
  return static_cast<Holder*>(p)->ctx;
}

double num(napi_env env, napi_value v) {
  double d = 0;
  napi_get_value_double(env, v, &d);
  return d;
}

// A word offset or count from JS: a finite integer in [0, 2^53); anything
// else throws a RangeError (false returned).
bool index_arg(napi_env env, napi_value v, uint64_t* out) {
  double d = 0;
  if (napi_get_value_double(env, v, &d) != napi_ok || !(d >= 0.0) || !(d < 9007199254740992.0) ||
      d != (double)(uint64_t)d) {
    napi_throw_range_error(env, nullptr, "offsets and counts must be integers in [0, 2^53)");
    return false;
  }
  *out = (uint64_t)d;
  return true;
}

napi_value make_num(napi_env env, double d) {
  napi_value r;
  napi_create_double(env, d, &r);
  return r;
}

napi_value make_str(napi_env env, const char* s) {
  napi_value r;
  napi_create_string_utf8(env, s ? s : "", NAPI_AUTO_LENGTH, &r);
  return r;
}

// typed array (any element type) -> data pointer and byte length
bool typed(napi_env env, napi_value v, void** data, size_t* bytes) {
  napi_typedarray_type t;
  size_t len = 0, off = 0;
  napi_value ab;
  if (napi_get_typedarray_info(env, v, &t, &len, data, &ab, &off) != napi_ok) {
    napi_throw_type_error(env, nullptr, "expected a typed array");
    return false;
  }
  size_t el = (t == napi_int8_array || t == napi_uint8_array || t == napi_uint8_clamped_array) ? 1
              : (t == napi_int16_array || t == napi_uint16_array) ? 2
              : (t == napi_float64_array || t == napi_bigint64_array || t == napi_biguint64_array) ? 8 : 4;
  *bytes = len * el;
  return true;
}

napi_value wrap(napi_env env, Holder* h) {
  napi_value r;
  if (napi_create_external(env, h, finalize, nullptr, &r) != napi_ok) {
    finalize(env, h, nullptr);
    napi_throw_error(env, nullptr, "napi_create_external");
    return nullptr;
  }
  return r;
}

// ---- te_config <-> JS object
enum Kind { I32, F32, U64, F64 };
struct FieldDesc { const char* name; Kind kind; size_t off; };
#define FD(n, k) {#n, k, offsetof(te_config, n)}
const FieldDesc kFields[] = {
    FD(nx, I32), FD(ny, I32), FD(nz, I32), FD(dx, F32), FD(dy, F32), FD(dz, F32), FD(dt, F32), FD(nsound, I32),
    FD(moist, I32), FD(open_x, I32), FD(open_y, I32), FD(weno_scalars, I32), FD(weno_momentum, I32), FD(tke, I32),
    FD(explicit_vdiff, I32), FD(dissipative_heating, I32), FD(rayleigh, I32), FD(rayleigh_zd, F32),
    FD(rayleigh_alpha, F32), FD(alph, F32), FD(kdiv, F32), FD(smeps, F32), FD(umove, F32), FD(vmove, F32),
    FD(th_sfc, F32), FD(th_trop, F32), FD(t_trop, F32), FD(z_trop, F32), FD(qv_pbl, F32), FD(hodo_radius, F32),
    FD(hodo_depth, F32), FD(shear_umax, F32), FD(shear_depth, F32), FD(bubble_amp, F32), FD(bubble_rh, F32),
    FD(bubble_rv, F32), FD(bubble_zc, F32), FD(noise_sigma, F32), FD(seed, U64), FD(initial_tke, I32), FD(fuse, I32), FD(sounding_seed, U64), FD(th_pert_sigma, F32), FD(qv_pert_sigma, F32),
    FD(species, I32), FD(microphysics, I32), FD(mass_conservation, I32), FD(cache_moisture_turbulence, I32),
    FD(stretch_x, I32), FD(stretch_y, I32), FD(dx_inner, F64), FD(dx_outer, F64), FD(nos_x_len, F64),
    FD(tot_x_len, F64), FD(dy_inner, F64), FD(dy_outer, F64), FD(nos_y_len, F64), FD(tot_y_len, F64),
    FD(adaptive_dt, I32), FD(tapfrq, F32), FD(statfrq, F32), FD(rstfrq, F32), FD(sounding_winds, I32),
};
#undef FD

napi_value config_to_js(napi_env env, const te_config& c) {
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  for (const FieldDesc& f : kFields) {
    const char* p = reinterpret_cast<const char*>(&c) + f.off;
    double d = f.kind == I32   ? *reinterpret_cast<const int32_t*>(p)
               : f.kind == F32 ? *reinterpret_cast<const float*>(p)
               : f.kind == F64 ? *reinterpret_cast<const double*>(p)
                               : (double)*reinterpret_cast<const uint64_t*>(p);
    NAPI_OK(napi_set_named_property(env, o, f.name, make_num(env, d)));
  }
  return o;
}

// Reads every known key present on `o` into c; zf (array of nz+1 face
// heights) and arena_fields (array of names) into the holders' storage.
bool config_from_js(napi_env env, napi_value o, te_config* c, std::vector<float>* zf, std::vector<std::string>* names,
                    std::vector<const char*>* cnames, std::string* sounding) {
  te_config_defaults(c);
  for (const FieldDesc& f : kFields) {
    bool has = false;
    napi_has_named_property(env, o, f.name, &has);
    if (!has) continue;
    napi_value v;
    napi_get_named_property(env, o, f.name, &v);
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t == napi_undefined || t == napi_null) continue;
    if (t != napi_number && t != napi_boolean) {
      napi_throw_type_error(env, nullptr, (std::string("config.") + f.name + " must be a number").c_str());
      return false;
    }
    double d = 0;
    if (t == napi_boolean) {
      bool b = false;
      napi_get_value_bool(env, v, &b);
      d = b ? 1 : 0;
    } else {
      napi_get_value_double(env, v, &d);
    }
    char* p = reinterpret_cast<char*>(c) + f.off;
    if (f.kind == I32) *reinterpret_cast<int32_t*>(p) = (int32_t)d;
    else if (f.kind == F32) *reinterpret_cast<float*>(p) = (float)d;
    else if (f.kind == F64) *reinterpret_cast<double*>(p) = d;
    else *reinterpret_cast<uint64_t*>(p) = (uint64_t)d;
  }
  bool has = false;
  napi_has_named_property(env, o, "zf", &has);
  if (has) {
    napi_value a;
    napi_get_named_property(env, o, "zf", &a);
    uint32_t n = 0;
    if (napi_get_array_length(env, a, &n) == napi_ok && n > 0) {
      zf->resize(n);
      for (uint32_t i = 0; i < n; ++i) {
        napi_value e;
        napi_get_element(env, a, i, &e);
        (*zf)[i] = (float)num(env, e);
      }
      c->zf = zf->data();
    }
  }
  napi_has_named_property(env, o, "input_sounding", &has);
  if (has) {
    napi_value t;
    napi_get_named_property(env, o, "input_sounding", &t);
    napi_valuetype ty;
    napi_typeof(env, t, &ty);
    if (ty == napi_string) {
      size_t len = 0;
      napi_get_value_string_utf8(env, t, nullptr, 0, &len);
      sounding->assign(len, '\0');
      napi_get_value_string_utf8(env, t, &(*sounding)[0], len + 1, &len);
      c->input_sounding = sounding->c_str();
    } else if (ty != napi_undefined && ty != napi_null) {
      napi_throw_type_error(env, nullptr, "config.input_sounding must be a string");
      return false;
    }
  }
  napi_has_named_property(env, o, "arena_fields", &has);
  if (has) {
    napi_value a;
    napi_get_named_property(env, o, "arena_fields", &a);
    uint32_t n = 0;
    if (napi_get_array_length(env, a, &n) == napi_ok && n > 0) {
      names->resize(n);
      for (uint32_t i = 0; i < n; ++i) {
        napi_value e;
        napi_get_element(env, a, i, &e);
        size_t len = 0;
        napi_get_value_string_utf8(env, e, nullptr, 0, &len);
        (*names)[i].resize(len);
        napi_get_value_string_utf8(env, e, &(*names)[i][0], len + 1, &len);
      }
      for (auto& s : *names) cnames->push_back(s.c_str());
      c->arena_fields = cnames->data();
      c->n_arena_fields = (int32_t)n;
    }
  }
  return true;
}

// ---- exports
napi_value ConfigDefaults(napi_env env, napi_callback_info) {
  te_config c;
  te_config_defaults(&c);
  return config_to_js(env, c);
}

napi_value Create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  size_t argc = 2;
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  te_config c;
  std::vector<float> zf;
  std::vector<std::string> names;
  std::vector<const char*> cnames;
  std::string sounding;
  if (argc >= 1) {
    if (!config_from_js(env, argv[0], &c, &zf, &names, &cnames, &sounding)) return nullptr;
  } else {
    te_config_defaults(&c);
  }
  bool encodeOnly = false;
  if (argc >= 2) napi_get_value_bool(env, argv[1], &encodeOnly);
  Holder* h = new Holder();
  int rc = encodeOnly ? te_encode(&c, &h->ctx) : te_create(&c, &h->ctx);
  if (rc != TE_OK) {
    napi_value r = throw_status(env, h->ctx, rc, encodeOnly ? "te_encode" : "te_create");
    finalize(env, h, nullptr);
    return r;
  }
  return wrap(env, h);
}

napi_value ProgramCreate(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  double memWords = num(env, argv[0]);
  uint32_t n = 0;
  NAPI_OK(napi_get_array_length(env, argv[1], &n));
  std::vector<std::string> names(n);
  std::vector<const char*> cn(n);
  for (uint32_t i = 0; i < n; ++i) {
    napi_value e;
    NAPI_OK(napi_get_element(env, argv[1], i, &e));
    size_t len = 0;
    NAPI_OK(napi_get_value_string_utf8(env, e, nullptr, 0, &len));
    names[i].resize(len);
    NAPI_OK(napi_get_value_string_utf8(env, e, &names[i][0], len + 1, &len));
    cn[i] = names[i].c_str();
  }
  void* data = nullptr;
  size_t bytes = 0;
  if (!typed(env, argv[2], &data, &bytes)) return nullptr;
  if (bytes != (size_t)n * 64 * 4) {
    napi_throw_range_error(env, nullptr, "args must hold 64 words per op");
    return nullptr;
  }
  Holder* h = new Holder();
  int rc = te_program_create((uint64_t)memWords, (int32_t)n, cn.data(), static_cast<const uint32_t*>(data), &h->ctx);
  if (rc != TE_OK) {
    napi_value r = throw_status(env, h->ctx, rc, "te_program_create");
    finalize(env, h, nullptr);
    return r;
  }
  return wrap(env, h);
}

napi_value Upload(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  void* data = nullptr;
  size_t bytes = 0;
  if (!typed(env, argv[2], &data, &bytes)) return nullptr;
  uint64_t off = 0;
  if (!index_arg(env, argv[1], &off)) return nullptr;
  int rc = te_upload_words(c, off, static_cast<const uint32_t*>(data), bytes / 4);
  if (rc != TE_OK) return throw_status(env, c, rc, "te_upload_words");
  return nullptr;
}

napi_value Download(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  uint64_t off = 0, n = 0;
  if (!index_arg(env, argv[1], &off) || !index_arg(env, argv[2], &n)) return nullptr;
  if (n > te_memory_words(c) || off > te_memory_words(c) - n) {
    napi_throw_range_error(env, nullptr, "download outside memory");
    return nullptr;
  }
  void* data = nullptr;
  napi_value ab, arr;
  NAPI_OK(napi_create_arraybuffer(env, n * 4, &data, &ab));
  int rc = te_download_words(c, off, static_cast<uint32_t*>(data), n);
  if (rc != TE_OK) return throw_status(env, c, rc, "te_download_words");
  NAPI_OK(napi_create_typedarray(env, napi_uint32_array, n, ab, 0, &arr));
  return arr;
}

napi_value Run(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int rc = te_run(c, (int32_t)num(env, argv[1]), (int32_t)num(env, argv[2]), (int32_t)num(env, argv[3]));
  if (rc != TE_OK) return throw_status(env, c, rc, "te_run");
  return nullptr;
}

napi_value Step(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int rc = te_step(c, (int32_t)num(env, argv[1]));
  if (rc != TE_OK) return throw_status(env, c, rc, "te_step");
  return nullptr;
}

napi_value Sync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int rc = te_sync(c);
  if (rc != TE_OK) return throw_status(env, c, rc, "te_sync");
  return nullptr;
}

napi_value InitState(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int rc = te_init_state(c);
  if (rc != TE_OK) return throw_status(env, c, rc, "te_init_state");
  return nullptr;
}

napi_value LastRunMs(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  float ms = 0;
  int rc = te_last_run_ms(c, &ms);
  if (rc != TE_OK) return throw_status(env, c, rc, "te_last_run_ms");
  return make_num(env, ms);
}

// autotune(ctx, reps, force) -> number of dispatches left on a tuned shape
napi_value Autotune(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int rc = te_autotune(c, (int32_t)num(env, argv[1]), (int32_t)num(env, argv[2]));
  if (rc != TE_OK) return throw_status(env, c, rc, "te_autotune");
  return make_num(env, te_tuned_ops(c));
}

// ---- adapt_dt = 1
napi_value clock_to_js(napi_env env, const te_clock_state& st) {
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  const std::pair<const char*, double> kv[] = {
      {"mtime", st.mtime}, {"dbldt", st.dbldt}, {"dt", st.dt}, {"dtlast", st.dtlast}, {"cflmax", st.cflmax},
      {"ksmax", st.ksmax}, {"nsound", (double)st.nsound}, {"steps", (double)st.steps}, {"fired", (double)st.fired}};
  for (const auto& p : kv) NAPI_OK(napi_set_named_property(env, o, p.first, make_num(env, p.second)));
  return o;
}

napi_value SetTimeStep(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int rc = te_set_time_step(c, (float)num(env, argv[1]), (int32_t)num(env, argv[2]));
  if (rc != TE_OK) return throw_status(env, c, rc, "te_set_time_step");
  return nullptr;
}

napi_value AdaptiveStart(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int rc = te_adaptive_start(c);
  if (rc != TE_OK) return throw_status(env, c, rc, "te_adaptive_start");
  return nullptr;
}

napi_value AdaptiveStep(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int rc = te_adaptive_step(c, (int32_t)num(env, argv[1]));
  if (rc != TE_OK) return throw_status(env, c, rc, "te_adaptive_step");
  return nullptr;
}

napi_value AdaptiveClock(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  te_clock_state st;
  int rc = te_adaptive_clock(c, &st);
  if (rc != TE_OK) return throw_status(env, c, rc, "te_adaptive_clock");
  return clock_to_js(env, st);
}

// te_clock as an external: clockCreate(dtl, minDx, minDz, twoD, ksCheck, [6 freqs, NaN = absent])
void clock_finalize(napi_env, void* p, void*) { te_clock_destroy(static_cast<te_clock*>(p)); }
te_clock* clock_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a clock");
    return nullptr;
  }
  return static_cast<te_clock*>(p);
}
napi_value ClockCreate(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return nullptr;
  float fr[6];
  for (uint32_t i = 0; i < 6; ++i) {
    napi_value e;
    NAPI_OK(napi_get_element(env, argv[5], i, &e));
    fr[i] = (float)num(env, e);
  }
  te_clock* c = te_clock_create(num(env, argv[0]), num(env, argv[1]), num(env, argv[2]), (int32_t)num(env, argv[3]),
                                (int32_t)num(env, argv[4]), fr);
  if (!c) {
    napi_throw_error(env, nullptr, "te_clock_create failed");
    return nullptr;
  }
  napi_value r;
  if (napi_create_external(env, c, clock_finalize, nullptr, &r) != napi_ok) {
    te_clock_destroy(c);
    napi_throw_error(env, nullptr, "napi_create_external");
    return nullptr;
  }
  return r;
}
napi_value ClockCall(napi_env env, napi_callback_info info) {
  // clockCall(clock, 'start' | 'begin' | 'end' | 'get', cflmax, ksmax)
  napi_value argv[4];
  size_t argc = 4;
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 2) {
    napi_throw_type_error(env, nullptr, "clockCall(clock, what, ...)");
    return nullptr;
  }
  te_clock* c = clock_of(env, argv[0]);
  if (!c) return nullptr;
  char what[8] = {0};
  size_t len = 0;
  napi_get_value_string_utf8(env, argv[1], what, sizeof(what), &len);
  const std::string w(what);
  if ((w == "start" || w == "end") && argc < 4) {
    napi_throw_type_error(env, nullptr, "clockCall: cflmax and ksmax expected");
    return nullptr;
  }
  if (w == "start") te_clock_start(c, (float)num(env, argv[2]), (float)num(env, argv[3]));
  else if (w == "begin") te_clock_begin_step(c);
  else if (w == "end") te_clock_end_step(c, (float)num(env, argv[2]), (float)num(env, argv[3]));
  else if (w != "get") {
    napi_throw_type_error(env, nullptr, "clockCall: unknown call");
    return nullptr;
  }
  te_clock_state st;
  te_clock_get(c, &st);
  return clock_to_js(env, st);
}

napi_value HashCall(napi_env env, napi_callback_info info, bool fnv) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  uint32_t h = 0;
  uint64_t off = 0, n = 0;
  if (!index_arg(env, argv[1], &off) || !index_arg(env, argv[2], &n)) return nullptr;
  int rc = fnv ? te_fnv1a(c, off, n, &h) : te_region_hash(c, off, n, &h);
  if (rc != TE_OK) return throw_status(env, c, rc, fnv ? "te_fnv1a" : "te_region_hash");
  return make_num(env, h);
}
napi_value RegionHash(napi_env env, napi_callback_info info) { return HashCall(env, info, false); }
napi_value Fnv1a(napi_env env, napi_callback_info info) { return HashCall(env, info, true); }

napi_value MemoryWords(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  return make_num(env, (double)te_memory_words(c));
}

napi_value Program(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int32_t n = te_op_count(c);
  napi_value names, ab, args, out;
  void* data = nullptr;
  NAPI_OK(napi_create_array_with_length(env, n, &names));
  NAPI_OK(napi_create_arraybuffer(env, (size_t)n * 256, &data, &ab));
  for (int32_t i = 0; i < n; ++i) {
    const char* nm = nullptr;
    te_op_get(c, i, &nm, static_cast<uint32_t*>(data) + (size_t)i * 64);
    NAPI_OK(napi_set_element(env, names, i, make_str(env, nm)));
  }
  NAPI_OK(napi_create_typedarray(env, napi_uint32_array, (size_t)n * 64, ab, 0, &args));
  NAPI_OK(napi_create_object(env, &out));
  NAPI_OK(napi_set_named_property(env, out, "names", names));
  NAPI_OK(napi_set_named_property(env, out, "args", args));
  return out;
}

napi_value Layout(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int32_t g[5];
  int rc = te_grid(c, g);
  if (rc != TE_OK) return throw_status(env, c, rc, "te_grid");
  napi_value out, fields;
  NAPI_OK(napi_create_object(env, &out));
  const char* keys[] = {"nx", "ny", "nz", "strideJ", "strideK"};
  for (int q = 0; q < 5; ++q) NAPI_OK(napi_set_named_property(env, out, keys[q], make_num(env, g[q])));
  NAPI_OK(napi_set_named_property(env, out, "fieldStrideWords", make_num(env, (double)te_field_stride_words(c))));
  NAPI_OK(napi_create_object(env, &fields));
  for (int32_t i = 0; i < te_field_count(c); ++i) {
    const char* nm = te_field_name(c, i);
    uint64_t off = 0;
    te_field_offset(c, nm, &off);
    NAPI_OK(napi_set_named_property(env, fields, nm, make_num(env, (double)off)));
  }
  NAPI_OK(napi_set_named_property(env, out, "fields", fields));
  return out;
}

// [{name, off}] of the context's regions in memory order
napi_value Regions(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  const int32_t n = te_region_count(c);
  napi_value arr;
  NAPI_OK(napi_create_array_with_length(env, n, &arr));
  for (int32_t i = 0; i < n; ++i) {
    const char* nm = te_region_name(c, i);
    uint64_t off = 0;
    te_region_offset(c, nm, &off);
    napi_value o;
    NAPI_OK(napi_create_object(env, &o));
    NAPI_OK(napi_set_named_property(env, o, "name", make_str(env, nm)));
    NAPI_OK(napi_set_named_property(env, o, "off", make_num(env, (double)off)));
    NAPI_OK(napi_set_element(env, arr, i, o));
  }
  return arr;
}

napi_value KernelNames(napi_env env, napi_callback_info) {
  napi_value arr;
  int32_t n = te_kernel_count();
  NAPI_OK(napi_create_array_with_length(env, n, &arr));
  for (int32_t i = 0; i < n; ++i) NAPI_OK(napi_set_element(env, arr, i, make_str(env, te_kernel_name(i))));
  return arr;
}

napi_value ProfileOps(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  int32_t n = te_op_count(c);
  void* data = nullptr;
  napi_value ab, arr;
  NAPI_OK(napi_create_arraybuffer(env, (size_t)n * 4, &data, &ab));
  int rc = te_profile_ops(c, (int32_t)num(env, argv[1]), static_cast<float*>(data));
  if (rc != TE_OK) return throw_status(env, c, rc, "te_profile_ops");
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr));
  return arr;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  void* p = nullptr;
  if (napi_get_value_external(env, argv[0], &p) == napi_ok && p) {
    Holder* h = static_cast<Holder*>(p);
    if (h->ctx) te_destroy(h->ctx);
    h->ctx = nullptr;
  }
  return nullptr;
}

// ---- adjoint (te_adjoint_*): externals owning a te_adjoint
struct AdjHolder {
  te_adjoint* adj = nullptr;
};

void finalize_adj(napi_env, void* data, void*) {
  AdjHolder* h = static_cast<AdjHolder*>(data);
  if (h->adj) te_adjoint_destroy(h->adj);
  delete h;
}

te_adjoint* adj_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p || !static_cast<AdjHolder*>(p)->adj) {
    napi_throw_type_error(env, nullptr, "expected a te adjoint");
    return nullptr;
  }
  return static_cast<AdjHolder*>(p)->adj;
}

napi_value throw_adj(napi_env env, const te_adjoint* a, int rc, const char* what) {
  std::string msg = std::string(what) + " failed (" + std::to_string(rc) + "): " + te_adjoint_last_error(a);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

napi_value AdjointCreate(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  te_ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  AdjHolder* h = new AdjHolder();
  int rc = te_adjoint_create(c, (int32_t)num(env, argv[1]), &h->adj);
  if (rc != TE_OK) {
    napi_value r = throw_adj(env, h->adj, rc, "te_adjoint_create");
    if (h->adj) te_adjoint_destroy(h->adj);
    delete h;
    return r;
  }
  napi_value r;
  NAPI_OK(napi_create_external(env, h, finalize_adj, nullptr, &r));
  return r;
}

napi_value FwdCkpt(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return nullptr;
  te_adjoint* a = adj_of(env, argv[0]);
  if (!a) return nullptr;
  int rc = te_fwd_ckpt(a, (int32_t)num(env, argv[1]));
  if (rc != TE_OK) return throw_adj(env, a, rc, "te_fwd_ckpt");
  return nullptr;
}

napi_value Bwd(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_adjoint* a = adj_of(env, argv[0]);
  if (!a) return nullptr;
  int rc = te_bwd(a);
  if (rc != TE_OK) return throw_adj(env, a, rc, "te_bwd");
  return nullptr;
}

napi_value AdjZero(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_adjoint* a = adj_of(env, argv[0]);
  if (!a) return nullptr;
  int rc = te_adj_zero(a);
  if (rc != TE_OK) return throw_adj(env, a, rc, "te_adj_zero");
  return nullptr;
}

napi_value AdjUpload(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  te_adjoint* a = adj_of(env, argv[0]);
  if (!a) return nullptr;
  void* data = nullptr;
  size_t bytes = 0;
  if (!typed(env, argv[2], &data, &bytes)) return nullptr;
  uint64_t off = 0;
  if (!index_arg(env, argv[1], &off)) return nullptr;
  int rc = te_adj_upload(a, off, static_cast<const float*>(data), bytes / 4);
  if (rc != TE_OK) return throw_adj(env, a, rc, "te_adj_upload");
  return nullptr;
}

napi_value AdjDownload(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return nullptr;
  te_adjoint* a = adj_of(env, argv[0]);
  if (!a) return nullptr;
  uint64_t off = 0, n = 0;
  if (!index_arg(env, argv[1], &off) || !index_arg(env, argv[2], &n)) return nullptr;
  const uint64_t words = te_memory_words(te_adjoint_program(a));
  if (n > words || off > words - n) {
    napi_throw_range_error(env, nullptr, "download outside memory");
    return nullptr;
  }
  void* data = nullptr;
  napi_value ab, arr;
  NAPI_OK(napi_create_arraybuffer(env, n * 4, &data, &ab));
  int rc = te_adj_download(a, off, static_cast<float*>(data), n);
  if (rc != TE_OK) return throw_adj(env, a, rc, "te_adj_download");
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr));
  return arr;
}

napi_value AdjLastMs(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  te_adjoint* a = adj_of(env, argv[0]);
  if (!a) return nullptr;
  float f = 0.0f, b = 0.0f;
  te_adj_last_ms(a, &f, &b);
  napi_value out;
  NAPI_OK(napi_create_array_with_length(env, 2, &out));
  NAPI_OK(napi_set_element(env, out, 0, make_num(env, f)));
  NAPI_OK(napi_set_element(env, out, 1, make_num(env, b)));
  return out;
}

napi_value AdjDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return nullptr;
  void* p = nullptr;
  if (napi_get_value_external(env, argv[0], &p) == napi_ok && p) {
    AdjHolder* h = static_cast<AdjHolder*>(p);
    if (h->adj) te_adjoint_destroy(h->adj);
    h->adj = nullptr;
  }
  return nullptr;
}

napi_value Init(napi_env env, napi_value exports) {
  struct { const char* name; napi_callback fn; } fns[] = {
      {"configDefaults", ConfigDefaults}, {"create", Create}, {"programCreate", ProgramCreate},
      {"upload", Upload}, {"download", Download}, {"run", Run}, {"step", Step}, {"sync", Sync},
      {"initState", InitState}, {"lastRunMs", LastRunMs}, {"autotune", Autotune}, {"regionHash", RegionHash}, {"fnv1a", Fnv1a},
      {"memoryWords", MemoryWords}, {"program", Program}, {"layout", Layout}, {"kernelNames", KernelNames},
      {"profileOps", ProfileOps}, {"destroy", Destroy},
      {"adjointCreate", AdjointCreate}, {"fwdCkpt", FwdCkpt}, {"bwd", Bwd}, {"adjZero", AdjZero},
      {"adjUpload", AdjUpload}, {"adjDownload", AdjDownload}, {"adjLastMs", AdjLastMs}, {"adjDestroy", AdjDestroy},
      {"setTimeStep", SetTimeStep}, {"adaptiveStart", AdaptiveStart}, {"adaptiveStep", AdaptiveStep},
      {"adaptiveClock", AdaptiveClock}, {"clockCreate", ClockCreate}, {"clockCall", ClockCall},
      {"regions", Regions},
  };
  for (auto& f : fns) {
    napi_value v;
    NAPI_OK(napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.fn, nullptr, &v));
    NAPI_OK(napi_set_named_property(env, exports, f.name, v));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
