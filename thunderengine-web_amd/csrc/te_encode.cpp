// The large-step encoder: builds the op list of one CM1 large step for a
// configuration, in the dispatch order, launch grids, parameter bindings and
// parameter bytes of the reference's TimeStep.encode (web/core/timestep.js:
// 797-2201).  Only the hot-path flags are built here (TKE, Rayleigh, open or
// periodic lateral boundaries, WENO, pressure tendency, mass-conservation and
// Morrison call sites for recorded configurations); the memory layout is this
// library's own: profiles, then the FieldArena, then parameter bytes.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <map>
#include <set>
#include <stdexcept>
#include <string>

#include "te_internal.hpp"
#include "te_scenario.hpp"

namespace te {
namespace {

using Bytes = std::vector<uint32_t>;
static uint32_t fbits(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
static Bytes F1(float v) { return {fbits(v)}; }
static Bytes I1(int32_t v) { return {(uint32_t)v}; }
static Bytes I2(int32_t a, int32_t b) { return {(uint32_t)a, (uint32_t)b}; }
static Bytes Fx2(float a, float b) { return {fbits(a), fbits(b)}; }
static Bytes Fx3(const float* v) { return {fbits(v[0]), fbits(v[1]), fbits(v[2]), 0u}; }
static Bytes Fx4(float a, float b, float c, float d) { return {fbits(a), fbits(b), fbits(c), fbits(d)}; }

// AcousticSchedule (timestep.js:54-85)
struct Schedule {
  int count[3];
  double dts[3];
};
static double round_half_away(double x) { return x < 0 ? -std::round(-x) : std::round(x); }
static int schedule_nsound(double dt, double dmin) {
  const float dtsm = (0.60f * (float)dmin) / 360.0f;
  int n = std::max((int)round_half_away(dt / dtsm), 4);
  if (n % 2 != 0) n += 1;
  if (dt / n > dtsm) n += 2;
  if (n > 192) throw std::runtime_error("time step needs more than 192 acoustic substeps; use a smaller dt");
  return n;
}
static Schedule make_schedule(int n, double dt) {
  Schedule s;
  int n1 = std::max((int)round_half_away(n / 3.0), 1);
  double d1 = dt / (n1 * 3.0);
  if (d1 > dt / n) { n1 += 1; d1 = dt / (n1 * 3.0); }
  s.count[0] = n1; s.dts[0] = d1;
  s.count[1] = n / 2; s.dts[1] = dt / n;
  s.count[2] = n; s.dts[2] = dt / n;
  return s;
}

struct Bind {
  int idx;
  uint64_t off;
};

class Encoder {
 public:
  explicit Encoder(const te_config& c) : cfg(c) {}
  bool withArena = true;  // false: the image leaves the arena out (Encoded::skipped)
  Encoded run();

 private:
  const te_config& cfg;
  std::vector<std::string> scalars;
  const std::vector<std::string> momentum{"u", "v", "w"};
  bool moist = false, tke = false, rayleigh = false, openX = false, openY = false, periodicX = false,
       periodicY = false, wenoS = true, wenoM = false, explicitZ = false, diss = false, cacheMT = false,
       hasMicro = false, massCons = false;
  const bool usePT = true;
  Vertical v;
  Base b;
  Profiles pr;
  Encoded E;
  std::map<Bytes, uint64_t> bytesAt;
  uint64_t dimsOff = 0, hmOff = 0;
  double minSpacing = 0;
  float dtf = 0;

  uint64_t region(const std::string& n) const {
    auto it = E.L.region.find(n);
    if (it == E.L.region.end()) throw std::runtime_error("encoder: no region " + n);
    return it->second;
  }
  uint64_t off(const std::string& n) const {
    uint64_t o = E.L.field(n);
    if (o == UINT64_MAX) throw std::runtime_error("encoder: field '" + n + "' is not in the arena");
    return o;
  }
  bool has(const std::string& n) const { return E.L.field(n) != UINT64_MAX; }
  std::string fieldName(uint64_t o) const {
    if (o < E.L.arena || E.L.field_words == 0) throw std::runtime_error("encoder: offset outside the arena");
    const uint64_t q = (o - E.L.arena) / E.L.field_words;
    if (q >= E.L.fields.size() || E.L.arena + q * E.L.field_words != o) throw std::runtime_error("encoder: not a field start");
    return E.L.fields[q];
  }
  uint64_t bytes(const Bytes& w) {
    auto it = bytesAt.find(w);
    if (it != bytesAt.end()) return it->second;
    uint64_t o = E.image.size() + E.skipped;
    E.image.insert(E.image.end(), w.begin(), w.end());
    while (E.image.size() % 4) E.image.push_back(0u);
    E.image.push_back(0u); E.image.push_back(0u); E.image.push_back(0u); E.image.push_back(0u);
    bytesAt[w] = o;
    return o;
  }
  Bytes acousticBytes(double dts) const {
    const float alph = cfg.alph, beta = 1.0f - cfg.alph;
    return {fbits((float)dts), fbits(alph), fbits(beta), fbits(cfg.kdiv), fbits(cfg.smeps)};
  }
  Bytes acoustic;  // current stage's acoustic parameter bytes

  // Periodic halo fills queue here when both directions are periodic and
  // fusion is on; they are written by one bc_periodic_xy launch (up to 8
  // fields) just before the next op that is not a halo fill.
  bool mergeHalo = false;
  // fuse = 1 on a periodic domain (nx, ny >= 4): the halo fill after a
  // scalar's RK update, TKE's advance and the stage's pressure is written by
  // the launch that produces the field (each cell writes its periodic images)
  // instead of by a bc_periodic_xy launch
  bool foldHalo = false;
  std::vector<std::string> pendingHalo, pendingZero;
  // turb_zero runs queue likewise (fusion on): one zero_fields launch per 8.
  void flushZero() {
    std::vector<std::string> q;
    q.swap(pendingZero);
    for (size_t s = 0; s < q.size(); s += 8) {
      const int n = (int)std::min<size_t>(8, q.size() - s);
      Op o;
      std::memset(&o, 0, sizeof(o));
      o.kid = kernel_id("zero_fields");
      o.A[0] = (uint32_t)(cfg.nx + 7); o.A[1] = (uint32_t)(cfg.ny + 7); o.A[2] = (uint32_t)(cfg.nz + 1);
      o.A[4 + 0] = (uint32_t)dimsOff;
      o.A[4 + 30] = (uint32_t)hmOff;
      for (int q2 = 0; q2 < n; ++q2) o.A[4 + 1 + q2] = (uint32_t)off(q[s + q2]);
      o.A[4 + 9] = (uint32_t)bytes(I1(n));
      E.ops.push_back(o);
      E.names.push_back("zero_fields");
    }
  }
  // turb_add_tendency runs queue likewise: up to 4 independent (src, dst,
  // scale) triples per add_fields launch.
  struct AddT { uint64_t src, dst, scale; };
  std::vector<AddT> pendingAdd;
  void flushAdd() {
    std::vector<AddT> q;
    q.swap(pendingAdd);
    for (size_t s = 0; s < q.size(); s += 4) {
      const int n = (int)std::min<size_t>(4, q.size() - s);
      Op o;
      std::memset(&o, 0, sizeof(o));
      o.kid = kernel_id("add_fields");
      o.A[0] = (uint32_t)(cfg.nx + 7); o.A[1] = (uint32_t)(cfg.ny + 7); o.A[2] = (uint32_t)(cfg.nz + 1);
      o.A[4 + 0] = (uint32_t)dimsOff;
      o.A[4 + 30] = (uint32_t)hmOff;
      for (int q2 = 0; q2 < n; ++q2) {
        o.A[4 + 1 + q2] = (uint32_t)q[s + q2].src;
        o.A[4 + 5 + q2] = (uint32_t)q[s + q2].dst;
        o.A[4 + 9 + q2] = (uint32_t)q[s + q2].scale;
      }
      o.A[4 + 13] = (uint32_t)bytes(I1(n));
      E.ops.push_back(o);
      E.names.push_back("add_fields");
    }
  }
  void flushHalo() {
    std::vector<std::string> q;
    q.swap(pendingHalo);
    for (size_t s = 0; s < q.size(); s += 8) {
      const int n = (int)std::min<size_t>(8, q.size() - s);
      Op o;
      std::memset(&o, 0, sizeof(o));
      o.kid = kernel_id("bc_periodic_xy");
      o.A[0] = (uint32_t)(7 * (cfg.nx + 7) + 7 * cfg.ny); o.A[1] = (uint32_t)(cfg.nz + 1); o.A[2] = 1;
      o.A[4 + 0] = (uint32_t)dimsOff;
      o.A[4 + 30] = (uint32_t)hmOff;
      for (int q2 = 0; q2 < n; ++q2) o.A[4 + 1 + q2] = (uint32_t)off(q[s + q2]);
      o.A[4 + 9] = (uint32_t)bytes(I1(n));
      E.ops.push_back(o);
      E.names.push_back("bc_periodic_xy");
    }
  }

  void op(const char* kname, std::array<int, 3> grid, std::initializer_list<Bind> binds, bool withAcoustic = false) {
    if (cfg.fuse != 0 && std::strcmp(kname, "turb_add_tendency") == 0 && binds.size() == 3) {
      if (!pendingHalo.empty()) flushHalo();
      if (!pendingZero.empty()) flushZero();
      uint64_t v[4] = {0, 0, 0, 0};
      for (const Bind& bd : binds) v[bd.idx] = bd.off;
      for (const AddT& t : pendingAdd)  // a dependent triple starts a new launch
        if (t.dst == v[1] || t.src == v[2] || t.dst == v[2]) { flushAdd(); break; }
      pendingAdd.push_back({v[1], v[2], v[3]});
      return;
    }
    if (!pendingAdd.empty()) flushAdd();
    if (cfg.fuse != 0 && std::strcmp(kname, "turb_zero") == 0 && binds.size() == 1) {
      if (!pendingHalo.empty()) flushHalo();
      const std::string f = fieldName(binds.begin()->off);
      if (std::find(pendingZero.begin(), pendingZero.end(), f) == pendingZero.end()) pendingZero.push_back(f);
      return;
    }
    if (!pendingHalo.empty()) flushHalo();
    if (!pendingZero.empty()) flushZero();
    // column sweeps keep their elimination coefficients (dead scratch) in LDS
    static const char* const ldsCols[] = {"acoustic_w_solve", "turb_scalar_vert_implicit", "turb_u_tendency",
                                          "turb_v_tendency", "turb_w_tendency", "tke_transport"};
    std::string kn = kname;
    if (cfg.fuse != 0)
      for (const char* c : ldsCols)
        if (kn == c) kn += "_lds";
    Op o;
    std::memset(&o, 0, sizeof(o));
    o.kid = kernel_id(kn.c_str());
    o.A[0] = (uint32_t)grid[0]; o.A[1] = (uint32_t)grid[1]; o.A[2] = (uint32_t)grid[2];
    o.A[4 + 0] = (uint32_t)dimsOff;
    o.A[4 + 30] = (uint32_t)hmOff;
    if (withAcoustic) o.A[4 + 1] = (uint32_t)bytes(acoustic);
    for (const Bind& bd : binds) o.A[4 + bd.idx] = (uint32_t)bd.off;
    E.ops.push_back(o);
    E.names.push_back(kn);
  }
  Bind fb(int i, const std::string& n) const { return {i, off(n)}; }
  Bind rb(int i, const std::string& n) const { return {i, region(n)}; }
  Bind bb(int i, const Bytes& w) { return {i, bytes(w)}; }

  std::string species(const std::string& n, const std::string& suf) const {
    return std::find(scalars.begin(), scalars.end(), n) != scalars.end() ? n + suf : "qzero";
  }
  static bool isU(const std::string& n) { return n[0] == 'u' || n == "rru"; }
  static bool isV(const std::string& n) { return n[0] == 'v' || n == "rrv"; }
  float wenoEps(const std::string& n) const {
    const float eps = 1.0e-18f;
    if (n == "th" || n == "tke") return (float)(10.0 * (double)eps);
    if (n == "nci" || n == "ncs" || n == "ncr" || n == "ncg") return (float)(1.0e6 * (double)eps);
    return (float)((double)0.01f * (double)eps);
  }

  void setup();
  void openHalo(const std::string& n);
  void haloFill(const std::string& n);
  void openTurbFlux();
  void openStress();
  void bindSurfaceFlux(std::vector<Bind>& bs, int plane);
  void turbulence();
  void advect(int nrk, const std::string& s, int stagger, const std::string& into, bool accumulate,
              const std::string& state);
  void advectFluxes(int nrk, const std::string& s, int stagger, const std::string& target);
  void moistureTurbulence(int nrk, const std::string& target);
  void afterAssemble(int nrk, const std::string& s, int stagger, const std::string& into, bool accumulate,
                     const std::string& target, const std::array<int, 3>& asmSize, const Bytes& dttmp,
                     const Bytes& open);
  void stages();
  void end();
};

void Encoder::setup() {
  const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
  if (nx < 4 || ny < 4 || nz < 4) throw std::runtime_error("grid must be at least 4 x 4 x 4");
  if (cfg.adaptive_dt && !cfg.tke)
    throw std::runtime_error("adaptive_dt needs tke = 1 (ksmax comes from the TKE scheme's K, config.js:725-729)");
  if (!(cfg.dx > 0 && cfg.dy > 0 && cfg.dt > 0)) throw std::runtime_error("dx, dy and dt must be positive");
  moist = cfg.moist != 0;
  if (moist) {
    const int sp = cfg.species ? cfg.species : 3;
    const char* all[] = {"th", "qv", "qc", "qr", "qi", "qs", "qg", "nci", "ncs", "ncr", "ncg"};
    if (sp != 3 && sp != 7 && sp != 11) throw std::runtime_error("species must be 3, 7 or 11");
    scalars.assign(all, all + sp);
  } else {
    scalars = {"th"};
  }
  tke = cfg.tke != 0;
  rayleigh = cfg.rayleigh != 0;
  openX = cfg.open_x != 0; openY = cfg.open_y != 0;
  periodicX = !openX; periodicY = !openY;
  mergeHalo = cfg.fuse != 0 && periodicX && periodicY;
  foldHalo = mergeHalo && cfg.fuse == 1 && cfg.nx >= 4 && cfg.ny >= 4;
  wenoS = cfg.weno_scalars != 0;
  wenoM = cfg.weno_momentum != 0;
  explicitZ = cfg.explicit_vdiff != 0;
  diss = cfg.dissipative_heating != 0 && tke;
  cacheMT = cfg.cache_moisture_turbulence != 0;
  hasMicro = cfg.microphysics != 0;
  if (hasMicro && !moist) throw std::runtime_error("microphysics needs a moist configuration");
  massCons = cfg.mass_conservation != 0;
  dtf = cfg.dt;

  // vertical grid and profiles
  std::vector<double> F(nz + 1);
  for (int k = 0; k <= nz; ++k) F[k] = cfg.zf ? (double)cfg.zf[k] : k * (double)cfg.dz;
  v = make_vertical(F, cfg.dz > 0 ? cfg.dz : 0.0);
  b = make_base(cfg, v);
  pr = make_profiles(cfg, v, b);

  Layout& L = E.L;
  L.nx = nx; L.ny = ny; L.nz = nz; L.sJ = nx + 7; L.sK = ny + 7;
  const uint64_t count = (uint64_t)L.sJ * L.sK * (nz + 1);
  L.field_words = ((count * 4 + 255) / 256) * 256 / 4;
  auto put = [&](const std::string& n, const std::vector<float>& vals, size_t min_words) {
    uint64_t o = E.image.size();
    L.region[n] = o;
    size_t w = std::max(vals.size(), min_words);
    w = (w + 63) / 64 * 64;
    E.image.resize(o + w, 0u);
    std::memcpy(E.image.data() + o, vals.data(), vals.size() * 4);
  };
  const size_t pz = nz + 2;
  {
    Bytes d = {(uint32_t)nx, (uint32_t)ny, (uint32_t)nz, (uint32_t)L.sJ, (uint32_t)L.sK, fbits(1.0f / cfg.dx),
               fbits(1.0f / cfg.dy), fbits((float)(1.0 / v.dz_nominal))};
    std::vector<float> df(8);
    std::memcpy(df.data(), d.data(), 32);
    put("dims", df, 8);
    dimsOff = L.region["dims"];
  }
  std::vector<float> zf32(nz + 1);
  for (int k = 0; k <= nz; ++k) zf32[k] = (float)v.zf[k];
  put("c1", v.c1, pz); put("c2", v.c2, pz); put("mh", v.mh, pz); put("mf", v.mf, pz); put("zf", zf32, pz);
  put("rho0", pr.rho0, pz); put("rr0", pr.rr0, pz); put("rf0", pr.rf0, pz); put("th0", pr.th0, pz);
  put("pi0", pr.pi0, pz); put("rth0", pr.rth0, pz); put("qv0", pr.qv0, pz); put("qc0", pr.qc0, pz);
  put("qi0", pr.qi0, pz); put("tauh", pr.tauh, pz); put("tauf", pr.tauf, pz); put("u0", pr.u0, pz);
  put("v0", pr.v0, pz);
  {
    const HAxis hx = axis_x(cfg), hy = axis_y(cfg);
    // AcousticSchedule's minimum horizontal spacing (timestep.js:403)
    minSpacing = std::min(hx.min_spacing, hy.min_spacing);
    E.L.min_spacing = minSpacing;
    put("hmetric", make_metric(hx, hy), 0);
  }
  hmOff = L.region["hmetric"];
  put("radbcw", {}, std::max(ny, 1) + 8); put("radbce", {}, std::max(ny, 1) + 8);
  put("radbcs", {}, std::max(nx, 1) + 8); put("radbcn", {}, std::max(nx, 1) + 8);
  if (cfg.adaptive_dt) put("stability", {}, 2);  // TimeStep.stabilityb (timestep.js:706-709)
  put("sfc", {}, (size_t)11 * L.sJ * L.sK);
  if (massCons) {
    put("mass_rows", {}, (size_t)2 * L.sK * nz * 2 * 2);
    put("mass_levels", {}, (size_t)2 * nz * 2 * 2);
    put("mass_ref", {}, 4); put("mass_out", {}, 8);
  }

  // FieldArena (grid.js:322-346): fields sorted by byte order
  std::vector<std::string> fl;
  if (cfg.n_arena_fields > 0 && cfg.arena_fields) {
    for (int i = 0; i < cfg.n_arena_fields; ++i) fl.push_back(cfg.arena_fields[i]);
  } else {
    std::vector<std::string> adv = scalars;
    adv.insert(adv.end(), momentum.begin(), momentum.end());
    for (auto& s : adv) fl.push_back(s);
    for (auto& s : adv) fl.push_back(s + "_a");
    for (const char* n : {"ppterm", "piadv", "thv", "pp3d", "ppi", "ppd", "ppx", "uten", "vten", "wten", "ppten",
                          "buoy", "prs", "rho"})
      fl.push_back(n);
    if (moist) { fl.push_back("thterm"); fl.push_back("div3d"); }
    if (moist && usePT)
      for (const char* n : {"qpten", "qtten", "qvten", "qcten", "qvten_t"}) fl.push_back(n);
    if (tke)
      for (const char* n : {"tke", "tke_a", "tketen", "rr", "rf", "tketen_s", "s11", "s22", "s33", "s12", "s13", "s23",
                            "defv", "defh", "nm", "lenscl", "dissten", "epst", "kmh", "kmv", "khh", "khv", "dwdx"})
        fl.push_back(n);
    if (tke || rayleigh)
      for (const char* n : {"uten_t", "vten_t", "wten_t", "thten_t"}) fl.push_back(n);
    if (cacheMT && moist && tke)
      for (size_t q = 1; q < scalars.size(); ++q) fl.push_back("qturb_" + scalars[q]);
    if (moist) {
      for (const char* n : {"qv", "qc", "qr", "qi", "qs", "qg"})
        if (std::find(scalars.begin(), scalars.end(), n) == scalars.end()) { fl.push_back("qzero"); break; }
    }
    for (const char* n : {"rru", "rrv", "rrw", "fx", "fy", "fz", "tend", "divx", "qk", "pk1", "pk2", "mm", "fpk", "cA",
                          "cB", "sadv"})
      fl.push_back(n);
    std::sort(fl.begin(), fl.end());
    fl.erase(std::unique(fl.begin(), fl.end()), fl.end());
  }
  L.fields = fl;
  L.arena = (E.image.size() + 63) / 64 * 64;
  if (withArena) {
    E.image.resize(L.arena + L.field_words * fl.size(), 0u);
  } else {  // the image holds the words below the arena and the payloads above it
    E.image.resize(L.arena, 0u);
    E.skipped = L.field_words * fl.size();
  }
}

void Encoder::openHalo(const std::string& n) {
  const int H = 3;
  if (openX) op("bc_open_x", {H + 1, cfg.ny + 1 + 2 * H, cfg.nz + 1}, {fb(1, n), bb(2, I1(isU(n) ? 1 : 0))});
  if (openY) op("bc_open_y", {H + 1, cfg.nx + 1 + 2 * H, cfg.nz + 1}, {fb(1, n), bb(2, I1(isV(n) ? 1 : 0))});
}

void Encoder::haloFill(const std::string& n) {
  const int H = 3;
  if (mergeHalo) {
    if (!pendingAdd.empty()) flushAdd();
    if (!pendingZero.empty()) flushZero();
    if (std::find(pendingHalo.begin(), pendingHalo.end(), n) == pendingHalo.end()) pendingHalo.push_back(n);
    return;
  }
  if (periodicX) op("bc_periodic_x", {H + 1, cfg.ny + 1 + 2 * H, cfg.nz + 1}, {fb(1, n)});
  if (periodicY) op("bc_periodic_y", {H + 1, cfg.nx + 1 + 2 * H, cfg.nz + 1}, {fb(1, n)});
  openHalo(n);
}

void Encoder::openTurbFlux() {
  if (!openX && !openY) return;
  op("turb_openbc_flux", {std::max(cfg.nx, cfg.ny) + 1, cfg.nz + 1, 1},
     {fb(1, "fx"), fb(2, "fy"), bb(3, I2(openX ? 1 : 0, openY ? 1 : 0))});
}

void Encoder::openStress() {
  if (openX) op("turb_openbc_stress_x", {cfg.ny + 2, cfg.nz, 1}, {fb(1, "s11"), fb(2, "s12")});
  if (openY) op("turb_openbc_stress_y", {cfg.nx + 2, cfg.nz, 1}, {fb(1, "s22"), fb(2, "s12")});
}

// bindSurfaceFlux (timestep.js:550-564): no surface fluxes on this path.
void Encoder::bindSurfaceFlux(std::vector<Bind>& bs, int plane) {
  const uint64_t planeWords = (uint64_t)E.L.sJ * E.L.sK;
  bs.push_back({14, region("sfc") + (plane >= 0 ? (uint64_t)plane * planeWords : 0)});
  bs.push_back(bb(15, I1(0)));
  bs.push_back(bb(16, I1(explicitZ ? 1 : 0)));
  bs.push_back(bb(17, Fx4(0, 0, 0, 0)));
  bs.push_back(bb(18, I1(1)));
}

void Encoder::turbulence() {
  const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
  const std::array<int, 3> full{nx, ny, nz}, whole{nx + 7, ny + 7, nz + 1}, haloWide{nx + 2, ny + 2, nz + 1},
      stag{nx + 1, ny + 1, nz}, stagCols{nx + 1, ny + 1, 1}, cols{nx, ny, 1};
  // turb_zero of the per-step SGS fields (timestep.js:1070-1072)
  std::vector<std::string> zl;
  for (const char* n : {"tketen", "tketen_s", "s11", "s22", "s33", "s12", "s13", "s23", "defv", "defh", "nm", "lenscl",
                        "dissten", "epst", "kmh", "kmv", "khh", "khv", "dwdx"})
    zl.push_back(n);
  if (!rayleigh)
    for (const char* n : {"uten_t", "vten_t", "wten_t", "thten_t"}) zl.push_back(n);
  for (auto& n : zl) op("turb_zero", whole, {fb(1, n)});
  op("turb_rho_faces", {nx + 4, ny + 4, nz + 1},
     {fb(1, "rho"), rb(2, "c1"), rb(3, "c2"), fb(4, "rr"), fb(5, "rf"), bb(6, Fx3(v.cgs)), bb(7, Fx3(v.cgt))});
  op("turb_def_centers", haloWide,
     {fb(1, "u_a"), fb(2, "v_a"), fb(3, "w_a"), fb(4, "rho"), rb(5, "mh"), fb(6, "s11"), fb(7, "s22"), fb(8, "s33"),
      fb(9, "s12")});
  op("turb_def_faces", haloWide,
     {fb(1, "u_a"), fb(2, "v_a"), fb(3, "w_a"), fb(4, "rf"), rb(5, "mf"), fb(6, "s13"), fb(7, "s23")});
  op("turb_def_assemble", full,
     {fb(1, "s11"), fb(2, "s22"), fb(3, "s33"), fb(4, "s12"), fb(5, "s13"), fb(6, "s23"), rb(7, "c1"), rb(8, "c2"),
      fb(9, "rf"), fb(10, "defv"), fb(11, "defh")});
  openStress();
  {
    const char* ks[] = {"qv", "qc", "qr", "qi", "qs", "qg"};
    std::vector<Bind> bs = {fb(1, "th_a"), fb(2, "pp3d"), rb(3, "th0"), rb(4, "pi0"), fb(5, "prs")};
    for (int q = 0; q < 6; ++q) bs.push_back(fb(6 + q, moist ? species(ks[q], "_a") : "th_a"));
    bs.push_back(rb(12, "c1")); bs.push_back(rb(13, "c2")); bs.push_back(rb(14, "mf")); bs.push_back(fb(15, "nm"));
    bs.push_back(bb(16, I1(moist ? 1 : 0))); bs.push_back(fb(17, "nm")); bs.push_back(bb(18, I1(0)));
    Op o;
    op("turb_nm", full, {});
    o = E.ops.back();
    for (const Bind& bd : bs) o.A[4 + bd.idx] = (uint32_t)bd.off;
    E.ops.back() = o;
  }
  op(foldHalo ? "turb_km_kh_xy" : "turb_km_kh", full,
     {fb(1, "tke_a"), fb(2, "nm"), rb(3, "mf"), fb(4, "kmh"), fb(5, "kmv"), fb(6, "khh"), fb(7, "khv"), fb(8, "lenscl"),
      fb(9, "dissten"), bb(10, F1(1.0f)), bb(11, F1(0.0f)), bb(12, F1(1.0f))});
  if (!foldHalo)
    for (const char* n : {"kmh", "kmv", "khh", "khv"}) haloFill(n);
  if (diss) {
    const char* ks[] = {"qv", "qc", "qr", "qi", "qs", "qg"};
    const float dz1 = (float)(v.zf[1] - v.zf[0]);
    Bytes dp = {(uint32_t)(moist ? 1 : 0), 0u, fbits(1.0f), fbits(dz1), 0u};
    std::vector<Bind> bs = {fb(1, "dissten"), rb(2, "pi0"), fb(3, "ppi")};
    for (int q = 0; q < 6; ++q) bs.push_back(fb(4 + q, moist ? species(ks[q], "_a") : "th_a"));
    bs.push_back(fb(10, "thten_t")); bs.push_back(bb(11, dp)); bs.push_back(rb(12, "sfc"));
    op("turb_dissipative_heating", full, {});
    for (const Bind& bd : bs) E.ops.back().A[4 + bd.idx] = (uint32_t)bd.off;
  }
  op("turb_gettau", haloWide,
     {fb(1, "kmh"), fb(2, "kmv"), fb(3, "s11"), fb(4, "s22"), fb(5, "s33"), fb(6, "s12"), fb(7, "s13"), fb(8, "s23")});
  op("turb_sfcstress", stagCols,
     {rb(1, "sfc"), fb(2, "rf"), fb(3, "s13"), fb(4, "s23"), bb(5, I1(0)), fb(6, "u_a"), fb(7, "v_a"),
      bb(8, Fx2(cfg.umove, cfg.vmove)), fb(9, "kmv"), bb(10, I1(0)), rb(11, "mf")});
  for (const char* n : {"s11", "s22", "s33", "s12", "s13", "s23"}) haloFill(n);
  openStress();
  op("tke_epst", full,
     {fb(1, "u_a"), fb(2, "v_a"), fb(3, "w_a"), fb(4, "s11"), fb(5, "s12"), fb(6, "s13"), fb(7, "s22"), fb(8, "s23"),
      fb(9, "s33"), fb(10, "rf"), rb(11, "mh"), rb(12, "mf"), rb(13, "c1"), rb(14, "c2"), fb(15, "epst")});
  op("tke_sources", full, {fb(1, "khv"), fb(2, "nm"), fb(3, "dissten"), fb(4, "epst"), fb(5, "tketen")});
  const Bytes dtBig = F1(dtf), exZ = I1(explicitZ ? 1 : 0), open = I2(openX ? 1 : 0, openY ? 1 : 0);
  op("turb_u_tendency", stagCols,
     {fb(1, "u_a"), fb(2, "w_a"), fb(3, "rho"), fb(4, "rf"), fb(5, "kmv"), fb(6, "s11"), fb(7, "s12"), fb(8, "s13"),
      rb(9, "mh"), rb(10, "mf"), bb(11, dtBig), fb(12, "cA"), fb(13, "cB"), fb(14, "dwdx"), fb(15, "uten_t"),
      bb(16, open), bb(17, exZ)});
  op("turb_v_tendency", stagCols,
     {fb(1, "v_a"), fb(2, "w_a"), fb(3, "rho"), fb(4, "rf"), fb(5, "kmv"), fb(6, "s12"), fb(7, "s22"), fb(8, "s23"),
      rb(9, "mh"), rb(10, "mf"), bb(11, dtBig), fb(12, "cA"), fb(13, "cB"), fb(14, "dwdx"), fb(15, "vten_t"),
      bb(16, open), bb(17, exZ)});
  op("turb_w_tendency", cols,
     {fb(1, "w_a"), fb(2, "rho"), fb(3, "rf"), fb(4, "kmh"), fb(5, "s13"), fb(6, "s23"), fb(7, "s33"), rb(8, "mh"),
      rb(9, "mf"), bb(10, dtBig), fb(11, "cA"), fb(12, "cB"), fb(13, "wten_t"), bb(14, open), bb(15, exZ)});
  op("build_sadv", whole, {rb(1, "th0"), fb(2, "th_a"), fb(3, "sadv")});
  auto scalarTurb = [&](const std::string& src, const std::string& sink, int plane) {
    op("turb_scalar_flux_h", stag,
       {fb(1, src), fb(2, "rho"), fb(3, "khh"), bb(4, F1(1.0f)), fb(5, "fx"), fb(6, "fy")});
    openTurbFlux();
    std::vector<Bind> bs = {fb(1, src), fb(2, "rho"), fb(3, "rr"), fb(4, "rf"), fb(5, "khv"), rb(6, "mh"), rb(7, "mf"),
                            fb(8, "fx"), fb(9, "fy"), bb(10, dtBig), fb(11, "cA"), fb(12, "cB"), fb(13, sink)};
    bindSurfaceFlux(bs, plane);
    op("turb_scalar_vert_implicit", cols, {});
    for (const Bind& bd : bs) E.ops.back().A[4 + bd.idx] = (uint32_t)bd.off;
  };
  scalarTurb("sadv", "thten_t", 9);
  if (moist && usePT) {
    op("turb_zero", whole, {fb(1, "qvten_t")});
    scalarTurb("qv_a", "qvten_t", 10);
  }
  op("tke_flux_h", stag, {fb(1, "tke_a"), fb(2, "rf"), fb(3, "kmh"), fb(4, "fx"), fb(5, "fy")});
  openTurbFlux();
  op("tke_transport", stagCols,
     {fb(1, "tke_a"), fb(2, "rho"), fb(3, "rf"), fb(4, "kmh"), fb(5, "kmv"), rb(6, "mh"), rb(7, "mf"), bb(8, dtBig),
      fb(9, "cA"), fb(10, "cB"), fb(11, "fx"), fb(12, "fy"), fb(13, "tketen"), bb(14, open), bb(15, exZ)});
}

// advect (timestep.js:1629-1900): fluxes, assembly, pdef, SGS and RK update of
// one field.  stagger 1 = scalar, 2 = u, 3 = v, 4 = w / TKE.
void Encoder::advect(int nrk, const std::string& s, int stagger, const std::string& into, bool accumulate,
                     const std::string& state) {
  const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
  const std::array<int, 3> full{nx, ny, nz}, whole{nx + 7, ny + 7, nz + 1}, fullU{nx + 1, ny, nz}, fullV{nx, ny + 1, nz},
      fullW{nx, ny, nz + 1}, stag{nx + 1, ny + 1, nz}, cols{nx, ny, 1};
  const Bytes open = I2(openX ? 1 : 0, openY ? 1 : 0);
  const Bytes dttmp = F1((float)((double)dtf / (3 - nrk)));
  const std::string target = state.empty() ? s : state;
  const std::array<int, 3> asmSize{nx + (stagger == 2 ? 1 : 0), ny + (stagger == 3 ? 1 : 0), nz};
  const bool pdefAfter = !accumulate && stagger == 1 && moist && nrk == 2 && target != scalars[0];
  const bool wenoUVm = (stagger == 2 || stagger == 3) && wenoM && nrk == 2;
  if (cfg.fuse != 0 && !pdefAfter && !wenoUVm) {
    // fluxes + assembly in one launch (pdef needs the fluxes in memory; the
    // WENO momentum fluxes of u and v keep their own dispatches)
    int kind;
    Bytes weps = F1(0.0f), pdef = I1(0), pdw = I1(0);
    if (stagger == 1) {
      const bool isTheta = target == scalars[0];
      if (wenoS && nrk == 2) {
        kind = 1;
        weps = F1(wenoEps(target));
        pdef = I1(isTheta ? 0 : 1);
        pdw = I1((isTheta || target == "qv") ? 0 : 1);
      } else {
        kind = 0;
      }
    } else {
      const bool wenoTke = stagger == 4 && s == "tke" && wenoS && nrk == 2;
      const bool wenoW = stagger == 4 && s == "w" && wenoM && nrk == 2;
      if (wenoTke || wenoW) {
        kind = 3;
        weps = F1(wenoW ? (float)(100.0 * 1.0e-18) : wenoEps("tke"));
      } else {
        kind = 2;
      }
    }
    std::vector<Bind> bs = {bb(1, I1(stagger)), fb(2, s), bb(3, I1(kind)), fb(4, "rrw"), bb(5, weps), fb(6, "divx"),
                            rb(7, "rho0"), rb(8, "rf0"), rb(9, "mh"), rb(10, "mf"), fb(11, into),
                            bb(12, I1(accumulate ? 1 : 0)), fb(13, "rru"), fb(14, "rrv"), rb(15, "c1"), rb(16, "c2"),
                            bb(17, open), rb(18, "rr0"), fb(19, "tend"), bb(20, pdef), bb(21, pdw), fb(22, "fx"),
                            fb(23, "fy"), fb(24, "fz"), bb(27, F1(0.0f))};
    // A scalar's pointwise follow-ups (apply_thterm, the SGS / damping adds,
    // rk_update) join the launch when nothing else writes its tendency in
    // between: moisture SGS fluxes then go to their cache first (they read
    // no tendency), which needs the cached mode.  rk_update joins only when
    // the advected field is not the one it writes (theta advects sadv): other
    // workgroups of the launch still read the advected field's neighbours.
    const bool moistSgs = moist && tke && target != scalars[0];
    const bool chained = stagger == 1 && !accumulate && (!moistSgs || cacheMT) && !(moist && hasMicro && usePT);
    const bool update = chained && s != target;
    const bool tkeChained = stagger == 4 && s == "tke" && accumulate && into == "tketen_s";
    if (cfg.fuse == 1 && stagger == 4 && s == "w" && into == "wten") {
      // buoyancy_apply (after the advection of TKE, which touches neither
      // wten nor buoy) joins the assembly of w
      bs.push_back(bb(32, I1(64)));
      bs.push_back(fb(42, "buoy"));
    } else if (tkeChained) {
      // turb_zero(tketen_s), this assembly and turb_add_tendency(tketen) in
      // one launch; tke_advance (which writes the advected tke) follows
      bs.push_back(bb(32, I1(16 | 2)));
      bs.push_back(fb(35, "tketen"));
      bs.push_back(bb(36, F1(1.0f)));
    } else if (chained) {
      if (moistSgs) moistureTurbulence(nrk, target);
      int flags = 0;
      if (moist && target == scalars[0]) {
        flags |= 1;
        bs.push_back(fb(33, "div3d"));
        bs.push_back(fb(34, "thterm"));
      }
      std::string add;
      if ((tke || rayleigh) && target == scalars[0]) add = "thten_t";
      if (moistSgs) add = "qturb_" + target;
      if (!add.empty()) {
        flags |= 2;
        bs.push_back(fb(35, add));
        bs.push_back(bb(36, F1(1.0f)));
      }
      if (update) {
        flags |= 8 | (foldHalo ? 32 : 0);
        bs.push_back(fb(39, target + "_a"));
        bs.push_back(fb(40, target));
        bs.push_back(bb(41, dttmp));
      }
      bs.push_back(bb(32, I1(flags)));
    } else {
      bs.push_back(bb(32, I1(0)));
    }
    op("adv_fused", asmSize, {});
    for (const Bind& bd : bs) E.ops.back().A[4 + bd.idx] = (uint32_t)bd.off;
    if (chained) {
      if (!update)
        op(foldHalo ? "rk_update_xy" : "rk_update", asmSize,
           {bb(1, I1(stagger)), fb(2, target + "_a"), fb(3, "tend"), fb(4, target), bb(5, dttmp)});
      if (!foldHalo) haloFill(target);
      return;
    }
    if (tkeChained) return;
  } else {
    advectFluxes(nrk, s, stagger, target);
    std::vector<Bind> bs = {bb(1, I1(stagger)), fb(2, s), fb(3, "fx"), fb(4, "fy"), fb(5, "fz"), fb(6, "divx"),
                            rb(7, "rho0"), rb(8, "rf0"), rb(9, "mh"), rb(10, "mf"), fb(11, into),
                            bb(12, I1(accumulate ? 1 : 0)), fb(13, "rru"), fb(14, "rrv"), rb(15, "c1"), rb(16, "c2"),
                            bb(17, open), rb(18, "rr0"), fb(19, "tend"), bb(27, F1(0.0f))};
    op("adv_assemble", asmSize, {});
    for (const Bind& bd : bs) E.ops.back().A[4 + bd.idx] = (uint32_t)bd.off;
  }
  afterAssemble(nrk, s, stagger, into, accumulate, target, asmSize, dttmp, open);
}

void Encoder::advectFluxes(int nrk, const std::string& s, int stagger, const std::string& target) {
  const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
  const std::array<int, 3> fullU{nx + 1, ny, nz}, fullV{nx, ny + 1, nz}, fullW{nx, ny, nz + 1};
  if (stagger == 1) {
    const bool isTheta = target == scalars[0];
    if (wenoS && nrk == 2) {
      const Bytes weps = F1(wenoEps(target)), pdef = I1(isTheta ? 0 : 1), pdw = I1((isTheta || target == "qv") ? 0 : 1);
      op("adv_flux_x_weno", fullU, {fb(1, s), fb(2, "rru"), bb(3, weps), bb(4, pdef), fb(5, "fx")});
      op("adv_flux_y_weno", fullV, {fb(1, s), fb(2, "rrv"), bb(3, weps), bb(4, pdef), fb(5, "fy")});
      op("adv_flux_z_weno", fullW,
         {fb(1, s), fb(2, "rrw"), rb(3, "c1"), rb(4, "c2"), bb(5, weps), bb(6, pdef), bb(7, pdw), fb(8, "fz")});
    } else {
      op("adv_flux_x", fullU, {fb(1, s), fb(2, "rru"), fb(3, "fx")});
      op("adv_flux_y", fullV, {fb(1, s), fb(2, "rrv"), fb(3, "fy")});
      op("adv_flux_z", fullW, {fb(1, s), fb(2, "rrw"), rb(3, "c1"), rb(4, "c2"), fb(5, "fz")});
    }
  } else {
    const Bytes st = I1(stagger);
    const std::array<int, 3> stagLo{nx + 2, ny + 2, nz}, zg{nx + 1, ny + 1, nz + 1};
    const bool wenoTke = stagger == 4 && s == "tke" && wenoS && nrk == 2;
    const bool wenoW = stagger == 4 && s == "w" && wenoM && nrk == 2;
    const bool wenoUV = (stagger == 2 || stagger == 3) && wenoM && nrk == 2;
    if (wenoUV) {
      const Bytes weps = F1((float)(100.0 * 1.0e-18));
      op("adv_flux_x_weno_uv", stagLo, {bb(1, st), fb(2, s), fb(3, "rru"), bb(4, weps), fb(5, "fx")});
      op("adv_flux_y_weno_uv", stagLo, {bb(1, st), fb(2, s), fb(3, "rrv"), bb(4, weps), fb(5, "fy")});
      op("adv_flux_z_weno_uv", zg,
         {bb(1, st), fb(2, s), fb(3, "rrw"), rb(4, "c1"), rb(5, "c2"), bb(6, weps), fb(7, "fz")});
    } else if (wenoTke || wenoW) {
      const Bytes weps = F1(wenoW ? (float)(100.0 * 1.0e-18) : wenoEps("tke"));
      op("adv_flux_x_weno_w", stagLo, {fb(1, s), fb(2, "rru"), rb(3, "c2"), bb(4, weps), fb(5, "fx")});
      op("adv_flux_y_weno_w", stagLo, {fb(1, s), fb(2, "rrv"), rb(3, "c2"), bb(4, weps), fb(5, "fy")});
      op("adv_flux_z_weno_w", zg, {fb(1, s), fb(2, "rrw"), bb(3, weps), fb(4, "fz")});
    } else {
      for (const char* k : {"adv_flux_x_stag", "adv_flux_y_stag", "adv_flux_z_stag"}) {
        const bool z = k[9] == 'z';
        const std::string fl = k[9] == 'x' ? "fx" : (k[9] == 'y' ? "fy" : "fz");
        op(k, z ? zg : stagLo,
           {bb(1, st), fb(2, s), fb(3, "rru"), fb(4, "rrv"), fb(5, "rrw"), rb(6, "c1"), rb(7, "c2"), fb(8, fl)});
      }
    }
  }
}

// Moisture SGS fluxes of one scalar (timestep.js:1826-1860): into its cache
// (computed at the first stage only) or straight into tend.
void Encoder::moistureTurbulence(int nrk, const std::string& target) {
  const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
  const std::array<int, 3> whole{nx + 7, ny + 7, nz + 1}, stag{nx + 1, ny + 1, nz}, cols{nx, ny, 1};
  const std::string qa = target + "_a";
  const std::string sink = cacheMT ? "qturb_" + target : "tend";
  if (cacheMT && nrk == 0) op("turb_zero", whole, {fb(1, sink)});
  if (!cacheMT || nrk == 0) {
    op("turb_scalar_flux_h", stag, {fb(1, qa), fb(2, "rho"), fb(3, "khh"), bb(4, F1(1.0f)), fb(5, "fx"), fb(6, "fy")});
    openTurbFlux();
    std::vector<Bind> bs = {fb(1, qa), fb(2, "rho"), fb(3, "rr"), fb(4, "rf"), fb(5, "khv"), rb(6, "mh"),
                            rb(7, "mf"), fb(8, "fx"), fb(9, "fy"), bb(10, F1(dtf)), fb(11, "cA"), fb(12, "cB"),
                            fb(13, sink)};
    bindSurfaceFlux(bs, qa == "qv_a" ? 10 : -1);
    op("turb_scalar_vert_implicit", cols, {});
    for (const Bind& bd : bs) E.ops.back().A[4 + bd.idx] = (uint32_t)bd.off;
  }
}

void Encoder::afterAssemble(int nrk, const std::string& s, int stagger, const std::string& into, bool accumulate,
                            const std::string& target, const std::array<int, 3>& asmSize, const Bytes& dttmp,
                            const Bytes& open) {
  const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
  const std::array<int, 3> full{nx, ny, nz}, whole{nx + 7, ny + 7, nz + 1}, stag{nx + 1, ny + 1, nz}, cols{nx, ny, 1};
  if (accumulate) return;
  if (stagger == 1 && moist && nrk == 2 && target != scalars[0]) {
    const std::string s0 = target + "_a";
    const char* names[3][3] = {{"fx", "rru", "pk1"}, {"fy", "rrv", "pk2"}, {"fz", "rrw", "fpk"}};
    for (int dd = 0; dd < 3; ++dd)
      op(foldHalo && dd < 2 ? "pdef_dum_xy" : "pdef_dum", asmSize,
         {bb(1, I1(dd)), fb(2, s), fb(3, s0), fb(4, names[dd][0]), fb(5, names[dd][1]), rb(6, "rho0"), rb(7, "mh"),
          bb(8, dttmp), bb(9, open), fb(10, names[dd][2])});
    if (!foldHalo) {
      haloFill("pk1");
      haloFill("pk2");
    }
    op("pdef_apply", asmSize,
       {fb(1, s), fb(2, s0), fb(3, "fx"), fb(4, "fy"), fb(5, "fz"), fb(6, "rru"), fb(7, "rrv"), fb(8, "rrw"),
        fb(9, "pk1"), fb(10, "pk2"), fb(11, "fpk"), rb(12, "rho0"), rb(13, "mh"), bb(14, dttmp), bb(15, open),
        fb(16, "tend")});
  }
  if (moist && target == scalars[0]) op("apply_thterm", full, {fb(1, "div3d"), fb(2, "thterm"), fb(3, "tend")});
  if ((tke || rayleigh) && target == scalars[0])
    op("turb_add_tendency", whole, {fb(1, "thten_t"), fb(2, "tend"), bb(3, F1(1.0f))});
  if (moist && tke && target != scalars[0]) {
    moistureTurbulence(nrk, target);
    if (cacheMT) op("turb_add_tendency", whole, {fb(1, "qturb_" + target), fb(2, "tend"), bb(3, F1(1.0f))});
  }
  if (moist && hasMicro && usePT) {
    const char* src = target == "th" ? "qtten" : (target == "qv" ? "qvten" : (target == "qc" ? "qcten" : nullptr));
    if (src) op("turb_add_tendency", whole, {fb(1, src), fb(2, "tend"), bb(3, F1(1.0f))});
  }
  const bool fold = foldHalo && stagger == 1;
  op(fold ? "rk_update_xy" : "rk_update", asmSize,
     {bb(1, I1(stagger)), fb(2, target + "_a"), fb(3, "tend"), fb(4, target), bb(5, dttmp)});
  if (!fold) haloFill(target);
}

void Encoder::stages() {
  const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
  const std::array<int, 3> full{nx, ny, nz}, whole{nx + 7, ny + 7, nz + 1}, stag{nx + 1, ny + 1, nz}, cols{nx, ny, 1};
  const Bytes open = I2(openX ? 1 : 0, openY ? 1 : 0);
  const int ns = cfg.nsound > 0 ? cfg.nsound : schedule_nsound(dtf, minSpacing);
  const Schedule sch = make_schedule(ns, dtf);
  const char* ks[] = {"qv", "qc", "qr", "qi", "qs", "qg"};
  for (int nrk = 0; nrk < 3; ++nrk) {
    const Bytes dttmp = F1((float)((double)dtf / (3 - nrk)));
    if (moist) {
      std::vector<Bind> bs = {fb(1, scalars[0]), fb(2, "pp3d"), rb(3, "th0"), rb(4, "pi0"), rb(5, "qv0"),
                              rb(6, "qc0"), rb(7, "qi0"), rb(8, "rth0")};
      for (int q = 0; q < 6; ++q) bs.push_back(fb(9 + q, species(ks[q], "")));
      int i = 15;
      for (const char* n : {"thv", "ppterm", "piadv", "thterm", "buoy"}) bs.push_back(fb(i++, n));
      op("update_diagnostics_moist", whole, {});
      for (const Bind& bd : bs) E.ops.back().A[4 + bd.idx] = (uint32_t)bd.off;
    } else {
      op("update_diagnostics", whole,
         {fb(1, scalars[0]), fb(2, "pp3d"), rb(3, "th0"), rb(4, "pi0"), fb(5, "thv"), fb(6, "ppterm"), fb(7, "piadv"),
          rb(8, "rth0"), fb(9, "buoy")});
    }
    if (cfg.fuse != 0 && (tke || rayleigh)) {  // reset + add of the SGS / damping tendencies in one pass
      op("tendency_reset_add", whole,
         {fb(1, "uten"), fb(2, "vten"), fb(3, "wten"), fb(4, "uten_t"), fb(5, "vten_t"), fb(6, "wten_t"),
          bb(7, F1(1.0f))});
    } else {
      op("stage_tendency_reset", whole, {fb(1, "uten"), fb(2, "vten"), fb(3, "wten")});
      if (tke || rayleigh)
        for (const char* p : {"u", "v", "w"})
          op("turb_add_tendency", whole,
             {fb(1, std::string(p) + "ten_t"), fb(2, std::string(p) + "ten"), bb(3, F1(1.0f))});
    }
    auto computeDiv = [&] {
      op("div_massflux", whole, {fb(1, "rru"), fb(2, "rrv"), fb(3, "rrw"), rb(4, "mh"), fb(5, "divx")});
    };
    if (cfg.fuse == 1) {  // the mass fluxes and their divergence in one launch
      op("massflux_div", whole,
         {fb(1, "u"), fb(2, "v"), fb(3, "w"), rb(4, "rho0"), rb(5, "rf0"), fb(6, "rru"), fb(7, "rrv"), fb(8, "rrw"),
          rb(9, "mh"), fb(10, "divx")});
    } else {
      op("massflux_from_state", whole,
         {fb(1, "u"), fb(2, "v"), fb(3, "w"), rb(4, "rho0"), rb(5, "rf0"), fb(6, "rru"), fb(7, "rrv"), fb(8, "rrw")});
      computeDiv();
    }
    advect(nrk, "u", 2, "uten", true, "");
    advect(nrk, "v", 3, "vten", true, "");
    advect(nrk, "w", 4, "wten", true, "");
    if (tke) {
      if (cfg.fuse != 0) {  // zero, assembly and add in one launch
        advect(nrk, "tke", 4, "tketen_s", true, "");
        op(foldHalo ? "tke_advance_xy" : "tke_advance", full,
           {fb(1, "tke_a"), fb(2, "tketen_s"), bb(3, dttmp), fb(4, "tke")});
      } else {
        op("turb_zero", whole, {fb(1, "tketen_s")});
        advect(nrk, "tke", 4, "tketen_s", true, "");
        op("turb_add_tendency", whole, {fb(1, "tketen"), fb(2, "tketen_s"), bb(3, F1(1.0f))});
        op("tke_advance", full, {fb(1, "tke_a"), fb(2, "tketen_s"), bb(3, dttmp), fb(4, "tke")});
      }
      if (!foldHalo) haloFill("tke");
    }
    if (cfg.fuse != 1)  // fuse = 1: in the assembly of w (advect)
      op("buoyancy_apply", {nx, ny, std::max(nz - 1, 1)}, {fb(1, "buoy"), rb(2, "c1"), rb(3, "c2"), fb(4, "wten")});
    op("stage_state_reset", whole,
       {fb(1, "u_a"), fb(2, "v_a"), fb(3, "w_a"), fb(4, "ppi"), fb(5, "u"), fb(6, "v"), fb(7, "w"), fb(8, "pp3d"),
        fb(9, "ppd"), fb(10, "ppx")});
    acoustic = acousticBytes((double)(float)sch.dts[nrk]);
    // fuse = 1 on a periodic domain: after acoustic_uv, one ac_column launch
    // per substep (fpk, column solve, pp, accumulate, halo fill; prepare's
    // terms formed in it, begin folded into the stage's first substep)
    const bool column = cfg.fuse == 1 && !openX && !openY && nx >= 4 && ny >= 4 && nz >= 4 && (nz + 7) / 8 <= 16;
    if (!column) {
      op("acoustic_prepare", full,
         {fb(2, "ppterm"), fb(3, "piadv"), fb(4, "thv"), rb(5, "rr0"), rb(6, "rf0"), rb(7, "mh"), rb(8, "mf"),
          rb(9, "c1"), rb(10, "c2"), fb(11, "qk"), fb(12, "pk1"), fb(13, "pk2"), fb(14, "mm")},
         true);
      op("acoustic_begin", whole, {fb(2, "w"), fb(3, "rru"), fb(4, "rrv"), fb(5, "rrw")}, true);
    }
    const int nloop = sch.count[nrk];
    for (int n = 0; n < nloop; ++n) {
      if (openX || openY)
        op("openbc_uv_tendency", {std::max(nx, ny), nz, 1},
           {rb(2, "radbcw"), rb(3, "radbce"), rb(4, "radbcs"), rb(5, "radbcn"), fb(6, "uten"), fb(7, "vten"),
            fb(8, "u"), fb(9, "v"), bb(10, open)},
           true);
      op("acoustic_uv", stag,
         {fb(2, "ppd"), fb(3, "thv"), fb(4, "uten"), fb(5, "vten"), fb(6, "u"), fb(7, "v"), bb(8, open)}, true);
      if (column) {
        const float tavg = n == nloop - 1 ? (float)(1.0 / nloop) : 0.0f;
        op("ac_column", {nx, ny, 1},
           {fb(2, "u"), fb(3, "v"), fb(4, "w"), fb(5, "piadv"), fb(6, "ppten"), fb(7, "ppterm"), fb(8, "thv"),
            fb(9, "wten"), fb(10, "pp3d"), fb(11, "ppd"), fb(12, "ppx"), fb(13, "rru"), fb(14, "rrv"),
            fb(15, "rrw"), bb(16, I1(nrk == 2 ? 1 : 0)), bb(17, F1(tavg)), rb(18, "rr0"), rb(19, "rf0"),
            rb(20, "mh"), rb(21, "mf"), rb(22, "c1"), rb(23, "c2"), fb(24, "qk"), fb(25, "pk1"), fb(26, "pk2"),
            fb(27, "mm"), fb(28, "fpk"), fb(29, "cA"), fb(31, "cB"), rb(32, "rho0"), bb(33, I1(n == 0 ? 1 : 0)),
            bb(34, I1(foldHalo && n == nloop - 1 ? 1 : 0))},
           true);
        continue;
      }
      if (cfg.fuse != 0) {  // pp + accumulate in one pass; accumulate reads no ppd, so ppd's
                            // halo joins u, v, w's
        op("acoustic_fpk", full,
           {fb(2, "u"), fb(3, "v"), fb(4, "w"), fb(5, "piadv"), fb(6, "ppten"), fb(7, "ppterm"), fb(8, "pk1"),
            fb(9, "pk2"), fb(10, "qk"), fb(11, "fpk"), rb(15, "rho0"), rb(16, "rr0"), rb(17, "rf0")},
           true);
        op("acoustic_w_solve", cols,
           {fb(2, "pk1"), fb(3, "pk2"), fb(4, "qk"), fb(5, "mm"), fb(6, "fpk"), fb(7, "wten"), fb(8, "pp3d"),
            fb(9, "w"), fb(10, "cA"), fb(11, "cB")},
           true);
        const float tavg = n == nloop - 1 ? (float)(1.0 / nloop) : 0.0f;
        op("ac_pp_accumulate", whole,
           {fb(2, "w"), fb(3, "pk1"), fb(4, "pk2"), fb(5, "qk"), fb(6, "fpk"), fb(7, "pp3d"), fb(8, "ppd"),
            fb(9, "ppx"), bb(10, I1(nrk == 2 ? 1 : 0)), fb(11, "u"), fb(12, "v"), fb(13, "rru"), fb(14, "rrv"),
            fb(15, "rrw"), bb(16, F1(tavg))},
           true);
        for (const char* h : {"ppd", "u", "v", "w"}) haloFill(h);
        continue;
      }
      op("acoustic_fpk", full,
         {fb(2, "u"), fb(3, "v"), fb(4, "w"), fb(5, "piadv"), fb(6, "ppten"), fb(7, "ppterm"), fb(8, "pk1"),
          fb(9, "pk2"), fb(10, "qk"), fb(11, "fpk"), rb(15, "rho0"), rb(16, "rr0"), rb(17, "rf0")},
         true);
      op("acoustic_w_solve", cols,
         {fb(2, "pk1"), fb(3, "pk2"), fb(4, "qk"), fb(5, "mm"), fb(6, "fpk"), fb(7, "wten"), fb(8, "pp3d"), fb(9, "w"),
          fb(10, "cA"), fb(11, "cB")},
         true);
      op("acoustic_pp", full,
         {fb(2, "w"), fb(3, "pk1"), fb(4, "pk2"), fb(5, "qk"), fb(6, "fpk"), fb(7, "pp3d"), fb(8, "ppd"), fb(9, "ppx"),
          bb(10, I1(nrk == 2 ? 1 : 0))},
         true);
      haloFill("ppd");
      const float tavg = n == nloop - 1 ? (float)(1.0 / nloop) : 0.0f;
      op("acoustic_accumulate", whole,
         {fb(2, "u"), fb(3, "v"), fb(4, "w"), fb(5, "rru"), fb(6, "rrv"), fb(7, "rrw"), bb(8, F1(tavg))}, true);
      haloFill("u");
      haloFill("v");
      haloFill("w");
    }
    if (!(column && foldHalo)) haloFill("pp3d");
    if (openX)
      op("radbc_ew", {ny, 1, 1}, {fb(1, "u_a"), fb(2, "u"), bb(3, dttmp), rb(4, "radbcw"), rb(5, "radbce")});
    if (openY)
      op("radbc_ns", {nx, 1, 1}, {fb(1, "v_a"), fb(2, "v"), bb(3, dttmp), rb(4, "radbcs"), rb(5, "radbcn")});
    if (moist) op("divergence3d", full, {fb(1, "rru"), fb(2, "rrv"), fb(3, "rrw"), rb(4, "mh"), fb(5, "div3d")});
    op("velocity_to_massflux", whole, {rb(1, "rho0"), rb(2, "rf0"), fb(3, "rru"), fb(4, "rrv"), fb(5, "rrw")});
    computeDiv();
    for (size_t q = 0; q < scalars.size(); ++q) {
      if (q == 0) {
        op("build_sadv", whole, {rb(1, "th0"), fb(2, scalars[0]), fb(3, "sadv")});
        advect(nrk, "sadv", 1, "tend", false, scalars[0]);
      } else {
        advect(nrk, scalars[q], 1, "tend", false, "");
      }
    }
    if (nrk == 2) {
      if (moist && usePT)
        op("diabatic_unapply", full,
           {bb(1, F1(dtf)), fb(2, "pp3d"), fb(3, "th"), fb(4, "qv"), fb(5, "qc"), fb(6, "qpten"), fb(7, "qtten"),
            fb(8, "qvten"), fb(9, "qcten")});
      op("update_rho_prs", full,
         {bb(1, I1(moist ? 1 : 0)), fb(2, scalars[0]), fb(3, "pp3d"), rb(4, "th0"), rb(5, "pi0"),
          fb(6, moist ? "qv" : "th"), fb(7, "prs"), fb(8, "rho")});
      if (massCons) {
        // apmasscon call sites (timestep.js:776-794,2142-2161): encoded for
        // comparison with recorded tapes; these kernels are not on this path.
        const uint64_t rows = region("mass_rows"), rowWords = (uint64_t)E.L.sK * nz * 2;
        op("mass_row_sums", {ny, nz, 1},
           {fb(1, "rho"), rb(2, "pi0"), fb(3, "pp3d"), rb(4, "mh"), {5, rows}, {6, rows + rowWords}});
        const uint64_t lev = region("mass_levels");
        op("mass_level_sums", {nz, 1, 1}, {{1, rows}, {2, rows + rowWords}, {3, lev}, {4, lev + (uint64_t)nz * 2}});
        op("mass_reference_init", {1, 1, 1}, {{1, lev}, rb(2, "mass_ref")});
        op("mass_correction_apply", full,
           {bb(1, I1(moist ? 1 : 0)), {2, lev}, {3, lev + (uint64_t)nz * 2}, rb(4, "mass_ref"), fb(5, scalars[0]),
            fb(6, "pp3d"), rb(7, "th0"), rb(8, "pi0"), fb(9, moist ? "qv" : "th"), fb(10, "prs"), fb(11, "rho"),
            rb(12, "mass_out")});
      }
    }
  }
}

void Encoder::end() {
  const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
  const std::array<int, 3> full{nx, ny, nz};
  if (hasMicro) {
    // microphysics call site (timestep.js:2167-2177); not on this path
    op("mp_pdefq", {nx, ny, 1}, {fb(1, "rho"), fb(2, "qv"), bb(3, F1(0.0f))});
    op("microphysics", {nx, ny, 1}, {});
  }
  if (moist && usePT)
    op("diabatic_rates", full,
       {bb(1, F1(dtf)), fb(2, "pp3d"), fb(3, "th"), fb(4, "qv"), fb(5, "qc"), fb(6, "qpten"), fb(7, "qtten"),
        fb(8, "qvten"), fb(9, "qcten")});
  op(foldHalo ? "ppx_finish_xy" : "ppx_finish", full, {fb(1, "pp3d"), fb(2, "ppx")});
  if (!foldHalo) haloFill("ppx");
  flushAdd();
  flushHalo();
  flushZero();
  if (cfg.adaptive_dt) {  // encodeStability (timestep.js:742-764): the step's stability maxima
    const float d = dtf, rdx = 1.0f / cfg.dx, rdy = 1.0f / cfg.dy, rdz = (float)(1.0 / v.dz_nominal);
    op("stability_max", full,
       {bb(1, {fbits((0.5f * d) * rdx), fbits((0.5f * d) * rdy), fbits((0.5f * d) * rdz), fbits((d * rdx) * rdx),
               fbits((d * rdy) * rdy), 0u, tke ? 1u : 0u}),
        fb(2, "u"), fb(3, "v"), fb(4, "w"), rb(5, "mh"), fb(6, tke ? "kmh" : "u"), fb(7, tke ? "kmv" : "u"),
        fb(8, tke ? "khh" : "u"), rb(9, "stability")});
  }
  // state swap (timestep.js:766-774)
  const uint64_t count = (uint64_t)E.L.sJ * E.L.sK * (nz + 1);
  std::vector<std::pair<std::string, std::string>> pairs;
  for (auto& s : scalars) pairs.push_back({s, s + "_a"});
  for (auto& s : momentum) pairs.push_back({s, s + "_a"});
  pairs.push_back({"pp3d", "ppi"});
  if (tke) pairs.push_back({"tke", "tke_a"});
  if (cfg.fuse != 0) {  // up to 4 copies of equal length per copy_fields launch
    for (size_t s = 0; s < pairs.size(); s += 4) {
      const int n = (int)std::min<size_t>(4, pairs.size() - s);
      Op o;
      std::memset(&o, 0, sizeof(o));
      o.kid = kernel_id("copy_fields");
      o.A[0] = (uint32_t)count; o.A[1] = 1; o.A[2] = 1;
      o.A[4 + 0] = (uint32_t)dimsOff;
      for (int q = 0; q < n; ++q) {
        o.A[4 + 1 + q] = (uint32_t)off(pairs[s + q].first);
        o.A[4 + 5 + q] = (uint32_t)off(pairs[s + q].second);
      }
      o.A[4 + 9] = (uint32_t)bytes(I1(n));
      E.ops.push_back(o);
      E.names.push_back("copy_fields");
    }
    pairs.clear();
  }
  for (auto& p : pairs) {
    Op o;
    std::memset(&o, 0, sizeof(o));
    o.kid = kernel_id("_copy");
    o.A[0] = (uint32_t)count; o.A[1] = 1; o.A[2] = 1;
    o.A[4] = (uint32_t)off(p.first);
    o.A[5] = (uint32_t)off(p.second);
    E.ops.push_back(o);
    E.names.push_back("_copy");
  }
}

Encoded Encoder::run() {
  setup();
  if (rayleigh) {
    op("rayleigh_damping", {cfg.nx + 1, cfg.ny + 1, cfg.nz + 1},
       {fb(1, "u_a"), fb(2, "v_a"), fb(3, "w_a"), fb(4, "th_a"), rb(5, "u0"), rb(6, "v0"), rb(7, "tauh"), rb(8, "tauf"),
        bb(9, F1(cfg.rayleigh_alpha)), fb(10, "uten_t"), fb(11, "vten_t"), fb(12, "wten_t"), fb(13, "thten_t"),
        rb(14, "tauf"), rb(15, "tauf"), bb(16, I1(0)), bb(17, I1(0)), rb(18, "tauf"), rb(19, "th0")});
  }
  if (tke) turbulence();
  {
    const bool hasThten = tke || rayleigh;
    op("pressure_tendency", {cfg.nx, cfg.ny, cfg.nz},
       {fb(1, "th_a"), fb(2, moist ? "qv_a" : "th_a"), fb(3, "rho"), fb(4, "ppi"), fb(5, hasThten ? "thten_t" : "ppten"),
        fb(6, moist ? "qpten" : "ppten"), rb(7, "th0"), rb(8, "pi0"), bb(9, F1(dtf)),
        bb(10, I1(moist && hasMicro ? 1 : 0)), fb(11, "ppten"), bb(12, I1(hasThten ? 1 : 0)),
        fb(13, moist ? "qvten_t" : "ppten"), bb(14, I1(moist ? 1 : 0))});
  }
  stages();
  end();
  E.L.words = (E.image.size() + E.skipped + 63) / 64 * 64;
  E.image.resize(E.L.words - E.skipped, 0u);
  return std::move(E);
}

}  // namespace

Encoded encode_step(const te_config& cfg, bool with_arena) {
  Encoder e(cfg);
  e.withArena = with_arena;
  return e.run();
}

void fill_initial_state(const te_config& cfg, const Layout& L, std::vector<uint32_t>& image) {
  std::vector<double> F(cfg.nz + 1);
  for (int k = 0; k <= cfg.nz; ++k) F[k] = cfg.zf ? (double)cfg.zf[k] : k * (double)cfg.dz;
  Vertical v = make_vertical(F, cfg.dz > 0 ? cfg.dz : 0.0);
  Base b = make_base(cfg, v);
  std::fill(image.begin() + L.arena, image.begin() + L.arena + L.field_words * L.fields.size(), 0u);
  fill_state(cfg, v, b, L, image.data());
}

}  // namespace te

extern "C" void te_config_defaults(te_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->nx = 64; c->ny = 64; c->nz = 40;
  c->dx = 1000.0f; c->dy = 1000.0f; c->dz = 500.0f;
  c->dt = 6.0f; c->nsound = 6;
  c->moist = 1; c->species = 3;
  c->weno_scalars = 1;
  c->rayleigh_zd = 15000.0f; c->rayleigh_alpha = (float)3.3333333333e-3;
  c->alph = 0.6f; c->kdiv = 0.1f; c->smeps = 0.0f;
  c->th_sfc = 300.0f; c->th_trop = 343.0f; c->t_trop = 213.0f; c->z_trop = 12000.0f; c->qv_pbl = 0.014f;
  c->hodo_radius = 10.0f; c->hodo_depth = 2000.0f; c->shear_umax = 30.0f; c->shear_depth = 6000.0f;
  c->bubble_amp = 2.0f; c->bubble_rh = 10000.0f; c->bubble_rv = 1400.0f; c->bubble_zc = 1400.0f;
  c->noise_sigma = 0.01f; c->seed = 20260101ull;
  c->fuse = 1;
  // SupercellScenario defaults (scenario.js:67-71): Rayleigh damping above
  // 15 km and moisture SGS fluxes computed once per step
  c->rayleigh = 1;
  c->cache_moisture_turbulence = 1;
}
