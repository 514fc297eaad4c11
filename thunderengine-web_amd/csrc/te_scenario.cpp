// Grid, base state, step profiles and the t=0 arena of a configuration.
//
// Vertical grid and step profiles follow the reference's binary32 recipes
// (web/core/grid.js:218-293, web/core/scenario.js:122-136) because their values
// become kernel inputs.  The sounding (web/core/base.js:181-216 Weisman-Klemp,
// :368-442 winds, :475-489 bubble) is this library's own construction of the
// same analytic profiles: an initial state, not part of the per-step parity.
#include <algorithm>
#include <cmath>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "te_scenario.hpp"

namespace te {

static inline float f32(double x) { return (float)x; }

// Fornberg weights for 3-point extrapolation to x0 (grid.js:219-235).
static void fornberg(float x0f, const float af[3], float out[3]) {
  double x0 = x0f, alpha[3] = {af[0], af[1], af[2]};
  double d[3][3] = {{1, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double b1 = 1.0;
  for (int n = 1; n <= 2; ++n) {
    double b2 = 1.0;
    for (int nu = 0; nu < n; ++nu) {
      double b3 = alpha[n] - alpha[nu];
      b2 = b2 * b3;
      d[n][nu] = ((alpha[n] - x0) * d[n - 1][nu]) / b3;
    }
    d[n][n] = (b1 / b2) * (-((alpha[n - 1] - x0) * d[n - 1][n - 1]));
    b1 = b2;
  }
  for (int q = 0; q < 3; ++q) out[q] = f32(d[2][q]);
}
static void lagrange(const double* x, int n, double x0, float out[3]) {
  double w[3] = {0, 0, 0};
  for (int i = 0; i < n; ++i) {
    double p = 1.0;
    for (int j = 0; j < n; ++j)
      if (j != i) p *= (x0 - x[j]) / (x[i] - x[j]);
    w[i] = p;
  }
  for (int q = 0; q < 3; ++q) out[q] = f32(w[q]);
}

Vertical make_vertical(const std::vector<double>& F, double nominal) {
  const int n = (int)F.size() - 1;
  if (n < 2) throw std::runtime_error("vertical grid needs at least two cells");
  for (int k = 0; k < n; ++k)
    if (!(F[k] < F[k + 1])) throw std::runtime_error("face heights must increase strictly");
  Vertical v;
  v.nz = n;
  v.zf = F;
  v.zh.resize(n);
  for (int k = 0; k < n; ++k) v.zh[k] = f32(0.5 * (F[k] + F[k + 1]));
  v.dz_nominal = nominal > 0 ? nominal : (F[n] - F[0]) / n;
  const float* c = v.zh.data();
  const float ghostLo = -c[0];
  const float ghostHi = c[n - 1] + 2.0f * (f32(F[n]) - c[n - 1]);
  const float dz32 = f32(v.dz_nominal);
  v.mf.assign(n + 1, 0.0f);
  v.mf[0] = dz32 / (c[0] - ghostLo);
  for (int k = 1; k < n; ++k) v.mf[k] = dz32 / (c[k] - c[k - 1]);
  v.mf[n] = dz32 / (ghostHi - c[n - 1]);
  v.c2.assign(n + 1, 0.0f);
  v.c2[0] = (f32(F[0]) - ghostLo) / (c[0] - ghostLo);
  for (int k = 1; k < n; ++k) v.c2[k] = (f32(F[k]) - c[k - 1]) / (c[k] - c[k - 1]);
  v.c2[n] = (f32(F[n]) - c[n - 1]) / (ghostHi - c[n - 1]);
  v.c1.resize(n + 1);
  for (int k = 0; k <= n; ++k) v.c1[k] = f32(1.0 - (double)v.c2[k]);
  v.mh.resize(n);
  for (int k = 0; k < n; ++k) v.mh[k] = dz32 / (f32(F[k + 1]) - f32(F[k]));
  if (n >= 3) {
    const float dBot = f32(F[1]) - f32(F[0]), dTop = f32(F[n]) - f32(F[n - 1]);
    const float a[3] = {c[0] / dBot, c[1] / dBot, c[2] / dBot};
    fornberg(f32(F[0]) / dBot, a, v.cgs);
    const float b[3] = {c[n - 1] / dTop, c[n - 2] / dTop, c[n - 3] / dTop};
    fornberg(f32(F[n]) / dTop, b, v.cgt);
  } else {
    double lo[3] = {v.zh[0], n > 1 ? v.zh[1] : 0.0, 0.0}, hi[3] = {v.zh[n - 1], n > 1 ? v.zh[n - 2] : 0.0, 0.0};
    lagrange(lo, std::min(3, n), F[0], v.cgs);
    lagrange(hi, std::min(3, n), F[n], v.cgt);
  }
  return v;
}

// ---- thermodynamic base state (constants of web/core/grid.js:13-20)
static const double G = 9.81, RDd = 287.04, CPd = 1005.7, RVd = 461.5, P00d = 1.0e5;

static float rslf(float p, float t, float eps) {
  float esl = 611.2f * f32(std::exp((double)((17.67f * (t - 273.15f)) / (t - 29.65f))));
  esl = std::min(esl, p * 0.5f);
  return (eps * esl) / (p - esl);
}

static double normal(uint64_t seed, uint64_t cell);

// Perturbed sounding (ensemble members): theta'(z) = sigma * N(0,1) white
// noise over the levels, smoothed by a Gaussian of 1 km width and rescaled
// to unit variance; one N(0,1) draw scales qv.  Draws are keyed by
// sounding_seed (cells 0..nz-1 for theta, nz for qv).
static std::vector<double> theta_perturbation(const te_config& cfg, const Vertical& v) {
  const int nk = v.nz;
  std::vector<double> wn(nk), out(nk, 0.0);
  if (!(cfg.th_pert_sigma > 0.0f)) return out;
  // unit variance: each smoothed value is a unit-norm combination of N(0,1)
  for (int k = 0; k < nk; ++k) wn[k] = normal(cfg.sounding_seed, (uint64_t)k);
  for (int k = 0; k < nk; ++k) {
    double s = 0.0, wsum = 0.0;
    for (int q = 0; q < nk; ++q) {
      const double d = (v.zh[k] - v.zh[q]) / 1000.0;
      const double wt = std::exp(-0.5 * d * d);
      s += wt * wn[q];
      wsum += wt * wt;
    }
    out[k] = s / std::sqrt(wsum);
  }
  for (int k = 0; k < nk; ++k) out[k] *= cfg.th_pert_sigma;
  return out;
}

// ---- caller's sounding (isnd = 7 / 17, CM1's input_sounding file)
static float powf_(float x, float y) { return f32(std::pow((double)x, (double)y)); }

// One token of the file as a binary32 (parseFloat32, base.js:15-52: the
// decimal string rounded once to the nearest float; strtof rounds once);
// false for what is not a decimal number.
static bool token_f32(const std::string& t, float* out) {
  size_t i = 0, n = t.size();
  if (i < n && (t[i] == '+' || t[i] == '-')) ++i;
  size_t d0 = i;
  while (i < n && std::isdigit((unsigned char)t[i])) ++i;
  size_t dig = i - d0;
  if (i < n && t[i] == '.') {
    ++i;
    size_t f0 = i;
    while (i < n && std::isdigit((unsigned char)t[i])) ++i;
    dig += i - f0;
  }
  if (dig == 0) return false;
  if (i < n && (t[i] == 'e' || t[i] == 'E')) {
    ++i;
    if (i < n && (t[i] == '+' || t[i] == '-')) ++i;
    size_t e0 = i;
    while (i < n && std::isdigit((unsigned char)t[i])) ++i;
    if (i == e0) return false;
  }
  if (i != n) return false;
  *out = std::strtof(t.c_str(), nullptr);
  return true;
}

// Sounding (base.js:659-714): header p_sfc (mb), theta_sfc (K), qv_sfc (g/kg);
// rows z (m), theta (K), qv (g/kg) and, with winds, u v (m/s).  A surface
// level with winds extrapolated from the first two rows is put below them;
// pi is integrated hydrostatically with the virtual temperature, and levels
// drier than 1e-12 kg/kg get 0.1 % of saturation.
struct SoundingLevels {
  std::vector<float> z, th, qv, u, v, pi, prs, t, rh;
  float pS = 0, thS = 0, qvS = 0;
};
static SoundingLevels parse_sounding(const char* text, bool winds, bool moist) {
  const float rd = f32(RDd), cp = f32(CPd), rv = f32(RVd), p00 = f32(P00d), g = f32(G);
  const float reps = rv / rd, eps = rd / rv;
  std::vector<std::vector<float>> rows;
  std::string all(text), line;
  for (char& ch : all)
    if (ch == '\r') ch = '\n';
  size_t pos = 0;
  while (pos <= all.size()) {
    size_t e = all.find('\n', pos);
    if (e == std::string::npos) e = all.size();
    line = all.substr(pos, e - pos);
    pos = e + 1;
    std::vector<float> nums;
    size_t q = 0;
    bool any = false;
    while (q < line.size()) {
      while (q < line.size() && (line[q] == ' ' || line[q] == '\t')) ++q;
      size_t b = q;
      while (q < line.size() && line[q] != ' ' && line[q] != '\t') ++q;
      if (q > b) {
        any = true;
        float x;
        if (token_f32(line.substr(b, q - b), &x)) nums.push_back(x);
      }
    }
    if (any) rows.push_back(nums);
  }
  if (rows.empty()) throw std::runtime_error("input_sounding is empty");
  const std::vector<float>& h = rows[0];
  if (h.size() < 3) throw std::runtime_error("input_sounding header: expected p_sfc (mb), theta_sfc (K), qv_sfc (g/kg)");
  SoundingLevels S;
  S.pS = f32((double)h[0] * 100.0);
  S.thS = h[1];
  S.qvS = moist ? f32((double)h[2] / 1000.0) : 0.0f;
  const size_t need = winds ? 5 : 3;
  std::vector<float> z, th, qv, u, vv;
  for (size_t r = 1; r < rows.size(); ++r) {
    if (rows[r].size() < need)
      throw std::runtime_error("input_sounding line " + std::to_string(r + 1) + ": expected z, theta, qv" +
                               (winds ? ", u, v" : ""));
    z.push_back(rows[r][0]);
    th.push_back(rows[r][1]);
    qv.push_back(moist ? f32((double)rows[r][2] / 1000.0) : 0.0f);
    u.push_back(winds ? rows[r][3] : 0.0f);
    vv.push_back(winds ? rows[r][4] : 0.0f);
  }
  if (z.size() < 2) throw std::runtime_error("input_sounding needs at least 2 levels (base.F:627-634)");
  for (size_t i = 1; i < z.size(); ++i)
    if (!(z[i] > z[i - 1])) throw std::runtime_error("input_sounding heights must increase");
  if (!(z[0] > 0.0f)) throw std::runtime_error("input_sounding heights must start above the surface");
  const float dz = z[1] - z[0];
  const float uS = dz != 0.0f ? u[0] - ((z[0] - 0.0f) * (u[1] - u[0])) / dz : u[0];
  const float vS = dz != 0.0f ? vv[0] - ((z[0] - 0.0f) * (vv[1] - vv[0])) / dz : vv[0];
  S.z.push_back(0.0f); S.th.push_back(S.thS); S.qv.push_back(S.qvS); S.u.push_back(uS); S.v.push_back(vS);
  S.z.insert(S.z.end(), z.begin(), z.end());
  S.th.insert(S.th.end(), th.begin(), th.end());
  S.qv.insert(S.qv.end(), qv.begin(), qv.end());
  S.u.insert(S.u.end(), u.begin(), u.end());
  S.v.insert(S.v.end(), vv.begin(), vv.end());
  const size_t n = S.z.size();
  std::vector<float> thv(n);
  for (size_t k = 0; k < n; ++k) thv[k] = (S.th[k] * std::fma(reps, S.qv[k], 1.0f)) / (1.0f + S.qv[k]);
  S.pi.assign(n, 0.0f);
  S.pi[0] = powf_(S.pS / p00, rd / cp);
  for (size_t k = 1; k < n; ++k)
    S.pi[k] = S.pi[k - 1] - (g * (S.z[k] - S.z[k - 1])) / ((cp * 0.5f) * (thv[k] + thv[k - 1]));
  const float cpOverRd = cp / rd;
  S.prs.resize(n); S.t.resize(n); S.rh.assign(n, 0.0f);
  for (size_t k = 0; k < n; ++k) {
    S.prs[k] = p00 * powf_(S.pi[k], cpOverRd);
    S.t[k] = S.th[k] * S.pi[k];
  }
  if (moist)
    for (size_t k = 0; k < n; ++k) {
      if (S.qv[k] < 1.0e-12f) S.qv[k] = 0.001f * rslf(S.prs[k], S.th[k] * S.pi[k], eps);
      S.rh[k] = S.qv[k] / rslf(S.prs[k], S.t[k], eps);
    }
  return S;
}

// Sounding.interpolated (base.js:716-757) and BaseState.initProfiles / tail
// (base.js:148-178): theta, u, v, p, T and relative humidity interpolated
// linearly to the scalar levels, qv from the humidity at the interpolated
// p and T, pi integrated up from the sounding's surface, winds relative to
// the domain's motion.
static Base sounding_base(const te_config& cfg, const Vertical& v) {
  const float rd = f32(RDd), cp = f32(CPd), rv = f32(RVd), p00 = f32(P00d), g = f32(G);
  const float reps = rv / rd, eps = rd / rv, cpdrd = cp / rd;
  const bool moist = cfg.moist != 0;
  const SoundingLevels S = parse_sounding(cfg.input_sounding, cfg.sounding_winds != 0, moist);
  const int nk = v.nz;
  if (!(S.z.back() >= v.zh[nk - 1]))
    throw std::runtime_error("input_sounding reaches " + std::to_string(S.z.back()) + " m, below the top scalar level " +
                             std::to_string(v.zh[nk - 1]) + " m (base.F:734-739)");
  Base b;
  b.th0.assign(nk, 0); b.qv0.assign(nk, 0); b.qc0.assign(nk, 0); b.pi0.assign(nk, 0); b.prs0.assign(nk, 0);
  b.rho0.assign(nk, 0); b.u0.assign(nk, 0); b.v0.assign(nk, 0);
  std::vector<float> pI(nk), tI(nk), rhI(nk), thv(nk);
  for (int k = 0; k < nk; ++k) {
    const float z = v.zh[k];
    size_t kk = 1;
    while (S.z[kk] < z) ++kk;
    const size_t dn = kk - 1, up = kk;
    const float fr = (z - S.z[dn]) / (S.z[up] - S.z[dn]);
    auto lerp = [&](const std::vector<float>& a) { return std::fma(a[up] - a[dn], fr, a[dn]); };
    b.th0[k] = lerp(S.th); b.u0[k] = lerp(S.u); b.v0[k] = lerp(S.v);
    pI[k] = lerp(S.prs); tI[k] = lerp(S.t); rhI[k] = lerp(S.rh);
  }
  for (int k = 0; k < nk; ++k) {
    b.qv0[k] = moist ? rhI[k] * rslf(pI[k], tI[k], eps) : 0.0f;
    thv[k] = (b.th0[k] * std::fma(reps, b.qv0[k], 1.0f)) / (1.0f + b.qv0[k]);
  }
  const float piSfc = powf_(S.pS / p00, rd / cp);
  const float thvSfc = (S.thS * std::fma(S.qvS, reps, 1.0f)) / (1.0f + S.qvS);
  b.pi0[0] = piSfc - (g * v.zh[0]) / ((cp * 0.5f) * (thvSfc + thv[0]));
  for (int k = 1; k < nk; ++k)
    b.pi0[k] = b.pi0[k - 1] - (g * (v.zh[k] - v.zh[k - 1])) / ((cp * 0.5f) * (thv[k] + thv[k - 1]));
  for (int k = 0; k < nk; ++k) {
    b.prs0[k] = p00 * powf_(b.pi0[k], cpdrd);
    b.rho0[k] = b.prs0[k] / (((rd * b.th0[k]) * b.pi0[k]) * std::fma(b.qv0[k], reps, 1.0f));
    b.u0[k] = b.u0[k] - cfg.umove;
    b.v0[k] = b.v0[k] - cfg.vmove;
  }
  return b;
}

Base make_base(const te_config& cfg, const Vertical& v) {
  if (cfg.input_sounding) return sounding_base(cfg, v);
  const int nk = v.nz;
  const float rd = f32(RDd), cp = f32(CPd), rv = f32(RVd), p00 = f32(P00d), g = f32(G);
  const float reps = rv / rd, eps = rd / rv, cpdrd = cp / rd;
  Base b;
  b.th0.assign(nk, 0); b.qv0.assign(nk, 0); b.qc0.assign(nk, 0); b.pi0.assign(nk, 0); b.prs0.assign(nk, 0);
  b.rho0.assign(nk, 0); b.u0.assign(nk, 0); b.v0.assign(nk, 0);
  std::vector<float> rh(nk, 0.0f), thv(nk, 0.0f);
  // Weisman & Klemp (1982) analytic sounding
  const float zTrop = cfg.z_trop, thTrop = cfg.th_trop, tTrop = cfg.t_trop, thSfc = cfg.th_sfc, prsSfc = 100000.0f;
  const float piSfc = f32(std::pow((double)(prsSfc / p00), (double)(rd / cp)));
  const float qvSfc = cfg.moist ? rslf(prsSfc, thSfc * piSfc, eps) : 0.0f;
  const float thvSfc = (thSfc * std::fma(qvSfc, reps, 1.0f)) / (1.0f + qvSfc);
  for (int k = 0; k < nk; ++k) {
    const float z = v.zh[k];
    if (z < zTrop) {
      const float fr = f32(std::pow((double)(z / zTrop), 1.25));
      b.th0[k] = std::fma(thTrop - thSfc, fr, thSfc);
      rh[k] = cfg.moist ? std::fma(-0.75f, fr, 1.0f) : 0.0f;
    } else {
      b.th0[k] = thTrop * f32(std::exp((double)((g / (tTrop * cp)) * (z - zTrop))));
      rh[k] = cfg.moist ? 0.25f : 0.0f;
    }
  }
  const std::vector<double> dth = theta_perturbation(cfg, v);
  for (int k = 0; k < nk; ++k) b.th0[k] = f32(b.th0[k] + dth[k]);
  const float qvScale = cfg.qv_pert_sigma > 0.0f
                            ? f32(1.0 + cfg.qv_pert_sigma * normal(cfg.sounding_seed, (uint64_t)nk)) : 1.0f;
  for (int it = 0; it < 20; ++it) {
    for (int k = 0; k < nk; ++k) thv[k] = (b.th0[k] * std::fma(reps, b.qv0[k], 1.0f)) / (1.0f + b.qv0[k]);
    b.pi0[0] = piSfc - (g * v.zh[0]) / ((cp * 0.5f) * (thvSfc + thv[0]));
    for (int k = 1; k < nk; ++k)
      b.pi0[k] = b.pi0[k - 1] - (g * (v.zh[k] - v.zh[k - 1])) / ((cp * 0.5f) * (thv[k] + thv[k - 1]));
    for (int k = 0; k < nk; ++k) b.prs0[k] = p00 * f32(std::pow((double)b.pi0[k], (double)cpdrd));
    for (int k = 0; k < nk; ++k) {
      b.qv0[k] = cfg.moist ? (rh[k] * rslf(b.prs0[k], b.th0[k] * b.pi0[k], eps)) * qvScale : 0.0f;
      if (b.qv0[k] > cfg.qv_pbl) b.qv0[k] = cfg.qv_pbl;
    }
  }
  for (int k = 0; k < nk; ++k) {
    b.prs0[k] = p00 * f32(std::pow((double)b.pi0[k], (double)cpdrd));
    b.rho0[k] = b.prs0[k] / (((rd * b.th0[k]) * b.pi0[k]) * std::fma(b.qv0[k], reps, 1.0f));
  }
  // quarter-circle hodograph of radius R over 0..zq, then straight to umax2 at zs
  const float R = cfg.hodo_radius, zq = cfg.hodo_depth, umax2 = cfg.shear_umax, zs = cfg.shear_depth;
  const double deg = M_PI / 180.0;
  for (int k = 0; k < nk; ++k) {
    const float z = v.zh[k];
    float u, w;
    if (z <= zq) {
      const double ang = 90.0 * (z / zq) * deg;
      u = f32(R - R * std::cos(ang));
      w = f32(R * std::sin(ang));
    } else if (z <= zs) {
      u = R + ((z - zq) * (umax2 - R)) / (zs - zq);
      w = R;
    } else {
      u = umax2;
      w = R;
    }
    b.u0[k] = u - cfg.umove;
    b.v0[k] = w - cfg.vmove;
  }
  return b;
}

Profiles make_profiles(const te_config& cfg, const Vertical& v, const Base& b) {
  const int nz = v.nz;
  Profiles p;
  p.rho0 = b.rho0;
  p.rr0.resize(nz);
  for (int k = 0; k < nz; ++k) p.rr0[k] = 1.0f / b.rho0[k];
  // rf0 at faces by extrapolation/interpolation (scenario.js:122-136)
  p.rf0.resize(nz + 1);
  const float* r = b.rho0.data();
  for (int k = 0; k <= nz; ++k) {
    if (k == 0) p.rf0[k] = std::fma(v.cgs[2], r[2], std::fma(v.cgs[0], r[0], v.cgs[1] * r[1]));
    else if (k == nz) p.rf0[k] = std::fma(v.cgt[2], r[nz - 3], std::fma(v.cgt[0], r[nz - 1], v.cgt[1] * r[nz - 2]));
    else p.rf0[k] = std::fma(v.c1[k], r[k - 1], v.c2[k] * r[k]);
  }
  p.th0 = b.th0; p.pi0 = b.pi0; p.qv0 = b.qv0; p.qc0 = b.qc0; p.qi0.assign(nz, 0.0f);
  p.rth0.resize(nz);
  for (int k = 0; k < nz; ++k) p.rth0[k] = 1.0f / b.th0[k];
  p.u0 = b.u0; p.v0 = b.v0;
  // Rayleigh damping weights above zd (timestep.js:657-677)
  p.tauh.assign(nz, 0.0f);
  p.tauf.assign(nz + 1, 0.0f);
  const float zd = cfg.rayleigh_zd, ztop = f32(v.zf[nz]), pi = f32(3.1415926535897932384626433);
  auto tau = [&](float z) -> float {
    if (!(z > zd)) return 0.0f;
    const float arg = (pi * (z - zd)) / (ztop - zd);
    return 0.5f * (1.0f - f32(std::cos((double)arg)));
  };
  for (int k = 0; k < nz; ++k) p.tauh[k] = tau(v.zh[k]);
  for (int k = 0; k <= nz; ++k) p.tauf[k] = tau(f32(v.zf[k]));
  return p;
}

// One horizontal axis with its origin at the domain centre (iorigin = 2).
// Uniform: faces i dx - n dx / 2 (grid.js:77-85).  Stretched (grid.js:102-143,
// CM1 param.F stretch_x = 1 / 2): a central zone of ni2 cells of `inner`
// flanked by ni1 (west) and ni3 (east) cells whose spacing grows
// quadratically from `inner` to `outer` with mean `nominal`; halo faces
// continue at `outer` (at `inner` on the unstretched west side of
// stretch 2).  The map factors are dx over the local spacing: uh at centres,
// uf at faces, 1 on the outermost faces (and copied inwards on open sides).
// The face positions are the reference's double expressions, operation for
// operation, because their binary32 roundings are kernel inputs.
HAxis make_axis(int n, double dx, int stretch, double inner, double outer, double nos, double tot, bool open_lo,
                bool open_hi, const char* name) {
  const int H = 3, W = n + 2 * H + 1;  // halo index -3 .. n+3
  HAxis a;
  if (stretch == 0) {
    const double half = 0.5 * (n * dx);
    for (int i = -H; i <= n + H; ++i) {
      a.xf.push_back(i * dx - half);
      a.xh.push_back((i + 0.5) * dx - half);
    }
    a.uh.assign(W, 1.0);
    a.uf.assign(W, 1.0);
    a.min_spacing = dx;
    return a;
  }
  const std::string what = std::string(name);
  if (stretch != 1 && stretch != 2) throw std::runtime_error(what + ": stretch must be 0, 1 or 2");
  if (!(inner > 0 && outer > 0 && nos >= 0 && tot > nos))
    throw std::runtime_error(what + ": stretching needs inner, outer > 0 and total length > unstretched length");
  const double nominal = 0.5 * (inner + outer);
  auto rha = [](double x) { return x < 0 ? -std::round(-x) : std::round(x); };
  int n1, n2, n3;
  if (stretch == 1) {
    n1 = (int)rha((tot - nos) * 0.5 / nominal);
    n2 = (int)rha(nos / inner);
    n3 = n1;
  } else {
    n1 = 0;
    n2 = (int)rha(nos / inner);
    n3 = (int)rha((tot - nos) / nominal);
  }
  if (!(n == n1 + n2 + n3 && n1 >= 0 && n2 >= 0 && n3 >= 0))
    throw std::runtime_error(what + ": " + std::to_string(n) + " points given, the stretching lengths need " +
                             std::to_string(n1) + " + " + std::to_string(n2) + " + " + std::to_string(n3) +
                             " (CM1 stops here, param.F:6805)");
  if (n3 < 2) throw std::runtime_error(what + ": the stretched zone needs at least 2 cells");
  const double c2 = (nominal - inner) / (nominal * nominal * (n3 - 1));
  const double c1 = (inner / nominal) - c2 * nominal;
  const double shift = 0.5 * tot;
  // F[i] = face with CM1 index i (1-based interior faces 1..n+1), i = 1-H .. n+1+H
  std::vector<double> Fv(n + 2 * H + 2, 0.0);
  auto F = [&](int i) -> double& { return Fv[(size_t)(i + H - 1)]; };
  for (int i = n1 + 1; i <= n1 + n2 + 1; ++i) F(i) = n1 * nominal + (i - n1 - 1) * inner - shift;
  const int east_end = n1 + n2 + n3 + (stretch == 1 ? H + 1 : H);
  for (int i = n1 + n2 + 2; i <= east_end; ++i) {
    const double m = (i - 1 - n1 - n2) * nominal;
    F(i) = n1 * nominal + n2 * inner + (c1 + c2 * m) * m - shift;
  }
  for (int i = 1 - H; i <= n1; ++i) {
    const double m = (n1 + 1 - i) * nominal;
    F(i) = n1 * nominal - (c1 + c2 * m) * m - shift;
  }
  const double west = stretch == 1 ? outer : inner;
  for (int i = 1; i <= H; ++i) {
    F(1 - i) = F(1) - i * west;
    F(n + 1 + i) = F(n + 1) + i * outer;
  }
  // HorizontalGrid.derive (grid.js:145-164)
  for (int i = -H; i <= n + H; ++i) a.xf.push_back(F(i + 1));
  for (int i = -H; i <= n - 1 + H; ++i) {
    const double lo = F(i + 1), hi = F(i + 2);
    a.xh.push_back(0.5 * (hi + lo));
    a.uh.push_back(dx / (hi - lo));
  }
  a.xh.push_back(a.xh.back() + (F(n + 1 + H) - F(n + H)));
  a.uh.push_back(a.uh.back());
  a.uf.push_back(1.0);
  for (int i = -H + 1; i <= n + H - 1; ++i) {
    const double c = 0.5 * (F(i + 1) + F(i + 2)), cm = 0.5 * (F(i) + F(i + 1));
    a.uf.push_back(dx / (c - cm));
  }
  a.uf.push_back(1.0);
  if (open_lo)
    for (int i = 0; i < H; ++i) a.uf[i] = a.uf[H];
  if (open_hi)
    for (int i = n + 1 + H; i < W; ++i) a.uf[i] = a.uf[n + H];
  a.uniform = false;
  a.min_spacing = INFINITY;
  for (int i = 0; i < n; ++i) a.min_spacing = std::min(a.min_spacing, a.xf[i + 1 + H] - a.xf[i + H]);
  return a;
}

HAxis axis_x(const te_config& c) {
  return make_axis(c.nx, c.dx, c.stretch_x, c.dx_inner, c.dx_outer, c.nos_x_len, c.tot_x_len, c.open_x != 0,
                   c.open_x != 0, "stretch_x");
}
HAxis axis_y(const te_config& c) {
  return make_axis(c.ny, c.dy, c.stretch_y, c.dy_inner, c.dy_outer, c.nos_y_len, c.tot_y_len, c.open_y != 0,
                   c.open_y != 0, "stretch_y");
}

// Metric buffer (Grid.metricValues, grid.js:200-214): uh | uf (strideJ) |
// vh | vf (strideK) | xf (strideJ) | yf (strideK) | axisymmetric terms
// arh1 arh2 arf1 arf2 (1), rxh (0) and xh (strideJ each).
std::vector<float> make_metric(const HAxis& x, const HAxis& y) {
  const int sJ = (int)x.xf.size(), sK = (int)y.xf.size();
  std::vector<float> m((size_t)9 * sJ + 3 * sK, 1.0f);
  for (int i = 0; i < sJ; ++i) {
    m[i] = f32(x.uh[i]);
    m[sJ + i] = f32(x.uf[i]);
    m[2 * sJ + 2 * sK + i] = f32(x.xf[i]);
  }
  for (int j = 0; j < sK; ++j) {
    m[2 * sJ + j] = f32(y.uh[j]);
    m[2 * sJ + sK + j] = f32(y.uf[j]);
    m[3 * sJ + 2 * sK + j] = f32(y.xf[j]);
  }
  const int base = 3 * sJ + 3 * sK;
  for (int i = 0; i < sJ; ++i) {
    m[base + 4 * sJ + i] = 0.0f;  // rxh (axisymmetric only)
    m[base + 5 * sJ + i] = f32(x.xh[i]);
  }
  return m;
}

// Counter-based normal deviates: SplitMix64 of (seed, cell) -> two uniforms ->
// Box-Muller.
static double unif(uint64_t key) {
  uint64_t z = key + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
static double normal(uint64_t seed, uint64_t cell) {
  const double u1 = unif(seed * 0x100000001B3ull ^ (cell * 2 + 0)), u2 = unif(seed * 0x100000001B3ull ^ (cell * 2 + 1));
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

void fill_state(const te_config& cfg, const Vertical& v, const Base& b, const Layout& L, uint32_t* M) {
  const int nx = L.nx, ny = L.ny, nz = L.nz, sJ = L.sJ, sK = L.sK;
  auto field = [&](const char* n) -> float* {
    uint64_t o = L.field(n);
    if (o == UINT64_MAX) return nullptr;
    return reinterpret_cast<float*>(M + o);
  };
  auto idx = [&](int i, int j, int k) { return (size_t)(i + 3) + (size_t)sJ * ((j + 3) + (size_t)sK * k); };
  const float rd = f32(RDd), cp = f32(CPd), rv = f32(RVd), p00 = f32(P00d), cpdrd = cp / rd;
  float *th = field("th"), *tha = field("th_a"), *u = field("u"), *ua = field("u_a"), *vv = field("v"), *va = field("v_a"),
        *prs = field("prs"), *rho = field("rho"), *qv = field("qv"), *qva = field("qv_a"), *qc = field("qc"),
        *qca = field("qc_a");
  const HAxis ax = axis_x(cfg), ay = axis_y(cfg);
  std::vector<float> thp((size_t)sJ * sK * (nz + 1), 0.0f);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double x = ax.xh[i + 3], y = ay.xh[j + 3], z = v.zh[k];
        const double bx = x / cfg.bubble_rh, by = y / cfg.bubble_rh, bz = (z - cfg.bubble_zc) / cfg.bubble_rv;
        const double beta = std::sqrt(bx * bx + by * by + bz * bz);
        double t = 0.0;
        if (beta < 1.0) {
          const double cb = std::cos(0.5 * M_PI * beta);
          t = cfg.bubble_amp * cb * cb;
        }
        if (cfg.noise_sigma > 0.0f)
          t += cfg.noise_sigma * normal(cfg.seed, (uint64_t)i + (uint64_t)nx * ((uint64_t)j + (uint64_t)ny * k));
        thp[idx(i, j, k)] = f32(t);
      }
  // halo: periodic images (open boundaries: nearest interior value)
  for (int k = 0; k < nz; ++k)
    for (int j = -3; j < ny + 4; ++j)
      for (int i = -3; i < nx + 4; ++i) {
        if (i >= 0 && i < nx && j >= 0 && j < ny) continue;
        int ii = cfg.open_x ? std::min(std::max(i, 0), nx - 1) : ((i % nx) + nx) % nx;
        int jj = cfg.open_y ? std::min(std::max(j, 0), ny - 1) : ((j % ny) + ny) % ny;
        thp[idx(i, j, k)] = thp[idx(ii, jj, k)];
      }
  for (int k = 0; k < nz; ++k) {
    const float pi = b.pi0[k];
    const float p = p00 * f32(std::pow((double)pi, (double)cpdrd));
    for (int j = -3; j < ny + 4; ++j)
      for (int i = -3; i < nx + 4; ++i) {
        const size_t c = idx(i, j, k);
        const float t = thp[c];
        th[c] = t; tha[c] = t;
        u[c] = b.u0[k]; ua[c] = b.u0[k];
        vv[c] = b.v0[k]; va[c] = b.v0[k];
        float q = 0.0f;
        if (cfg.moist) {
          q = b.qv0[k];
          qv[c] = q; qva[c] = q;
          if (qc) { qc[c] = b.qc0[k]; qca[c] = b.qc0[k]; }
        }
        const float rmix = rd + std::max(0.0f, q) * rv;
        prs[c] = p;
        rho[c] = p / (((b.th0[k] + t) * pi) * rmix);
      }
  }
  // t=0 TKE profile of init3d testcases 3, 4/5, 9/11 (scenario.js:347-369),
  // on faces k = 0..nz, zero at the surface and the top
  float *tke = field("tke"), *tkea = field("tke_a");
  const int tc = cfg.initial_tke;
  if (tke && (tc == 3 || tc == 4 || tc == 5 || tc == 9 || tc == 11)) {
    std::vector<float> prof(nz + 1, 0.0f);
    for (int k = 1; k < nz; ++k) {
      const double zf = v.zf[k], zh = v.zh[std::min(k, nz - 1)];
      if (tc == 3) prof[k] = f32(std::max(0.0, 1.0 - zf / 3000.0));
      else if (tc == 4 || tc == 5) prof[k] = 1.0f;
      else if (zh < 255.0) {
        const double r = 1.0 - zf / 255.0;
        prof[k] = f32(0.4 * r * r * r);
      }
    }
    for (int k = 0; k <= nz; ++k)
      for (size_t q = 0; q < (size_t)sJ * sK; ++q) {
        const size_t c = (size_t)k * sJ * sK + q;
        tke[c] = prof[k];
        if (tkea) tkea[c] = prof[k];
      }
  }
}

}  // namespace te
