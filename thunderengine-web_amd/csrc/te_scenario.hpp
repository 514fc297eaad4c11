// Grid, base state and profiles of a configuration (host side).
#pragma once
#include <cstdint>
#include <vector>

#include "te_internal.hpp"

namespace te {

struct Vertical {
  int nz = 0;
  std::vector<double> zf;                       // face heights
  std::vector<float> zh, mh, mf, c1, c2;        // centres, metric terms, face weights
  float cgs[3] = {0, 0, 0}, cgt[3] = {0, 0, 0};  // surface / top extrapolation weights
  double dz_nominal = 0;
};
struct Base {
  std::vector<float> th0, qv0, qc0, pi0, prs0, rho0, u0, v0;
};
struct Profiles {
  std::vector<float> rho0, rr0, rf0, th0, pi0, rth0, qv0, qc0, qi0, tauh, tauf, u0, v0;
};

Vertical make_vertical(const std::vector<double>& faces, double nominal_dz);
Base make_base(const te_config& cfg, const Vertical& v);
Profiles make_profiles(const te_config& cfg, const Vertical& v, const Base& b);
// One horizontal axis: n + 7 values from halo index -3 of the faces xf,
// centres xh and map factors uh (centres) and uf (faces); uniform or CM1's
// stretched layout (HorizontalGrid, web/core/grid.js:52-172).
struct HAxis {
  std::vector<double> xf, xh, uh, uf;
  bool uniform = true;
  double min_spacing = 0;  // smallest interior face spacing (m)
};
HAxis make_axis(int n, double dx, int stretch, double inner, double outer, double nos, double tot, bool open_lo,
                bool open_hi, const char* name);
HAxis axis_x(const te_config& cfg);
HAxis axis_y(const te_config& cfg);
std::vector<float> make_metric(const HAxis& x, const HAxis& y);
void fill_state(const te_config& cfg, const Vertical& v, const Base& b, const Layout& L, uint32_t* M);

}  // namespace te
