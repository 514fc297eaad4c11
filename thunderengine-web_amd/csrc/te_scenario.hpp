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
std::vector<float> make_metric(int nx, int ny, float dx, float dy);
void fill_state(const te_config& cfg, const Vertical& v, const Base& b, const Layout& L, uint32_t* M);

}  // namespace te
