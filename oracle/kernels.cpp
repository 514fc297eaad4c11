// Oracle kernel registry and program executor — TEST INFRASTRUCTURE ONLY.
//
// A program is the reference's replay list (replay.js:85-121 after plan()):
// one entry per dispatch or copy, each with a kernel name and a 64-word
// argument slot (A[0..2] thread grid, A[4+n] word offset of parameter n).
// "_copy" entries copy A[0] words from A[4] to A[5] (replay.js:403-406).
#ifdef _OPENMP
#include <omp.h>
#endif
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "oracle.hpp"

namespace orc {
#include "k_dyn.inc"
#include "k_adv.inc"
#include "k_ac.inc"
#include "k_bc.inc"
#include "k_turb.inc"
#include "k_mass.inc"

#define KE(n) {#n, &n<float>, &n<double>, nullptr}
#define KT(n) {#n, &n<float>, &n<double>, &n<TL>}
static const KernelEntry kTable[] = {
    KT(update_diagnostics_moist), KT(update_diagnostics), KT(stage_tendency_reset), KT(stage_state_reset),
    KT(massflux_from_state), KT(velocity_to_massflux), KT(div_massflux), KT(divergence3d), KT(buoyancy_apply),
    KT(rk_update), KT(build_sadv), KT(apply_thterm), KT(turb_add_tendency), KT(turb_zero), KE(tke_advance),
    KT(ppx_finish), KT(diabatic_unapply), KT(diabatic_rates), KT(pressure_tendency), KT(update_rho_prs),
    KT(adv_flux_x), KT(adv_flux_y), KT(adv_flux_z), KT(adv_flux_x_weno), KT(adv_flux_y_weno), KT(adv_flux_z_weno),
    KT(adv_flux_x_stag), KT(adv_flux_y_stag), KT(adv_flux_z_stag), KT(adv_flux_x_weno_w), KT(adv_flux_y_weno_w),
    KT(adv_flux_z_weno_w), KT(adv_flux_x_weno_uv), KT(adv_flux_y_weno_uv), KT(adv_flux_z_weno_uv), KT(adv_assemble),
    KT(acoustic_prepare), KT(acoustic_begin), KT(acoustic_uv), KT(acoustic_fpk), KT(acoustic_w_solve), KT(acoustic_pp),
    KT(acoustic_accumulate), KT(pdef_dum), KT(pdef_apply), KT(bc_periodic_x), KT(bc_periodic_y), KE(bc_open_x),
    KE(bc_open_y), KE(openbc_uv_tendency), KE(radbc_ew), KE(radbc_ns), KE(turb_openbc_flux),
    KE(turb_openbc_stress_x), KE(turb_openbc_stress_y), KT(rayleigh_damping), KE(turb_rho_faces),
    KE(turb_def_centers), KE(turb_def_faces), KE(turb_def_assemble), KE(turb_nm), KE(turb_km_kh),
    KE(turb_dissipative_heating), KE(turb_gettau), KE(turb_sfcstress), KE(tke_epst), KE(tke_sources),
    KE(turb_u_tendency), KE(turb_v_tendency), KE(turb_w_tendency), KE(turb_scalar_flux_h),
    KE(turb_scalar_vert_implicit), KE(tke_flux_h), KE(tke_transport), KE(mass_row_sums), KE(mass_level_sums),
    KE(mass_reference_init), KE(mass_correction_apply), KE(stability_max),
};
#undef KE
#undef KT

int kernel_count() { return (int)(sizeof(kTable) / sizeof(kTable[0])); }
const KernelEntry* kernel_at(int i) { return &kTable[i]; }
const KernelEntry* find_kernel(const std::string& name) {
  for (const auto& e : kTable)
    if (name == e.name) return &e;
  return nullptr;
}

inline void call(const KernelEntry* e, const K<float>& k) { e->f32(k); }
inline void call(const KernelEntry* e, const K<double>& k) { e->f64(k); }
inline void call(const KernelEntry* e, const K<TL>& k) {
  static thread_local std::string msg;
  if (!e->ftl) {
    msg = std::string("oracle: no tangent-linear build of kernel '") + e->name + "'";
    throw msg.c_str();
  }
  e->ftl(k);
}

// The periodic halo fill of field f (timestep.js:1593 haloFill):
// bc_periodic_x then bc_periodic_y, with the dims and metric slots of A.
template <class T>
static void fill_periodic(uint32_t* w, T* v, const uint32_t* A, uint32_t f) {
  const uint32_t* dm = w + A[4];
  const uint32_t nx = dm[0], ny = dm[1], nz = dm[2];
  uint32_t Ax[64] = {0}, Ay[64] = {0};
  Ax[0] = 4; Ax[1] = ny + 7; Ax[2] = nz + 1;
  Ay[0] = 4; Ay[1] = nx + 7; Ay[2] = nz + 1;
  Ax[4] = Ay[4] = A[4];
  Ax[5] = Ay[5] = f;
  Ax[34] = Ay[34] = A[34];
  K<T> kx{w, v, Ax}, ky{w, v, Ay};
  call(find_kernel("bc_periodic_x"), kx);
  call(find_kernel("bc_periodic_y"), ky);
}

template <class T>
static int run_program(uint32_t* w, T* v, const char* const* names, const uint32_t* args, int nops, int first,
                       int last, char* err, int errlen) {
  for (int o = first; o < nops && o <= last; ++o) {
    const uint32_t* A = args + (size_t)o * 64;
    std::string nm(names[o]);
    {
      // the library's pointwise updates with the halo fill of their output
      // folded in: the update, then bc_periodic_x and bc_periodic_y of the
      // field it writes (parameter slot of that field)
      struct Folded { const char* name; int out[4]; int n; };
      static const Folded kFolded[] = {{"rk_update_xy", {4}, 1},    {"tke_advance_xy", {4}, 1},
                                       {"pdef_dum_xy", {10}, 1},    {"ppx_finish_xy", {2}, 1},
                                       {"turb_km_kh_xy", {4, 5, 6, 7}, 4}};
      const Folded* fo = nullptr;
      for (const auto& f : kFolded)
        if (nm == f.name) fo = &f;
      if (fo) {
        K<T> k1{w, v, A};
        call(find_kernel(nm.substr(0, nm.size() - 3)), k1);
        for (int q = 0; q < fo->n; ++q) fill_periodic(w, v, A, A[4 + fo->out[q]]);
        continue;
      }
    }
    if (nm == "_copy") {
      const uint32_t n = A[0], src = A[4], dst = A[5];
#pragma omp parallel for schedule(static)
      for (int64_t q = 0; q < (int64_t)n; ++q) v[dst + q] = v[src + q];
      continue;
    }
    if (nm == "bc_periodic_xy") {
      // the library's merged periodic halo launch, checked as what it stands
      // for: bc_periodic_x then bc_periodic_y on each of its fields, in order
      const KernelEntry* ex = find_kernel("bc_periodic_x");
      const KernelEntry* ey = find_kernel("bc_periodic_y");
      const uint32_t* dm = w + A[4];
      const uint32_t nx = dm[0], ny = dm[1], nz = dm[2];
      const uint32_t nf = w[A[4 + 9]];
      for (uint32_t q = 0; q < nf && q < 8; ++q) {
        uint32_t Ax[64] = {0}, Ay[64] = {0};
        Ax[0] = 4; Ax[1] = ny + 7; Ax[2] = nz + 1;
        Ay[0] = 4; Ay[1] = nx + 7; Ay[2] = nz + 1;
        Ax[4] = Ay[4] = A[4];
        Ax[5] = Ay[5] = A[5 + q];
        Ax[34] = Ay[34] = A[34];
        K<T> kx{w, v, Ax}, ky{w, v, Ay};
        call(ex, kx); call(ey, ky);
      }
      continue;
    }
    if (nm == "ac_column") {
      // the dispatches one acoustic substep launch stands for: on the stage's
      // first substep acoustic_prepare and acoustic_begin, then acoustic_fpk,
      // acoustic_w_solve, acoustic_pp, acoustic_accumulate and the periodic
      // halo fill of ppd, u, v, w (bc_periodic_x then bc_periodic_y each)
      const uint32_t* dm = w + A[4];
      const uint32_t nx = dm[0], ny = dm[1], nz = dm[2];
      auto run1 = [&](const char* kn, std::array<uint32_t, 3> g3, std::vector<std::pair<int, uint32_t>> ps) {
        uint32_t X[64] = {0};
        X[0] = g3[0]; X[1] = g3[1]; X[2] = g3[2];
        X[4] = A[4];      // dims
        X[5] = A[5];      // acoustic parameters
        X[34] = A[34];    // metric
        for (auto& pr : ps) X[4 + pr.first] = pr.second;
        const KernelEntry* e1 = find_kernel(kn);
        K<T> k1{w, v, X};
        call(e1, k1);
      };
      const std::array<uint32_t, 3> full = {nx, ny, nz}, whole = {nx + 7, ny + 7, nz + 1}, cols = {nx, ny, 1};
      auto P = [&](int q) { return A[4 + q]; };
      const uint32_t U = P(2), V = P(3), W = P(4), PIADV = P(5), PPTEN = P(6), PPTERM = P(7), THV = P(8), WTEN = P(9),
                     PP3D = P(10), PPD = P(11), PPX = P(12), RRU = P(13), RRV = P(14), RRW = P(15), QK = P(24),
                     PK1 = P(25), PK2 = P(26), MM = P(27), FPK = P(28), CA = P(29), CB = P(31);
      if (w[P(33)] != 0u) {
        run1("acoustic_prepare", full, {{2, PPTERM}, {3, PIADV}, {4, THV}, {5, P(18)}, {6, P(19)}, {7, P(20)}, {8, P(21)},
                                        {9, P(22)}, {10, P(23)}, {11, QK}, {12, PK1}, {13, PK2}, {14, MM}});
        run1("acoustic_begin", whole, {{2, W}, {3, RRU}, {4, RRV}, {5, RRW}});
      }
      run1("acoustic_fpk", full, {{2, U}, {3, V}, {4, W}, {5, PIADV}, {6, PPTEN}, {7, PPTERM}, {8, PK1}, {9, PK2},
                                  {10, QK}, {11, FPK}, {15, P(32)}, {16, P(18)}, {17, P(19)}});
      run1("acoustic_w_solve", cols, {{2, PK1}, {3, PK2}, {4, QK}, {5, MM}, {6, FPK}, {7, WTEN}, {8, PP3D}, {9, W},
                                      {10, CA}, {11, CB}});
      run1("acoustic_pp", full, {{2, W}, {3, PK1}, {4, PK2}, {5, QK}, {6, FPK}, {7, PP3D}, {8, PPD}, {9, PPX},
                                 {10, P(16)}});
      run1("acoustic_accumulate", whole, {{2, U}, {3, V}, {4, W}, {5, RRU}, {6, RRV}, {7, RRW}, {8, P(17)}});
      for (uint32_t f : {PPD, U, V, W}) {
        run1("bc_periodic_x", {4, ny + 7, nz + 1}, {{1, f}});
        run1("bc_periodic_y", {4, nx + 7, nz + 1}, {{1, f}});
      }
      if (P(34) != 0 && w[P(34)] != 0u) fill_periodic(w, v, A, PP3D);  // the stage's pp3d fill, folded in
      continue;
    }
    if (nm == "massflux_div") {
      // massflux_from_state, then div_massflux of the fluxes it wrote
      K<T> k1{w, v, A};
      call(find_kernel("massflux_from_state"), k1);
      uint32_t X[64] = {0};
      X[0] = A[0]; X[1] = A[1]; X[2] = A[2];
      X[4] = A[4]; X[34] = A[34];
      X[5] = A[4 + 6]; X[6] = A[4 + 7]; X[7] = A[4 + 8]; X[8] = A[4 + 9]; X[9] = A[4 + 10];
      K<T> k2{w, v, X};
      call(find_kernel("div_massflux"), k2);
      continue;
    }
    if (nm == "ac_pp_accumulate") {
      // acoustic_pp then acoustic_accumulate, the dispatches it stands for
      const uint32_t* dm = w + A[4];
      const uint32_t nx = dm[0], ny = dm[1], nz = dm[2];
      uint32_t X[64] = {0}, Y[64] = {0};
      X[4] = Y[4] = A[4];      // dims
      X[5] = Y[5] = A[5];      // acoustic parameters
      X[34] = Y[34] = A[34];   // metric
      X[0] = nx; X[1] = ny; X[2] = nz;
      for (int q = 2; q <= 10; ++q) X[4 + q] = A[4 + q];  // w pk1 pk2 qk fpk pp3d ppd ppx saveX
      Y[0] = A[0]; Y[1] = A[1]; Y[2] = A[2];
      const int as[7] = {11, 12, 2, 13, 14, 15, 16};         // u v w rru rrv rrw tavg
      for (int q = 0; q < 7; ++q) Y[4 + 2 + q] = A[4 + as[q]];
      const KernelEntry *e1 = find_kernel("acoustic_pp"), *e2 = find_kernel("acoustic_accumulate");
      K<T> k1{w, v, X}, k2{w, v, Y};
      call(e1, k1); call(e2, k2);
      continue;
    }
    if (nm == "adv_fused") {
      // the three flux dispatches it stands for, then adv_assemble
      const uint32_t* dm = w + A[4];
      const uint32_t nx = dm[0], ny = dm[1], nz = dm[2];
      const uint32_t kind = w[A[4 + 3]];
      const uint32_t S = A[4 + 2], RU = A[4 + 13], RV = A[4 + 14], RW = A[4 + 4], C1 = A[4 + 15], C2 = A[4 + 16];
      const uint32_t ST = A[4 + 1], WE = A[4 + 5], PD = A[4 + 20], PW = A[4 + 21];
      const uint32_t FX = A[4 + 22], FY = A[4 + 23], FZ = A[4 + 24];
      struct Slot { const char* k; uint32_t g[3]; std::vector<std::pair<int, uint32_t>> p; };
      std::vector<Slot> fl;
      if (kind == 0) {
        fl = {{"adv_flux_x", {nx + 1, ny, nz}, {{1, S}, {2, RU}, {3, FX}}},
              {"adv_flux_y", {nx, ny + 1, nz}, {{1, S}, {2, RV}, {3, FY}}},
              {"adv_flux_z", {nx, ny, nz + 1}, {{1, S}, {2, RW}, {3, C1}, {4, C2}, {5, FZ}}}};
      } else if (kind == 1) {
        fl = {{"adv_flux_x_weno", {nx + 1, ny, nz}, {{1, S}, {2, RU}, {3, WE}, {4, PD}, {5, FX}}},
              {"adv_flux_y_weno", {nx, ny + 1, nz}, {{1, S}, {2, RV}, {3, WE}, {4, PD}, {5, FY}}},
              {"adv_flux_z_weno", {nx, ny, nz + 1}, {{1, S}, {2, RW}, {3, C1}, {4, C2}, {5, WE}, {6, PD}, {7, PW}, {8, FZ}}}};
      } else if (kind == 2) {
        std::vector<std::pair<int, uint32_t>> base = {{1, ST}, {2, S}, {3, RU}, {4, RV}, {5, RW}, {6, C1}, {7, C2}};
        auto with = [&](uint32_t f) { auto p = base; p.push_back({8, f}); return p; };
        fl = {{"adv_flux_x_stag", {nx + 2, ny + 2, nz}, with(FX)},
              {"adv_flux_y_stag", {nx + 2, ny + 2, nz}, with(FY)},
              {"adv_flux_z_stag", {nx + 1, ny + 1, nz + 1}, with(FZ)}};
      } else {
        fl = {{"adv_flux_x_weno_w", {nx + 2, ny + 2, nz}, {{1, S}, {2, RU}, {3, C2}, {4, WE}, {5, FX}}},
              {"adv_flux_y_weno_w", {nx + 2, ny + 2, nz}, {{1, S}, {2, RV}, {3, C2}, {4, WE}, {5, FY}}},
              {"adv_flux_z_weno_w", {nx + 1, ny + 1, nz + 1}, {{1, S}, {2, RW}, {3, WE}, {4, FZ}}}};
      }
      for (const Slot& sl : fl) {
        uint32_t X[64] = {0};
        X[0] = sl.g[0]; X[1] = sl.g[1]; X[2] = sl.g[2];
        X[4] = A[4];
        X[34] = A[34];
        for (auto& pr : sl.p) X[4 + pr.first] = pr.second;
        const KernelEntry* e1 = find_kernel(sl.k);
        K<T> k1{w, v, X};
        call(e1, k1);
      }
      if (w[A[4 + 32]] & 16u) {  // turb_zero of the accumulating sink
        uint32_t Z[64] = {0};
        Z[0] = nx + 7; Z[1] = ny + 7; Z[2] = nz + 1;
        Z[4] = A[4]; Z[5] = A[4 + 11]; Z[34] = A[34];
        const KernelEntry* ez = find_kernel("turb_zero");
        K<T> kz{w, v, Z};
        call(ez, kz);
      }
      uint32_t Y[64];
      std::memcpy(Y, A, sizeof(Y));
      Y[4 + 3] = FX; Y[4 + 4] = FY; Y[4 + 5] = FZ;
      const KernelEntry* ea = find_kernel("adv_assemble");
      K<T> ka{w, v, Y};
      call(ea, ka);
      // the scalar follow-ups chained into the launch, as their own dispatches
      const uint32_t chain = w[A[4 + 32]], TEND = A[4 + 11];
      auto run1 = [&](const char* kn, const uint32_t* g3, std::vector<std::pair<int, uint32_t>> ps) {
        uint32_t X[64] = {0};
        X[0] = g3[0]; X[1] = g3[1]; X[2] = g3[2];
        X[4] = A[4];
        X[34] = A[34];
        for (auto& pr : ps) X[4 + pr.first] = pr.second;
        const KernelEntry* e1 = find_kernel(kn);
        K<T> k1{w, v, X};
        call(e1, k1);
      };
      const uint32_t full[3] = {nx, ny, nz}, whole[3] = {nx + 7, ny + 7, nz + 1}, asz[3] = {A[0], A[1], A[2]};
      if (chain & 1) run1("apply_thterm", full, {{1, A[4 + 33]}, {2, A[4 + 34]}, {3, TEND}});
      if (chain & 2) run1("turb_add_tendency", whole, {{1, A[4 + 35]}, {2, TEND}, {3, A[4 + 36]}});
      if (chain & 4) run1("turb_add_tendency", whole, {{1, A[4 + 37]}, {2, TEND}, {3, A[4 + 38]}});
      if (chain & 8) run1("rk_update", asz, {{1, A[4 + 1]}, {2, A[4 + 39]}, {3, TEND}, {4, A[4 + 40]}, {5, A[4 + 41]}});
      if (chain & 32) fill_periodic(w, v, A, A[4 + 40]);  // the updated field's halo fill, folded in
      if (chain & 64) {  // buoyancy_apply on the assembled tendency of w
        const uint32_t bg[3] = {nx, ny, std::max<uint32_t>(nz - 1, 1)};
        run1("buoyancy_apply", bg, {{1, A[4 + 42]}, {2, A[4 + 15]}, {3, A[4 + 16]}, {4, TEND}});
      }
      continue;
    }
    if (nm == "tendency_reset_add") {  // stage_tendency_reset, then turb_add_tendency x3
      const KernelEntry *er = find_kernel("stage_tendency_reset"), *ea = find_kernel("turb_add_tendency");
      uint32_t R[64] = {0};
      R[0] = A[0]; R[1] = A[1]; R[2] = A[2]; R[4] = A[4]; R[5] = A[5]; R[6] = A[6]; R[7] = A[7]; R[34] = A[34];
      K<T> kr{w, v, R};
      call(er, kr);
      for (int q = 0; q < 3; ++q) {
        uint32_t X[64] = {0};
        X[0] = A[0]; X[1] = A[1]; X[2] = A[2]; X[4] = A[4];
        X[5] = A[4 + 4 + q]; X[6] = A[4 + 1 + q]; X[7] = A[4 + 7];
        X[34] = A[34];
        K<T> kx{w, v, X};
        call(ea, kx);
      }
      continue;
    }
    if (nm == "add_fields") {  // turb_add_tendency on each triple, in order
      const KernelEntry* ea = find_kernel("turb_add_tendency");
      const uint32_t nf = w[A[4 + 13]];
      for (uint32_t q = 0; q < nf && q < 4; ++q) {
        uint32_t X[64] = {0};
        X[0] = A[0]; X[1] = A[1]; X[2] = A[2];
        X[4] = A[4];
        X[5] = A[5 + q]; X[6] = A[9 + q]; X[7] = A[13 + q];
        X[34] = A[34];
        K<T> kx{w, v, X};
        call(ea, kx);
      }
      continue;
    }
    if (nm == "zero_fields") {  // turb_zero on each of its fields
      const KernelEntry* ez = find_kernel("turb_zero");
      const uint32_t nf = w[A[4 + 9]];
      for (uint32_t q = 0; q < nf && q < 8; ++q) {
        uint32_t Az[64] = {0};
        Az[0] = A[0]; Az[1] = A[1]; Az[2] = A[2];
        Az[4] = A[4];
        Az[5] = A[5 + q];
        Az[34] = A[34];
        K<T> kz{w, v, Az};
        call(ez, kz);
      }
      continue;
    }
    if (nm == "copy_fields") {  // the buffer copies it stands for, in order
      const uint32_t nf = w[A[4 + 9]], n = A[0];
      for (uint32_t f = 0; f < nf && f < 4; ++f) {
        const uint32_t src = A[5 + f], dst = A[9 + f];
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < (int64_t)n; ++q) v[dst + q] = v[src + q];
      }
      continue;
    }
    // the library's column sweeps with LDS coefficients run as the reference
    // kernels they stand for (their cA / cB fields are scratch)
    if (nm.size() > 4 && nm.compare(nm.size() - 4, 4, "_lds") == 0) nm.resize(nm.size() - 4);
    const KernelEntry* e = find_kernel(nm);
    if (!e) {
      std::snprintf(err, errlen, "oracle: no kernel '%s' (op %d)", names[o], o);
      return -1;
    }
    K<T> k{w, v, A};
    call(e, k);
  }
  return 0;
}

// reference region hash (replay.js:408-418): sum over the region of a
// 32-bit mix of (word ^ index*0x9E3779B1).
static uint32_t region_hash(const uint32_t* m, size_t len) {
  uint32_t s = 0;
  for (size_t i = 0; i < len; ++i) {
    uint32_t h = (m[i] ^ ((uint32_t)i * 0x9E3779B1u)) * 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    s += h;
  }
  return s;
}

}  // namespace orc

extern "C" {

// OpenMP threads the oracle's loops use.
int oracle_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Runs ops [first, last] of a program on a float word memory.
int oracle_run(uint32_t* mem, const char* const* names, const uint32_t* args, int nops, int first, int last, char* err,
               int errlen) {
  try {
    return orc::run_program<float>(mem, reinterpret_cast<float*>(mem), names, args, nops, first, last, err, errlen);
  } catch (const char* msg) {
    std::snprintf(err, errlen, "%s", msg);
    return -2;
  }
}

// FP64 build: `mem` keeps the byte parameters, `vals` the double values of
// every word (same indexing).
int oracle_run_f64(uint32_t* mem, double* vals, const char* const* names, const uint32_t* args, int nops, int first,
                   int last, char* err, int errlen) {
  try {
    return orc::run_program<double>(mem, vals, names, args, nops, first, last, err, errlen);
  } catch (const char* msg) {
    std::snprintf(err, errlen, "%s", msg);
    return -2;
  }
}

// FP64 tangent-linear build: `vals` holds (value, tangent) double pairs of
// every word (same indexing); the tangents after the ops are J times those
// before.
int oracle_run_tl(uint32_t* mem, double* vals, const char* const* names, const uint32_t* args, int nops, int first,
                  int last, char* err, int errlen) {
  try {
    return orc::run_program<orc::TL>(mem, reinterpret_cast<orc::TL*>(vals), names, args, nops, first, last, err,
                                      errlen);
  } catch (const char* msg) {
    std::snprintf(err, errlen, "%s", msg);
    return -2;
  }
}

uint32_t oracle_region_hash(const uint32_t* mem, uint64_t len) { return orc::region_hash(mem, len); }

uint32_t oracle_fnv1a(const uint32_t* mem, uint64_t len) {
  uint32_t h = 0x811c9dc5u;
  for (uint64_t i = 0; i < len; ++i) h = (h ^ mem[i]) * 0x01000193u;
  return h;
}

int oracle_kernel_count() { return orc::kernel_count(); }
const char* oracle_kernel_name(int i) { return orc::kernel_at(i)->name; }

// Elementary functions exposed for the closed-form / brute-force tests.
float oracle_rr_pow(float x, float p) { return orc::rr_pow(x, p); }
float oracle_rr_log(float x) { return orc::rr_log(x); }
float oracle_rr_exp(float x) { return orc::rr_exp(x); }
float oracle_weno5(float s1, float s2, float s3, float s4, float s5, float weps) {
  return orc::weno5<float>(s1, s2, s3, s4, s5, weps);
}
double oracle_weno5_f64(double s1, double s2, double s3, double s4, double s5, double weps) {
  return orc::weno5<double>(s1, s2, s3, s4, s5, weps);
}
float oracle_flx5(float s1, float s2, float s3, float s4, float s5) { return orc::flx5<float>(s1, s2, s3, s4, s5); }
float oracle_flx3(float s1, float s2, float s3) { return orc::flx3<float>(s1, s2, s3); }
float oracle_upstrpd(float s1, float s2, float s3, float weps) { return orc::upstrpd<float>(s1, s2, s3, weps); }
}
