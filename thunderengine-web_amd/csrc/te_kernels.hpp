// Registry of the step kernels: every kernel has the signature
// (uint32_t* M, KArg a) and is addressed by the reference kernel name.
#pragma once
#include "te_device.hpp"

#define TE_KERNEL_LIST(X)                                                                                      \
  X(update_diagnostics_moist) X(update_diagnostics) X(stage_tendency_reset) X(stage_state_reset)              \
  X(massflux_from_state) X(velocity_to_massflux) X(div_massflux) X(divergence3d) X(buoyancy_apply) X(rk_update) \
  X(build_sadv) X(apply_thterm) X(turb_add_tendency) X(turb_zero) X(tke_advance) X(ppx_finish)                 \
  X(diabatic_unapply) X(diabatic_rates) X(pressure_tendency) X(update_rho_prs) X(adv_flux_x) X(adv_flux_y)     \
  X(adv_flux_z) X(adv_flux_x_weno) X(adv_flux_y_weno) X(adv_flux_z_weno) X(adv_flux_x_stag) X(adv_flux_y_stag) \
  X(adv_flux_z_stag) X(adv_flux_x_weno_w) X(adv_flux_y_weno_w) X(adv_flux_z_weno_w) X(adv_flux_x_weno_uv)      \
  X(adv_flux_y_weno_uv) X(adv_flux_z_weno_uv) X(adv_assemble) X(acoustic_prepare) X(acoustic_begin)            \
  X(acoustic_uv) X(acoustic_fpk) X(acoustic_w_solve) X(acoustic_pp) X(acoustic_accumulate) X(pdef_dum)         \
  X(pdef_apply) X(bc_periodic_x) X(bc_periodic_y) X(bc_open_x) X(bc_open_y) X(openbc_uv_tendency) X(radbc_ew)   \
  X(radbc_ns) X(turb_openbc_flux) X(turb_openbc_stress_x) X(turb_openbc_stress_y) X(rayleigh_damping)          \
  X(turb_rho_faces) X(turb_def_centers) X(turb_def_faces) X(turb_def_assemble) X(turb_nm) X(turb_km_kh)        \
  X(turb_dissipative_heating) X(turb_gettau) X(turb_sfcstress) X(tke_epst) X(tke_sources) X(turb_u_tendency)   \
  X(turb_v_tendency) X(turb_w_tendency) X(turb_scalar_flux_h) X(turb_scalar_vert_implicit) X(tke_flux_h)       \
  X(tke_transport) X(mass_row_sums) X(mass_level_sums) X(mass_reference_init) X(mass_correction_apply) X(_copy)  \
  X(stability_max) \
  TE_FUSED_LIST(X)

// Launches of this build with no counterpart in the reference's kernel list:
// each replaces a run of reference dispatches with the same result on the
// fields the step carries.
#define TE_FUSED_LIST(X) X(bc_periodic_xy) X(zero_fields) X(copy_fields) X(ac_pp_accumulate) \
  X(adv_fused) X(add_fields) X(tendency_reset_add) X(acoustic_w_solve_lds) X(turb_scalar_vert_implicit_lds) X(turb_u_tendency_lds) \
  X(turb_v_tendency_lds) X(turb_w_tendency_lds) X(tke_transport_lds) X(rk_update_xy) X(tke_advance_xy) \
  X(pdef_dum_xy) X(ppx_finish_xy) X(turb_km_kh_xy) X(massflux_div)

// Fused launches of 64 x AC_WAVES threads per workgroup (one acoustic
// substep per launch, k_fused.hip).
#define AC_WAVES 8
#define TE_WIDE_LIST(X) X(ac_column)

namespace te {
#define TE_DECL(n) __global__ __launch_bounds__(256) void k_##n(uint32_t* M, KArg a);
TE_KERNEL_LIST(TE_DECL)
#undef TE_DECL
#define TE_DECL_WIDE(n) __global__ __launch_bounds__(64 * AC_WAVES) void k_##n(uint32_t* M, KArg a);
TE_WIDE_LIST(TE_DECL_WIDE)
#undef TE_DECL_WIDE
// k_ac_column instantiated for 1..16 levels per wave (nullptr outside)
const void* ac_column_kernel(int levels_per_wave);
// turb_scalar_vert_implicit_lds in 64 x AC_WAVES workgroups (k_turb.hip),
// instantiated for 1..16 levels per wave (nullptr outside)
const void* svi_waves_kernel(int levels_per_wave);
// adv_fused specialised on (flux kind, stagger) (k_adv.hip); null for the
// pairs only the run-time dispatch of k_adv_fused serves
const void* adv_fused_kernel(int kind, int stag);
// turb_u_tendency_lds (dir 0) / turb_v_tendency_lds (dir 1) in the same form
const void* uv_waves_kernel(int dir, int levels_per_wave);
// turb_w_tendency_lds (which 0) / tke_transport_lds (which 1): levels 1 .. nz-1
const void* face_waves_kernel(int which, int levels_per_wave);
}  // namespace te
