// Per-point bodies of the reference kernels on dual numbers, for the adjoint
// (see adj_core.hpp).  Each body restates one reference kernel: the same
// thread grid, bounds, branches and operation order as the forward kernel of
// the same name (the citation names the WGSL body it follows); active loads go
// through r.ld(), outputs through r.out<o>().
#pragma once
#include "adj_core.hpp"

namespace te {
namespace adj {

// ---- advection stencils (adv_flux_x.wgsl:106-113, adv_flux_z.wgsl:106-111,
// adv_flux_x_weno.wgsl:106-151, adv_flux_z_weno.wgsl:111-120).  Quotient
// constants such as 13/12 are formed in binary64 here.
TE_HD D1 flx5(D1 s1, D1 s2, D1 s3, D1 s4, D1 s5) {
  return fma_(D1(-3.0 / 6e+01), s5,
              fma_(D1(27.0 / 6e+01), s4,
                   fma_(D1(47.0 / 6e+01), s3, fma_(D1(2.0 / 6e+01), s1, D1(-13.0 / 6e+01) * s2))));
}
TE_HD D1 flx3(D1 s1, D1 s2, D1 s3) {
  return fma_(D1(2.0 / 6.0), s3, fma_(D1(-1.0 / 6.0), s1, D1(5.0 / 6.0) * s2));
}
// WENO5 with its weights differentiated through the smoothness indicators
TE_HD D1 weno5(D1 s1, D1 s2, D1 s3, D1 s4, D1 s5, float weps) {
  const D1 c1312 = D1(13.0 / 12.0);
  D1 a1 = fma_(D1(-2.0f), s2, s1) + s3, q1 = fma_(D1(3.0f), s3, fma_(D1(-4.0f), s2, s1));
  D1 b1 = fma_(c1312 * a1, a1, (D1(0.25f) * q1) * q1);
  D1 a2 = fma_(D1(-2.0f), s3, s2) + s4, q2 = s2 - s4;
  D1 b2 = fma_(c1312 * a2, a2, (D1(0.25f) * q2) * q2);
  D1 a3 = fma_(D1(-2.0f), s4, s3) + s5, q3 = fma_(D1(3.0f), s3, -(D1(4.0f) * s4)) + s5;
  D1 b3 = fma_(c1312 * a3, a3, (D1(0.25f) * q3) * q3);
  D1 tau = ab(b1 - b3);
  D1 r1 = mn(D1(1e+30f), tau / (b1 + D1(weps)));
  D1 r2 = mn(D1(1e+30f), tau / (b2 + D1(weps)));
  D1 r3 = mn(D1(1e+30f), tau / (b3 + D1(weps)));
  D1 m = mx(r1, mx(r2, r3));
  D1 one = D1(1.0f);
  if (m.v > 1e+15f) {
    D1 im = D1(1.0f) / m;
    r1 = r1 * im; r2 = r2 * im; r3 = r3 * im;
    one = im * im;
  }
  D1 w1 = D1(0.1f) * fma_(r1, r1, one);
  D1 w2 = D1(0.6f) * fma_(r2, r2, one);
  D1 w3 = D1(0.3f) * fma_(r3, r3, one);
  D1 p1 = fma_(D1(11.0 / 6.0), s3, fma_(D1(2.0 / 6.0), s1, D1(-7.0 / 6.0) * s2));
  D1 p2 = fma_(D1(2.0 / 6.0), s4, fma_(D1(-1.0 / 6.0), s2, D1(5.0 / 6.0) * s3));
  D1 p3 = fma_(D1(-1.0 / 6.0), s5, fma_(D1(2.0 / 6.0), s3, D1(5.0 / 6.0) * s4));
  return fma_(w3, p3, fma_(w1, p1, w2 * p2)) / ((w1 + w2) + w3);
}
TE_HD bool weno_pdef_ok(D1 s1, D1 s2, D1 s3, D1 s4, D1 s5, float weps, bool pdef) {
  if (!pdef) return true;
  D1 bmax = mx(mx(mx(ab(s1), ab(s2)), mx(ab(s3), ab(s4))), ab(s5));
  return bmax.v > weps;
}
TE_HD D1 upstrpd(D1 s1, D1 s2, D1 s3, float weps) {
  D1 dd = s2 - s1;
  D1 rr = (s3 - s2) / (copysign_mag(D1(weps), dd) + dd);
  D1 phi = mx(D1(0.0f), mn(rr, mn(fma_(D1(2.0 / 6.0), rr, D1(1.0 / 6.0)), D1(1.0f))));
  return fma_(phi, s2 - s1, s2);
}

// ---- diagnostics, resets, mass fluxes

// update_diagnostics_moist.wgsl:128-184
struct B_update_diagnostics_moist {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int i = x - 3, j = y - 3, kk = z;
    if (i > g.nx + 3 || j > g.ny + 3 || kk >= g.nz) return;
    const P thP = r.f(1), ppP = r.f(2), th0 = r.f(3), pi0 = r.f(4), qv0 = r.f(5), qc0 = r.f(6),
                qi0 = r.f(7), rth0 = r.f(8), qv = r.f(9), qc = r.f(10), qr = r.f(11), qi = r.f(12),
                qs = r.f(13), qg = r.f(14), thv = r.f(15), ppterm = r.f(16), piadv = r.f(17), thterm = r.f(18),
                buoy = r.f(19);
    const int64_t c = g.idx(i, j, kk);
    const D1 rd = RD, rv = RV, cp = CP, cv = CV, cpv = CPV, cvv = CVV, rovcp = ROVCP;
    const D1 reps = rv / rd;
    const D1 repsm1 = reps - D1(1.0f);
    const D1 tP = r.ld(thP, c);
    const D1 qvv = mx(r.ld(qv, c), D1(0.0f));
    const D1 liq = r.ld(qc, c) + r.ld(qr, c);
    const D1 ice = (r.ld(qi, c) + r.ld(qs, c)) + r.ld(qg, c);
    const D1 qli = mx(D1(0.0f), liq + ice);
    const D1 cpli = fma_(D1(CPL), mx(D1(0.0f), liq), D1(CPI) * mx(D1(0.0f), ice));
    const D1 th = D1(th0[kk]) + tP;
    const D1 pi = D1(pi0[kk]) + r.ld(ppP, c);
    const D1 cpm = fma_(cpv, qvv, cp) + cpli;
    const D1 cvm = D1(1.0f) / (fma_(cvv, qvv, cv) + cpli);
    r.out<0>(thv, c, (th * fma_(reps, qvv, D1(1.0f))) / ((D1(1.0f) + qvv) + qli));
    r.out<1>(piadv, c, pi);
    r.out<2>(ppterm, c, ((pi * rovcp) * cpm) * cvm);
    r.out<3>(thterm, c, (th * fma_(-rovcp, cpm, fma_(rv, qvv, rd))) * cvm);
    r.out<4>(buoy, c, D1(GRAV) * (fma_(tP, D1(rth0[kk]), repsm1 * (qvv - D1(qv0[kk]))) - ((qli - D1(qc0[kk])) - D1(qi0[kk]))));
  }
};

// update_diagnostics.wgsl:114-144 (dry)
struct B_update_diagnostics {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int i = x - 3, j = y - 3, kk = z;
    if (i > g.nx + 3 || j > g.ny + 3 || kk >= g.nz) return;
    const P thP = r.f(1), ppP = r.f(2), th0 = r.f(3), pi0 = r.f(4), thv = r.f(5), ppterm = r.f(6),
                piadv = r.f(7), rth0 = r.f(8), buoy = r.f(9);
    const D1 rddcv = D1(RD) / (D1(CP) - D1(RD));
    const int64_t c = g.idx(i, j, kk);
    const D1 tP = r.ld(thP, c);
    const D1 pi = D1(pi0[kk]) + r.ld(ppP, c);
    r.out<0>(thv, c, D1(th0[kk]) + tP);
    r.out<1>(piadv, c, pi);
    r.out<2>(ppterm, c, pi * rddcv);
    r.out<3>(buoy, c, (D1(GRAV) * tP) * D1(rth0[kk]));
  }
};

// stage_tendency_reset.wgsl:110-129
struct B_stage_tendency_reset {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int i = x - 3, j = y - 3, kk = z;
    if (i > g.nx + 3 || j > g.ny + 3 || kk > g.nz) return;
    const int64_t c = g.idx(i, j, kk);
    r.out<0>(r.f(1), c, D1(0.0f));
    r.out<1>(r.f(2), c, D1(0.0f));
    r.out<2>(r.f(3), c, D1(0.0f));
  }
};

// stage_state_reset.wgsl:110-138
// five independent copies: one body per output
template <int O>
struct B_stage_state_reset_part {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int i = x - 3, j = y - 3, kk = z;
    if (i > g.nx + 3 || j > g.ny + 3 || kk > g.nz) return;
    const int64_t c = g.idx(i, j, kk);
    r.out<0>(r.f(5 + O), c, r.ld(r.f(O == 4 ? 10 : 1 + O), c));
  }
};
using B_stage_state_reset = Parts<B_stage_state_reset_part<0>, B_stage_state_reset_part<1>,
                                  B_stage_state_reset_part<2>, B_stage_state_reset_part<3>,
                                  B_stage_state_reset_part<4>>;

// massflux_from_state.wgsl:111-137
// three independent products: one body per output
template <int O>
struct B_massflux_from_state_part {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int i = x - 3, j = y - 3, kk = z;
    if (i > g.nx + 3 || j > g.ny + 3 || kk > g.nz) return;
    const P u = r.f(1), v = r.f(2), w = r.f(3), rho0 = r.f(4), rf0 = r.f(5);
    const int64_t c = g.idx(i, j, kk);
    if (O == 0 && kk < g.nz) r.out<0>(r.f(6), c, D1(rho0[kk]) * r.ld(u, c));
    if (O == 1 && kk < g.nz) r.out<0>(r.f(7), c, D1(rho0[kk]) * r.ld(v, c));
    if (O == 2) r.out<0>(r.f(8), c, (kk == 0 || kk == g.nz) ? D1(0.0f) : D1(rf0[kk]) * r.ld(w, c));
  }
};
using B_massflux_from_state =
    Parts<B_massflux_from_state_part<0>, B_massflux_from_state_part<1>, B_massflux_from_state_part<2>>;

// velocity_to_massflux.wgsl:111-134
// three independent in-place products: one body per output
template <int O>
struct B_velocity_to_massflux_part {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int i = x - 3, j = y - 3, kk = z;
    if (i > g.nx + 3 || j > g.ny + 3 || kk > g.nz) return;
    const P rho0 = r.f(1), rf0 = r.f(2), rru = r.f(3), rrv = r.f(4), rrw = r.f(5);
    const int64_t c = g.idx(i, j, kk);
    if (O == 0 && kk < g.nz) r.out<0>(rru, c, r.ld(rru, c) * D1(rho0[kk]));
    if (O == 1 && kk < g.nz) r.out<0>(rrv, c, r.ld(rrv, c) * D1(rho0[kk]));
    if (O == 2) r.out<0>(rrw, c, (kk == 0 || kk == g.nz) ? D1(0.0f) : r.ld(rrw, c) * D1(rf0[kk]));
  }
};
using B_velocity_to_massflux =
    Parts<B_velocity_to_massflux_part<0>, B_velocity_to_massflux_part<1>, B_velocity_to_massflux_part<2>>;

// div_massflux.wgsl:125-147
// The divergence is the sum of three one-direction differences: one body
// per direction adds its share (keep), then a body with no loads takes the
// output's adjoint away (the kernel overwrites it).
template <int D>
struct B_div_massflux_part {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int i = x - 3, j = y - 3, kk = z;
    if (i > g.nx + 3 - 1 || j > g.ny + 3 - 1 || kk >= g.nz) return;
    const P rru = r.f(1), rrv = r.f(2), rrw = r.f(3), mh = r.f(4), hm = r.f(30);
    const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, kk);
    if (D == 0) r.keep<0>(r.f(5), c, ((r.ld(rru, c + 1) - r.ld(rru, c)) * D1(g.rdx)) * D1(hm_uh(hm, g, i)));
    else if (D == 1) r.keep<0>(r.f(5), c, ((r.ld(rrv, c + sj) - r.ld(rrv, c)) * D1(g.rdy)) * D1(hm_vh(hm, g, j)));
    else if (D == 2) r.keep<0>(r.f(5), c, ((r.ld(rrw, c + sk) - r.ld(rrw, c)) * D1(g.rdz)) * D1(mh[kk]));
    else r.out<0>(r.f(5), c, D1(0.0f));
  }
};
using B_div_massflux = Parts<B_div_massflux_part<0>, B_div_massflux_part<1>, B_div_massflux_part<2>,
                             B_div_massflux_part<3>>;

// divergence3d.wgsl:125-147
struct B_divergence3d {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int i = x, j = y, kk = z;
    if (i >= g.nx || j >= g.ny || kk >= g.nz) return;
    const P u = r.f(1), v = r.f(2), w = r.f(3), mh = r.f(4), hm = r.f(30);
    const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, kk);
    const D1 dz = (r.ld(w, c + sk) - r.ld(w, c)) * D1(g.rdz);
    const D1 dx = (r.ld(u, c + 1) - r.ld(u, c)) * D1(g.rdx);
    const D1 dy = (r.ld(v, c + sj) - r.ld(v, c)) * D1(g.rdy);
    const D1 d = fma_(dz, D1(mh[kk]), fma_(dx, D1(hm_uh(hm, g, i)), dy * D1(hm_vh(hm, g, j))));
    r.out<0>(r.f(5), c, ab(d).v < 1e-30f ? D1(0.0f) : d);
  }
};

// buoyancy_apply.wgsl:111-129
struct B_buoyancy_apply {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int i = x, j = y, kk = z + 1;
    if (i >= g.nx || j >= g.ny || kk > g.nz - 1) return;
    const P buoy = r.f(1), c1 = r.f(2), c2 = r.f(3), wten = r.f(4);
    const int64_t sk = g.sk(), c = g.idx(i, j, kk);
    const D1 b = fma_(D1(c1[kk]), r.ld(buoy, c - sk), D1(c2[kk]) * r.ld(buoy, c));
    r.out<0>(wten, c, r.ld(wten, c) + b);
  }
};

// rk_update.wgsl:111-132
struct B_rk_update {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int stag = r.I(1);
    const int i = x, j = y, kk = z;
    const int iMax = stag == 2 ? g.nx : g.nx - 1, jMax = stag == 3 ? g.ny : g.ny - 1, kLo = stag == 4 ? 1 : 0;
    if (i > iMax || j > jMax || kk < kLo || kk >= g.nz) return;
    const int64_t c = g.idx(i, j, kk);
    r.out<0>(r.f(4), c, fma_(D1(r.F(5)), r.ld(r.f(3), c), r.ld(r.f(2), c)));
  }
};

// build_sadv.wgsl:111-127
struct B_build_sadv {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int i = x - 3, j = y - 3, kk = z;
    if (i > g.nx + 3 || j > g.ny + 3 || kk >= g.nz) return;
    const int64_t c = g.idx(i, j, kk);
    r.out<0>(r.f(3), c, (D1(r.f(1)[kk]) - D1(3e+02f)) + r.ld(r.f(2), c));
  }
};

// apply_thterm.wgsl:111-127
struct B_apply_thterm {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    const int64_t c = g.idx(x, y, z);
    // tend += -div3d * thterm: the identity on tend's adjoint (keep)
    r.keep<0>(r.f(3), c, -r.ld(r.f(1), c) * r.ld(r.f(2), c));
  }
};

// turb_add_tendency.wgsl:111-127
struct B_turb_add_tendency {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    const int i = x - 3, j = y - 3, kk = z;
    if (i > g.nx + 3 || j > g.ny + 3 || kk > g.nz) return;
    const int64_t c = g.idx(i, j, kk);
    // dst += scale * src: the identity on dst's adjoint (keep; src may be dst)
    r.keep<0>(r.f(2), c, D1(r.F(3)) * r.ld(r.f(1), c));
  }
};

// turb_zero.wgsl:103-116 (raw word index over the whole field)
struct B_turb_zero {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    if (x >= g.sJ || y >= g.sK || z > g.nz + 1) return;
    r.out<0>(r.f(1), x + (int64_t)g.sJ * (y + (int64_t)g.sK * z), D1(0.0f));
  }
};

// ppx_finish.wgsl:111-126
struct B_ppx_finish {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    const int64_t c = g.idx(x, y, z);
    const P ppx = r.f(2);
    r.out<0>(ppx, c, r.ld(r.f(1), c) + r.ld(ppx, c));
  }
};

// diabatic_unapply.wgsl:111-140
struct B_diabatic_unapply {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    const int64_t c = g.idx(x, y, z);
    const D1 mdt = D1(-r.F(1));
    const P pp3d = r.f(2), th = r.f(3), qv = r.f(4), qc = r.f(5), qpten = r.f(6), qtten = r.f(7),
                qvten = r.f(8), qcten = r.f(9);
    const D1 p = fma_(mdt, r.ld(qpten, c), r.ld(pp3d, c));
    const D1 t = fma_(mdt, r.ld(qtten, c), r.ld(th, c));
    const D1 v = fma_(mdt, r.ld(qvten, c), r.ld(qv, c));
    const D1 q = fma_(mdt, r.ld(qcten, c), r.ld(qc, c));
    r.out<0>(pp3d, c, p); r.out<1>(qpten, c, p);
    r.out<2>(th, c, t); r.out<3>(qtten, c, t);
    r.out<4>(qv, c, v); r.out<5>(qvten, c, v);
    r.out<6>(qc, c, q); r.out<7>(qcten, c, q);
  }
};

// diabatic_rates.wgsl:111-137
struct B_diabatic_rates {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    const int64_t c = g.idx(x, y, z);
    const D1 rdt = D1(1.0f) / D1(r.F(1));
    const P pp3d = r.f(2), th = r.f(3), qv = r.f(4), qc = r.f(5), qpten = r.f(6), qtten = r.f(7),
                qvten = r.f(8), qcten = r.f(9);
    r.out<0>(qpten, c, (r.ld(pp3d, c) - r.ld(qpten, c)) * rdt);
    r.out<1>(qtten, c, (r.ld(th, c) - r.ld(qtten, c)) * rdt);
    r.out<2>(qvten, c, (r.ld(qv, c) - r.ld(qvten, c)) * rdt);
    r.out<3>(qcten, c, (r.ld(qc, c) - r.ld(qcten, c)) * rdt);
  }
};

// pressure_tendency.wgsl:234-291
struct B_pressure_tendency {
  static constexpr bool kReload = true;
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    const int64_t c = g.idx(x, y, z);
    const P th = r.f(1), qv = r.f(2), rho = r.f(3), ppi = r.f(4), thten = r.f(5), qpten = r.f(6),
                th0 = r.f(7), pi0 = r.f(8), ppten = r.f(11), qvt = r.f(13);
    const D1 dt = D1(r.F(9));
    const int hasDiabatic = r.I(10), hasStepTend = r.I(12), moist = r.I(14);
    const D1 rd = RD, rv = RV, cp = CP, p00 = P00;
    const D1 rp00 = D1(1.0f) / p00;
    const D1 rddcv = rd / (cp - rd);
    const float tem = 0.0001f * 0.0001f, qsmall = 1e-14f;
    const D1 tht = hasStepTend != 0 ? r.ld(thten, c) : D1(0.0f);
    if (moist == 0) {
      D1 p = D1(0.0f);
      if (ab(dt * tht).v > tem) {
        const D1 thnew = fma_(dt, tht, r.ld(th, c));
        const D1 pinew = powc(((r.ld(rho, c) * (D1(th0[z]) + thnew)) * rd) * rp00, rddcv.v) - D1(pi0[z]);
        p = (pinew - r.ld(ppi, c)) / dt;
      }
      r.out<0>(ppten, c, p);
      return;
    }
    const D1 qvten = r.ld(qvt, c);
    D1 p = D1(0.0f);
    if (ab(dt * tht).v > tem || ab(dt * qvten).v > qsmall) {
      const D1 thnew = fma_(dt, tht, r.ld(th, c));
      const D1 qvnew = fma_(dt, qvten, r.ld(qv, c));
      const D1 pinew = powc(((r.ld(rho, c) * (D1(th0[z]) + thnew)) * fma_(rv, qvnew, rd)) * rp00, rddcv.v) - D1(pi0[z]);
      p = (pinew - r.ld(ppi, c)) / dt;
    }
    if (hasDiabatic != 0) p = p + r.ld(qpten, c);
    r.out<0>(ppten, c, p);
  }
};

// update_rho_prs.wgsl:234-267
struct B_update_rho_prs {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    const ADims g = r.dims();
    if (x >= g.nx || y >= g.ny || z >= g.nz) return;
    const int64_t c = g.idx(x, y, z);
    const int moist = r.I(1);
    const P thP = r.f(2), ppP = r.f(3), th0 = r.f(4), pi0 = r.f(5), qv = r.f(6);
    const float cpdrd = CP / RD;
    const D1 pi = D1(pi0[z]) + r.ld(ppP, c);
    const D1 th = D1(th0[z]) + r.ld(thP, c);
    const D1 p = D1(P00) * powc(pi, cpdrd);
    r.out<0>(r.f(7), c, p);
    const D1 rmix = moist != 0 ? fma_(mx(D1(0.0f), r.ld(qv, c)), D1(RV), D1(RD)) : D1(RD);
    r.out<1>(r.f(8), c, p / ((th * pi) * rmix));
  }
};

// ---- advection fluxes and assembly

// adv_flux_x.wgsl:120-143 / adv_flux_y.wgsl:120-144 (DIR 0 / 1)
template <int DIR>
struct B_adv_flux_h {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (DIR == 0 ? (i > g.nx || j >= g.ny || kk >= g.nz) : (i >= g.nx || j > g.ny || kk >= g.nz)) return;
    const P s = r.f(1), rr = r.f(2);
    const int64_t st = DIR == 0 ? 1 : (int64_t)g.sJ, c = g.idx(i, j, kk);
    const D1 u = r.ld(rr, c);
    D1 f;
    if (u.v >= 0.0f) f = flx5(r.ld(s, c - 3 * st), r.ld(s, c - 2 * st), r.ld(s, c - st), r.ld(s, c), r.ld(s, c + st));
    else f = flx5(r.ld(s, c + 2 * st), r.ld(s, c + st), r.ld(s, c), r.ld(s, c - st), r.ld(s, c - 2 * st));
    r.out<0>(r.f(3), c, u * f);
  }
};

// adv_flux_z.wgsl:127-200
struct B_adv_flux_z {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i >= g.nx || j >= g.ny || kk > g.nz) return;
    const P s = r.f(1), rrw = r.f(2), c1 = r.f(3), c2 = r.f(4), fz = r.f(5);
    const int64_t sk = g.sk(), c = g.idx(i, j, kk);
    if (kk == 0 || kk == g.nz) { r.out<0>(fz, c, D1(0.0f)); return; }
    const D1 w = r.ld(rrw, c);
    const bool up = w.v >= 0.0f;
    auto S = [&](int n) { return r.ld(s, c + n * sk); };
    D1 f;
    if (kk == 1) f = up ? fma_(D1(c1[kk]), S(-1), D1(c2[kk]) * S(0)) : flx3(S(1), S(0), S(-1));
    else if (kk == 2) f = up ? flx3(S(-2), S(-1), S(0)) : flx5(S(2), S(1), S(0), S(-1), S(-2));
    else if (kk == g.nz - 2) f = up ? flx5(S(-3), S(-2), S(-1), S(0), S(1)) : flx3(S(1), S(0), S(-1));
    else if (kk == g.nz - 1) f = up ? flx3(S(-2), S(-1), S(0)) : fma_(D1(c1[kk]), S(-1), D1(c2[kk]) * S(0));
    else f = up ? flx5(S(-3), S(-2), S(-1), S(0), S(1)) : flx5(S(2), S(1), S(0), S(-1), S(-2));
    r.out<0>(fz, c, w * f);
  }
};

// adv_flux_x_weno.wgsl:158-200 / adv_flux_y_weno.wgsl:158-201 (DIR 0 / 1)
template <int DIR>
struct B_adv_flux_h_weno {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i > g.nx || j > g.ny || kk >= g.nz) return;
    const P a = r.f(1), rr = r.f(2), fl = r.f(5);
    const float weps = r.F(3);
    const bool pdef = r.I(4) == 1;
    const int64_t st = DIR == 0 ? 1 : (int64_t)g.sJ, c = g.idx(i, j, kk);
    const D1 m = r.ld(rr, c);
    D1 s1, s2, s3, s4, s5;
    if (m.v >= 0.0f) {
      s1 = r.ld(a, c - 3 * st); s2 = r.ld(a, c - 2 * st); s3 = r.ld(a, c - st); s4 = r.ld(a, c); s5 = r.ld(a, c + st);
    } else {
      s1 = r.ld(a, c + 2 * st); s2 = r.ld(a, c + st); s3 = r.ld(a, c); s4 = r.ld(a, c - st); s5 = r.ld(a, c - 2 * st);
    }
    r.out<0>(fl, c, weno_pdef_ok(s1, s2, s3, s4, s5, weps, pdef) ? m * weno5(s1, s2, s3, s4, s5, weps) : D1(0.0f));
  }
};

// adv_flux_z_weno.wgsl:174-279
struct B_adv_flux_z_weno {
  static constexpr bool kReload = true;
  static TE_HD void run(Acc& r, int i, int j, int z) {
    const ADims g = r.dims();
    const int kf = z + 1;
    if (i >= g.nx || j >= g.ny || kf > g.nz + 1) return;
    const int kk = kf - 1;
    const P a = r.f(1), rrw = r.f(2), c1 = r.f(3), c2 = r.f(4), fz = r.f(8);
    const float weps = r.F(5);
    const int pdefweno = r.I(7);
    const int64_t sk = g.sk(), c = g.idx(i, j, kk);
    if (kf == 1 || kf == g.nz + 1) { r.out<0>(fz, c, D1(0.0f)); return; }
    auto L = [&](int n) { return r.ld(a, c + n * sk); };
    const D1 w = r.ld(rrw, c);
    const int nk = g.nz;
    if (kf == 2) {
      r.out<0>(fz, c, (w.v >= 0.0f) ? w * fma_(D1(c1[kk]), L(-1), D1(c2[kk]) * L(0)) : w * upstrpd(L(1), L(0), L(-1), weps));
      return;
    }
    if (kf == nk) {
      r.out<0>(fz, c, (w.v > 0.0f) ? w * upstrpd(L(-2), L(-1), L(0), weps) : w * fma_(D1(c1[kk]), L(-1), D1(c2[kk]) * L(0)));
      return;
    }
    D1 s1, s2, s3, s4, s5;
    if (w.v >= 0.0f) { s1 = L(-3); s2 = L(-2); s3 = L(-1); s4 = L(0); s5 = L(1); }
    else { s1 = L(2); s2 = L(1); s3 = L(0); s4 = L(-1); s5 = L(-2); }
    if (pdefweno == 1) {
      bool doit = weno_pdef_ok(s1, s2, s3, s4, s5, weps, true);
      if (kf == 3 && w.v > 0.0f) doit = false;
      if (kf == nk - 1 && w.v < 0.0f) doit = false;
      if (doit) { r.out<0>(fz, c, w * weno5(s1, s2, s3, s4, s5, weps)); return; }
      if (kf != 3 && kf != nk - 1) { r.out<0>(fz, c, D1(0.0f)); return; }
    }
    if (kf == 3) {
      r.out<0>(fz, c, (w.v >= 0.0f) ? w * upstrpd(L(-2), L(-1), L(0), weps) : w * weno5(L(2), L(1), L(0), L(-1), L(-2), weps));
      return;
    }
    if (kf == nk - 1) {
      r.out<0>(fz, c, (w.v >= 0.0f) ? w * weno5(L(-3), L(-2), L(-1), L(0), L(1), weps) : w * upstrpd(L(1), L(0), L(-1), weps));
      return;
    }
    r.out<0>(fz, c, w * weno5(s1, s2, s3, s4, s5, weps));
  }
};

// f_flx5_upwind (adv_flux_x_stag.wgsl:106-128)
TE_HD D1 flx5_upwind(Acc& r, P a, int64_t c, int64_t s, bool up, bool shifted) {
  if (shifted)
    return up ? flx5(r.ld(a, c - 2 * s), r.ld(a, c - s), r.ld(a, c), r.ld(a, c + s), r.ld(a, c + 2 * s))
              : flx5(r.ld(a, c + 3 * s), r.ld(a, c + 2 * s), r.ld(a, c + s), r.ld(a, c), r.ld(a, c - s));
  return up ? flx5(r.ld(a, c - 3 * s), r.ld(a, c - 2 * s), r.ld(a, c - s), r.ld(a, c), r.ld(a, c + s))
            : flx5(r.ld(a, c + 2 * s), r.ld(a, c + s), r.ld(a, c), r.ld(a, c - s), r.ld(a, c - 2 * s));
}
// f_weno5_upwind (adv_flux_x_weno_uv.wgsl:106-129)
TE_HD D1 weno5_upwind(Acc& r, P a, int64_t c, int64_t s, bool up, bool shifted, float weps) {
  if (shifted)
    return up ? weno5(r.ld(a, c - 2 * s), r.ld(a, c - s), r.ld(a, c), r.ld(a, c + s), r.ld(a, c + 2 * s), weps)
              : weno5(r.ld(a, c + 3 * s), r.ld(a, c + 2 * s), r.ld(a, c + s), r.ld(a, c), r.ld(a, c - s), weps);
  return up ? weno5(r.ld(a, c - 3 * s), r.ld(a, c - 2 * s), r.ld(a, c - s), r.ld(a, c), r.ld(a, c + s), weps)
            : weno5(r.ld(a, c + 2 * s), r.ld(a, c + s), r.ld(a, c), r.ld(a, c - s), r.ld(a, c - 2 * s), weps);
}

// adv_flux_x_stag.wgsl:144-186 / adv_flux_y_stag.wgsl (DIR 0 / 1)
template <int DIR>
struct B_adv_flux_h_stag {
  static TE_HD void run(Acc& r, int x, int y, int kk) {
    const ADims g = r.dims();
    const int i = x - 1, j = y - 1;
    if (i > g.nx || j > g.ny || kk >= g.nz) return;
    const int stag = r.I(1);
    const P a = r.f(2), rr = DIR == 0 ? r.f(3) : r.f(4), c2 = r.f(7), fl = r.f(8);
    const int64_t sj = g.sJ, sk = g.sk(), st = DIR == 0 ? 1 : sj, c = g.idx(i, j, kk);
    if (stag == 4 && kk == 0) { r.out<0>(fl, c, D1(0.0f)); return; }
    D1 m;
    if (stag == 2) m = D1(0.5f) * (r.ld(rr, c) + r.ld(rr, c + (DIR == 0 ? 1 : -1)));
    else if (stag == 3) m = D1(0.5f) * (r.ld(rr, c) + r.ld(rr, c + (DIR == 0 ? -sj : sj)));
    else if (stag == 4) { const D1 cc2 = D1(c2[kk]); m = fma_(cc2, r.ld(rr, c), (D1(1.0f) - cc2) * r.ld(rr, c - sk)); }
    else m = r.ld(rr, c);
    r.out<0>(fl, c, m * flx5_upwind(r, a, c, st, m.v >= 0.0f, stag == (DIR == 0 ? 2 : 3)));
  }
};

// adv_flux_z_stag.wgsl:127-268
struct B_adv_flux_z_stag {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i > g.nx || j > g.ny || kk > g.nz) return;
    const int stag = r.I(1);
    const P a = r.f(2), rrw = r.f(5), c1 = r.f(6), c2 = r.f(7), fz = r.f(8);
    const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, kk);
    // the forward kernel forms the face mass flux before its range checks
    D1 m;
    if (stag == 2) m = D1(0.5f) * (r.ld(rrw, c) + r.ld(rrw, c - 1));
    else if (stag == 3) m = D1(0.5f) * (r.ld(rrw, c) + r.ld(rrw, c - sj));
    else m = D1(0.5f) * (r.ld(rrw, c) + r.ld(rrw, c + sk));
    const bool up = m.v >= 0.0f;
    auto A_ = [&](int n) { return r.ld(a, c + n * sk); };
    D1 f;
    if (stag == 4) {
      if (kk > g.nz - 1) { r.out<0>(fz, c, D1(0.0f)); return; }
      if (kk == 0) f = up ? D1(0.5f) * (A_(0) + A_(1)) : flx3(A_(2), A_(1), A_(0));
      else if (kk == 1) f = up ? flx3(A_(-1), A_(0), A_(1)) : flx5(A_(3), A_(2), A_(1), A_(0), A_(-1));
      else if (kk == g.nz - 2) f = up ? flx5(A_(-2), A_(-1), A_(0), A_(1), A_(2)) : flx3(A_(2), A_(1), A_(0));
      else if (kk == g.nz - 1) f = up ? flx3(A_(-1), A_(0), A_(1)) : D1(0.5f) * (A_(0) + A_(1));
      else f = up ? flx5(A_(-2), A_(-1), A_(0), A_(1), A_(2)) : flx5(A_(3), A_(2), A_(1), A_(0), A_(-1));
    } else {
      if (kk == 0 || kk == g.nz) { r.out<0>(fz, c, D1(0.0f)); return; }
      if (kk == 1) f = up ? fma_(D1(c1[kk]), A_(-1), D1(c2[kk]) * A_(0)) : flx3(A_(1), A_(0), A_(-1));
      else if (kk == 2) f = up ? flx3(A_(-2), A_(-1), A_(0)) : flx5(A_(2), A_(1), A_(0), A_(-1), A_(-2));
      else if (kk == g.nz - 2) f = up ? flx5(A_(-3), A_(-2), A_(-1), A_(0), A_(1)) : flx3(A_(1), A_(0), A_(-1));
      else if (kk == g.nz - 1) f = up ? flx3(A_(-2), A_(-1), A_(0)) : fma_(D1(c1[kk]), A_(-1), D1(c2[kk]) * A_(0));
      else f = up ? flx5(A_(-3), A_(-2), A_(-1), A_(0), A_(1)) : flx5(A_(2), A_(1), A_(0), A_(-1), A_(-2));
    }
    r.out<0>(fz, c, m * f);
  }
};

// adv_flux_x_weno_w.wgsl:143-175 / adv_flux_y_weno_w.wgsl (DIR 0 / 1)
template <int DIR>
struct B_adv_flux_h_weno_w {
  static TE_HD void run(Acc& r, int x, int y, int kk) {
    const ADims g = r.dims();
    const int i = x - 1, j = y - 1;
    if (i < 0 || j < 0 || i > g.nx || j > g.ny || kk >= g.nz) return;
    const P a = r.f(1), rr = r.f(2), c2 = r.f(3), fl = r.f(5);
    const float weps = r.F(4);
    const int64_t sk = g.sk(), st = DIR == 0 ? 1 : (int64_t)g.sJ, c = g.idx(i, j, kk);
    if (kk == 0) { r.out<0>(fl, c, D1(0.0f)); return; }
    const D1 cc2 = D1(c2[kk]), cc1 = D1(1.0f) - cc2;
    const D1 ub = fma_(cc2, r.ld(rr, c), cc1 * r.ld(rr, c - sk));
    r.out<0>(fl, c, ub * weno5_upwind(r, a, c, st, ub.v >= 0.0f, false, weps));
  }
};

// adv_flux_z_weno_w.wgsl:159-231
struct B_adv_flux_z_weno_w {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i > g.nx || j > g.ny || kk > g.nz) return;
    const P a = r.f(1), rrw = r.f(2), fz = r.f(4);
    const float weps = r.F(3);
    const int64_t sk = g.sk(), c = g.idx(i, j, kk);
    if (kk > g.nz - 1) { r.out<0>(fz, c, D1(0.0f)); return; }
    auto L = [&](int n) { return r.ld(a, c + n * sk); };
    const D1 wbar = D1(0.5f) * (r.ld(rrw, c) + r.ld(rrw, c + sk));
    const int nz = g.nz;
    D1 f;
    if (kk == 0) f = (wbar.v >= 0.0f) ? D1(0.5f) * (L(0) + L(1)) : upstrpd(L(2), L(1), L(0), weps);
    else if (kk == nz - 1) f = (wbar.v > 0.0f) ? upstrpd(L(-1), L(0), L(1), weps) : D1(0.5f) * (L(0) + L(1));
    else if (kk == 1) f = (wbar.v >= 0.0f) ? upstrpd(L(-1), L(0), L(1), weps) : weno5(L(3), L(2), L(1), L(0), L(-1), weps);
    else if (kk == nz - 2) f = (wbar.v > 0.0f) ? weno5(L(-2), L(-1), L(0), L(1), L(2), weps) : upstrpd(L(2), L(1), L(0), weps);
    else f = (wbar.v >= 0.0f) ? weno5(L(-2), L(-1), L(0), L(1), L(2), weps) : weno5(L(3), L(2), L(1), L(0), L(-1), weps);
    r.out<0>(fz, c, wbar * f);
  }
};

// adv_flux_x_weno_uv.wgsl:168-188 / adv_flux_y_weno_uv.wgsl (DIR 0 / 1)
template <int DIR>
struct B_adv_flux_h_weno_uv {
  static TE_HD void run(Acc& r, int x, int y, int kk) {
    const ADims g = r.dims();
    const int i = x - 1, j = y - 1;
    if (i > g.nx || j > g.ny || kk >= g.nz) return;
    const int stag = r.I(1);
    const P a = r.f(2), rr = r.f(3), fl = r.f(5);
    const float weps = r.F(4);
    const int64_t sj = g.sJ, st = DIR == 0 ? 1 : sj, c = g.idx(i, j, kk);
    D1 m;
    if (DIR == 0) m = (stag == 2) ? D1(0.5f) * (r.ld(rr, c) + r.ld(rr, c + 1)) : D1(0.5f) * (r.ld(rr, c) + r.ld(rr, c - sj));
    else m = (stag == 2) ? D1(0.5f) * (r.ld(rr, c) + r.ld(rr, c - 1)) : D1(0.5f) * (r.ld(rr, c) + r.ld(rr, c + sj));
    r.out<0>(fl, c, m * weno5_upwind(r, a, c, st, m.v >= 0.0f, stag == (DIR == 0 ? 2 : 3), weps));
  }
};

// adv_flux_z_weno_uv.wgsl:159-236
struct B_adv_flux_z_weno_uv {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i > g.nx || j > g.ny || kk > g.nz) return;
    const int stag = r.I(1);
    const P a = r.f(2), rrw = r.f(3), c1 = r.f(4), c2 = r.f(5), fz = r.f(7);
    const float weps = r.F(6);
    const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, kk);
    if (kk == 0 || kk == g.nz) { r.out<0>(fz, c, D1(0.0f)); return; }
    auto L = [&](int n) { return r.ld(a, c + n * sk); };
    const D1 m = (stag == 2) ? D1(0.5f) * (r.ld(rrw, c) + r.ld(rrw, c - 1)) : D1(0.5f) * (r.ld(rrw, c) + r.ld(rrw, c - sj));
    const bool up = m.v >= 0.0f;
    D1 f;
    if (kk == 1) f = up ? fma_(D1(c1[kk]), L(-1), D1(c2[kk]) * L(0)) : upstrpd(L(1), L(0), L(-1), weps);
    else if (kk == 2) f = up ? upstrpd(L(-2), L(-1), L(0), weps) : weno5(L(2), L(1), L(0), L(-1), L(-2), weps);
    else if (kk == g.nz - 2) f = up ? weno5(L(-3), L(-2), L(-1), L(0), L(1), weps) : upstrpd(L(1), L(0), L(-1), weps);
    else if (kk == g.nz - 1) f = up ? upstrpd(L(-2), L(-1), L(0), weps) : fma_(D1(c1[kk]), L(-1), D1(c2[kk]) * L(0));
    else f = up ? weno5(L(-3), L(-2), L(-1), L(0), L(1), weps) : weno5(L(2), L(1), L(0), L(-1), L(-2), weps);
    r.out<0>(fz, c, m * f);
  }
};

// adv_assemble.wgsl:139-279
struct B_adv_assemble {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    const int stag = r.I(1);
    const int st = stag < 0 ? -stag : stag;
    const int iMax = st == 2 ? g.nx : g.nx - 1, jMax = st == 3 ? g.ny : g.ny - 1, kLo = st == 4 ? 1 : 0,
              kMax = g.nz - 1;
    if (i > iMax || j > jMax || kk < kLo || kk > kMax) return;
    const int accumulate = r.I(12);
    const int openx = r.I(17, 0), openy = r.I(17, 1);
    if (openx != 0 && st == 2 && (i == 0 || i == g.nx)) return;
    if (openy != 0 && st == 3 && (j == 0 || j == g.ny)) return;
    const P s = r.f(2), fx = r.f(3), fy = r.f(4), fz = r.f(5), divx = r.f(6), rho0 = r.f(7), rf0 = r.f(8),
                mh = r.f(9), mf = r.f(10), tend = r.f(11), rru = r.f(13), rrv = r.f(14), c1 = r.f(15),
                c2 = r.f(16), hm = r.f(30);
    const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, kk);
    D1 dFx = st == 2 ? r.ld(fx, c) - r.ld(fx, c - 1) : r.ld(fx, c + 1) - r.ld(fx, c);
    D1 dFy = st == 3 ? r.ld(fy, c) - r.ld(fy, c - sj) : r.ld(fy, c + sj) - r.ld(fy, c);
    const D1 dFz = st == 4 ? r.ld(fz, c) - r.ld(fz, c - sk) : r.ld(fz, c + sk) - r.ld(fz, c);
    const D1 mz = stag < 0 ? D1(1.0f) : (stag == 4 ? D1(mf[kk]) : D1(mh[kk]));
    D1 dv;
    if (st == 2) dv = D1(0.5f) * (r.ld(divx, c) + r.ld(divx, c - 1));
    else if (st == 3) dv = D1(0.5f) * (r.ld(divx, c) + r.ld(divx, c - sj));
    else if (st == 4) dv = fma_(D1(c2[kk]), r.ld(divx, c), D1(c1[kk]) * r.ld(divx, c - sk));
    else dv = r.ld(divx, c);
    if (openx != 0 && st != 2) {
      const bool west = i == 0, east = i == g.nx - 1;
      if (west || east) {
        const int64_t cf = west ? c : c + 1;
        D1 un, piece;
        if (st == 3) {
          un = r.ld(rru, cf) + r.ld(rru, cf - sj);
          piece = ((D1(0.5f) * ((r.ld(rru, c + 1 - sj) - r.ld(rru, c - sj)) + (r.ld(rru, c + 1) - r.ld(rru, c)))) * D1(g.rdx)) *
                  D1(hm_uh(hm, g, i));
        } else if (st == 4) {
          const D1 cc2 = D1(0.5f) * (D1(c2[kk - 1]) + D1(c2[kk])), cc1 = D1(1.0f) - cc2;
          un = fma_(cc2, r.ld(rru, cf), cc1 * r.ld(rru, cf - sk));
          piece = (fma_(D1(c1[kk]), r.ld(rru, c + 1 - sk) - r.ld(rru, c - sk), D1(c2[kk]) * (r.ld(rru, c + 1) - r.ld(rru, c))) *
                   D1(g.rdx)) * D1(hm_uh(hm, g, i));
        } else {
          un = r.ld(rru, cf);
          piece = ((r.ld(rru, c + 1) - r.ld(rru, c)) * D1(g.rdx)) * D1(hm_uh(hm, g, i));
        }
        const bool inflow = west ? un.v >= 0.0f : un.v <= 0.0f;
        if (inflow) { dFx = D1(0.0f); dv = dv - piece; }
      }
    }
    if (openy != 0 && st != 3) {
      const bool south = j == 0, north = j == g.ny - 1;
      if (south || north) {
        const int64_t cf = south ? c : c + sj;
        D1 vn, piece;
        if (st == 2) {
          vn = r.ld(rrv, cf) + r.ld(rrv, cf - 1);
          piece = ((D1(0.5f) * ((r.ld(rrv, c + sj - 1) - r.ld(rrv, c - 1)) + (r.ld(rrv, c + sj) - r.ld(rrv, c)))) * D1(g.rdy)) *
                  D1(hm_vh(hm, g, j));
        } else if (st == 4) {
          const D1 cc2 = D1(0.5f) * (D1(c2[kk - 1]) + D1(c2[kk])), cc1 = D1(1.0f) - cc2;
          vn = fma_(cc2, r.ld(rrv, cf), cc1 * r.ld(rrv, cf - sk));
          piece = (fma_(D1(c1[kk]), r.ld(rrv, c + sj - sk) - r.ld(rrv, c - sk), D1(c2[kk]) * (r.ld(rrv, c + sj) - r.ld(rrv, c))) *
                   D1(g.rdy)) * D1(hm_vh(hm, g, j));
        } else {
          vn = r.ld(rrv, cf);
          piece = ((r.ld(rrv, c + sj) - r.ld(rrv, c)) * D1(g.rdy)) * D1(hm_vh(hm, g, j));
        }
        const bool inflow = south ? vn.v >= 0.0f : vn.v <= 0.0f;
        if (inflow) { dFy = D1(0.0f); dv = dv - piece; }
      }
    }
    const D1 mxm = D1(st == 2 ? hm_uf(hm, g, i) : hm_uh(hm, g, i));
    const D1 mym = D1(st == 3 ? hm_vf(hm, g, j) : hm_vh(hm, g, j));
    const D1 divF = fma_(dFz * D1(g.rdz), mz, fma_(dFx * D1(g.rdx), mxm, (dFy * D1(g.rdy)) * mym));
    const D1 rho = D1(st == 4 ? rf0[kk] : rho0[kk]);
    const D1 t = fma_(r.ld(s, c), dv, -divF) / rho;
    r.out<0>(tend, c, accumulate != 0 ? r.ld(tend, c) + t : t);
  }
};

// ---- acoustic substeps

// acoustic_prepare.wgsl:130-184 (pk1/pk2 operand order switched by vec_v).
// Its four outputs read disjoint loads (ppterm; piadv at k, k+1; piadv at
// k, k-1; thv), so each output is a body of its own: a seeded pass then
// evaluates one output, not all four.
template <int O>
struct B_acoustic_prepare_part {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i >= g.nx || j >= g.ny || kk >= g.nz) return;
    const float dts = r.F(1, 0);
    const P ppterm = r.f(2), piadv = r.f(3), thv = r.f(4), rr0 = r.f(5), rf0 = r.f(6), mh = r.f(7),
            mf = r.f(8), c1 = r.f(9), c2 = r.f(10);
    const int64_t sk = g.sk(), c = g.idx(i, j, kk);
    const D1 tem = (D1(dts) * D1(g.rdz)) * D1(mh[kk]);
    const D1 c1a = D1(c1[kk]), c1b = D1(c1[kk + 1]), c2a = D1(c2[kk]), c2b = D1(c2[kk + 1]);
    const bool vec_v = g.nx > 3 && i < (g.nx / 4) * 4;
    if (O == 0) {
      r.out<0>(r.f(11), c, tem * r.ld(ppterm, c));
    } else if (O == 1) {
      const D1 r2 = (tem * D1(rr0[kk])) * D1(rf0[kk + 1]);
      D1 p2 = D1(0.0f);
      if (kk != g.nz - 1) {
        p2 = vec_v ? fma_(-c2b, r.ld(piadv, c + sk), (D1(1.0f) - c1b) * r.ld(piadv, c))
                   : fma_(D1(1.0f) - c1b, r.ld(piadv, c), -(c2b * r.ld(piadv, c + sk)));
      }
      r.out<0>(r.f(13), c, kk == g.nz - 1 ? D1(0.0f) : r2 * p2);
    } else if (O == 2) {
      const D1 r1 = (tem * D1(rr0[kk])) * D1(rf0[kk]);
      D1 p1 = D1(0.0f);
      if (kk != 0) {
        p1 = vec_v ? fma_(c2a - D1(1.0f), r.ld(piadv, c), c1a * r.ld(piadv, c - sk))
                   : fma_(c1a, r.ld(piadv, c - sk), (c2a - D1(1.0f)) * r.ld(piadv, c));
      }
      r.out<0>(r.f(12), c, kk == 0 ? D1(0.0f) : r1 * p1);
    } else {
      const D1 tem1 = ((D1(dts) * D1(g.rdz)) * D1(CP)) * D1(mf[kk]);
      const D1 tc = r.ld(thv, c);
      r.out<0>(r.f(14), c, tem1 * fma_(c2a, tc, c1a * (kk == 0 ? tc : r.ld(thv, c - sk))));
    }
  }
};
using B_acoustic_prepare = Parts<B_acoustic_prepare_part<0>, B_acoustic_prepare_part<1>, B_acoustic_prepare_part<2>,
                                 B_acoustic_prepare_part<3>>;

// acoustic_begin.wgsl:128-149
struct B_acoustic_begin {
  static TE_HD void run(Acc& r, int x, int y, int kk) {
    const ADims g = r.dims();
    const int i = x - 3, j = y - 3;
    if (i > g.nx + 3 || j > g.ny + 3 || kk > g.nz) return;
    const int64_t c = g.idx(i, j, kk);
    const float beta = r.F(1, 2);
    r.out<0>(r.f(3), c, D1(0.0f));
    r.out<1>(r.f(4), c, D1(0.0f));
    r.out<2>(r.f(5), c, r.ld(r.f(2), c) * D1(beta));
  }
};

// acoustic_uv.wgsl:144-181
struct B_acoustic_uv {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i > g.nx || j > g.ny || kk >= g.nz) return;
    const D1 dts = D1(r.F(1, 0));
    const P ppd = r.f(2), thv = r.f(3), uten = r.f(4), vten = r.f(5), u = r.f(6), v = r.f(7), hm = r.f(30);
    const int openx = r.I(8, 0), openy = r.I(8, 1);
    const int64_t sj = g.sJ, c = g.idx(i, j, kk);
    const D1 cp = CP;
    const D1 tem1 = (D1(g.rdx) * cp) * D1(0.5f);
    const D1 tem2 = (D1(g.rdy) * cp) * D1(0.5f);
    const bool uEdge = openx != 0 && (i == 0 || i == g.nx);
    const bool vEdge = openy != 0 && (j == 0 || j == g.ny);
    if (!uEdge)
      r.out<0>(u, c, fma_(dts, fma_(-((tem1 * (r.ld(ppd, c) - r.ld(ppd, c - 1))) * D1(hm_uf(hm, g, i))),
                                    r.ld(thv, c) + r.ld(thv, c - 1), r.ld(uten, c)), r.ld(u, c)));
    if (!vEdge)
      r.out<1>(v, c, fma_(dts, fma_(-((tem2 * (r.ld(ppd, c) - r.ld(ppd, c - sj))) * D1(hm_vf(hm, g, j))),
                                    r.ld(thv, c) + r.ld(thv, c - sj), r.ld(vten, c)), r.ld(v, c)));
  }
};

// acoustic_fpk.wgsl:167-207
struct B_acoustic_fpk {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i >= g.nx || j >= g.ny || kk >= g.nz) return;
    const D1 dts = D1(r.F(1, 0)), beta = D1(r.F(1, 2));
    const float smeps = r.F(1, 4);
    const P u = r.f(2), v = r.f(3), w = r.f(4), piadv = r.f(5), ppten = r.f(6), ppterm = r.f(7),
                pk1 = r.f(8), pk2 = r.f(9), qk = r.f(10), hm = r.f(30);
    const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, kk);
    const D1 temx = (dts * D1(0.5f)) * D1(g.rdx);
    const D1 temy = (dts * D1(0.5f)) * D1(g.rdy);
    const D1 u0 = r.ld(u, c), u1 = r.ld(u, c + 1), v0 = r.ld(v, c), v1 = r.ld(v, c + sj);
    D1 div = fma_((u1 - u0) * D1(g.rdx), D1(hm_uh(hm, g, i)), ((v1 - v0) * D1(g.rdy)) * D1(hm_vh(hm, g, j)));
    if (ab(div).v < smeps) div = D1(0.0f);
    const D1 pc = r.ld(piadv, c);
    const D1 Au = fma_(u0, pc - r.ld(piadv, c - 1), u1 * (r.ld(piadv, c + 1) - pc));
    const D1 Av = fma_(v0, pc - r.ld(piadv, c - sj), v1 * (r.ld(piadv, c + sj) - pc));
    const D1 advPi = fma_((-Au) * temx, D1(hm_uh(hm, g, i)), -((Av * temy) * D1(hm_vh(hm, g, j))));
    const D1 q = r.ld(qk, c);
    r.out<0>(r.f(11), c, fma_(beta, fma_(r.ld(w, c), r.ld(pk1, c) + q, r.ld(w, c + sk) * (r.ld(pk2, c) - q)),
                              fma_(dts, fma_(-r.ld(ppterm, c), div, r.ld(ppten, c)), advPi)));
  }
};

// acoustic_pp.wgsl:131-167 (K_TERRAIN=false)
struct B_acoustic_pp {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i >= g.nx || j >= g.ny || kk >= g.nz) return;
    const D1 alph = D1(r.F(1, 1)), kdiv = D1(r.F(1, 3));
    const float smeps = r.F(1, 4);
    const P w = r.f(2), pk1 = r.f(3), pk2 = r.f(4), qk = r.f(5), fpk = r.f(6), pp3d = r.f(7), ppd = r.f(8),
                ppx = r.f(9);
    const int saveX = r.I(10);
    const int64_t sk = g.sk(), c = g.idx(i, j, kk);
    const D1 prev = r.ld(pp3d, c);
    const D1 q = r.ld(qk, c);
    D1 p = fma_(alph, fma_(r.ld(w, c), r.ld(pk1, c) + q, r.ld(w, c + sk) * (r.ld(pk2, c) - q)), prev + r.ld(fpk, c));
    if (ab(p).v < smeps) p = D1(0.0f);
    r.out<0>(pp3d, c, p);
    const D1 dum1 = kdiv * (p - prev);
    r.out<1>(ppd, c, p + dum1);
    if (saveX != 0) r.out<2>(ppx, c, dum1);
  }
};

// acoustic_accumulate.wgsl:129-158
struct B_acoustic_accumulate {
  static TE_HD void run(Acc& r, int x, int y, int kk) {
    const ADims g = r.dims();
    const int i = x - 3, j = y - 3;
    if (i > g.nx + 3 || j > g.ny + 3 || kk > g.nz) return;
    const D1 alph = D1(r.F(1, 1));
    const P u = r.f(2), v = r.f(3), w = r.f(4), rru = r.f(5), rrv = r.f(6), rrw = r.f(7);
    const float tavg = r.F(8);
    const int64_t c = g.idx(i, j, kk);
    if (tavg > 0.0f) {
      r.out<0>(rru, c, (r.ld(rru, c) + r.ld(u, c)) * D1(tavg));
      r.out<1>(rrv, c, (r.ld(rrv, c) + r.ld(v, c)) * D1(tavg));
      r.out<2>(rrw, c, fma_(r.ld(w, c), alph, r.ld(rrw, c)) * D1(tavg));
    } else {
      r.out<0>(rru, c, r.ld(rru, c) + r.ld(u, c));
      r.out<1>(rrv, c, r.ld(rrv, c) + r.ld(v, c));
      r.out<2>(rrw, c, r.ld(rrw, c) + r.ld(w, c));
    }
  }
};

// ---- positive-definite flux correction (pdef_dum.wgsl, pdef_apply.wgsl)
// f_pdef_inflow (pdef_dum.wgsl:108-134)
TE_HD bool pdef_inflow(const ADims& g, int dir, int i, int j, int64_t c, int64_t step, const float* rr, int openx,
                       int openy) {
  if (dir == 0 && openx != 0) {
    if (i == 0 && rr[c] >= 0.0f) return true;
    if (i == g.nx - 1 && rr[c + step] <= 0.0f) return true;
  }
  if (dir == 1 && openy != 0) {
    if (j == 0 && rr[c] >= 0.0f) return true;
    if (j == g.ny - 1 && rr[c + step] <= 0.0f) return true;
  }
  return false;
}

// pdef_dum.wgsl:177-231 (K_AXISYMM=false)
struct B_pdef_dum {
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i >= g.nx || j >= g.ny || kk >= g.nz) return;
    const int dir = r.I(1);
    const P s = r.f(2), s0 = r.f(3), f = r.f(4), rr = r.f(5), rho0 = r.f(6), mh = r.f(7), dum = r.f(10),
                hm = r.f(30);
    const D1 dt = D1(r.F(8));
    const int openx = r.I(9, 0), openy = r.I(9, 1);
    const int64_t c = g.idx(i, j, kk);
    const int64_t step = dir == 0 ? 1 : (dir == 1 ? (int64_t)g.sJ : g.sk());
    const D1 rd = D1(dir == 0 ? g.rdx : (dir == 1 ? g.rdy : g.rdz));
    const D1 base = D1(rho0[kk]) * r.ld(s0, c);
    if (pdef_inflow(g, dir, i, j, c, step, rr, openx, openy)) { r.out<0>(dum, c, base); return; }
    D1 adv = (-(r.ld(f, c + step) - r.ld(f, c))) * rd;
    D1 sdiv = (r.ld(s, c) * (r.ld(rr, c + step) - r.ld(rr, c))) * rd;
    const D1 m = D1(dir == 0 ? hm_uh(hm, g, i) : (dir == 1 ? hm_vh(hm, g, j) : mh[kk]));
    adv = adv * m;
    sdiv = sdiv * m;
    r.out<0>(dum, c, fma_(dt, adv + sdiv, base));
  }
};

// f_pdef_cell / f_pdef_mass (pdef_apply.wgsl:115-158)
struct PdefCell {
  bool neg;
  D1 f1, f2, foo3;
};
TE_HD PdefCell pdef_cell(D1 d, D1 dm, D1 dp) {
  PdefCell r{false, D1(0.0f), D1(0.0f), D1(0.0f)};
  if (d.v < 0.0f) {
    const D1 f1 = mx(D1(0.0f), dm), f2 = mx(D1(0.0f), dp);
    if ((f1 + f2).v > 1e-30f) {
      r.neg = true; r.f1 = f1; r.f2 = f2;
      r.foo3 = mx(d, -(f1 + f2)) / (f1 + f2);
    }
  }
  return r;
}
TE_HD D1 pdef_mass(D1 dm2, D1 dm1, D1 d0, D1 dp1, D1 dp2, bool& flag) {
  const PdefCell l = pdef_cell(dm1, dm2, d0), c = pdef_cell(d0, dm1, dp1), rr = pdef_cell(dp1, d0, dp2);
  D1 mass = D1(0.0f);
  if (l.neg) mass = fma_(l.f2, l.foo3, mass);
  if (c.neg) mass = fma_(-(c.f1 + c.f2), c.foo3, mass);
  if (rr.neg) mass = fma_(rr.f1, rr.foo3, mass);
  flag = c.neg || (l.neg && d0.v > 1e-30f) || (rr.neg && d0.v > 1e-30f);
  return mass;
}

// pdef_apply.wgsl:207-282.  The kernel adds the corrections of the three
// directions to the tendency (tend + (dx + dy + dz) / rho0), so its adjoint
// is the identity on tend plus the adjoint of each direction's share: three
// bodies with about a third of the active loads each.
template <int D>
struct B_pdef_part {
  static constexpr bool kReload = true;
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i >= g.nx || j >= g.ny || kk >= g.nz) return;
    const P s = r.f(1), s0 = r.f(2), fx = r.f(3), fy = r.f(4), fz = r.f(5), rru = r.f(6), rrv = r.f(7),
            rrw = r.f(8), dx = r.f(9), dy = r.f(10), dz = r.f(11), rho0 = r.f(12), mh = r.f(13), tend = r.f(16),
            hm = r.f(30);
    const float dt = r.F(14);
    const int openx = r.I(15, 0), openy = r.I(15, 1);
    const int64_t sj = g.sJ, sk = g.sk(), c = g.idx(i, j, kk);
    const D1 rdt = D1(1.0f) / D1(dt);
    bool flag;
    D1 delta;
    if (D == 0) {
      const D1 d0 = r.ld(dx, c);
      const D1 mxs = pdef_mass(r.ld(dx, c - 2), r.ld(dx, c - 1), d0, r.ld(dx, c + 1), r.ld(dx, c + 2), flag);
      if (!flag) return;
      const D1 base = D1(rho0[kk]) * r.ld(s0, c);
      const D1 sdiv = ((r.ld(s, c) * (r.ld(rru, c + 1) - r.ld(rru, c))) * D1(g.rdx)) * D1(hm_uh(hm, g, i));
      const D1 advOld = pdef_inflow(g, 0, i, j, c, 1, rru, openx, openy)
                            ? -sdiv
                            : ((-(r.ld(fx, c + 1) - r.ld(fx, c))) * D1(g.rdx)) * D1(hm_uh(hm, g, i));
      delta = fma_((d0 + mxs) - base, rdt, -sdiv) - advOld;
    } else if (D == 1) {
      const D1 d0 = r.ld(dy, c);
      const D1 mys = pdef_mass(r.ld(dy, c - 2 * sj), r.ld(dy, c - sj), d0, r.ld(dy, c + sj), r.ld(dy, c + 2 * sj), flag);
      if (!flag) return;
      const D1 base = D1(rho0[kk]) * r.ld(s0, c);
      const D1 sdiv = ((r.ld(s, c) * (r.ld(rrv, c + sj) - r.ld(rrv, c))) * D1(g.rdy)) * D1(hm_vh(hm, g, j));
      const D1 advOld = pdef_inflow(g, 1, i, j, c, sj, rrv, openx, openy)
                            ? -sdiv
                            : ((-(r.ld(fy, c + sj) - r.ld(fy, c))) * D1(g.rdy)) * D1(hm_vh(hm, g, j));
      delta = fma_((d0 + mys) - base, rdt, -sdiv) - advOld;
    } else {
      const D1 zm2 = kk >= 2 ? r.ld(dz, c - 2 * sk) : D1(0.0f);
      const D1 zm1 = kk >= 1 ? r.ld(dz, c - sk) : D1(0.0f);
      const D1 d0 = r.ld(dz, c);
      const D1 zp1 = kk + 1 < g.nz ? r.ld(dz, c + sk) : D1(0.0f);
      const D1 zp2 = kk + 2 < g.nz ? r.ld(dz, c + 2 * sk) : D1(0.0f);
      const D1 mzs = pdef_mass(zm2, zm1, d0, zp1, zp2, flag);
      if (!flag) return;
      const D1 base = D1(rho0[kk]) * r.ld(s0, c);
      const D1 sdiv = ((r.ld(s, c) * (r.ld(rrw, c + sk) - r.ld(rrw, c))) * D1(g.rdz)) * D1(mh[kk]);
      const D1 advOld = ((-(r.ld(fz, c + sk) - r.ld(fz, c))) * D1(g.rdz)) * D1(mh[kk]);
      delta = fma_((d0 + mzs) - base, rdt, -sdiv) - advOld;
    }
    r.keep<0>(tend, c, delta / D1(rho0[kk]));
  }
};
using B_pdef_apply = Parts<B_pdef_part<0>, B_pdef_part<1>, B_pdef_part<2>>;

// ---- halos, copies, damping

// bc_periodic_x.wgsl:111-133 / bc_periodic_y.wgsl:111-133 (DIR 0 / 1)
template <int DIR>
struct B_bc_periodic {
  static TE_HD void run(Acc& r, int h, int y, int kk) {
    const ADims g = r.dims();
    const int q = y - 3, n = DIR == 0 ? g.nx : g.ny;
    if (h > 3 || q > (DIR == 0 ? g.ny : g.nx) + 3 || kk > g.nz) return;
    const P f = r.f(1);
    const int lo = (((-1 - h) % n) + n) % n;
    const int hi = (((n + h) % n) + n) % n;
    auto at = [&](int p) { return DIR == 0 ? g.idx(p, q, kk) : g.idx(q, p, kk); };
    if (h < 3) r.out<0>(f, at(-1 - h), r.ld(f, at(lo)));
    r.out<1>(f, at(n + h), r.ld(f, at(hi)));
  }
};

// replay.js:403-406: copy A[0] words from parameter 0 to parameter 1
struct B_copy {
  static TE_HD void run(Acc& r, int x, int y, int z) {
    if ((uint32_t)x >= r.a->n[0] || y != 0 || z != 0) return;
    r.out<0>(r.f(1), x, r.ld(r.f(0), x));
  }
};

// rayleigh_damping.wgsl:111-164 (relaxation to the u0 / v0 profiles)
struct B_rayleigh_damping {
  static constexpr bool kReload = true;
  static TE_HD void run(Acc& r, int i, int j, int kk) {
    const ADims g = r.dims();
    if (i > g.nx || j > g.ny || kk > g.nz) return;
    const P u = r.f(1), v = r.f(2), w = r.f(3), th = r.f(4), u0 = r.f(5), v0 = r.f(6), tauh = r.f(7),
                tauf = r.f(8), hdx = r.f(14), hdy = r.f(15), avg = r.f(18), th0 = r.f(19);
    const D1 rdalpha = D1(r.F(9));
    const int lateral = r.I(16), toAverage = r.I(17);
    const int64_t c = g.idx(i, j, kk);
    if (kk < g.nz) {
      const D1 t = D1(tauh[kk]);
      const D1 ur = D1(toAverage != 0 ? avg[kk] : u0[kk]);
      const D1 vr = D1(toAverage != 0 ? avg[g.nz + kk] : v0[kk]);
      if (j < g.ny) r.out<0>(r.f(10), c, ((-rdalpha) * t) * (r.ld(u, c) - ur));
      if (i < g.nx) r.out<1>(r.f(11), c, ((-rdalpha) * t) * (r.ld(v, c) - vr));
      if (i < g.nx && j < g.ny) {
        const D1 d = toAverage != 0 ? (D1(th0[kk]) - D1(avg[2 * g.nz + kk])) + r.ld(th, c) : r.ld(th, c);
        r.out<2>(r.f(13), c, ((-rdalpha) * t) * d);
      }
    }
    if (i < g.nx && j < g.ny) {
      D1 tf = D1(tauf[kk]);
      if (lateral != 0 && kk < g.nz) tf = mx(mx(tf, D1(hdx[i])), D1(hdy[j]));
      r.out<3>(r.f(12), c, ((-rdalpha) * tf) * r.ld(w, c));
    }
  }
};


// ---- acoustic_w_solve (acoustic_w_solve.wgsl:124-212): adjoint of the
// per-column tridiagonal solve T w = d, rows k = 1..nz-1 (a_k w[k-1] + b_k w[k]
// + c_k w[k+1] = d_k, w[0] = w[nz] = 0; the forward sweep never uses a_1 or
// c_{nz-1}).  With wbar the adjoint of the new w:
//   dbar = T^-T wbar, through the forward factors T = L U (L: diagonal
//   1/r_k = b_k + a_k cA_{k-1}, subdiagonal a_k; U: unit diagonal,
//   superdiagonal -cA_k): U^T z = wbar, then L^T dbar = z;
//   abar_k = -dbar_k w[k-1], bbar_k = -dbar_k w[k], cbar_k = -dbar_k w[k+1];
// then a, b, c, d of each level are pushed back to their 17 loads at levels
// k and k-1 (seeded dual passes over the coefficient formulas).  One thread
// per column; every word it touches belongs to its own column, so no atomics.
// The adjoint words of cA and cB (the forward's elimination scratch, which no
// later dispatch reads) hold the recomputed factors and the new w meanwhile
// and are cleared at the end.
struct WCoef {
  D1 aa, bb, cc, dd;
};
// lam_1 (acoustic_w_solve.wgsl:124-134) on dual numbers
TE_HD WCoef w_coef(Acc& r, int64_t c, int64_t sk, float asq, float dts, float alph, float beta) {
  const P pk1 = r.f(2), pk2 = r.f(3), qk = r.f(4), mm = r.f(5), fpk = r.f(6), wten = r.f(7), pp3d = r.f(8),
              w = r.f(9);
  const int64_t cm = c - sk;
  WCoef q;
  const D1 m = r.ld(mm, c);
  q.aa = D1(-asq) * (m * (r.ld(pk1, cm) + r.ld(qk, cm)));
  q.cc = D1(-asq) * ((-m) * (r.ld(pk2, c) - r.ld(qk, c)));
  q.bb = fma_(D1(-asq), m * ((-(r.ld(pk1, c) + r.ld(qk, c))) + (r.ld(pk2, cm) - r.ld(qk, cm))), D1(1.0f));
  const D1 fw = fma_(D1(dts), r.ld(wten, c), D1(beta) * ((-m) * (r.ld(pp3d, c) - r.ld(pp3d, cm))));
  q.dd = fma_(D1(alph), (-m) * ((r.ld(pp3d, c) + r.ld(fpk, c)) - (r.ld(pp3d, cm) + r.ld(fpk, cm))), r.ld(w, c) + fw);
  return q;
}

// Coefficients of row k (k = 1..nz-1) at one (i, j), into four planes of the
// caller's scratch (a, b, c, d; plane stride fw words, field indexing).
TE_HD void w_solve_coef(const float* M, const AdjArg& a, float* scr, uint64_t fw, int i, int j, int k) {
  Acc r(M, nullptr, &a);
  const ADims g = r.dims();
  if (i >= g.nx || j >= g.ny || k < 1 || k > g.nz - 1) return;
  const float dts = r.F(1, 0), alph = r.F(1, 1), beta = r.F(1, 2);
  const int64_t sk = g.sk(), c = g.idx(i, j, k);
  const WCoef q = w_coef(r, c, sk, alph * alph, dts, alph, beta);
  scr[c] = (float)q.aa.v;
  scr[fw + c] = (float)q.bb.v;
  scr[2 * fw + c] = (float)q.cc.v;
  scr[3 * fw + c] = (float)q.dd.v;
}

// Column part: forward factors, the new w, U^T z = wbar and L^T dbar = z,
// from the coefficient planes (binary32, the forward's recurrences).  tA / tB
// (stride ts) hold cA and cB, then z, for the column (LDS on the device);
// the reciprocal pivots replace c in its plane.  Leaves dbar_k in the adjoint
// words of cA and the new w in those of cB (levels 1..nz-1) and clears wbar
// for the inputs' adjoints.
TE_HD void w_solve_sweep(const float* M, float* Mb, const AdjArg& a, float* scr, uint64_t fw, float* tA, float* tB,
                         int64_t ts, int i, int j, float* pre = nullptr, int64_t ps = 0) {
  Acc r(M, Mb, &a);
  const ADims g = r.dims();
  if (i >= g.nx || j >= g.ny) return;
  const int nz = g.nz;
  const int64_t sk = g.sk(), c0 = g.idx(i, j, 0);
  float* __restrict__ wb = Mb + a.p[9] + c0;   // adjoint of w (column)
  float* __restrict__ S1 = Mb + a.p[10] + c0;  // adjoint words of cA: dbar_k
  float* __restrict__ S2 = Mb + a.p[11] + c0;  // adjoint words of cB: the new w
  // the column's a, b, c (then 1 / pivot), d and wbar: the scratch planes and
  // wb by default, or a copy of them (5 arrays of stride ps, level-major
  // groups of nz + 1) the caller preloaded, e.g. into LDS
  const float *AA, *BB, *DD, *WB;
  float* RR;
  int64_t cs;
  if (pre) {
    AA = pre; BB = pre + (nz + 1) * ps; RR = pre + 2 * (nz + 1) * ps; DD = pre + 3 * (nz + 1) * ps;
    WB = pre + 4 * (nz + 1) * ps;
    cs = ps;
  } else {
    AA = scr + c0; BB = scr + fw + c0; RR = scr + 2 * fw + c0; DD = scr + 3 * fw + c0; WB = wb;
    cs = sk;
  }
  float cAp = 0.0f, cBp = 0.0f;
  for (int k = 1; k <= nz - 1; ++k) {
    const float aa = AA[k * cs], bb = BB[k * cs], cc = RR[k * cs], dd = DD[k * cs];
    const float rr = 1.0f / (k == 1 ? bb : fmaf_(aa, cAp, bb));
    cAp = (-cc) * rr;
    cBp = (k == 1 ? dd : fmaf_(-aa, cBp, dd)) * rr;
    RR[k * cs] = rr;
    tA[k * ts] = cAp;
    tB[k * ts] = cBp;
  }
  float wn = cBp;  // w[nz-1]
  S2[(nz - 1) * sk] = wn;
  for (int k = nz - 2; k >= 1; --k) {
    wn = fmaf_(tA[k * ts], wn, tB[k * ts]);
    S2[k * sk] = wn;
  }
  // U^T z = wbar (z in tB), levels 0 and nz are constant outputs
  float z = WB[cs];
  tB[ts] = z;
  for (int k = 2; k <= nz - 1; ++k) {
    z = fmaf_(tA[(k - 1) * ts], z, WB[k * cs]);
    tB[k * ts] = z;
  }
  // L^T dbar = z, from the top
  float aNext = 0.0f, yNext = 0.0f;
  for (int k = nz - 1; k >= 1; --k) {
    const float y = (k == nz - 1 ? tB[k * ts] : fmaf_(-aNext, yNext, tB[k * ts])) * RR[k * cs];
    S1[k * sk] = y;
    aNext = AA[k * cs];
    yNext = y;
  }
  for (int k = 0; k <= nz; ++k) wb[k * sk] = 0.0f;
}

// Level part, one word (i, j, k), k = 0..nz-1, after every column's sweep:
// the adjoints of row k's and row k+1's coefficients pushed back to the words
// of level k, in closed form (lam_1, acoustic_w_solve.wgsl:124-134, with
// m = mm[k], cm the level below a row):
//   a = -asq m (pk1[cm] + qk[cm]),  c = -asq (-m) (pk2 - qk),
//   b = 1 - asq m (-(pk1 + qk) + (pk2[cm] - qk[cm])),
//   d = w + dts wten - beta m (pp - pp[cm]) - alph m ((pp + fpk) - (pp[cm] + fpk[cm])),
// abar = -dbar w[k-1], bbar = -dbar w[k], cbar = -dbar w[k+1] (the new w).
TE_HD void w_solve_level(const float* M, float* Mb, const AdjArg& a, int i, int j, int k) {
  Acc r(M, Mb, &a);
  const ADims g = r.dims();
  const int nz = g.nz;
  if (i >= g.nx || j >= g.ny || k < 0 || k > nz - 1) return;
  const float dts = r.F(1, 0), alph = r.F(1, 1), beta = r.F(1, 2);
  const float asq = alph * alph, ab_ = alph + beta;
  const P pk1 = r.f(2), pk2 = r.f(3), qk = r.f(4), mm = r.f(5), fpk = r.f(6), pp = r.f(8);
  const int64_t sk = g.sk(), c0 = g.idx(i, j, 0), c = c0 + k * sk;
  const float* S1 = Mb + a.p[10] + c0;  // dbar
  const float* S2 = Mb + a.p[11] + c0;  // new w
  // coefficient adjoints of row kk (0 outside rows 1..nz-1)
  auto row = [&](int kk, float& ab, float& bb, float& cb, float& db) {
    ab = bb = cb = db = 0.0f;
    if (kk < 1 || kk > nz - 1) return;
    const float y = S1[kk * sk];
    db = y;
    bb = -y * S2[kk * sk];
    if (kk > 1) ab = -y * S2[(kk - 1) * sk];
    if (kk < nz - 1) cb = -y * S2[(kk + 1) * sk];
  };
  float a0, b0, c0b, d0, a1, b1, c1b, d1;
  row(k, a0, b0, c0b, d0);
  row(k + 1, a1, b1, c1b, d1);
  if (a0 == 0.0f && b0 == 0.0f && c0b == 0.0f && d0 == 0.0f && a1 == 0.0f && b1 == 0.0f && c1b == 0.0f && d1 == 0.0f)
    return;
  float pk1b = 0.0f, pk2b = 0.0f, qkb = 0.0f, ppb = 0.0f, fpkb = 0.0f, mmb = 0.0f, wtb = 0.0f, wb = 0.0f;
  if (k >= 1) {  // row k: this level is c
    const int64_t cm = c - sk;
    const float m = mm[c];
    const float am = asq * m;
    pk1b += b0 * am;
    qkb += am * (b0 - c0b);
    pk2b += c0b * am;
    wtb += d0 * dts;
    ppb += -d0 * m * ab_;
    fpkb += -d0 * alph * m;
    wb += d0;
    const float dpf = (pp[c] + fpk[c]) - (pp[cm] + fpk[cm]);
    mmb += a0 * (-asq) * (pk1[cm] + qk[cm]) + c0b * asq * (pk2[c] - qk[c]) +
           b0 * (-asq) * ((-(pk1[c] + qk[c])) + (pk2[cm] - qk[cm])) + d0 * (-alph * dpf - beta * (pp[c] - pp[cm]));
  }
  if (k + 1 <= nz - 1) {  // row k+1: this level is its cm
    const float m = mm[c + sk];
    const float am = asq * m;
    pk1b += -a1 * am;
    qkb += -am * (a1 - b1);
    pk2b += -b1 * am;
    ppb += d1 * m * ab_;
    fpkb += d1 * alph * m;
  }
  float* B = Mb;
  if (pk1b != 0.0f) B[a.p[2] + c] += pk1b;
  if (pk2b != 0.0f) B[a.p[3] + c] += pk2b;
  if (qkb != 0.0f) B[a.p[4] + c] += qkb;
  if (mmb != 0.0f) B[a.p[5] + c] += mmb;
  if (fpkb != 0.0f) B[a.p[6] + c] += fpkb;
  if (wtb != 0.0f) B[a.p[7] + c] += wtb;
  if (ppb != 0.0f) B[a.p[8] + c] += ppb;
  if (wb != 0.0f) B[a.p[9] + c] += wb;
}

// Clears the adjoint words of cA and cB (the forward's dead scratch).
TE_HD void w_solve_clear(float* Mb, const AdjArg& a, const ADims& g, int i, int j) {
  if (i >= g.nx || j >= g.ny) return;
  const int64_t sk = g.sk(), c0 = g.idx(i, j, 0);
  for (int k = 0; k <= g.nz; ++k) {
    Mb[a.p[10] + c0 + k * sk] = 0.0f;
    Mb[a.p[11] + c0 + k * sk] = 0.0f;
  }
}

// ---- gather-form adjoints of linear stencils.  One thread per word q of the
// field grid (all (nx+7)(ny+7)(nz+1) words, and lead(g) words before the
// field's first one that a window reaches by address) adds to each input field's adjoint
// at q the contributions of every cell whose stencil reads q: no atomics,
// single pass, binary32 (the local Jacobian entries are the forward's own
// values).  Outputs the dispatch overwrites are first copied to a scratch
// plane and cleared (prep, over the forward grid), since neighbouring threads
// read them.  Periodic boundaries only (programs with open boundaries have no
// adjoint).

// adv_assemble.wgsl:139-279: tend = (s dv - divF) / rho (+ tend when accumulating)
struct G_adv_assemble {
  static TE_HD bool cell(const AdjArg& a, const ADims& g, int st, int i, int j, int k) {
    const int iMax = st == 2 ? g.nx : g.nx - 1, jMax = st == 3 ? g.ny : g.ny - 1, kLo = st == 4 ? 1 : 0;
    return i >= 0 && j >= 0 && i <= iMax && j <= jMax && k >= kLo && k <= g.nz - 1 && i < (int)a.n[0] &&
           j < (int)a.n[1] && k < (int)a.n[2];
  }
  static TE_HD void prep(const float* M, float* Mb, const AdjArg& a, float* scr, int i, int j, int k) {
    Acc r(M, Mb, &a);
    const ADims g = r.dims();
    const int stag = r.I(1), st = stag < 0 ? -stag : stag;
    if (!cell(a, g, st, i, j, k)) return;
    const int64_t c = g.idx(i, j, k), w = (int64_t)a.p[11] + c;
    scr[c] = Mb[w];
    if (r.I(12) == 0) Mb[w] = 0.0f;
  }
  static TE_HD void gather(const float* M, float* Mb, const AdjArg& a, const float* scr, int64_t qw) {
    Acc r(M, Mb, &a);
    const ADims g = r.dims();
    int x, y, k;  // the word's coordinates, offset by the halo
    {
      int i_, j_;
      g.coords(qw, i_, j_, k);
      x = i_ + 3;
      y = j_ + 3;
    }
    const int i = x - 3, j = y - 3;
    const int stag = r.I(1), st = stag < 0 ? -stag : stag;
    const P s = r.f(2), divx = r.f(6), rho0 = r.f(7), rf0 = r.f(8), mh = r.f(9), mf = r.f(10),
                c1 = r.f(15), c2 = r.f(16), hm = r.f(30);
    const int64_t sj = g.sJ, sk = g.sk(), q = g.idx(i, j, k);
    auto ok = [&](int ii, int jj, int kk) { return cell(a, g, st, ii, jj, kk); };
    auto rho = [&](int kk) { return st == 4 ? rf0[kk] : rho0[kk]; };
    // -tbar / rho times the flux difference's factor, per direction (0 off the cells)
    auto gx = [&](int ii) { return ok(ii, j, k) ? ((-scr[g.idx(ii, j, k)] / rho(k)) * g.rdx) * (st == 2 ? hm_uf(hm, g, ii) : hm_uh(hm, g, ii)) : 0.0f; };
    auto gy = [&](int jj) { return ok(i, jj, k) ? ((-scr[g.idx(i, jj, k)] / rho(k)) * g.rdy) * (st == 3 ? hm_vf(hm, g, jj) : hm_vh(hm, g, jj)) : 0.0f; };
    auto gz = [&](int kk) {
      if (!ok(i, j, kk)) return 0.0f;
      const float mz = stag < 0 ? 1.0f : (stag == 4 ? mf[kk] : mh[kk]);
      return ((-scr[g.idx(i, j, kk)] / rho(kk)) * g.rdz) * mz;
    };
    const float fxb = st == 2 ? gx(i) - gx(i + 1) : gx(i - 1) - gx(i);
    const float fyb = st == 3 ? gy(j) - gy(j + 1) : gy(j - 1) - gy(j);
    const float fzb = st == 4 ? gz(k) - gz(k + 1) : gz(k - 1) - gz(k);
    // tbar s / rho at a cell: the adjoint of its dv
    auto dvb = [&](int ii, int jj, int kk) { return ok(ii, jj, kk) ? (scr[g.idx(ii, jj, kk)] * s[g.idx(ii, jj, kk)]) / rho(kk) : 0.0f; };
    float divb;
    if (st == 2) divb = 0.5f * dvb(i, j, k) + 0.5f * dvb(i + 1, j, k);
    else if (st == 3) divb = 0.5f * dvb(i, j, k) + 0.5f * dvb(i, j + 1, k);
    else if (st == 4) divb = (ok(i, j, k) ? c2[k] * dvb(i, j, k) : 0.0f) + (ok(i, j, k + 1) ? c1[k + 1] * dvb(i, j, k + 1) : 0.0f);
    else divb = dvb(i, j, k);
    float sb = 0.0f;
    if (ok(i, j, k)) {
      float dv;
      if (st == 2) dv = 0.5f * (divx[q] + divx[q - 1]);
      else if (st == 3) dv = 0.5f * (divx[q] + divx[q - sj]);
      else if (st == 4) dv = fmaf_(c2[k], divx[q], c1[k] * divx[q - sk]);
      else dv = divx[q];
      sb = (scr[q] * dv) / rho(k);
    }
    if (fxb != 0.0f) Mb[a.p[3] + q] += fxb;
    if (fyb != 0.0f) Mb[a.p[4] + q] += fyb;
    if (fzb != 0.0f) Mb[a.p[5] + q] += fzb;
    if (divb != 0.0f) Mb[a.p[6] + q] += divb;
    if (sb != 0.0f) Mb[a.p[2] + q] += sb;
  }
  static constexpr bool kPrep = true;
  static TE_HD int64_t lead(const ADims&) { return 0; }
};

// acoustic_uv.wgsl:144-181: u += dts (uten - P (thv + thv[-1])), P = tem1 (ppd - ppd[-1]) uf
// (likewise v); u and v are updated in place, so their adjoints pass through.
struct G_acoustic_uv {
  static TE_HD void prep(const float*, float*, const AdjArg&, float*, int, int, int) {}
  static TE_HD void gather(const float* M, float* Mb, const AdjArg& a, const float*, int64_t qw) {
    Acc r(M, Mb, &a);
    const ADims g = r.dims();
    int x, y, k;  // the word's coordinates, offset by the halo
    {
      int i_, j_;
      g.coords(qw, i_, j_, k);
      x = i_ + 3;
      y = j_ + 3;
    }
    const int i = x - 3, j = y - 3;
    if (k >= g.nz) return;
    const float dts = r.F(1, 0);
    const P ppd = r.f(2), thv = r.f(3), hm = r.f(30);
    const float *ub = Mb + a.p[6], *vb = Mb + a.p[7];
    const int64_t sj = g.sJ, q = g.idx(i, j, k);
    const float tem1 = (g.rdx * CP) * 0.5f, tem2 = (g.rdy * CP) * 0.5f;
    auto ok = [&](int ii, int jj) { return ii >= 0 && jj >= 0 && ii <= g.nx && jj <= g.ny && ii < (int)a.n[0] && jj < (int)a.n[1] && k < (int)a.n[2]; };
    // cell c: -dts ubar (P at c) and -dts ubar (thv + thv[-1]) tem1 uf
    float ppdb = 0.0f, thvb = 0.0f, utb = 0.0f, vtb = 0.0f;
    for (int d = 0; d < 2; ++d) {  // u of cells q (d = 0) and q + 1 (d = 1)
      const int ii = i + d;
      if (!ok(ii, j)) continue;
      const int64_t c = q + d;
      const float u_b = ub[c];
      if (u_b == 0.0f) continue;
      const float P = (tem1 * (ppd[c] - ppd[c - 1])) * hm_uf(hm, g, ii);
      const float T = thv[c] + thv[c - 1];
      const float gP = -dts * u_b;
      ppdb += (d == 0 ? 1.0f : -1.0f) * ((gP * T) * tem1) * hm_uf(hm, g, ii);
      thvb += gP * P;
      if (d == 0) utb = dts * u_b;
    }
    for (int d = 0; d < 2; ++d) {  // v of cells q (d = 0) and q + sj (d = 1)
      const int jj = j + d;
      if (!ok(i, jj)) continue;
      const int64_t c = q + d * sj;
      const float v_b = vb[c];
      if (v_b == 0.0f) continue;
      const float P = (tem2 * (ppd[c] - ppd[c - sj])) * hm_vf(hm, g, jj);
      const float T = thv[c] + thv[c - sj];
      const float gP = -dts * v_b;
      ppdb += (d == 0 ? 1.0f : -1.0f) * ((gP * T) * tem2) * hm_vf(hm, g, jj);
      thvb += gP * P;
      if (d == 0) vtb = dts * v_b;
    }
    if (ppdb != 0.0f) Mb[a.p[2] + q] += ppdb;
    if (thvb != 0.0f) Mb[a.p[3] + q] += thvb;
    if (utb != 0.0f) Mb[a.p[4] + q] += utb;
    if (vtb != 0.0f) Mb[a.p[5] + q] += vtb;
  }
  static constexpr bool kPrep = false;
  static TE_HD int64_t lead(const ADims&) { return 0; }
};

// acoustic_fpk.wgsl:167-207: fpk = beta A + (dts B + advPi) with A = w pk1+qk-part,
// B = ppten - ppterm div, advPi from u, v and piadv differences.
struct G_acoustic_fpk {
  static TE_HD bool cell(const AdjArg& a, const ADims& g, int i, int j, int k) {
    return i >= 0 && j >= 0 && k >= 0 && i < g.nx && j < g.ny && k < g.nz && i < (int)a.n[0] && j < (int)a.n[1] &&
           k < (int)a.n[2];
  }
  static TE_HD void prep(const float* M, float* Mb, const AdjArg& a, float* scr, int i, int j, int k) {
    Acc r(M, Mb, &a);
    const ADims g = r.dims();
    if (!cell(a, g, i, j, k)) return;
    const int64_t c = g.idx(i, j, k), w = (int64_t)a.p[11] + c;
    scr[c] = Mb[w];
    Mb[w] = 0.0f;
  }
  static TE_HD void gather(const float* M, float* Mb, const AdjArg& a, const float* scr, int64_t qw) {
    Acc r(M, Mb, &a);
    const ADims g = r.dims();
    int x, y, k;  // the word's coordinates, offset by the halo
    {
      int i_, j_;
      g.coords(qw, i_, j_, k);
      x = i_ + 3;
      y = j_ + 3;
    }
    const int i = x - 3, j = y - 3;
    const float dts = r.F(1, 0), beta = r.F(1, 2), smeps = r.F(1, 4);
    const P u = r.f(2), v = r.f(3), w = r.f(4), pi = r.f(5), ppterm = r.f(7), pk1 = r.f(8), pk2 = r.f(9),
                qk = r.f(10), hm = r.f(30);
    const int64_t sj = g.sJ, sk = g.sk(), q = g.idx(i, j, k);
    const float temx = (dts * 0.5f) * g.rdx, temy = (dts * 0.5f) * g.rdy;
    auto fb = [&](int ii, int jj, int kk) { return cell(a, g, ii, jj, kk) ? scr[g.idx(ii, jj, kk)] : 0.0f; };
    // adjoint of div at a cell (0 where the forward zeroed div)
    auto divb = [&](int ii, int jj, int kk, int64_t c) {
      const float f = fb(ii, jj, kk);
      if (f == 0.0f) return 0.0f;
      const float div = fmaf_((u[c + 1] - u[c]) * g.rdx, hm_uh(hm, g, ii), ((v[c + sj] - v[c]) * g.rdy) * hm_vh(hm, g, jj));
      if (ab(D1(div)).v < smeps) return 0.0f;
      return -(dts * f) * ppterm[c];
    };
    auto aub = [&](int ii, int jj, int kk) { return -(fb(ii, jj, kk) * temx) * hm_uh(hm, g, ii); };
    auto avb = [&](int ii, int jj, int kk) { return -(fb(ii, jj, kk) * temy) * hm_vh(hm, g, jj); };
    float ubs = 0.0f, vbs = 0.0f, wbs = 0.0f, pib = 0.0f, pk1b = 0.0f, pk2b = 0.0f, qkb = 0.0f, pptenb = 0.0f,
          pptermb = 0.0f;
    // cell q: u0, v0, w0, pc, pk1, pk2, qk, ppten, ppterm
    if (cell(a, g, i, j, k)) {
      const float f = scr[q];
      if (f != 0.0f) {
        const float A = beta * f, B = dts * f;
        const float w0 = w[q], w1 = w[q + sk];
        wbs += A * (pk1[q] + qk[q]);
        pk1b += A * w0;
        pk2b += A * w1;
        qkb += A * (w0 - w1);
        pptenb += B;
        const float div = fmaf_((u[q + 1] - u[q]) * g.rdx, hm_uh(hm, g, i), ((v[q + sj] - v[q]) * g.rdy) * hm_vh(hm, g, j));
        const bool zeroed = ab(D1(div)).v < smeps;
        if (!zeroed) pptermb += -B * div;
        const float db = zeroed ? 0.0f : -B * ppterm[q];
        ubs += -(db * g.rdx) * hm_uh(hm, g, i);
        vbs += -(db * g.rdy) * hm_vh(hm, g, j);
        const float Au = aub(i, j, k), Av = avb(i, j, k);
        ubs += Au * (pi[q] - pi[q - 1]);
        vbs += Av * (pi[q] - pi[q - sj]);
        pib += Au * (u[q] - u[q + 1]) + Av * (v[q] - v[q + sj]);
      }
    }
    // cell q - 1: u1 = u[q], pip = pi[q]
    if (fb(i - 1, j, k) != 0.0f) {
      const int64_t c = q - 1;
      const float db = divb(i - 1, j, k, c);
      ubs += (db * g.rdx) * hm_uh(hm, g, i - 1);
      const float Au = aub(i - 1, j, k);
      ubs += Au * (pi[c + 1] - pi[c]);
      pib += Au * u[c + 1];
    }
    // cell q + 1: pim = pi[q]
    if (fb(i + 1, j, k) != 0.0f) pib -= aub(i + 1, j, k) * u[q + 1];
    // cell q - sj: v1 = v[q], pjp = pi[q]
    if (fb(i, j - 1, k) != 0.0f) {
      const int64_t c = q - sj;
      const float db = divb(i, j - 1, k, c);
      vbs += (db * g.rdy) * hm_vh(hm, g, j - 1);
      const float Av = avb(i, j - 1, k);
      vbs += Av * (pi[c + sj] - pi[c]);
      pib += Av * v[c + sj];
    }
    // cell q + sj: pjm = pi[q]
    if (fb(i, j + 1, k) != 0.0f) pib -= avb(i, j + 1, k) * v[q + sj];
    // cell q - sk: w1 = w[q]
    if (fb(i, j, k - 1) != 0.0f) {
      const int64_t c = q - sk;
      wbs += (beta * scr[c]) * (pk2[c] - qk[c]);
    }
    if (ubs != 0.0f) Mb[a.p[2] + q] += ubs;
    if (vbs != 0.0f) Mb[a.p[3] + q] += vbs;
    if (wbs != 0.0f) Mb[a.p[4] + q] += wbs;
    if (pib != 0.0f) Mb[a.p[5] + q] += pib;
    if (pptenb != 0.0f) Mb[a.p[6] + q] += pptenb;
    if (pptermb != 0.0f) Mb[a.p[7] + q] += pptermb;
    if (pk1b != 0.0f) Mb[a.p[8] + q] += pk1b;
    if (pk2b != 0.0f) Mb[a.p[9] + q] += pk2b;
    if (qkb != 0.0f) Mb[a.p[10] + q] += qkb;
  }
  static constexpr bool kPrep = true;
  static TE_HD int64_t lead(const ADims&) { return 0; }
};

// acoustic_pp.wgsl:131-167: p = alph (w pk1+qk-part) + (prev + fpk), zeroed
// below smeps; pp3d = p (in place), ppd = p + kdiv (p - prev), ppx = kdiv (p - prev).
// prep: everything local to the cell, and alph pbar into the scratch plane;
// gather: w[k] as the w[k+1] of the cell below.
struct G_acoustic_pp {
  static TE_HD void prep(const float* M, float* Mb, const AdjArg& a, float* scr, int i, int j, int k) {
    Acc r(M, Mb, &a);
    const ADims g = r.dims();
    if (i >= g.nx || j >= g.ny || k >= g.nz) return;
    const float alph = r.F(1, 1), kdiv = r.F(1, 3), smeps = r.F(1, 4);
    const P w = r.f(2), pk1 = r.f(3), pk2 = r.f(4), qk = r.f(5), fpk = r.f(6), pp = r.f(7);
    const int saveX = r.I(10);
    const int64_t sk = g.sk(), c = g.idx(i, j, k);
    float* ppb = Mb + a.p[7] + c;
    float* ppdb = Mb + a.p[8] + c;
    const float xb = saveX != 0 ? Mb[a.p[9] + c] : 0.0f;
    const float pbar_out = *ppb + (1.0f + kdiv) * *ppdb + kdiv * xb;
    const float w1 = w[c], w2 = w[c + sk], q = qk[c];
    const float prev = pp[c];
    const float pval = fmaf_(alph, fmaf_(w1, pk1[c] + q, w2 * (pk2[c] - q)), prev + fpk[c]);
    const bool zeroed = ab(D1(pval)).v < smeps;
    const float pb = zeroed ? 0.0f : pbar_out;
    *ppb = pb - kdiv * (*ppdb + xb);  // adjoint of the old pp3d
    *ppdb = 0.0f;
    if (saveX != 0) Mb[a.p[9] + c] = 0.0f;
    const float A = alph * pb;
    scr[c] = A;
    if (pb == 0.0f) return;
    Mb[a.p[6] + c] += pb;
    Mb[a.p[3] + c] += A * w1;
    Mb[a.p[4] + c] += A * w2;
    Mb[a.p[5] + c] += A * (w1 - w2);
    Mb[a.p[2] + c] += A * (pk1[c] + q);
  }
  static TE_HD void gather(const float* M, float* Mb, const AdjArg& a, const float* scr, int64_t qw) {
    Acc r(M, Mb, &a);
    const ADims g = r.dims();
    int x, y, k;  // the word's coordinates, offset by the halo
    {
      int i_, j_;
      g.coords(qw, i_, j_, k);
      x = i_ + 3;
      y = j_ + 3;
    }
    const int i = x - 3, j = y - 3;
    if (i < 0 || j < 0 || i >= g.nx || j >= g.ny || k < 1 || k > g.nz) return;
    const P pk2 = r.f(4), qk = r.f(5);
    const int64_t cb = g.idx(i, j, k - 1);  // the cell below
    const float A = scr[cb];
    if (A != 0.0f) Mb[a.p[2] + cb + g.sk()] += A * (pk2[cb] - qk[cb]);
  }
  static constexpr bool kPrep = true;
  static TE_HD int64_t lead(const ADims&) { return 0; }
};

// Horizontal flx5 fluxes, scalar (adv_flux_x.wgsl:120-143, adv_flux_y.wgsl:120-144)
// and staggered momentum (adv_flux_x_stag.wgsl:144-186, _y_): fl = m f, f the
// upwind (2, -13, 47, 27, -3)/60 face value of a 5-point window along the
// direction, m the face mass flux.  A word of the advected field collects
// fbar m w_p over the (at most 7) cells whose window holds it at position p;
// a mass-flux word collects fbar f (times the interpolation weight) over the
// cells whose m reads it.
template <int DIR, bool STAG>
struct G_flux_h {
  static constexpr bool kPrep = true;
  // a y window at row j = -1 of level 0 reaches one row before the field
  static TE_HD int64_t lead(const ADims& g) { return (STAG && DIR == 1) ? g.sJ : 0; }
  static TE_HD float wt(int p) {
    return p == 0 ? 2.0f / 6e+01f : p == 1 ? -13.0f / 6e+01f : p == 2 ? 47.0f / 6e+01f : p == 3 ? 27.0f / 6e+01f : -3.0f / 6e+01f;
  }
  // cells the forward writes / whose value depends on its inputs
  static TE_HD bool cell_out(const AdjArg& a, const ADims& g, int i, int j, int k) {
    if (STAG) {
      if (i < -1 || j < -1 || i > g.nx || j > g.ny || k < 0 || k >= g.nz) return false;
      return i + 1 < (int)a.n[0] && j + 1 < (int)a.n[1] && k < (int)a.n[2];
    }
    if (i < 0 || j < 0 || k < 0 || k >= g.nz) return false;
    if (DIR == 0 ? (i > g.nx || j >= g.ny) : (i >= g.nx || j > g.ny)) return false;
    return i < (int)a.n[0] && j < (int)a.n[1] && k < (int)a.n[2];
  }
  static TE_HD bool cell_dep(const AdjArg& a, const ADims& g, int stag, int i, int j, int k) {
    return cell_out(a, g, i, j, k) && !(STAG && stag == 4 && k == 0);
  }
  static TE_HD int pa() { return STAG ? 2 : 1; }                  // advected field
  static TE_HD int prr() { return STAG ? (DIR == 0 ? 3 : 4) : 2; }  // mass flux
  static TE_HD int pout() { return STAG ? 8 : 3; }
  static TE_HD float mass(const float* rr, const float* c2, const ADims& g, int stag, int64_t c, int k) {
    if (!STAG) return rr[c];
    const int64_t sj = g.sJ, sk = g.sk();
    if (stag == 2) return 0.5f * (rr[c] + rr[c + (DIR == 0 ? 1 : -1)]);
    if (stag == 3) return 0.5f * (rr[c] + rr[c + (DIR == 0 ? -sj : sj)]);
    if (stag == 4) return fmaf_(c2[k], rr[c], (1.0f - c2[k]) * rr[c - sk]);
    return rr[c];
  }
  // window offset of position p (0..4) for an upwind / shifted stencil
  static TE_HD int offset(bool up, bool shifted, int p) { return up ? (shifted ? p - 2 : p - 3) : (shifted ? 3 - p : 2 - p); }
  static TE_HD void prep(const float* M, float* Mb, const AdjArg& a, float* scr, int x, int y, int k) {
    Acc r(M, Mb, &a);
    const ADims g = r.dims();
    const int i = STAG ? x - 1 : x, j = STAG ? y - 1 : y;
    if (!cell_out(a, g, i, j, k)) return;
    const int stag = STAG ? r.I(1) : 1;
    const int64_t c = g.idx(i, j, k), w = (int64_t)a.p[pout()] + c;
    scr[c] = cell_dep(a, g, stag, i, j, k) ? Mb[w] : 0.0f;
    Mb[w] = 0.0f;
  }
  static TE_HD void gather(const float* M, float* Mb, const AdjArg& a, const float* scr, int64_t qw) {
    Acc r(M, Mb, &a);
    const ADims g = r.dims();
    int x, y, k;  // the word's coordinates, offset by the halo
    {
      int i_, j_;
      g.coords(qw, i_, j_, k);
      x = i_ + 3;
      y = j_ + 3;
    }
    const int i = x - 3, j = y - 3;
    if (k >= g.nz) return;
    const int stag = STAG ? r.I(1) : 1;
    const bool shifted = STAG && stag == (DIR == 0 ? 2 : 3);
    const P av = r.f(pa()), rr = r.f(prr());
    const float* c2 = STAG ? (const float*)r.f(7) : nullptr;
    const int64_t sj = g.sJ, sk = g.sk(), st = DIR == 0 ? 1 : sj, q = g.idx(i, j, k);
    auto face = [&](int64_t c, bool up) {
      float v[5];
      for (int p = 0; p < 5; ++p) v[p] = av[c + offset(up, shifted, p) * st];
      return fmaf_(-3.0f / 6e+01f, v[4],
                   fmaf_(27.0f / 6e+01f, v[3], fmaf_(47.0f / 6e+01f, v[2], fmaf_(2.0f / 6e+01f, v[0], (-13.0f / 6e+01f) * v[1]))));
    };
    // the advected field.  Candidate cells by linear index, decoded: a window
    // at the grid's first row / column reaches into the previous one by
    // address, exactly as the forward reads it.
    float ab_ = 0.0f;
    for (int o = -3; o <= 3; ++o) {
      const int64_t c = q - o * st;
      int ci = i, cj = j, ck = k;
      g.step(DIR, -o, ci, cj, ck);
      if (!cell_dep(a, g, stag, ci, cj, ck)) continue;
      const float fb = scr[c];
      if (fb == 0.0f) continue;
      const float m = mass(rr, c2, g, stag, c, ck);
      const bool up = m >= 0.0f;
      for (int p = 0; p < 5; ++p)
        if (offset(up, shifted, p) == o) ab_ += (fb * m) * wt(p);
    }
    // the mass flux: the cells whose m reads word q, with their weights
    float rb = 0.0f;
    auto add_m = [&](int ci, int cj, int ck, float lam) {
      if (ck < 0 || !cell_dep(a, g, stag, ci, cj, ck)) return;
      const int64_t c = g.idx(ci, cj, ck);
      const float fb = scr[c];
      if (fb == 0.0f) return;
      rb += lam * (fb * face(c, mass(rr, c2, g, stag, c, ck) >= 0.0f));
    };
    if (!STAG || (stag != 2 && stag != 3 && stag != 4)) {
      add_m(i, j, k, 1.0f);
    } else if (stag == 2) {  // m at c uses rr[c] and rr[c + 1] (x) / rr[c - 1] (y)
      add_m(i, j, k, 0.5f);
      if (DIR == 0) add_m(i - 1, j, k, 0.5f); else add_m(i + 1, j, k, 0.5f);
    } else if (stag == 3) {  // rr[c] and rr[c - sj] (x) / rr[c + sj] (y)
      add_m(i, j, k, 0.5f);
      if (DIR == 0) add_m(i, j + 1, k, 0.5f); else add_m(i, j - 1, k, 0.5f);
    } else {  // stag 4: c2[k] rr[c] + (1 - c2[k]) rr[c - sk]
      add_m(i, j, k, c2[k]);
      if (k + 1 < g.nz) add_m(i, j, k + 1, 1.0f - c2[k + 1]);
    }
    if (ab_ != 0.0f) Mb[a.p[pa()] + q] += ab_;
    if (rb != 0.0f) Mb[a.p[prr()] + q] += rb;
  }
};

// acoustic_accumulate.wgsl:129-158: rr += u (times tavg on the last
// substep, rrw with alph w), all in place and pointwise: per word, no
// neighbours, no prep.
struct G_acoustic_accumulate {
  static constexpr bool kPrep = false;
  static TE_HD int64_t lead(const ADims&) { return 0; }
  static TE_HD void prep(const float*, float*, const AdjArg&, float*, int, int, int) {}
  static TE_HD void gather(const float* M, float* Mb, const AdjArg& a, const float*, int64_t qw) {
    Acc r(M, Mb, &a);
    const ADims g = r.dims();
    int i, j, k;
    g.coords(qw, i, j, k);
    if (i < -3 || j < -3 || k < 0 || i > g.nx + 3 || j > g.ny + 3 || k > g.nz) return;
    if (i + 3 >= (int)a.n[0] || j + 3 >= (int)a.n[1] || k >= (int)a.n[2]) return;
    const float alph = r.F(1, 1), tavg = r.F(8);
    const int64_t c = qw;
    float* rub = Mb + a.p[5] + c;
    float* rvb = Mb + a.p[6] + c;
    float* rwb = Mb + a.p[7] + c;
    if (tavg > 0.0f) {
      const float u = *rub * tavg, v = *rvb * tavg, w = *rwb * tavg;
      *rub = u; *rvb = v; *rwb = w;
      if (u != 0.0f) Mb[a.p[2] + c] += u;
      if (v != 0.0f) Mb[a.p[3] + c] += v;
      if (w != 0.0f) Mb[a.p[4] + c] += w * alph;
    } else {
      if (*rub != 0.0f) Mb[a.p[2] + c] += *rub;
      if (*rvb != 0.0f) Mb[a.p[3] + c] += *rvb;
      if (*rwb != 0.0f) Mb[a.p[4] + c] += *rwb;
    }
  }
};

}  // namespace adj
}  // namespace te

// Kernels with a per-point adjoint body: X(name, body, count of written
// parameters, {written parameters}).
#define TE_ADJ_POINT_KERNELS(X) \
  X("update_diagnostics_moist", B_update_diagnostics_moist, 5, {15, 16, 17, 18, 19}) \
  X("update_diagnostics", B_update_diagnostics, 4, {5, 6, 7, 9}) \
  X("stage_tendency_reset", B_stage_tendency_reset, 3, {1, 2, 3}) \
  X("stage_state_reset", B_stage_state_reset, 5, {5, 6, 7, 8, 9}) \
  X("massflux_from_state", B_massflux_from_state, 3, {6, 7, 8}) \
  X("velocity_to_massflux", B_velocity_to_massflux, 3, {3, 4, 5}) \
  X("div_massflux", B_div_massflux, 1, {5}) \
  X("divergence3d", B_divergence3d, 1, {5}) \
  X("buoyancy_apply", B_buoyancy_apply, 1, {4}) \
  X("rk_update", B_rk_update, 1, {4}) \
  X("build_sadv", B_build_sadv, 1, {3}) \
  X("apply_thterm", B_apply_thterm, 1, {3}) \
  X("turb_add_tendency", B_turb_add_tendency, 1, {2}) \
  X("turb_zero", B_turb_zero, 1, {1}) \
  X("ppx_finish", B_ppx_finish, 1, {2}) \
  X("diabatic_unapply", B_diabatic_unapply, 8, {2, 3, 4, 5, 6, 7, 8, 9}) \
  X("diabatic_rates", B_diabatic_rates, 4, {6, 7, 8, 9}) \
  X("pressure_tendency", B_pressure_tendency, 1, {11}) \
  X("update_rho_prs", B_update_rho_prs, 2, {7, 8}) \
  X("adv_flux_x", B_adv_flux_h<0>, 1, {3}) \
  X("adv_flux_y", B_adv_flux_h<1>, 1, {3}) \
  X("adv_flux_z", B_adv_flux_z, 1, {5}) \
  X("adv_flux_x_weno", B_adv_flux_h_weno<0>, 1, {5}) \
  X("adv_flux_y_weno", B_adv_flux_h_weno<1>, 1, {5}) \
  X("adv_flux_z_weno", B_adv_flux_z_weno, 1, {8}) \
  X("adv_flux_x_stag", B_adv_flux_h_stag<0>, 1, {8}) \
  X("adv_flux_y_stag", B_adv_flux_h_stag<1>, 1, {8}) \
  X("adv_flux_z_stag", B_adv_flux_z_stag, 1, {8}) \
  X("adv_flux_x_weno_w", B_adv_flux_h_weno_w<0>, 1, {5}) \
  X("adv_flux_y_weno_w", B_adv_flux_h_weno_w<1>, 1, {5}) \
  X("adv_flux_z_weno_w", B_adv_flux_z_weno_w, 1, {4}) \
  X("adv_flux_x_weno_uv", B_adv_flux_h_weno_uv<0>, 1, {5}) \
  X("adv_flux_y_weno_uv", B_adv_flux_h_weno_uv<1>, 1, {5}) \
  X("adv_flux_z_weno_uv", B_adv_flux_z_weno_uv, 1, {7}) \
  X("adv_assemble", B_adv_assemble, 1, {11}) \
  X("acoustic_prepare", B_acoustic_prepare, 4, {11, 12, 13, 14}) \
  X("acoustic_begin", B_acoustic_begin, 3, {3, 4, 5}) \
  X("acoustic_uv", B_acoustic_uv, 2, {6, 7}) \
  X("acoustic_fpk", B_acoustic_fpk, 1, {11}) \
  X("acoustic_pp", B_acoustic_pp, 3, {7, 8, 9}) \
  X("acoustic_accumulate", B_acoustic_accumulate, 3, {5, 6, 7}) \
  X("pdef_dum", B_pdef_dum, 1, {10}) \
  X("pdef_apply", B_pdef_apply, 1, {16}) \
  X("bc_periodic_x", B_bc_periodic<0>, 1, {1}) \
  X("bc_periodic_y", B_bc_periodic<1>, 1, {1}) \
  X("_copy", B_copy, 1, {1}) \
  X("rayleigh_damping", B_rayleigh_damping, 4, {10, 11, 12, 13})

// Kernels whose adjoint runs in gather form (replacing the per-point body's
// scatter form): X(name, gather struct); COMMA separates template arguments.
#define COMMA ,
#define TE_ADJ_GATHER_KERNELS(X) \
  X("adv_assemble", G_adv_assemble) \
  X("acoustic_uv", G_acoustic_uv) \
  X("acoustic_fpk", G_acoustic_fpk) \
  X("acoustic_pp", G_acoustic_pp) \
  X("adv_flux_x", G_flux_h<0 COMMA false>) \
  X("adv_flux_y", G_flux_h<1 COMMA false>) \
  X("adv_flux_x_stag", G_flux_h<0 COMMA true>) \
  X("adv_flux_y_stag", G_flux_h<1 COMMA true>) \
  X("acoustic_accumulate", G_acoustic_accumulate)
