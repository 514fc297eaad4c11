// Variable large step (adapt_dt = 1): CM1's getnewdt and dtsmall
// (misclibs.F:543-847) with the output clocks of cm1.F:2721-3878, as
// web/core/adaptive.js:15-109 states them.  The arithmetic follows the
// reference's mix of binary32 roundings (Fortran `real`) and double
// (`double precision` and JavaScript numbers) operation for operation, since
// the chosen dt becomes the step's parameter bytes.
#include <cmath>
#include <cstdint>
#include <new>
#include <vector>

#include "../../include/te.h"

namespace {

inline float f(double x) { return (float)x; }
const int kMaxNsound = 192;  // AcousticSchedule.maxNsound (config.js:238)
enum Kind { Tape = 0, Restart = 1, Stat = 2, DomainDiag = 3, AzimAvg = 4, Hifrq = 5 };

struct OutClock {
  int kind;
  float freq;
  double time;
  bool fired;
};

}  // namespace

struct te_clock {
  double dtl;  // binary32 value of the namelist dtl
  bool ks_check = true;
  float dtsm = 0, ks_limit = 0;
  double dbldt = 0, mtime = 0;
  float dt = 0, dtlast = 0, cflmax = 0, ksmax = 0;
  int nsound = 0, steps = 0;
  std::vector<OutClock> clocks;

  // AcousticSchedule.smallStep (config.js:239-243)
  static float small_step(double dx, double dz, bool two_d) {
    if (!(dz > 0)) return f((double)f((double)f(0.60) * (double)f(dx)) / (double)f(360.0));
    const float m = f(std::fmin(dx, dz));
    return f((double)f((double)(two_d ? f(0.60) : f(0.50)) * (double)m) / (double)f(360.0));
  }

  void get_new_dt(float c_in, float k_in, bool doinit) {
    const float cfl_limit = f(1.00), max_change = f(0.10);
    dtlast = dt;
    float cfl = f(std::fmax((double)c_in, (double)f(1.0e-10)));
    float ks = f((double)k_in);
    if (cfl > cfl_limit) dbldt = dbldt * (double)f((double)cfl_limit / (double)cfl);
    else dbldt = dbldt * std::fmin((double)f((double)f(1.0) + (double)max_change), (double)f((double)cfl_limit / cfl));
    if (ks_check && ((double)ks * dbldt / (double)dt) > (double)ks_limit)
      dbldt = std::fmin(dbldt, (double)f((double)f((double)dt * (double)ks_limit) /
                                         std::fmax((double)f(1.0e-10), (double)ks)));
    dbldt = std::fmin(dbldt, (double)f((double)f(2.0) * dtl));
    if (!doinit) {
      for (const OutClock& c : clocks) {
        if (!(c.time > 0)) continue;
        const double tout = c.fired ? ((c.time - mtime) + (double)c.freq) : (c.time - mtime);
        if (tout > 2.0 * dbldt && tout <= 3.0 * dbldt) dbldt = tout / 3.0;
        else if (tout > dbldt && tout <= 2.0 * dbldt) dbldt = tout / 2.0;
        else if (tout <= dbldt) dbldt = tout;
      }
    }
    dt = f(dbldt);
    dt_small();
    const float tem = f((double)dt / (double)dtlast);
    cfl = f((double)cfl * (double)tem);
    ks = f((double)ks * (double)tem);
    cflmax = cfl;
    ksmax = ks;
  }

  void dt_small() {
    const double x = dbldt / (double)dtsm, ax = std::fabs(x);
    double r = std::floor(ax);  // Math.round(|x|): nearest, halves up
    if (ax - r >= 0.5) r += 1.0;
    int n = (int)std::fmax(std::trunc(x < 0 ? -r : r), 4.0);
    if (n % 2 != 0) n += 1;
    if (dbldt / (double)f((double)n) > (double)dtsm) n += 2;
    if (n > kMaxNsound) {
      n = kMaxNsound;
      dbldt = (double)f((double)f((double)n) * (double)dtsm);
      dt = f(dbldt);
    }
    nsound = n;
  }

  void advance_clocks() {
    for (OutClock& c : clocks) {
      if (!(c.fired && c.freq > 0)) continue;
      do {
        c.time = c.time + (double)c.freq;
      } while (!(c.time > mtime));
    }
  }
};

extern "C" {

te_clock* te_clock_create(double dtl, double min_dx, double min_dz, int32_t two_dimensional, int32_t ks_check,
                          const float freq[6]) {
  te_clock* c = new (std::nothrow) te_clock();
  if (!c) return nullptr;
  c->dtl = (double)f(dtl);
  c->ks_check = ks_check != 0;
  c->dtsm = te_clock::small_step(min_dx, min_dz, two_dimensional != 0);
  c->ks_limit = two_dimensional ? f(0.18) : f((double)f(0.18) / (double)f(2.0));
  c->dbldt = c->dtl;
  c->dt = f(c->dtl);
  c->dtlast = c->dt;
  for (int k = 0; freq && k < 6; ++k) {
    if (std::isnan(freq[k])) continue;
    const float fr = freq[k];
    c->clocks.push_back(OutClock{k, fr, k == Restart ? (double)fr : 0.0, false});
  }
  return c;
}

void te_clock_start(te_clock* c, float cflmax, float ksmax) {
  if (!c) return;
  c->get_new_dt(cflmax, ksmax, true);
  for (OutClock& o : c->clocks) o.fired = o.kind != Restart;
  c->advance_clocks();
}

void te_clock_begin_step(te_clock* c) {
  if (!c) return;
  const float tenth = f((double)f(0.1) * (double)c->dt);
  for (OutClock& o : c->clocks) {
    const bool due = (c->mtime + c->dbldt) >= (o.time - (double)tenth);
    if (o.kind == Tape || o.kind == Stat) o.fired = due || o.freq < 0;
    else if (o.kind == Restart) o.fired = due && o.freq > 0;
    else o.fired = due || o.freq <= 0;
  }
  c->mtime = c->mtime + c->dbldt;
  c->steps += 1;
}

void te_clock_end_step(te_clock* c, float cflmax, float ksmax) {
  if (!c) return;
  c->get_new_dt(cflmax, ksmax, false);
  c->advance_clocks();
}

void te_clock_get(const te_clock* c, te_clock_state* o) {
  if (!c || !o) return;
  o->mtime = c->mtime;
  o->dbldt = c->dbldt;
  o->dt = c->dt;
  o->dtlast = c->dtlast;
  o->cflmax = c->cflmax;
  o->ksmax = c->ksmax;
  o->nsound = c->nsound;
  o->steps = c->steps;
  o->fired = 0;
  for (const OutClock& k : c->clocks)
    if (k.fired) o->fired |= 1u << k.kind;
}

void te_clock_destroy(te_clock* c) { delete c; }

}  // extern "C"
