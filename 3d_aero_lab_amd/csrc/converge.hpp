// ============================================================================
// converge.hpp — convergence rules of a monitored solve, one update per
// readback; host and device code (the served form decides on the device, the
// tests run the same code on the host).
//
//   AERO3D_RULE_RES_DROP   done when res / res0 <= res_drop (res0: the first
//                          positive residual), or at `cap` iterations
//   AERO3D_RULE_REFERENCE  the reference controller's rule (CFD3.readback,
//                          src/air_cfd.js:555, 573-576): more than six
//                          readbacks, res / res0 < 2e-3 and the lift
//                          coefficients of the last four readbacks within
//                          2e-3 of each other; or more than `cap` iterations
//                          (the reference caps at 30000)
// Both rules also end a solve whose residual is not finite after its first
// readback (diverged); the reference keeps iterating such a solve until the
// cap.
// ============================================================================
#pragma once
#include <math.h>

#ifdef __HIPCC__
#define AERO3D_HD __host__ __device__
#else
#define AERO3D_HD
#endif

namespace aero3d {

enum { kRuleResDrop = 0, kRuleReference = 1 };

struct ConvState {
  double res0;      // first positive residual (NaN before)
  double cl[4];     // lift coefficients of the last four readbacks (ring)
  int nreads;       // readbacks so far, the first (snapshot-only) one included
  int conv, diverged, done;
};

AERO3D_HD inline void conv_init(ConvState &s) {
  s.res0 = NAN;
  for (int k = 0; k < 4; k++) s.cl[k] = NAN;
  s.nreads = 0; s.conv = 0; s.diverged = 0; s.done = 0;
}

// One readback: residual res (NaN on the first readback), lift coefficient CL
// and the iteration count it of the solve at the readback.
AERO3D_HD inline void conv_update(ConvState &s, int rule, double res_drop, int cap, double res, double CL, int it) {
  s.nreads++;
  if (!(s.res0 > 0) && res > 0) s.res0 = res;          // src/air_cfd.js:554
  s.cl[(s.nreads - 1) & 3] = CL;
  if (rule == kRuleReference) {
    bool ok = s.nreads > 6 && s.res0 > 0 && res / s.res0 < 2e-3;
    if (ok) {                                           // spread of the last four CL; a NaN fails
      double lo = s.cl[0], hi = s.cl[0];
      bool nan = false;
      for (int k = 0; k < 4; k++) {
        const double c = s.cl[k];
        if (c != c) nan = true;
        lo = c < lo ? c : lo; hi = c > hi ? c : hi;
      }
      ok = !nan && hi - lo < 2e-3;
    }
    s.conv = ok ? 1 : 0;
    s.diverged = (s.nreads > 1 && !isfinite(res)) ? 1 : 0;
    s.done = (s.conv || s.diverged || it > cap) ? 1 : 0;
  } else {
    s.conv = (s.res0 > 0 && res / s.res0 <= res_drop) ? 1 : 0;
    s.diverged = (s.nreads > 1 && !isfinite(res)) ? 1 : 0;
    s.done = (s.conv || s.diverged || it >= cap) ? 1 : 0;
  }
}

}  // namespace aero3d
