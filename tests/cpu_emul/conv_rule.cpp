// Host build of the served form's convergence rules (csrc/converge.hpp), the
// same code the device runs, for the CPU tests: a readback sequence in, the
// rule's flags after every readback out.
#include <cstdint>

#include "converge.hpp"

extern "C" int conv_run(int rule, double res_drop, int cap, int n, const double *res, const double *cl, const int *it, int *conv,
                        int *diverged, int *done, double *res0) {
  aero3d::ConvState s;
  aero3d::conv_init(s);
  for (int k = 0; k < n; k++) {
    aero3d::conv_update(s, rule, res_drop, cap, res[k], cl[k], it[k]);
    conv[k] = s.conv; diverged[k] = s.diverged; done[k] = s.done; res0[k] = s.res0;
  }
  return s.nreads;
}
