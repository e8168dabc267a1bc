// Host build of the device hysteresis control's rule (sweep_ctl_update of
// cupymag_amd/csrc/sweep_ctl.h, compiled with hipcc --offload-host-only), for
// tests/test_host.py.
//
// stdin: the 64-double scalar block with the control attached, npoints, the
// field list (npoints x 3), nsteps, then per step sumsq and avg_m (4 doubles)
// stdout: per step use_init, point_done, state, point, nstep, H of the step
// (3), H of the next step (3): 11 doubles
#include <cstdio>
#include <vector>

#include "sweep_ctl.h"

int main() {
  double c[cpm::S_NSCAL];
  double np = 0, ns = 0;
  if (std::fread(c, sizeof(double), cpm::S_NSCAL, stdin) != (size_t)cpm::S_NSCAL ||
      std::fread(&np, sizeof(double), 1, stdin) != 1)
    return 2;
  std::vector<double> fields(3 * (size_t)np + 3);
  if (np > 0 && std::fread(fields.data(), sizeof(double), 3 * (size_t)np, stdin) != 3 * (size_t)np) return 2;
  if (std::fread(&ns, sizeof(double), 1, stdin) != 1) return 2;
  for (long k = 0; k < (long)ns; ++k) {
    double in[4];
    if (std::fread(in, sizeof(double), 4, stdin) != 4) return 2;
    c[cpm::S_SUMSQ] = in[0];
    for (int d = 0; d < 3; ++d) c[cpm::S_AVGM + d] = in[1 + d];
    cpm::sweep_ctl_update(c, fields.data());
    const double out[11] = {c[cpm::C_USEINIT], c[cpm::C_PDONE], c[cpm::C_STATE], c[cpm::C_POINT], c[cpm::C_NSTEP],
                            c[cpm::C_PH], c[cpm::C_PH + 1], c[cpm::C_PH + 2], c[cpm::C_H], c[cpm::C_H + 1],
                            c[cpm::C_H + 2]};
    std::fwrite(out, sizeof(double), 11, stdout);
  }
  return 0;
}
