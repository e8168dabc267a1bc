// Device scalar slots of a simulation (double offsets in its scalar block)
// and the hysteresis control that runs on them after every step.
//
// The block holds the step's scalars, then the state of the reference's
// loop control (Micromagnetics.py:109,299-335) when a control is attached
// (cpm_sim_control); the whole block reaches the host in one transfer at the
// end of a step.  sweep_ctl_update is written for host and device so that
// the rule is tested on the CPU (tests/sweep_ctl_host.hip).
#pragma once
#include <cstdint>

#ifndef __HIPCC__
#define __host__
#define __device__
#endif

namespace cpm {

enum : int {
  S_AVGM = 0, S_AVGM100 = 3, S_AVGE = 6, S_EBAR = 12, S_SUMSQ = 18,
  C_ON = 32,       // 1: the control below drives the field and use_init
  C_H = 33,        // (3) field of the step being run, reduced units
  C_DH = 36,       // (3) field step on LLG convergence (H -= dH); field lists: unused
  C_ACC = 39,      // LLG accuracy: nDOF * llg_accuracy_factor
  C_INITF = 40,    // use_init when sumsq < initial_factor * accuracy
  C_STOPDIR = 41,  // (3) stop condition: avg_m . dir <= value
  C_STOPVAL = 44,
  C_MINSTEP = 45,  // LLG convergence needs more than this many steps at a point
  C_SPP = 46,      // steps per point cap (0: none)
  C_NPTS = 47,     // field points of a list (0: H -= dH stepping)
  C_NSTEP = 48,    // steps at the current point (after the last step)
  C_POINT = 49,    // index of the current point
  C_USEINIT = 50,  // use_init of the next step
  C_PDONE = 51,    // the last step completed a field point
  C_STATE = 52,    // 0 running, 1 stop condition met, 2 field list done
  C_PH = 53,       // (3) field of the last step
  C_STEPS = 56,    // steps run under the control
  S_NSCAL = 64
};

// After a step whose sumsq and new average magnetization are in the block:
// use_init for the next step (:300); LLG convergence, sumsq < accuracy after
// more than llg_min_step steps at the point (:309,324), or the per-point step
// cap completes the field point; on convergence the field steps by -dH
// (:329-333), with a field list it moves to the list's next point (after
// the last one the state is 2); then the stop condition of the loop
// (:109) on the new average (state 1).
__host__ __device__ inline void sweep_ctl_update(double* c, const double* fields) {
  if (c[C_ON] == 0.0 || c[C_STATE] != 0.0) return;
  const double sumsq = c[S_SUMSQ];
  c[C_STEPS] += 1.0;
  const double nstep = c[C_NSTEP] + 1.0;
  c[C_USEINIT] = (sumsq < c[C_INITF] * c[C_ACC]) ? 1.0 : 0.0;
  for (int k = 0; k < 3; ++k) c[C_PH + k] = c[C_H + k];
  const bool conv = sumsq < c[C_ACC] && nstep > c[C_MINSTEP];
  const bool cap = c[C_SPP] > 0.0 && nstep >= c[C_SPP];
  c[C_PDONE] = (conv || cap) ? 1.0 : 0.0;
  c[C_NSTEP] = nstep;
  if (conv || cap) {
    c[C_NSTEP] = 0.0;
    c[C_POINT] += 1.0;
    if (c[C_NPTS] > 0.0) {
      if (c[C_POINT] >= c[C_NPTS]) {
        c[C_STATE] = 2.0;
      } else {
        const int64_t q = (int64_t)c[C_POINT];
        for (int k = 0; k < 3; ++k) c[C_H + k] = fields[3 * q + k];
      }
    } else if (conv) {
      for (int k = 0; k < 3; ++k) c[C_H + k] -= c[C_DH + k];
    }
  }
  const double proj = c[S_AVGM] * c[C_STOPDIR] + c[S_AVGM + 1] * c[C_STOPDIR + 1] + c[S_AVGM + 2] * c[C_STOPDIR + 2];
  if (c[C_STATE] == 0.0 && !(proj > c[C_STOPVAL])) c[C_STATE] = 1.0;
}

}  // namespace cpm
