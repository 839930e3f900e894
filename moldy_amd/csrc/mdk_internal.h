// mdk_internal.h -- device data layout shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdk_pot.h"

// 3x3 tensors are Fortran column-major: T(i,j) == t[(j-1)*3 + (i-1)].
#define TT(a, i, j) (a)[((j)-1) * 3 + ((i)-1)]

// Box and thermodynamic state of one system, resident on the device.  Two
// copies exist; a step reads copy `cur` in its drift kernel and writes the
// advanced box into the other copy, which every later kernel of the step uses.
struct BoxState {
  double b0[9], b1[9], b2[9], b3[9];  // b3: box jolt of the predictor-corrector
  double tk[9], tc[9], tg[9], tp[9], tginv[9], tgdot[9], tgid[9], fb[9];
  double tgid_lag[9];   // tgid seen by the velocity half-kick (quench_m.F90:595)
  double binv[9];       // cofactor(b0)/vol used by update_therm_avs (dynamics_m.F90:692-700)
  double bref[9];       // b0 when the current neighbour list was built (skin bound)
  double vol, pe, pe_emb, ke, s1, s2, s3, te, h, th, boxke;
  double maxdisp2;      // max |B0 (s - s_ref)|^2 since the last full-list build
  double maxstep2;      // max |B0 (x1 + x2)|^2 of the last step (next drift)
  int status;           // MDK_ERR_* raised on the device (singular box)
  int pad_;
};

// Kernel classes for the HIP-event timing accumulators.
enum { TC_DRIFT = 0, TC_DENSITY, TC_FORCE, TC_KICK, TC_REBUILD, TC_NCLASS };

// Per-block partial sums of force() (value-major, one column of nblocks per
// value): [0] pair energy and [1..6] virial W (xx yy zz xy xz yz) from the
// force pass; [7..15] kinten's tk (column-major) from the kernel that last
// set the velocities (half kick, drift, predictor, or k_kin); [16] embedding
// energy from the density pass
#define NPART_FORCE 7
#define PART_KIN 7
#define PART_EMB 16
#define NPART_F 17
// Partials of the half-kick kernel: s1, s2, s3, maxdisp2, maxstep2
#define NPART_KICK 5

#define MDK_BLOCK 256
