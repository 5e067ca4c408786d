// rd_dev.h -- device records of the frozen reaction-diffusion window and
// its adjoint (k_adjoint.hip, host.hip).
//
// A window is the sub-step of calc_ca2p (ca2p.cpp:392-399) run n_sub times
// from a snapshot of the tissue, with positions, lists, states and voxel maps
// frozen and four parameter groups made runtime values: the Ca2+ and IP3
// gap-junction coefficients (Tis::ca_du, Tis::ca_dp), the starting
// conductance of every edge, and the per-cell IP3 production Tis::ca_kpa.
#pragma once
#include "epi_dev.h"

namespace epi {

struct RdTis {
  int nafm;                    // AFM cells of the tissue (loss normaliser)
  int izb;                     // iz_bound of the window
  long long box;               // grid entries of planes 0..izb (saved per state)
  // snapshot: [0] = the current buffer, [1] = the other one
  double *u0[2], *p0[2], *A0[2];
  double *h0, *avg0, *kpa0, *g0;
  // saved states of one segment (slot-major) and the checkpoints.  Of the
  // ATP grid a segment state keeps only each AFM cell's grid_avg8 (sAv), the
  // one value the reverse sub-step reads from it
  double *su, *sp, *sh, *sg, *sAv;
  double *cu, *cp_, *ch, *cg, *cA;   // checkpoint-major, two buffers each for u, p, A
  // adjoint: [q] = the state after the sub-step, [1-q] = before it
  double *lu[2], *lp[2], *lh[2], *lg[2];
  // ATP adjoint in three rotating buffers: [a] after the sub-step, [a+1]
  // before it, [a+2] dead (cleared for the next sub-step), a = step mod 3
  double* lA[3];
  double *gkpa;                // dL/dkpa per cell
  // the voxels each AFM cell owns in the window's frozen voxel map (its fa
  // source), up to oc per cell, in stamp-box order
  int* own;
  int* own_n;
  int oc;
  double* scal;                // [0] loss, [1] dL/ddu, [2] dL/ddp, [3] dL/dgj0
  double* target;              // per-cell target ca2p
  double* kpa_in;              // per-cell kpa for the next run (set by the host)
};

}  // namespace epi
