// mech_dev.h -- device record of the frozen mechanics window and its
// adjoint (k_mech.hip, host.hip).
//
// A window is cell_interaction + pos_copy (cell_interaction.cpp:648-685,
// cellmanager.cpp:19-26) run n_steps times from a snapshot of the positions,
// with lists, states, pairs and dermis frozen, without the periodic fix, and
// with eps_m, K_TOTAL, Kspring and KBEND (Tis::m_*) as runtime values.
#pragma once
#include "epi_dev.h"

namespace epi {

struct MechTis {
  double* x0[2];      // snapshot of both position buffers, [0] = current
  double* slot;       // positions of the states of one segment (slot-major, 3 * cap each)
  double* ck;         // checkpoints (3 * cap each)
  double* lam[2];     // position adjoints (3 * cap each)
  double* target;     // 3 * cap: x, y, z
  double* scal;       // [0] loss, [1] dL/deps, [2] dL/dK_TOTAL, [3] dL/dKspring, [4] dL/dKBEND
};

}  // namespace epi
