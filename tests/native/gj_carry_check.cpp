// Host build of the row assembly's gap-junction carry-over search
// (epi_amd/csrc/gj_carry.h), for tests/test_gj_carry.py: the accessor records
// every slot the search reads, so the test can check that it stays inside the
// previous row.
#include <cstdint>

#include "../../epi_amd/csrc/gj_carry.h"

extern "C" int carry_slot(int e, int olen, const int64_t* prev_uids, int64_t u, int* lo, int* hi) {
  *lo = 1 << 30;
  *hi = -1;
  return epi::gj_carry_slot(e, olen, (long long)u, [&](int q) {
    if (q < *lo) *lo = q;
    if (q > *hi) *hi = q;
    return (long long)prev_uids[q];
  });
}
