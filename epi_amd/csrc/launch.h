// launch.h -- host-side launchers of the kernel families (one per .hip file).
#pragma once
#include <hip/hip_runtime.h>

namespace epi {
struct Tis;

void launch_forces(hipStream_t s, const Tis* dT, int ntis, int pb, int cb, int max_memb, int max_other);
void launch_renew(hipStream_t s, const Tis* dT, int ntis, int pb, int cb, int cpar, int max_n);
void launch_initialize_sc(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int pb, int max_n);
void launch_periodic_fix(hipStream_t s, const Tis* dT, int ntis, int pb, int max_n);
void launch_connect(hipStream_t s, const Tis* dT, int ntis, int pb, int nb, int max_n, int max_buckets);
void launch_zzmax(hipStream_t s, const Tis* dT, int ntis, int pb, int max_n);
void launch_zz_set(hipStream_t s, const Tis* dT, int ntis, double v);
void launch_maps(hipStream_t s, const Tis* dT, int ntis, int pb, int max_n, long long max_gsize, int max_halo);
void launch_ext_stim(hipStream_t s, const Tis* dT, int ntis, int bp, int max_nx, int max_ny, int izb_est);
void launch_ca_begin(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int cb, int max_n, long long max_box);
void launch_ca_substep(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int cb, int cp, int ap, int bp,
                       int max_n, int max_nx, int max_ny, int izb_est);
void launch_ca_touch(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int cb, int max_n, long long max_box);
void launch_ca_diffu_zero(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int max_n);
void launch_ca_end(hipStream_t s, const Tis* dT, const int* tlist, int ntis, int max_n);
}  // namespace epi
