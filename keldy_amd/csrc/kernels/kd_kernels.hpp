// kd_kernels.hpp -- launchers of the device kernels (internal to libkeldy_amd).
#pragma once

#include <hip/hip_runtime.h>

#include "kd_device.hpp"
#include "kd_flow_device.hpp"

namespace kd {

// prep_kernel<order> (grid_prep blocks of 256) then integrand_kernel<order, T>
// (grid blocks of 64) over p.n_points points (modes MODE_*).
hipError_t launch_integrand(int order, int precision, const KdParams &p, const PrepBuf &buf, int grid_prep,
                            int grid, hipStream_t st);
// acc[0..3] += sum over blocks of part[4*blk + 0..3] (fixed order).
hipError_t launch_reduce(const double *part, int nblk, double *acc, hipStream_t st);

} // namespace kd
