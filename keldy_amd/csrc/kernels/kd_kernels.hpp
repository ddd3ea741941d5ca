// kd_kernels.hpp -- launchers of the device kernels (internal to libkeldy_amd).
#pragma once

#include <hip/hip_runtime.h>

#include "kd_device.hpp"
#include "kd_flow_device.hpp"

namespace kd {

// prep_kernel<order>: grid blocks of 256 threads over p.n_points points.
hipError_t launch_prep(int order, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st);
// integrand_kernel<order, T> (the Keldysh sum): grid blocks of 64 threads.
hipError_t launch_lu(int order, int precision, const KdParams &p, const PrepBuf &buf, int grid, hipStream_t st);
// acc[0..3] += sum over blocks of part[4*blk + 0..3] (fixed order).
hipError_t launch_reduce(const double *part, int nblk, double *acc, hipStream_t st);

} // namespace kd
