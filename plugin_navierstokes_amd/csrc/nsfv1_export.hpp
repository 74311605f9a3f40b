// nsfv1_export.hpp -- launcher of the data-export kernel (values and
// gradients at SCVF ips, corners or given local points of every element).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsfv1 {

enum { EX_VELOCITY = 0, EX_VELOCITY_GRAD = 1, EX_VELOCITY_IP = 2, EX_PRESSURE = 3, EX_PRESSURE_GRAD = 4 };
enum { EX_AT_SCVF = 0, EX_AT_SCV = 1, EX_AT_LOCAL = 2 };
constexpr int EX_MAX_LOCAL = 32;  // local points per call

struct ExportArgs {
    const double* xyz;
    const int32_t* conn;
    int64_t ne;
    int npt;    // points per element
    int ncomp;  // doubles per point
    int what, at;
    double lips[EX_MAX_LOCAL][3];
};

hipError_t launch_export_points(int et, const ExportArgs& a, const double* u, double* out, int* err, hipStream_t s);

}  // namespace nsfv1
