// nsfv1_march.hpp -- host interface of the plane-marching assembly of
// structured hexahedral grids (nsfv1_march.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsfv1_device.hpp"

namespace nsfv1 {

// A lexicographic grid of nx x ny x nz hexahedra: node (i, j, k) is
// i + NX (j + NY k) with NX = nx + 1, NY = ny + 1; element (i, j, k) is
// i + nx (j + ny k) with the UG4 corner order (bottom face counter-
// clockwise from (i, j, k), then the top face).
struct MarchGrid {
    int nx, ny, nz;
    int64_t NX, NXY;      // node strides in y and z
    int tiles_x, tiles_y;
    int zc;               // node planes per work unit
    int nzc;              // work units along z
    int64_t units;
};

// tile of a work unit: TX x TY node columns
constexpr int MARCH_TX = 7, MARCH_TY = 8;

// Assembly of the A part (plus the mass part and source, by parts) of a
// structured grid of parallelepipeds with FIELDS + one upwind object of
// type `fast` (no, full, skewed) and the RAW diffusion length; J is FP32
// and 16-byte aligned (or null), d FP64 and 16-byte aligned (or null).
hipError_t launch_assemble_march(int fast, const MarchGrid& g, const double* xyz, const KParams& p, const double* u,
                                 const double* uold, uint8_t* upw_dec, float* J, double* d, int* err,
                                 hipStream_t stream);

}  // namespace nsfv1
