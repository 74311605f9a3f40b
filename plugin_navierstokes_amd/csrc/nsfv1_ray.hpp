// nsfv1_ray.hpp -- device helpers shared by the assembly kernels: small
// vector operations, the element centre, the side/ray intersection of the
// skewed upwinds and the LDS row accumulation.
#pragma once

#include <hip/hip_runtime.h>

#include "nsfv1_topo.hpp"

namespace nsfv1 {

#define NSFV1_DBL_MAX 1.7976931348623157e308

// The _nc helpers round every product and sum separately (no FMA
// contraction), as the oracle's arithmetic does: a cross-product component
// a*b - c*d whose products are equal is then exactly zero.
template <int D>
__device__ inline double dot_nc(const double* a, const double* b)
{
#pragma clang fp contract(off)
    double s = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) s += a[i] * b[i];
    return s;
}

template <int D>
__device__ inline double dist2_nc(const double* a, const double* b)
{
#pragma clang fp contract(off)
    double s = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) s += (a[i] - b[i]) * (a[i] - b[i]);
    return s;
}

__device__ inline void cross_nc(const double* a, const double* b, double* c)
{
#pragma clang fp contract(off)
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// element centre (average of the corners, oracle operation order)
template <int ET>
__device__ inline void elem_centre(const double (&co)[Topo<ET>::nsh][Topo<ET>::dim], double* c)
{
    constexpr int N = Topo<ET>::nsh, D = Topo<ET>::dim;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double s = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) s += co[k][i];
        c[i] = s / N;
    }
}

// ---------------------------------------------------------------------
// side/ray intersection (ugcore ElementSideRayIntersection, restated):
// the first side in reference order hit at ray parameter t <= 0 with side
// parameters inside up to 1e-10, searched upwind from the ip along the ip
// velocity (upwind.cpp:351-353, 547-549).
// ---------------------------------------------------------------------
#define NSFV1_RAY_EPS 1e-10

__device__ inline bool ray_tri_nc(const double* p0, const double* p1, const double* p2, const double* from,
                                  const double* dir, double& b1o, double& b2o)
{
    // no FMA contraction: the oracle's rounding, operation for operation
#pragma clang fp contract(off)
    double e1[3], e2[3], pv[3], tv[3], qv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) { e1[i] = p1[i] - p0[i]; e2[i] = p2[i] - p0[i]; }
    cross_nc(dir, e2, pv);
    const double det = dot_nc<3>(e1, pv);
    if (det == 0.0) return false;
    const double inv = 1.0 / det;
#pragma unroll
    for (int i = 0; i < 3; ++i) tv[i] = from[i] - p0[i];
    const double b1 = dot_nc<3>(tv, pv) * inv;
    cross_nc(tv, e1, qv);
    const double b2 = dot_nc<3>(dir, qv) * inv;
    const double t = dot_nc<3>(e2, qv) * inv;
    if (b1 < -NSFV1_RAY_EPS || b2 < -NSFV1_RAY_EPS || b1 + b2 > 1.0 + NSFV1_RAY_EPS) return false;
    if (t > 0.0) return false;
    b1o = b1;
    b2o = b2;
    return true;
}

template <int ET, int S, int TR>
__device__ inline void try_side(const double (&co)[Topo<ET>::nsh][Topo<ET>::dim], const double* from,
                                const double* dir, int& side, double* gi, double* li)
{
#pragma clang fp contract(off)
    using T = Topo<ET>;
    if (side >= 0) return;
    if constexpr (T::dim == 2) {
        if constexpr (TR == 0) {
            constexpr int a = T::side[S][0], b = T::side[S][1];
            const double ex = co[b][0] - co[a][0], ey = co[b][1] - co[a][1];
            const double rx = co[a][0] - from[0], ry = co[a][1] - from[1];
            const double det = -dir[0] * ey + ex * dir[1];
            if (det == 0.0) return;
            const double t = (-rx * ey + ex * ry) / det;
            const double lam = (dir[0] * ry - dir[1] * rx) / det;
            if (lam < -NSFV1_RAY_EPS || lam > 1.0 + NSFV1_RAY_EPS) return;
            if (t > 0.0) return;
            side = S;
            gi[0] = co[a][0] + lam * ex;
            gi[1] = co[a][1] + lam * ey;
#pragma unroll
            for (int i = 0; i < 2; ++i) li[i] = (1.0 - lam) * T::lc[a][i] + lam * T::lc[b][i];
        }
    } else {
        if constexpr (TR == 0 || T::side_n[S] == 4) {
            constexpr int a = T::side[S][0];
            constexpr int b = T::side[S][TR == 0 ? 1 : 2];
            constexpr int c = T::side[S][TR == 0 ? 2 : 3];
            double b1, b2;
            if (!ray_tri_nc(co[a], co[b], co[c], from, dir, b1, b2)) return;
            side = S;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                gi[i] = co[a][i] + b1 * (co[b][i] - co[a][i]) + b2 * (co[c][i] - co[a][i]);
                li[i] = (1.0 - b1 - b2) * T::lc[a][i] + b1 * T::lc[b][i] + b2 * T::lc[c][i];
            }
        }
    }
}

template <int ET, int S>
__device__ inline void side_loop(const double (&co)[Topo<ET>::nsh][Topo<ET>::dim], const double* from,
                                 const double* dir, int& side, double* gi, double* li)
{
    if constexpr (S < Topo<ET>::nside) {
        try_side<ET, S, 0>(co, from, dir, side, gi, li);
        try_side<ET, S, 1>(co, from, dir, side, gi, li);
        side_loop<ET, S + 1>(co, from, dir, side, gi, li);
    }
}

// Adds a column block of NC values to a row in LDS.  Blocks of four are
// 16-byte aligned (rows are padded to 16 bytes), so they move as 128-bit
// LDS loads and stores.
template <int NC, typename T>
__device__ inline void lds_add_block(T* p, const T (&x)[NC])
{
    if constexpr (NC == 4 && sizeof(T) == 4) {
        float4* q = reinterpret_cast<float4*>(p);
        float4 t = *q;
        t.x += x[0]; t.y += x[1]; t.z += x[2]; t.w += x[3];
        *q = t;
    } else if constexpr (NC == 4 && sizeof(T) == 8) {
        double2* q = reinterpret_cast<double2*>(p);
        double2 t0 = q[0], t1 = q[1];
        t0.x += x[0]; t0.y += x[1]; t1.x += x[2]; t1.y += x[3];
        q[0] = t0;
        q[1] = t1;
    } else {
#pragma unroll
        for (int b = 0; b < NC; ++b) p[b] += x[b];
    }
}

}  // namespace nsfv1
