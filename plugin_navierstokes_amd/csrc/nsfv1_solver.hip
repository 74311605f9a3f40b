// nsfv1_solver.hip -- device linear algebra for the Newton loop around the
// FV1 assembly: Dirichlet rows, CSR SpMV with the (FP32 or FP64) Jacobian on
// FP64 vectors, point-block Jacobi (one (dim+1)x(dim+1) block per node), and
// the BLAS-1 kernels of restarted GMRES.  All reductions use a fixed
// decomposition and order, so results are bitwise reproducible.
//
// The reference delegates this to ugcore (DirichletBoundary, NewtonSolver,
// linear solvers), which is not part of the plugin; these kernels provide the
// same service for the device-resident Jacobian and defect.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsfv1_device.hpp"
#include "nsfv1_solver.hpp"

namespace nsfv1 {

// ---------------------------------------------------------------------
// Dirichlet rows: J row -> identity, d = u - g (ugcore DirichletBoundary
// row treatment, restated)
// ---------------------------------------------------------------------
template <typename JT>
__global__ void dirichlet_rows(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                               const int64_t* __restrict__ rows, int64_t nrows, JT* __restrict__ J,
                               double* __restrict__ d, const double* __restrict__ u, const double* __restrict__ g)
{
    const int64_t r = (int64_t)blockIdx.x;
    if (r >= nrows) return;
    const int64_t row = rows[r];
    if (J)
        for (int64_t k = rowptr[row] + threadIdx.x; k < rowptr[row + 1]; k += blockDim.x)
            J[k] = colidx[k] == row ? JT(1) : JT(0);
    if (threadIdx.x == 0 && d) d[row] = u[row] - (g ? g[r] : 0.0);
}

// ---------------------------------------------------------------------
// y = A x, 16 lanes per row, fixed-order shuffle reduction
// ---------------------------------------------------------------------
template <typename JT>
__global__ void __launch_bounds__(256)
spmv_csr(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colidx, const JT* __restrict__ A,
         const double* __restrict__ x, double* __restrict__ y, int64_t n)
{
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int sub = threadIdx.x & 15;
    double s = 0.0;
    if (row < n) {
        const int64_t a = rowptr[row], b = rowptr[row + 1];
        for (int64_t k = a + sub; k < b; k += 16) s += (double)A[k] * x[colidx[k]];
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
    if (row < n && sub == 0) y[row] = s;
}

// ---------------------------------------------------------------------
// point-block Jacobi: invert the (nc x nc) diagonal block of every node
// (Gauss-Jordan with partial pivoting in FP64)
// ---------------------------------------------------------------------
template <typename JT, int NC>
__global__ void block_jacobi_setup(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                   const JT* __restrict__ A, int64_t nv, double* __restrict__ Binv,
                                   int* __restrict__ err)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double M[NC][2 * NC];
    const int64_t r0 = rowptr[v * NC];
    const int deg = (int)((rowptr[v * NC + 1] - r0) / NC);
    // slot of the diagonal block in the row
    int lo = 0, hi = deg - 1, j = 0;
    const int32_t key = (int32_t)(v * NC);
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int32_t c = colidx[r0 + (int64_t)mid * NC];
        if (c == key) { j = mid; break; }
        if (c < key) lo = mid + 1; else hi = mid - 1;
    }
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        const int64_t ra = rowptr[v * NC + a] + (int64_t)j * NC;
#pragma unroll
        for (int b = 0; b < NC; ++b) {
            M[a][b] = (double)A[ra + b];
            M[a][NC + b] = a == b ? 1.0 : 0.0;
        }
    }
    bool ok = true;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        int p = c;
        double best = fabs(M[c][c]);
#pragma unroll
        for (int r = c + 1; r < NC; ++r)
            if (fabs(M[r][c]) > best) { best = fabs(M[r][c]); p = r; }
        if (best == 0.0) { ok = false; break; }
        // swap rows c and p with compile-time indices only, so that M
        // stays in registers
#pragma unroll
        for (int r = c + 1; r < NC; ++r)
            if (r == p)
#pragma unroll
                for (int k = 0; k < 2 * NC; ++k) { const double t = M[c][k]; M[c][k] = M[r][k]; M[r][k] = t; }
        const double ip = 1.0 / M[c][c];
#pragma unroll
        for (int k = 0; k < 2 * NC; ++k) M[c][k] *= ip;
#pragma unroll
        for (int r = 0; r < NC; ++r)
            if (r != c) {
                const double f = M[r][c];
#pragma unroll
                for (int k = 0; k < 2 * NC; ++k) M[r][k] -= f * M[c][k];
            }
    }
#pragma unroll
    for (int a = 0; a < NC; ++a)
#pragma unroll
        for (int b = 0; b < NC; ++b) Binv[(v * NC + a) * NC + b] = ok ? M[a][NC + b] : (a == b ? 1.0 : 0.0);
    if (!ok) kernel_error(err, KERR_SINGULAR_BLOCK);
}

template <int NC>
__global__ void block_jacobi_apply(const double* __restrict__ Binv, const double* __restrict__ x,
                                   double* __restrict__ y, int64_t nv)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double xv[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) xv[a] = x[v * NC + a];
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        double s = 0;
#pragma unroll
        for (int b = 0; b < NC; ++b) s += Binv[(v * NC + a) * NC + b] * xv[b];
        y[v * NC + a] = s;
    }
}

// ---------------------------------------------------------------------
// BLAS-1 with deterministic two-stage reductions
// ---------------------------------------------------------------------
constexpr int RED_BLOCKS = 1024;

__device__ inline double block_sum(double s, double* sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = s;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
    __syncthreads();
    return t;
}

// partial[b][i] = sum over block b's share of V_i . w, i < k
__global__ void __launch_bounds__(256)
multi_dot_partial(const double* __restrict__ V, int64_t ld, int k, const double* __restrict__ w, int64_t n,
                  double* __restrict__ partial)
{
    __shared__ double sh[4];
    const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const int64_t a = (int64_t)blockIdx.x * chunk, b = a + chunk < n ? a + chunk : n;
    for (int i = 0; i < k; ++i) {
        double s = 0;
        for (int64_t t = a + threadIdx.x; t < b; t += blockDim.x) s += V[i * ld + t] * w[t];
        const double tot = block_sum(s, sh);
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * 64 + i] = tot;
    }
}

// one workgroup per dot product: every thread sums a fixed strided subset
// of the partials, then the fixed-order block reduction (deterministic)
__global__ void multi_dot_final(const double* __restrict__ partial, int nblk, int k, double* __restrict__ out)
{
    __shared__ double sh[4];
    const int i = blockIdx.x;
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partial[(int64_t)b * 64 + i];
    const double tot = block_sum(s, sh);
    if (threadIdx.x == 0) out[i] = tot;
}

// w -= sum_i h_i V_i
__global__ void multi_axpy(const double* __restrict__ V, int64_t ld, int k, const double* __restrict__ h,
                           double* __restrict__ w, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double s = w[t];
    for (int i = 0; i < k; ++i) s -= h[i] * V[i * ld + t];
    w[t] = s;
}

// y = a x + b y; with b = 0, y is not read (it may hold anything, NaN
// bit patterns of a fresh allocation included)
__global__ void axpby(double a, const double* __restrict__ x, double b, double* __restrict__ y, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = b == 0.0 ? a * x[t] : a * x[t] + b * y[t];
}

// ---------------------------------------------------------------------
// device-side GMRES of the capturable Newton step: the Hessenberg matrix,
// the Givens rotations and the back substitution stay on the device, so a
// fixed number of iterations runs without host synchronisation
// ---------------------------------------------------------------------
__global__ void fill_value(double* __restrict__ x, double v, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] = v;
}

// zeroes n 4-byte words (a kernel, not a memset node, so that a captured
// assembly replays it faithfully)
__global__ void zero_words(uint32_t* __restrict__ x, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] = 0u;
}

// beta = sqrt(||r||^2); the squared norm is also copied to res2 if given
__global__ void gm_beta(const double* __restrict__ n2, double* __restrict__ beta, double* __restrict__ res2)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *beta = sqrt(*n2);
        if (res2) *res2 = *n2;
    }
}

// column j of the (m+1) x m Hessenberg matrix H (row-major) from the two
// classical Gram-Schmidt passes hd[0..j], hd[hs..hs+j] and ||w||^2 = hd[2 hs]
__global__ void gm_hcol(const double* __restrict__ hd, int hs, double* __restrict__ H, int m, int j)
{
    const int i = threadIdx.x;
    if (i <= j) H[i * m + j] = hd[i] + hd[hs + i];
    if (i == 0) H[(j + 1) * m + j] = sqrt(hd[2 * hs]);
}

// y = x / h, or 0 when h = 0 (a breakdown leaves a zero basis vector, whose
// Hessenberg column then vanishes)
__global__ void gm_scale(const double* __restrict__ x, const double* __restrict__ h, double* __restrict__ y, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double hv = *h;
    y[t] = hv > 0.0 ? x[t] / hv : 0.0;
}

// min ||beta e1 - H y|| over y (H overwritten by its rotated triangle):
// the Givens rotations of the host GMRES, then back substitution; a zero
// diagonal gives y_i = 0.  One thread (m <= GMRES_MAX_RESTART).
__global__ void gm_lsq(double* __restrict__ H, const double* __restrict__ beta, int m, double* __restrict__ y)
{
    __shared__ double g[GMRES_MAX_RESTART + 1], cs[GMRES_MAX_RESTART], sn[GMRES_MAX_RESTART];
    if (threadIdx.x != 0) return;
    for (int i = 0; i <= m; ++i) g[i] = 0.0;
    g[0] = *beta;
    for (int j = 0; j < m; ++j) {
        for (int i = 0; i < j; ++i) {
            const double t = cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
            H[(i + 1) * m + j] = -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
            H[i * m + j] = t;
        }
        const double a = H[j * m + j], c = H[(j + 1) * m + j];
        const double r = hypot(a, c);
        cs[j] = r > 0 ? a / r : 1.0;
        sn[j] = r > 0 ? c / r : 0.0;
        H[j * m + j] = r;
        H[(j + 1) * m + j] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
    }
    for (int i = m - 1; i >= 0; --i) {
        double t = g[i];
        for (int k = i + 1; k < m; ++k) t -= H[i * m + k] * y[k];
        y[i] = H[i * m + i] != 0.0 ? t / H[i * m + i] : 0.0;
    }
}

// x += sum_i c_i V_i
__global__ void multi_add(const double* __restrict__ V, int64_t ld, int k, const double* __restrict__ c,
                          double* __restrict__ x, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double s = x[t];
    for (int i = 0; i < k; ++i) s += c[i] * V[i * ld + t];
    x[t] = s;
}


// ---------------------------------------------------------------------
// geometric multigrid pieces: residual, damped block-Jacobi update,
// node-level transfer operators, Dirichlet masking, state injection and
// the dense coarsest-level solve
// ---------------------------------------------------------------------

// Node-block products.  The pattern consists of (dim+1)x(dim+1) node
// blocks: node v has nb neighbours adj[off(v) + s] (off(v) = rowptr[v NC] /
// NC^2), and its NC rows all have nb NC entries, so entry (v a, s NC + c)
// sits at rowptr[v NC] + a nb NC + s NC + c.  A group of G lanes handles one
// node: lane j takes the slots s = j, j + G, ..., reads the neighbour index
// once and the NC x NC block as NC row segments (one 16-byte load per FP32
// row when NC = 4), and keeps NC FP64 partial sums.  The G partials are then
// combined by a reduce-scatter across the group (two exchange steps that
// halve the set of rows a lane carries, then a butterfly), in a fixed order.
// y = A x (RES = 0), y = b - A x (RES = 1), or one damped point-block
// Jacobi sweep y = x + omega Binv (b - A x) (RES = 2; y must not alias x).
template <typename JT, int NC>
__device__ __forceinline__ void block_row(const JT* __restrict__ p, double (&q)[NC])
{
#pragma unroll
    for (int c = 0; c < NC; ++c) q[c] = (double)p[c];
}
template <>
__device__ __forceinline__ void block_row<float, 4>(const float* __restrict__ p, double (&q)[4])
{
    const float4 t = *(const float4*)p;
    q[0] = t.x, q[1] = t.y, q[2] = t.z, q[3] = t.w;
}
template <>
__device__ __forceinline__ void block_row<double, 4>(const double* __restrict__ p, double (&q)[4])
{
    const double2 t0 = *(const double2*)p, t1 = *(const double2*)(p + 2);
    q[0] = t0.x, q[1] = t0.y, q[2] = t1.x, q[3] = t1.y;
}

template <typename JT, int NC, int G, int RES>
__global__ void __launch_bounds__(256)
nodeblock_mv(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ adj, const JT* __restrict__ A,
             const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ y, int64_t nv,
             const double* __restrict__ Binv, double omega)
{
    static_assert(NC == 3 || NC == 4, "node blocks of 3 or 4 unknowns");
    static_assert(G >= 4 && G <= 64 && (G & (G - 1)) == 0, "group size");
    const int64_t v = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int sub = threadIdx.x & (G - 1);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (v < nv) {
        const int64_t base = rowptr[v * NC];
        const int nb = (int)((rowptr[v * NC + 1] - base) / NC);
        const int32_t* ad = adj + base / (NC * NC);
        const JT* Av = A + base;
#pragma unroll 2
        for (int sl = sub; sl < nb; sl += G) {
            const int64_t col = ad[sl];
            double xv[NC];
            if constexpr (NC == 4) {
                const double2 x0 = *(const double2*)(x + col * 4), x1 = *(const double2*)(x + col * 4 + 2);
                xv[0] = x0.x, xv[1] = x0.y, xv[2] = x1.x, xv[3] = x1.y;
            } else {
#pragma unroll
                for (int c = 0; c < NC; ++c) xv[c] = x[col * NC + c];
            }
#pragma unroll
            for (int a = 0; a < NC; ++a) {
                double q[NC];
                block_row<JT, NC>(Av + (int64_t)a * nb * NC + sl * NC, q);
                double t = acc[a];
#pragma unroll
                for (int c = 0; c < NC; ++c) t += q[c] * xv[c];
                acc[a] = t;
            }
        }
    }
    // reduce-scatter over the G lanes of the group: after the first step a
    // lane carries rows {0,1} or {2,3}, after the second one row
    const bool h1 = sub & (G / 2), h2 = sub & (G / 4);
    double w0, w1;
    {
        const double s0 = h1 ? acc[0] : acc[2], s1 = h1 ? acc[1] : acc[3];
        const double r0 = __shfl_xor(s0, G / 2, G), r1 = __shfl_xor(s1, G / 2, G);
        w0 = (h1 ? acc[2] : acc[0]) + r0;
        w1 = (h1 ? acc[3] : acc[1]) + r1;
    }
    double z = (h2 ? w1 : w0) + __shfl_xor(h2 ? w0 : w1, G / 4, G);
#pragma unroll
    for (int o = G / 8; o > 0; o >>= 1) z += __shfl_xor(z, o, G);
    const int a = (h1 ? 2 : 0) + (h2 ? 1 : 0);
    if constexpr (RES == 2) {
        // row c's sum sits on lane (c >> 1) G/2 + (c & 1) G/4 of the group
        double rr[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) rr[c] = __shfl(z, (c >> 1) * (G / 2) + (c & 1) * (G / 4), G);
        if (v < nv && (sub & (G / 4 - 1)) == 0 && a < NC) {
#pragma unroll
            for (int c = 0; c < NC; ++c) rr[c] = b[v * NC + c] - rr[c];
            const int64_t row = v * NC + a;
            double t = 0;
#pragma unroll
            for (int c = 0; c < NC; ++c) t += Binv[row * NC + c] * rr[c];
            y[row] = x[row] + omega * t;
        }
    } else if (v < nv && (sub & (G / 4 - 1)) == 0 && a < NC) {
        const int64_t row = v * NC + a;
        y[row] = RES ? b[row] - z : z;
    }
}

template <typename JT>
__global__ void __launch_bounds__(256)
residual_csr(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colidx, const JT* __restrict__ A,
             const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r, int64_t n)
{
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int sub = threadIdx.x & 15;
    double s = 0.0;
    if (row < n) {
        const int64_t a = rowptr[row], e = rowptr[row + 1];
        for (int64_t k = a + sub; k < e; k += 16) s += (double)A[k] * x[colidx[k]];
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
    if (row < n && sub == 0) r[row] = b[row] - s;
}

// x = (first ? 0 : x) + omega * Binv r, per node block
template <int NC>
__global__ void bj_update(const double* __restrict__ Binv, const double* __restrict__ r, double* __restrict__ x,
                          double omega, int first, int64_t nv)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double rv[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) rv[a] = r[v * NC + a];
#pragma unroll
    for (int a = 0; a < NC; ++a) {
        double s = 0;
#pragma unroll
        for (int b = 0; b < NC; ++b) s += Binv[(v * NC + a) * NC + b] * rv[b];
        x[v * NC + a] = (first ? 0.0 : x[v * NC + a]) + omega * s;
    }
}

// y[row, c] (= or +=) sum_k w_k x[col_k, c] for every component c < nc
__global__ void node_transfer(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                              const double* __restrict__ w, const double* __restrict__ x, double* __restrict__ y,
                              int64_t nrows, int nc, int accumulate)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows * nc) return;
    const int64_t row = t / nc;
    const int c = (int)(t - row * nc);
    double s = 0.0;
    const int64_t a = rowptr[row], b = rowptr[row + 1];
    // unrolled so that several index -> value load chains are in flight
#pragma unroll 4
    for (int64_t k = a; k < b; ++k) s += w[k] * x[(int64_t)colidx[k] * nc + c];
    y[t] = accumulate ? y[t] + s : s;
}

__global__ void zero_rows(const int64_t* __restrict__ rows, int64_t nrows, double* __restrict__ y)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nrows) y[rows[t]] = 0.0;
}

// coarse state: uc[w, c] = uf[inj[w], c]
__global__ void inject_state(const int64_t* __restrict__ inj, const double* __restrict__ uf, double* __restrict__ uc,
                             int64_t nvc, int nc)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvc * nc) return;
    const int64_t w = t / nc;
    const int c = (int)(t - w * nc);
    uc[t] = uf[inj[w] * nc + c];
}

// y = A x with a dense row-major n x n matrix, one wavefront per row
__global__ void __launch_bounds__(256)
dense_matvec(const double* __restrict__ A, const double* __restrict__ x, double* __restrict__ y, int n)
{
    const int row = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    if (row < n)
        for (int k = lane; k < n; k += 64) s += A[(int64_t)row * n + k] * x[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (row < n && lane == 0) y[row] = s;
}

// ---------------------------------------------------------------------
// coarsest-level dense inverse: Gauss-Jordan in exchange form with partial
// pivoting and no row swaps.  The tableau y = M x starts as the coarse
// matrix; step k takes the unpivoted row p with the largest |M[p][k]| and
// exchanges y_p with x_k:
//   M'[p][k] = 1/a, M'[p][j] = -M[p][j]/a,
//   M'[i][k] = M[i][k]/a, M'[i][j] = M[i][j] - M'[i][k] M[p][j]   (i != p),
// a = M[p][k].  After n steps x_k = sum_j M[p_k][j] y_{p_j}, so
// inv[k][p_j] = M[p_k][j].  Ping-pong tableaux; every workgroup repeats the
// pivot search with the same fixed-order reduction (ties go to the smaller
// row).  A column that vanishes on every unpivoted row (a null mode of the
// coarse operator, e.g. the constant pressure of a domain with Dirichlet
// velocity everywhere) pins its unknown to zero in the coarse correction
// (a = 0 is replaced by 1/a := 0), instead of dividing by zero; a column
// with no finite entry left is reported.
// ---------------------------------------------------------------------
template <typename JT>
__global__ void __launch_bounds__(256)
gj_load(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colidx, const JT* __restrict__ J, int n,
        double* __restrict__ M, double* __restrict__ cabs, int* __restrict__ done, int* __restrict__ err)
{
    const int r = blockIdx.x;
    double* row = M + (int64_t)r * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) row[j] = 0.0;
    __syncthreads();
    for (int64_t k = rowptr[r] + threadIdx.x; k < rowptr[r + 1]; k += blockDim.x) row[colidx[k]] = (double)J[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        cabs[r] = fabs(row[0]);
        done[r] = 0;
        if (r == 0) *err = 0;
    }
}

// GJ_BLOCK exchange steps per launch (same pivots, same arithmetic as one
// step per launch, with every update an explicit fma).  Each workgroup
// loads columns k0 .. k0+nb-1 of every row (the panel) and replays the
// steps on it: the pivot search of step s reads the panel after steps
// 0 .. s-1, the pivot row of step s is row p_s of the input evolved through
// those steps, and the multiplier of row r at step s (its column-k value
// after the step, f = M[r][k] / a) is kept for the row updates.  The panel
// and the full rows take the same operations in the same order, so the
// pivots are those the full rows give.  Then each workgroup applies the nb
// steps to its rows.  Pivoted flags ping-pong between launches (din ->
// dout), so no workgroup reads a flag that another one writes.
constexpr int GJ_BLOCK = 4;

__global__ void __launch_bounds__(256)
gj_block(const double* __restrict__ Min, double* __restrict__ Mout, const int* __restrict__ din,
         int* __restrict__ dout, int* __restrict__ piv, int* __restrict__ err, int n, int k0, int nb)
{
    __shared__ double pan[MG_DENSE_MAX][GJ_BLOCK];   // panel columns of every row
    __shared__ double prow[GJ_BLOCK][MG_DENSE_MAX];  // pivot row of step s, before step s
    __shared__ double fm[GJ_BLOCK][MG_DENSE_MAX];    // multiplier of every row at step s
    __shared__ int sdone[MG_DENSE_MAX];
    __shared__ double sv[4], sia[GJ_BLOCK];
    __shared__ int si[4], sp[GJ_BLOCK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int r = t; r < n; r += 256) {
        sdone[r] = din[r];
#pragma unroll
        for (int q = 0; q < GJ_BLOCK; ++q) pan[r][q] = q < nb ? Min[(int64_t)r * n + k0 + q] : 0.0;
    }
    __syncthreads();
    for (int q = 0; q < nb; ++q) {
        const int k = k0 + q;
        // pivot search over the unpivoted rows (ties to the smaller row)
        double bv = -2.0;
        int bi = 0;
        for (int r = t; r < n; r += 256) {
            const double v = sdone[r] ? -1.0 : fabs(pan[r][q]);
            if (v > bv) bv = v, bi = r;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(bv, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            if (ov > bv || (ov == bv && oi < bi)) bv = ov, bi = oi;
        }
        if (lane == 0) sv[w] = bv, si[w] = bi;
        __syncthreads();
        bv = sv[0], bi = si[0];
#pragma unroll
        for (int x = 1; x < 4; ++x)
            if (sv[x] > bv || (sv[x] == bv && si[x] < bi)) bv = sv[x], bi = si[x];
        const int p = bi;
        const bool pin = bv == 0.0;
        if (t == 0 && blockIdx.x == 0 && !(bv >= 0.0)) *err = 1;
        // the pivot row before this step: row p evolved through steps 0 .. q-1
        for (int j = t; j < n; j += 256) {
            double v = Min[(int64_t)p * n + j];
            for (int x = 0; x < q; ++x) {
                const double f = fm[x][p];
                v = j == k0 + x ? f : fma(-f, prow[x][j], v);
            }
            prow[q][j] = v;
        }
        double pv[GJ_BLOCK];
#pragma unroll
        for (int x = 0; x < GJ_BLOCK; ++x) pv[x] = pan[p][x];
        __syncthreads();
        const double ia = pin ? 0.0 : 1.0 / prow[q][k];
        if (t == 0) sp[q] = p, sia[q] = ia;
        for (int r = t; r < n; r += 256) {
            if (r == p) {
                fm[q][r] = ia;
#pragma unroll
                for (int x = 0; x < GJ_BLOCK; ++x)
                    if (x > q) pan[r][x] = -pv[x] * ia;
            } else {
                const double f = pan[r][q] * ia;
                fm[q][r] = f;
#pragma unroll
                for (int x = 0; x < GJ_BLOCK; ++x)
                    if (x > q) pan[r][x] = fma(-f, pv[x], pan[r][x]);
            }
        }
        if (t == 0) sdone[p] = 1;
        __syncthreads();
    }
    if (blockIdx.x == 0 && t < nb) piv[k0 + t] = sp[t];
    // the workgroup's rows through the nb steps
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        for (int j = t; j < n; j += 256) {
            double v = Min[(int64_t)i * n + j];
            for (int x = 0; x < nb; ++x) {
                const int kx = k0 + x;
                if (i == sp[x]) {
                    v = j == kx ? sia[x] : -v * sia[x];
                } else {
                    const double f = fm[x][i];
                    v = j == kx ? f : fma(-f, prow[x][j], v);
                }
            }
            Mout[(int64_t)i * n + j] = v;
        }
        if (t == 0) dout[i] = sdone[i];
    }
}

__global__ void gj_unpermute(const double* __restrict__ M, const int* __restrict__ piv, int n, double* __restrict__ inv)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * n) return;
    const int k = (int)(e / n), j = (int)(e - (int64_t)k * n);
    inv[(int64_t)k * n + piv[j]] = M[(int64_t)piv[k] * n + j];
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
static inline int blocks_for(int64_t n, int t) { return (int)((n + t - 1) / t); }

hipError_t launch_dirichlet(int jprec, const int64_t* rowptr, const int32_t* colidx, const int64_t* rows,
                            int64_t nrows, void* J, double* d, const double* u, const double* g, hipStream_t s)
{
    if (nrows <= 0) return hipSuccess;
    if (jprec == 64)
        dirichlet_rows<double><<<(unsigned)nrows, 64, 0, s>>>(rowptr, colidx, rows, nrows, (double*)J, d, u, g);
    else
        dirichlet_rows<float><<<(unsigned)nrows, 64, 0, s>>>(rowptr, colidx, rows, nrows, (float*)J, d, u, g);
    return hipGetLastError();
}

// node-block launch: G lanes per node, G the power of two (8, 16 or 32) that
// covers the mean number of neighbour blocks per node
template <typename JT, int NC, int RES>
static void launch_nodeblock(const int64_t* rowptr, const int32_t* adj, const JT* A, const double* x, const double* b,
                             double* y, int64_t nv, int64_t nnz, const double* Binv, double omega, hipStream_t s)
{
    const int64_t slots = nv > 0 ? nnz / ((int64_t)NC * NC * nv) : 0;
    if (slots > 16)
        nodeblock_mv<JT, NC, 32, RES><<<blocks_for(nv * 32, 256), 256, 0, s>>>(rowptr, adj, A, x, b, y, nv, Binv, omega);
    else if (slots > 8)
        nodeblock_mv<JT, NC, 16, RES><<<blocks_for(nv * 16, 256), 256, 0, s>>>(rowptr, adj, A, x, b, y, nv, Binv, omega);
    else
        nodeblock_mv<JT, NC, 8, RES><<<blocks_for(nv * 8, 256), 256, 0, s>>>(rowptr, adj, A, x, b, y, nv, Binv, omega);
}

// with 4 unknowns per node the node-block kernels take 16-byte vector loads
// of A and x
static bool nodeblock_ok(const int32_t* adj, int nc, const void* A, const double* x)
{
    return adj && (nc == 3 || (nc == 4 && (((uintptr_t)A | (uintptr_t)x) & 15) == 0));
}

template <int RES>
static void launch_mv(int jprec, const int64_t* rowptr, const int32_t* adj, int nc, const void* A, const double* x,
                      const double* b, double* y, int64_t n, int64_t nnz, hipStream_t s,
                      const double* Binv = nullptr, double omega = 0.0)
{
    const int64_t nv = n / nc;
    if (nc == 4) {
        if (jprec == 64) launch_nodeblock<double, 4, RES>(rowptr, adj, (const double*)A, x, b, y, nv, nnz, Binv, omega, s);
        else launch_nodeblock<float, 4, RES>(rowptr, adj, (const float*)A, x, b, y, nv, nnz, Binv, omega, s);
    } else {
        if (jprec == 64) launch_nodeblock<double, 3, RES>(rowptr, adj, (const double*)A, x, b, y, nv, nnz, Binv, omega, s);
        else launch_nodeblock<float, 3, RES>(rowptr, adj, (const float*)A, x, b, y, nv, nnz, Binv, omega, s);
    }
}

hipError_t launch_spmv(int jprec, const int64_t* rowptr, const int32_t* colidx, const int32_t* adj, int nc,
                       const void* A, const double* x, double* y, int64_t n, int64_t nnz, hipStream_t s)
{
    if (nodeblock_ok(adj, nc, A, x)) {
        launch_mv<0>(jprec, rowptr, adj, nc, A, x, nullptr, y, n, nnz, s);
    } else {
        const int bl = blocks_for(n * 16, 256);
        if (jprec == 64) spmv_csr<double><<<bl, 256, 0, s>>>(rowptr, colidx, (const double*)A, x, y, n);
        else spmv_csr<float><<<bl, 256, 0, s>>>(rowptr, colidx, (const float*)A, x, y, n);
    }
    return hipGetLastError();
}

hipError_t launch_block_jacobi_setup(int jprec, int nc, const int64_t* rowptr, const int32_t* colidx, const void* A,
                                     int64_t nv, double* Binv, int* err, hipStream_t s)
{
    const int bl = blocks_for(nv, 128);
    if (nc == 3) {
        if (jprec == 64) block_jacobi_setup<double, 3><<<bl, 128, 0, s>>>(rowptr, colidx, (const double*)A, nv, Binv, err);
        else block_jacobi_setup<float, 3><<<bl, 128, 0, s>>>(rowptr, colidx, (const float*)A, nv, Binv, err);
    } else {
        if (jprec == 64) block_jacobi_setup<double, 4><<<bl, 128, 0, s>>>(rowptr, colidx, (const double*)A, nv, Binv, err);
        else block_jacobi_setup<float, 4><<<bl, 128, 0, s>>>(rowptr, colidx, (const float*)A, nv, Binv, err);
    }
    return hipGetLastError();
}

hipError_t launch_block_jacobi_apply(int nc, const double* Binv, const double* x, double* y, int64_t nv, hipStream_t s)
{
    const int bl = blocks_for(nv, 256);
    if (nc == 3) block_jacobi_apply<3><<<bl, 256, 0, s>>>(Binv, x, y, nv);
    else block_jacobi_apply<4><<<bl, 256, 0, s>>>(Binv, x, y, nv);
    return hipGetLastError();
}

hipError_t launch_multi_dot(const double* V, int64_t ld, int k, const double* w, int64_t n, double* partial,
                            double* out, hipStream_t s)
{
    multi_dot_partial<<<RED_BLOCKS, 256, 0, s>>>(V, ld, k, w, n, partial);
    if (k > 0) multi_dot_final<<<k, 256, 0, s>>>(partial, RED_BLOCKS, k, out);
    return hipGetLastError();
}

hipError_t launch_multi_axpy(const double* V, int64_t ld, int k, const double* h, double* w, int64_t n, hipStream_t s)
{
    multi_axpy<<<blocks_for(n, 256), 256, 0, s>>>(V, ld, k, h, w, n);
    return hipGetLastError();
}

hipError_t launch_multi_add(const double* V, int64_t ld, int k, const double* c, double* x, int64_t n, hipStream_t s)
{
    multi_add<<<blocks_for(n, 256), 256, 0, s>>>(V, ld, k, c, x, n);
    return hipGetLastError();
}

hipError_t launch_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s)
{
    axpby<<<blocks_for(n, 256), 256, 0, s>>>(a, x, b, y, n);
    return hipGetLastError();
}

hipError_t launch_zero_words(void* x, int64_t nwords, hipStream_t s)
{
    if (nwords <= 0) return hipSuccess;
    zero_words<<<blocks_for(nwords, 256), 256, 0, s>>>(static_cast<uint32_t*>(x), nwords);
    return hipGetLastError();
}

hipError_t launch_fill(double* x, double v, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    fill_value<<<blocks_for(n, 256), 256, 0, s>>>(x, v, n);
    return hipGetLastError();
}

hipError_t launch_gm_beta(const double* n2, double* beta, double* res2, hipStream_t s)
{
    gm_beta<<<1, 64, 0, s>>>(n2, beta, res2);
    return hipGetLastError();
}

hipError_t launch_gm_hcol(const double* hd, int hs, double* H, int m, int j, hipStream_t s)
{
    if (j < 0 || j >= m || m > GMRES_MAX_RESTART || hs < m) return hipErrorInvalidValue;
    gm_hcol<<<1, 64, 0, s>>>(hd, hs, H, m, j);
    return hipGetLastError();
}

hipError_t launch_gm_scale(const double* x, const double* h, double* y, int64_t n, hipStream_t s)
{
    gm_scale<<<blocks_for(n, 256), 256, 0, s>>>(x, h, y, n);
    return hipGetLastError();
}

hipError_t launch_gm_lsq(double* H, const double* beta, int m, double* y, hipStream_t s)
{
    if (m < 1 || m > GMRES_MAX_RESTART) return hipErrorInvalidValue;
    gm_lsq<<<1, 64, 0, s>>>(H, beta, m, y);
    return hipGetLastError();
}

hipError_t launch_residual(int jprec, const int64_t* rowptr, const int32_t* colidx, const int32_t* adj, int nc,
                           const void* A, const double* x, const double* b, double* r, int64_t n, int64_t nnz,
                           hipStream_t s)
{
    if (nodeblock_ok(adj, nc, A, x)) {
        launch_mv<1>(jprec, rowptr, adj, nc, A, x, b, r, n, nnz, s);
    } else {
        const int bl = blocks_for(n * 16, 256);
        if (jprec == 64) residual_csr<double><<<bl, 256, 0, s>>>(rowptr, colidx, (const double*)A, x, b, r, n);
        else residual_csr<float><<<bl, 256, 0, s>>>(rowptr, colidx, (const float*)A, x, b, r, n);
    }
    return hipGetLastError();
}

hipError_t launch_smooth(int jprec, const int64_t* rowptr, const int32_t* colidx, const int32_t* adj, int nc,
                         const void* A, const double* x, const double* b, const double* Binv, double omega, double* r,
                         double* y, int64_t n, int64_t nnz, hipStream_t s)
{
    if (nodeblock_ok(adj, nc, A, x)) {
        launch_mv<2>(jprec, rowptr, adj, nc, A, x, b, y, n, nnz, s, Binv, omega);
        return hipGetLastError();
    }
    hipError_t e = launch_residual(jprec, rowptr, colidx, adj, nc, A, x, b, r, n, nnz, s);
    if (e == hipSuccess) e = hipMemcpyAsync(y, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = launch_bj_update(nc, Binv, r, y, omega, 0, n / nc, s);
    return e;
}

hipError_t launch_bj_update(int nc, const double* Binv, const double* r, double* x, double omega, int first,
                            int64_t nv, hipStream_t s)
{
    const int bl = blocks_for(nv, 256);
    if (nc == 3) bj_update<3><<<bl, 256, 0, s>>>(Binv, r, x, omega, first, nv);
    else bj_update<4><<<bl, 256, 0, s>>>(Binv, r, x, omega, first, nv);
    return hipGetLastError();
}

hipError_t launch_node_transfer(const int64_t* rowptr, const int32_t* colidx, const double* w, const double* x,
                                double* y, int64_t nrows, int nc, int accumulate, hipStream_t s)
{
    node_transfer<<<blocks_for(nrows * nc, 256), 256, 0, s>>>(rowptr, colidx, w, x, y, nrows, nc, accumulate);
    return hipGetLastError();
}

hipError_t launch_zero_rows(const int64_t* rows, int64_t nrows, double* y, hipStream_t s)
{
    if (nrows <= 0) return hipSuccess;
    zero_rows<<<blocks_for(nrows, 256), 256, 0, s>>>(rows, nrows, y);
    return hipGetLastError();
}

hipError_t launch_inject(const int64_t* inj, const double* uf, double* uc, int64_t nvc, int nc, hipStream_t s)
{
    inject_state<<<blocks_for(nvc * nc, 256), 256, 0, s>>>(inj, uf, uc, nvc, nc);
    return hipGetLastError();
}

size_t dense_inverse_work_bytes(int n)
{
    return sizeof(double) * (2 * (size_t)n * n + 2 * (size_t)n) + sizeof(int) * (2 * (size_t)n + 4);
}

hipError_t launch_dense_inverse(int jprec, const int64_t* rowptr, const int32_t* colidx, const void* J, int n,
                                void* work, double* inv, int* err, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    double* M[2] = {(double*)work, (double*)work + (size_t)n * n};
    double* c[2] = {M[1] + (size_t)n * n, M[1] + (size_t)n * n + n};
    int* done = (int*)(c[1] + n);
    int* piv = done + n;
    if (jprec == 64) gj_load<double><<<n, 256, 0, s>>>(rowptr, colidx, (const double*)J, n, M[0], c[0], done, err);
    else gj_load<float><<<n, 256, 0, s>>>(rowptr, colidx, (const float*)J, n, M[0], c[0], done, err);
    if (n > MG_DENSE_MAX) return hipErrorInvalidValue;
    // GJ_BLOCK steps per launch; the pivoted flags alternate between `done`
    // and the (unused here) second pivot-magnitude array
    int* dflag[2] = {done, reinterpret_cast<int*>(c[1])};
    const int nl = (n + GJ_BLOCK - 1) / GJ_BLOCK;
    const unsigned g = n < 256 ? (unsigned)n : 256u;
    for (int l = 0; l < nl; ++l) {
        const int k0 = l * GJ_BLOCK, nb = n - k0 < GJ_BLOCK ? n - k0 : GJ_BLOCK;
        gj_block<<<g, 256, 0, s>>>(M[l & 1], M[(l + 1) & 1], dflag[l & 1], dflag[(l + 1) & 1], piv, err, n, k0, nb);
    }
    gj_unpermute<<<blocks_for((int64_t)n * n, 256), 256, 0, s>>>(M[nl & 1], piv, n, inv);
    return hipGetLastError();
}

hipError_t launch_dense_matvec(const double* A, const double* x, double* y, int n, hipStream_t s)
{
    dense_matvec<<<blocks_for((int64_t)n * 64, 256), 256, 0, s>>>(A, x, y, n);
    return hipGetLastError();
}

}  // namespace nsfv1
