// kernels_fp64.hip -- plain FP64 kernels of the time step, bit-faithful to the reference.
//
// Compiled with -ffp-contract=off (and the pragma below): every a*b+c stays two roundings, as in
// the reference's x86-64 build, so that the stencil sums, the Euler update and the SOR update
// reproduce the reference binary64 results bit for bit.
//
// Stencil sums: the reference multiplies dense Kronecker matrices by the flattened field
// (src/main.c:142-152, src/linearalg.c:236-285): y[r] = 0.0 + A[r][0]*x[0] + A[r][1]*x[1] + ...
// Adding the +-0 products of structural zeros never changes a sum, so the ascending sum over a
// row's 7-column window (zero-padded, starting from 0.0) is the same number.
#pragma clang fp contract(off)

#include <float.h>

#include "cnv_device.h"
#include "sor_rb.h"

namespace cnv {

// ---- velocity boundary values (src/main.c:283-296: rows first, columns win at corners) ----
__device__ inline double bc_u(const double* u, int i, int j, int n, int ld, const WallBC& b) {
    if (j == 0) return b.u1;
    if (j == n - 1) return b.u2;
    if (i == 0) return b.u3;
    if (i == n - 1) return b.u4;
    return u[(size_t)i * ld + j];
}
__device__ inline double bc_v(const double* v, int i, int j, int n, int ld, const WallBC& b) {
    if (j == 0) return b.v1;
    if (j == n - 1) return b.v2;
    if (i == 0) return b.v3;
    if (i == n - 1) return b.v4;
    return v[(size_t)i * ld + j];
}

// (I (x) d) f at (i, j): sum over l of d[j][l] f[i][l]   (acts along j, i.e. x)
__device__ inline double sum_x(const double* coef, const int* start, const double* f, int i, int j, int ld) {
    const double* c = coef + (size_t)j * kStencilWidth;
    const double* row = f + (size_t)i * ld + start[j];
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m) s += c[m] * row[m];
    return s;
}
// (d (x) I) f at (i, j): sum over k of d[i][k] f[k][j]   (acts along i, i.e. y)
__device__ inline double sum_y(const double* coef, const int* start, const double* f, int i, int j, int ld) {
    const double* c = coef + (size_t)i * kStencilWidth;
    const double* col = f + (size_t)start[i] * ld + j;
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m) s += c[m] * col[(size_t)m * ld];
    return s;
}

// ---- wall vorticity: w[edge] = (Dx v - Dy u)[edge] on BC-applied u, v (src/main.c:298-320) --
__global__ void __launch_bounds__(kBlock) wall_vorticity_kernel(double* w, const double* u, const double* v,
                                                               int n, int ld, Stencils s, WallBC b) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    int i, j;
    if (t < n) { i = 0; j = t; }
    else if (t < 2 * n) { i = n - 1; j = t - n; }
    else if (t < 3 * n - 2) { i = t - 2 * n + 1; j = 0; }
    else if (t < 4 * n - 4) { i = t - (3 * n - 2) + 1; j = n - 1; }
    else return;
    const double* c1x = s.d1x + (size_t)j * kStencilWidth;
    const double* c1y = s.d1y + (size_t)i * kStencilWidth;
    const int sj = s.start[j], si = s.start[i];
    double dvdx = 0.0, dudy = 0.0;
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m) dvdx += c1x[m] * bc_v(v, i, sj + m, n, ld, b);
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m) dudy += c1y[m] * bc_u(u, si + m, j, n, ld, b);
    w[(size_t)i * ld + j] = dvdx - dudy;
}

// ---- derivatives of w and the explicit Euler update, every point (src/main.c:323-344,
//      src/fluiddyn.c:98):  w' = (-u*dwdx - v*dwdy + (1./Re)*(d2wdx2+d2wdy2))*dt + w ---------
__global__ void __launch_bounds__(kBlock) transport_fp64_kernel(const double* __restrict__ w,
                                                               double* __restrict__ wout,
                                                               const double* __restrict__ u,
                                                               const double* __restrict__ v, int n,
                                                               int ld, Stencils s, WallBC b,
                                                               double rRe, double dt) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= n || j >= n) return;
    const double dwdx = sum_x(s.d1x, s.start, w, i, j, ld);
    const double dwdy = sum_y(s.d1y, s.start, w, i, j, ld);
    const double d2wdx2 = sum_x(s.d2x, s.start, w, i, j, ld);
    const double d2wdy2 = sum_y(s.d2y, s.start, w, i, j, ld);
    const double uu = bc_u(u, i, j, n, ld, b);
    const double vv = bc_v(v, i, j, n, ld, b);
    wout[(size_t)i * ld + j] = (-uu * dwdx - vv * dwdy + rRe * (d2wdx2 + d2wdy2)) * dt + w[(size_t)i * ld + j];
}

// ---- start of a Poisson solve: psi = 0 (src/poisson.c:229-230), reset the control block ----
__global__ void __launch_bounds__(kBlock) poisson_init_kernel(double* psi0, SorCtl* ctl, Diag* diag,
                                                             long long count, double tol, int itmax) {
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < count; q += stride) psi0[q] = 0.0;
    if (blockIdx.x == 0 && threadIdx.x < 2) {  // both state slots; launch 0 reads slot 1
        SorCtl* sl = ctl + threadIdx.x;
        sl->k = 0;
        sl->parity = 0;
        sl->status = itmax > 0 ? SOR_RUNNING : SOR_FAILED;
        sl->s_done = 0;
        sl->launches = 0;
        sl->itmax = itmax;
        sl->e_last = 0.0;
        sl->tol = tol;
        sl->rtol = 0.0;
        sl->n_last = 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (diag) {
            diag->kmax = dkey(-DBL_MAX);
            diag->kmin = dkey(DBL_MAX);
            diag->outer_norm = 0.0;
        }
    }
}

// ---- velocities u = DY psi, v = -(DX psi) at every point (src/main.c:366-383) --------------
__global__ void __launch_bounds__(kBlock) velocity_kernel(const double* psi0, const double* psi1,
                                                         const SorCtl* fin, double* __restrict__ u,
                                                         double* __restrict__ v, int n, int ld,
                                                         Stencils s) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= n || j >= n) return;
    const double* psi = fin->parity ? psi1 : psi0;
    u[(size_t)i * ld + j] = sum_y(s.d1y, s.start, psi, i, j, ld);
    v[(size_t)i * ld + j] = -sum_x(s.d1x, s.start, psi, i, j, ld);
}

// ---- continuity diagnostic: max/min of DX u + DY v (src/main.c:387-408, fluiddyn.c:149,
//      linearalg.c:475-545; comparisons skip NaN exactly as "if (A > max)") ------------------
__global__ void __launch_bounds__(kBlock) continuity_kernel(const double* __restrict__ u,
                                                           const double* __restrict__ v, Diag* diag,
                                                           int n, int ld, Stencils s) {
    double mx = -DBL_MAX, mn = DBL_MAX;
    const long long count = (long long)n * n;
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < count; q += stride) {
        const int i = (int)(q / n), j = (int)(q - (long long)i * n);
        const double c = sum_x(s.d1x, s.start, u, i, j, ld) + sum_y(s.d1y, s.start, v, i, j, ld);
        if (c > mx) mx = c;
        if (c < mn) mn = c;
    }
    unsigned long long kx = dkey(mx), kn = dkey(mn);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long ox = __shfl_xor(kx, off, 64), on = __shfl_xor(kn, off, 64);
        kx = ox > kx ? ox : kx;
        kn = on < kn ? on : kn;
    }
    __shared__ unsigned long long sx[kBlock / 64], sn[kBlock / 64];
    if ((threadIdx.x & 63) == 0) { sx[threadIdx.x >> 6] = kx; sn[threadIdx.x >> 6] = kn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        kx = sx[0];
        kn = sn[0];
        for (int w = 1; w < kBlock / 64; ++w) { kx = sx[w] > kx ? sx[w] : kx; kn = sn[w] < kn ? sn[w] : kn; }
        atomicMax(&diag->kmax, kx);
        atomicMin(&diag->kmin, kn);
    }
}

// ---- legacy elementwise entry points -------------------------------------------------------
__global__ void __launch_bounds__(kBlock) euler_kernel(double* w, const double* dwdx, const double* dwdy,
                                                      const double* d2x, const double* d2y,
                                                      const double* u, const double* v, double rRe,
                                                      double dt, long long count) {
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q < count) w[q] = (-u[q] * dwdx[q] - v[q] * dwdy[q] + rRe * (d2x[q] + d2y[q])) * dt + w[q];
}

__global__ void __launch_bounds__(kBlock) add_kernel(const double* a, const double* b, double* out,
                                                    long long count, int sub) {
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q < count) out[q] = sub ? b[q] - a[q] : a[q] + b[q];
}

__global__ void __launch_bounds__(kBlock) copy_kernel(double* dst, const double* src, long long count) {
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q < count) dst[q] = src[q];
}

// ---- host wrappers -------------------------------------------------------------------------
static inline int launched() { return hipGetLastError() == hipSuccess ? 0 : 1; }
static inline unsigned blocks_for(long long count) { return (unsigned)((count + kBlock - 1) / kBlock); }

int sor_fp64_launch(const SorArgs<double>& a, TileShape shape, bool gs, hipStream_t st) {
    return sor_dispatch<double>(a, shape, gs, st);
}

bool sor_shape_supported(TileShape sh) {
#define CNV_SOR_OK(s_, nw_, rb_) if (sh.s == s_ && sh.nw == nw_ && sh.rb == rb_) return true;
    CNV_SOR_SHAPES(CNV_SOR_OK)
#undef CNV_SOR_OK
    return false;
}

int wall_vorticity_launch(double* w, const double* u, const double* v, int n, int ld, const Stencils& s,
                          const WallBC& bc, hipStream_t st) {
    hipLaunchKernelGGL(wall_vorticity_kernel, dim3(blocks_for(4LL * n)), dim3(kBlock), 0, st, w, u, v, n, ld, s, bc);
    return launched();
}

int transport_fp64_launch(const double* w_in, double* w_out, const double* u, const double* v, int n, int ld,
                          const Stencils& s, const WallBC& bc, double rRe, double dt, hipStream_t st) {
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    hipLaunchKernelGGL(transport_fp64_kernel, grid, dim3(kBlock), 0, st, w_in, w_out, u, v, n, ld, s, bc, rRe, dt);
    return launched();
}

int poisson_init_launch(double* psi0, SorCtl* ctl, Diag* diag, int n, int ld, double tol, int itmax,
                        hipStream_t st) {
    const long long count = (long long)n * ld;
    unsigned nb = blocks_for(count);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(poisson_init_kernel, dim3(nb), dim3(kBlock), 0, st, psi0, ctl, diag, count, tol, itmax);
    return launched();
}

int velocity_launch(const double* psi0, const double* psi1, const SorCtl* fin, double* u, double* v, int n,
                    int ld, const Stencils& s, hipStream_t st) {
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    hipLaunchKernelGGL(velocity_kernel, grid, dim3(kBlock), 0, st, psi0, psi1, fin, u, v, n, ld, s);
    return launched();
}

int continuity_launch(const double* u, const double* v, Diag* diag, int n, int ld, const Stencils& s,
                      hipStream_t st) {
    unsigned nb = blocks_for((long long)n * n);
    if (nb > 512) nb = 512;
    hipLaunchKernelGGL(continuity_kernel, dim3(nb), dim3(kBlock), 0, st, u, v, diag, n, ld, s);
    return launched();
}

int euler_launch(double* w, const double* dwdx, const double* dwdy, const double* d2x, const double* d2y,
                 const double* u, const double* v, double rRe, double dt, long long count, hipStream_t st) {
    hipLaunchKernelGGL(euler_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, st, w, dwdx, dwdy, d2x, d2y, u, v,
                       rRe, dt, count);
    return launched();
}

int add_launch(const double* a, const double* b, double* out, long long count, int sub, hipStream_t st) {
    hipLaunchKernelGGL(add_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, st, a, b, out, count, sub);
    return launched();
}

int copy_launch(double* dst, const double* src, long long count, hipStream_t st) {
    hipLaunchKernelGGL(copy_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, st, dst, src, count);
    return launched();
}

}  // namespace cnv
