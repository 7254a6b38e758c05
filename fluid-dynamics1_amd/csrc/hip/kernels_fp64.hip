// kernels_fp64.hip -- plain FP64 kernels of the time step, bit-faithful to the reference.
//
// Compiled with -ffp-contract=off (and the pragma below): every a*b+c stays two roundings, as in
// the reference's x86-64 build, so that the stencil sums, the Euler update and the SOR update
// reproduce the reference binary64 results bit for bit.
// The per-point bodies live in step_kernels.h (shared with the batched kernels).
#pragma clang fp contract(off)

#include <float.h>

#include "cnv_device.h"
#include "sor_rb.h"
#include "step_kernels.h"

namespace cnv {

__global__ void __launch_bounds__(kBlock) wall_vorticity_kernel(double* w, const double* u, const double* v,
                                                               int n, int ld, Stencils s, WallBC b) {
    wall_vorticity_at(w, u, v, n, ld, s, b, blockIdx.x * kBlock + threadIdx.x);
}

__global__ void __launch_bounds__(kBlock) transport_fp64_kernel(const double* __restrict__ w,
                                                               double* __restrict__ wout,
                                                               const double* __restrict__ u,
                                                               const double* __restrict__ v, int n,
                                                               int ld, Stencils s, WallBC b,
                                                               double rRe, double dt) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i < n && j < n) transport_fp64_at(w, wout, u, v, n, ld, s, b, rRe, dt, i, j);
}

// ---- start of a Poisson solve: psi = 0 (src/poisson.c:229-230), reset the control block ----
__global__ void __launch_bounds__(kBlock) poisson_init_kernel(double* psi0, SorCtl* ctl, Diag* diag,
                                                             long long count, double tol, int itmax,
                                                             int zero) {
    const long long stride = (long long)gridDim.x * kBlock;
    if (zero)
        for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < count; q += stride) psi0[q] = 0.0;
    if (blockIdx.x == 0 && threadIdx.x < 2) ctl_reset(ctl + threadIdx.x, tol, 0.0, itmax);
    if (blockIdx.x == 0 && threadIdx.x == 0 && diag) diag_reset(diag);
}

// warm start (opt-in): the previous solve's final iterate, psi[fin->parity], becomes psi0, the
// first iterate of the next solve (read before the slots are reset by the init kernel)
__global__ void __launch_bounds__(kBlock) warm_start_kernel(double* psi0, Planes psi, const SorCtl* fin,
                                                           long long count) {
    const double* src = psi.p[fin->parity];
    if (src == psi0) return;  // already in psi0
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < count; q += stride) psi0[q] = src[q];
}

// ---- velocities u = DY psi, v = -(DX psi) at every point (src/main.c:366-383) --------------
__global__ void __launch_bounds__(kBlock) velocity_kernel(Planes psi, const SorCtl* fin, double* __restrict__ u,
                                                         double* __restrict__ v, int n, int ld, Stencils s) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= n || j >= n) return;
    velocity_at(psi.p[fin->parity], u, v, ld, s, i, j);
}

// ---- continuity diagnostic: max/min of DX u + DY v (src/main.c:387-408, fluiddyn.c:149,
//      linearalg.c:475-545; comparisons skip NaN exactly as "if (A > max)"), and the largest
//      |u| or |v| (the blow-up check) --------------------------------------------------------
__global__ void __launch_bounds__(kBlock) continuity_kernel(const double* __restrict__ u,
                                                           const double* __restrict__ v, Diag* diag,
                                                           int n, int ld, Stencils s) {
    double mx = -DBL_MAX, mn = DBL_MAX;
    unsigned long long ks = dkey(0.0);
    const long long count = (long long)n * n;
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < count; q += stride) {
        const int i = (int)(q / n), j = (int)(q - (long long)i * n);
        const double c = continuity_at(u, v, ld, s, i, j);
        if (c > mx) mx = c;
        if (c < mn) mn = c;
        const size_t p = (size_t)i * ld + j;
        const unsigned long long kq = speed_key(u[p], v[p]);
        ks = kq > ks ? kq : ks;
    }
    diag_fold_block(mx, mn, ks, diag);
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

// State snapshot for output: dst = [w | psi | u | v], each `count` doubles, with psi read from the
// buffer that holds the solve's final iterate (fin->parity), or from psi0 when psi_fixed (the
// refined form accumulates psi there).
__global__ void __launch_bounds__(kBlock) snapshot_kernel(const double* __restrict__ w, Planes planes,
                                                         const double* __restrict__ u,
                                                         const double* __restrict__ v, const SorCtl* fin,
                                                         double* __restrict__ dst, long long count) {
    const double* __restrict__ psi = planes.p[fin->parity];
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q < count) {
        dst[q] = w[q];
        dst[count + q] = psi[q];
        dst[2 * count + q] = u[q];
        dst[3 * count + q] = v[q];
    }
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
                        hipStream_t st, int zero) {
    const long long count = (long long)n * ld;
    unsigned nb = blocks_for(count);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(poisson_init_kernel, dim3(nb), dim3(kBlock), 0, st, psi0, ctl, diag, count, tol, itmax,
                       zero);
    return launched();
}

int warm_start_launch(double* psi0, const Planes& psi, const SorCtl* fin, int n, int ld, hipStream_t st) {
    const long long count = (long long)n * ld;
    unsigned nb = blocks_for(count);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(warm_start_kernel, dim3(nb), dim3(kBlock), 0, st, psi0, psi, fin, count);
    return launched();
}

int velocity_launch(const Planes& psi, const SorCtl* fin, double* u, double* v, int n, int ld, const Stencils& s,
                    hipStream_t st) {
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    hipLaunchKernelGGL(velocity_kernel, grid, dim3(kBlock), 0, st, psi, fin, u, v, n, ld, s);
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

int snapshot_launch(const double* w, const Planes& psi, const double* u, const double* v, const SorCtl* fin,
                    double* dst, long long count, hipStream_t st) {
    hipLaunchKernelGGL(snapshot_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, st, w, psi, u, v, fin, dst,
                       count);
    return launched();
}

}  // namespace cnv
