// kernels_dc.hip -- the FP32 + FP64 defect-correction form of the step.
//
// Poisson solve as defect correction around the same SOR iteration as the plain path:
//   psi64 = 0
//   repeat:  r   = f - L psi64                         (FP64 residual, rounded to FP32)
//            e   = SOR_fp32(L e = r) from e = 0         (same red-black kernel on FP32 planes)
//            psi64 += e                                 (FP64 accumulation)
//   until an inner sweep's L1 update fell below tol (the plain solver's own stopping rule).
// L is the operator whose fixed point the reference SOR iterates to (src/poisson.c:268):
//   L psi = (dy*dy*(N+S) + dx*dx*(E+W) - 2*(dx*dx+dy*dy)*psi) / (dx*dx*dy*dy).
// SOR is a one-step stationary iteration, so SOR on L e = r from e = 0 produces exactly the
// iterates that SOR on psi would produce from psi64: the inner sweeps continue the plain
// solver's iteration and their update norm is its stopping quantity.  An inner solve ends early
// ("fold") once its update norm fell below inner_rtol times the L1 norm of e, where FP32 stops
// resolving the update; the correction is then folded into psi64 and the residual refreshed in
// FP64, which restores full resolution.  The residual, the psi accumulation, the omega update and
// the velocities (small differences of a large psi) are FP64.  The transport stencils are FP64 by
// default (transport_fp32 = 0): their FP32 rounding, added to omega step after step, makes the
// refined trajectory drift from the FP64 one; transport_fp32 = 1 runs them in FP32
// (transport_fp32_at) for short runs.
#include <float.h>

#include "cnv_device.h"
#include "sor_rb.h"
#include "step_kernels.h"

namespace cnv {

int sor_fp32_launch(const SorArgs<float>& a, TileShape shape, bool gs, hipStream_t st) {
    return sor_dispatch<float>(a, shape, gs, st);
}

// r32 = (float)(f - L psi) on the interior, f = -w; 0 on the edges.
__global__ void __launch_bounds__(kBlock) dc_residual_kernel(const double* __restrict__ w,
                                                            const double* __restrict__ psi,
                                                            float* __restrict__ r32, int n, int ld,
                                                            double c1, double c2, double c3,
                                                            double B) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i < n && j < n) dc_residual_at(w, psi, r32, n, ld, c1, c2, c3, B, i, j);
}

// psi64 += e32 (the inner solve's final iterate, chosen by the control block's parity)
__global__ void __launch_bounds__(kBlock) dc_correct_kernel(double* __restrict__ psi, const float* e0,
                                                           const float* e1, const SorCtl* fin,
                                                           long long count) {
    const float* e = fin->parity ? e1 : e0;
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q < count) psi[q] += (double)e[q];
}

// start of an inner solve: e = 0, control block reset with the absolute tolerance (the plain
// solver's stopping rule) and the FP32 fold ratio
__global__ void __launch_bounds__(kBlock) dc_inner_init_kernel(float* e0, SorCtl* ctl, long long count,
                                                              double rtol, double tol, int itmax,
                                                              const int* budget) {
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < count; q += stride) e0[q] = 0.0f;
    // the sweep budget comes from the device counter when there is one (captured step)
    if (blockIdx.x == 0 && threadIdx.x < 2) ctl_reset(ctl + threadIdx.x, tol, rtol, budget ? *budget : itmax);
}

// device sweep budget of a refined solve: fin == nullptr resets it to itmax, otherwise the inner
// solve that ended in slot fin spent fin->k sweeps of it
__global__ void dc_budget_kernel(int* budget, int itmax, const SorCtl* fin) {
    if (threadIdx.x == 0) *budget = fin ? *budget - fin->k : itmax;
}

__global__ void __launch_bounds__(kBlock) transport_fp32_kernel(const double* __restrict__ w,
                                                               double* __restrict__ wout,
                                                               const double* __restrict__ u,
                                                               const double* __restrict__ v, int n,
                                                               int ld, Stencils s, WallBC b,
                                                               double rRe, double dt) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i < n && j < n) transport_fp32_at(w, wout, u, v, n, ld, s, b, rRe, dt, i, j);
}

// refined-form velocities from the FP64 accumulator with FP32 stencils (velocity_fp32_at)
__global__ void __launch_bounds__(kBlock) velocity_fp32_kernel(const double* __restrict__ psi, double* __restrict__ u,
                                                              double* __restrict__ v, int n, int ld, Stencils s) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i < n && j < n) velocity_fp32_at(psi, u, v, ld, s, i, j);
}

static inline int launched() { return hipGetLastError() == hipSuccess ? 0 : 1; }
static inline unsigned blocks_for(long long count) { return (unsigned)((count + kBlock - 1) / kBlock); }

int velocity_fp32_launch(const double* psi, double* u, double* v, int n, int ld, const Stencils& s,
                         hipStream_t st) {
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    hipLaunchKernelGGL(velocity_fp32_kernel, grid, dim3(kBlock), 0, st, psi, u, v, n, ld, s);
    return launched();
}

int dc_residual_launch(const double* w, const double* psi, float* r32, int n, int ld, double c1,
                       double c2, double c3, double B, hipStream_t st) {
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    hipLaunchKernelGGL(dc_residual_kernel, grid, dim3(kBlock), 0, st, w, psi, r32, n, ld, c1, c2, c3, B);
    return launched();
}

int dc_correct_launch(double* psi, const float* e0, const float* e1, const SorCtl* fin, int n, int ld,
                      hipStream_t st) {
    const long long count = (long long)n * ld;
    hipLaunchKernelGGL(dc_correct_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, st, psi, e0, e1, fin, count);
    return launched();
}

int dc_inner_init_launch(float* e0, SorCtl* ctl, int n, int ld, double rtol, double tol, int itmax,
                         hipStream_t st, const int* budget) {
    const long long count = (long long)n * ld;
    unsigned nb = blocks_for(count);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(dc_inner_init_kernel, dim3(nb), dim3(kBlock), 0, st, e0, ctl, count, rtol, tol, itmax,
                       budget);
    return launched();
}

int dc_budget_launch(int* budget, int itmax, const SorCtl* fin, hipStream_t st) {
    hipLaunchKernelGGL(dc_budget_kernel, dim3(1), dim3(64), 0, st, budget, itmax, fin);
    return launched();
}

int transport_fp32_launch(const double* w_in, double* w_out, const double* u, const double* v, int n, int ld,
                          const Stencils& s, const WallBC& bc, double rRe, double dt, hipStream_t st) {
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    hipLaunchKernelGGL(transport_fp32_kernel, grid, dim3(kBlock), 0, st, w_in, w_out, u, v, n, ld, s, bc, rRe, dt);
    return launched();
}

}  // namespace cnv
