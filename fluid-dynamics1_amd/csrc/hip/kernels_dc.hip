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
// FP64, which restores full resolution.  Only the residual, the psi accumulation and the omega
// update are FP64; the transport and velocity stencils run in FP32.
#include <float.h>

#include "cnv_device.h"
#include "sor_rb.h"

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
    if (i >= n || j >= n) return;
    const size_t q = (size_t)i * ld + j;
    float r = 0.0f;
    if (i >= 1 && i <= n - 2 && j >= 1 && j <= n - 2) {
        const double lap = (c1 * (psi[q + ld] + psi[q - ld]) + c2 * (psi[q + 1] + psi[q - 1]) - B * psi[q]) / c3;
        r = (float)(-w[q] - lap);
    }
    r32[q] = r;
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
                                                              double rtol, double tol, int itmax) {
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long q = (long long)blockIdx.x * kBlock + threadIdx.x; q < count; q += stride) e0[q] = 0.0f;
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
        sl->rtol = rtol;
        sl->n_last = 0.0;
    }
}

__global__ void __launch_bounds__(kBlock) zero_f32_kernel(float* p, long long count) {
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q < count) p[q] = 0.0f;
}

// ---- FP32 stencils ---------------------------------------------------------------------------
__device__ inline float bcf(const double* a, int i, int j, int n, int ld, double e0, double e1,
                            double e2, double e3) {
    if (j == 0) return (float)e0;
    if (j == n - 1) return (float)e1;
    if (i == 0) return (float)e2;
    if (i == n - 1) return (float)e3;
    return (float)a[(size_t)i * ld + j];
}

__device__ inline float fsum_x(const float* coef, const int* start, const double* f, int i, int j, int ld) {
    const float* c = coef + (size_t)j * kStencilWidth;
    const double* row = f + (size_t)i * ld + start[j];
    float s = 0.0f;
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m) s += c[m] * (float)row[m];
    return s;
}
__device__ inline float fsum_y(const float* coef, const int* start, const double* f, int i, int j, int ld) {
    const float* c = coef + (size_t)i * kStencilWidth;
    const double* col = f + (size_t)start[i] * ld + j;
    float s = 0.0f;
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m) s += c[m] * (float)col[(size_t)m * ld];
    return s;
}

// omega update: derivatives and right-hand side in FP32, the state update w + dt*rhs in FP64
__global__ void __launch_bounds__(kBlock) transport_fp32_kernel(const double* __restrict__ w,
                                                               double* __restrict__ wout,
                                                               const double* __restrict__ u,
                                                               const double* __restrict__ v, int n,
                                                               int ld, Stencils s, WallBC b,
                                                               double rRe, double dt) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= n || j >= n) return;
    const float dwdx = fsum_x(s.f1x, s.start, w, i, j, ld);
    const float dwdy = fsum_y(s.f1y, s.start, w, i, j, ld);
    const float d2 = fsum_x(s.f2x, s.start, w, i, j, ld) + fsum_y(s.f2y, s.start, w, i, j, ld);
    const float uu = bcf(u, i, j, n, ld, b.u1, b.u2, b.u3, b.u4);
    const float vv = bcf(v, i, j, n, ld, b.v1, b.v2, b.v3, b.v4);
    const float rhs = -uu * dwdx - vv * dwdy + (float)rRe * d2;
    const size_t q = (size_t)i * ld + j;
    wout[q] = w[q] + (double)rhs * dt;
}

// u = DY psi, v = -(DX psi) with FP32 stencil arithmetic on the FP64 psi
__global__ void __launch_bounds__(kBlock) velocity_fp32_kernel(const double* __restrict__ psi,
                                                              double* __restrict__ u,
                                                              double* __restrict__ v, int n, int ld,
                                                              Stencils s) {
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= n || j >= n) return;
    u[(size_t)i * ld + j] = (double)fsum_y(s.f1y, s.start, psi, i, j, ld);
    v[(size_t)i * ld + j] = -(double)fsum_x(s.f1x, s.start, psi, i, j, ld);
}

static inline int launched() { return hipGetLastError() == hipSuccess ? 0 : 1; }
static inline unsigned blocks_for(long long count) { return (unsigned)((count + kBlock - 1) / kBlock); }

int dc_residual_launch(const double* w, const double* psi, float* r32, Diag*, int n, int ld, double c1,
                       double c2, double c3, double B, double, hipStream_t st) {
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
                         hipStream_t st) {
    const long long count = (long long)n * ld;
    unsigned nb = blocks_for(count);
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(dc_inner_init_kernel, dim3(nb), dim3(kBlock), 0, st, e0, ctl, count, rtol, tol, itmax);
    return launched();
}

int zero_f32_launch(float* p, long long count, hipStream_t st) {
    hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, st, p, count);
    return launched();
}

int transport_fp32_launch(const double* w_in, double* w_out, const double* u, const double* v, int n, int ld,
                          const Stencils& s, const WallBC& bc, double rRe, double dt, hipStream_t st) {
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    hipLaunchKernelGGL(transport_fp32_kernel, grid, dim3(kBlock), 0, st, w_in, w_out, u, v, n, ld, s, bc, rRe, dt);
    return launched();
}

int velocity_fp32_launch(const double* psi, double* u, double* v, int n, int ld, const Stencils& s,
                         hipStream_t st) {
    dim3 grid((n + 63) / 64, (n + 3) / 4);
    hipLaunchKernelGGL(velocity_fp32_kernel, grid, dim3(kBlock), 0, st, psi, u, v, n, ld, s);
    return launched();
}

}  // namespace cnv
