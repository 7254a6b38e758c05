// step_kernels.h -- per-point bodies of the time-step kernels, shared by the one-cavity kernels
// (kernels_fp64.hip, kernels_dc.hip) and the batched kernels (batch.hip).
//
// Each body states its floating-point contraction rule itself, so it compiles to the same
// arithmetic in every file that includes it:
//   * FP64 bodies: contract(off) -- every a*b+c stays two roundings, as in the reference's x86-64
//     build, which makes the stencil sums and the Euler update bit-faithful;
//   * defect-correction bodies (FP32 transport, FP64 residual): contract(fast), so a*b+c may
//     become one fused multiply-add.
//
// Stencil sums: the reference multiplies dense Kronecker matrices by the flattened field
// (src/main.c:142-152, src/linearalg.c:236-285): y[r] = 0.0 + A[r][0]*x[0] + A[r][1]*x[1] + ...
// Adding the +-0 products of structural zeros never changes a sum, so the ascending sum over a
// row's 7-column window (zero-padded, starting from 0.0) is the same number.
#pragma once

#include <float.h>

#include "cnv_device.h"

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
#pragma clang fp contract(off)
    const double* c = coef + (size_t)j * kStencilWidth;
    const double* row = f + (size_t)i * ld + start[j];
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m) s += c[m] * row[m];
    return s;
}
// (d (x) I) f at (i, j): sum over k of d[i][k] f[k][j]   (acts along i, i.e. y)
__device__ inline double sum_y(const double* coef, const int* start, const double* f, int i, int j, int ld) {
#pragma clang fp contract(off)
    const double* c = coef + (size_t)i * kStencilWidth;
    const double* col = f + (size_t)start[i] * ld + j;
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m) s += c[m] * col[(size_t)m * ld];
    return s;
}

// ---- wall vorticity: w[edge] = (Dx v - Dy u)[edge] on BC-applied u, v (src/main.c:298-320);
//      t enumerates the 4n-4 edge points --------------------------------------------------
__device__ inline void wall_vorticity_at(double* w, const double* u, const double* v, int n, int ld,
                                         const Stencils& s, const WallBC& b, int t) {
#pragma clang fp contract(off)
    int i, j;
    if (t < n) { i = 0; j = t; }
    else if (t < 2 * n) { i = n - 1; j = t - n; }
    else if (t < 3 * n - 2) { i = t - 2 * n + 1; j = 0; }
    else if (t < 4 * n - 4) { i = t - (3 * n - 2) + 1; j = n - 1; }
    else return;
    const double* c1x = s.d1x + (size_t)j * kStencilWidth;
    const double* c1y = s.d1y + (size_t)i * kStencilWidth;
    const int sj = s.start[j], si = s.start[i];
    // the same values as bc_u / bc_v, chosen among copies of the wall velocities (selecting the
    // struct's fields by reference in this loop made the compiler spill the struct to scratch);
    // every stencil point lies inside the grid, so its interior value is always loadable
    const double u1 = b.u1, u2 = b.u2, u3 = b.u3, u4 = b.u4;
    const double v1 = b.v1, v2 = b.v2, v3 = b.v3, v4 = b.v4;
    auto pick = [n](double in, int pi, int pj, double w1, double w2, double w3, double w4) {
        double r = pi == n - 1 ? w4 : in;
        r = pi == 0 ? w3 : r;
        r = pj == n - 1 ? w2 : r;
        return pj == 0 ? w1 : r;
    };
    double dvdx = 0.0, dudy = 0.0;
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m)
        dvdx += c1x[m] * pick(v[(size_t)i * ld + sj + m], i, sj + m, v1, v2, v3, v4);
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m)
        dudy += c1y[m] * pick(u[(size_t)(si + m) * ld + j], si + m, j, u1, u2, u3, u4);
    w[(size_t)i * ld + j] = dvdx - dudy;
}

// ---- derivatives of w and the explicit Euler update at (i, j) (src/main.c:323-344,
//      src/fluiddyn.c:98):  w' = (-u*dwdx - v*dwdy + (1./Re)*(d2wdx2+d2wdy2))*dt + w ---------
__device__ inline void transport_fp64_at(const double* __restrict__ w, double* __restrict__ wout,
                                         const double* __restrict__ u, const double* __restrict__ v,
                                         int n, int ld, const Stencils& s, const WallBC& b, double rRe,
                                         double dt, int i, int j) {
#pragma clang fp contract(off)
    const double dwdx = sum_x(s.d1x, s.start, w, i, j, ld);
    const double dwdy = sum_y(s.d1y, s.start, w, i, j, ld);
    const double d2wdx2 = sum_x(s.d2x, s.start, w, i, j, ld);
    const double d2wdy2 = sum_y(s.d2y, s.start, w, i, j, ld);
    const double uu = bc_u(u, i, j, n, ld, b);
    const double vv = bc_v(v, i, j, n, ld, b);
    wout[(size_t)i * ld + j] = (-uu * dwdx - vv * dwdy + rRe * (d2wdx2 + d2wdy2)) * dt + w[(size_t)i * ld + j];
}

// ---- one state slot of a fresh solve (launch 0 reads slot 1; both are written) -------------
__device__ inline void ctl_reset(SorCtl* sl, double tol, double rtol, int itmax) {
    sl->k = 0;
    sl->parity = 0;
    sl->status = itmax > 0 ? SOR_RUNNING : SOR_FAILED;
    sl->s_done = 0;
    sl->launches = 0;
    sl->itmax = itmax;
    sl->rb_src = 0;
    sl->rb_n = 0;
    sl->e_last = 0.0;
    sl->tol = tol;
    sl->rtol = rtol;
    sl->n_last = 0.0;
}
__device__ inline void diag_reset(Diag* d) {
    d->kmax = dkey(-DBL_MAX);
    d->kmin = dkey(DBL_MAX);
    d->kspd = dkey(0.0);
}

// ---- velocities u = DY psi, v = -(DX psi) at (i, j) (src/main.c:366-383) -------------------
__device__ inline void velocity_at(const double* psi, double* __restrict__ u, double* __restrict__ v,
                                   int ld, const Stencils& s, int i, int j) {
#pragma clang fp contract(off)
    u[(size_t)i * ld + j] = sum_y(s.d1y, s.start, psi, i, j, ld);
    v[(size_t)i * ld + j] = -sum_x(s.d1x, s.start, psi, i, j, ld);
}

// ---- continuity Dx u + Dy v at (i, j) (src/main.c:387-408, src/fluiddyn.c:149) -------------
__device__ inline double continuity_at(const double* u, const double* v, int ld, const Stencils& s,
                                       int i, int j) {
#pragma clang fp contract(off)
    return sum_x(s.d1x, s.start, u, i, j, ld) + sum_y(s.d1y, s.start, v, i, j, ld);
}

// The speed key of a point: the larger key of |u| and |v|.  fabs clears the sign of a NaN, whose
// key then orders above +infinity's, so a NaN velocity wins every max.
__device__ inline unsigned long long speed_key(double uu, double vv) {
    const unsigned long long a = dkey(fabs(uu)), b = dkey(fabs(vv));
    return a > b ? a : b;
}

// Block-wide max/min of per-thread extrema as order-preserving keys, folded into *diag with
// atomics (exact and order independent), with the largest speed key.  kBlock threads.
__device__ inline void diag_fold_block(double mx, double mn, unsigned long long ks, Diag* diag) {
    unsigned long long kx = dkey(mx), kn = dkey(mn);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long ox = __shfl_xor(kx, off, 64), on = __shfl_xor(kn, off, 64);
        const unsigned long long os = __shfl_xor(ks, off, 64);
        kx = ox > kx ? ox : kx;
        kn = on < kn ? on : kn;
        ks = os > ks ? os : ks;
    }
    __shared__ unsigned long long sx[kBlock / 64], sn[kBlock / 64], ss[kBlock / 64];
    if ((threadIdx.x & 63) == 0) {
        sx[threadIdx.x >> 6] = kx;
        sn[threadIdx.x >> 6] = kn;
        ss[threadIdx.x >> 6] = ks;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        kx = sx[0];
        kn = sn[0];
        ks = ss[0];
        for (int w = 1; w < kBlock / 64; ++w) {
            kx = sx[w] > kx ? sx[w] : kx;
            kn = sn[w] < kn ? sn[w] : kn;
            ks = ss[w] > ks ? ss[w] : ks;
        }
        atomicMax(&diag->kmax, kx);
        atomicMin(&diag->kmin, kn);
        atomicMax(&diag->kspd, ks);
    }
}

// ---- defect correction: r32 = (float)(f - L psi) on the interior, f = -w; 0 on the edges ----
__device__ inline void dc_residual_at(const double* __restrict__ w, const double* __restrict__ psi,
                                      float* __restrict__ r32, int n, int ld, double c1, double c2,
                                      double c3, double B, int i, int j) {
#pragma clang fp contract(fast)
    const size_t q = (size_t)i * ld + j;
    float r = 0.0f;
    if (i >= 1 && i <= n - 2 && j >= 1 && j <= n - 2) {
        const double lap = (c1 * (psi[q + ld] + psi[q - ld]) + c2 * (psi[q + 1] + psi[q - 1]) - B * psi[q]) / c3;
        r = (float)(-w[q] - lap);
    }
    r32[q] = r;
}

// ---- FP32 transport stencils (defect-correction form) ---------------------------------------
__device__ inline float bcf(const double* a, int i, int j, int n, int ld, double e0, double e1,
                            double e2, double e3) {
    if (j == 0) return (float)e0;
    if (j == n - 1) return (float)e1;
    if (i == 0) return (float)e2;
    if (i == n - 1) return (float)e3;
    return (float)a[(size_t)i * ld + j];
}
// Each window's coefficients sum to zero (a derivative of a constant is 0), so the stencil of f
// equals the stencil of f - f(i, j): the taps enter as FP64 differences from the point's own value,
// rounded to FP32, and the sums run in FP32.  Rounding f itself to FP32 would leave an error of
// FP32 epsilon times |f| / h in every derivative (the differences cancel the leading digits), which
// accumulates step after step; rounding the differences leaves FP32 epsilon times the local change.
__device__ inline float fsum_x(const float* coef, const int* start, const double* f, int i, int j, int ld) {
#pragma clang fp contract(fast)
    const float* c = coef + (size_t)j * kStencilWidth;
    const double* row = f + (size_t)i * ld + start[j];
    const double fc = f[(size_t)i * ld + j];
    float s = 0.0f;
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m) s += c[m] * (float)(row[m] - fc);
    return s;
}
__device__ inline float fsum_y(const float* coef, const int* start, const double* f, int i, int j, int ld) {
#pragma clang fp contract(fast)
    const float* c = coef + (size_t)i * kStencilWidth;
    const double* col = f + (size_t)start[i] * ld + j;
    const double fc = f[(size_t)i * ld + j];
    float s = 0.0f;
#pragma unroll
    for (int m = 0; m < kStencilWidth; ++m) s += c[m] * (float)(col[(size_t)m * ld] - fc);
    return s;
}
// velocities of the refined form, u = DY psi, v = -(DX psi), with the stencils in FP32 (taps as
// FP64 differences from the point's own psi, so the result keeps FP32 precision relative to the
// derivative, not to |psi| / h)
__device__ inline void velocity_fp32_at(const double* psi, double* __restrict__ u, double* __restrict__ v,
                                        int ld, const Stencils& s, int i, int j) {
    u[(size_t)i * ld + j] = (double)fsum_y(s.f1y, s.start, psi, i, j, ld);
    v[(size_t)i * ld + j] = -(double)fsum_x(s.f1x, s.start, psi, i, j, ld);
}
// omega update: derivatives and right-hand side in FP32, the state update w + dt*rhs in FP64
__device__ inline void transport_fp32_at(const double* __restrict__ w, double* __restrict__ wout,
                                         const double* __restrict__ u, const double* __restrict__ v,
                                         int n, int ld, const Stencils& s, const WallBC& b, double rRe,
                                         double dt, int i, int j) {
#pragma clang fp contract(fast)
    const float dwdx = fsum_x(s.f1x, s.start, w, i, j, ld);
    const float dwdy = fsum_y(s.f1y, s.start, w, i, j, ld);
    const float d2 = fsum_x(s.f2x, s.start, w, i, j, ld) + fsum_y(s.f2y, s.start, w, i, j, ld);
    const float uu = bcf(u, i, j, n, ld, b.u1, b.u2, b.u3, b.u4);
    const float vv = bcf(v, i, j, n, ld, b.v1, b.v2, b.v3, b.v4);
    const float rhs = -uu * dwdx - vv * dwdy + (float)rRe * d2;
    const size_t q = (size_t)i * ld + j;
    wout[q] = w[q] + (double)rhs * dt;
}

}  // namespace cnv
