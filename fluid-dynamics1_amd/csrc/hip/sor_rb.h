// sor_rb.h -- red-black SOR sweeps for the 5-point Poisson stencil, temporally blocked in LDS.
//
// One launch advances the whole grid by up to S full sweeps (red half-sweep, then black),
// computing exactly what S successive global sweeps compute, point for point:
//   * each workgroup owns a TI x TJ tile and loads it with a halo of H = 2S cells (one cell per
//     half-sweep) from the source buffer into LDS;
//   * half-sweep h updates the points of its colour inside the tile shrunk by h+1 on every side,
//     whose four neighbours were all updated correctly by half-sweep h-1; after 2S half-sweeps
//     the owned tile is exact, and it is written to the other buffer (ping-pong);
//   * LDS stores each row as two colour planes (red-black split), so a wave reads its 4
//     neighbours as contiguous runs: N/S at the same index of the other plane in rows i+-1,
//     E/W at index m+p and m+p-1 of the other plane in the same row, p = (colour + row) & 1.
// Point update (src/poisson.c:268, SOR; src/poisson.c:205, Gauss-Seidel), same association:
//   u = beta*(dy*dy*(N+S) + dx*dx*(E+W) - dx*dx*dy*dy*f) / (2*(dx*dx+dy*dy)) + (1-beta)*u0
// where u0, the value before this sweep, is the point's current value in either colouring.
// The L1 norm of each sweep's update (src/poisson.c:34-60; sqrt(pow(d,2)) == |d|) is summed
// over owned points per workgroup in a fixed order and written to partial[s][tile]; the last
// workgroup to finish (atomic ticket) reduces them in tile order, so the sum is deterministic.
//
// Convergence control lives on the device (SorCtl) so that launches can be queued blindly:
//   * a launch reads the status first and returns at once after convergence or failure;
//   * if sweep s < S-1 of a launch reaches e < tol, the launch's output buffer is discarded and
//     the next launch redoes s+1 sweeps from the unchanged source buffer (ROLLBACK), so the
//     final iterate is exactly the one of the first sweep with e < tol (src/poisson.c:273);
//   * the sweep budget itmax is honoured exactly (src/poisson.c:234).
#pragma once

#include "cnv_device.h"

namespace cnv {

__device__ inline int sor_tile_of_block(int b, int ntiles) {
    // Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  Give each XCD a
    // contiguous band of tiles so that the halo rows two neighbouring tiles share are fetched
    // into the same L2.
    if ((ntiles & 7) == 0) {
        const int per = ntiles >> 3;
        return (b & 7) * per + (b >> 3);
    }
    return b;
}

template <typename T>
__device__ inline T wave_sum(T x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

template <typename T, int TI, int TJ, int S, bool GS>
__global__ void __launch_bounds__(kBlock) sor_rb_kernel(SorArgs<T> a) {
    constexpr int H = 2 * S;
    constexpr int EI = TI + 2 * H;
    constexpr int EJ = TJ + 2 * H;
    constexpr int EJ2 = EJ / 2;
    constexpr int NPC = EI * EJ2;  // points per colour in the extended tile
    constexpr int R = (NPC + kBlock - 1) / kBlock;
    constexpr int NW = kBlock / 64;
    static_assert(TI % 2 == 0 && TJ % 2 == 0, "even tiles keep the colour parity tile-local");

    __shared__ T lds[EI * 2 * EJ2];
    __shared__ double red[S][NW];
    __shared__ double red2[NW];
    __shared__ double etot[S];
    __shared__ int last_flag;

    SorCtl* ctl = a.ctl;
    const int status = ctl->status;
    if (status == SOR_CONVERGED || status == SOR_FAILED) return;
    const int k0 = ctl->k;
    const int budget = ctl->itmax - k0;
    const int s_eff = (status == SOR_ROLLBACK) ? ctl->s_rerun : (budget < S ? budget : S);
    const int parity = ctl->parity;
    const T* __restrict__ src = parity ? a.psi1 : a.psi0;
    T* __restrict__ dst = parity ? a.psi0 : a.psi1;

    const int ntiles = a.tiles_i * a.tiles_j;
    const int tile = sor_tile_of_block(blockIdx.x, ntiles);
    const int ti = tile / a.tiles_j;
    const int tj = tile - ti * a.tiles_j;
    const int gi0 = ti * TI - H;  // even: colour of (li, lj) is (li + lj) & 1
    const int gj0 = tj * TJ - H;
    const int nr = a.nr, nc = a.nc;
    const int ld = a.ld;
    const int tid = threadIdx.x;

    for (int e = tid; e < EI * EJ; e += kBlock) {
        const int li = e / EJ, lj = e - li * EJ;
        const int gi = gi0 + li, gj = gj0 + lj;
        T val = T(0);
        if ((unsigned)gi < (unsigned)nr && (unsigned)gj < (unsigned)nc) val = src[(size_t)gi * ld + gj];
        lds[(li * 2 + ((li + lj) & 1)) * EJ2 + (lj >> 1)] = val;
    }

    T freg[2][R];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int q = tid + r * kBlock;
            const int li = q / EJ2, m = q - li * EJ2;
            const int lj = 2 * m + ((c + li) & 1);
            const int gi = gi0 + li, gj = gj0 + lj;
            T fv = T(0);
            if (q < NPC && gi >= 1 && gi <= nr - 2 && gj >= 1 && gj <= nc - 2) {
                const T s = a.src_f[(size_t)gi * ld + gj];
                fv = a.negate ? -s : s;
            }
            freg[c][r] = fv;
        }
    }
    __syncthreads();

    for (int s = 0; s < s_eff; ++s) {
        double eacc = 0.0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int h = 2 * s + c;
            const int oc = 1 - c;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int q = tid + r * kBlock;
                const int li = q / EJ2, m = q - li * EJ2;
                const int p = (c + li) & 1;
                const int lj = 2 * m + p;
                const int gi = gi0 + li, gj = gj0 + lj;
                const bool upd = q < NPC && li > h && li < EI - 1 - h && lj > h && lj < EJ - 1 - h &&
                                 gi >= 1 && gi <= nr - 2 && gj >= 1 && gj <= nc - 2;
                if (upd) {
                    const T up = lds[((li + 1) * 2 + oc) * EJ2 + m];
                    const T dn = lds[((li - 1) * 2 + oc) * EJ2 + m];
                    const T rt = lds[(li * 2 + oc) * EJ2 + m + p];
                    const T lt = lds[(li * 2 + oc) * EJ2 + m + p - 1];
                    T* self = &lds[(li * 2 + c) * EJ2 + m];
                    const T old = *self;
                    const T x = a.c1 * (up + dn) + a.c2 * (rt + lt) - a.c3 * freg[c][r];
                    const T nv = GS ? x / a.B : (a.beta * x) / a.B + a.omb * old;
                    *self = nv;
                    if (li >= H && li < H + TI && lj >= H && lj < H + TJ) eacc += (double)fabs(nv - old);
                }
            }
            __syncthreads();
        }
        const double ws = wave_sum(eacc);
        if ((tid & 63) == 0) red[s][tid >> 6] = ws;
    }

    for (int e = tid; e < TI * TJ; e += kBlock) {
        const int oi = e / TJ, oj = e - oi * TJ;
        const int li = oi + H, lj = oj + H;
        const int gi = gi0 + li, gj = gj0 + lj;
        if (gi < nr && gj < nc)
            dst[(size_t)gi * ld + gj] = lds[(li * 2 + ((li + lj) & 1)) * EJ2 + (lj >> 1)];
    }
    __syncthreads();
    if (tid < s_eff) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) t += red[tid][w];
        a.partial[(size_t)tid * ntiles + tile] = t;
    }

    // ---- last workgroup: reduce the per-tile updates, decide, reset the ticket ------------
    __threadfence();
    __syncthreads();
    if (tid == 0) last_flag = (atomicAdd(&ctl->ticket, 1u) == (unsigned)(ntiles - 1));
    __syncthreads();
    if (!last_flag) return;
    __threadfence();
    for (int s = 0; s < s_eff; ++s) {
        double acc = 0.0;
        for (int t = tid; t < ntiles; t += kBlock) acc += a.partial[(size_t)s * ntiles + t];
        acc = wave_sum(acc);
        if ((tid & 63) == 0) red2[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            double e = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) e += red2[w];
            etot[s] = e;
        }
        __syncthreads();
    }
    if (tid == 0) {
        double thr = ctl->tol;
        int s_conv = -1;
        for (int s = 0; s < s_eff; ++s) {
            const double e = etot[s];
            if (ctl->rel_first && k0 + s == 0) {  // defect-correction inner solve
                thr = ctl->rtol * e;
                if (thr < ctl->tol_floor) thr = ctl->tol_floor;
                ctl->tol = thr;
                ctl->rel_first = 0;
            }
            if (e < thr) {
                s_conv = s;
                break;
            }
        }
        if (s_conv >= 0 && s_conv < s_eff - 1) {
            ctl->status = SOR_ROLLBACK;
            ctl->s_rerun = s_conv + 1;
        } else {
            ctl->k = k0 + s_eff;
            ctl->parity = parity ^ 1;
            ctl->e_last = etot[s_eff - 1];
            if (s_conv >= 0)
                ctl->status = SOR_CONVERGED;
            else
                ctl->status = (k0 + s_eff >= ctl->itmax) ? SOR_FAILED : SOR_RUNNING;
        }
        ctl->launches += 1;
        __threadfence();
        ctl->ticket = 0u;
    }
}

template <typename T, int TI, int TJ, int S>
inline int sor_launch_shape(const SorArgs<T>& a, bool gs, hipStream_t st) {
    const int ntiles = a.tiles_i * a.tiles_j;
    if (gs)
        hipLaunchKernelGGL((sor_rb_kernel<T, TI, TJ, S, true>), dim3(ntiles), dim3(kBlock), 0, st, a);
    else
        hipLaunchKernelGGL((sor_rb_kernel<T, TI, TJ, S, false>), dim3(ntiles), dim3(kBlock), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// The tile shapes compiled in; TileShape {ti, tj, s}.
#define CNV_SOR_SHAPES(X) \
    X(32, 64, 1)          \
    X(32, 64, 2)          \
    X(32, 64, 4)          \
    X(64, 64, 4)          \
    X(32, 64, 8)          \
    X(64, 64, 8)

template <typename T>
inline int sor_dispatch(const SorArgs<T>& a, TileShape sh, bool gs, hipStream_t st) {
#define CNV_SOR_CASE(ti_, tj_, s_) \
    if (sh.ti == ti_ && sh.tj == tj_ && sh.s == s_) return sor_launch_shape<T, ti_, tj_, s_>(a, gs, st);
    CNV_SOR_SHAPES(CNV_SOR_CASE)
#undef CNV_SOR_CASE
    return 2;
}

}  // namespace cnv
