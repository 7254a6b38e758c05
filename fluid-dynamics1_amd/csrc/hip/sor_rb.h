// sor_rb.h -- red-black SOR sweeps for the 5-point Poisson stencil, register-blocked and
// temporally blocked.
//
// One launch advances the whole grid by up to S full sweeps (red half-sweep, then black),
// computing exactly what S successive global sweeps compute, point for point:
//   * each workgroup owns a TI x TJ tile and loads it with a halo of H = 2S cells (one cell per
//     half-sweep) into registers; half-sweep h updates the points of its colour inside the
//     extended tile shrunk by h+1 on every side, whose four neighbours were all updated
//     correctly by half-sweep h-1; after 2S half-sweeps the owned tile is exact and is written to
//     the other buffer (ping-pong);
//   * the extended tile is 128 columns wide: lane m of every wave holds columns 2m and 2m+1,
//     and wave w holds RB consecutive rows.  A point's N/S neighbours are in the same lane's
//     registers (rows r+-1), one of its E/W neighbours is the lane's other column and the other
//     comes from lane m+-1 through a DPP wave shift; only the first and last row of each wave
//     are exchanged through LDS, once per half-sweep;
//   * every point of a row is computed and those outside the valid rectangle or on the domain
//     boundary keep their value (select), so the sweep is straight-line code.
// Point update (src/poisson.c:268 SOR, src/poisson.c:205 Gauss-Seidel), same association:
//   u = beta*(dy*dy*(N+S) + dx*dx*(E+W) - dx*dx*dy*dy*f) / (2*(dx*dx+dy*dy)) + (1-beta)*u0
// where u0 (the value before this sweep) is the point's current value in either colouring.
// In FP64 the division by the constant B = 2*(dx*dx+dy*dy) is q0 = x*RN(1/B),
// q = fma(fma(-q0, B, x), RN(1/B), q0): Markstein's correction step, which returns the
// correctly rounded quotient RN(x/B) -- the same double as x/B -- for finite x.
// The L1 norm of each sweep's update (src/poisson.c:34-60; sqrt(pow(d,2)) == |d|) is summed
// over owned points per workgroup in a fixed order into partial[L&1][s][tile].  The next launch
// decides on them in its prologue, overlapped with its own (speculative) tile load: every
// workgroup reduces the partials in tile order, so all reach the same decision, and workgroup 0
// records it in the state slot L&1 (SorCtl) that the host polls:
//   * a launch whose previous state is terminal copies it forward and returns;
//   * if sweep s < S-1 of the previous launch reached e < tol, that launch's output is
//     discarded: this launch redoes s+1 sweeps from its (untouched) input buffer, so the final
//     iterate is exactly the one of the first sweep with e < tol (src/poisson.c:273);
//   * the sweep budget itmax is honoured exactly (src/poisson.c:234);
//   * defect correction (FP32 planes): the solve also stops ("fold") once the update norm fell
//     below rtol times the iterate's own L1 norm, the resolution limit of FP32.
#pragma once

#include <type_traits>

#include "cnv_device.h"

namespace cnv {

template <typename T>
__device__ inline T wave_sum(T x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

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

__device__ inline double sor_div(double x, double B, double rB) {
    const double q0 = x * rB;
    const double r = __fma_rn(-q0, B, x);
    return __fma_rn(r, rB, q0);
}
// FP32 planes (defect correction): no bit-parity requirement, one multiply by RN(1/B)
__device__ inline float sor_div(float x, float, float rB) { return x * rB; }

// Lane i receives lane i+1's value (wave_shl:1) / lane i-1's value (wave_shr:1).
template <int CTRL>
__device__ inline double dpp_shift(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ inline float dpp_shift(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}
constexpr int kDppFromNext = 0x130;  // wave_shl:1
constexpr int kDppFromPrev = 0x138;  // wave_shr:1

template <typename T, int S, int NW, int RB, bool GS>
__global__ void __launch_bounds__(NW * 64) sor_rb_kernel(SorArgs<T> a) {
    constexpr int H = 2 * S;
    constexpr int EJ = 128;          // extended tile columns; lane m holds columns 2m, 2m+1
    constexpr int TJ = EJ - 2 * H;
    constexpr int EI = NW * RB;      // extended tile rows; wave w holds rows [w*RB, w*RB+RB)
    constexpr int TI = EI - 2 * H;
    static_assert(TI > 0 && TJ > 0, "halo larger than the tile");
    static_assert(RB % 2 == 0, "even row blocks keep each wave's colour pattern static");
    typedef typename std::conditional<sizeof(T) == 4, float, double>::type Acc;

    __shared__ T xch[2][NW][2][64];  // [half-sweep parity][wave][first/last row][lane]
    __shared__ double red[S][NW][2];
    __shared__ double lred[S][2][NW * 64];  // previous launch's partials, one column per thread
    __shared__ double etot[kMaxSweepsPerLaunch][2];

    const int L = a.launch;
    const SorCtl* prev = a.ctl + ((L + 1) & 1);
    SorCtl* mine = a.ctl + (L & 1);
    const int tid = threadIdx.x;
    const int pstatus = prev->status;
    if (pstatus != SOR_RUNNING) {  // the solve already ended: carry its state forward
        if (blockIdx.x == 0 && tid == 0) *mine = *prev;
        return;
    }
    const int sd = prev->s_done;   // sweeps the previous launch performed (0: first launch)
    const int k_prev = prev->k;
    const int itmax = prev->itmax;
    const double tol = prev->tol, rtol = prev->rtol;

    const int ntiles = a.tiles_i * a.tiles_j;
    const int tile = sor_tile_of_block(blockIdx.x, ntiles);
    const int ti = tile / a.tiles_j;
    const int tj = tile - ti * a.tiles_j;
    const int gi0 = ti * TI - H;  // even, so the colour of local (li, lj) is (li + lj) & 1
    const int gj0 = tj * TJ - H;
    const int nr = a.nr, nc = a.nc, ld = a.ld;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform (SGPR)
    const int row0 = wv * RB;        // even: the colour of (row0 + r, 2m + q) is (r + q) & 1

    // the previous launch's per-tile partials: thread t sums tiles t, t + NT, ... in that order
    const size_t pstride = (size_t)kMaxSweepsPerLaunch * ntiles * 2;
    double pv[S][2];
#pragma unroll
    for (int s = 0; s < S; ++s) pv[s][0] = pv[s][1] = 0.0;
    if (sd > 0) {
        const double* pp = a.partial + ((L + 1) & 1) * pstride;
        for (int t = tid; t < ntiles; t += NW * 64) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (s < sd) {
                    pv[s][0] += pp[((size_t)s * ntiles + t) * 2];
                    pv[s][1] += pp[((size_t)s * ntiles + t) * 2 + 1];
                }
            }
        }
    }

    // speculative tile load for the common case (sweep on from the previous launch's output)
    const T* __restrict__ srcN = (L & 1) ? a.psi1 : a.psi0;
    // column facts: q = 0 -> column 2m ("A"), q = 1 -> column 2m+1 ("B").  colh[q]: the point is
    // updated in half-sweep h iff colh[q] > h (distance to the extended tile's side, or -1 on
    // the domain boundary and outside the domain).
    int colh[2];
    bool ownc[2];
    int gjq[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int lj = 2 * lane + q;
        gjq[q] = gj0 + lj;
        const int dj = lj < EJ - 1 - lj ? lj : EJ - 1 - lj;
        colh[q] = (gjq[q] >= 1 && gjq[q] <= nc - 2) ? dj : -1;
        ownc[q] = lj >= H && lj < H + TJ && gjq[q] < nc;
    }

    T v[2][RB];   // v[q][r]: value at (row0 + r, 2*lane + q)
    T fc[2][RB];  // fc[q][r]: c3 * f at the same point (the hoisted c3*f product of the update)
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int gi = gi0 + row0 + r;
        const bool rin = (unsigned)gi < (unsigned)nr;
        const bool rint = gi >= 1 && gi <= nr - 2;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int gj = gjq[q];
            T val = T(0), fv = T(0);
            if (rin && (unsigned)gj < (unsigned)nc) val = srcN[(size_t)gi * ld + gj];
            if (rint && colh[q] >= 0) {
                const T sv = a.src_f[(size_t)gi * ld + gj];
                fv = a.negate ? -sv : sv;
            }
            v[q][r] = val;
            fc[q][r] = a.c3 * fv;
        }
    }
    // ---- decision on the previous launch: every workgroup reduces its per-tile partials in
    //      the same fixed order and reaches the same decision (src/poisson.c:272-278) --------
    if (sd > 0) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (s < sd) {
                lred[s][0][tid] = pv[s][0];
                lred[s][1][tid] = pv[s][1];
            }
        }
        __syncthreads();
        for (int row = wv; row < 2 * sd; row += NW) {  // wave w reduces (sweep, quantity) rows
            const int s = row >> 1, q = row & 1;
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < NW; ++j) acc += lred[s][q][lane + 64 * j];
            acc = wave_sum(acc);
            if (lane == 0) etot[s][q] = acc;
        }
        __syncthreads();
    }
    int mode = 0;          // 0: sweep on, 1: stop (terminal), 2: redo s_run sweeps (rollback)
    int s_run = 0;
    SorCtl next;
    next.k = k_prev;
    next.parity = L & 1;   // where the previous launch left its output
    next.status = SOR_RUNNING;
    next.itmax = itmax;
    next.tol = tol;
    next.rtol = rtol;
    next.launches = prev->launches;
    next.e_last = prev->e_last;
    next.n_last = prev->n_last;
    if (sd > 0) {
        int s_conv = -1, s_fold = -1;
        for (int s = 0; s < sd; ++s) {
            if (etot[s][0] < tol) {
                s_conv = s;
                break;
            }
            if (rtol > 0.0 && etot[s][0] < rtol * etot[s][1]) {
                s_fold = s;
                break;
            }
        }
        if (s_conv >= 0 && s_conv < sd - 1) {
            // converged inside the previous launch: redo s_conv+1 sweeps from its (untouched)
            // input buffer; that iterate is the final one (src/poisson.c:273)
            mode = 2;
            s_run = s_conv + 1;
            next.status = SOR_CONVERGED;
            next.k = k_prev + s_run;
            next.e_last = etot[s_conv][0];
            next.n_last = etot[s_conv][1];
        } else {
            next.k = k_prev + sd;
            next.e_last = etot[sd - 1][0];
            next.n_last = etot[sd - 1][1];
            if (s_conv >= 0) {
                next.status = SOR_CONVERGED;
                mode = 1;
            } else if (s_fold >= 0) {
                next.status = SOR_FOLD;
                mode = 1;
            } else if (next.k >= itmax) {
                next.status = SOR_FAILED;
                mode = 1;
            }
        }
    } else if (itmax <= k_prev) {
        next.status = SOR_FAILED;
        mode = 1;
    }
    if (mode == 0) s_run = (itmax - next.k) < S ? (itmax - next.k) : S;
    next.s_done = (mode == 0) ? s_run : 0;
    if (mode != 1) next.launches += 1;
    if (blockIdx.x == 0 && tid == 0) *mine = next;
    if (mode == 1) return;
    const int s_eff = s_run;
    // normal: read the previous launch's output, write the other buffer; rollback: read the
    // previous launch's input, write over its output
    T* __restrict__ dst = ((L & 1) ^ (mode == 2 ? 0 : 1)) ? a.psi1 : a.psi0;
    if (mode == 2) {
        const T* __restrict__ srcR = (L & 1) ? a.psi0 : a.psi1;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int gi = gi0 + row0 + r;
#pragma unroll
            for (int q = 0; q < 2; ++q)
                v[q][r] = ((unsigned)gi < (unsigned)nr && (unsigned)gjq[q] < (unsigned)nc)
                              ? srcR[(size_t)gi * ld + gjq[q]] : T(0);
        }
    }

    // half-sweep 0 updates colour 0 and reads its neighbours' colour-1 boundary rows:
    // colour 1 sits in column q = (r + 1) & 1 of row r
    xch[1][wv][0][lane] = v[1][0];
    xch[1][wv][1][lane] = v[(RB - 1 + 1) & 1][RB - 1];
    __syncthreads();

    const T c1 = a.c1, c2 = a.c2, B = a.B, rB = a.rB, beta = a.beta, omb = a.omb;
    for (int s = 0; s < s_eff; ++s) {
        Acc acc[2] = {Acc(0), Acc(0)}, nrm[2] = {Acc(0), Acc(0)};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int h = 2 * s + c;
            // boundary rows of the neighbouring waves (colour 1-c, written by half-sweep h-1)
            const T sb = xch[c ^ 1][wv > 0 ? wv - 1 : 0][1][lane];
            const T nb = xch[c ^ 1][wv < NW - 1 ? wv + 1 : NW - 1][0][lane];
            T nv[RB];
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                const int q = (c + r) & 1;  // column of colour c in row r (compile-time)
                const int li = row0 + r;    // wave-uniform from here on
                const int gi = gi0 + li;
                const int rowd = li < EI - 1 - li ? li : EI - 1 - li;
                const bool rowok = rowd > h && gi >= 1 && gi <= nr - 2;
                const T old = v[q][r];
                const T up = (r < RB - 1) ? v[q][r + 1] : nb;
                const T dn = (r > 0) ? v[q][r - 1] : sb;
                // E + W: the lane's other column plus the neighbouring lane's nearer column
                const T own = v[q ^ 1][r];
                const T far = q ? dpp_shift<kDppFromNext>(v[0][r]) : dpp_shift<kDppFromPrev>(v[1][r]);
                const T rt = q ? far : own;
                const T lt = q ? own : far;
                const T x = c1 * (up + dn) + c2 * (rt + lt) - fc[q][r];
                const T cand = GS ? sor_div(x, B, rB) : sor_div(beta * x, B, rB) + omb * old;
                nv[r] = (rowok && colh[q] > h) ? cand : old;
                if (li >= H && li < H + TI) {  // owned row (uniform branch); columns masked below
                    acc[q] += (Acc)fabs(nv[r] - old);
                    if (sizeof(T) == 4) nrm[q] += (Acc)fabs(nv[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < RB; ++r) v[(c + r) & 1][r] = nv[r];
            xch[c][wv][0][lane] = v[c & 1][0];
            xch[c][wv][1][lane] = v[(c + RB - 1) & 1][RB - 1];
            __syncthreads();
        }
        const Acc eacc = (ownc[0] ? acc[0] : Acc(0)) + (ownc[1] ? acc[1] : Acc(0));
        const Acc nacc = (ownc[0] ? nrm[0] : Acc(0)) + (ownc[1] ? nrm[1] : Acc(0));
        const double we = wave_sum((double)eacc);
        const double wn = sizeof(T) == 4 ? wave_sum((double)nacc) : 0.0;
        if (lane == 0) {
            red[s][wv][0] = we;
            red[s][wv][1] = wn;
        }
    }

#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int li = row0 + r;
        const int gi = gi0 + li;
        if (li >= H && li < H + TI && gi < nr) {
#pragma unroll
            for (int q = 0; q < 2; ++q)
                if (ownc[q]) dst[(size_t)gi * ld + gjq[q]] = v[q][r];
        }
    }
    __syncthreads();
    if (mode == 0 && tid < 2 * s_eff) {  // a rollback launch ends the solve: no partials needed
        const int s = tid >> 1, q = tid & 1;
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) t += red[s][w][q];
        a.partial[(L & 1) * pstride + ((size_t)s * ntiles + tile) * 2 + q] = t;
    }
}

template <typename T, int S, int NW, int RB>
inline int sor_launch_shape(const SorArgs<T>& a, bool gs, hipStream_t st) {
    const int ntiles = a.tiles_i * a.tiles_j;
    if (gs)
        hipLaunchKernelGGL((sor_rb_kernel<T, S, NW, RB, true>), dim3(ntiles), dim3(NW * 64), 0, st, a);
    else
        hipLaunchKernelGGL((sor_rb_kernel<T, S, NW, RB, false>), dim3(ntiles), dim3(NW * 64), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// The launch shapes compiled in, TileShape {s, nw, rb}: S sweeps per launch, nw waves of rb rows;
// the tile owns (nw*rb - 4s) rows x (128 - 4s) columns.
#define CNV_SOR_SHAPES(X) \
    X(1, 4, 10)           \
    X(2, 4, 10)           \
    X(4, 4, 12)           \
    X(4, 8, 8)            \
    X(6, 8, 8)            \
    X(8, 8, 10)           \
    X(8, 8, 12)

template <typename T>
inline int sor_dispatch(const SorArgs<T>& a, TileShape sh, bool gs, hipStream_t st) {
#define CNV_SOR_CASE(s_, nw_, rb_) \
    if (sh.s == s_ && sh.nw == nw_ && sh.rb == rb_) return sor_launch_shape<T, s_, nw_, rb_>(a, gs, st);
    CNV_SOR_SHAPES(CNV_SOR_CASE)
#undef CNV_SOR_CASE
    return 2;
}

}  // namespace cnv
