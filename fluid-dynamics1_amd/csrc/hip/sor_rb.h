// sor_rb.h -- red-black SOR sweeps for the 5-point Poisson stencil, register-blocked and
// temporally blocked.
//
// One launch advances the whole grid by up to S full sweeps (red half-sweep, then black),
// computing exactly what S successive global sweeps compute, point for point:
//   * each workgroup owns a TI x TJ tile and loads it with a halo of H = 2S cells (one cell per
//     half-sweep) into registers; after half-sweep h the points of the extended tile at distance
//     > h from its edge hold exactly what h+1 global half-sweeps produce (their four neighbours
//     were exact after half-sweep h-1); after 2S half-sweeps the owned tile is exact and is
//     written to the other buffer (ping-pong);
//   * the extended tile is 128 columns wide: lane m of every wave holds columns 2m and 2m+1,
//     and wave w holds RB consecutive rows.  A point's N/S neighbours are in the same lane's
//     registers (rows r+-1), one of its E/W neighbours is the lane's other column and the other
//     comes from lane m+-1 through a DPP wave shift; only the first and last row of each wave
//     are exchanged through LDS, once per half-sweep;
//   * every point of a row is computed, so the sweep is straight-line code with no selects.  A
//     point outside the valid rectangle takes a wrong (finite) value, which only points outside
//     the valid rectangle ever read.  Domain-boundary points must keep psi = 0: their relaxation
//     factor is 0 instead of beta, which leaves them at +-0, and the final store writes +0 there.
//     Points outside the domain have factor 0 too and start from a copy of the nearest data
//     (the tile loads are clamped into the plane): finite, and read only by the boundary and by
//     other points outside;
//   * the L1 norms are accumulated per lane with a 0/1 weight per row (owned rows only) and a
//     per-column ownership mask, written to LDS once per sweep and reduced once per launch
//     (FP32 planes: in the launch's last sweep only).
// Point update (src/poisson.c:268 SOR, src/poisson.c:205 Gauss-Seidel), same association:
//   u = beta*(dy*dy*(N+S) + dx*dx*(E+W) - dx*dx*dy*dy*f) / (2*(dx*dx+dy*dy)) + (1-beta)*u0
// where u0 (the value before this sweep) is the point's current value in either colouring.
// In FP64 the division by the constant B = 2*(dx*dx+dy*dy) is q0 = x*RN(1/B),
// q = fma(fma(-q0, B, x), RN(1/B), q0): Markstein's correction step, which returns the
// correctly rounded quotient RN(x/B) -- the same double as x/B -- for finite x.  When B is a
// power of two (dx = dy = 2^-k), 1/B is exact and x*(1/B) is already RN(x/B): one multiply (XR).
// The L1 norm of each sweep's update (src/poisson.c:34-60; sqrt(pow(d,2)) == |d|) is summed
// over owned points per workgroup in a fixed order into partial[L&1][s][tile] (FP32 planes: the
// last sweep's only, into s = 0).  The next launch decides on them in one extra workgroup beside
// its sweeping ones (the decider), so that neither the partials' latency nor the reduction delays
// any sweep: the decider of launch L+1 reduces launch L's partials in a fixed order and writes
// state slot (L+1)&1 (SorCtl), which launch L+2's workgroups and the host read; the sweepers of
// launch L+1 read slot L&1 only, complete before they started, and sweep speculatively:
//   * a launch whose slot is terminal returns after its tile load (its decider copies the slot
//     forward);
//   * if launch L's sweeps ended the solve, launch L+1's speculative sweeps are discarded;
//   * if sweep s < S-1 of launch L reached e < tol, launch L's output is discarded too: launch
//     L+2 redoes s+1 sweeps from launch L's input plane, which FP64 keeps intact by rotating
//     through three planes, so the final iterate is exactly the one of the first sweep with
//     e < tol (src/poisson.c:273); FP32 planes, a preconditioner, test the last sweep of a
//     launch only, and only of the launches with (L & test_mask) == test_mask (the others sum
//     nothing, leave no partials, and their decider carries the state forward), and rotate
//     through two;
//   * the sweep budget itmax is honoured exactly (src/poisson.c:234);
//   * a non-finite L1 update (or iterate norm) ends the solve with SOR_DIVERGED;
//   * relative rule: with rtol > 0 the solve also stops ("fold") once the update norm of a sweep
//     fell below rtol times the iterate's own L1 norm after that launch's last sweep (measured
//     once per launch) -- the resolution limit of an FP32 defect-correction inner solve, or the
//     relative tolerance of the adjoint problem's FP64 solves (NORM).
#pragma once

#include <float.h>

#include <type_traits>

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

__device__ inline double sor_div(double x, double B, double rB) {
    const double q0 = x * rB;
    const double r = __fma_rn(-q0, B, x);
    return __fma_rn(r, rB, q0);
}
// FP32 planes (defect correction): no bit-parity requirement, one multiply by RN(1/B)
__device__ inline float sor_div(float x, float, float rB) { return x * rB; }

// Two consecutive elements of a row, aligned to one element: rows of any pitch are accessed in
// pairs, one multi-dword access each (the hardware needs dword alignment only).
template <typename T> struct PairOf;
template <> struct PairOf<double> { typedef double type __attribute__((ext_vector_type(2), aligned(8))); };
template <> struct PairOf<float> { typedef float type __attribute__((ext_vector_type(2), aligned(4))); };

// The pair at byte offset boff of a row: the row pointer is wave-uniform and the lane's offset a
// 32-bit value, so the address is a scalar base plus a vector offset with no 64-bit lane arithmetic.
template <typename T>
__device__ inline typename PairOf<T>::type load_pair(const T* __restrict__ row, unsigned boff) {
    return *reinterpret_cast<const typename PairOf<T>::type*>(reinterpret_cast<const char*>(row) + boff);
}
// The pair (x0, x1) at byte offset boff of a row when both, else x0 alone.  FP32 tiles are stored
// non-temporally, FP64 tiles with ordinary stores: each measured 0.8% (FP32) and 2.5% (FP64)
// faster per 1024^2 step than the other kind (the FP64 tile read back by the next launch then
// comes from nearer).
template <typename T>
__device__ inline void store_pair(T* __restrict__ row, unsigned boff, bool both, T x0, T x1) {
    typedef typename PairOf<T>::type P;
    constexpr bool kNT = sizeof(T) == 4;
    T* p = reinterpret_cast<T*>(reinterpret_cast<char*>(row) + boff);
    if (both) {
        P d;
        d.x = x0;
        d.y = x1;
        if (kNT) __builtin_nontemporal_store(d, reinterpret_cast<P*>(p));
        else *reinterpret_cast<P*>(p) = d;
    } else {
        if (kNT) __builtin_nontemporal_store(x0, p);
        else *p = x0;
    }
}

// Lane i receives the value of the lane the DPP control CTRL names, e.g. lane i+1's
// (wave_shl:1) or lane i-1's (wave_shr:1); a lane without a source (63 / 0 there) reads 0
// (bound_ctrl), so no old value has to be materialised.
template <int CTRL>
__device__ inline double dpp_shift(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ inline float dpp_shift(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, true));
}
constexpr int kDppFromNext = 0x130;  // wave_shl:1
constexpr int kDppFromPrev = 0x138;  // wave_shr:1

// One step of a wave reduction of N values: x[i] += the lane value that CTRL selects.
template <int CTRL, int N>
__device__ inline void dpp_accumulate(double (&x)[N]) {
    double y[N];
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = dpp_shift<CTRL>(x[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] += y[i];
}
// Sums over the 64 lanes of a wave of N independent values, returned to every lane, in a fixed
// association (the same in every workgroup): pairs and quads (quad_perm [1,0,3,2] and
// [2,3,0,1]), the four quads of each 16-lane row (row_shr 4 and 8: lane 15 of a row holds the
// row's sum), then the rows (row_bcast 15: lane 63 gets rows 3 + 2, lane 31 rows 1 + 0;
// row_bcast 31: lane 63 gets all four); lane 63's value is broadcast.  Cross-lane moves only, no
// LDS round trips; each step is applied to all N values before the next, so that their chains
// interleave (no idle cycles between a value's update and its next cross-lane read).
template <int N>
__device__ inline void wave_sum_n(double (&x)[N]) {
    dpp_accumulate<0xb1>(x);
    dpp_accumulate<0x4e>(x);
    dpp_accumulate<0x114>(x);
    dpp_accumulate<0x118>(x);
    dpp_accumulate<0x142>(x);
    dpp_accumulate<0x143>(x);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const int lo = __builtin_amdgcn_readlane(__double2loint(x[i]), 63);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(x[i]), 63);
        x[i] = __hiloint2double(hi, lo);
    }
}
__device__ inline double wave_sum(double x) {
    double v[1] = {x};
    wave_sum_n(v);
    return v[0];
}

// Plane idx of a solve's iterate planes (FP64: 0, 1, 2; FP32: 0, 1 -- psi2 is not used).
template <typename T>
__device__ inline T* sor_plane(const SorArgs<T>& a, int idx) {
    if constexpr (SorPlanes<T>::value == 2) return (idx & 1) ? a.psi1 : a.psi0;
    else return idx == 0 ? a.psi0 : (idx == 1 ? a.psi1 : a.psi2);
}

// One sweeping workgroup's launch on tile `tile` of the solve described by `a` (NW*64 threads).
// It reads the state slot the previous launch's decider wrote: RUNNING -- sweep on from plane
// L % NB (the previous launch's output) into plane (L+1) % NB, as many sweeps as the budget
// leaves, and leave this tile's per-sweep L1 sums for the next launch's decider; ROLLBACK -- the
// solve converged inside the launch before last: redo rb_n sweeps from plane rb_src into plane
// `parity`, which then holds the final iterate; anything else -- the solve has ended: return.
template <typename T, int S, int NW, int RB, bool GS, bool XR, bool NORM>
__device__ __forceinline__ void sor_rb_sweeper(const SorArgs<T>& a, const int L, const int tile) {
#pragma clang fp contract(off)
    constexpr int NB = SorPlanes<T>::value;
    constexpr int H = 2 * S;
    // the iterate's L1 norm is needed by the relative stopping rule (rtol): always on FP32
    // planes, on request (NORM) for FP64.  It is measured once per launch, after its last sweep.
    constexpr bool kNorm = sizeof(T) == 4 || NORM;
    // FP32 planes (the inner solves of defect correction) are a preconditioner with no bit-parity
    // requirement: their stopping rule is applied to a launch's last sweep only (and only on the
    // launches the test mask selects), so the other sweeps skip the L1 update (a solve runs at
    // most (test_mask + 1) * S - 1 sweeps past the first sweep that met the rule).  FP64 applies
    // it to every sweep, as src/poisson.c:272-278 does.
    constexpr bool kLastOnly = sizeof(T) == 4;
    constexpr int SE = kLastOnly ? 1 : S;  // L1 update sums per launch
    constexpr int EJ = 128;          // extended tile columns; lane m holds columns 2m, 2m+1
    constexpr int TJ = EJ - 2 * H;
    constexpr int EI = NW * RB;      // extended tile rows; wave w holds rows [w*RB, w*RB+RB)
    constexpr int TI = EI - 2 * H;
    static_assert(TI > 0 && TJ > 0, "halo larger than the tile");
    static_assert(RB % 2 == 0, "even row blocks keep each wave's colour pattern static");
    typedef typename std::conditional<sizeof(T) == 4, float, double>::type Acc;

    __shared__ T xch[2][NW][2][64];  // [half-sweep parity][wave][first/last row][lane]
    // one column per thread: this launch's per-lane L1 update sums, one row per summed sweep,
    // and the per-lane iterate norm after the last sweep
    __shared__ double lred[SE][NW * 64];
    __shared__ double lnrm[kNorm ? NW * 64 : 1];

    const SorCtl* prev = a.ctl + ((L + 1) & 1);
    const int tid = threadIdx.x;
    const int ntiles = a.tiles_i * a.tiles_j;
    const int ti = tile / a.tiles_j;
    const int tj = tile - ti * a.tiles_j;
    const int gi0 = ti * TI - H;  // even, so the colour of local (li, lj) is (li + lj) & 1
    const int gj0 = tj * TJ - H;
    const int nr = a.nr, nc = a.nc;
    const long long ld = a.ld;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform (SGPR)
    const int row0 = wv * RB;        // even: the colour of (row0 + r, 2m + q) is (r + q) & 1

    // Column facts of the lane's pair (gj, gj+1), gj even: q = 0 -> column gj, q = 1 -> gj+1.
    const int gj = gj0 + 2 * lane;
    bool colin[2], ownc[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        colin[q] = gj + q >= 1 && gj + q <= nc - 2;
        ownc[q] = 2 * lane + q >= H && 2 * lane + q < H + TJ && gj + q < nc;  // ownc[1] implies ownc[0]
    }
    // Every load reads a pair inside the data: rows clamped into [0, nr), the pair's first column
    // into [0, nc-2].  So no lane needs a select on what it loads: a point outside the domain may
    // start from any finite value (it is read only by points outside the domain and by the domain
    // boundary, whose relaxation factor is 0), and each domain-boundary point is read from its own
    // position (+-0 there).  With nc odd, the lane of column nc-1 reads the pair (nc-2, nc-1) and
    // takes its second element (fixed after the loads); its column nc is outside.
    const unsigned bld = (unsigned)(gj < 0 ? 0 : (gj > nc - 2 ? nc - 2 : gj)) * sizeof(T);
    auto row_of = [&](const T* plane, int gi) -> const T* {  // wave-uniform row pointer
        return plane + (gi < 0 ? 0 : (gi >= nr ? nr - 1 : gi)) * ld;
    };

    // relaxation factor of an updated point: FP64 beta (1 for Gauss-Seidel), applied before the
    // division; FP32 beta/B (1/B), which folds the division in.  Per column, 0 off the interior
    // columns; per point, also 0 on the boundary rows and outside the domain.
    T bcoef;
    if constexpr (sizeof(T) == 8) bcoef = GS ? T(1) : a.beta;
    else bcoef = GS ? a.rB : a.beta * a.rB;
    const T c3s = a.negate ? -a.c3 : a.c3;  // f = -src folded in: c3*(-x) == (-c3)*x exactly
    T bcol[2], cncol[2], cecol[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        bcol[q] = colin[q] ? bcoef : T(0);
        cncol[q] = colin[q] ? bcoef * a.c1 : T(0);
        cecol[q] = colin[q] ? bcoef * a.c2 : T(0);
    }

    // Speculative tile load for the common case (sweep on from the previous launch's output),
    // issued before the state slot is read so that both are in flight together.
    const T* __restrict__ srcN = sor_plane(a, L % NB);
    T v[2][RB];   // v[q][r]: value at (row0 + r, 2*lane + q)
    T fc[2][RB];  // FP64: c3*f (the hoisted c3*f product of the update); FP32: -be*c3*f
    T be[2][RB];  // FP64: bcoef at interior points, 0 on the domain boundary and outside it
    T cn[2][RB];  // FP32: be*c1 and be*c2, the relaxation folded into the stencil coefficients
    T ce[2][RB];
    Acc w[RB];    // 1 for owned rows inside the domain, else 0 (wave-uniform)
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int li = row0 + r;
        const int gi = gi0 + li;
        const bool rint = gi >= 1 && gi <= nr - 2;
        w[r] = (li >= H && li < H + TI && gi < nr) ? Acc(1) : Acc(0);
        const typename PairOf<T>::type d = load_pair(row_of(srcN, gi), bld);
        const typename PairOf<T>::type f = load_pair(row_of(a.src_f, gi), bld);
        v[0][r] = d.x;
        v[1][r] = d.y;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            // a non-interior point's source is not read by the reference: 0, whatever it holds
            const bool inter = rint && colin[q];
            const T fq = q ? f.y : f.x;
            if constexpr (sizeof(T) == 8) {
                fc[q][r] = inter ? c3s * fq : T(0);
                be[q][r] = rint ? bcol[q] : T(0);
            } else {
                fc[q][r] = inter ? bcoef * -(c3s * fq) : T(0);
                cn[q][r] = rint ? cncol[q] : T(0);
                ce[q][r] = rint ? cecol[q] : T(0);
            }
        }
    }

    // what this launch does: the state after launch L-2 and the sweeps launch L-1 performed
    const int pstatus = prev->status;
    bool rollback = false;
    int s_run = 0;
    T* __restrict__ dst = nullptr;
    if (pstatus == SOR_RUNNING) {
        const int budget = prev->itmax - (prev->k + prev->s_done);
        s_run = budget < S ? budget : S;
        dst = sor_plane(a, (L + 1) % NB);
    } else if (pstatus == SOR_ROLLBACK) {
        rollback = true;
        s_run = prev->rb_n;
        dst = sor_plane(a, prev->parity);
    }
    if (s_run <= 0) return;  // the solve has ended, or its budget is spent
    // FP32: whether this launch's last sweep is tested (FP64: every sweep of every launch)
    const bool tested = sizeof(T) == 8 || (L & a.test_mask) == a.test_mask;
    if (rollback) {
        // the converged launch's input replaces the speculatively loaded tile
        const T* __restrict__ srcR = sor_plane(a, prev->rb_src);
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const typename PairOf<T>::type d = load_pair(row_of(srcR, gi0 + row0 + r), bld);
            v[0][r] = d.x;
            v[1][r] = d.y;
        }
    }
    if (nc & 1) {  // wave-uniform: the lane of column nc-1 read the pair (nc-2, nc-1)
        if (gj == nc - 1) {
#pragma unroll
            for (int r = 0; r < RB; ++r) v[0][r] = v[1][r];
        }
    }

    const T c1 = a.c1, c2 = a.c2, B = a.B, rB = a.rB, omb = a.omb;
    // half-sweep 0 updates colour 0 and reads its neighbours' colour-1 boundary rows: colour 1
    // sits in column q = (r + 1) & 1 of row r
    auto sweeps_begin = [&]() {
        xch[1][wv][0][lane] = v[1][0];
        xch[1][wv][1][lane] = v[(RB - 1 + 1) & 1][RB - 1];
        __syncthreads();
    };
    // sweep s (red half-sweep, then black) of a pass; with_norm: also sum the iterate's norm
    // (FP32: and the update, which only the last sweep sums); l1 (std::true_type or false_type):
    // sum the L1 norms at all, which a wave holding no owned row skips (its sums are 0)
    auto one_sweep = [&](const int s, const bool with_norm, auto l1) {
        constexpr bool kL1 = decltype(l1)::value;
        Acc acc[2] = {Acc(0), Acc(0)}, nrm[2] = {Acc(0), Acc(0)};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            // boundary rows of the neighbouring waves (colour 1-c, written by half-sweep h-1)
            const T sb = xch[c ^ 1][wv > 0 ? wv - 1 : 0][1][lane];
            const T nb = xch[c ^ 1][wv < NW - 1 ? wv + 1 : NW - 1][0][lane];
            // rows 1..RB-2 first: they do not wait for the LDS reads of sb / nb
#pragma unroll
            for (int rr = 0; rr < RB; ++rr) {
                const int r = rr < RB - 2 ? rr + 1 : (rr == RB - 2 ? 0 : RB - 1);
                const int q = (c + r) & 1;  // column of colour c in row r (compile-time)
                const T old = v[q][r];
                const T up = (r < RB - 1) ? v[q][r + 1] : nb;
                const T dn = (r > 0) ? v[q][r - 1] : sb;
                // E + W: the lane's other column plus the neighbouring lane's nearer column
                const T own = v[q ^ 1][r];
                const T far = q ? dpp_shift<kDppFromNext>(v[0][r]) : dpp_shift<kDppFromPrev>(v[1][r]);
                T cand;
                if constexpr (sizeof(T) == 8) {
                    // the reference's association: ((c1*(N+S) + c2*(E+W)) - c3*f), times beta,
                    // divided by B, plus (1-beta)*u0; + is commutative, so E/W order is free
                    const T x = c1 * (up + dn) + c2 * (own + far) - fc[q][r];
                    const T y = be[q][r] * x;
                    const T qt = XR ? y * rB : sor_div(y, B, rB);
                    cand = GS ? qt : qt + omb * old;
                    v[q][r] = cand;
                    if constexpr (kL1) {
                        acc[q] = __fma_rn(w[r], fabs(cand - old), acc[q]);
                        if (kNorm && with_norm) nrm[q] = __fma_rn(w[r], fabs(cand), nrm[q]);
                    }
                } else {
                    // relaxation and 1/B folded into cn, ce and fc (0 where the point is fixed)
                    const T x = __fmaf_rn(cn[q][r], up + dn, __fmaf_rn(ce[q][r], own + far, fc[q][r]));
                    cand = GS ? x : __fmaf_rn(omb, old, x);
                    v[q][r] = cand;
                    if (kL1 && with_norm) {
                        acc[q] = __fmaf_rn(w[r], fabsf(cand - old), acc[q]);
                        nrm[q] = __fmaf_rn(w[r], fabsf(cand), nrm[q]);
                    }
                }
            }
            xch[c][wv][0][lane] = v[c & 1][0];
            xch[c][wv][1][lane] = v[(c + RB - 1) & 1][RB - 1];
            __syncthreads();
        }
        // per-lane sums of this sweep, owned columns only; reduced once after the last sweep
        if (!kLastOnly || with_norm)
            lred[kLastOnly ? 0 : s][tid] = (double)((ownc[0] ? acc[0] : Acc(0)) + (ownc[1] ? acc[1] : Acc(0)));
        if (kNorm && with_norm)
            lnrm[tid] = (double)((ownc[0] ? nrm[0] : Acc(0)) + (ownc[1] ? nrm[1] : Acc(0)));
    };
    // the common full pass, unrolled: the values stay in their registers from sweep to sweep (a
    // rolled loop makes the compiler rotate them through copies at every iteration); not for the
    // FP64 shapes whose unrolled pass no longer fits the register budget of their wave count.
    // A rollback or a budget-limited last launch takes the loop, which sums the FP64 norm in
    // every sweep (the last one's stays).
    // The waves that hold only halo rows (a quarter of them at 1024^2) run it without the FP64
    // L1 sums, two of the point update's 14 instructions.
    constexpr bool kUnroll = sizeof(T) == 4 || (S * RB <= 40 && NW != 12 && !(kNorm && S * RB > 20));
    bool owns = false;  // wave-uniform: the wave holds an owned row
#pragma unroll
    for (int r = 0; r < RB; ++r) owns = owns || w[r] != Acc(0);
    sweeps_begin();
    if (kUnroll && s_run == S) {
        // FP32: an untested launch skips the last sweep's sums.  Small shapes get a pass of their
        // own without them (6.10 against 6.29 us per launch at 1024^2 with (5, 16, 4)); large ones
        // branch on the uniform flag in the last sweep, since a second unrolled pass of their
        // size made the kernel slower (the batched (8, 8, 10) FP32 sweep 4.05 -> 4.7 us)
        constexpr bool kSplit = sizeof(T) == 4 && S * RB <= 24;
        if (kSplit && !tested) {
#pragma unroll
            for (int s = 0; s < S; ++s) one_sweep(s, false, std::false_type{});
        } else if (owns || sizeof(T) == 4) {
#pragma unroll
            for (int s = 0; s < S; ++s) one_sweep(s, s == S - 1 && (kSplit || tested), std::true_type{});
        } else {
#pragma unroll
            for (int s = 0; s < S; ++s) one_sweep(s, s == S - 1, std::false_type{});
        }
    } else {
        for (int s = 0; s < s_run; ++s) one_sweep(s, kLastOnly ? tested && s == s_run - 1 : kNorm, std::true_type{});
    }

    // every wave's lane sums are in LDS; the stores below do not wait for this barrier's peers,
    // so the reducing waves start as soon as their own stores are issued
    if (!rollback && tested) __syncthreads();
    // the owned tile to its plane; domain-boundary points hold +-0: store +0
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int li = row0 + r;
        const int gi = gi0 + li;
        if (li >= H && li < H + TI && gi < nr) {  // owned row (wave-uniform)
            const bool rint = gi >= 1 && gi <= nr - 2;
            // owned columns lie inside the row: 0 <= gj < nc
            if (ownc[0])
                store_pair(dst + gi * ld, (unsigned)gj * sizeof(T), ownc[1], (rint && colin[0]) ? v[0][r] : T(0),
                           (rint && colin[1]) ? v[1][r] : T(0));
        }
    }
    if (rollback || !tested) return;  // the final iterate, or an untested launch: no partials

    // this tile's per-sweep partials (the per-lane sums in lred, the norm in lnrm) for the next
    // launch's decider: wave w reduces rows w, w + NW, ..., each in a fixed order
    const size_t pstride = (size_t)kMaxSweepsPerLaunch * ntiles * 2;
    const int nupd = kLastOnly ? 1 : s_run;  // update rows (FP32: the last sweep's)
    const int nrows = nupd + (kNorm ? 1 : 0);
    for (int row = wv; row < nrows; row += NW) {
        const double* col = row < nupd ? lred[row] : lnrm;
        // the NW lane sums of this column, pairwise (depth log2 NW, not a chain of NW adds)
        double t[NW];
#pragma unroll
        for (int j = 0; j < NW; ++j) t[j] = col[lane + 64 * j];
#pragma unroll
        for (int d = 1; d < NW; d *= 2)
#pragma unroll
            for (int j = 0; j + d < NW; j += 2 * d) t[j] += t[j + d];
        const double acc = wave_sum(t[0]);
        const size_t slot = row < nupd ? ((size_t)row * ntiles + tile) * 2 : (size_t)tile * 2 + 1;
        if (lane == 0) a.partial[(L & 1) * pstride + slot] = acc;
    }
}

// The deciding workgroup of launch L (NW*64 threads): decides on launch L-1's sweeps from their
// per-tile partial sums (src/poisson.c:272-278) and writes state slot L&1, which launch L+1's
// workgroups and the host read.  Every sum runs in one fixed order: thread t adds the partials of
// tiles t, t + NT, ... in that order, each wave reduces its threads' sums with DPP moves, and lane
// i adds the per-wave totals of quantity i in wave order.
template <typename T, int S, int NW, bool NORM>
__device__ __forceinline__ void sor_rb_decider(const SorArgs<T>& a, const int L) {
    constexpr int NB = SorPlanes<T>::value;
    constexpr bool kNorm = sizeof(T) == 4 || NORM;
    constexpr bool kLastOnly = sizeof(T) == 4;
    constexpr int SE = kLastOnly ? 1 : S;
    constexpr int nq = SE + (kNorm ? 1 : 0);  // reduced quantities: SE updates (and the norm)
    __shared__ double lw[NW][nq];             // per-wave totals

    const SorCtl* prev = a.ctl + ((L + 1) & 1);
    SorCtl* mine = a.ctl + (L & 1);
    const int tid = threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ntiles = a.tiles_i * a.tiles_j;
    // FP32: launch L-1 left partials only if it was a tested launch
    const bool tested = sizeof(T) == 8 || ((L - 1) & a.test_mask) == a.test_mask;

    // the previous launch's per-tile partials (all S rows are read, before the slot is known, so
    // that these loads are in flight together with it; only the first sd are used).  Layout per
    // tile: [s][tile][0] the L1 update of sweep s (FP32: [0][tile][0], the launch's last sweep),
    // [0][tile][1] the iterate's norm after the last sweep.
    const size_t pstride = (size_t)kMaxSweepsPerLaunch * ntiles * 2;
    double pv[nq];
#pragma unroll
    for (int i = 0; i < nq; ++i) pv[i] = 0.0;
    if (tested) {
        const double* pp = a.partial + ((L + 1) & 1) * pstride;
        constexpr int NT = NW * 64;
        if (ntiles <= NT) {  // one tile per thread at most (1024^2: 240 tiles)
            if (tid < ntiles) {
#pragma unroll
                for (int s = 0; s < SE; ++s) pv[s] = pp[((size_t)s * ntiles + tid) * 2];
                if (kNorm) pv[nq - 1] = pp[(size_t)tid * 2 + 1];
            }
        } else {
            // several tiles per thread: U at a time, all loads issued before the sums, so that
            // the round trips overlap instead of following one another (same order)
            constexpr int U = 4;
            for (int t0 = tid; t0 < ntiles; t0 += U * NT) {
                double x[U][nq];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = t0 + u * NT;
                    const int tc = t < ntiles ? t : 0;  // in range; its value is discarded
#pragma unroll
                    for (int s = 0; s < SE; ++s) x[u][s] = pp[((size_t)s * ntiles + tc) * 2];
                    if (kNorm) x[u][nq - 1] = pp[(size_t)tc * 2 + 1];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (t0 + u * NT < ntiles) {
#pragma unroll
                        for (int i = 0; i < nq; ++i) pv[i] += x[u][i];
                    }
                }
            }
        }
    }
    const SorCtl P = *prev;
    if (P.status != SOR_RUNNING) {
        // ended: carry the state forward; a rollback is being done by this launch's sweepers,
        // after which plane P.parity holds the final iterate
        if (tid == 0) {
            SorCtl next = P;
            if (P.status == SOR_ROLLBACK) {
                next.status = SOR_CONVERGED;
                next.launches += 1;
            }
            next.s_done = 0;
            *mine = next;
        }
        return;
    }
    const int sd = P.s_done;  // sweeps the previous launch performed (0: none, e.g. launch 0)
    if (!tested) {
        // an untested launch: its sweeps are accepted, the solve runs on unless its budget is spent
        if (tid == 0) {
            SorCtl next = P;
            next.parity = L % NB;
            next.k = P.k + sd;
            if (next.k >= P.itmax) next.status = SOR_FAILED;
            const int budget = P.itmax - next.k;
            next.s_done = next.status == SOR_RUNNING ? (budget < S ? (budget > 0 ? budget : 0) : S) : 0;
            if (next.status == SOR_RUNNING && next.s_done > 0) next.launches = P.launches + 1;
            *mine = next;
        }
        return;
    }
    const int nwp = min(NW, (ntiles + 63) >> 6);  // waves holding partials
    if (sd > 0 && wv < nwp) {
        wave_sum_n(pv);
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < nq; ++i) lw[wv][i] = pv[i];
        }
    }
    __syncthreads();
    if (wv != 0) return;
    // lane i of wave 0 sums the per-wave totals of quantity i in wave order; broadcast
    double etot[SE];
    double ntot = 0.0;
    {
        double e = 0.0;
        if (sd > 0 && (lane < (kLastOnly ? 1 : sd) || (kNorm && lane == nq - 1))) {
            for (int j = 0; j < nwp; ++j) e += lw[j][lane];
        }
        const int lo = __double2loint(e), hi = __double2hiint(e);
#pragma unroll
        for (int s = 0; s < SE; ++s)
            etot[s] = __hiloint2double(__builtin_amdgcn_readlane(hi, s), __builtin_amdgcn_readlane(lo, s));
        if (kNorm)
            ntot = __hiloint2double(__builtin_amdgcn_readlane(hi, nq - 1), __builtin_amdgcn_readlane(lo, nq - 1));
    }
    const int k_prev = P.k, itmax = P.itmax;
    const double tol = P.tol, rtol = P.rtol;
    SorCtl next = P;
    next.parity = L % NB;  // where the previous launch left its output
    next.status = SOR_RUNNING;
    if (sd > 0) {
        // the first sweep s < sd with e < tol (converged) or, failing that, below rtol times the
        // iterate's norm after the launch (fold); static indices keep etot in registers.  Sum i
        // belongs to sweep i (FP32: to the last sweep, sd - 1, so there is never a rollback).
        // A sum that is NaN or +infinity (sums of |delta| are never negative) ends the solve as
        // diverged at that sweep: no later sweep can bring a non-finite iterate below tol, and the
        // reference's loop would spend its whole budget before failing (src/poisson.c:280-284).
        int s_conv = -1, s_fold = -1, s_bad = -1;
#pragma unroll
        for (int i = 0; i < SE; ++i) {
            const int s = kLastOnly ? sd - 1 : i;
            if (s < sd && s_conv < 0 && s_fold < 0 && s_bad < 0) {
                if (!(etot[i] <= DBL_MAX)) s_bad = s;
                else if (etot[i] < tol) s_conv = s;
                else if (rtol > 0.0 && etot[i] < rtol * ntot) s_fold = s;
            }
        }
        if (kNorm && s_conv < 0 && s_fold < 0 && s_bad < 0 && !(ntot <= DBL_MAX)) s_bad = sd - 1;
        const int s_rep = (s_conv >= 0 && s_conv < sd - 1) ? s_conv : sd - 1;  // reported sweep
        double e_rep = 0.0;
#pragma unroll
        for (int i = 0; i < SE; ++i) {
            if ((kLastOnly ? sd - 1 : i) == s_rep) e_rep = etot[i];
        }
        next.e_last = e_rep;
        next.n_last = ntot;
        if (s_conv >= 0 && s_conv < sd - 1) {
            // converged inside the previous launch: the next launch redoes s_conv+1 sweeps from
            // that launch's input (plane (L-1) % NB, untouched since), into plane L % NB (its
            // output, no longer needed); that iterate is the final one (src/poisson.c:273)
            next.status = SOR_ROLLBACK;
            next.k = k_prev + s_conv + 1;
            next.rb_src = (L + NB - 1) % NB;
            next.rb_n = s_conv + 1;
        } else {
            next.k = k_prev + sd;
            if (s_conv >= 0) next.status = SOR_CONVERGED;
            else if (s_fold >= 0) next.status = SOR_FOLD;
            else if (s_bad >= 0) next.status = SOR_DIVERGED;
            else if (next.k >= itmax) next.status = SOR_FAILED;
        }
    } else if (itmax <= k_prev) {
        next.status = SOR_FAILED;
    }
    // the sweeps this launch's sweepers perform (they computed the same from the slot they read)
    const int budget = itmax - (k_prev + sd);
    next.s_done = next.status == SOR_RUNNING ? (budget < S ? (budget > 0 ? budget : 0) : S) : 0;
    if (next.status == SOR_RUNNING && next.s_done > 0) next.launches = P.launches + 1;
    if (lane == 0) *mine = next;
}

template <typename T, int S, int NW, int RB, bool GS, bool XR, bool NORM = false>
__global__ void __launch_bounds__(NW * 64) sor_rb_kernel(SorArgs<T> a) {
    const int ntiles = a.tiles_i * a.tiles_j;
    if ((int)blockIdx.x == ntiles) sor_rb_decider<T, S, NW, NORM>(a, a.launch);
    else sor_rb_sweeper<T, S, NW, RB, GS, XR, NORM>(a, a.launch, sor_tile_of_block(blockIdx.x, ntiles));
}

// Batched form: workgroup b runs tile map[b].tile of solve map[b].inst (inst < 0: padding; tile
// == the solve's tile count: its decider).  All solves of a batch share the launch index L.  The
// solve's arguments are read in place from global memory: a private copy would be indexed by the
// plane number and live in scratch memory.
template <typename T, int S, int NW, int RB, bool GS, bool XR>
__global__ void __launch_bounds__(NW * 64) sor_rb_batch_kernel(const SorArgs<T>* __restrict__ args,
                                                               const SorBlock* __restrict__ map, int L) {
    const SorBlock m = map[blockIdx.x];
    if (m.inst < 0) return;
    const SorArgs<T>& a = args[m.inst];
    if (m.tile == a.tiles_i * a.tiles_j) sor_rb_decider<T, S, NW, false>(a, L);
    else sor_rb_sweeper<T, S, NW, RB, GS, XR, false>(a, L, m.tile);
}

template <typename T, int S, int NW, int RB>
inline int sor_launch_shape(const SorArgs<T>& a, bool gs, hipStream_t st) {
    const int ntiles = a.tiles_i * a.tiles_j;
    const dim3 grid(ntiles + 1), block(NW * 64);  // the tiles' sweepers and one decider
    // FP32 always multiplies by RN(1/B); FP64 does when 1/B is exact
    const bool xr = sizeof(T) == 4 || a.exact_rcp;
    if constexpr (sizeof(T) == 8) {
        if (a.norm && !gs) {  // FP64 with the iterate norm (relative stopping rule)
            if (xr) hipLaunchKernelGGL((sor_rb_kernel<T, S, NW, RB, false, true, true>), grid, block, 0, st, a);
            else hipLaunchKernelGGL((sor_rb_kernel<T, S, NW, RB, false, false, true>), grid, block, 0, st, a);
            return hipGetLastError() == hipSuccess ? 0 : 1;
        }
    }
    if (gs && xr) hipLaunchKernelGGL((sor_rb_kernel<T, S, NW, RB, true, true>), grid, block, 0, st, a);
    else if (gs) hipLaunchKernelGGL((sor_rb_kernel<T, S, NW, RB, true, false>), grid, block, 0, st, a);
    else if (xr) hipLaunchKernelGGL((sor_rb_kernel<T, S, NW, RB, false, true>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((sor_rb_kernel<T, S, NW, RB, false, false>), grid, block, 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// The launch shapes compiled in, TileShape {s, nw, rb}: S sweeps per launch, nw waves of rb rows;
// the tile owns (nw*rb - 4s) rows x (128 - 4s) columns.  The default is the first of (5, 4, 6),
// (5, 8, 4), (6, 16, 4), (5, 16, 4) whose tiles fit the device's CUs one per CU: on a small
// grid the GPU is otherwise idle, so thin tiles (more halo, shorter workgroups) win -- per sweep
// at 256^2 FP64 1.50 us with (5, 4, 6) (192 workgroups of 4 waves) against 2.21 us with
// (5, 16, 4) (18 workgroups); at 1024^2 (5, 16, 4) gives 240 workgroups of 16 waves (four per
// SIMD), 1.72x halo redundancy.  The others serve the batched autotuner and the sweep tests.
#define CNV_SOR_SHAPES(X) \
    X(1, 4, 10)           \
    X(2, 4, 10)           \
    X(4, 8, 8)            \
    X(6, 8, 8)            \
    X(8, 8, 10)           \
    X(4, 16, 4)           \
    X(5, 16, 4)           \
    X(6, 16, 4)           \
    X(5, 12, 6)           \
    X(5, 8, 4)            \
    X(5, 8, 8)            \
    X(4, 8, 4)            \
    X(5, 4, 6)

template <typename T>
inline int sor_dispatch(const SorArgs<T>& a, TileShape sh, bool gs, hipStream_t st) {
#define CNV_SOR_CASE(s_, nw_, rb_) \
    if (sh.s == s_ && sh.nw == nw_ && sh.rb == rb_) return sor_launch_shape<T, s_, nw_, rb_>(a, gs, st);
    CNV_SOR_SHAPES(CNV_SOR_CASE)
#undef CNV_SOR_CASE
    return 2;
}

// Batched launch: FP64 uses the corrected division (exact for every B), FP32 the multiply.
template <typename T, int S, int NW, int RB>
inline int sor_batch_launch_shape(const SorArgs<T>* args, const SorBlock* map, int nblocks, int L, bool gs,
                                  hipStream_t st) {
    constexpr bool XR = sizeof(T) == 4;
    const dim3 grid(nblocks), block(NW * 64);
    if (gs) hipLaunchKernelGGL((sor_rb_batch_kernel<T, S, NW, RB, true, XR>), grid, block, 0, st, args, map, L);
    else hipLaunchKernelGGL((sor_rb_batch_kernel<T, S, NW, RB, false, XR>), grid, block, 0, st, args, map, L);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

template <typename T>
inline int sor_batch_dispatch(const SorArgs<T>* args, const SorBlock* map, int nblocks, int L, TileShape sh,
                              bool gs, hipStream_t st) {
#define CNV_SOR_CASE(s_, nw_, rb_) \
    if (sh.s == s_ && sh.nw == nw_ && sh.rb == rb_) return sor_batch_launch_shape<T, s_, nw_, rb_>(args, map, nblocks, L, gs, st);
    CNV_SOR_SHAPES(CNV_SOR_CASE)
#undef CNV_SOR_CASE
    return 2;
}

}  // namespace cnv
