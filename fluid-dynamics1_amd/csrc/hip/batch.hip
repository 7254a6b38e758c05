// batch.hip -- many independent cavities of different sizes advanced together on one GPU.
//
// A batch holds one context per cavity (its fields, stencil tables and Diff tables, exactly as
// cnv_create builds them) and advances all of them with batched launches: every phase of the
// time step (wall vorticity, transport, Poisson init, each SOR launch, velocity, continuity,
// and the defect-correction residual / correction) is ONE launch whose workgroups are mapped to
// (cavity, local block) by a table built at creation.  Each cavity keeps its own Poisson state
// slots, so one batched SOR launch advances every cavity whose solve is still running and the
// workgroups of finished cavities return at once; the host polls all cavities' slots with one
// copy per chunk of launches.  The per-point arithmetic is the one-cavity kernels' own
// (step_kernels.h, sor_rb.h), so each cavity's result equals the same cavity run alone with the
// same launch shape, bit for bit.
//
// Defect correction runs its outer iterations in lockstep over the batch: residual and inner
// init for every cavity still iterating, SOR launches until every inner solve has ended, then
// the FP64 correction.  Cavities that converged drop out of the later outer iterations.
//
// A cavity that diverges -- a non-finite L1 update ends its solve (SOR_DIVERGED), or its largest
// |u| or |v| after a step exceeds its blowup_speed -- is taken out of the workgroup maps: later
// steps launch nothing for it and the other cavities go on (cnv_batch_status reports it).
//
// The SOR launch shape (sweeps per launch, waves, rows per wave) is one per batch; the autotuner
// (cnv_batch_autotune) times every compiled shape on the batch itself and keeps the fastest.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "cnv.h"
#include "cnv_device.h"
#include "ctx.h"
#include "sor_rb.h"
#include "step_kernels.h"

using namespace cnv;

namespace {

// Per-cavity description read by the batched kernels.
struct StepItem {
    double* w[2];
    double* psi[3];
    double* u;
    double* v;
    float* e32[2];
    float* r32;
    SorCtl* ctl;  // the cavity's two state slots (inside the batch's slot array)
    Diag* diag;
    Stencils st;
    WallBC bc;
    int n, ld;
    double rRe, dt, tol, c1, c2, c3, B;
    int itmax;
};

// Workgroup b of a batched launch: cavity inst (< 0: padding), local block lb of nb.
struct BlockRef {
    int inst, lb, nb;
};

// Defect correction, per cavity and outer iteration: does it take part, and its sweep budget.
struct DcSlot {
    int active;
    int itmax;
};

constexpr int kLinBlocks = 64;  // workgroups per cavity for fills and reductions

__global__ void __launch_bounds__(kBlock) batch_wall_vorticity_kernel(const StepItem* __restrict__ it,
                                                                     const BlockRef* __restrict__ map, int wcur) {
    const BlockRef m = map[blockIdx.x];
    if (m.inst < 0) return;
    const StepItem& I = it[m.inst];
    wall_vorticity_at(I.w[wcur], I.u, I.v, I.n, I.ld, I.st, I.bc, m.lb * kBlock + threadIdx.x);
}

// local block lb of a cavity covers 64 columns x 4 rows, blocks ordered row-major
__device__ inline bool cell_of(const BlockRef& m, int n, int& i, int& j) {
    const int gx = (n + 63) / 64;
    const int by = m.lb / gx, bx = m.lb - by * gx;
    j = bx * 64 + (threadIdx.x & 63);
    i = by * 4 + (threadIdx.x >> 6);
    return i < n && j < n;
}

template <bool DC>
__global__ void __launch_bounds__(kBlock) batch_transport_kernel(const StepItem* __restrict__ it,
                                                                 const BlockRef* __restrict__ map, int wcur) {
    const BlockRef m = map[blockIdx.x];
    if (m.inst < 0) return;
    const StepItem& I = it[m.inst];
    int i, j;
    if (!cell_of(m, I.n, i, j)) return;
    if (DC)
        transport_fp32_at(I.w[wcur], I.w[wcur ^ 1], I.u, I.v, I.n, I.ld, I.st, I.bc, I.rRe, I.dt, i, j);
    else
        transport_fp64_at(I.w[wcur], I.w[wcur ^ 1], I.u, I.v, I.n, I.ld, I.st, I.bc, I.rRe, I.dt, i, j);
}

// psi = 0 (src/poisson.c:229-230), both state slots reset, diagnostic reset
__global__ void __launch_bounds__(kBlock) batch_poisson_init_kernel(const StepItem* __restrict__ it,
                                                                    const BlockRef* __restrict__ map) {
    const BlockRef m = map[blockIdx.x];
    if (m.inst < 0) return;
    const StepItem& I = it[m.inst];
    const long long count = (long long)I.n * I.ld;
    for (long long q = (long long)m.lb * kBlock + threadIdx.x; q < count; q += (long long)m.nb * kBlock)
        I.psi[0][q] = 0.0;
    if (m.lb == 0 && threadIdx.x < 2) ctl_reset(I.ctl + threadIdx.x, I.tol, 0.0, I.itmax);
    if (m.lb == 0 && threadIdx.x == 0) diag_reset(I.diag);
}

// DC: the refined form's accumulator psi[0]; F32: its FP32 stencils (velocity_fp32_at)
template <bool DC, bool F32>
__global__ void __launch_bounds__(kBlock) batch_velocity_kernel(const StepItem* __restrict__ it,
                                                                const BlockRef* __restrict__ map, int fin_slot) {
    const BlockRef m = map[blockIdx.x];
    if (m.inst < 0) return;
    const StepItem& I = it[m.inst];
    int i, j;
    if (!cell_of(m, I.n, i, j)) return;
    const double* psi = DC ? I.psi[0] : I.psi[I.ctl[fin_slot].parity];
    if (F32) velocity_fp32_at(psi, I.u, I.v, I.ld, I.st, i, j);
    else velocity_at(psi, I.u, I.v, I.ld, I.st, i, j);
}

__global__ void __launch_bounds__(kBlock) batch_continuity_kernel(const StepItem* __restrict__ it,
                                                                  const BlockRef* __restrict__ map) {
    const BlockRef m = map[blockIdx.x];
    if (m.inst < 0) return;  // uniform per workgroup: the block-wide fold below is safe
    const StepItem& I = it[m.inst];
    double mx = -DBL_MAX, mn = DBL_MAX;
    unsigned long long ks = dkey(0.0);
    const int n = I.n;
    const long long count = (long long)n * n;
    for (long long q = (long long)m.lb * kBlock + threadIdx.x; q < count; q += (long long)m.nb * kBlock) {
        const int i = (int)(q / n), j = (int)(q - (long long)i * n);
        const double c = continuity_at(I.u, I.v, I.ld, I.st, i, j);
        if (c > mx) mx = c;
        if (c < mn) mn = c;
        const size_t p = (size_t)i * I.ld + j;
        const unsigned long long kq = speed_key(I.u[p], I.v[p]);
        ks = kq > ks ? kq : ks;
    }
    diag_fold_block(mx, mn, ks, I.diag);
}

__global__ void __launch_bounds__(kBlock) batch_dc_residual_kernel(const StepItem* __restrict__ it,
                                                                   const BlockRef* __restrict__ map,
                                                                   const DcSlot* __restrict__ dc, int wcur) {
    const BlockRef m = map[blockIdx.x];
    if (m.inst < 0 || !dc[m.inst].active) return;
    const StepItem& I = it[m.inst];
    int i, j;
    if (!cell_of(m, I.n, i, j)) return;
    dc_residual_at(I.w[wcur], I.psi[0], I.r32, I.n, I.ld, I.c1, I.c2, I.c3, I.B, i, j);
}

// inner solve start: e = 0 and fresh slots for active cavities; inactive cavities get terminal
// slots, so the batched SOR launches pass over them
__global__ void __launch_bounds__(kBlock) batch_dc_inner_init_kernel(const StepItem* __restrict__ it,
                                                                     const BlockRef* __restrict__ map,
                                                                     const DcSlot* __restrict__ dc, double rtol) {
    const BlockRef m = map[blockIdx.x];
    if (m.inst < 0) return;
    const StepItem& I = it[m.inst];
    const DcSlot d = dc[m.inst];
    if (d.active) {
        const long long count = (long long)I.n * I.ld;
        for (long long q = (long long)m.lb * kBlock + threadIdx.x; q < count; q += (long long)m.nb * kBlock)
            I.e32[0][q] = 0.0f;
    }
    if (m.lb == 0 && threadIdx.x < 2) {
        ctl_reset(I.ctl + threadIdx.x, I.tol, rtol, d.itmax);
        if (!d.active) I.ctl[threadIdx.x].status = SOR_CONVERGED;
    }
}

// psi64 += e32 (the inner solve's final iterate) for active cavities
__global__ void __launch_bounds__(kBlock) batch_dc_correct_kernel(const StepItem* __restrict__ it,
                                                                  const BlockRef* __restrict__ map,
                                                                  const DcSlot* __restrict__ dc, int fin_slot) {
    const BlockRef m = map[blockIdx.x];
    if (m.inst < 0 || !dc[m.inst].active) return;
    const StepItem& I = it[m.inst];
    const float* e = I.e32[I.ctl[fin_slot].parity & 1];
    const long long count = (long long)I.n * I.ld;
    for (long long q = (long long)m.lb * kBlock + threadIdx.x; q < count; q += (long long)m.nb * kBlock)
        I.psi[0][q] += (double)e[q];
}

// synthetic source for the autotuner: a smooth field with a few sign changes
template <typename T>
__global__ void __launch_bounds__(kBlock) batch_fill_source_kernel(const StepItem* __restrict__ it,
                                                                   const BlockRef* __restrict__ map, T** dst) {
    const BlockRef m = map[blockIdx.x];
    if (m.inst < 0) return;
    const StepItem& I = it[m.inst];
    T* p = dst[m.inst];
    const long long count = (long long)I.n * I.ld;
    for (long long q = (long long)m.lb * kBlock + threadIdx.x; q < count; q += (long long)m.nb * kBlock) {
        const int i = (int)(q / I.ld), j = (int)(q - (long long)i * I.ld);
        const double x = (double)j / I.n, y = (double)i / I.n;
        p[q] = (T)(sin(6.283185307179586 * x) * cos(9.42477796076938 * y));
    }
}

// autotuner: running slots that no sweep can stop (tol < 0, unreachable budget)
__global__ void __launch_bounds__(kBlock) batch_tune_slots_kernel(SorCtl* ctl, int nslots) {
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q < nslots) ctl_reset(ctl + q, -1.0, 0.0, 1 << 30);
}

__global__ void __launch_bounds__(kBlock) batch_zero_kernel(const StepItem* __restrict__ it,
                                                            const BlockRef* __restrict__ map) {
    const BlockRef m = map[blockIdx.x];
    if (m.inst < 0) return;
    const StepItem& I = it[m.inst];
    const long long count = (long long)I.n * I.ld;
    for (long long q = (long long)m.lb * kBlock + threadIdx.x; q < count; q += (long long)m.nb * kBlock) {
        I.psi[0][q] = 0.0;
        I.psi[1][q] = 0.0;
        I.psi[2][q] = 0.0;
        if (I.e32[0]) {
            I.e32[0][q] = 0.0f;
            I.e32[1][q] = 0.0f;
        }
    }
}

thread_local char g_berr[256] = "";  // last error of this thread

}  // namespace

struct cnv_batch {
    int count = 0;
    int dc = 0, gs = 0;
    std::vector<int> alive;        // 1: the cavity takes part in the steps; 0: it diverged
    std::vector<double> blowup;    // per cavity: blowup_speed (0 = no check)
    bool maps_dirty = false;       // alive changed: rebuild the workgroup maps before the next step
    bool transport_fp32 = true;  // refined form: FP32 transport and velocity stencils (the default)
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<cnv_ctx*> ctx;
    TileShape shape{5, 16, 4};
    int launches_per_poll = 2;
    int wcur = 0;
    int fin_slot = 0;
    int k_prev = 0;  // most sweeps any cavity needed in the previous solve
    static constexpr int kDcPred = 64;
    int k_dc[kDcPred] = {};  // FP32_DC: most sweeps any cavity's inner solve of outer iteration o took
    double rtol = 1e-5;
    int max_outer = 50;
    // device tables
    StepItem* d_items = nullptr;
    SorCtl* d_ctl = nullptr;  // [count][2]
    Diag* d_diag = nullptr;   // [count]
    double* d_partial = nullptr;
    SorArgs<double>* d_args64[2] = {nullptr, nullptr};  // source w[wcur] for wcur = 0, 1
    SorArgs<float>* d_args32 = nullptr;
    DcSlot* d_dc = nullptr;
    BlockRef *d_map_edge = nullptr, *d_map_cell = nullptr, *d_map_lin = nullptr;
    int nb_edge = 0, nb_cell = 0, nb_lin = 0;
    SorBlock* d_map_sor = nullptr;
    int nb_sor = 0;
    std::vector<size_t> part_off;  // per cavity, doubles into d_partial
    size_t part_total = 0;
    // host mirrors (pinned)
    SorCtl* h_ctl = nullptr;  // [kSlots][count][2]
    Diag* h_diag = nullptr;
    DcSlot* h_dc = nullptr;
    hipEvent_t ev[kSlots] = {};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    char err[256] = "";
};

static int bfail(cnv_batch* b, int code, const char* fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (b) snprintf(b->err, sizeof(b->err), "%s", buf);
    snprintf(g_berr, sizeof(g_berr), "%s", buf);
    return code;
}
#define BHIP(b, x)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return bfail((b), CNV_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)
#define BLAUNCH(b)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = hipGetLastError();                                                            \
        if (e_ != hipSuccess) return bfail((b), CNV_ERR_HIP, "batched launch failed: %s", hipGetErrorString(e_)); \
    } while (0)

static int tiles_of(int n, TileShape sh, int* ti, int* tj) {
    *ti = (n + tile_rows(sh) - 1) / tile_rows(sh);
    *tj = (n + tile_cols(sh) - 1) / tile_cols(sh);
    return *ti * *tj;
}

template <typename V>
static int upload(cnv_batch* b, V** dst, const std::vector<V>& src) {
    if (*dst) BHIP(b, hipFree(*dst));
    *dst = nullptr;
    BHIP(b, hipMalloc(dst, sizeof(V) * std::max<size_t>(src.size(), 1)));
    if (!src.empty()) BHIP(b, hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice));
    return CNV_OK;
}

// Workgroup maps of the cell-wise, edge-wise and fill/reduction launches (cavity-major).
static int build_maps(cnv_batch* b) {
    std::vector<BlockRef> edge, cell, lin;
    for (int i = 0; i < b->count; ++i) {
        if (!b->alive[i]) continue;
        const int n = b->ctx[i]->n;
        const int ne = (4 * n + kBlock - 1) / kBlock;
        for (int l = 0; l < ne; ++l) edge.push_back({i, l, ne});
        const int nc = ((n + 63) / 64) * ((n + 3) / 4);
        for (int l = 0; l < nc; ++l) cell.push_back({i, l, nc});
        const long long cnt = (long long)n * b->ctx[i]->ld;
        const int nl = (int)std::min<long long>(kLinBlocks, (cnt + kBlock - 1) / kBlock);
        for (int l = 0; l < nl; ++l) lin.push_back({i, l, nl});
    }
    b->nb_edge = (int)edge.size();
    b->nb_cell = (int)cell.size();
    b->nb_lin = (int)lin.size();
    int rc = upload(b, &b->d_map_edge, edge);
    if (!rc) rc = upload(b, &b->d_map_cell, cell);
    if (!rc) rc = upload(b, &b->d_map_lin, lin);
    return rc;
}

// SOR workgroup map and per-cavity SOR arguments for launch shape sh.  Workgroups are dispatched
// round-robin over the 8 XCDs; the (cavity, tile) list is cut into 8 contiguous pieces, one per
// XCD, so that neighbouring tiles of a cavity share an L2.
static int build_sor(cnv_batch* b, TileShape sh) {
    std::vector<SorBlock> list;
    std::vector<int> ti(b->count), tj(b->count);
    for (int i = 0; i < b->count; ++i) {
        const int nt = tiles_of(b->ctx[i]->n, sh, &ti[i], &tj[i]);
        if (b->alive[i])
            for (int t = 0; t <= nt; ++t) list.push_back({i, t});  // t == nt: the cavity's decider
    }
    while (list.size() % 8) list.push_back({-1, 0});
    const int per = (int)list.size() / 8;
    std::vector<SorBlock> map(list.size());
    for (size_t blk = 0; blk < list.size(); ++blk) map[blk] = list[(blk & 7) * per + (blk >> 3)];
    b->nb_sor = (int)map.size();
    int rc = upload(b, &b->d_map_sor, map);
    if (rc) return rc;
    // partial buffers sized for this shape
    b->part_off.assign(b->count, 0);
    size_t total = 0;
    for (int i = 0; i < b->count; ++i) {
        b->part_off[i] = total;
        total += (size_t)2 * 2 * kMaxSweepsPerLaunch * ti[i] * tj[i];
    }
    if (total > b->part_total) {
        if (b->d_partial) BHIP(b, hipFree(b->d_partial));
        b->d_partial = nullptr;
        BHIP(b, hipMalloc(&b->d_partial, sizeof(double) * total));
        b->part_total = total;
    }
    std::vector<SorArgs<double>> a64[2];
    std::vector<SorArgs<float>> a32;
    for (int i = 0; i < b->count; ++i) {
        cnv_ctx* c = b->ctx[i];
        for (int w = 0; w < 2; ++w) {
            SorArgs<double> a = sor_args64(c, c->w[w], 1);
            a.ctl = b->d_ctl + 2 * i;
            a.partial = b->d_partial + b->part_off[i];
            a.tiles_i = ti[i];
            a.tiles_j = tj[i];
            a64[w].push_back(a);
        }
        if (b->dc) {
            SorArgs<float> a = sor_args32(c);
            // by default every launch tests its last sweep: the batch's outer iterations run in
            // lockstep, so a late test on any cavity's short inner solve holds up every cavity
            a.test_mask = inner_test_every(c->p, 1) - 1;
            a.ctl = b->d_ctl + 2 * i;
            a.partial = b->d_partial + b->part_off[i];
            a.tiles_i = ti[i];
            a.tiles_j = tj[i];
            a32.push_back(a);
        }
    }
    rc = upload(b, &b->d_args64[0], a64[0]);
    if (!rc) rc = upload(b, &b->d_args64[1], a64[1]);
    if (!rc && b->dc) rc = upload(b, &b->d_args32, a32);
    if (!rc) b->shape = sh;
    return rc;
}

static int sor_launch(cnv_batch* b, int L, bool fp32) {
    const int rc = fp32 ? sor_batch_dispatch<float>(b->d_args32, b->d_map_sor, b->nb_sor, L, b->shape, b->gs, b->stream)
                        : sor_batch_dispatch<double>(b->d_args64[b->wcur], b->d_map_sor, b->nb_sor, L, b->shape,
                                                     b->gs, b->stream);
    if (rc) return bfail(b, CNV_ERR_HIP, "batched SOR launch failed: %s", hipGetErrorString(hipGetLastError()));
    return CNV_OK;
}

// Queue batched SOR launches until every cavity's solve has ended (same protocol as the
// one-cavity driver: one chunk queued behind the one being waited on).  snaps[i]: final slot.
static int run_sor_batch(cnv_batch* b, bool fp32, std::vector<SorCtl>& snaps) {
    const int S = b->shape.s;
    const int M = b->launches_per_poll;
    int itmax = 0;
    for (cnv_ctx* c : b->ctx) itmax = std::max(itmax, c->p.itmax);
    const long long cap = (long long)itmax + 4LL * M + 8;
    int first = M;
    if (b->k_prev > 0) first = std::max(first, b->k_prev / S - M);
    long long queued = 0;
    int slot = 0;
    int pending[kSlots], pend_fin[kSlots];
    int npend = 0;
    const size_t per = (size_t)2 * b->count;
    while (true) {
        const int chunk = queued == 0 ? first : M;
        for (int l = 0; l < chunk; ++l) {
            const int rc = sor_launch(b, (int)(queued + l), fp32);
            if (rc) return rc;
        }
        queued += chunk;
        b->fin_slot = (int)((queued - 1) & 1);
        BHIP(b, hipMemcpyAsync(b->h_ctl + slot * per, b->d_ctl, sizeof(SorCtl) * per, hipMemcpyDeviceToHost, b->stream));
        BHIP(b, hipEventRecord(b->ev[slot], b->stream));
        pending[npend] = slot;
        pend_fin[npend] = b->fin_slot;
        ++npend;
        slot = (slot + 1) % kSlots;
        if (npend >= 2) {
            const int s0 = pending[0], f0 = pend_fin[0];
            for (int q = 1; q < npend; ++q) {
                pending[q - 1] = pending[q];
                pend_fin[q - 1] = pend_fin[q];
            }
            --npend;
            BHIP(b, hipEventSynchronize(b->ev[s0]));
            const SorCtl* h = b->h_ctl + s0 * per;
            bool all = true;
            for (int i = 0; i < b->count && all; ++i) all = !b->alive[i] || sor_done(h[2 * i + f0].status);
            if (all) {
                snaps.resize(b->count);
                for (int i = 0; i < b->count; ++i) snaps[i] = h[2 * i + f0];
                // launches queued after the snapshot carry the terminal slots forward
                return CNV_OK;
            }
        }
        if (queued > cap + first) return bfail(b, CNV_ERR_HIP, "batched SOR did not terminate after %lld launches", queued);
    }
}

static int batch_one_step(cnv_batch* b, cnv_step_stats* stats) {
    const int cnt = b->count;
    const int win = b->wcur;
    if (b->maps_dirty) {  // a cavity diverged (or came back): maps of the live cavities only
        int rc = build_maps(b);
        if (!rc) rc = build_sor(b, b->shape);
        if (rc) return rc;
        b->maps_dirty = false;
    }
    int nalive = 0;
    for (int i = 0; i < cnt; ++i) nalive += b->alive[i];
    if (nalive == 0) {  // nothing left to step
        for (int i = 0; i < cnt && stats; ++i) {
            stats[i] = cnv_step_stats{};
            stats[i].status = CNV_DIVERGED;
        }
        return CNV_OK;
    }
    hipLaunchKernelGGL(batch_wall_vorticity_kernel, dim3(b->nb_edge), dim3(kBlock), 0, b->stream, b->d_items,
                       b->d_map_edge, win);
    BLAUNCH(b);
    if (b->dc && b->transport_fp32)
        hipLaunchKernelGGL(batch_transport_kernel<true>, dim3(b->nb_cell), dim3(kBlock), 0, b->stream, b->d_items,
                           b->d_map_cell, win);
    else
        hipLaunchKernelGGL(batch_transport_kernel<false>, dim3(b->nb_cell), dim3(kBlock), 0, b->stream, b->d_items,
                           b->d_map_cell, win);
    BLAUNCH(b);
    b->wcur ^= 1;
    for (int i = 0; i < cnt; ++i)  // a diverged cavity keeps the buffers it was left in
        if (b->alive[i]) b->ctx[i]->wcur = b->wcur;
    hipLaunchKernelGGL(batch_poisson_init_kernel, dim3(b->nb_lin), dim3(kBlock), 0, b->stream, b->d_items,
                       b->d_map_lin);
    BLAUNCH(b);

    std::vector<SorCtl> fin(cnt), snaps;
    std::vector<int> outer(cnt, 0), total(cnt, 0), active(b->alive);
    if (!b->dc) {
        int rc = run_sor_batch(b, false, snaps);
        if (rc) return rc;
        fin = snaps;
    } else {
        for (int i = 0; i < cnt; ++i) fin[i].status = SOR_FAILED;
        for (int o = 0; o < b->max_outer; ++o) {
            int nact = 0;
            for (int i = 0; i < cnt; ++i) {
                if (active[i] && total[i] >= b->ctx[i]->p.itmax) {  // sweep budget spent: failed
                    active[i] = 0;
                    fin[i].status = SOR_FAILED;
                }
                b->h_dc[i].active = active[i];
                b->h_dc[i].itmax = b->ctx[i]->p.itmax - total[i];
                nact += active[i];
            }
            if (!nact) break;
            BHIP(b, hipMemcpyAsync(b->d_dc, b->h_dc, sizeof(DcSlot) * cnt, hipMemcpyHostToDevice, b->stream));
            hipLaunchKernelGGL(batch_dc_residual_kernel, dim3(b->nb_cell), dim3(kBlock), 0, b->stream, b->d_items,
                               b->d_map_cell, b->d_dc, b->wcur);
            BLAUNCH(b);
            hipLaunchKernelGGL(batch_dc_inner_init_kernel, dim3(b->nb_lin), dim3(kBlock), 0, b->stream, b->d_items,
                               b->d_map_lin, b->d_dc, b->rtol);
            BLAUNCH(b);
            b->k_prev = o < cnv_batch::kDcPred ? b->k_dc[o] : 0;  // this outer iteration's last length
            int rc = run_sor_batch(b, true, snaps);
            if (rc) return rc;
            hipLaunchKernelGGL(batch_dc_correct_kernel, dim3(b->nb_lin), dim3(kBlock), 0, b->stream, b->d_items,
                               b->d_map_lin, b->d_dc, b->fin_slot);
            BLAUNCH(b);
            // the slot mirror below is read after the next synchronisation, not now: b->h_dc
            // stays untouched until then (the copy above is ordered before the SOR launches
            // that run_sor_batch waited for)
            int kmax_o = 0;
            for (int i = 0; i < cnt; ++i)
                if (active[i]) kmax_o = std::max(kmax_o, snaps[i].k);
            if (o < cnv_batch::kDcPred) b->k_dc[o] = kmax_o;
            for (int i = 0; i < cnt; ++i) {
                if (!active[i]) continue;
                total[i] += snaps[i].k;
                outer[i] = o + 1;
                fin[i] = snaps[i];
                fin[i].k = total[i];
                if (dc_final(snaps[i].status)) active[i] = 0;
            }
        }
        for (int i = 0; i < cnt; ++i)
            if (active[i]) fin[i].status = SOR_FAILED;  // ran out of outer iterations
    }
    int kmax = 0;
    for (int i = 0; i < cnt; ++i)
        if (b->alive[i]) kmax = std::max(kmax, b->dc ? 0 : fin[i].k);
    if (!b->dc) b->k_prev = kmax;

    if (b->dc && b->transport_fp32)
        hipLaunchKernelGGL((batch_velocity_kernel<true, true>), dim3(b->nb_cell), dim3(kBlock), 0, b->stream,
                           b->d_items, b->d_map_cell, b->fin_slot);
    else if (b->dc)
        hipLaunchKernelGGL((batch_velocity_kernel<true, false>), dim3(b->nb_cell), dim3(kBlock), 0, b->stream,
                           b->d_items, b->d_map_cell, b->fin_slot);
    else
        hipLaunchKernelGGL((batch_velocity_kernel<false, false>), dim3(b->nb_cell), dim3(kBlock), 0, b->stream,
                           b->d_items, b->d_map_cell, b->fin_slot);
    BLAUNCH(b);
    hipLaunchKernelGGL(batch_continuity_kernel, dim3(b->nb_lin), dim3(kBlock), 0, b->stream, b->d_items, b->d_map_lin);
    BLAUNCH(b);
    BHIP(b, hipMemcpyAsync(b->h_diag, b->d_diag, sizeof(Diag) * cnt, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    int status = CNV_OK;
    for (int i = 0; i < cnt; ++i) {
        cnv_step_stats s{};
        if (!b->alive[i]) {  // diverged in an earlier step: not stepped
            s.status = CNV_DIVERGED;
            if (stats) stats[i] = s;
            continue;
        }
        s.status = solve_status(fin[i].status);
        s.sweeps = fin[i].k;
        s.k = fin[i].k - 1;
        s.e = fin[i].e_last;
        s.outer = outer[i];
        s.cont_max = dkey_inv(b->h_diag[i].kmax);
        s.cont_min = dkey_inv(b->h_diag[i].kmin);
        const double speed = dkey_inv(b->h_diag[i].kspd);
        if (s.status == CNV_OK && b->blowup[i] > 0 && !(speed <= b->blowup[i])) s.status = CNV_DIVERGED;
        if (stats) stats[i] = s;
        if (s.status == CNV_DIVERGED) {
            b->alive[i] = 0;
            b->maps_dirty = true;
        } else if (s.status != CNV_OK) {
            status = CNV_NOT_CONVERGED;
        }
        b->ctx[i]->fin_slot = b->fin_slot;
    }
    if (status != CNV_OK)
        return bfail(b, CNV_NOT_CONVERGED, "Error: maximum number of iterations achieved for Poisson equation.");
    return CNV_OK;
}

// ---- C ABI -----------------------------------------------------------------------------------

const char* cnv_batch_last_error(const cnv_batch* b) { return b ? b->err : g_berr; }

int cnv_batch_size(const cnv_batch* b) { return b ? b->count : 0; }

void cnv_batch_destroy(cnv_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (cnv_ctx* c : b->ctx) cnv_destroy(c);
    void* dev[] = {b->d_items, b->d_ctl, b->d_diag, b->d_partial, b->d_args64[0], b->d_args64[1], b->d_args32,
                   b->d_dc, b->d_map_edge, b->d_map_cell, b->d_map_lin, b->d_map_sor};
    for (void* p : dev)
        if (p) (void)hipFree(p);
    if (b->h_ctl) (void)hipHostFree(b->h_ctl);
    if (b->h_diag) (void)hipHostFree(b->h_diag);
    if (b->h_dc) (void)hipHostFree(b->h_dc);
    for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int cnv_batch_create(const cnv_params* items, int count, cnv_batch** out) {
    if (!out) return bfail(nullptr, CNV_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!items || count <= 0) return bfail(nullptr, CNV_ERR_ARG, "empty batch");
    const cnv_params& p0 = items[0];
    for (int i = 1; i < count; ++i) {
        const cnv_params& p = items[i];
        if (p.precision != p0.precision || p.order != p0.order || p.poisson_type != p0.poisson_type ||
            p.device != p0.device || p.sweeps_per_launch != p0.sweeps_per_launch ||
            p.tile_waves != p0.tile_waves || p.wave_rows != p0.wave_rows || p.transport_fp32 != p0.transport_fp32)
            return bfail(nullptr, CNV_ERR_ARG,
                         "batch item %d: precision, order, poisson_type, device, launch shape and transport_fp32 "
                         "must match item 0", i);
    }
    for (int i = 0; i < count; ++i)
        if (items[i].warm_start)
            return bfail(nullptr, CNV_ERR_ARG, "batch item %d: warm_start is for one-cavity contexts only", i);
    cnv_batch* b = new cnv_batch();
    b->count = count;
    b->alive.assign(count, 1);
    b->blowup.resize(count);
    for (int i = 0; i < count; ++i) b->blowup[i] = items[i].blowup_speed;
    b->dc = p0.precision == CNV_PREC_FP32_DC;
    b->transport_fp32 = p0.transport_fp32 != 0;
    b->gs = p0.poisson_type == 1;
    b->device = p0.device;
    b->rtol = p0.inner_rtol;
    b->max_outer = p0.max_outer > 0 ? p0.max_outer : 50;
    // 2 per poll: a batch launch covers every cavity's tiles (tens of microseconds), each
    // solve's end wastes up to two chunks of them, and the first chunk already covers the
    // predicted length (the previous step's, per outer iteration for the refined form); the
    // 64-cavity workload measured 7% faster per step (FP32-refined) and 2% (FP64) than with 16
    b->launches_per_poll = p0.launches_per_poll > 0 ? p0.launches_per_poll : 2;
    auto setup = [&]() -> int {
        for (int i = 0; i < count; ++i) {
            cnv_ctx* c = nullptr;
            const int rc = cnv_create(&items[i], &c);
            if (rc) return bfail(b, rc, "batch item %d: %s", i, cnv_last_error(nullptr));
            b->ctx.push_back(c);
        }
        BHIP(b, hipSetDevice(b->device));
        BHIP(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        BHIP(b, hipMalloc(&b->d_ctl, sizeof(SorCtl) * 2 * count));
        BHIP(b, hipMemset(b->d_ctl, 0, sizeof(SorCtl) * 2 * count));
        BHIP(b, hipMalloc(&b->d_diag, sizeof(Diag) * count));
        BHIP(b, hipMemset(b->d_diag, 0, sizeof(Diag) * count));
        BHIP(b, hipMalloc(&b->d_dc, sizeof(DcSlot) * count));
        BHIP(b, hipHostMalloc(&b->h_ctl, sizeof(SorCtl) * 2 * count * kSlots, hipHostMallocDefault));
        BHIP(b, hipHostMalloc(&b->h_diag, sizeof(Diag) * count, hipHostMallocDefault));
        BHIP(b, hipHostMalloc(&b->h_dc, sizeof(DcSlot) * count, hipHostMallocDefault));
        for (int s = 0; s < kSlots; ++s) BHIP(b, hipEventCreateWithFlags(&b->ev[s], hipEventDisableTiming));
        BHIP(b, hipEventCreate(&b->t0));
        BHIP(b, hipEventCreate(&b->t1));
        std::vector<StepItem> it(count);
        for (int i = 0; i < count; ++i) {
            cnv_ctx* c = b->ctx[i];
            StepItem& I = it[i];
            memset(&I, 0, sizeof(I));
            I.w[0] = c->w[0];
            I.w[1] = c->w[1];
            I.psi[0] = c->psi[0];
            I.psi[1] = c->psi[1];
            I.psi[2] = c->psi[2];
            I.u = c->u;
            I.v = c->v;
            I.e32[0] = c->e32[0];
            I.e32[1] = c->e32[1];
            I.r32 = c->r32;
            I.ctl = b->d_ctl + 2 * i;
            I.diag = b->d_diag + i;
            I.st = c->st;
            I.bc = c->bc;
            I.n = c->n;
            I.ld = c->ld;
            I.rRe = 1. / c->p.Re;  // src/fluiddyn.c:98 "(1. / Re)"
            I.dt = c->p.dt;
            I.tol = c->p.tol;
            I.itmax = c->p.itmax;
            const double dx = c->p.dx, dy = c->p.dy;
            I.c1 = dy * dy;
            I.c2 = dx * dx;
            I.c3 = dx * dx * dy * dy;
            I.B = 2 * (dx * dx + dy * dy);
        }
        int rc = upload(b, &b->d_items, it);
        if (!rc) rc = build_maps(b);
        // the items' launch shape when they name one (cnv_create checked it), else (5, 16, 4):
        // a one-cavity default follows that cavity's size, a batch's spans its cavities
        const bool named = p0.sweeps_per_launch > 0 || p0.tile_waves > 0 || p0.wave_rows > 0;
        if (!rc) rc = build_sor(b, named ? b->ctx[0]->shape : TileShape{5, 16, 4});
        if (!rc) BHIP(b, hipDeviceSynchronize());
        return rc;
    };
    const int rc = setup();
    if (rc) {
        char keep[256];
        snprintf(keep, sizeof(keep), "%s", b->err);
        cnv_batch_destroy(b);
        snprintf(g_berr, sizeof(g_berr), "%s", keep);
        return rc;
    }
    *out = b;
    return CNV_OK;
}

int cnv_batch_step(cnv_batch* b, int nsteps, cnv_step_stats* stats) {
    if (!b || nsteps < 0) return bfail(b, CNV_ERR_ARG, "bad arguments");
    BHIP(b, hipSetDevice(b->device));
    for (int t = 0; t < nsteps; ++t) {
        const int rc = batch_one_step(b, stats ? stats + (size_t)t * b->count : nullptr);
        if (rc) return rc;
    }
    return CNV_OK;
}

int cnv_batch_step_timed(cnv_batch* b, int nsteps, cnv_step_stats* stats, double* elapsed_ms) {
    if (!b) return bfail(b, CNV_ERR_ARG, "null batch");
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipStreamSynchronize(b->stream));
    BHIP(b, hipEventRecord(b->t0, b->stream));
    const int rc = cnv_batch_step(b, nsteps, stats);
    BHIP(b, hipEventRecord(b->t1, b->stream));
    BHIP(b, hipEventSynchronize(b->t1));
    float ms = 0.f;
    BHIP(b, hipEventElapsedTime(&ms, b->t0, b->t1));
    if (elapsed_ms) *elapsed_ms = ms;
    return rc;
}

int cnv_batch_get_state(cnv_batch* b, int i, double* w, double* psi, double* u, double* v) {
    if (!b || i < 0 || i >= b->count) return bfail(b, CNV_ERR_ARG, "batch index %d out of range", i);
    BHIP(b, hipStreamSynchronize(b->stream));
    cnv_ctx* c = b->ctx[i];
    // the cavity's final iterate: psi[0] for defect correction, else the buffer its last slot names
    // (c->fin_slot and c->wcur: as the cavity's last step left them, also after it diverged)
    SorCtl sl{};
    BHIP(b, hipMemcpy(&sl, b->d_ctl + 2 * i + c->fin_slot, sizeof(SorCtl), hipMemcpyDeviceToHost));
    const double* psrc = b->dc ? c->psi[0] : c->psi[sl.parity];
    const size_t bytes = sizeof(double) * c->cells;
    if (psi) BHIP(b, hipMemcpy(psi, psrc, bytes, hipMemcpyDeviceToHost));
    if (w) BHIP(b, hipMemcpy(w, c->w[c->wcur], bytes, hipMemcpyDeviceToHost));
    if (u) BHIP(b, hipMemcpy(u, c->u, bytes, hipMemcpyDeviceToHost));
    if (v) BHIP(b, hipMemcpy(v, c->v, bytes, hipMemcpyDeviceToHost));
    return CNV_OK;
}

int cnv_batch_set_state(cnv_batch* b, int i, const double* w, const double* psi, const double* u, const double* v) {
    if (!b || i < 0 || i >= b->count) return bfail(b, CNV_ERR_ARG, "batch index %d out of range", i);
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipStreamSynchronize(b->stream));
    cnv_ctx* c = b->ctx[i];
    c->wcur = b->wcur;  // the batch's current w buffer
    const int rc = cnv_set_state(c, w, psi, u, v);
    if (rc) return bfail(b, rc, "batch item %d: %s", i, cnv_last_error(c));
    if (!b->alive[i]) {
        b->alive[i] = 1;
        b->maps_dirty = true;
    }
    return CNV_OK;
}

int cnv_batch_status(const cnv_batch* b, int* status) {
    if (!b) return -bfail(nullptr, CNV_ERR_ARG, "null batch");
    int n = 0;
    for (int i = 0; i < b->count; ++i) {
        if (status) status[i] = b->alive[i] ? CNV_OK : CNV_DIVERGED;
        n += b->alive[i];
    }
    return n;
}

int cnv_batch_autotune(cnv_batch* b, int launches, int* sweeps_per_launch, int* tile_waves, int* wave_rows,
                       double* ms_per_sweep) {
    if (!b || launches < 4) return bfail(b, CNV_ERR_ARG, "autotune needs at least 4 launches");
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipStreamSynchronize(b->stream));
    if (b->maps_dirty) {
        int rc = build_maps(b);
        if (!rc) rc = build_sor(b, b->shape);
        if (rc) return rc;
        b->maps_dirty = false;
    }
    if (cnv_batch_status(b, nullptr) == 0) return bfail(b, CNV_ERR_ARG, "autotune: every cavity has diverged");
    // the slots are overwritten below; a diverged cavity's (which say where its last iterate
    // is) are put back afterwards
    std::vector<SorCtl> ctl_keep((size_t)2 * b->count);
    BHIP(b, hipMemcpy(ctl_keep.data(), b->d_ctl, sizeof(SorCtl) * ctl_keep.size(), hipMemcpyDeviceToHost));
    const TileShape cands[] = {
#define CNV_SOR_LIST(s_, nw_, rb_) TileShape{s_, nw_, rb_},
        CNV_SOR_SHAPES(CNV_SOR_LIST)
#undef CNV_SOR_LIST
    };
    const TileShape keep = b->shape;
    const int wkeep = b->wcur;
    TileShape best = keep;
    double best_ms = 1e30;
    void** d_dst = nullptr;
    auto tune = [&]() -> int {
        // synthetic sources in the planes the solve reads: FP64 w[wcur] (negated by the solve) or
        // the FP32 residual r32
        std::vector<void*> dst(b->count);
        for (int i = 0; i < b->count; ++i)
            dst[i] = b->dc ? (void*)b->ctx[i]->r32 : (void*)b->ctx[i]->w[b->wcur ^ 1];
        BHIP(b, hipMalloc(&d_dst, sizeof(void*) * b->count));
        BHIP(b, hipMemcpy(d_dst, dst.data(), sizeof(void*) * b->count, hipMemcpyHostToDevice));
        if (b->dc)
            hipLaunchKernelGGL(batch_fill_source_kernel<float>, dim3(b->nb_lin), dim3(kBlock), 0, b->stream,
                               b->d_items, b->d_map_lin, (float**)d_dst);
        else
            hipLaunchKernelGGL(batch_fill_source_kernel<double>, dim3(b->nb_lin), dim3(kBlock), 0, b->stream,
                               b->d_items, b->d_map_lin, (double**)d_dst);
        BLAUNCH(b);
        b->wcur ^= 1;  // the solve reads w[wcur]: point it at the synthetic field
        // every launch sweeps S times: slots that no sweep can stop
        for (const TileShape& sh : cands) {
            int rc = build_sor(b, sh);
            if (rc) return rc;
            hipLaunchKernelGGL(batch_tune_slots_kernel, dim3((2 * b->count + kBlock - 1) / kBlock), dim3(kBlock), 0,
                               b->stream, b->d_ctl, 2 * b->count);
            BLAUNCH(b);
            for (int l = 0; l < 2; ++l) {
                rc = sor_launch(b, l, b->dc);
                if (rc) return rc;
            }
            BHIP(b, hipEventRecord(b->t0, b->stream));
            for (int l = 2; l < launches; ++l) {
                rc = sor_launch(b, l, b->dc);
                if (rc) return rc;
            }
            BHIP(b, hipEventRecord(b->t1, b->stream));
            BHIP(b, hipEventSynchronize(b->t1));
            float ms = 0.f;
            BHIP(b, hipEventElapsedTime(&ms, b->t0, b->t1));
            const double per_sweep = ms / ((launches - 2) * (double)sh.s);
            if (per_sweep < best_ms) {
                best_ms = per_sweep;
                best = sh;
            }
        }
        return CNV_OK;
    };
    const int rc_tune = tune();
    // on every exit: the batch's own w buffer, the chosen (or, after a failure, the kept) shape,
    // zeroed iterates of the live cavities, the diverged cavities' slots, no scratch left behind
    char msg[256];
    snprintf(msg, sizeof(msg), "%s", b->err);
    (void)hipStreamSynchronize(b->stream);
    b->wcur = wkeep;
    if (d_dst) (void)hipFree(d_dst);
    int rc = build_sor(b, rc_tune ? keep : best);
    if (!rc) {
        hipLaunchKernelGGL(batch_zero_kernel, dim3(b->nb_lin), dim3(kBlock), 0, b->stream, b->d_items, b->d_map_lin);
        BLAUNCH(b);
        for (int i = 0; i < b->count; ++i)
            if (!b->alive[i])
                BHIP(b, hipMemcpyAsync(b->d_ctl + 2 * i, ctl_keep.data() + 2 * i, 2 * sizeof(SorCtl),
                                       hipMemcpyHostToDevice, b->stream));
        BHIP(b, hipStreamSynchronize(b->stream));
    }
    b->k_prev = 0;
    if (rc_tune) return bfail(b, rc_tune, "%s", msg);
    if (rc) return rc;
    if (sweeps_per_launch) *sweeps_per_launch = best.s;
    if (tile_waves) *tile_waves = best.nw;
    if (wave_rows) *wave_rows = best.rb;
    if (ms_per_sweep) *ms_per_sweep = best_ms;
    return CNV_OK;
}
