// context.hip -- the device-resident C ABI (include/cnv.h): contexts, state, the time step loop.
//
// One context = one cavity on one GPU.  The step (src/main.c:276-451) is a fixed sequence of
// stream-ordered launches; only the Poisson solve has a data-dependent length:
//   wall vorticity (edges of w, in place)           src/main.c:283-320
//   transport: derivatives + explicit Euler, w -> w' (ping-pong)   src/main.c:323-344
//   Poisson: init (psi = 0, control block reset), then SOR launches queued in chunks; each
//            launch checks the device control block and turns into a no-op once the solve
//            has converged, so the host only polls a pinned copy of that block between chunks
//            and stops queueing                   src/main.c:347-363, src/poisson.c:224-285
//   velocity: u = DY psi, v = -(DX psi)            src/main.c:366-383
//   continuity max/min (optional diagnostic)       src/main.c:387-408
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "cnv.h"
#include "cnv_device.h"
#include "ctx.h"
#include "../host/cnv_host.h"

using namespace cnv;

namespace {

thread_local char g_err[256] = "";  // last error of this thread (contexts may live on different threads)

}  // namespace

int cnv_fail(cnv_ctx* c, int code, const char* fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) snprintf(c->err, sizeof(c->err), "%s", buf);
    snprintf(g_err, sizeof(g_err), "%s", buf);
    return code;
}
// The public functions below are declared extern "C" in cnv.h, which gives them C linkage.

double cnv_reference_beta(int nx, int ny) {
    const double pi = 3.14159265359;  // include/poisson.h:9 of the reference
    return 0.5 * (2 / (1 + sin(pi / (nx + 1))) + 2 / (1 + sin(pi / (ny + 1))));  // src/main.c:134
}

size_t cnv_params_size(void) { return sizeof(cnv_params); }

void cnv_default_params(cnv_params* p, int n) {
    memset(p, 0, sizeof(*p));
    p->n = n;
    p->order = 6;
    p->poisson_type = 2;
    p->precision = CNV_PREC_FP64;
    p->itmax = 10000;
    p->device = 0;
    p->dx = 1.0 / n;
    p->dy = 1.0 / n;
    p->Re = 1000.0;
    p->dt = 0.005;
    p->beta = cnv_reference_beta(n, n);
    p->tol = 1e-3;
    p->u4 = 1.0;
    p->inner_rtol = 1e-5;
    p->max_outer = 50;
    p->transport_fp32 = 1;
    p->diagnostics = 1;
}

const char* cnv_status_string(int s) {
    switch (s) {
        case CNV_OK: return "ok";
        case CNV_ERR_ARG: return "invalid argument";
        case CNV_ERR_HIP: return "HIP runtime error";
        case CNV_NOT_CONVERGED: return "Poisson solve did not converge within poisson_max_it";
        case CNV_ERR_NOMEM: return "out of memory";
        case CNV_ERR_NO_DEVICE: return "no HIP device";
        case CNV_DIVERGED: return "diverged (non-finite Poisson update, or blow-up speed exceeded)";
        default: return "unknown status";
    }
}

const char* cnv_last_error(const cnv_ctx* c) { return c ? c->err : g_err; }

int cnv_device_name(int device, char* buf, int len) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return cnv_fail(nullptr, CNV_ERR_NO_DEVICE, "no device %d", device);
    // some drivers report no marketing name: name the ordinal instead
    if (prop.name[0]) snprintf(buf, len, "%s (%s)", prop.name, prop.gcnArchName);
    else snprintf(buf, len, "HIP device %d (%s)", device, prop.gcnArchName);
    return CNV_OK;
}

// The default shape for an n x n grid on a device with `cus` compute units: the first of these
// tiles whose count fits one workgroup per CU -- thin tiles first (on a small grid the GPU is
// otherwise idle), then six sweeps per launch on the 16-wave tile, then five.  Measured per
// sweep, FP64: 128^2 and 256^2 best with (5, 4, 6), 512^2 with (5, 8, 4), 768^2 with (6, 16, 4)
// (2.21 us against 2.42 us with (5, 16, 4)), 1024^2 with (5, 16, 4) ((6, 16, 4) needs 260 tiles
// there).  Grids beyond take (5, 16, 4), the thickest tile, which has the least halo.
TileShape cnv_auto_shape(int n, int cus) {
    const TileShape cand[] = {{5, 4, 6}, {5, 8, 4}, {6, 16, 4}, {5, 16, 4}};
    for (const TileShape& t : cand) {
        const long long tiles = (long long)((n + tile_rows(t) - 1) / tile_rows(t)) * ((n + tile_cols(t) - 1) / tile_cols(t));
        if (tiles <= cus) return t;
    }
    return cand[3];
}

// Default launch shape per sweeps-per-launch (tuned on MI355X at 1024^2; 0 = by grid size).
static TileShape shape_for(int spl) {
    switch (spl) {
        case 0: return TileShape{5, 16, 4};
        case 1: return TileShape{1, 4, 10};
        case 2: return TileShape{2, 4, 10};
        case 4: return TileShape{4, 8, 8};
        case 5: return TileShape{5, 16, 4};
        case 6: return TileShape{6, 8, 8};
        default: return TileShape{spl, 8, 10};
    }
}

static int upload_tables(cnv_ctx* c) {
    const int n = c->n;
    const int o = c->p.order;
    std::vector<double> D((size_t)n * n);
    std::vector<int> start(n), start2(n);
    std::vector<double> tab((size_t)4 * n * kStencilWidth);
    const double hs[4] = {c->p.dx, c->p.dy, c->p.dx, c->p.dy};
    const int which[4] = {1, 1, 2, 2};
    for (int t = 0; t < 4; ++t) {
        int rc = cnv_fd_dense(which[t], n, o, hs[t], D.data());
        if (rc) return cnv_fail(c, CNV_ERR_ARG, "** Error: valid orders are 2, 4 or 6 ** (order=%d, n=%d)", o, n);
        if (cnv_fd_windows(n, D.data(), t == 0 ? start.data() : start2.data(), tab.data() + (size_t)t * n * kStencilWidth))
            return cnv_fail(c, CNV_ERR_ARG, "stencil window does not cover a row (n=%d)", n);
    }
    std::vector<float> ftab(tab.size());
    for (size_t q = 0; q < tab.size(); ++q) ftab[q] = (float)tab[q];
    HIPC(c, hipMalloc(&c->d_start, sizeof(int) * n));
    HIPC(c, hipMalloc(&c->d_tab, sizeof(double) * tab.size()));
    HIPC(c, hipMalloc(&c->d_ftab, sizeof(float) * ftab.size()));
    HIPC(c, hipMemcpy(c->d_start, start.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->d_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->d_ftab, ftab.data(), sizeof(float) * ftab.size(), hipMemcpyHostToDevice));
    const size_t stride = (size_t)n * kStencilWidth;
    c->st.start = c->d_start;
    c->st.d1x = c->d_tab;
    c->st.d1y = c->d_tab + stride;
    c->st.d2x = c->d_tab + 2 * stride;
    c->st.d2y = c->d_tab + 3 * stride;
    c->st.f1x = c->d_ftab;
    c->st.f1y = c->d_ftab + stride;
    c->st.f2x = c->d_ftab + 2 * stride;
    c->st.f2y = c->d_ftab + 3 * stride;
    return CNV_OK;
}

int cnv_set_state(cnv_ctx* c, const double* w, const double* psi, const double* u, const double* v) {
    if (!c) return cnv_fail(nullptr, CNV_ERR_ARG, "null context");
    const size_t bytes = sizeof(double) * c->cells;
    if (w) HIPC(c, hipMemcpyAsync(c->w[c->wcur], w, bytes, hipMemcpyHostToDevice, c->stream));
    if (psi) {
        // psi lives in psi[parity]; make every plane hold it so that any parity reads it
        for (int b = 0; b < 3; ++b) HIPC(c, hipMemcpyAsync(c->psi[b], psi, bytes, hipMemcpyHostToDevice, c->stream));
    }
    if (u) HIPC(c, hipMemcpyAsync(c->u, u, bytes, hipMemcpyHostToDevice, c->stream));
    if (v) HIPC(c, hipMemcpyAsync(c->v, v, bytes, hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return CNV_OK;
}

static int current_psi(cnv_ctx* c, const double** out) {
    if (c->p.precision == CNV_PREC_FP32_DC) {
        *out = c->psi[0];
        return CNV_OK;
    }
    HIPC(c, hipMemcpyAsync(c->h_ctl, c->ctl + c->fin_slot, sizeof(SorCtl), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    *out = c->psi[c->h_ctl[0].parity];  // the plane of the final iterate
    return CNV_OK;
}

int cnv_get_state(cnv_ctx* c, double* w, double* psi, double* u, double* v) {
    if (!c) return cnv_fail(nullptr, CNV_ERR_ARG, "null context");
    const size_t bytes = sizeof(double) * c->cells;
    if (psi) {
        const double* src = nullptr;
        int rc = current_psi(c, &src);
        if (rc) return rc;
        HIPC(c, hipMemcpyAsync(psi, src, bytes, hipMemcpyDeviceToHost, c->stream));
    }
    if (w) HIPC(c, hipMemcpyAsync(w, c->w[c->wcur], bytes, hipMemcpyDeviceToHost, c->stream));
    if (u) HIPC(c, hipMemcpyAsync(u, c->u, bytes, hipMemcpyDeviceToHost, c->stream));
    if (v) HIPC(c, hipMemcpyAsync(v, c->v, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return CNV_OK;
}

int cnv_snapshot_begin(cnv_ctx* c) {
    if (!c) return cnv_fail(nullptr, CNV_ERR_ARG, "null context");
    if (hipSetDevice(c->p.device) != hipSuccess) return cnv_fail(c, CNV_ERR_HIP, "hipSetDevice failed");
    const size_t bytes = sizeof(double) * c->cells * 4;
    if (!c->d_snap) {  // first snapshot: its buffers, stream and events
        HIPC(c, hipMalloc(&c->d_snap, bytes));
        HIPC(c, hipHostMalloc(&c->h_snap, bytes, hipHostMallocDefault));
        HIPC(c, hipStreamCreateWithFlags(&c->snap_stream, hipStreamNonBlocking));
        HIPC(c, hipEventCreateWithFlags(&c->ev_snap_ready, hipEventDisableTiming));
        HIPC(c, hipEventCreateWithFlags(&c->ev_snap_done, hipEventDisableTiming));
    }
    // the previous snapshot's transfer reads the device copy that is about to be overwritten
    if (c->snap_begun) HIPC(c, hipEventSynchronize(c->ev_snap_done));
    LAUNCH(c, snapshot_launch(c->w[c->wcur], planes_of(c), c->u, c->v, c->ctl + c->fin_slot,
                              c->d_snap, (long long)c->cells, c->stream));
    HIPC(c, hipEventRecord(c->ev_snap_ready, c->stream));
    HIPC(c, hipStreamWaitEvent(c->snap_stream, c->ev_snap_ready, 0));
    HIPC(c, hipMemcpyAsync(c->h_snap, c->d_snap, bytes, hipMemcpyDeviceToHost, c->snap_stream));
    HIPC(c, hipEventRecord(c->ev_snap_done, c->snap_stream));
    c->snap_begun = true;
    return CNV_OK;
}

int cnv_snapshot_wait(cnv_ctx* c, const double** w, const double** psi, const double** u, const double** v) {
    // may run on another thread while c steps: it touches only the snapshot's event and buffer,
    // and reports failures through this thread's message only
    if (!c || !c->snap_begun) return cnv_fail(nullptr, CNV_ERR_ARG, "no snapshot was begun on this context");
    const hipError_t e = hipEventSynchronize(c->ev_snap_done);
    if (e != hipSuccess) return cnv_fail(nullptr, CNV_ERR_HIP, "snapshot transfer failed: %s", hipGetErrorString(e));
    const double* h = c->h_snap;
    if (w) *w = h;
    if (psi) *psi = h + c->cells;
    if (u) *u = h + 2 * c->cells;
    if (v) *v = h + 3 * c->cells;
    return CNV_OK;
}

int cnv_device_state(cnv_ctx* c, double** w, double** psi, double** u, double** v) {
    if (!c) return cnv_fail(nullptr, CNV_ERR_ARG, "null context");
    if (psi) {
        const double* src = nullptr;
        int rc = current_psi(c, &src);
        if (rc) return rc;
        *psi = const_cast<double*>(src);
    }
    if (w) *w = c->w[c->wcur];
    if (u) *u = c->u;
    if (v) *v = c->v;
    return CNV_OK;
}

static int initial_state(cnv_ctx* c) {
    // w = psi = 0; u = ui, v = vi on the interior and 0 on the edges (src/main.c:178-221)
    std::vector<double> a(c->cells, 0.0), uu(c->cells, 0.0), vv(c->cells, 0.0);
    const int n = c->n;
    for (int i = 1; i < n - 1; ++i)
        for (int j = 1; j < n - 1; ++j) {
            uu[(size_t)i * n + j] = c->p.ui;
            vv[(size_t)i * n + j] = c->p.vi;
        }
    return cnv_set_state(c, a.data(), a.data(), uu.data(), vv.data());
}

static void drop_graphs(cnv_ctx* c) {
    hipGraphExec_t* gs[] = {&c->g_pre[0],  &c->g_pre[1],   &c->g_chunk[0], &c->g_chunk[1], &c->g_tail[0],
                            &c->g_tail[1], &c->g_resid[0], &c->g_resid[1], &c->g_correct,  &c->g_post};
    for (hipGraphExec_t* g : gs) {
        if (*g) (void)hipGraphExecDestroy(*g);
        *g = nullptr;
    }
}

void cnv_destroy(cnv_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->snap_stream) (void)hipStreamSynchronize(c->snap_stream);
    drop_graphs(c);
    void* dev[] = {c->w[0], c->w[1], c->psi[0], c->psi[1], c->psi[2], c->u, c->v, c->d_start, c->d_tab, c->d_ftab,
                   c->ctl, c->diag, c->partial, c->e32[0], c->e32[1], c->r32, c->d_budget, c->d_snap};
    for (void* p : dev)
        if (p) (void)hipFree(p);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    if (c->h_diag) (void)hipHostFree(c->h_diag);
    if (c->h_snap) (void)hipHostFree(c->h_snap);
    if (c->ev_snap_ready) (void)hipEventDestroy(c->ev_snap_ready);
    if (c->ev_snap_done) (void)hipEventDestroy(c->ev_snap_done);
    if (c->snap_stream) (void)hipStreamDestroy(c->snap_stream);
    for (int s = 0; s < kSlots; ++s)
        if (c->ev[s]) (void)hipEventDestroy(c->ev[s]);
    if (c->t0) (void)hipEventDestroy(c->t0);
    if (c->t1) (void)hipEventDestroy(c->t1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int cnv_create(const cnv_params* p, cnv_ctx** out) {
    if (!out) return cnv_fail(nullptr, CNV_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!p) return cnv_fail(nullptr, CNV_ERR_ARG, "null params");
    if (p->n < 8) return cnv_fail(nullptr, CNV_ERR_ARG, "grid side n=%d: need n >= 8", p->n);
    if (p->order != 2 && p->order != 4 && p->order != 6)
        return cnv_fail(nullptr, CNV_ERR_ARG, "** Error: valid orders are 2, 4 or 6 ** (order=%d)", p->order);
    if (p->poisson_type != 1 && p->poisson_type != 2)
        return cnv_fail(nullptr, CNV_ERR_ARG, "Error - invalid option for Poisson solver (poisson_type=%d)", p->poisson_type);
    if (p->precision != CNV_PREC_FP64 && p->precision != CNV_PREC_FP32_DC)
        return cnv_fail(nullptr, CNV_ERR_ARG, "invalid precision %d", p->precision);
    if (!(p->dx > 0) || !(p->dy > 0)) return cnv_fail(nullptr, CNV_ERR_ARG, "grid spacing must be positive");
    if (p->inner_test_every < 0 || p->inner_test_every > 64 || (p->inner_test_every & (p->inner_test_every - 1)))
        return cnv_fail(nullptr, CNV_ERR_ARG, "inner_test_every=%d: need 0 (default) or a power of two up to 64",
                        p->inner_test_every);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= p->device || p->device < 0)
        return cnv_fail(nullptr, CNV_ERR_NO_DEVICE, "HIP device %d not available (%d devices)", p->device, ndev);
    if (hipSetDevice(p->device) != hipSuccess) return cnv_fail(nullptr, CNV_ERR_NO_DEVICE, "hipSetDevice(%d) failed", p->device);

    cnv_ctx* c = new cnv_ctx();
    c->p = *p;
    c->n = p->n;
    c->ld = p->n;
    c->cells = (size_t)p->n * p->n;
    c->bc = WallBC{p->u1, p->u2, p->u3, p->u4, p->v1, p->v2, p->v3, p->v4};
    const int spl = p->sweeps_per_launch;
    c->shape = shape_for(spl);
    if (spl == 0 && p->tile_waves <= 0 && p->wave_rows <= 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess || cus <= 0)
            cus = 256;
        c->shape = cnv_auto_shape(p->n, cus);
    }
    if (p->tile_waves > 0) c->shape.nw = p->tile_waves;
    if (p->wave_rows > 0) c->shape.rb = p->wave_rows;
    if ((p->sweeps_per_launch && c->shape.s != p->sweeps_per_launch) || !sor_shape_supported(c->shape)) {
        cnv_destroy(c);
        return cnv_fail(nullptr, CNV_ERR_ARG, "unsupported SOR launch shape (sweeps_per_launch %d, waves %d, rows per wave %d)",
                    spl, p->tile_waves, p->wave_rows);
    }
    // captured chunks of 66 launches: each replay costs a graph launch and a poll copy in the
    // stream, and a solve's end up to two chunks of no-op launches; at 1024^2 chunks of 64
    // measured 0.6% faster per step than 32 or 128 (the chunk length is a multiple of 6, see
    // build_graphs).  Eager solves choose their chunk per solve (run_sor).
    c->launches_per_poll = p->launches_per_poll > 0 ? p->launches_per_poll : 66;
    c->tail_chunk = p->precision == CNV_PREC_FP32_DC ? std::lcm(6, inner_test_every(*p, kFp32TestEvery)) : 6;
    c->poll_fixed = p->launches_per_poll > 0 ? p->launches_per_poll : 0;
    c->tiles_i = (c->n + tile_rows(c->shape) - 1) / tile_rows(c->shape);
    c->tiles_j = (c->n + tile_cols(c->shape) - 1) / tile_cols(c->shape);
    auto setup = [&]() -> int {
        HIPC(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        const size_t bytes = sizeof(double) * c->cells;
        for (int b = 0; b < 2; ++b) HIPC(c, hipMalloc(&c->w[b], bytes));
        for (int b = 0; b < 3; ++b) HIPC(c, hipMalloc(&c->psi[b], bytes));
        HIPC(c, hipMalloc(&c->u, bytes));
        HIPC(c, hipMalloc(&c->v, bytes));
        HIPC(c, hipMalloc(&c->ctl, 2 * sizeof(SorCtl)));
        HIPC(c, hipMalloc(&c->diag, sizeof(Diag)));
        HIPC(c, hipMemset(c->ctl, 0, 2 * sizeof(SorCtl)));
        HIPC(c, hipMemset(c->diag, 0, sizeof(Diag)));
        HIPC(c, hipMalloc(&c->partial, sizeof(double) * 2 * 2 * kMaxSweepsPerLaunch * (size_t)c->tiles_i * c->tiles_j));
        HIPC(c, hipHostMalloc(&c->h_ctl, sizeof(SorCtl) * kSlots, hipHostMallocDefault));
        HIPC(c, hipHostMalloc(&c->h_diag, sizeof(Diag), hipHostMallocDefault));
        for (int s = 0; s < kSlots; ++s) HIPC(c, hipEventCreateWithFlags(&c->ev[s], hipEventDisableTiming));
        HIPC(c, hipEventCreate(&c->t0));
        HIPC(c, hipEventCreate(&c->t1));
        if (p->precision == CNV_PREC_FP32_DC) {
            for (int b = 0; b < 2; ++b) HIPC(c, hipMalloc(&c->e32[b], sizeof(float) * c->cells));
            HIPC(c, hipMalloc(&c->r32, sizeof(float) * c->cells));
            HIPC(c, hipMalloc(&c->d_budget, sizeof(int)));
        }
        int rc = upload_tables(c);
        if (rc) return rc;
        return initial_state(c);
    };
    int rc = setup();
    if (rc) {
        char keep[256];
        snprintf(keep, sizeof(keep), "%s", c->err);
        cnv_destroy(c);
        snprintf(g_err, sizeof(g_err), "%s", keep);
        return rc;
    }
    *out = c;
    return CNV_OK;
}

int cnv_sweeps_per_launch(const cnv_ctx* c) { return c ? c->shape.s : 0; }

// ---- SOR driver ---------------------------------------------------------------------------------
static int poisson_fp64(cnv_ctx* c, const double* f_src, int negate, SorCtl* fin) {
    if (c->p.warm_start)
        LAUNCH(c, warm_start_launch(c->psi[0], planes_of(c), c->ctl + c->fin_slot, c->n, c->ld, c->stream));
    LAUNCH(c, poisson_init_launch(c->psi[0], c->ctl, c->diag, c->n, c->ld, c->p.tol, c->p.itmax, c->stream,
                                  !c->p.warm_start));
    SorArgs<double> a = sor_args64(c, f_src, negate);
    int rc = run_sor<double>(c, a, fin);
    if (rc) return rc;
    c->k_prev = fin->k;
    return CNV_OK;
}

// defect correction: psi[0] accumulates in FP64, e32 ping-pong holds the inner FP32 iterates
static int poisson_dc(cnv_ctx* c, const double* w_src, SorCtl* fin, int* outer) {
    // psi[0] is the FP64 accumulator: a warm start keeps the previous step's final iterate there
    LAUNCH(c, poisson_init_launch(c->psi[0], c->ctl, c->diag, c->n, c->ld, c->p.tol, c->p.itmax, c->stream,
                                  !c->p.warm_start));
    const double dx = c->p.dx, dy = c->p.dy;
    const double c1 = dy * dy, c2 = dx * dx, c3 = dx * dx * dy * dy, B = 2 * (dx * dx + dy * dy);
    SorArgs<float> a = sor_args32(c);
    int total = 0;
    const int max_outer = c->p.max_outer > 0 ? c->p.max_outer : 50;
    *fin = SorCtl{};
    fin->status = SOR_FAILED;
    for (int o = 0; o < max_outer && total < c->p.itmax; ++o) {
        LAUNCH(c, dc_residual_launch(w_src, c->psi[0], c->r32, c->n, c->ld, c1, c2, c3, B, c->stream));
        LAUNCH(c, dc_inner_init_launch(c->e32[0], c->ctl, c->n, c->ld, c->p.inner_rtol, c->p.tol, c->p.itmax - total,
                                       c->stream));
        SorCtl snap{};
        c->k_prev = o < cnv_ctx::kDcPred ? c->k_dc[o] : 0;  // this outer iteration's last length
        int rc = run_sor<float>(c, a, &snap);
        if (rc) return rc;
        if (o < cnv_ctx::kDcPred) c->k_dc[o] = snap.k;
        LAUNCH(c, dc_correct_launch(c->psi[0], c->e32[0], c->e32[1], c->ctl + c->fin_slot, c->n, c->ld, c->stream));
        total += snap.k;
        *outer = o + 1;
        *fin = snap;
        fin->k = total;
        if (dc_final(snap.status)) return CNV_OK;
    }
    fin->status = SOR_FAILED;
    return CNV_OK;
}

// ---- captured step (use_graph) -------------------------------------------------------------
// Launch index L only matters through L mod 2 (state slots, partial sums), L mod 3 (FP64
// iterate planes) and L mod 4 (the FP32 inner solves' stopping tests), and every solve's state
// lives in the device slots, so a captured chunk of M (a multiple of tail_chunk: 6 FP64, 12
// FP32) SOR launches with L = 0..M-1 is replayed until the solve ends: each replay continues the
// previous one's rotation, and after whole chunks the final slot is 1.  Two chunk sizes:
// launches_per_poll for the predicted length of the solve and tail_chunk for its end, so that
// the no-op launches queued after convergence stay few.
static int capture(cnv_ctx* c, hipGraphExec_t* out, int (*body)(cnv_ctx*, int), int arg) {
    hipGraph_t g = nullptr;
    HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = body(c, arg);
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return cnv_fail(c, CNV_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return cnv_fail(c, CNV_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
    return CNV_OK;
}

static bool is_dc(const cnv_ctx* c) { return c->p.precision == CNV_PREC_FP32_DC; }
// the continuity kernel also finds the largest |u| or |v|, which the blow-up check reads
static bool wants_diag(const cnv_ctx* c) { return c->p.diagnostics || c->p.blowup_speed > 0; }

// wall vorticity and transport of w[wc] into w[wc^1], then a fresh Poisson solve
static int body_pre(cnv_ctx* c, int wc) {
    const int n = c->n, ld = c->ld;
    LAUNCH(c, wall_vorticity_launch(c->w[wc], c->u, c->v, n, ld, c->st, c->bc, c->stream));
    if (is_dc(c)) {
        if (c->p.transport_fp32)
            LAUNCH(c, transport_fp32_launch(c->w[wc], c->w[wc ^ 1], c->u, c->v, n, ld, c->st, c->bc, 1. / c->p.Re,
                                            c->p.dt, c->stream));
        else
            LAUNCH(c, transport_fp64_launch(c->w[wc], c->w[wc ^ 1], c->u, c->v, n, ld, c->st, c->bc, 1. / c->p.Re,
                                            c->p.dt, c->stream));
        // psi[0] is the FP64 accumulator (a warm start keeps it); the solve's sweep budget
        LAUNCH(c, poisson_init_launch(c->psi[0], c->ctl, c->diag, n, ld, c->p.tol, c->p.itmax, c->stream,
                                      !c->p.warm_start));
        LAUNCH(c, dc_budget_launch(c->d_budget, c->p.itmax, nullptr, c->stream));
        return CNV_OK;
    }
    LAUNCH(c, transport_fp64_launch(c->w[wc], c->w[wc ^ 1], c->u, c->v, n, ld, c->st, c->bc, 1. / c->p.Re, c->p.dt,
                                    c->stream));
    // warm start: whole chunks end on slot 1, which holds the previous solve's final state
    if (c->p.warm_start) LAUNCH(c, warm_start_launch(c->psi[0], planes_of(c), c->ctl + 1, n, ld, c->stream));
    LAUNCH(c, poisson_init_launch(c->psi[0], c->ctl, c->diag, n, ld, c->p.tol, c->p.itmax, c->stream,
                                  !c->p.warm_start));
    return CNV_OK;
}
// m SOR launches, L = 0..m-1: FP64 on the source w[src] (f = -w), or the refined form's FP32
// inner solve on r32
static int body_chunk_of(cnv_ctx* c, int src, int m) {
    const bool gs = c->p.poisson_type == 1;
    if (is_dc(c)) {
        const SorArgs<float> a = sor_args32(c);
        for (int l = 0; l < m; ++l)
            if (launch_sor(c, a, l, gs)) return cnv_fail(c, CNV_ERR_HIP, "SOR launch failed while capturing");
        return CNV_OK;
    }
    const SorArgs<double> a = sor_args64(c, c->w[src], 1);
    for (int l = 0; l < m; ++l)
        if (launch_sor(c, a, l, gs)) return cnv_fail(c, CNV_ERR_HIP, "SOR launch failed while capturing");
    return CNV_OK;
}
static int body_chunk(cnv_ctx* c, int src) { return body_chunk_of(c, src, c->launches_per_poll); }
static int body_tail(cnv_ctx* c, int src) { return body_chunk_of(c, src, c->tail_chunk); }
// refined form, one outer iteration: FP64 residual of psi[0] against w[src], inner solve init
// with the remaining budget (device counter) ...
static int body_resid(cnv_ctx* c, int src) {
    const double dx = c->p.dx, dy = c->p.dy;
    LAUNCH(c, dc_residual_launch(c->w[src], c->psi[0], c->r32, c->n, c->ld, dy * dy, dx * dx,
                                 dx * dx * dy * dy, 2 * (dx * dx + dy * dy), c->stream));
    LAUNCH(c, dc_inner_init_launch(c->e32[0], c->ctl, c->n, c->ld, c->p.inner_rtol, c->p.tol, 0, c->stream,
                                   c->d_budget));
    return CNV_OK;
}
// ... and after the inner solve (final slot 1): psi[0] += e, budget -= its sweeps
static int body_correct(cnv_ctx* c, int) {
    LAUNCH(c, dc_correct_launch(c->psi[0], c->e32[0], c->e32[1], c->ctl + 1, c->n, c->ld, c->stream));
    LAUNCH(c, dc_budget_launch(c->d_budget, 0, c->ctl + 1, c->stream));
    return CNV_OK;
}
// velocities from the final iterate (slot 1 after whole chunks; psi[0] in the refined form),
// continuity diagnostic
static int body_post(cnv_ctx* c, int) {
    LAUNCH(c, velocities(c, c->ctl + 1));
    if (wants_diag(c)) LAUNCH(c, continuity_launch(c->u, c->v, c->diag, c->n, c->ld, c->st, c->stream));
    return CNV_OK;
}

static int build_graphs(cnv_ctx* c) {
    // whole chunks end on slot 1 and on the launch index's phase (plane rotation, stopping tests)
    const int u = c->tail_chunk;
    c->launches_per_poll = (c->launches_per_poll + u - 1) / u * u;
    const int nsrc = is_dc(c) ? 1 : 2;  // the refined form's inner solves always read r32
    for (int wc = 0; wc < 2; ++wc) {
        int rc = capture(c, &c->g_pre[wc], body_pre, wc);
        if (!rc && wc < nsrc) rc = capture(c, &c->g_chunk[wc], body_chunk, wc);
        if (!rc && wc < nsrc) rc = capture(c, &c->g_tail[wc], body_tail, wc);
        if (!rc && is_dc(c)) rc = capture(c, &c->g_resid[wc], body_resid, wc);
        if (rc) return rc;
    }
    if (is_dc(c)) {
        const int rc = capture(c, &c->g_correct, body_correct, 0);
        if (rc) return rc;
    }
    return capture(c, &c->g_post, body_post, 0);
}

// Replays chunk graphs until a polled slot is terminal.  With a prediction of P launches, big
// chunks and then tail chunks up to about P - 2*tail_chunk launches go before the first poll, and
// tail chunks follow, as many per poll as make up the eager solve's chunk for that prediction
// (poll_chunk); without one, big chunks one per poll.
static int replay_solve(cnv_ctx* c, int src, int predicted_sweeps, SorCtl* final_ctl) {
    const int M = c->launches_per_poll, Mt = c->tail_chunk;
    const int S = c->shape.s;
    const int P = predicted_sweeps > 0 ? predicted_sweeps / S : 0;
    const long long cap = (long long)c->p.itmax / S + 4LL * M + 8;
    const int first = P > 0 ? std::max(0, (P - 2 * Mt) / M) : 1;  // big chunks before the first poll
    const int extra = P > 0 ? std::max(0, (P - 2 * Mt - first * M) / Mt) : 0;  // and tail chunks
    const int tails = P > 0 ? std::max(1, (poll_chunk(c, P) + Mt / 2) / Mt) : 0;  // tail chunks per poll
    long long queued = 0;
    int slot = 0, round = 0;
    int pending[kSlots];
    int npend = 0;
    c->fin_slot = 1;
    while (true) {
        if (round == 0 && first + extra > 0) {
            for (int k = 0; k < first; ++k) HIPC(c, hipGraphLaunch(c->g_chunk[src], c->stream));
            for (int k = 0; k < extra; ++k) HIPC(c, hipGraphLaunch(c->g_tail[src], c->stream));
            queued += (long long)first * M + (long long)extra * Mt;
        } else if (P > 0) {
            for (int k = 0; k < tails; ++k) HIPC(c, hipGraphLaunch(c->g_tail[src], c->stream));
            queued += (long long)tails * Mt;
        } else {
            HIPC(c, hipGraphLaunch(c->g_chunk[src], c->stream));
            queued += M;
        }
        ++round;
        HIPC(c, hipMemcpyAsync(&c->h_ctl[slot], c->ctl + 1, sizeof(SorCtl), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipEventRecord(c->ev[slot], c->stream));
        pending[npend++] = slot;
        slot = (slot + 1) % kSlots;
        if (npend >= 2) {
            const int s0 = pending[0];
            for (int q = 1; q < npend; ++q) pending[q - 1] = pending[q];
            --npend;
            HIPC(c, hipEventSynchronize(c->ev[s0]));
            const SorCtl snap = c->h_ctl[s0];
            if (sor_done(snap.status)) {
                *final_ctl = snap;
                return CNV_OK;
            }
        }
        if (queued > cap + (long long)first * M)
            return cnv_fail(c, CNV_ERR_HIP, "SOR state did not terminate after %lld launches", queued);
    }
}

// the refined form's solve from captured graphs: outer iterations driven by the host, each a
// residual graph, replayed inner chunks and a correction graph (as poisson_dc)
static int poisson_dc_graph(cnv_ctx* c, int src, SorCtl* fin, int* outer) {
    int total = 0;
    const int max_outer = c->p.max_outer > 0 ? c->p.max_outer : 50;
    *fin = SorCtl{};
    fin->status = SOR_FAILED;
    for (int o = 0; o < max_outer && total < c->p.itmax; ++o) {
        HIPC(c, hipGraphLaunch(c->g_resid[src], c->stream));
        SorCtl snap{};
        int rc = replay_solve(c, 0, o < cnv_ctx::kDcPred ? c->k_dc[o] : 0, &snap);
        if (rc) return rc;
        if (o < cnv_ctx::kDcPred) c->k_dc[o] = snap.k;
        HIPC(c, hipGraphLaunch(c->g_correct, c->stream));
        total += snap.k;
        *outer = o + 1;
        *fin = snap;
        fin->k = total;
        if (dc_final(snap.status)) return CNV_OK;
    }
    fin->status = SOR_FAILED;
    return CNV_OK;
}

static int one_step(cnv_ctx* c, cnv_step_stats* stats);

// A step's statistics from its final solve slot and the diagnostics; the blow-up check.
static int finish_step(cnv_ctx* c, const SorCtl& fin, int outer, cnv_step_stats* stats) {
    cnv_step_stats s{};
    s.status = solve_status(fin.status);
    s.sweeps = fin.k;
    s.k = fin.k - 1;
    s.e = fin.e_last;
    s.outer = outer;
    double speed = 0.0;
    if (wants_diag(c)) {
        HIPC(c, hipMemcpyAsync(c->h_diag, c->diag, sizeof(Diag), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        s.cont_max = dkey_inv(c->h_diag->kmax);
        s.cont_min = dkey_inv(c->h_diag->kmin);
        speed = dkey_inv(c->h_diag->kspd);
    }
    if (s.status == CNV_OK && c->p.blowup_speed > 0 && !(speed <= c->p.blowup_speed)) s.status = CNV_DIVERGED;
    if (stats) *stats = s;
    if (s.status == CNV_DIVERGED) {
        if (fin.status == SOR_DIVERGED)
            return cnv_fail(c, CNV_DIVERGED, "Poisson solve diverged: non-finite L1 update after %d sweeps", fin.k);
        return cnv_fail(c, CNV_DIVERGED, "blow-up: largest |u| or |v| %g exceeds blowup_speed %g", speed,
                        c->p.blowup_speed);
    }
    if (s.status != CNV_OK)
        return cnv_fail(c, CNV_NOT_CONVERGED, "Error: maximum number of iterations achieved for Poisson equation.");
    return CNV_OK;
}

static int one_step_graph(cnv_ctx* c, cnv_step_stats* stats) {
    if (!c->g_post) {
        const int rc = build_graphs(c);
        if (rc) {
            drop_graphs(c);
            return rc;
        }
    }
    HIPC(c, hipGraphLaunch(c->g_pre[c->wcur], c->stream));
    c->wcur ^= 1;
    SorCtl fin{};
    int outer = 0;
    int rc = is_dc(c) ? poisson_dc_graph(c, c->wcur, &fin, &outer) : replay_solve(c, c->wcur, c->k_prev, &fin);
    if (rc) return rc;
    if (!is_dc(c)) c->k_prev = fin.k;
    HIPC(c, hipGraphLaunch(c->g_post, c->stream));
    return finish_step(c, fin, outer, stats);
}

static int one_step(cnv_ctx* c, cnv_step_stats* stats) {
    if (c->p.use_graph) return one_step_graph(c, stats);
    const int n = c->n, ld = c->ld;
    const double rRe = 1. / c->p.Re;  // src/fluiddyn.c:98 "(1. / Re)"
    double* w_in = c->w[c->wcur];
    double* w_out = c->w[c->wcur ^ 1];
    const bool dc = c->p.precision == CNV_PREC_FP32_DC;
    LAUNCH(c, wall_vorticity_launch(w_in, c->u, c->v, n, ld, c->st, c->bc, c->stream));
    if (dc && c->p.transport_fp32)
        LAUNCH(c, transport_fp32_launch(w_in, w_out, c->u, c->v, n, ld, c->st, c->bc, rRe, c->p.dt, c->stream));
    else
        LAUNCH(c, transport_fp64_launch(w_in, w_out, c->u, c->v, n, ld, c->st, c->bc, rRe, c->p.dt, c->stream));
    c->wcur ^= 1;
    SorCtl fin{};
    int outer = 0;
    int rc = dc ? poisson_dc(c, w_out, &fin, &outer) : poisson_fp64(c, w_out, 1, &fin);
    if (rc) return rc;
    LAUNCH(c, velocities(c, c->ctl + c->fin_slot));
    if (wants_diag(c)) LAUNCH(c, continuity_launch(c->u, c->v, c->diag, n, ld, c->st, c->stream));
    return finish_step(c, fin, outer, stats);
}

int cnv_step(cnv_ctx* c, int nsteps, cnv_step_stats* stats) {
    if (!c || nsteps < 0) return cnv_fail(c, CNV_ERR_ARG, "bad arguments to cnv_step");
    if (hipSetDevice(c->p.device) != hipSuccess) return cnv_fail(c, CNV_ERR_HIP, "hipSetDevice failed");
    for (int t = 0; t < nsteps; ++t) {
        int rc = one_step(c, stats ? &stats[t] : nullptr);
        if (rc) return rc;
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    return CNV_OK;
}

int cnv_step_timed(cnv_ctx* c, int nsteps, cnv_step_stats* stats, double* elapsed_ms) {
    if (!c || nsteps < 0) return cnv_fail(c, CNV_ERR_ARG, "bad arguments to cnv_step_timed");
    if (hipSetDevice(c->p.device) != hipSuccess) return cnv_fail(c, CNV_ERR_HIP, "hipSetDevice failed");
    HIPC(c, hipStreamSynchronize(c->stream));
    HIPC(c, hipEventRecord(c->t0, c->stream));
    for (int t = 0; t < nsteps; ++t) {
        int rc = one_step(c, stats ? &stats[t] : nullptr);
        if (rc) return rc;
    }
    HIPC(c, hipEventRecord(c->t1, c->stream));
    HIPC(c, hipEventSynchronize(c->t1));
    float ms = 0.f;
    HIPC(c, hipEventElapsedTime(&ms, c->t0, c->t1));
    if (elapsed_ms) *elapsed_ms = ms;
    return CNV_OK;
}

int cnv_poisson_device(cnv_ctx* c, const double* d_f, double* d_psi, int* k, double* e) {
    if (!c || !d_f || !d_psi) return cnv_fail(c, CNV_ERR_ARG, "bad arguments to cnv_poisson_device");
    if (hipSetDevice(c->p.device) != hipSuccess) return cnv_fail(c, CNV_ERR_HIP, "hipSetDevice failed");
    SorCtl fin{};
    int rc = poisson_fp64(c, d_f, 0, &fin);
    if (rc) return rc;
    HIPC(c, hipMemcpyAsync(d_psi, c->psi[fin.parity], sizeof(double) * c->cells, hipMemcpyDeviceToDevice, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    if (k) *k = fin.k - 1;
    if (e) *e = fin.e_last;
    if (fin.status == SOR_DIVERGED)
        return cnv_fail(c, CNV_DIVERGED, "Poisson solve diverged: non-finite L1 update after %d sweeps", fin.k);
    if (fin.status != SOR_CONVERGED)
        return cnv_fail(c, CNV_NOT_CONVERGED, "Error: maximum number of iterations achieved for Poisson equation.");
    return CNV_OK;
}

int cnv_poisson_host(cnv_ctx* c, const double* f, double* psi, int* k, double* e) {
    if (!c || !f || !psi) return cnv_fail(c, CNV_ERR_ARG, "bad arguments to cnv_poisson_host");
    if (hipSetDevice(c->p.device) != hipSuccess) return cnv_fail(c, CNV_ERR_HIP, "hipSetDevice failed");
    const size_t bytes = sizeof(double) * c->cells;
    // stage f in the spare w buffer: the solve reads it as its source term directly
    double* d_f = c->w[c->wcur ^ 1];
    HIPC(c, hipMemcpyAsync(d_f, f, bytes, hipMemcpyHostToDevice, c->stream));
    SorCtl fin{};
    int rc = poisson_fp64(c, d_f, 0, &fin);
    if (rc) return rc;
    HIPC(c, hipMemcpyAsync(psi, c->psi[fin.parity], bytes, hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    if (k) *k = fin.k - 1;
    if (e) *e = fin.e_last;
    if (fin.status == SOR_DIVERGED)
        return cnv_fail(c, CNV_DIVERGED, "Poisson solve diverged: non-finite L1 update after %d sweeps", fin.k);
    if (fin.status != SOR_CONVERGED)
        return cnv_fail(c, CNV_NOT_CONVERGED, "Error: maximum number of iterations achieved for Poisson equation.");
    return CNV_OK;
}

int cnv_time_sor(cnv_ctx* c, int launches, double* mean_ms) {
    if (!c || launches <= 0) return cnv_fail(c, CNV_ERR_ARG, "bad arguments to cnv_time_sor");
    if (hipSetDevice(c->p.device) != hipSuccess) return cnv_fail(c, CNV_ERR_HIP, "hipSetDevice failed");
    const bool dc = c->p.precision == CNV_PREC_FP32_DC;
    // a solve that never converges (tol < 0) and never runs out of budget, on the current w
    if (dc) {
        LAUNCH(c, dc_residual_launch(c->w[c->wcur], c->psi[0], c->r32, c->n, c->ld, c->p.dy * c->p.dy,
                                     c->p.dx * c->p.dx, c->p.dx * c->p.dx * c->p.dy * c->p.dy,
                                     2 * (c->p.dx * c->p.dx + c->p.dy * c->p.dy), c->stream));
        LAUNCH(c, dc_inner_init_launch(c->e32[0], c->ctl, c->n, c->ld, 0.0, -1.0, 0x7fffffff, c->stream));
    } else {
        LAUNCH(c, poisson_init_launch(c->psi[0], c->ctl, c->diag, c->n, c->ld, -1.0, 0x7fffffff, c->stream));
    }
    SorArgs<double> a64 = sor_args64(c, c->w[c->wcur], 1);
    SorArgs<float> a32 = sor_args32(c);
    const bool gs = c->p.poisson_type == 1;
    // Time back-to-back launches (each decides on its predecessor, as in a solve) in the steady
    // state of the solve: its first few hundred launches, while the iterate spreads out from
    // zero through tiny values, run about 5% slower and are left out (1000 untimed launches).
    // The timed launches are captured in a graph and replayed, as a solve's chunks are, so that
    // the host's launch rate stays out of the measurement.
    int L = 0;
    for (int l = 0; l < 1000; ++l, ++L)
        LAUNCH(c, dc ? launch_sor(c, a32, L, gs) : launch_sor(c, a64, L, gs));
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = CNV_OK;
    for (int l = 0; l < launches && rc == CNV_OK; ++l, ++L)
        if (dc ? launch_sor(c, a32, L, gs) : launch_sor(c, a64, L, gs))
            rc = cnv_fail(c, CNV_ERR_HIP, "SOR launch failed while capturing");
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    if (rc == CNV_OK && ec != hipSuccess)
        rc = cnv_fail(c, CNV_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ec));
    if (rc == CNV_OK && hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess)
        rc = cnv_fail(c, CNV_ERR_HIP, "hipGraphInstantiate failed");
    if (g) (void)hipGraphDestroy(g);
    if (rc != CNV_OK) return rc;
    float ms = 0.f;
    if (hipEventRecord(c->t0, c->stream) != hipSuccess || hipGraphLaunch(ge, c->stream) != hipSuccess ||
        hipEventRecord(c->t1, c->stream) != hipSuccess || hipEventSynchronize(c->t1) != hipSuccess ||
        hipEventElapsedTime(&ms, c->t0, c->t1) != hipSuccess)
        rc = cnv_fail(c, CNV_ERR_HIP, "timed SOR replay failed");
    (void)hipGraphExecDestroy(ge);
    if (rc != CNV_OK) return rc;
    *mean_ms = ms / launches;
    return CNV_OK;
}

int cnv_euler_host(int n, double* w, const double* dwdx, const double* dwdy, const double* d2wdx2,
                   const double* d2wdy2, const double* u, const double* v, double Re, double dt) {
    if (n <= 0) return cnv_fail(nullptr, CNV_ERR_ARG, "bad size");
    const size_t bytes = sizeof(double) * (size_t)n;
    double* d[7] = {};
    const double* h[7] = {w, dwdx, dwdy, d2wdx2, d2wdy2, u, v};
    int rc = CNV_OK;
    for (int q = 0; q < 7 && rc == CNV_OK; ++q) {
        if (hipMalloc(&d[q], bytes) != hipSuccess || hipMemcpy(d[q], h[q], bytes, hipMemcpyHostToDevice) != hipSuccess)
            rc = cnv_fail(nullptr, CNV_ERR_HIP, "euler: device allocation/upload failed");
    }
    if (rc == CNV_OK) {
        if (euler_launch(d[0], d[1], d[2], d[3], d[4], d[5], d[6], 1. / Re, dt, n, nullptr) ||
            hipMemcpy(w, d[0], bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = cnv_fail(nullptr, CNV_ERR_HIP, "euler: kernel failed");
    }
    for (int q = 0; q < 7; ++q)
        if (d[q]) (void)hipFree(d[q]);
    return rc;
}

static int elementwise_host(int m, int n, const double* a, const double* b, double* out, int sub) {
    const size_t cnt = (size_t)m * n;
    const size_t bytes = sizeof(double) * cnt;
    double *da = nullptr, *db = nullptr, *dc = nullptr;
    int rc = CNV_OK;
    if (hipMalloc(&da, bytes) != hipSuccess || hipMalloc(&db, bytes) != hipSuccess || hipMalloc(&dc, bytes) != hipSuccess ||
        hipMemcpy(da, a, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess || add_launch(da, db, dc, (long long)cnt, sub, nullptr) ||
        hipMemcpy(out, dc, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        rc = cnv_fail(nullptr, CNV_ERR_HIP, "elementwise kernel failed");
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dc) (void)hipFree(dc);
    return rc;
}

int cnv_add_host(int m, int n, const double* a, const double* b, double* out) {
    if (m <= 0 || n <= 0) return cnv_fail(nullptr, CNV_ERR_ARG, "bad size");
    return elementwise_host(m, n, a, b, out, 0);
}
int cnv_sub_host(int m, int n, const double* a, const double* b, double* out) {
    if (m <= 0 || n <= 0) return cnv_fail(nullptr, CNV_ERR_ARG, "bad size");
    return elementwise_host(m, n, a, b, out, 1);
}

int cnv_calibration_copy(cnv_ctx* c, long long nbytes, double* ms) {
    if (!c || nbytes <= 0 || nbytes % 8) return cnv_fail(c, CNV_ERR_ARG, "calibration copy needs a positive multiple of 8 bytes");
    if (hipSetDevice(c->p.device) != hipSuccess) return cnv_fail(c, CNV_ERR_HIP, "hipSetDevice failed");
    double *a = nullptr, *b = nullptr;
    HIPC(c, hipMalloc(&a, nbytes));
    if (hipMalloc(&b, nbytes) != hipSuccess) {
        (void)hipFree(a);
        return cnv_fail(c, CNV_ERR_HIP, "hipMalloc of %lld bytes failed", nbytes);
    }
    int rc = CNV_OK;
    if (copy_launch(b, a, nbytes / 8, c->stream)) rc = cnv_fail(c, CNV_ERR_HIP, "copy kernel failed");  // warm
    if (rc == CNV_OK && (hipEventRecord(c->t0, c->stream) != hipSuccess || copy_launch(b, a, nbytes / 8, c->stream) ||
                         hipEventRecord(c->t1, c->stream) != hipSuccess || hipEventSynchronize(c->t1) != hipSuccess))
        rc = cnv_fail(c, CNV_ERR_HIP, "copy kernel failed");
    float t = 0.f;
    if (rc == CNV_OK && hipEventElapsedTime(&t, c->t0, c->t1) == hipSuccess && ms) *ms = t;
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int cnv_set_poisson_params(cnv_ctx* c, double dx, double dy, double beta, double tol, int itmax, int poisson_type) {
    if (!c) return cnv_fail(nullptr, CNV_ERR_ARG, "null context");
    if (poisson_type != 1 && poisson_type != 2)
        return cnv_fail(c, CNV_ERR_ARG, "Error - invalid option for Poisson solver (poisson_type=%d)", poisson_type);
    if (!(dx > 0) || !(dy > 0)) return cnv_fail(c, CNV_ERR_ARG, "grid spacing must be positive");
    c->p.dx = dx;
    c->p.dy = dy;
    c->p.beta = beta;
    c->p.tol = tol;
    c->p.itmax = itmax;
    c->p.poisson_type = poisson_type;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);  // the captured launches hold the old coefficients
    return CNV_OK;
}
