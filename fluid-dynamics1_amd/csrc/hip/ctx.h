// ctx.h -- internal to libcnv: the context structure behind the opaque cnv_ctx of cnv.h and
// the helpers the one-cavity API (context.hip) and the batched API (batch.hip) share.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "cnv.h"
#include "cnv_device.h"

constexpr int kSlots = 4;  // pinned copies of the control block in flight

struct cnv_ctx {
    cnv_params p;
    int n, ld;
    size_t cells;
    hipStream_t stream = nullptr;
    // state
    double* w[2] = {nullptr, nullptr};
    int wcur = 0;
    double* psi[3] = {nullptr, nullptr, nullptr};  // FP64 SOR planes; FP32_DC: psi[0] accumulates
    double* u = nullptr;
    double* v = nullptr;
    // tables
    int* d_start = nullptr;
    double* d_tab = nullptr;  // 4 operators x n x 7
    float* d_ftab = nullptr;
    cnv::Stencils st{};
    // solver control
    cnv::SorCtl* ctl = nullptr;  // the Poisson solve's two state slots
    int fin_slot = 0;            // slot written by the last SOR launch of the last solve
    cnv::Diag* diag = nullptr;
    double* partial = nullptr;
    cnv::SorCtl* h_ctl = nullptr;  // pinned, kSlots entries
    cnv::Diag* h_diag = nullptr;   // pinned
    hipEvent_t ev[kSlots] = {};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    // defect correction
    float* e32[2] = {nullptr, nullptr};
    float* r32 = nullptr;
    // launch shape
    cnv::TileShape shape{5, 16, 4};
    int tiles_i = 0, tiles_j = 0;
    int launches_per_poll = 66;  // SOR launches per captured chunk (graph path), a multiple of tail_chunk
    int poll_fixed = 0;   // launches per poll as requested; 0: chosen per solve (eager path)
    int k_prev = 0;       // sweeps the next solve is predicted to take (the previous one's)
    static constexpr int kDcPred = 64;
    int k_dc[kDcPred] = {};  // FP32_DC: sweeps of the inner solve of each outer iteration
    cnv::WallBC bc{};
    // captured step (use_graph): pre-solve launches per w parity, a chunk of launches_per_poll
    // SOR launches and a tail chunk of tail_chunk per source plane (FP64; the refined form's
    // inner solves use index 0), post-solve launches; the refined form adds, per outer
    // iteration, a residual + inner-init graph per w parity and a correction graph.  Chunk
    // lengths are multiples of the launch index's period: 6 for FP64 (state slots mod 2, planes
    // mod 3), 12 for the FP32 inner solves (slots and planes mod 2, stopping tests mod 4).
    int tail_chunk = 6;
    hipGraphExec_t g_pre[2] = {nullptr, nullptr};
    hipGraphExec_t g_chunk[2] = {nullptr, nullptr};
    hipGraphExec_t g_tail[2] = {nullptr, nullptr};
    hipGraphExec_t g_resid[2] = {nullptr, nullptr};
    hipGraphExec_t g_correct = nullptr;
    hipGraphExec_t g_post = nullptr;
    int* d_budget = nullptr;  // refined form, captured: the solve's remaining sweep budget
    // output snapshots (cnv_snapshot_begin/wait): a device copy of (w, psi, u, v) taken in stream
    // order, then transferred to pinned host memory on a stream of its own
    double* d_snap = nullptr;
    double* h_snap = nullptr;
    hipStream_t snap_stream = nullptr;
    hipEvent_t ev_snap_ready = nullptr, ev_snap_done = nullptr;
    bool snap_begun = false;
    char err[256] = "";
};

// Records the message (per context, and process-wide for cnv_last_error(NULL)); returns code.
int cnv_fail(cnv_ctx* c, int code, const char* fmt, ...);

#define HIPC(c, x)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) return cnv_fail((c), CNV_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)
#define LAUNCH(c, x)                                                                           \
    do {                                                                                        \
        if ((x) != 0) return cnv_fail((c), CNV_ERR_HIP, "launch failed: %s (%s)", #x,         \
                                      hipGetErrorString(hipGetLastError()));                    \
    } while (0)

// SOR launch arguments for an n x n grid (pitch ld) with spacing (dx, dy), relaxation beta, the
// given tiling, scratch (partial sums, state slots), source plane src and iterate planes b0, b1,
// b2 (b2: FP64 only, nullptr for FP32).
template <typename T>
inline cnv::SorArgs<T> sor_args_for(int n, int ld, double dx, double dy, double beta, int tiles_i, int tiles_j,
                                    double* partial, cnv::SorCtl* ctl, const T* src, T* b0, T* b1, T* b2,
                                    int negate) {
    cnv::SorArgs<T> a{};
    a.src_f = src;
    a.psi0 = b0;
    a.psi1 = b1;
    a.psi2 = b2;
    a.partial = partial;
    a.ctl = ctl;
    a.nr = n;
    a.nc = n;
    a.ld = ld;
    a.tiles_i = tiles_i;
    a.tiles_j = tiles_j;
    a.negate = negate;
    // the reference evaluates dy*dy, dx*dx, dx*dx*dy*dy and 2*(dx*dx+dy*dy) left to right
    a.c1 = (T)(dy * dy);
    a.c2 = (T)(dx * dx);
    a.c3 = (T)(dx * dx * dy * dy);
    a.B = (T)(2 * (dx * dx + dy * dy));
    a.beta = (T)beta;
    a.omb = (T)(1 - beta);
    a.rB = (T)(1.0 / (2 * (dx * dx + dy * dy)));
    int e2 = 0;
    const double Bd = 2 * (dx * dx + dy * dy);
    a.exact_rcp = std::frexp(Bd, &e2) == 0.5 && std::isnormal(1.0 / Bd) && (1.0 / Bd) * Bd == 1.0;
    a.test_mask = 0;
    return a;
}

// The context's FP64 solve: source src, planes psi[0..2]; f = -src when negate.
inline cnv::SorArgs<double> sor_args64(cnv_ctx* c, const double* src, int negate) {
    return sor_args_for<double>(c->n, c->ld, c->p.dx, c->p.dy, c->p.beta, c->tiles_i, c->tiles_j, c->partial, c->ctl,
                                src, c->psi[0], c->psi[1], c->psi[2], negate);
}
// Launches per stopping test of FP32 inner solves: the parameter, or the given default.
inline int inner_test_every(const cnv_params& p, int dflt) { return p.inner_test_every > 0 ? p.inner_test_every : dflt; }
// The context's FP32 inner solve of defect correction: source r32, planes e32[0..1]; the inner
// solves may relax with their own factor and test their stopping rule every few launches.
inline cnv::SorArgs<float> sor_args32(cnv_ctx* c) {
    const double beta = c->p.inner_beta > 0 ? c->p.inner_beta : c->p.beta;
    cnv::SorArgs<float> a = sor_args_for<float>(c->n, c->ld, c->p.dx, c->p.dy, beta, c->tiles_i, c->tiles_j,
                                                c->partial, c->ctl, c->r32, c->e32[0], c->e32[1], nullptr, 0);
    a.test_mask = inner_test_every(c->p, cnv::kFp32TestEvery) - 1;
    return a;
}
// The stencils of the refined form's transport and velocities run in FP32 (transport_fp32).
inline bool fp32_stencils(const cnv_ctx* c) {
    return c->p.precision == CNV_PREC_FP32_DC && c->p.transport_fp32;
}
// The planes of the context's solve for the kernels that read its final iterate: FP64 the three
// SOR planes; the refined form's final iterate is always the FP64 accumulator psi[0].
inline cnv::Planes planes_of(const cnv_ctx* c) {
    if (c->p.precision == CNV_PREC_FP32_DC) return cnv::Planes{{c->psi[0], c->psi[0], c->psi[0]}};
    return cnv::Planes{{c->psi[0], c->psi[1], c->psi[2]}};
}
// Velocities from the solve's final iterate (slot fin): FP64 stencils, or FP32 ones from the refined
// form's FP64 accumulator psi[0].
inline int velocities(cnv_ctx* c, const cnv::SorCtl* fin) {
    if (fp32_stencils(c)) return cnv::velocity_fp32_launch(c->psi[0], c->u, c->v, c->n, c->ld, c->st, c->stream);
    return cnv::velocity_launch(planes_of(c), fin, c->u, c->v, c->n, c->ld, c->st, c->stream);
}

// The default SOR launch shape of an n x n grid on a device with `cus` compute units.
cnv::TileShape cnv_auto_shape(int n, int cus);

// Queue SOR launches until the device control block reports convergence or failure.  Keeps
// one chunk queued behind the one being waited on, so the GPU never idles on the host.
inline int launch_sor(cnv_ctx* c, cnv::SorArgs<double> a, int L, bool gs) {
    a.launch = L;
    return cnv::sor_fp64_launch(a, c->shape, gs, c->stream);
}
inline int launch_sor(cnv_ctx* c, cnv::SorArgs<float> a, int L, bool gs) {
    a.launch = L;
    return cnv::sor_fp32_launch(a, c->shape, gs, c->stream);
}

// A solve has ended (ROLLBACK is not an end: the next launch still writes the final iterate).
inline bool sor_done(int status) { return status != cnv::SOR_RUNNING && status != cnv::SOR_ROLLBACK; }

// A finished solve's slot status as a step status of cnv.h.
inline int solve_status(int sor_status) {
    return sor_status == cnv::SOR_CONVERGED ? CNV_OK : (sor_status == cnv::SOR_DIVERGED ? CNV_DIVERGED : CNV_NOT_CONVERGED);
}
// Ends the outer iterations of a defect-correction solve.
inline bool dc_final(int sor_status) {
    return sor_status == cnv::SOR_CONVERGED || sor_status == cnv::SOR_FAILED || sor_status == cnv::SOR_DIVERGED;
}

// Queue SOR launches (launch index L = 0, 1, ...) until the state slot written by a queued
// launch reports a terminal status.  Keeps one chunk queued behind the one being waited on, so
// the GPU never idles on the host; launches queued after the end are no-ops.
// With a prediction of P launches (k_prev), the first chunk is P - M launches and the rest come
// M at a time.  Each poll costs a copy in the stream (about 5.6 us), and the end of a solve wastes
// M to 2M no-op launches (about 3.8 us each at 1024^2); with a prediction off by about 2% of P
// the cost is smallest near M = sqrt(P / 50), kept within [4, 32].  Without a prediction M = 16.
inline int poll_chunk(const cnv_ctx* c, int predicted_launches) {
    if (c->poll_fixed > 0) return c->poll_fixed;
    if (predicted_launches <= 0) return 16;
    const int m = (int)std::lround(std::sqrt(predicted_launches / 50.0));
    return m < 4 ? 4 : (m > 32 ? 32 : m);
}
template <typename T>
inline int run_sor(cnv_ctx* c, const cnv::SorArgs<T>& a, cnv::SorCtl* final_ctl) {
    const bool gs = (c->p.poisson_type == 1);
    const int S = c->shape.s;
    const int P = c->k_prev > 0 ? c->k_prev / S : 0;  // predicted launches
    const int M = poll_chunk(c, P);
    const long long cap = (long long)c->p.itmax + 4LL * M + 8;  // working launches <= itmax/S + 2
    int first = M;
    if (P - M > first) first = P - M;
    long long queued = 0;
    int slot = 0;
    int pending[kSlots];
    int npend = 0;
    while (true) {
        const int chunk = queued == 0 ? first : M;
        for (int l = 0; l < chunk; ++l) {
            if (launch_sor(c, a, (int)(queued + l), gs))
                return cnv_fail(c, CNV_ERR_HIP, "SOR launch failed: %s", hipGetErrorString(hipGetLastError()));
        }
        queued += chunk;
        c->fin_slot = (int)((queued - 1) & 1);
        HIPC(c, hipMemcpyAsync(&c->h_ctl[slot], c->ctl + c->fin_slot, sizeof(cnv::SorCtl), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipEventRecord(c->ev[slot], c->stream));
        pending[npend++] = slot;
        slot = (slot + 1) % kSlots;
        if (npend >= 2) {
            const int s0 = pending[0];
            for (int q = 1; q < npend; ++q) pending[q - 1] = pending[q];
            --npend;
            HIPC(c, hipEventSynchronize(c->ev[s0]));
            const cnv::SorCtl snap = c->h_ctl[s0];
            if (sor_done(snap.status)) {
                *final_ctl = snap;
                return CNV_OK;
            }
        }
        if (queued > cap + first) return cnv_fail(c, CNV_ERR_HIP, "SOR state did not terminate after %lld launches", queued);
    }
}
