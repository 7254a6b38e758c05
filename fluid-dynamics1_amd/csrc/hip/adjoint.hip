// adjoint.hip -- gradient of the final kinetic energy of a cavity run with respect to the initial
// vorticity field and the lid velocity: hand-written adjoint kernels of the full FP64 time step
// with checkpoint-and-recompute.
//
// The step (src/main.c:276-451) carries the state s_t = (w_t, u_t, v_t); psi is recomputed from w
// every step.  With ub, vb the BC-applied velocities (src/main.c:283-296; the lid row i = n-1,
// columns 1..n-2, holds U), it is
//   wt      = w_t with its edges replaced by (Dx vb - Dy ub)             (src/main.c:298-320)
//   w_{t+1} = wt + dt*(-ub.(Dx wt) - vb.(Dy wt) + (1/Re)(Dxx wt + Dyy wt)) (src/fluiddyn.c:98)
//   psi     = A^-1 (c3 w_{t+1}) on the interior, 0 on the edges: A is the operator whose fixed
//             point the SOR iteration reaches, A psi = B psi - c1(N+S) - c2(E+W)
//             (src/poisson.c:268, f = -w)
//   u_{t+1} = Dy psi,  v_{t+1} = -(Dx psi)                               (src/main.c:366-383)
// and the objective is E = 1/2 dx dy sum over all points of (u_T^2 + v_T^2).
// The reverse step, given the adjoints (lw, lu, lv) of s_{t+1}:
//   lpsi = Dy^T lu - Dx^T lv, interior only
//   lw  += c3 A^-1 lpsi        (A is symmetric: the same SOR solve, the self-adjoint Poisson step)
//   lwt  = lw + dt*(-Dx^T(ub.lw) - Dy^T(vb.lw) + (1/Re)(Dxx^T lw + Dyy^T lw))
//   lub  = -dt (Dx wt).lw,  lvb = -dt (Dy wt).lw
//   lvb += Dx^T(edge(lwt)),  lub -= Dy^T(edge(lwt))              (wall vorticity)
//   (lw, lu, lv) of s_t = interior parts of (lwt, lub, lvb);  dE/dU += sum of lub on the lid
// The transposed operators use windowed tables built on the host from the same dense Diff
// matrices (exact transposes).  The forward steps stop their Poisson solves on a relative rule
// (L1 update < rtol * L1 norm of the iterate) so that the recomputed trajectory, the adjoint solves
// and finite differences all see fully converged solves; an absolute 1e-14 update is below the
// resolution of an FP64 sum over the grid.
//
// Checkpoints: the forward pass stores s_t every K steps in device memory; the backward pass
// recomputes each K-step segment from its checkpoint (storing its states) and runs the reverse
// steps through it, last segment first; the recompute of the segment before runs at the same time,
// on a second context and stream driven by a second host thread (cnv_adj_backward).
//
// Warm start (params warm_start): each forward solve starts from the previous step's psi and each
// adjoint solve from the previous reverse step's solution, instead of from 0.  The checkpoints
// then also hold the psi entering their step, and the adjoint solution is carried across
// segments, so the gradient still does not depend on the checkpoint interval.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "cnv.h"
#include "cnv_device.h"
#include "ctx.h"
#include "step_kernels.h"
#include "../host/cnv_host.h"

using namespace cnv;

namespace {

constexpr int kTW = 10;  // transposed windows: Diff1/Diff2 columns span at most 10 rows

struct TStencils {
    const int* start;   // start[l]: first row of column l's window
    const double* d1x;  // [l][m] = Diff1(dx)[start[l] + m][l]
    const double* d1y;
    const double* d2x;
    const double* d2y;
};

// (Dx^T g)[i][l] = sum_j d1x[j][l] g[i][j];  optional pointwise weight p on g (ub) and edge mask
template <int MODE>  // 0: g, 1: ub.g, 2: edge(g)
__device__ inline double tsum_x(const double* cT, const TStencils& T, const double* g, const double* u, int n,
                                int ld, const WallBC& bc, int i, int l) {
    const double* c = cT + (size_t)l * kTW;
    const int j0 = T.start[l];
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < kTW; ++m) {
        const int j = j0 + m;
        double x = g[(size_t)i * ld + j];
        if (MODE == 1) x *= bc_u(u, i, j, n, ld, bc);
        if (MODE == 2 && !(i == 0 || i == n - 1 || j == 0 || j == n - 1)) x = 0.0;
        s += c[m] * x;
    }
    return s;
}
// (Dy^T g)[k][j] = sum_i d1y[i][k] g[i][j]
template <int MODE>  // 0: g, 1: vb.g, 2: edge(g)
__device__ inline double tsum_y(const double* cT, const TStencils& T, const double* g, const double* v, int n,
                                int ld, const WallBC& bc, int k, int j) {
    const double* c = cT + (size_t)k * kTW;
    const int i0 = T.start[k];
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < kTW; ++m) {
        const int i = i0 + m;
        double x = g[(size_t)i * ld + j];
        if (MODE == 1) x *= bc_v(v, i, j, n, ld, bc);
        if (MODE == 2 && !(i == 0 || i == n - 1 || j == 0 || j == n - 1)) x = 0.0;
        s += c[m] * x;
    }
    return s;
}

__device__ inline bool cell(int n, int& i, int& j) {
    j = blockIdx.x * 64 + (threadIdx.x & 63);
    i = blockIdx.y * 4 + (threadIdx.x >> 6);
    return i < n && j < n;
}
__device__ inline bool interior(int n, int i, int j) { return i >= 1 && i <= n - 2 && j >= 1 && j <= n - 2; }

// seed of the objective: lu = dxdy u_T, lv = dxdy v_T, lw = 0
__global__ void __launch_bounds__(kBlock) adj_seed_kernel(const double* u, const double* v, double* lw, double* lu,
                                                         double* lv, int n, int ld, double h2) {
    int i, j;
    if (!cell(n, i, j)) return;
    const size_t q = (size_t)i * ld + j;
    lw[q] = 0.0;
    lu[q] = h2 * u[q];
    lv[q] = h2 * v[q];
}

// adjoint Poisson source: src = (Dy^T lu - Dx^T lv) / c3 on the interior, 0 on the edges; the
// SOR solve with f = -src reaches A z = lpsi
__global__ void __launch_bounds__(kBlock) adj_vel_src_kernel(const double* lu, const double* lv, double* src, int n,
                                                            int ld, TStencils T, WallBC bc, double c3) {
    int i, j;
    if (!cell(n, i, j)) return;
    double s = 0.0;
    if (interior(n, i, j))
        s = (tsum_y<0>(T.d1y, T, lu, nullptr, n, ld, bc, i, j) - tsum_x<0>(T.d1x, T, lv, nullptr, n, ld, bc, i, j)) / c3;
    src[(size_t)i * ld + j] = s;
}

// lw += c3 z on the interior (z: the adjoint solve's final iterate)
__global__ void __launch_bounds__(kBlock) adj_accum_kernel(double* lw, Planes zp, const SorCtl* fin, int n, int ld,
                                                          double c3) {
    int i, j;
    if (!cell(n, i, j) || !interior(n, i, j)) return;
    const double* z = zp.p[fin->parity];
    const size_t q = (size_t)i * ld + j;
    lw[q] += c3 * z[q];
}

// transport adjoint: lwt, lub, lvb from lw, wt and the state's velocities
__global__ void __launch_bounds__(kBlock) adj_transport_kernel(const double* lw, const double* wt, const double* u,
                                                              const double* v, double* lwt, double* lub, double* lvb,
                                                              int n, int ld, Stencils S, TStencils T, WallBC bc,
                                                              double rRe, double dt) {
    int i, l;
    if (!cell(n, i, l)) return;
    const size_t q = (size_t)i * ld + l;
    const double adv = -tsum_x<1>(T.d1x, T, lw, u, n, ld, bc, i, l) - tsum_y<1>(T.d1y, T, lw, v, n, ld, bc, i, l);
    const double dif = tsum_x<0>(T.d2x, T, lw, nullptr, n, ld, bc, i, l) + tsum_y<0>(T.d2y, T, lw, nullptr, n, ld, bc, i, l);
    lwt[q] = lw[q] + dt * (adv + rRe * dif);
    lub[q] = -dt * sum_x(S.d1x, S.start, wt, i, l, ld) * lw[q];
    lvb[q] = -dt * sum_y(S.d1y, S.start, wt, i, l, ld) * lw[q];
}

// wall-vorticity adjoint: lvb += Dx^T(edge(lwt)), lub -= Dy^T(edge(lwt))
__global__ void __launch_bounds__(kBlock) adj_wall_kernel(const double* lwt, double* lub, double* lvb, int n, int ld,
                                                         TStencils T, WallBC bc) {
    int i, l;
    if (!cell(n, i, l)) return;
    const size_t q = (size_t)i * ld + l;
    lvb[q] += tsum_x<2>(T.d1x, T, lwt, nullptr, n, ld, bc, i, l);
    lub[q] -= tsum_y<2>(T.d1y, T, lwt, nullptr, n, ld, bc, i, l);
}

// adjoint of s_t: interior parts (the BCs overwrite the edges of u, v and the wall vorticity those
// of w before they are read)
__global__ void __launch_bounds__(kBlock) adj_finalize_kernel(const double* lwt, const double* lub, const double* lvb,
                                                             double* lw, double* lu, double* lv, int n, int ld) {
    int i, j;
    if (!cell(n, i, j)) return;
    const size_t q = (size_t)i * ld + j;
    const bool in = interior(n, i, j);
    lw[q] = in ? lwt[q] : 0.0;
    lu[q] = in ? lub[q] : 0.0;
    lv[q] = in ? lvb[q] : 0.0;
}

// dE/dU += sum of lub over the lid (row n-1, columns 1..n-2), in column order (one thread)
__global__ void adj_lid_kernel(const double* lub, double* dU, int n, int ld) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int j = 1; j <= n - 2; ++j) s += lub[(size_t)(n - 1) * ld + j];
    *dU += s;
}

// both state slots of a solve that stops on the relative rule
__global__ void adj_slots_kernel(SorCtl* ctl, double tol, double rtol, int itmax) {
    if (threadIdx.x < 2) ctl_reset(ctl + threadIdx.x, tol, rtol, itmax);
}

__global__ void adj_zero_kernel(double* p, long long count) {
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (q < count) p[q] = 0.0;
}

thread_local char g_aerr[256] = "";  // last error of this thread

}  // namespace

struct cnv_adjoint {
    cnv_ctx* c = nullptr;   // forward runs and the backward pass's recompute
    cnv_ctx* cz = nullptr;  // the reverse steps (their solves and kernels), on a stream of its own
    int K = 100;
    double rtol_fwd = 1e-14, rtol_adj = 1e-14;
    int T = 0;  // steps of the last forward run
    double U = 1.0;
    size_t cells = 0;
    // transposed tables
    int* d_tstart = nullptr;
    double* d_ttab = nullptr;
    TStencils ts{};
    // storage
    double* d_ckpt = nullptr;  // [nckpt][3][cells]
    int nckpt_cap = 0;
    double* d_seg[2] = {nullptr, nullptr};  // [K][3][cells] each: segment s in d_seg[s & 1]
    hipEvent_t ev_ready[2] = {nullptr, nullptr};  // d_seg[b] recomputed (recorded on c's stream)
    hipEvent_t ev_free[2] = {nullptr, nullptr};   // d_seg[b] read by the reverse steps (cz's stream)
    int seg_stored = -1;  // segment whose step states the last forward run left in its buffer (-1: none)
    double *lw = nullptr, *lu = nullptr, *lv = nullptr, *lwt = nullptr, *lub = nullptr, *lvb = nullptr;
    double *wt = nullptr, *src = nullptr, *dU = nullptr;
    double *uT = nullptr, *vT = nullptr;  // the last forward run's final velocities (the seed)
    int k_fwd = 0, k_adj = 0;  // queue-ahead predictions of the two kinds of solve
    // warm start (params warm_start): each forward solve starts from the previous step's psi and
    // each adjoint solve from the previous reverse step's solution; checkpoints then also hold
    // the psi that enters their step
    bool warm = false;
    double* d_ckpsi = nullptr;  // [nckpt][cells]
    double* zk = nullptr;       // the last adjoint solve's solution
    long long sweeps_fwd = 0, sweeps_adj = 0, sweeps_rec = 0;  // forward run; backward: adjoint, recompute
    char err[256] = "";
};

static std::mutex g_aerr_mu;  // a->err is written by both threads of a backward pass

static int afail(cnv_adjoint* a, int code, const char* fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (a) {
        std::lock_guard<std::mutex> l(g_aerr_mu);
        snprintf(a->err, sizeof(a->err), "%s", buf);
    }
    snprintf(g_aerr, sizeof(g_aerr), "%s", buf);
    return code;
}
#define AHIP(a, x)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return afail((a), CNV_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)
#define ALAUNCH(a)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = hipGetLastError();                                                            \
        if (e_ != hipSuccess) return afail((a), CNV_ERR_HIP, "adjoint launch failed: %s", hipGetErrorString(e_)); \
    } while (0)
// a launch helper of the step kernels (nonzero when the launch failed)
#define ALAUNCH_OF(a, x)                                                                              \
    do {                                                                                              \
        if ((x) != 0) return afail((a), CNV_ERR_HIP, "adjoint launch failed: %s", #x);               \
    } while (0)
// a call on the context ctx that returns a status and leaves its message in ctx->err
#define ACTX(a, ctx, x)                                                                               \
    do {                                                                                              \
        const int rc_ = (x);                                                                          \
        if (rc_) return afail((a), rc_, "%s", (ctx)->err);                                           \
    } while (0)

static dim3 cell_grid(int n) { return dim3((n + 63) / 64, (n + 3) / 4); }

// Poisson solve on the context's grid with the relative stopping rule: psi[parity] <- A^-1(...)
// for the source src read as f = -src.  Returns the sweeps taken.
// With warm start the solve starts from psi[0] as the caller left it, else from 0.
static int solve_rel(cnv_adjoint* a, cnv_ctx* c, const double* src, double rtol, int* kpred, int* sweeps) {
    ALAUNCH_OF(a, poisson_init_launch(c->psi[0], c->ctl, nullptr, c->n, c->ld, 0.0, c->p.itmax, c->stream, !a->warm));
    hipLaunchKernelGGL(adj_slots_kernel, dim3(1), dim3(64), 0, c->stream, c->ctl, 1e-300, rtol, c->p.itmax);
    ALAUNCH(a);
    SorArgs<double> sa = sor_args64(c, src, 1);
    sa.norm = 1;
    SorCtl fin{};
    c->k_prev = *kpred;
    ACTX(a, c, run_sor<double>(c, sa, &fin));
    if (fin.status == SOR_DIVERGED)
        return afail(a, CNV_DIVERGED, "adjoint-problem Poisson solve diverged (non-finite update) after %d sweeps", fin.k);
    if (fin.status == SOR_FAILED)
        return afail(a, CNV_NOT_CONVERGED, "adjoint-problem Poisson solve did not converge in %d sweeps", fin.k);
    *kpred = fin.k;
    *sweeps = fin.k;
    return CNV_OK;
}

// one forward step on the context state (w[wcur], u, v), Poisson solve on the relative rule;
// with warm start its first iterate is the last solve's final one, unless the caller has put
// the step's own in psi[0] (start_given)
static int fwd_step(cnv_adjoint* a, bool start_given, long long* sweep_count) {
    cnv_ctx* c = a->c;
    const int n = c->n, ld = c->ld;
    if (a->warm && !start_given)
        ALAUNCH_OF(a, warm_start_launch(c->psi[0], planes_of(c), c->ctl + c->fin_slot, n, ld, c->stream));
    double* w_in = c->w[c->wcur];
    double* w_out = c->w[c->wcur ^ 1];
    ALAUNCH_OF(a, wall_vorticity_launch(w_in, c->u, c->v, n, ld, c->st, c->bc, c->stream));
    ALAUNCH_OF(a, transport_fp64_launch(w_in, w_out, c->u, c->v, n, ld, c->st, c->bc, 1. / c->p.Re, c->p.dt, c->stream));
    c->wcur ^= 1;
    int k = 0;
    int rc = solve_rel(a, c, w_out, a->rtol_fwd, &a->k_fwd, &k);
    if (rc) return rc;
    *sweep_count += k;
    ALAUNCH_OF(a, velocity_launch(planes_of(c), c->ctl + c->fin_slot, c->u, c->v, n, ld, c->st, c->stream));
    return CNV_OK;
}

// warm start: psi[0] <- the psi entering checkpointed step idx (save: from psi[0])
static int ckpt_psi(cnv_adjoint* a, int idx, bool save) {
    cnv_ctx* c = a->c;
    double* ck = a->d_ckpsi + (size_t)idx * a->cells;
    const size_t bytes = sizeof(double) * a->cells;
    AHIP(a, hipMemcpyAsync(save ? ck : c->psi[0], save ? c->psi[0] : ck, bytes, hipMemcpyDeviceToDevice, c->stream));
    return CNV_OK;
}

static double* slot3(double* base, int idx, int f, size_t cells) { return base + ((size_t)idx * 3 + f) * cells; }

static int save_state(cnv_adjoint* a, double* base, int idx) {
    cnv_ctx* c = a->c;
    const size_t bytes = sizeof(double) * a->cells;
    AHIP(a, hipMemcpyAsync(slot3(base, idx, 0, a->cells), c->w[c->wcur], bytes, hipMemcpyDeviceToDevice, c->stream));
    AHIP(a, hipMemcpyAsync(slot3(base, idx, 1, a->cells), c->u, bytes, hipMemcpyDeviceToDevice, c->stream));
    AHIP(a, hipMemcpyAsync(slot3(base, idx, 2, a->cells), c->v, bytes, hipMemcpyDeviceToDevice, c->stream));
    return CNV_OK;
}
static int load_state(cnv_adjoint* a, double* base, int idx) {
    cnv_ctx* c = a->c;
    const size_t bytes = sizeof(double) * a->cells;
    AHIP(a, hipMemcpyAsync(c->w[c->wcur], slot3(base, idx, 0, a->cells), bytes, hipMemcpyDeviceToDevice, c->stream));
    AHIP(a, hipMemcpyAsync(c->u, slot3(base, idx, 1, a->cells), bytes, hipMemcpyDeviceToDevice, c->stream));
    AHIP(a, hipMemcpyAsync(c->v, slot3(base, idx, 2, a->cells), bytes, hipMemcpyDeviceToDevice, c->stream));
    return CNV_OK;
}

// reverse of one step whose input state is stored at slot idx of seg; (lw, lu, lv) in/out; on the
// reverse steps' context cz
static int rev_step(cnv_adjoint* a, double* seg, int idx) {
    cnv_ctx* c = a->cz;
    const int n = c->n, ld = c->ld;
    const double dx = c->p.dx, dy = c->p.dy;
    const double c3 = dx * dx * dy * dy;
    const dim3 g = cell_grid(n), blk(kBlock);
    hipLaunchKernelGGL(adj_vel_src_kernel, g, blk, 0, c->stream, a->lu, a->lv, a->src, n, ld, a->ts, c->bc, c3);
    ALAUNCH(a);
    int k = 0;
    if (a->warm)
        AHIP(a, hipMemcpyAsync(c->psi[0], a->zk, sizeof(double) * a->cells, hipMemcpyDeviceToDevice, c->stream));
    int rc = solve_rel(a, c, a->src, a->rtol_adj, &a->k_adj, &k);
    if (rc) return rc;
    a->sweeps_adj += k;
    if (a->warm) ALAUNCH_OF(a, warm_start_launch(a->zk, planes_of(c), c->ctl + c->fin_slot, n, ld, c->stream));
    hipLaunchKernelGGL(adj_accum_kernel, g, blk, 0, c->stream, a->lw, planes_of(c), c->ctl + c->fin_slot, n,
                       ld, c3);
    ALAUNCH(a);
    const double* ws = slot3(seg, idx, 0, a->cells);
    const double* us = slot3(seg, idx, 1, a->cells);
    const double* vs = slot3(seg, idx, 2, a->cells);
    AHIP(a, hipMemcpyAsync(a->wt, ws, sizeof(double) * a->cells, hipMemcpyDeviceToDevice, c->stream));
    ALAUNCH_OF(a, wall_vorticity_launch(a->wt, us, vs, n, ld, c->st, c->bc, c->stream));
    hipLaunchKernelGGL(adj_transport_kernel, g, blk, 0, c->stream, a->lw, a->wt, us, vs, a->lwt, a->lub, a->lvb, n, ld,
                       c->st, a->ts, c->bc, 1. / c->p.Re, c->p.dt);
    ALAUNCH(a);
    hipLaunchKernelGGL(adj_wall_kernel, g, blk, 0, c->stream, a->lwt, a->lub, a->lvb, n, ld, a->ts, c->bc);
    ALAUNCH(a);
    hipLaunchKernelGGL(adj_lid_kernel, dim3(1), dim3(64), 0, c->stream, a->lub, a->dU, n, ld);
    ALAUNCH(a);
    hipLaunchKernelGGL(adj_finalize_kernel, g, blk, 0, c->stream, a->lwt, a->lub, a->lvb, a->lw, a->lu, a->lv, n, ld);
    ALAUNCH(a);
    return CNV_OK;
}

static int build_transposed(cnv_adjoint* a) {
    cnv_ctx* c = a->c;
    const int n = c->n, o = c->p.order;
    std::vector<int> start(n), tstart(n);
    std::vector<double> D((size_t)n * n), win((size_t)n * kStencilWidth);
    // structural windows (every operator has the same ones): row r covers [start[r], start[r]+7)
    if (cnv_fd_dense(1, n, o, c->p.dx, D.data()) || cnv_fd_windows(n, D.data(), start.data(), win.data()))
        return afail(a, CNV_ERR_ARG, "bad stencil order %d for n = %d", o, n);
    for (int l = 0; l < n; ++l) {
        int jlo = n, jhi = -1;
        for (int j = 0; j < n; ++j)
            if (start[j] <= l && l < start[j] + kStencilWidth) {
                jlo = std::min(jlo, j);
                jhi = std::max(jhi, j);
            }
        if (jhi - jlo + 1 > kTW) return afail(a, CNV_ERR_ARG, "transposed stencil wider than %d rows", kTW);
        tstart[l] = std::min(jlo, n - kTW);
    }
    std::vector<double> tab((size_t)4 * n * kTW);
    const double hs[4] = {c->p.dx, c->p.dy, c->p.dx, c->p.dy};
    const int which[4] = {1, 1, 2, 2};
    for (int t = 0; t < 4; ++t) {
        if (cnv_fd_dense(which[t], n, o, hs[t], D.data())) return afail(a, CNV_ERR_ARG, "bad stencil order");
        for (int l = 0; l < n; ++l)
            for (int m = 0; m < kTW; ++m) tab[((size_t)t * n + l) * kTW + m] = D[(size_t)(tstart[l] + m) * n + l];
    }
    AHIP(a, hipMalloc(&a->d_tstart, sizeof(int) * n));
    AHIP(a, hipMalloc(&a->d_ttab, sizeof(double) * tab.size()));
    AHIP(a, hipMemcpy(a->d_tstart, tstart.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    AHIP(a, hipMemcpy(a->d_ttab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    const size_t stride = (size_t)n * kTW;
    a->ts = TStencils{a->d_tstart, a->d_ttab, a->d_ttab + stride, a->d_ttab + 2 * stride, a->d_ttab + 3 * stride};
    return CNV_OK;
}

// ---- C ABI -----------------------------------------------------------------------------------

const char* cnv_adj_last_error(const cnv_adjoint* a) { return a ? a->err : g_aerr; }

void cnv_adj_destroy(cnv_adjoint* a) {
    if (!a) return;
    for (cnv_ctx* c : {a->c, a->cz})
        if (c) {
            (void)hipStreamSynchronize(c->stream);
            cnv_destroy(c);
        }
    for (hipEvent_t e : {a->ev_ready[0], a->ev_ready[1], a->ev_free[0], a->ev_free[1]})
        if (e) (void)hipEventDestroy(e);
    void* dev[] = {a->d_tstart, a->d_ttab, a->d_ckpt, a->d_ckpsi, a->d_seg[0], a->d_seg[1], a->lw, a->lu, a->lv,
                   a->lwt, a->lub, a->lvb, a->wt, a->src, a->dU, a->uT, a->vT, a->zk};
    for (void* p : dev)
        if (p) (void)hipFree(p);
    delete a;
}

int cnv_adj_create(const cnv_params* p, int checkpoint_every, double rtol_forward, double rtol_adjoint,
                   cnv_adjoint** out) {
    if (!out) return afail(nullptr, CNV_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!p || checkpoint_every < 1 || !(rtol_forward > 0) || !(rtol_adjoint > 0))
        return afail(nullptr, CNV_ERR_ARG, "adjoint: need params, checkpoint_every >= 1 and positive tolerances");
    if (p->precision != CNV_PREC_FP64 || p->poisson_type != 2 || p->n < kTW)
        return afail(nullptr, CNV_ERR_ARG, "adjoint: FP64 SOR on n >= %d only", kTW);
    cnv_adjoint* a = new cnv_adjoint();
    a->K = checkpoint_every;
    a->rtol_fwd = rtol_forward;
    a->rtol_adj = rtol_adjoint;
    auto setup = [&]() -> int {
        cnv_params q = *p;
        q.diagnostics = 0;
        int rc = cnv_create(&q, &a->c);
        if (!rc) rc = cnv_create(&q, &a->cz);
        if (rc) return afail(a, rc, "%s", cnv_last_error(nullptr));
        a->cells = a->c->cells;
        const size_t bytes = sizeof(double) * a->cells;
        for (int b = 0; b < 2; ++b) {
            AHIP(a, hipMalloc(&a->d_seg[b], bytes * 3 * a->K));
            AHIP(a, hipEventCreateWithFlags(&a->ev_ready[b], hipEventDisableTiming));
            AHIP(a, hipEventCreateWithFlags(&a->ev_free[b], hipEventDisableTiming));
        }
        double** bufs[] = {&a->lw, &a->lu, &a->lv, &a->lwt, &a->lub, &a->lvb, &a->wt, &a->src, &a->uT, &a->vT};
        for (double** b : bufs) AHIP(a, hipMalloc(b, bytes));
        AHIP(a, hipMalloc(&a->dU, sizeof(double)));
        if (p->warm_start) {
            a->warm = true;
            AHIP(a, hipMalloc(&a->zk, bytes));
        }
        return build_transposed(a);
    };
    const int rc = setup();
    if (rc) {
        char keep[256];
        snprintf(keep, sizeof(keep), "%s", a->err);
        cnv_adj_destroy(a);
        snprintf(g_aerr, sizeof(g_aerr), "%s", keep);
        return rc;
    }
    *out = a;
    return CNV_OK;
}

int cnv_adj_forward(cnv_adjoint* a, const double* w0, double U, int nsteps, double* energy) {
    if (!a || !w0 || nsteps < 1) return afail(a, CNV_ERR_ARG, "adjoint forward: bad arguments");
    cnv_ctx* c = a->c;
    AHIP(a, hipSetDevice(c->p.device));
    const int n = c->n;
    const size_t bytes = sizeof(double) * a->cells;
    const int nck = (nsteps + a->K - 1) / a->K;
    if (nck > a->nckpt_cap) {
        if (a->d_ckpt) AHIP(a, hipFree(a->d_ckpt));
        if (a->d_ckpsi) AHIP(a, hipFree(a->d_ckpsi));
        a->d_ckpt = a->d_ckpsi = nullptr;
        AHIP(a, hipMalloc(&a->d_ckpt, bytes * 3 * nck));
        if (a->warm) AHIP(a, hipMalloc(&a->d_ckpsi, bytes * nck));
        a->nckpt_cap = nck;
    }
    // initial state: w0; u = ui, v = vi inside, 0 on the edges (src/main.c:178-221); psi = 0
    std::vector<double> u0(a->cells, 0.0), v0(a->cells, 0.0);
    for (int i = 1; i < n - 1; ++i)
        for (int j = 1; j < n - 1; ++j) {
            u0[(size_t)i * n + j] = c->p.ui;
            v0[(size_t)i * n + j] = c->p.vi;
        }
    AHIP(a, hipStreamSynchronize(c->stream));
    AHIP(a, hipMemcpy(c->w[c->wcur], w0, bytes, hipMemcpyHostToDevice));
    AHIP(a, hipMemcpy(c->u, u0.data(), bytes, hipMemcpyHostToDevice));
    AHIP(a, hipMemcpy(c->v, v0.data(), bytes, hipMemcpyHostToDevice));
    c->bc.u4 = U;
    a->U = U;
    a->T = 0;  // until this run has finished: a backward call after a failed run is refused
    a->sweeps_fwd = 0;
    a->seg_stored = -1;
    if (a->warm) AHIP(a, hipMemsetAsync(c->psi[0], 0, bytes, c->stream));  // step 0 starts from 0
    const int last = nck - 1;  // its step states go straight to its segment buffer: no recompute
    for (int t = 0; t < nsteps; ++t) {
        const bool ck = t % a->K == 0;
        if (t >= last * a->K) {
            int rc = save_state(a, a->d_seg[last & 1], t - last * a->K);
            if (rc) return rc;
        }
        if (ck) {
            int rc = save_state(a, a->d_ckpt, t / a->K);
            if (rc) return rc;
            // the psi entering step t: the last solve's final iterate (0 before step 0)
            if (a->warm && t > 0)
                ALAUNCH_OF(a, warm_start_launch(c->psi[0], planes_of(c), c->ctl + c->fin_slot, n, c->ld, c->stream));
            if (a->warm && (rc = ckpt_psi(a, t / a->K, true))) return rc;
        }
        int rc = fwd_step(a, ck, &a->sweeps_fwd);
        if (rc) return rc;
    }
    // the backward pass seeds from these: its recompute moves the context's own state
    AHIP(a, hipMemcpyAsync(a->uT, c->u, bytes, hipMemcpyDeviceToDevice, c->stream));
    AHIP(a, hipMemcpyAsync(a->vT, c->v, bytes, hipMemcpyDeviceToDevice, c->stream));
    // E = 1/2 dx dy sum (u^2 + v^2), summed on the host in row-major order
    AHIP(a, hipMemcpyAsync(u0.data(), c->u, bytes, hipMemcpyDeviceToHost, c->stream));
    AHIP(a, hipMemcpyAsync(v0.data(), c->v, bytes, hipMemcpyDeviceToHost, c->stream));
    AHIP(a, hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (size_t q = 0; q < a->cells; ++q) s += u0[q] * u0[q] + v0[q] * v0[q];
    if (energy) *energy = 0.5 * c->p.dx * c->p.dy * s;
    a->T = nsteps;
    a->seg_stored = last;
    return CNV_OK;
}

// recompute of segment s from its checkpoint on the context c, storing each step's input state
// in d_seg[b]
static int recompute(cnv_adjoint* a, int s, int b) {
    const int t0 = s * a->K, t1 = std::min(a->T, t0 + a->K);
    int rc = load_state(a, a->d_ckpt, s);
    if (!rc && a->warm) rc = ckpt_psi(a, s, false);
    for (int t = t0; t < t1 && !rc; ++t) {
        rc = save_state(a, a->d_seg[b], t - t0);
        if (!rc) rc = fwd_step(a, t == t0, &a->sweeps_rec);
    }
    return rc;
}

// The backward pass runs two pipelines at once.  A second thread recomputes the segments, last
// first, on the context c (the last one only when a backward pass since the forward run has reused
// its buffer); this thread runs each segment's reverse steps on the context cz as soon as its
// recompute is done.  Segment s goes to d_seg[s & 1], so the recompute of segment s - 1
// overlaps the reverse steps of segment s; the recompute of s - 2 waits until those reverse steps
// have read the buffer.  The two streams are ordered by events (ev_ready[b] after a recompute,
// ev_free[b] after its reverse steps), and each thread queues its wait only after the host has seen
// the other thread record the event it waits for.  Every kernel sees the same inputs as it would
// one pipeline after the other, so the gradient is the same bit for bit.
int cnv_adj_backward(cnv_adjoint* a, double* dE_dw0, double* dE_dU) {
    if (!a || a->T < 1) return afail(a, CNV_ERR_ARG, "adjoint backward: run cnv_adj_forward first");
    cnv_ctx* c = a->c;
    cnv_ctx* z = a->cz;
    AHIP(a, hipSetDevice(c->p.device));
    const int n = c->n, ld = c->ld;
    const dim3 g = cell_grid(n), blk(kBlock);
    AHIP(a, hipStreamSynchronize(c->stream));
    AHIP(a, hipStreamSynchronize(z->stream));
    z->bc = c->bc;  // the wall velocities of the last forward run (its lid velocity U)
    // seed from the last forward run's final state, kept apart from the context's state, which
    // each backward pass's recompute moves (to step min(T, K) by its end): a second backward
    // call after one forward run returns the same gradient
    hipLaunchKernelGGL(adj_seed_kernel, g, blk, 0, z->stream, a->uT, a->vT, a->lw, a->lu, a->lv, n, ld,
                       c->p.dx * c->p.dy);
    ALAUNCH(a);
    hipLaunchKernelGGL(adj_zero_kernel, dim3(1), dim3(kBlock), 0, z->stream, a->dU, 1LL);
    ALAUNCH(a);
    if (a->warm) AHIP(a, hipMemsetAsync(a->zk, 0, sizeof(double) * a->cells, z->stream));  // the last step's z starts from 0
    a->sweeps_adj = a->sweeps_rec = 0;
    const int nseg = (a->T + a->K - 1) / a->K;
    // the last segment's states, stored by the forward run, hold until this pass reuses its buffer
    const bool have_last = a->seg_stored == nseg - 1;
    a->seg_stored = -1;

    std::mutex m;
    std::condition_variable cv;
    int recomputed = nseg, reversed = nseg;  // lowest segment whose recompute / reverse steps are queued
    bool stop = false;
    auto post = [&](int& which, int sg) {
        {
            std::lock_guard<std::mutex> l(m);
            which = sg;
        }
        cv.notify_all();
    };
    auto halt = [&]() {
        {
            std::lock_guard<std::mutex> l(m);
            stop = true;
        }
        cv.notify_all();
    };
    auto wait_until = [&](const int& which, int sg) {  // false: the other thread stopped
        std::unique_lock<std::mutex> l(m);
        cv.wait(l, [&] { return stop || which <= sg; });
        return !stop;
    };
    int rc_rec = CNV_OK;
    char err_rec[256] = "";
    auto recompute_all = [&]() {
        int rc = CNV_OK;
        if (hipSetDevice(c->p.device) != hipSuccess) rc = afail(a, CNV_ERR_HIP, "hipSetDevice failed in the recompute");
        for (int sg = nseg - 1; sg >= 0 && !rc; --sg) {
            const int b = sg & 1;
            if (sg + 2 < nseg) {  // d_seg[b] holds segment sg + 2 until its reverse steps have run
                if (!wait_until(reversed, sg + 2)) break;
                if (hipStreamWaitEvent(c->stream, a->ev_free[b], 0) != hipSuccess) {
                    rc = afail(a, CNV_ERR_HIP, "hipStreamWaitEvent failed in the recompute");
                    break;
                }
            }
            if (!(have_last && sg == nseg - 1)) rc = recompute(a, sg, b);
            if (!rc && hipEventRecord(a->ev_ready[b], c->stream) != hipSuccess)
                rc = afail(a, CNV_ERR_HIP, "hipEventRecord failed in the recompute");
            if (!rc) post(recomputed, sg);
        }
        if (rc) {
            {
                std::lock_guard<std::mutex> l(g_aerr_mu);
                snprintf(err_rec, sizeof(err_rec), "%s", a->err);
            }
            rc_rec = rc;
            halt();
        }
    };
    std::thread rec;
    try {
        rec = std::thread(recompute_all);
    } catch (...) {
        return afail(a, CNV_ERR_NOMEM, "adjoint backward: could not start the recompute thread");
    }
    int rc = CNV_OK;
    for (int sg = nseg - 1; sg >= 0 && !rc; --sg) {
        const int b = sg & 1;
        const int t0 = sg * a->K, t1 = std::min(a->T, t0 + a->K);
        if (!wait_until(recomputed, sg)) break;
        if (hipStreamWaitEvent(z->stream, a->ev_ready[b], 0) != hipSuccess)
            rc = afail(a, CNV_ERR_HIP, "hipStreamWaitEvent failed in the reverse steps");
        for (int t = t1 - 1; t >= t0 && !rc; --t) rc = rev_step(a, a->d_seg[b], t - t0);
        if (!rc && hipEventRecord(a->ev_free[b], z->stream) != hipSuccess)
            rc = afail(a, CNV_ERR_HIP, "hipEventRecord failed in the reverse steps");
        if (rc) halt();
        else post(reversed, sg);
    }
    rec.join();
    // nothing of this call stays in flight, whatever ended it
    const hipError_t e1 = hipStreamSynchronize(c->stream), e2 = hipStreamSynchronize(z->stream);
    if (!rc && rc_rec) rc = afail(a, rc_rec, "%s", err_rec);
    if (rc) return rc;
    if (e1 != hipSuccess || e2 != hipSuccess)
        return afail(a, CNV_ERR_HIP, "backward pass: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    const size_t bytes = sizeof(double) * a->cells;
    if (dE_dw0) AHIP(a, hipMemcpyAsync(dE_dw0, a->lw, bytes, hipMemcpyDeviceToHost, z->stream));
    double dU = 0.0;
    AHIP(a, hipMemcpyAsync(&dU, a->dU, sizeof(double), hipMemcpyDeviceToHost, z->stream));
    AHIP(a, hipStreamSynchronize(z->stream));
    if (dE_dU) *dE_dU = dU;
    return CNV_OK;
}

int cnv_adj_sweeps(const cnv_adjoint* a, long long* forward, long long* adjoint) {
    if (!a) return afail(nullptr, CNV_ERR_ARG, "null adjoint");
    if (forward) *forward = a->sweeps_fwd;
    if (adjoint) *adjoint = a->sweeps_adj;
    return CNV_OK;
}
