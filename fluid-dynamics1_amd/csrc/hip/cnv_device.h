// cnv_device.h -- structures shared by the gfx950 kernels and the host side of libcnv.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cnv {

constexpr int kStencilWidth = 7;  // every Diff1/Diff2 row fits a 7-column window
constexpr int kBlock = 256;       // threads per workgroup (4 wave64)

// Device-side state of one Poisson solve, kept in two slots.  SOR launch L (launch index passed
// by the host, counted from 0 per solve) has one deciding workgroup, which reads slot (L-1)&1 and
// the previous launch's per-tile partial sums and writes slot L&1; its sweeping workgroups read
// slot (L-1)&1 only (the state after launch L-2, and how many sweeps launch L-1 did).  The init
// kernels write slot 1 for launch 0.
enum SorStatus : int {
    SOR_RUNNING = 0,
    SOR_CONVERGED = 1,  // a sweep's L1 update fell below tol
    SOR_FAILED = 2,     // itmax sweeps without convergence
    SOR_FOLD = 4,       // defect correction: the FP32 iterate reached its resolution (rtol)
    SOR_DIVERGED = 8,   // a sweep's L1 update (or the iterate's norm) was NaN or infinite
    SOR_ROLLBACK = 16   // converged inside a launch before its last sweep: the next launch redoes
                        // rb_n sweeps from plane rb_src into plane `parity` (not terminal yet)
};

struct SorCtl {
    int k;              // sweeps accepted before the launch that wrote the slot
    int parity;         // terminal states (and ROLLBACK): the plane holding the final iterate
    int status;         // SorStatus after deciding on the previous launch
    int s_done;         // sweeps the writing launch performs itself (0: none)
    int launches;       // launches that did work (diagnostic)
    int itmax;          // sweep budget of this solve
    int rb_src;         // ROLLBACK: the plane to sweep from (the input of the converged launch)
    int rb_n;           // ROLLBACK: sweeps to redo
    double e_last;      // L1 update of the last accepted sweep
    double tol;         // stop when a sweep's L1 update < tol
    double rtol;        // stop ("fold") when the L1 update < rtol * L1 norm of the iterate (0: never)
    double n_last;      // L1 norm of the iterate after the last accepted sweep (FP32 planes)
};
static_assert(sizeof(SorCtl) == 64, "one state slot per 64 bytes");

// Iterate planes a solve rotates through: launch L reads plane L % NB and writes (L+1) % NB.
// FP64 three, so that the input of the launch before last is still intact when the decision on
// it arrives (a rollback redoes sweeps from it); FP32 two (its stopping rule looks at each
// launch's last sweep only, so it never rolls back).
template <typename T>
struct SorPlanes {
    static constexpr int value = sizeof(T) == 8 ? 3 : 2;
};
// Launches per stopping test of a one-cavity context's FP32 inner solves (SorArgs::test_mask + 1):
// only the launches with L % 4 == 3 test their last sweep; a batch tests every launch.
constexpr int kFp32TestEvery = 4;

// The planes of an FP64 solve, for the kernels that read its final iterate (plane fin->parity).
struct Planes {
    const double* p[3];
};

// Continuity diagnostic: max/min as order-preserving unsigned keys (exact, order independent),
// and the largest |u| or |v| of the stored velocities (a NaN orders above +infinity), which the
// blow-up check reads.
struct Diag {
    unsigned long long kmax;
    unsigned long long kmin;
    unsigned long long kspd;
};

__host__ __device__ inline unsigned long long dkey(double x) {
    unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double dkey_inv(unsigned long long k) {
    unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __builtin_bit_cast(double, b);
}

// Per-row stencil tables: row r of an operator has its nonzeros inside columns
// [start[r], start[r] + 7); coef[r*7 + m] multiplies column start[r] + m (zero-padded).
struct Stencils {
    const int* start;
    const double* d1x;  // Diff1(n, order, dx)
    const double* d1y;  // Diff1(n, order, dy)
    const double* d2x;  // Diff2(n, order, dx)
    const double* d2y;  // Diff2(n, order, dy)
    const float* f1x;   // the same tables rounded to FP32 (defect-correction transport)
    const float* f1y;
    const float* f2x;
    const float* f2y;
};

struct WallBC {
    double u1, u2, u3, u4, v1, v2, v3, v4;
};

// SOR launch arguments (one instance per precision).
template <typename T>
struct SorArgs {
    const T* src_f;      // source term: w (plain path, negate = 1) or the FP32 residual
    T* psi0;             // iterate planes (psi2: FP64 only)
    T* psi1;
    T* psi2;
    double* partial;     // [2][kMaxSweepsPerLaunch][ntiles][2]: per-tile L1 update and
                         // iterate norm, double-buffered by launch parity
    SorCtl* ctl;         // the solve's two state slots
    int launch;          // launch index within the solve
    int nr, nc, ld, tiles_i, tiles_j;  // rows (i, y) and columns (j, x) of the grid
    int negate;          // 1: f = -src (the reference's invsig before the solve)
    T c1, c2, c3, B, beta, omb;  // dy*dy, dx*dx, dx*dx*dy*dy, 2*(dx*dx+dy*dy), beta, 1-beta
    T rB;                // RN(1/B)
    int exact_rcp;       // 1: 1/B is exact (B a power of two), so x*rB == RN(x/B)
    int norm;            // FP64: also sum the iterate's L1 norm (needed when rtol > 0)
    int test_mask;       // FP32: the stopping rule is tested on the last sweep of the launches
                         // with (launch & test_mask) == test_mask only (test_mask + 1, a power
                         // of two, launches per test); the others sum nothing, leave no
                         // partials, and their decision carries the state forward.  FP64
                         // tests every sweep (0)
};

// Batched SOR launch: workgroup -> (solve, tile); inst < 0 marks padding workgroups, tile ==
// the solve's tile count its deciding workgroup.
struct SorBlock {
    int inst;
    int tile;
};

// ---- host launch wrappers (defined in the .hip files) ---------------------------------
// SOR launch shape: s sweeps per launch, nw waves per workgroup holding rb rows each; the
// workgroup's tile owns (nw*rb - 4s) rows x (128 - 4s) columns.
struct TileShape { int s, nw, rb; };
constexpr int kMaxSweepsPerLaunch = 8;
inline int tile_cols(const TileShape& t) { return 128 - 4 * t.s; }
inline int tile_rows(const TileShape& t) { return t.nw * t.rb - 4 * t.s; }



int sor_fp64_launch(const SorArgs<double>& a, TileShape shape, bool gs, hipStream_t st);
int sor_fp32_launch(const SorArgs<float>& a, TileShape shape, bool gs, hipStream_t st);
bool sor_shape_supported(TileShape shape);

int wall_vorticity_launch(double* w, const double* u, const double* v, int n, int ld,
                          const Stencils& s, const WallBC& bc, hipStream_t st);
int transport_fp64_launch(const double* w_in, double* w_out, const double* u, const double* v,
                          int n, int ld, const Stencils& s, const WallBC& bc, double rRe,
                          double dt, hipStream_t st);
int transport_fp32_launch(const double* w_in, double* w_out, const double* u, const double* v,
                          int n, int ld, const Stencils& s, const WallBC& bc, double rRe,
                          double dt, hipStream_t st);
// zero = 0 keeps psi0 (warm start) and only resets the state slots and the diagnostic
int poisson_init_launch(double* psi0, SorCtl* ctl, Diag* diag, int n, int ld, double tol,
                        int itmax, hipStream_t st, int zero = 1);
// warm start: psi0 <- psi.p[fin->parity] (the previous solve's final iterate)
int warm_start_launch(double* psi0, const Planes& psi, const SorCtl* fin, int n, int ld, hipStream_t st);
int velocity_launch(const Planes& psi, const SorCtl* fin, double* u, double* v, int n, int ld,
                    const Stencils& s, hipStream_t st);
int continuity_launch(const double* u, const double* v, Diag* diag, int n, int ld,
                      const Stencils& s, hipStream_t st);
int euler_launch(double* w, const double* dwdx, const double* dwdy, const double* d2x,
                 const double* d2y, const double* u, const double* v, double rRe, double dt,
                 long long count, hipStream_t st);
int add_launch(const double* a, const double* b, double* out, long long count, int sub,
               hipStream_t st);
int copy_launch(double* dst, const double* src, long long count, hipStream_t st);
int snapshot_launch(const double* w, const Planes& psi, const double* u, const double* v, const SorCtl* fin,
                    double* dst, long long count, hipStream_t st);

// defect correction
// refined-form velocities from psi with FP32 stencils
int velocity_fp32_launch(const double* psi, double* u, double* v, int n, int ld, const Stencils& s,
                         hipStream_t st);
int dc_residual_launch(const double* w, const double* psi, float* r32, int n, int ld, double c1,
                       double c2, double c3, double B, hipStream_t st);
int dc_correct_launch(double* psi, const float* e0, const float* e1, const SorCtl* fin, int n,
                      int ld, hipStream_t st);
// budget != nullptr: the inner solve's sweep budget is read from that device counter
int dc_inner_init_launch(float* e0, SorCtl* ctl, int n, int ld, double rtol, double tol,
                         int itmax, hipStream_t st, const int* budget = nullptr);
// the device budget counter: fin == nullptr sets it to itmax, else subtracts fin->k
int dc_budget_launch(int* budget, int itmax, const SorCtl* fin, hipStream_t st);

}  // namespace cnv
