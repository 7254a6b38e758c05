// noalloc.hip -- the Poisson solve of cnv_poisson_device without a context and without any
// device allocation (include/cnv.h: cnv_poisson_device_noalloc).
//
// The solve's scratch -- the second and third SOR iterate planes, the per-tile L1 partial sums and
// the two state slots -- lives in `scratch`, a file-scope device array of the library
// sized for grids up to kNoAllocMaxN per side; each GPU the library runs on holds its own copy
// until the process ends.  Host code never passes the array's name to a kernel (on the host it
// is only a stand-in for the symbol): the first call on each device, with that device current,
// asks the runtime for the array's device address (hipGetSymbolAddress) and keeps it per device;
// every kernel argument and copy uses that address.  Every call on one GPU shares the array, so
// calls that overlap on one GPU race; each call resets what must start at zero (the first
// iterate and the state slots, by the Poisson init kernel) and needs nothing else from the last
// call's contents: launch L writes every point of plane (L+1) % 3 before any launch reads it, and
// a launch's decider reads only partial sums that the launch before it wrote.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cnv.h"
#include "cnv_device.h"
#include "ctx.h"

using namespace cnv;

namespace {

constexpr int kNoAllocMaxN = CNV_NOALLOC_MAX_N;
constexpr size_t kPlaneDoubles = (size_t)kNoAllocMaxN * kNoAllocMaxN;
constexpr size_t kIterDoubles = 2 * kPlaneDoubles;  // planes 1 and 2 (plane 0 is the caller's psi)
// tiles of the thinnest compiled default tile (4 x 108 owned points) at the largest side
constexpr size_t kMaxTiles = 2560;
constexpr size_t kPartialDoubles = (size_t)2 * 2 * kMaxSweepsPerLaunch * kMaxTiles;
constexpr size_t kCtlDoubles = 2 * sizeof(SorCtl) / sizeof(double);
constexpr size_t kScratchDoubles = kIterDoubles + kPartialDoubles + kCtlDoubles;
static_assert(sizeof(SorCtl) % sizeof(double) == 0, "state slots pack into doubles");
static_assert(((kIterDoubles + kPartialDoubles) * sizeof(double)) % 256 == 0, "slots stay 256-byte aligned");

constexpr int kMaxDevices = 64;
constexpr int kLaunchesPerPoll = 32;

}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};  // per device: the address of that device's `scratch`

int scratch_of(int dev, double** out) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return cnv_fail(nullptr, CNV_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                            hipGetErrorString(e));
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return CNV_OK;
}

}  // namespace

int cnv_poisson_device_noalloc(int n, double dx, double dy, double beta, double tol, int itmax, int poisson_type,
                               const double* d_f, double* d_psi, int* k, double* e) {
    if (!d_f || !d_psi) return cnv_fail(nullptr, CNV_ERR_ARG, "cnv_poisson_device_noalloc: null buffer");
    if (n < 8 || n > kNoAllocMaxN)
        return cnv_fail(nullptr, CNV_ERR_ARG, "cnv_poisson_device_noalloc: n = %d outside [8, %d] (the static scratch "
                        "holds grids up to %d x %d)", n, kNoAllocMaxN, kNoAllocMaxN, kNoAllocMaxN);
    if (poisson_type != 1 && poisson_type != 2)
        return cnv_fail(nullptr, CNV_ERR_ARG, "Error - invalid option for Poisson solver (poisson_type=%d)", poisson_type);
    if (!(dx > 0) || !(dy > 0)) return cnv_fail(nullptr, CNV_ERR_ARG, "grid spacing must be positive");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return cnv_fail(nullptr, CNV_ERR_NO_DEVICE, "cnv_poisson_device_noalloc: no current HIP device");
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    // the launch shape a context of this size takes by default, so that results, k and e equal
    // cnv_poisson_device's
    const TileShape sh = cnv_auto_shape(n, cus);
    const int ti = (n + tile_rows(sh) - 1) / tile_rows(sh), tj = (n + tile_cols(sh) - 1) / tile_cols(sh);
    if ((size_t)ti * tj > kMaxTiles)
        return cnv_fail(nullptr, CNV_ERR_ARG, "cnv_poisson_device_noalloc: %d tiles exceed the scratch's %zu", ti * tj,
                        kMaxTiles);
    double* base = nullptr;
    int rc = scratch_of(dev, &base);
    if (rc) return rc;
    double* psi1 = base;
    double* psi2 = base + kPlaneDoubles;
    double* partial = base + kIterDoubles;
    SorCtl* ctl = reinterpret_cast<SorCtl*>(base + kIterDoubles + kPartialDoubles);
    hipStream_t st = nullptr;  // the null stream
    if (poisson_init_launch(d_psi, ctl, nullptr, n, n, tol, itmax, st))
        return cnv_fail(nullptr, CNV_ERR_HIP, "Poisson init launch failed: %s", hipGetErrorString(hipGetLastError()));
    SorArgs<double> a = sor_args_for<double>(n, n, dx, dy, beta, ti, tj, partial, ctl, d_f, d_psi, psi1, psi2, 0);
    const bool gs = poisson_type == 1;
    const long long cap = (long long)itmax / sh.s + 4LL * kLaunchesPerPoll + 8;  // working launches <= itmax/S + 3
    long long queued = 0;
    SorCtl fin{};
    while (true) {
        for (int l = 0; l < kLaunchesPerPoll; ++l) {
            a.launch = (int)(queued + l);
            if (sor_fp64_launch(a, sh, gs, st))
                return cnv_fail(nullptr, CNV_ERR_HIP, "SOR launch failed: %s", hipGetErrorString(hipGetLastError()));
        }
        queued += kLaunchesPerPoll;
        // the slot the last queued launch wrote; a copy into pageable host memory returns when the
        // launches before it have run
        const hipError_t ec = hipMemcpyAsync(&fin, ctl + ((queued - 1) & 1), sizeof(SorCtl), hipMemcpyDeviceToHost, st);
        if (ec != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            return cnv_fail(nullptr, CNV_ERR_HIP, "reading the SOR state failed: %s", hipGetErrorString(hipGetLastError()));
        if (sor_done(fin.status)) break;
        if (queued > cap) return cnv_fail(nullptr, CNV_ERR_HIP, "SOR state did not terminate after %lld launches", queued);
    }
    if (fin.parity != 0) {  // the final iterate is in a scratch plane
        if (hipMemcpyAsync(d_psi, fin.parity == 1 ? psi1 : psi2, sizeof(double) * n * n, hipMemcpyDeviceToDevice,
                           st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return cnv_fail(nullptr, CNV_ERR_HIP, "copying the final iterate failed: %s",
                            hipGetErrorString(hipGetLastError()));
    }
    if (k) *k = fin.k - 1;
    if (e) *e = fin.e_last;
    if (fin.status == SOR_DIVERGED)
        return cnv_fail(nullptr, CNV_DIVERGED, "Poisson solve diverged: non-finite L1 update after %d sweeps", fin.k);
    if (fin.status != SOR_CONVERGED)
        return cnv_fail(nullptr, CNV_NOT_CONVERGED, "Error: maximum number of iterations achieved for Poisson equation.");
    return CNV_OK;
}
