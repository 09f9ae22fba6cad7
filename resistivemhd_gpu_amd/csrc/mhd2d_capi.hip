// mhd2d_capi.hip -- host side of the C ABI declared in include/mhd2d.h:
// device buffers, the per-step launch sequence and host transfers.
//
// Per RK2 step (the reference's ResistiveMHD2D::oneStepRK2,
// ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_periodicXY/resistiveMHD_2D.cu:133-195):
//   stage 1 (U -> U*, CT) ; boundary x, y of U*
//   stage 2 (U*, U -> U in place) ; boundary x, y of U (+ dt of the next step)
// dt never leaves the device: stage 2 and the boundary kernels reduce the
// per-cell CFL limit of the new state into a control block, and the next
// stage 1 turns it into dt.
//
// MHD_SCHEME_MULTI_SHARE runs the MPI library's step
// (ResistiveMHD_multiGPU_cuda/ResistiveMHD2D_GPU/resistiveMHD2D.cu:147-221)
// on one subdomain that is its own neighbour on every side:
//   stage 1 (U -> U*, face fluxes F0 F5 G0 G4 -> aux[0:4]) ; Ec of smoothed U
//   halo of U* ; stage 2 (U*, U -> U2, fluxes -> aux[4:8]) ; halo of U2 ;
//   Ec of U2 ; corner EMF ; final CT of U2 from U's face B (+ next dt) ;
//   halo of U2 ; U2 becomes U.
// The self-exchange halo (M/mpi.cu:104-188 with every neighbour = self) is
// the periodic fill in x then y.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mhd2d.h"
#include "mhd2d_kernels.h"

using mhd::Ctrl;
using mhd::Grid;
using mhd::Phys;

struct mhd_handle {
    mhd_config cfg{};
    Grid g{};
    Phys ph{};
    double *U = nullptr;      // state U^n (8 planes)
    double *Us = nullptr;     // stage-1 state U*
    double *U2 = nullptr;     // MULTI_SHARE: stage-2 state (8 planes)
    double *aux = nullptr;    // MULTI_SHARE: face fluxes F0 F5 G0 G4 of stage 1, then of stage 2
    double *ecE = nullptr;    // MULTI_SHARE: cell EMF of U^n and of U2
    double *eta = nullptr;    // optional per-cell resistivity plane
    double *eta_m = nullptr;  // MULTI_SHARE: getEta at x-faces, y-faces, cells (3 planes) or null
    double *side = nullptr;   // CT value of the bottom boundary-face By per column
    double *aos = nullptr;    // AoS staging for host transfers (allocated on first use)
    Ctrl *ctrl = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    int num_cu = 0;
    int seg_request = 0;      // 0: sweep length chosen from the occupancy
    int cur = 0;              // control slot holding the per-cell dt minimum of U
    bool dt_valid = false;    // ctrl->acc[cur] matches the current U
    bool exact = true;        // numerics: exact reference arithmetic (default) or fast
    int scheme = MHD_SCHEME_SINGLE;
    bool profiling = false;
    std::vector<hipEvent_t> ev;   // 4 events per timed step: stage-1 start/end, stage-2 start/end
    std::string err;
};

namespace {

// Last error of calls without a handle: a fixed buffer, so that setting it
// allocates nothing (mhd_step_rk2_noalloc reports through it).
thread_local char g_last_error[512];

void set_last_error(const char *msg)
{
    std::snprintf(g_last_error, sizeof g_last_error, "%s", msg);
}

int fail(mhd_handle *h, int code, const std::string &msg)
{
    if (h) h->err = msg;
    set_last_error(msg.c_str());
    return code;
}

int hip_fail(mhd_handle *h, hipError_t e, const char *what)
{
    return fail(h, MHD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHECK(h, call)                                   \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #call);\
    } while (0)

constexpr int kSweepWarmup = 4;   // columns a sweep loads before its first update

// Sweep length along x.  Every (strip, segment) item is one wavefront; a
// segment of L columns costs about L + kSweepWarmup column steps.  The
// launch runs in ceil(items / slots) rounds, slots = resident stage
// wavefronts on the chip, so L is chosen to minimise rounds * (L + warm-up).
int choose_seg(int nx_int, int nstrips, long long slots)
{
    int best = nx_int;
    double best_cost = 1e300;
    for (int L = 8; L <= nx_int; ++L) {
        const long long items = (long long)nstrips * ((nx_int + L - 1) / L);
        const long long rounds = (items + slots - 1) / slots;
        const double cost = (double)rounds * (L + kSweepWarmup);
        if (cost < best_cost - 1e-9) {
            best_cost = cost;
            best = L;
        }
    }
    return best;
}

// Work decomposition of one grid (nx, ny with ghosts) of a given plane
// layout: strips, segments of seg columns (0: from the occupancy, given the
// resident stage wavefronts per CU) and the XCD-contiguous item mapping.
void layout_grid(Grid &g, int nx, int ny, int pitch, long long plane, int num_cu, int waves_per_cu, int seg)
{
    g.nx = nx;
    g.ny = ny;
    g.pitch = pitch;
    g.plane = plane;
    const int nx_int = nx - 2 * mhd::kGhost;
    g.nstrips = (ny - 2 * mhd::kGhost + mhd::kStripRows - 1) / mhd::kStripRows;
    g.tail_g = 0;
    g.tail_items = 0;
    if (seg <= 0) {
        const long long slots = (long long)(num_cu > 0 ? num_cu : 256) * waves_per_cu;
        seg = choose_seg(nx_int, g.nstrips, slots);
    }
    g.seg = seg < nx_int ? seg : nx_int;
    g.nsegs = (nx_int + g.seg - 1) / g.seg;
    // MI355X: 256 CUs in 8 XCDs of 32
    g.xcds = num_cu >= 64 ? num_cu / 32 : 1;
    if (g.xcds < 1) g.xcds = 1;
    const int items = g.nstrips * g.nsegs;
    g.items_per_xcd = (items + g.xcds - 1) / g.xcds;
}

void layout(mhd_handle *h, int seg)
{
    const int pitch = (h->cfg.ny + 7) & ~7;
    const long long plane = ((long long)h->cfg.nx * pitch + 63) & ~63LL;
    layout_grid(h->g, h->cfg.nx, h->cfg.ny, pitch, plane, h->num_cu, mhd::stage_waves_per_cu(h->exact), seg);
}

void set_phys(mhd_handle *h)
{
    h->ph.gamma = h->cfg.gamma;
    h->ph.dx = h->cfg.dx;
    h->ph.dy = h->cfg.dy;
    h->ph.eps = h->cfg.eps;
    h->ph.cfl = h->cfg.cfl;
    h->ph.eta = h->cfg.eta;
}

// nullptr if the configuration is valid, else what is wrong with it.
const char *config_error(const mhd_config *c)
{
    if (!c) return "config is NULL";
    if (c->nx < 4 * MHD_GHOST || c->ny < 4 * MHD_GHOST) return "nx and ny must be >= 12 (6 interior cells)";
    if ((long long)c->nx * c->ny > (1LL << 31)) return "nx*ny must fit in 2^31 cells";
    if (!(c->dx > 0.0) || !(c->dy > 0.0)) return "dx and dy must be > 0";
    if (!(c->gamma > 1.0)) return "gamma must be > 1";
    if (!(c->cfl > 0.0)) return "cfl must be > 0";
    if (!(c->eta >= 0.0)) return "eta must be >= 0";
    if (!(c->eps >= 0.0)) return "eps must be >= 0";
    if ((c->bc_x != MHD_BC_PERIODIC && c->bc_x != MHD_BC_SYMMETRIC) ||
        (c->bc_y != MHD_BC_PERIODIC && c->bc_y != MHD_BC_SYMMETRIC))
        return "bc_x/bc_y must be MHD_BC_PERIODIC or MHD_BC_SYMMETRIC";
    return nullptr;
}

int check_config(const mhd_config *c, std::string &why)
{
    const char *e = config_error(c);
    if (e) why = e;
    return e == nullptr;
}

int ensure_aos(mhd_handle *h)
{
    if (h->aos) return MHD_OK;
    const size_t bytes = sizeof(double) * 8 * (size_t)h->cfg.nx * h->cfg.ny;
    HIPCHECK(h, hipMalloc(&h->aos, bytes));
    return MHD_OK;
}

int apply_bc_pair(mhd_handle *h, double *u, Ctrl *ctrl, int slot)
{
    HIPCHECK(h, mhd::launch_bc_x(h->cfg.bc_x, h->exact, u, ctrl, slot, h->g, h->ph, h->stream));
    HIPCHECK(h, mhd::launch_bc_y(h->cfg.bc_y, h->exact, u, ctrl, slot, h->g, h->ph, h->stream));
    return MHD_OK;
}

bool multi(const mhd_handle *h) { return h->scheme == MHD_SCHEME_MULTI_SHARE; }

int refresh_dt(mhd_handle *h)
{
    if (h->dt_valid) return MHD_OK;
    HIPCHECK(h, mhd::launch_dt_all(h->exact, multi(h), h->U, h->ctrl, h->cur, h->g, h->ph, h->stream));
    h->dt_valid = true;
    return MHD_OK;
}

int read_ctrl(mhd_handle *h, Ctrl *out)
{
    HIPCHECK(h, hipMemcpyAsync(out, h->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return MHD_OK;
}

// ----------------------------------------------- the allocation-free step
// Scratch of mhd_step_rk2_noalloc on each GPU: U* (8 planes of at most
// kNoAllocPlane doubles), the side row of the symmetric-x boundary face
// (at most kNoAllocSide doubles) and the control block.  The array lives in
// the library's code object, so every GPU the library runs on holds it
// (32 MiB) until the process ends; its address on a device comes from
// hipGetSymbolAddress with that device current (host code never uses the
// array's name as an address).
constexpr long long kNoAllocPlane = 1LL << 19;
constexpr long long kNoAllocSide = 1LL << 13;
constexpr long long kNoAllocCtrl = 64;   // doubles reserved for the control block
constexpr long long kNoAllocDoubles = 8 * kNoAllocPlane + kNoAllocSide + kNoAllocCtrl;
static_assert(sizeof(Ctrl) <= sizeof(double) * kNoAllocCtrl, "control block does not fit its scratch slot");

static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
struct NoAllocDevice {
    double *ptr = nullptr;   // this device's copy of scratch
    int num_cu = 0;
    int waves_per_cu = 0;
};
NoAllocDevice g_noalloc[kMaxDevices];
std::mutex g_noalloc_mutex;

// The control block of a fresh run: no dt yet, no frozen-dt replay.
__global__ void noalloc_ctrl_init_kernel(Ctrl *c)
{
    if (threadIdx.x == 0) {
        c->acc[0] = c->acc[1] = mhd::kInfBits;
        c->dt = 0.0;
        c->time = 0.0;
        c->steps = 0;
        c->crashed = 0;
        c->first_nan[0] = c->first_nan[1] = 0;
        c->pad = 0;
        c->dt_in = nullptr;
        c->dt_out = nullptr;
        c->seq0 = 0;
    }
}

// dt of the last step, the time the steps advanced and the crash flag.
__global__ void noalloc_info_kernel(const Ctrl *c, double *info)
{
    if (threadIdx.x == 0) {
        info[0] = c->dt;
        info[1] = c->time;
        info[2] = c->crashed ? 1.0 : 0.0;
    }
}

int noalloc_fail(int code, const char *msg)
{
    set_last_error(msg);
    return code;
}

}  // namespace

extern "C" {

int mhd_step_rk2_noalloc(const mhd_config *cfg, double *dev_state, long long pitch, long long plane,
                         const double *dev_eta, int nsteps, void *stream, double *dev_info)
{
    char msg[256];
    if (const char *bad = config_error(cfg)) {
        std::snprintf(msg, sizeof msg, "mhd_step_rk2_noalloc: %s", bad);
        return noalloc_fail(MHD_ERR_INVALID, msg);
    }
    if (!dev_state) return noalloc_fail(MHD_ERR_INVALID, "mhd_step_rk2_noalloc: dev_state is NULL");
    if (nsteps < 0) return noalloc_fail(MHD_ERR_INVALID, "mhd_step_rk2_noalloc: nsteps must be >= 0");
    if (pitch < cfg->ny || (pitch & 1) || plane < (long long)cfg->nx * pitch || (plane & 1) ||
        ((uintptr_t)dev_state & 15) || ((uintptr_t)dev_eta & 15))
        return noalloc_fail(MHD_ERR_INVALID, "mhd_step_rk2_noalloc: needs even pitch >= ny, even plane >= nx * pitch "
                                             "and 16-byte aligned buffers");
    if (plane > kNoAllocPlane || cfg->nx > kNoAllocSide) {
        std::snprintf(msg, sizeof msg,
                      "mhd_step_rk2_noalloc: a plane of %lld doubles (nx %d) exceeds the scratch maximum of %lld "
                      "doubles per plane and nx <= %lld; use mhd_create / mhd_step_rk2",
                      plane, cfg->nx, kNoAllocPlane, kNoAllocSide);
        return noalloc_fail(MHD_ERR_INVALID, msg);
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return noalloc_fail(MHD_ERR_HIP, "mhd_step_rk2_noalloc: hipGetDevice failed");
    if (dev < 0 || dev >= kMaxDevices) return noalloc_fail(MHD_ERR_INVALID, "mhd_step_rk2_noalloc: device index too large");
    NoAllocDevice d;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mutex);
        NoAllocDevice &slot = g_noalloc[dev];
        if (!slot.ptr) {
            void *p = nullptr;
            if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess || !p) {
                std::snprintf(msg, sizeof msg, "mhd_step_rk2_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
                return noalloc_fail(MHD_ERR_HIP, msg);
            }
            int cu = 0;
            if ((e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
                return noalloc_fail(MHD_ERR_HIP, "mhd_step_rk2_noalloc: hipDeviceGetAttribute failed");
            slot.num_cu = cu;
            slot.waves_per_cu = mhd::stage_waves_per_cu(true);
            slot.ptr = static_cast<double *>(p);
        }
        d = slot;
    }
    Grid g{};
    layout_grid(g, cfg->nx, cfg->ny, (int)pitch, plane, d.num_cu, d.waves_per_cu, 0);
    Phys ph{cfg->gamma, cfg->dx, cfg->dy, cfg->eps, cfg->cfl, cfg->eta};
    double *us = d.ptr;
    double *side = d.ptr + 8 * plane;
    Ctrl *ctrl = reinterpret_cast<Ctrl *>(d.ptr + 8 * kNoAllocPlane + kNoAllocSide);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const double *eta = dev_eta;
    // the same launches as mhd_step_rk2 (exact numerics, single-GPU scheme),
    // with U* and the control block in the scratch
    hipLaunchKernelGGL(noalloc_ctrl_init_kernel, dim3(1), dim3(64), 0, s, ctrl);
    int cur = 0;
    if ((e = mhd::launch_dt_all(true, false, dev_state, ctrl, cur, g, ph, s)) != hipSuccess) goto hip_error;
    for (int n = 0; n < nsteps; ++n) {
        if ((e = mhd::launch_stage(1, true, MHD_SCHEME_SINGLE, dev_state, us, dev_state, eta, side, nullptr, ctrl, g,
                                   ph, cur, s)) != hipSuccess ||
            (e = mhd::launch_bc_x(cfg->bc_x, true, us, nullptr, 0, g, ph, s)) != hipSuccess ||
            (e = mhd::launch_bc_y(cfg->bc_y, true, us, nullptr, 0, g, ph, s)) != hipSuccess ||
            (e = mhd::launch_stage(2, true, MHD_SCHEME_SINGLE, us, dev_state, dev_state, eta, side, nullptr, ctrl, g,
                                   ph, cur, s)) != hipSuccess ||
            (e = mhd::launch_bc_x(cfg->bc_x, true, dev_state, ctrl, cur ^ 1, g, ph, s)) != hipSuccess ||
            (e = mhd::launch_bc_y(cfg->bc_y, true, dev_state, ctrl, cur ^ 1, g, ph, s)) != hipSuccess)
            goto hip_error;
        cur ^= 1;
    }
    if (dev_info) hipLaunchKernelGGL(noalloc_info_kernel, dim3(1), dim3(64), 0, s, ctrl, dev_info);
    if ((e = hipGetLastError()) != hipSuccess) goto hip_error;
    return MHD_OK;
hip_error:
    std::snprintf(msg, sizeof msg, "mhd_step_rk2_noalloc: kernel launch: %s", hipGetErrorString(e));
    return noalloc_fail(MHD_ERR_HIP, msg);
}

const char *mhd_last_error(const mhd_handle *h)
{
    return h ? h->err.c_str() : g_last_error;
}

int mhd_create(const mhd_config *cfg, mhd_handle **out)
{
    if (!out) return fail(nullptr, MHD_ERR_INVALID, "out is NULL");
    *out = nullptr;
    std::string why;
    if (!check_config(cfg, why)) return fail(nullptr, MHD_ERR_INVALID, why);
    mhd_handle *h = new (std::nothrow) mhd_handle();
    if (!h) return fail(nullptr, MHD_ERR_NOMEM, "host allocation failed");
    h->cfg = *cfg;
    set_phys(h);
    hipError_t e;
    if ((e = hipGetDevice(&h->device)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&h->num_cu, hipDeviceAttributeMultiprocessorCount, h->device)) != hipSuccess) {
        const std::string msg = std::string("mhd_create: ") + hipGetErrorString(e);
        delete h;
        return fail(nullptr, MHD_ERR_HIP, msg);
    }
    layout(h, 0);
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&h->U, sizeof(double) * 8 * (size_t)h->g.plane)) != hipSuccess ||
        (e = hipMalloc(&h->Us, sizeof(double) * 8 * (size_t)h->g.plane)) != hipSuccess ||
        (e = hipMalloc(&h->side, sizeof(double) * (size_t)h->cfg.nx)) != hipSuccess ||
        (e = hipMalloc(&h->ctrl, sizeof(Ctrl))) != hipSuccess) {
        const std::string msg = std::string("mhd_create: ") + hipGetErrorString(e);
        mhd_destroy(h);
        return fail(nullptr, e == hipErrorOutOfMemory ? MHD_ERR_NOMEM : MHD_ERR_HIP, msg);
    }
    Ctrl c0{};
    c0.acc[0] = c0.acc[1] = mhd::kInfBits;
    if ((e = hipMemsetAsync(h->U, 0, sizeof(double) * 8 * (size_t)h->g.plane, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(h->Us, 0, sizeof(double) * 8 * (size_t)h->g.plane, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(h->side, 0, sizeof(double) * (size_t)h->cfg.nx, h->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(h->ctrl, &c0, sizeof(Ctrl), hipMemcpyHostToDevice, h->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(h->stream)) != hipSuccess) {
        const std::string msg = std::string("mhd_create: ") + hipGetErrorString(e);
        mhd_destroy(h);
        return fail(nullptr, MHD_ERR_HIP, msg);
    }
    *out = h;
    return MHD_OK;
}

int mhd_destroy(mhd_handle *h)
{
    if (!h) return MHD_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    (void)hipFree(h->U);
    (void)hipFree(h->Us);
    (void)hipFree(h->U2);
    (void)hipFree(h->aux);
    (void)hipFree(h->ecE);
    (void)hipFree(h->eta);
    (void)hipFree(h->eta_m);
    (void)hipFree(h->side);
    (void)hipFree(h->aos);
    (void)hipFree(h->ctrl);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return MHD_OK;
}

int mhd_get_config(const mhd_handle *h, mhd_config *cfg)
{
    if (!h || !cfg) return fail(nullptr, MHD_ERR_INVALID, "NULL argument");
    *cfg = h->cfg;
    return MHD_OK;
}

int mhd_set_state(mhd_handle *h, const double *host_aos)
{
    if (!h || !host_aos) return fail(h, MHD_ERR_INVALID, "NULL argument");
    int rc = ensure_aos(h);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * 8 * (size_t)h->cfg.nx * h->cfg.ny;
    HIPCHECK(h, hipMemcpyAsync(h->aos, host_aos, bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(h, mhd::launch_unpack(h->aos, h->U, h->g, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    h->dt_valid = false;
    return MHD_OK;
}

int mhd_get_state(mhd_handle *h, double *host_aos)
{
    if (!h || !host_aos) return fail(h, MHD_ERR_INVALID, "NULL argument");
    int rc = ensure_aos(h);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * 8 * (size_t)h->cfg.nx * h->cfg.ny;
    HIPCHECK(h, mhd::launch_pack(h->U, h->aos, h->g, h->stream));
    HIPCHECK(h, hipMemcpyAsync(host_aos, h->aos, bytes, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return MHD_OK;
}

int mhd_set_eta_field(mhd_handle *h, const double *host_eta)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    if (host_eta && multi(h))
        return fail(h, MHD_ERR_STATE, "MHD_SCHEME_MULTI_SHARE takes eta planes at its own sampling positions "
                                      "(mhd_set_eta_staggered), not a per-cell plane");
    if (!host_eta) {
        HIPCHECK(h, hipStreamSynchronize(h->stream));
        HIPCHECK(h, hipFree(h->eta));
        h->eta = nullptr;
        return MHD_OK;
    }
    if (!h->eta) HIPCHECK(h, hipMalloc(&h->eta, sizeof(double) * (size_t)h->g.plane));
    int rc = ensure_aos(h);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * (size_t)h->cfg.nx * h->cfg.ny;
    HIPCHECK(h, hipMemcpyAsync(h->aos, host_eta, bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(h, mhd::launch_scatter_plane(h->aos, h->eta, h->g, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return MHD_OK;
}

// Ghost entries of a plane equal to their periodic images (the one-rank
// subdomain's halo is the periodic wrap, so the library evaluates the
// ghost-layer EMFs and fluxes at the images).
static bool periodic_images(const double *p, int nx, int ny)
{
    const int lx = nx - 2 * mhd::kGhost, ly = ny - 2 * mhd::kGhost;
    auto wrap = [](int a, int n, int l) { return a < mhd::kGhost ? a + l : (a >= n - mhd::kGhost ? a - l : a); };
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const int wi = wrap(i, nx, lx), wj = wrap(j, ny, ly);
            if ((wi != i || wj != j) && p[j + (size_t)i * ny] != p[wj + (size_t)wi * ny]) return false;
        }
    return true;
}

int mhd_set_eta_staggered(mhd_handle *h, const double *host_xface, const double *host_yface,
                          const double *host_cell)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    if (!multi(h)) return fail(h, MHD_ERR_STATE, "staggered eta planes belong to MHD_SCHEME_MULTI_SHARE");
    if (!host_xface && !host_yface && !host_cell) {
        HIPCHECK(h, hipStreamSynchronize(h->stream));
        HIPCHECK(h, hipFree(h->eta_m));
        h->eta_m = nullptr;
        return MHD_OK;
    }
    if (!host_xface || !host_yface || !host_cell)
        return fail(h, MHD_ERR_INVALID, "give all three eta planes (x-face, y-face, cell) or none");
    const double *src[3] = {host_xface, host_yface, host_cell};
    const char *name[3] = {"x-face", "y-face", "cell"};
    for (int q = 0; q < 3; ++q) {
        const size_t n = (size_t)h->cfg.nx * h->cfg.ny;
        for (size_t k = 0; k < n; ++k)
            if (!(src[q][k] >= 0.0))
                return fail(h, MHD_ERR_INVALID, std::string("the ") + name[q] + " eta plane has a negative or NaN entry");
        if (!periodic_images(src[q], h->cfg.nx, h->cfg.ny))
            return fail(h, MHD_ERR_INVALID, std::string("the ") + name[q] + " eta plane's ghost entries differ from "
                        "their periodic images (the subdomain's halo is the periodic wrap)");
    }
    const size_t plane = (size_t)h->g.plane;
    if (!h->eta_m) HIPCHECK(h, hipMalloc(&h->eta_m, sizeof(double) * 3 * plane));
    int rc = ensure_aos(h);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * (size_t)h->cfg.nx * h->cfg.ny;
    for (int q = 0; q < 3; ++q) {
        HIPCHECK(h, hipMemcpyAsync(h->aos, src[q], bytes, hipMemcpyHostToDevice, h->stream));
        HIPCHECK(h, mhd::launch_scatter_plane(h->aos, h->eta_m + q * plane, h->g, h->stream));
        HIPCHECK(h, hipStreamSynchronize(h->stream));
    }
    return MHD_OK;
}

int mhd_set_eta(mhd_handle *h, double eta)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    if (!(eta >= 0.0)) return fail(h, MHD_ERR_INVALID, "eta must be >= 0");
    h->cfg.eta = eta;
    set_phys(h);
    return MHD_OK;
}

int mhd_apply_boundary(mhd_handle *h, int axis, int kind)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    if ((axis != 0 && axis != 1) || (kind != MHD_BC_PERIODIC && kind != MHD_BC_SYMMETRIC))
        return fail(h, MHD_ERR_INVALID, "axis must be 0/1 and kind MHD_BC_PERIODIC/MHD_BC_SYMMETRIC");
    if (axis == 0)
        HIPCHECK(h, mhd::launch_bc_x(kind, h->exact, h->U, nullptr, 0, h->g, h->ph, h->stream));
    else
        HIPCHECK(h, mhd::launch_bc_y(kind, h->exact, h->U, nullptr, 0, h->g, h->ph, h->stream));
    h->dt_valid = false;
    return MHD_OK;
}

int mhd_set_boundary(mhd_handle *h, int bc_x, int bc_y)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    if ((bc_x != MHD_BC_PERIODIC && bc_x != MHD_BC_SYMMETRIC) || (bc_y != MHD_BC_PERIODIC && bc_y != MHD_BC_SYMMETRIC))
        return fail(h, MHD_ERR_INVALID, "bc must be MHD_BC_PERIODIC or MHD_BC_SYMMETRIC");
    if (multi(h) && (bc_x != MHD_BC_PERIODIC || bc_y != MHD_BC_PERIODIC))
        return fail(h, MHD_ERR_STATE, "MHD_SCHEME_MULTI_SHARE fills its halo periodically");
    h->cfg.bc_x = bc_x;
    h->cfg.bc_y = bc_y;
    return MHD_OK;
}

int mhd_calculate_dt(mhd_handle *h, double *dt)
{
    if (!h || !dt) return fail(h, MHD_ERR_INVALID, "NULL argument");
    HIPCHECK(h, mhd::launch_dt_all(h->exact, multi(h), h->U, h->ctrl, h->cur, h->g, h->ph, h->stream));
    h->dt_valid = true;
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    const double m = c.first_nan[h->cur] ? std::nan("") : __builtin_bit_cast(double, c.acc[h->cur]);
    if (multi(h)) {
        *dt = m * h->cfg.cfl;   // M/resistiveMHD2D.cu:541: interior cells, no diffusion limit
        return MHD_OK;
    }
    const double lim = std::fmin(h->cfg.dx * h->cfg.dx / h->cfg.eta, h->cfg.dy * h->cfg.dy / h->cfg.eta);
    *dt = std::fmin(m, lim) * h->cfg.cfl;
    return MHD_OK;
}

int mhd_step_rk2(mhd_handle *h, int nsteps)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    if (nsteps < 0) return fail(h, MHD_ERR_INVALID, "nsteps must be >= 0");
    int rc = refresh_dt(h);
    if (rc) return rc;
    const bool m = multi(h);
    const mhd::Grid &g = h->g;
    const size_t plane = (size_t)g.plane;
    for (int n = 0; n < nsteps; ++n) {
        hipEvent_t ev[4] = {};
        if (h->profiling) {
            for (int q = 0; q < 4; ++q) {
                HIPCHECK(h, hipEventCreate(&ev[q]));
                h->ev.push_back(ev[q]);
            }
        }
        double *aux1 = m ? h->aux : nullptr, *aux2 = m ? h->aux + 4 * plane : nullptr;
        double *out2 = m ? h->U2 : h->U;
        if (ev[0]) HIPCHECK(h, hipEventRecord(ev[0], h->stream));
        const double *eta_st = m ? h->eta_m : h->eta;   // MULTI: x-face and y-face planes
        const double *eta_c = h->eta_m ? h->eta_m + 2 * plane : nullptr;
        HIPCHECK(h, mhd::launch_stage(1, h->exact, h->scheme, h->U, h->Us, h->U, eta_st, h->side, aux1, h->ctrl, g,
                                      h->ph, h->cur, h->stream));
        if (ev[1]) HIPCHECK(h, hipEventRecord(ev[1], h->stream));
        if (m) HIPCHECK(h, mhd::launch_m_ec(true, h->U, eta_c, h->ecE, g, h->ph, h->stream));
        if ((rc = apply_bc_pair(h, h->Us, nullptr, 0))) return rc;
        if (ev[2]) HIPCHECK(h, hipEventRecord(ev[2], h->stream));
        HIPCHECK(h, mhd::launch_stage(2, h->exact, h->scheme, h->Us, out2, h->U, eta_st, h->side, aux2, h->ctrl, g,
                                      h->ph, h->cur, h->stream));
        if (ev[3]) HIPCHECK(h, hipEventRecord(ev[3], h->stream));
        if (m) {
            if ((rc = apply_bc_pair(h, h->U2, nullptr, 0))) return rc;
            HIPCHECK(h, mhd::launch_m_ec(false, h->U2, eta_c, h->ecE + plane, g, h->ph, h->stream));
            HIPCHECK(h, mhd::launch_m_emf_ct(h->exact, aux1, aux2, h->ecE, h->ecE + plane, h->U, h->U2, h->ctrl,
                                             h->cur ^ 1, g, h->ph, h->stream));
            if ((rc = apply_bc_pair(h, h->U2, nullptr, 0))) return rc;
            std::swap(h->U, h->U2);
        } else {
            if ((rc = apply_bc_pair(h, h->U, h->ctrl, h->cur ^ 1))) return rc;
        }
        h->cur ^= 1;
    }
    return MHD_OK;
}

int mhd_set_scheme(mhd_handle *h, int scheme)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    if (scheme != MHD_SCHEME_SINGLE && scheme != MHD_SCHEME_MULTI_SHARE)
        return fail(h, MHD_ERR_INVALID, "scheme must be MHD_SCHEME_SINGLE or MHD_SCHEME_MULTI_SHARE");
    if (scheme == MHD_SCHEME_MULTI_SHARE) {
        if (h->cfg.bc_x != MHD_BC_PERIODIC || h->cfg.bc_y != MHD_BC_PERIODIC)
            return fail(h, MHD_ERR_STATE, "MHD_SCHEME_MULTI_SHARE needs periodic x and y (its halo is the self-exchange)");
        if (h->eta) return fail(h, MHD_ERR_STATE, "MHD_SCHEME_MULTI_SHARE takes a uniform eta; clear the eta field");
        HIPCHECK(h, hipStreamSynchronize(h->stream));
        const size_t plane = sizeof(double) * (size_t)h->g.plane;
        if (!h->U2) {
            hipError_t e;
            if ((e = hipMalloc(&h->U2, 8 * plane)) != hipSuccess || (e = hipMalloc(&h->aux, 8 * plane)) != hipSuccess ||
                (e = hipMalloc(&h->ecE, 2 * plane)) != hipSuccess) {
                (void)hipFree(h->U2);
                (void)hipFree(h->aux);
                (void)hipFree(h->ecE);
                h->U2 = h->aux = h->ecE = nullptr;
                return fail(h, e == hipErrorOutOfMemory ? MHD_ERR_NOMEM : MHD_ERR_HIP,
                            std::string("mhd_set_scheme: ") + hipGetErrorString(e));
            }
            HIPCHECK(h, hipMemsetAsync(h->U2, 0, 8 * plane, h->stream));
            HIPCHECK(h, hipMemsetAsync(h->aux, 0, 8 * plane, h->stream));
            HIPCHECK(h, hipMemsetAsync(h->ecE, 0, 2 * plane, h->stream));
            HIPCHECK(h, hipStreamSynchronize(h->stream));
        }
    }
    if (scheme == MHD_SCHEME_SINGLE && h->eta_m) {
        HIPCHECK(h, hipStreamSynchronize(h->stream));
        HIPCHECK(h, hipFree(h->eta_m));   // the MPI library's getEta planes have no meaning there
        h->eta_m = nullptr;
    }
    h->scheme = scheme;
    h->dt_valid = false;   // the two schemes reduce dt over different cells
    return MHD_OK;
}

int mhd_get_scheme(const mhd_handle *h, int *scheme)
{
    if (!h || !scheme) return fail(nullptr, MHD_ERR_INVALID, "NULL argument");
    *scheme = h->scheme;
    return MHD_OK;
}

int mhd_profile_enable(mhd_handle *h, int on)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    h->ev.clear();
    h->profiling = on != 0;
    return MHD_OK;
}

int mhd_profile_read(mhd_handle *h, double *ms_stage1, double *ms_stage2, int *steps)
{
    if (!h || !ms_stage1 || !ms_stage2 || !steps) return fail(h, MHD_ERR_INVALID, "NULL argument");
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    double s1 = 0.0, s2 = 0.0;
    for (size_t q = 0; q + 3 < h->ev.size(); q += 4) {
        float a = 0.f, b = 0.f;
        HIPCHECK(h, hipEventElapsedTime(&a, h->ev[q], h->ev[q + 1]));
        HIPCHECK(h, hipEventElapsedTime(&b, h->ev[q + 2], h->ev[q + 3]));
        s1 += a;
        s2 += b;
    }
    *ms_stage1 = s1;
    *ms_stage2 = s2;
    *steps = (int)(h->ev.size() / 4);
    return MHD_OK;
}

int mhd_one_step_rk2(mhd_handle *h) { return mhd_step_rk2(h, 1); }

int mhd_get_dt(mhd_handle *h, double *dt)
{
    if (!h || !dt) return fail(h, MHD_ERR_INVALID, "NULL argument");
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    *dt = c.dt;
    return MHD_OK;
}

int mhd_get_time(mhd_handle *h, double *t)
{
    if (!h || !t) return fail(h, MHD_ERR_INVALID, "NULL argument");
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    *t = c.time;
    return MHD_OK;
}

int mhd_set_time(mhd_handle *h, double t)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    c.time = t;
    HIPCHECK(h, hipMemcpyAsync(h->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return MHD_OK;
}

int mhd_check_crashed(mhd_handle *h, int *crashed)
{
    if (!h || !crashed) return fail(h, MHD_ERR_INVALID, "NULL argument");
    int rc = refresh_dt(h);   // a state set from outside is scanned for NaN here
    if (rc) return rc;
    Ctrl c;
    rc = read_ctrl(h, &c);
    if (rc) return rc;
    *crashed = c.crashed ? 1 : 0;
    if (multi(h) && c.dt < 0.0) *crashed = 1;   // M/resistiveMHD2D.cu:576
    return MHD_OK;
}

int mhd_save(mhd_handle *h, const char *dir, const char *name, int step)
{
    if (!h || !dir || !name) return fail(h, MHD_ERR_INVALID, "NULL argument");
    std::vector<double> buf;
    try {
        buf.resize(8 * (size_t)h->cfg.nx * h->cfg.ny);
    } catch (...) {
        return fail(h, MHD_ERR_NOMEM, "host allocation failed");
    }
    int rc = mhd_get_state(h, buf.data());
    if (rc) return rc;
    const std::string path = std::string(dir) + "/" + name + "_" + std::to_string(step) + ".bin";
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(h, MHD_ERR_IO, "cannot open " + path + ": " + std::strerror(errno));
    const size_t n = std::fwrite(buf.data(), sizeof(double), buf.size(), f);
    const int closed = std::fclose(f);
    if (n != buf.size() || closed != 0) return fail(h, MHD_ERR_IO, "short write to " + path);
    return MHD_OK;
}

int mhd_load(mhd_handle *h, const char *path)
{
    if (!h || !path) return fail(h, MHD_ERR_INVALID, "NULL argument");
    std::vector<double> buf;
    try {
        buf.resize(8 * (size_t)h->cfg.nx * h->cfg.ny);
    } catch (...) {
        return fail(h, MHD_ERR_NOMEM, "host allocation failed");
    }
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(h, MHD_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    const size_t n = std::fread(buf.data(), sizeof(double), buf.size(), f);
    const bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    if (n != buf.size() || extra)
        return fail(h, MHD_ERR_IO, std::string(path) + ": size does not match nx*ny*8 doubles");
    return mhd_set_state(h, buf.data());
}

// Per-rank file of the MPI library's ResistiveMHD2D::save
// (ResistiveMHD_multiGPU_cuda/ResistiveMHD2D_GPU/resistiveMHD2D.cu:457-486):
// {dir}/{name}_{step}_{rank}.bin holding the interior cells only, i-major,
// 8 doubles per cell.
int mhd_save_subdomain(mhd_handle *h, const char *dir, const char *name, int step, int rank)
{
    if (!h || !dir || !name) return fail(h, MHD_ERR_INVALID, "NULL argument");
    const int nx = h->cfg.nx, ny = h->cfg.ny, g = MHD_GHOST;
    std::vector<double> buf, out;
    try {
        buf.resize(8 * (size_t)nx * ny);
        out.reserve(8 * (size_t)(nx - 2 * g) * (ny - 2 * g));
    } catch (...) {
        return fail(h, MHD_ERR_NOMEM, "host allocation failed");
    }
    int rc = mhd_get_state(h, buf.data());
    if (rc) return rc;
    for (int i = g; i < nx - g; ++i)
        out.insert(out.end(), buf.begin() + 8 * ((size_t)i * ny + g), buf.begin() + 8 * ((size_t)i * ny + ny - g));
    const std::string path = std::string(dir) + "/" + name + "_" + std::to_string(step) + "_" + std::to_string(rank) +
                             ".bin";
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(h, MHD_ERR_IO, "cannot open " + path + ": " + std::strerror(errno));
    const size_t n = std::fwrite(out.data(), sizeof(double), out.size(), f);
    const int closed = std::fclose(f);
    if (n != out.size() || closed != 0) return fail(h, MHD_ERR_IO, "short write to " + path);
    return MHD_OK;
}

int mhd_load_subdomain(mhd_handle *h, const char *path)
{
    if (!h || !path) return fail(h, MHD_ERR_INVALID, "NULL argument");
    const int nx = h->cfg.nx, ny = h->cfg.ny, g = MHD_GHOST;
    std::vector<double> buf, in;
    try {
        buf.resize(8 * (size_t)nx * ny);
        in.resize(8 * (size_t)(nx - 2 * g) * (ny - 2 * g));
    } catch (...) {
        return fail(h, MHD_ERR_NOMEM, "host allocation failed");
    }
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(h, MHD_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    const size_t n = std::fread(in.data(), sizeof(double), in.size(), f);
    const bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    if (n != in.size() || extra)
        return fail(h, MHD_ERR_IO, std::string(path) + ": size does not match the interior's cells * 8 doubles");
    int rc = mhd_get_state(h, buf.data());
    if (rc) return rc;
    const size_t row = 8 * (size_t)(ny - 2 * g);
    for (int i = g; i < nx - g; ++i)
        std::copy(in.begin() + (i - g) * row, in.begin() + (i - g + 1) * row, buf.begin() + 8 * ((size_t)i * ny + g));
    return mhd_set_state(h, buf.data());
}

int mhd_synchronize(mhd_handle *h)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return MHD_OK;
}

int mhd_init_orszag_tang(mhd_handle *h)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    std::vector<double> buf(8 * (size_t)h->cfg.nx * h->cfg.ny);
    int rc = mhd_fill_orszag_tang(&h->cfg, buf.data());
    if (rc) return fail(h, rc, "mhd_fill_orszag_tang failed");
    if ((rc = mhd_set_state(h, buf.data()))) return rc;
    if ((rc = mhd_apply_boundary(h, 0, h->cfg.bc_x))) return rc;
    return mhd_apply_boundary(h, 1, h->cfg.bc_y);
}

int mhd_init_harris(mhd_handle *h, unsigned long long seed, double noise_amp, double eta0, double eta1)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    std::vector<double> buf(8 * (size_t)h->cfg.nx * h->cfg.ny);
    int rc = mhd_fill_harris(&h->cfg, seed, noise_amp, buf.data());
    if (rc) return fail(h, rc, "mhd_fill_harris failed");
    if ((rc = mhd_set_state(h, buf.data()))) return rc;
    std::vector<double> eta((size_t)h->cfg.nx * h->cfg.ny);
    if ((rc = mhd_fill_harris_eta(&h->cfg, eta0, eta1, eta.data()))) return fail(h, rc, "mhd_fill_harris_eta failed");
    if ((rc = mhd_set_eta_field(h, eta.data()))) return rc;
    if ((rc = mhd_apply_boundary(h, 0, h->cfg.bc_x))) return rc;
    return mhd_apply_boundary(h, 1, h->cfg.bc_y);
}

int mhd_init_turbulence(mhd_handle *h, unsigned long long seed, int kmax, double rms_v, double rms_b)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    std::vector<double> buf;
    try {
        buf.resize(8 * (size_t)h->cfg.nx * h->cfg.ny);
    } catch (...) {
        return fail(h, MHD_ERR_NOMEM, "host allocation failed");
    }
    int rc = mhd_fill_turbulence(&h->cfg, seed, kmax, rms_v, rms_b, buf.data());
    if (rc) return fail(h, rc, "mhd_fill_turbulence: invalid arguments (kmax must be >= 1 and < half the interior size)");
    if ((rc = mhd_set_state(h, buf.data()))) return rc;
    if ((rc = mhd_apply_boundary(h, 0, h->cfg.bc_x))) return rc;
    return mhd_apply_boundary(h, 1, h->cfg.bc_y);
}

int mhd_set_sweep_length(mhd_handle *h, int columns)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    if (columns < 0 || columns > 1 << 20) return fail(h, MHD_ERR_INVALID, "columns out of range");
    h->seg_request = columns;
    layout(h, columns);
    return MHD_OK;
}

int mhd_set_numerics(mhd_handle *h, int mode)
{
    if (!h) return fail(h, MHD_ERR_INVALID, "NULL handle");
    if (mode != MHD_NUMERICS_FAST && mode != MHD_NUMERICS_EXACT)
        return fail(h, MHD_ERR_INVALID, "mode must be MHD_NUMERICS_FAST or MHD_NUMERICS_EXACT");
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    h->exact = mode == MHD_NUMERICS_EXACT;
    h->dt_valid = false;   // the next dt is reduced with the new arithmetic
    layout(h, h->seg_request);   // occupancy differs between the two kernels
    return MHD_OK;
}

int mhd_face_regions(mhd_handle *h, int mode, unsigned char *host_rx, unsigned char *host_ry)
{
    if (!h || !host_rx || !host_ry) return fail(h, MHD_ERR_INVALID, "NULL argument");
    if (mode != MHD_NUMERICS_FAST && mode != MHD_NUMERICS_EXACT)
        return fail(h, MHD_ERR_INVALID, "mode must be MHD_NUMERICS_FAST or MHD_NUMERICS_EXACT");
    const size_t n = (size_t)h->cfg.nx * h->cfg.ny;
    unsigned char *d = nullptr;
    if (hipMalloc(&d, 2 * n) != hipSuccess) return fail(h, MHD_ERR_NOMEM, "face region buffers");
    hipError_t e = mhd::launch_face_regions(mode == MHD_NUMERICS_EXACT, h->U, d, d + n, h->g, h->ph, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host_rx, d, n, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host_ry, d + n, n, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(h, e, "mhd_face_regions");
    return MHD_OK;
}

int mhd_get_numerics(const mhd_handle *h, int *mode)
{
    if (!h || !mode) return fail(nullptr, MHD_ERR_INVALID, "NULL argument");
    *mode = h->exact ? MHD_NUMERICS_EXACT : MHD_NUMERICS_FAST;
    return MHD_OK;
}

int mhd_get_sweep_length(const mhd_handle *h, int *columns)
{
    if (!h || !columns) return fail(nullptr, MHD_ERR_INVALID, "NULL argument");
    *columns = h->g.seg;
    return MHD_OK;
}

}  // extern "C"
