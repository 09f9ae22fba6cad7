// mhd2d_batch.hip -- host side of the batched form (mhd_batch_* in
// include/mhd2d.h): many independent runs of uneven grid sizes, each with its
// own constants, resistivity, boundaries and dt, advanced together.
//
// Every RK stage is ONE launch over a work list holding the (strip, segment)
// items of all runs back to back (rk_stage_batch_kernel); each boundary
// pass is one launch over all runs.  A step is therefore 6 launches whatever
// the number of runs, against 6 per run for a loop over single runs, and
// small runs no longer leave the GPU idle: their items share the machine
// with those of the large runs.  Steps are replayed from a captured hipGraph
// (two steps per graph, so the control slot the graph starts from comes
// back).  The sweep length (items' extent along x) is chosen by
// mhd_batch_autotune, which times candidates on the batch itself.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/mhd2d.h"
#include "mhd2d_kernels.h"

using mhd::Ctrl;
using mhd::Grid;
using mhd::Phys;
using mhd::RunDesc;

#include "mhd2d_batch_impl.h"

namespace {

thread_local std::string g_batch_error;

int bfail(mhd_batch *b, int code, const std::string &msg)
{
    if (b) b->err = msg;
    g_batch_error = msg;
    return code;
}

#define BHIP(b, call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return bfail(b, MHD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kG = mhd::kGhost;

int plane_pitch(int ny) { return (ny + 7) & ~7; }
long long plane_size(int nx, int ny) { return ((long long)nx * plane_pitch(ny) + 63) & ~63LL; }

void drop_graph(mhd_batch *b)
{
    for (hipGraphExec_t &g : b->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

// Wavefronts per XCD of the stage launches for b->waves_per_cu.
void set_launch_shape(mhd_batch *b)
{
    const int cu_per_xcd = std::max(1, (b->num_cu > 0 ? b->num_cu : 256) / b->xcds);
    b->blocks_per_xcd = b->waves_per_cu <= 0 ? b->items_per_xcd
                                              : std::max(1, std::min(b->items_per_xcd, cu_per_xcd * b->waves_per_cu));
}

// Work lists for sweep length L: run r's full-strip items are [first_item,
// first_item + nstrips * nsegs), strips fastest as in the single-run kernel;
// its packed last strip's items are [first_tail, first_tail + tail_items) of
// the second list.  Last strips are packed only when the full-strip launch
// would take several rounds of resident wavefronts: packing saves lanes,
// which shortens a launch only by whole rounds, and costs a launch of its
// own.
int build_work_list(mhd_batch *b, int L)
{
    const long long slots = (long long)(b->num_cu > 0 ? b->num_cu : 256) * mhd::stage_batch_waves_per_cu(b->exact);
    long long unpacked = 0;
    for (int r = 0; r < b->nruns; ++r) {
        Grid g = b->desc[r].g;
        const int nx_int = g.nx - 2 * kG;
        mhd::set_strips(g, g.ny - 2 * kG, false);
        mhd::set_segs(g, nx_int, std::max(1, std::min(L, nx_int)));
        unpacked += mhd::grid_items(g);
    }
    const bool pack = b->strip_packing == MHD_STRIP_PACK_ALWAYS ||
                      (b->strip_packing == MHD_STRIP_PACK_AUTO && unpacked > 2 * slots);
    std::vector<int> item_run, tail_run;
    for (int r = 0; r < b->nruns; ++r) {
        Grid &g = b->desc[r].g;
        const int nx_int = g.nx - 2 * kG;
        mhd::set_strips(g, g.ny - 2 * kG, pack);
        mhd::set_segs(g, nx_int, std::max(1, std::min(L, nx_int)));
        b->desc[r].first_item = (int)item_run.size();
        item_run.insert(item_run.end(), (size_t)g.nstrips * g.nsegs, r);
        b->desc[r].first_tail = (int)tail_run.size();
        tail_run.insert(tail_run.end(), (size_t)g.tail_items, r);
    }
    b->tail_items = (int)tail_run.size();
    if (b->tail_items > b->tail_cap) {
        BHIP(b, hipStreamSynchronize(b->stream));
        BHIP(b, hipFree(b->d_tail_run));
        b->d_tail_run = nullptr;
        BHIP(b, hipMalloc(&b->d_tail_run, sizeof(int) * (size_t)b->tail_items));
        b->tail_cap = b->tail_items;
    }
    if (b->tail_items)
        BHIP(b, hipMemcpyAsync(b->d_tail_run, tail_run.data(), sizeof(int) * tail_run.size(), hipMemcpyHostToDevice,
                               b->stream));
    b->seg = L;
    b->total_items = (int)item_run.size();
    b->xcds = b->num_cu >= 64 ? b->num_cu / 32 : 1;
    b->items_per_xcd = (b->total_items + b->xcds - 1) / b->xcds;
    set_launch_shape(b);
    if (b->total_items > b->item_cap) {
        BHIP(b, hipStreamSynchronize(b->stream));
        BHIP(b, hipFree(b->d_item_run));
        b->d_item_run = nullptr;
        BHIP(b, hipMalloc(&b->d_item_run, sizeof(int) * (size_t)b->total_items));
        b->item_cap = b->total_items;
    }
    BHIP(b, hipMemcpyAsync(b->d_item_run, item_run.data(), sizeof(int) * item_run.size(), hipMemcpyHostToDevice,
                           b->stream));
    BHIP(b, hipMemcpyAsync(b->d_desc, b->desc.data(), sizeof(RunDesc) * b->desc.size(), hipMemcpyHostToDevice,
                           b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    drop_graph(b);
    return MHD_OK;
}

// Default sweep length: minimise rounds * (L + warm-up) over the whole work
// list, as the single-run layout does for one grid.
int default_seg(const mhd_batch *b)
{
    const long long slots = (long long)(b->num_cu > 0 ? b->num_cu : 256) * mhd::stage_batch_waves_per_cu(b->exact);
    int best = 8;
    double best_cost = 1e300;
    for (int L = 8; L <= 512; L += 4) {
        long long items = 0;
        for (int r = 0; r < b->nruns; ++r) {
            Grid g = b->desc[r].g;
            const int nx_int = g.nx - 2 * kG;
            mhd::set_strips(g, g.ny - 2 * kG, false);
            mhd::set_segs(g, nx_int, std::max(1, std::min(L, nx_int)));
            items += mhd::grid_items(g);
        }
        const long long rounds = (items + slots - 1) / slots;
        const double cost = (double)rounds * (std::min(L, b->max_nx) + 4);
        if (cost < best_cost - 1e-9) {
            best_cost = cost;
            best = L;
        }
    }
    return best;
}

}  // namespace

int mhd_batch_refresh_dt(mhd_batch *b)
{
    if (b->dt_valid) return MHD_OK;
    BHIP(b, mhd::launch_dt_all_batch(b->exact, b->d_desc, b->nruns, b->max_nx, b->cur, b->stream));
    b->dt_valid = true;
    return MHD_OK;
}

namespace {

int refresh_dt(mhd_batch *b) { return mhd_batch_refresh_dt(b); }

int enqueue_step(mhd_batch *b, int cur)
{
    BHIP(b, mhd::launch_stage_batch(1, b->exact, b->d_desc, b->d_item_run, b->total_items, b->xcds,
                                    b->items_per_xcd, b->blocks_per_xcd, cur, b->stream, b->d_tail_run, b->tail_items));
    BHIP(b, mhd::launch_bc_batch(b->exact, b->d_desc, b->nruns, b->max_nx, b->max_ny, true, false, 0, b->stream));
    BHIP(b, mhd::launch_stage_batch(2, b->exact, b->d_desc, b->d_item_run, b->total_items, b->xcds,
                                    b->items_per_xcd, b->blocks_per_xcd, cur, b->stream, b->d_tail_run, b->tail_items));
    BHIP(b, mhd::launch_bc_batch(b->exact, b->d_desc, b->nruns, b->max_nx, b->max_ny, false, true, cur ^ 1,
                                 b->stream));
    return MHD_OK;
}

constexpr int kGraphSteps = 2;   // even: the captured graph ends on the slot it started from

// Captures kGraphSteps steps starting from control slot b->cur (one graph per
// starting slot, so an odd number of steps between calls costs no recapture).
int capture_graph(mhd_batch *b)
{
    hipGraphExec_t &exec = b->graph[b->cur];
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    hipGraph_t graph = nullptr;
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = MHD_OK;
    for (int k = 0; k < kGraphSteps && rc == MHD_OK; ++k) rc = enqueue_step(b, b->cur ^ (k & 1));
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, MHD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        exec = nullptr;
        return bfail(b, MHD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return MHD_OK;
}

int check_run(mhd_batch *b, int run)
{
    if (!b) return bfail(nullptr, MHD_ERR_INVALID, "NULL batch");
    if (run < 0 || run >= b->nruns) return bfail(b, MHD_ERR_INVALID, "run index out of range");
    return MHD_OK;
}

int read_ctrl(mhd_batch *b, int run, Ctrl *out)
{
    BHIP(b, hipMemcpyAsync(out, b->desc[run].ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return MHD_OK;
}

}  // namespace

extern "C" {

const char *mhd_batch_last_error(const mhd_batch *b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

int mhd_batch_create(const mhd_config *cfgs, int nruns, mhd_batch **out)
{
    if (!out) return bfail(nullptr, MHD_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!cfgs || nruns < 1) return bfail(nullptr, MHD_ERR_INVALID, "need at least one run configuration");
    for (int r = 0; r < nruns; ++r) {
        const mhd_config &c = cfgs[r];
        if (c.nx < 4 * MHD_GHOST || c.ny < 4 * MHD_GHOST || (long long)c.nx * c.ny > (1LL << 31) || !(c.dx > 0.0) ||
            !(c.dy > 0.0) || !(c.gamma > 1.0) || !(c.cfl > 0.0) || !(c.eta >= 0.0) || !(c.eps >= 0.0) ||
            (c.bc_x != MHD_BC_PERIODIC && c.bc_x != MHD_BC_SYMMETRIC) ||
            (c.bc_y != MHD_BC_PERIODIC && c.bc_y != MHD_BC_SYMMETRIC))
            return bfail(nullptr, MHD_ERR_INVALID, "invalid configuration of run " + std::to_string(r));
    }
    mhd_batch *b = new (std::nothrow) mhd_batch();
    if (!b) return bfail(nullptr, MHD_ERR_NOMEM, "host allocation failed");
    b->nruns = nruns;
    b->cfg.assign(cfgs, cfgs + nruns);
    b->desc.resize(nruns);
    b->eta.assign(nruns, nullptr);
    hipError_t e;
    if ((e = hipGetDevice(&b->device)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&b->num_cu, hipDeviceAttributeMultiprocessorCount, b->device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) {
        const std::string msg = std::string("mhd_batch_create: ") + hipGetErrorString(e);
        mhd_batch_destroy(b);
        return bfail(nullptr, MHD_ERR_HIP, msg);
    }
    // one pool: per run U (8 planes), U* (8 planes), side row
    std::vector<size_t> off(nruns);
    size_t total = 0, max_cells = 0;
    for (int r = 0; r < nruns; ++r) {
        off[r] = total;
        const long long plane = plane_size(cfgs[r].nx, cfgs[r].ny);
        total += (size_t)16 * plane + (((size_t)cfgs[r].nx + 63) & ~(size_t)63);
        max_cells = std::max(max_cells, (size_t)cfgs[r].nx * cfgs[r].ny);
        b->max_nx = std::max(b->max_nx, cfgs[r].nx);
        b->max_ny = std::max(b->max_ny, cfgs[r].ny);
    }
    if ((e = hipMalloc(&b->pool, sizeof(double) * total)) != hipSuccess ||
        (e = hipMalloc(&b->ctrl, sizeof(Ctrl) * (size_t)nruns)) != hipSuccess ||
        (e = hipMalloc(&b->d_desc, sizeof(RunDesc) * (size_t)nruns)) != hipSuccess ||
        (e = hipMalloc(&b->aos, sizeof(double) * 8 * max_cells)) != hipSuccess ||
        (e = hipMemsetAsync(b->pool, 0, sizeof(double) * total, b->stream)) != hipSuccess) {
        const std::string msg = std::string("mhd_batch_create: ") + hipGetErrorString(e);
        mhd_batch_destroy(b);
        return bfail(nullptr, e == hipErrorOutOfMemory ? MHD_ERR_NOMEM : MHD_ERR_HIP, msg);
    }
    std::vector<Ctrl> c0(nruns);
    for (Ctrl &c : c0) {
        std::memset(&c, 0, sizeof c);
        c.acc[0] = c.acc[1] = mhd::kInfBits;
    }
    for (int r = 0; r < nruns; ++r) {
        const mhd_config &c = cfgs[r];
        RunDesc &d = b->desc[r];
        std::memset(&d, 0, sizeof d);
        d.g.nx = c.nx;
        d.g.ny = c.ny;
        d.g.pitch = plane_pitch(c.ny);
        d.g.plane = plane_size(c.nx, c.ny);
        mhd::set_strips(d.g, c.ny - 2 * kG, false);   // build_work_list decides on packing
        d.g.xcds = 1;
        d.ph.gamma = c.gamma;
        d.ph.dx = c.dx;
        d.ph.dy = c.dy;
        d.ph.eps = c.eps;
        d.ph.cfl = c.cfl;
        d.ph.eta = c.eta;
        d.u = b->pool + off[r];
        d.us = d.u + 8 * d.g.plane;
        d.side = d.us + 8 * d.g.plane;
        d.eta = nullptr;
        d.ctrl = b->ctrl + r;
        d.bc_x = c.bc_x;
        d.bc_y = c.bc_y;
    }
    if ((e = hipMemcpyAsync(b->ctrl, c0.data(), sizeof(Ctrl) * c0.size(), hipMemcpyHostToDevice, b->stream)) !=
        hipSuccess) {
        const std::string msg = std::string("mhd_batch_create: ") + hipGetErrorString(e);
        mhd_batch_destroy(b);
        return bfail(nullptr, MHD_ERR_HIP, msg);
    }
    int rc = build_work_list(b, default_seg(b));
    if (rc) {
        const std::string msg = b->err;
        mhd_batch_destroy(b);
        return bfail(nullptr, rc, msg);
    }
    *out = b;
    return MHD_OK;
}

int mhd_batch_destroy(mhd_batch *b)
{
    if (!b) return MHD_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (double *p : b->eta) (void)hipFree(p);
    (void)hipFree(b->pool);
    (void)hipFree(b->ctrl);
    (void)hipFree(b->d_desc);
    (void)hipFree(b->d_item_run);
    (void)hipFree(b->d_tail_run);
    (void)hipFree(b->aos);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return MHD_OK;
}

int mhd_batch_size(const mhd_batch *b, int *nruns)
{
    if (!b || !nruns) return bfail(nullptr, MHD_ERR_INVALID, "NULL argument");
    *nruns = b->nruns;
    return MHD_OK;
}

int mhd_batch_set_state(mhd_batch *b, int run, const double *host_aos)
{
    int rc = check_run(b, run);
    if (rc) return rc;
    if (!host_aos) return bfail(b, MHD_ERR_INVALID, "NULL state");
    const RunDesc &d = b->desc[run];
    const size_t bytes = sizeof(double) * 8 * (size_t)d.g.nx * d.g.ny;
    BHIP(b, hipMemcpyAsync(b->aos, host_aos, bytes, hipMemcpyHostToDevice, b->stream));
    BHIP(b, mhd::launch_unpack(b->aos, d.u, d.g, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    b->dt_valid = false;
    return MHD_OK;
}

int mhd_batch_get_state(mhd_batch *b, int run, double *host_aos)
{
    int rc = check_run(b, run);
    if (rc) return rc;
    if (!host_aos) return bfail(b, MHD_ERR_INVALID, "NULL state");
    const RunDesc &d = b->desc[run];
    const size_t bytes = sizeof(double) * 8 * (size_t)d.g.nx * d.g.ny;
    BHIP(b, mhd::launch_pack(d.u, b->aos, d.g, b->stream));
    BHIP(b, hipMemcpyAsync(host_aos, b->aos, bytes, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return MHD_OK;
}

// ResistiveMHD2D::save (P/resistiveMHD_2D.cu:198-244) for one run: the
// same file layout as mhd_save.
int mhd_batch_save(mhd_batch *b, int run, const char *dir, const char *name, int step)
{
    int rc = check_run(b, run);
    if (rc) return rc;
    if (!dir || !name) return bfail(b, MHD_ERR_INVALID, "NULL argument");
    const Grid &g = b->desc[run].g;
    std::vector<double> buf;
    try {
        buf.resize(8 * (size_t)g.nx * g.ny);
    } catch (...) {
        return bfail(b, MHD_ERR_NOMEM, "host allocation failed");
    }
    if ((rc = mhd_batch_get_state(b, run, buf.data()))) return rc;
    const std::string path = std::string(dir) + "/" + name + "_" + std::to_string(step) + ".bin";
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return bfail(b, MHD_ERR_IO, "cannot open " + path + ": " + std::strerror(errno));
    const size_t n = std::fwrite(buf.data(), sizeof(double), buf.size(), f);
    const int closed = std::fclose(f);
    if (n != buf.size() || closed != 0) return bfail(b, MHD_ERR_IO, "short write to " + path);
    return MHD_OK;
}

int mhd_batch_load(mhd_batch *b, int run, const char *path)
{
    int rc = check_run(b, run);
    if (rc) return rc;
    if (!path) return bfail(b, MHD_ERR_INVALID, "NULL argument");
    const Grid &g = b->desc[run].g;
    std::vector<double> buf;
    try {
        buf.resize(8 * (size_t)g.nx * g.ny);
    } catch (...) {
        return bfail(b, MHD_ERR_NOMEM, "host allocation failed");
    }
    FILE *f = std::fopen(path, "rb");
    if (!f) return bfail(b, MHD_ERR_IO, std::string("cannot open ") + path + ": " + std::strerror(errno));
    const size_t n = std::fread(buf.data(), sizeof(double), buf.size(), f);
    const bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    if (n != buf.size() || extra)
        return bfail(b, MHD_ERR_IO, std::string(path) + ": size does not match nx*ny*8 doubles");
    return mhd_batch_set_state(b, run, buf.data());
}

int mhd_batch_set_eta_field(mhd_batch *b, int run, const double *host_eta)
{
    int rc = check_run(b, run);
    if (rc) return rc;
    RunDesc &d = b->desc[run];
    BHIP(b, hipStreamSynchronize(b->stream));
    if (!host_eta) {
        BHIP(b, hipFree(b->eta[run]));
        b->eta[run] = nullptr;
    } else {
        if (!b->eta[run]) BHIP(b, hipMalloc(&b->eta[run], sizeof(double) * (size_t)d.g.plane));
        BHIP(b, hipMemcpyAsync(b->aos, host_eta, sizeof(double) * (size_t)d.g.nx * d.g.ny, hipMemcpyHostToDevice,
                               b->stream));
        BHIP(b, mhd::launch_scatter_plane(b->aos, b->eta[run], d.g, b->stream));
    }
    d.eta = b->eta[run];
    BHIP(b, hipMemcpyAsync(b->d_desc + run, &d, sizeof(RunDesc), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return MHD_OK;
}

int mhd_batch_apply_boundary(mhd_batch *b)
{
    if (!b) return bfail(nullptr, MHD_ERR_INVALID, "NULL batch");
    BHIP(b, mhd::launch_bc_batch(b->exact, b->d_desc, b->nruns, b->max_nx, b->max_ny, false, false, 0, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    b->dt_valid = false;
    return MHD_OK;
}

int mhd_batch_init_orszag_tang(mhd_batch *b, int run)
{
    int rc = check_run(b, run);
    if (rc) return rc;
    std::vector<double> buf(8 * (size_t)b->cfg[run].nx * b->cfg[run].ny);
    if ((rc = mhd_fill_orszag_tang(&b->cfg[run], buf.data()))) return bfail(b, rc, "mhd_fill_orszag_tang failed");
    return mhd_batch_set_state(b, run, buf.data());
}

int mhd_batch_init_harris(mhd_batch *b, int run, unsigned long long seed, double noise_amp)
{
    int rc = check_run(b, run);
    if (rc) return rc;
    std::vector<double> buf(8 * (size_t)b->cfg[run].nx * b->cfg[run].ny);
    if ((rc = mhd_fill_harris(&b->cfg[run], seed, noise_amp, buf.data()))) return bfail(b, rc, "mhd_fill_harris failed");
    return mhd_batch_set_state(b, run, buf.data());
}

int mhd_batch_step(mhd_batch *b, int nsteps)
{
    if (!b) return bfail(nullptr, MHD_ERR_INVALID, "NULL batch");
    if (nsteps < 0) return bfail(b, MHD_ERR_INVALID, "nsteps must be >= 0");
    int rc = refresh_dt(b);
    if (rc) return rc;
    int n = 0;
    if (b->use_graph && nsteps >= kGraphSteps) {
        if (!b->graph[b->cur] && (rc = capture_graph(b))) return rc;
        for (; n + kGraphSteps <= nsteps; n += kGraphSteps) BHIP(b, hipGraphLaunch(b->graph[b->cur], b->stream));
    }
    for (; n < nsteps; ++n) {
        if ((rc = enqueue_step(b, b->cur))) return rc;
        b->cur ^= 1;
    }
    return MHD_OK;
}

int mhd_batch_synchronize(mhd_batch *b)
{
    if (!b) return bfail(nullptr, MHD_ERR_INVALID, "NULL batch");
    BHIP(b, hipStreamSynchronize(b->stream));
    return MHD_OK;
}

int mhd_batch_get_dt(mhd_batch *b, int run, double *dt)
{
    int rc = check_run(b, run);
    if (rc) return rc;
    if (!dt) return bfail(b, MHD_ERR_INVALID, "NULL argument");
    Ctrl c;
    if ((rc = read_ctrl(b, run, &c))) return rc;
    *dt = c.dt;
    return MHD_OK;
}

int mhd_batch_get_time(mhd_batch *b, int run, double *t)
{
    int rc = check_run(b, run);
    if (rc) return rc;
    if (!t) return bfail(b, MHD_ERR_INVALID, "NULL argument");
    Ctrl c;
    if ((rc = read_ctrl(b, run, &c))) return rc;
    *t = c.time;
    return MHD_OK;
}

int mhd_batch_check_crashed(mhd_batch *b, int run, int *crashed)
{
    int rc = check_run(b, run);
    if (rc) return rc;
    if (!crashed) return bfail(b, MHD_ERR_INVALID, "NULL argument");
    if ((rc = refresh_dt(b))) return rc;
    Ctrl c;
    if ((rc = read_ctrl(b, run, &c))) return rc;
    *crashed = c.crashed ? 1 : 0;
    return MHD_OK;
}

int mhd_batch_set_numerics(mhd_batch *b, int mode)
{
    if (!b) return bfail(nullptr, MHD_ERR_INVALID, "NULL batch");
    if (mode != MHD_NUMERICS_FAST && mode != MHD_NUMERICS_EXACT)
        return bfail(b, MHD_ERR_INVALID, "mode must be MHD_NUMERICS_FAST or MHD_NUMERICS_EXACT");
    BHIP(b, hipStreamSynchronize(b->stream));
    b->exact = mode == MHD_NUMERICS_EXACT;
    b->dt_valid = false;
    return build_work_list(b, b->seg_request > 0 ? b->seg_request : default_seg(b));
}

int mhd_batch_set_sweep_length(mhd_batch *b, int columns)
{
    if (!b) return bfail(nullptr, MHD_ERR_INVALID, "NULL batch");
    if (columns < 0 || columns > 1 << 20) return bfail(b, MHD_ERR_INVALID, "columns out of range");
    b->seg_request = columns;
    return build_work_list(b, columns > 0 ? columns : default_seg(b));
}

int mhd_batch_get_sweep_length(const mhd_batch *b, int *columns)
{
    if (!b || !columns) return bfail(nullptr, MHD_ERR_INVALID, "NULL argument");
    *columns = b->seg;
    return MHD_OK;
}

int mhd_batch_set_graph(mhd_batch *b, int on)
{
    if (!b) return bfail(nullptr, MHD_ERR_INVALID, "NULL batch");
    BHIP(b, hipStreamSynchronize(b->stream));
    b->use_graph = on != 0;
    if (!b->use_graph) drop_graph(b);
    return MHD_OK;
}

int mhd_batch_autotune(mhd_batch *b, int steps, int *columns)
{
    if (!b) return bfail(nullptr, MHD_ERR_INVALID, "NULL batch");
    if (steps < 1) steps = 2;
    int rc = refresh_dt(b);
    if (rc) return rc;
    // save every run's state and control block; the candidates' steps are
    // undone afterwards
    size_t total = 0;
    for (const RunDesc &d : b->desc) total += (size_t)8 * d.g.plane;
    double *save = nullptr;
    Ctrl *save_ctrl = nullptr;
    BHIP(b, hipMalloc(&save, sizeof(double) * total));
    hipError_t e = hipMalloc(&save_ctrl, sizeof(Ctrl) * (size_t)b->nruns);
    if (e != hipSuccess) {
        (void)hipFree(save);
        return bfail(b, MHD_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    auto snapshot = [&](bool restore) -> int {
        size_t o = 0;
        for (const RunDesc &d : b->desc) {
            const size_t n = sizeof(double) * 8 * (size_t)d.g.plane;
            BHIP(b, hipMemcpyAsync(restore ? d.u : save + o, restore ? save + o : d.u, n, hipMemcpyDeviceToDevice,
                                   b->stream));
            o += (size_t)8 * d.g.plane;
        }
        BHIP(b, hipMemcpyAsync(restore ? (void *)b->ctrl : (void *)save_ctrl,
                               restore ? (const void *)save_ctrl : (const void *)b->ctrl,
                               sizeof(Ctrl) * (size_t)b->nruns, hipMemcpyDeviceToDevice, b->stream));
        return MHD_OK;
    };
    const int saved_cur = b->cur;
    // times steps * kGraphSteps steps with sweep length L and w resident
    // wavefronts per CU (0: one per item), then restores the snapshot
    auto trial = [&](int L, int w, double &t) -> int {
        b->waves_per_cu = w;
        int r;
        if ((r = build_work_list(b, L))) return r;
        if ((r = mhd_batch_step(b, kGraphSteps))) return r;   // warm-up and graph capture
        hipError_t err = hipStreamSynchronize(b->stream);
        if (err != hipSuccess) return bfail(b, MHD_ERR_HIP, std::string("autotune: ") + hipGetErrorString(err));
        const auto t0 = std::chrono::steady_clock::now();
        if ((r = mhd_batch_step(b, steps * kGraphSteps))) return r;
        err = hipStreamSynchronize(b->stream);
        if (err != hipSuccess) return bfail(b, MHD_ERR_HIP, std::string("autotune: ") + hipGetErrorString(err));
        t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (b->cur != saved_cur) b->cur ^= 1;   // keep the slot parity of the saved control blocks
        return snapshot(true);
    };
    // sweep lengths at the current launch shape, then launch shapes at the
    // best sweep length: one wavefront per item, or a persistent launch
    // with half or all of the resident wavefront slots per CU
    static const int cand[] = {8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256};
    const int w0 = b->waves_per_cu;
    int best = b->seg, best_w = w0;
    double best_t = 1e300, t = 0.0;
    rc = snapshot(false);
    for (int L : cand) {
        if (rc) break;
        if (L > b->max_nx && L != cand[0]) break;
        if ((rc = trial(L, w0, t))) break;
        if (t < best_t) {
            best_t = t;
            best = L;
        }
    }
    const int occ = std::max(1, mhd::stage_batch_waves_per_cu(b->exact));
    const int wcand[] = {0, std::max(1, occ / 2), occ};
    for (int k = 0; k < 3; ++k) {
        const int w = wcand[k];
        if (rc) break;
        if (w == w0 || (k == 2 && w == wcand[1])) continue;   // timed already
        if ((rc = trial(best, w, t))) break;
        if (t < best_t) {
            best_t = t;
            best_w = w;
        }
    }
    if (!rc) rc = snapshot(true);
    (void)hipStreamSynchronize(b->stream);
    (void)hipFree(save);
    (void)hipFree(save_ctrl);
    b->cur = saved_cur;
    b->waves_per_cu = best_w;
    if (rc) return rc;
    b->seg_request = best;
    if ((rc = build_work_list(b, best))) return rc;
    if (columns) *columns = best;
    return MHD_OK;
}

int mhd_batch_set_waves_per_cu(mhd_batch *b, int waves)
{
    if (!b) return bfail(nullptr, MHD_ERR_INVALID, "NULL batch");
    if (waves < 0 || waves > 64) return bfail(b, MHD_ERR_INVALID, "waves_per_cu must be in 0 .. 64");
    BHIP(b, hipStreamSynchronize(b->stream));
    b->waves_per_cu = waves;
    set_launch_shape(b);
    drop_graph(b);
    return MHD_OK;
}

int mhd_batch_get_waves_per_cu(const mhd_batch *b, int *waves)
{
    if (!b || !waves) return bfail(nullptr, MHD_ERR_INVALID, "NULL argument");
    *waves = b->waves_per_cu;
    return MHD_OK;
}

int mhd_batch_set_strip_packing(mhd_batch *b, int mode)
{
    if (!b) return bfail(nullptr, MHD_ERR_INVALID, "NULL batch");
    if (mode != MHD_STRIP_PACK_AUTO && mode != MHD_STRIP_PACK_ALWAYS && mode != MHD_STRIP_PACK_NEVER)
        return bfail(b, MHD_ERR_INVALID, "mode must be MHD_STRIP_PACK_AUTO, _ALWAYS or _NEVER");
    BHIP(b, hipStreamSynchronize(b->stream));
    b->strip_packing = mode;
    return build_work_list(b, b->seg);
}

int mhd_batch_get_packed_items(const mhd_batch *b, int *items)
{
    if (!b || !items) return bfail(nullptr, MHD_ERR_INVALID, "NULL argument");
    *items = b->tail_items;
    return MHD_OK;
}

}  // extern "C"
