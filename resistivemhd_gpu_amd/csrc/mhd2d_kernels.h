// mhd2d_kernels.h -- launch interface of the gfx950 kernels (internal to the
// library; the public ABI is include/mhd2d.h).
#pragma once
#include <hip/hip_runtime.h>
#include "mhd2d_device.h"

namespace mhd {

constexpr int kGhost = 3;        // ghost layers per side
constexpr int kStripRows = 60;   // interior rows a wavefront updates (64 lanes - 4 halo rows)

// A last strip of at most this many interior rows is packed: one wavefront
// sweeps 64 / (rows + 4) segments of it side by side, each in its own lane
// group, instead of one segment in a mostly idle strip.
constexpr int kPackedStripRows = 28;

struct Grid {
    int nx, ny;          // cells per axis including ghosts
    int pitch;           // doubles between consecutive x rows of a plane
    long long plane;     // doubles between planes
    int seg;             // columns one wavefront sweeps
    int nstrips, nsegs;  // work decomposition: full strips along y (not the packed one), segments along x
    int tail_g;          // packed last strip: lanes per segment group (its rows + 4 halo rows), 0 if none
    int tail_items;      // packed last strip: wavefronts, 64 / tail_g segments each
    int xcds;            // L2 domains workgroups are dealt to round-robin
    int items_per_xcd;   // ceil(items / xcds)
};

// Strips of ny_int interior rows; pack: whether a short last strip may be
// packed (single-GPU scheme only).
inline void set_strips(Grid &g, int ny_int, bool pack)
{
    const int n = (ny_int + kStripRows - 1) / kStripRows;
    const int last = ny_int - (n - 1) * kStripRows;
    g.nstrips = n;
    g.tail_g = 0;
    if (pack && last <= kPackedStripRows) {
        g.nstrips = n - 1;
        g.tail_g = last + 4;
    }
}

// Segments of seg columns over nx_int interior columns.
inline void set_segs(Grid &g, int nx_int, int seg)
{
    g.seg = seg;
    g.nsegs = (nx_int + seg - 1) / seg;
    const int per = g.tail_g ? 64 / g.tail_g : 1;
    g.tail_items = g.tail_g ? (g.nsegs + per - 1) / per : 0;
}

// Work items (wavefronts) of one stage over the grid: full strips, then the
// packed last strip.
inline int grid_items(const Grid &g) { return g.nstrips * g.nsegs + g.tail_items; }

// One run of a batch: its grid, constants, buffers and control block, and
// the index of its first item in the batch's work list (full strips) and in
// its packed-strip list.
struct RunDesc {
    Grid g;
    Phys ph;
    double *u, *us, *eta, *side;
    double *uo;   // stage 2 writes U^{n+1} here instead of over u (nullptr: in place)
    Ctrl *ctrl;
    int bc_x, bc_y;
    int first_item;
    int first_tail;
};

// exact = true: operation-for-operation reference arithmetic (built without
// contraction it rounds as the reference does); false: fast numerics.
// scheme 0: the single-GPU library's stage; 1: the MPI library's stage on
// one subdomain, which also stores the face fluxes F0, F5, G0, G4 in aux.
hipError_t launch_stage(int stage, bool exact, int scheme, const double *uin, double *uout, const double *uold,
                        const double *eta, double *side, double *aux, Ctrl *ctrl, const Grid &g, const Phys &ph,
                        int cur, hipStream_t s);
hipError_t launch_bc_x(int kind, bool exact, double *u, Ctrl *ctrl, int slot, const Grid &g, const Phys &ph,
                       hipStream_t s);
hipError_t launch_bc_y(int kind, bool exact, double *u, Ctrl *ctrl, int slot, const Grid &g, const Phys &ph,
                       hipStream_t s);
// interior: dt over interior cells only (the MPI library's calculateDt)
hipError_t launch_dt_all(bool exact, bool interior, const double *u, Ctrl *ctrl, int slot, const Grid &g,
                         const Phys &ph, hipStream_t s);
// MPI library's CT-contact EMF: cell-centred EMF of a state (smoothed or as
// stored); corner EMF from two flux sets and the two cell EMFs, with the
// final CT and the next dt in the same pass.
hipError_t launch_m_ec(bool smooth, const double *u, const double *eta_c, double *ec, const Grid &g, const Phys &ph,
                      hipStream_t s);
hipError_t launch_m_emf_ct(bool exact, const double *aux_old, const double *aux_now, const double *ec_old,
                           const double *ec_now, const double *uold, double *u, Ctrl *ctrl, int slot, const Grid &g,
                           const Phys &ph, hipStream_t s);
// Resident 64-lane stage workgroups per CU (the lower of the two stages).
int stage_waves_per_cu(bool exact);
// Batched stage over a work list of total_items items (item_run: run of
// each item), items_per_xcd per XCD taken by blocks_per_xcd wavefronts per
// XCD (equal: one wavefront per item; fewer: persistent), and the batched
// boundary pair of every run's U* (stage_state) or U, optionally reducing
// the ghost cells' dt into slot.
hipError_t launch_stage_batch(int stage, bool exact, const RunDesc *runs, const int *item_run, int total_items,
                              int xcds, int items_per_xcd, int blocks_per_xcd, int cur, hipStream_t s,
                              const int *tail_run, int tail_items);
hipError_t launch_bc_batch(bool exact, const RunDesc *runs, int nruns, int max_nx, int max_ny, bool stage_state,
                           bool with_dt, int slot, hipStream_t s);
int stage_batch_waves_per_cu(bool exact);
// Per-cell dt (and the NaN scan) of every run's state into its control
// block's slot, one launch for the whole batch.
hipError_t launch_dt_all_batch(bool exact, const RunDesc *runs, int nruns, int max_nx, int slot, hipStream_t s);
hipError_t launch_unpack(const double *aos, double *u, const Grid &g, hipStream_t s);
hipError_t launch_pack(const double *u, double *aos, const Grid &g, hipStream_t s);
hipError_t launch_scatter_plane(const double *src, double *dst, const Grid &g, hipStream_t s);
hipError_t launch_face_regions(bool exact, const double *u, unsigned char *rx, unsigned char *ry, const Grid &g,
                               const Phys &ph, hipStream_t s);

}  // namespace mhd
