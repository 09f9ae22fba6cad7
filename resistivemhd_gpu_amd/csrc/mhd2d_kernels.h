// mhd2d_kernels.h -- launch interface of the gfx950 kernels (internal to the
// library; the public ABI is include/mhd2d.h).
#pragma once
#include <hip/hip_runtime.h>
#include "mhd2d_device.h"

namespace mhd {

constexpr int kGhost = 3;        // ghost layers per side
constexpr int kStripRows = 60;   // interior rows a wavefront updates (64 lanes - 4 halo rows)

struct Grid {
    int nx, ny;          // cells per axis including ghosts
    int pitch;           // doubles between consecutive x rows of a plane
    long long plane;     // doubles between planes
    int seg;             // columns one wavefront sweeps
    int nstrips, nsegs;  // work decomposition: strips along y, segments along x
    int xcds;            // L2 domains workgroups are dealt to round-robin
    int items_per_xcd;   // ceil(nstrips * nsegs / xcds)
};

// One run of a batch: its grid, constants, buffers and control block, and
// the index of its first item in the batch's work list.
struct RunDesc {
    Grid g;
    Phys ph;
    double *u, *us, *eta, *side;
    double *uo;   // stage 2 writes U^{n+1} here instead of over u (nullptr: in place)
    Ctrl *ctrl;
    int bc_x, bc_y;
    int first_item;
    int pad;
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
// stored), corner EMF from two flux sets, final CT with the next dt.
hipError_t launch_m_ec(bool smooth, const double *u, double *ec, const Grid &g, const Phys &ph, hipStream_t s);
hipError_t launch_m_emf(const double *aux_old, const double *aux_now, const double *ec_old, const double *ec_now,
                        double *E, const Grid &g, hipStream_t s);
hipError_t launch_m_ct(bool exact, const double *uold, const double *E, double *u, Ctrl *ctrl, int slot,
                       const Grid &g, const Phys &ph, hipStream_t s);
// Resident 64-lane stage workgroups per CU (the lower of the two stages).
int stage_waves_per_cu(bool exact);
// Batched stage over a work list of total_items items (item_run: run of
// each item), items_per_xcd per XCD taken by blocks_per_xcd wavefronts per
// XCD (equal: one wavefront per item; fewer: persistent), and the batched
// boundary pair of every run's U* (stage_state) or U, optionally reducing
// the ghost cells' dt into slot.
hipError_t launch_stage_batch(int stage, bool exact, const RunDesc *runs, const int *item_run, int total_items,
                              int xcds, int items_per_xcd, int blocks_per_xcd, int cur, hipStream_t s);
hipError_t launch_bc_batch(bool exact, const RunDesc *runs, int nruns, int max_nx, int max_ny, bool stage_state,
                           bool with_dt, int slot, hipStream_t s);
int stage_batch_waves_per_cu(bool exact);
// Per-cell dt (and the NaN scan) of every run's state into its control
// block's slot, one launch for the whole batch.
hipError_t launch_dt_all_batch(bool exact, const RunDesc *runs, int nruns, int max_nx, int slot, hipStream_t s);
hipError_t launch_unpack(const double *aos, double *u, const Grid &g, hipStream_t s);
hipError_t launch_pack(const double *u, double *aos, const Grid &g, hipStream_t s);
hipError_t launch_scatter_plane(const double *src, double *dst, const Grid &g, hipStream_t s);

}  // namespace mhd
