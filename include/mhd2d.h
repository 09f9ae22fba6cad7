/*
 * mhd2d.h -- C ABI of the MI355X (gfx950) 2D resistive-MHD RK2 library.
 *
 * The state uses the reference's conventions: nx, ny count cells including
 * 3 ghost layers per side; cell (i, j) has linear index k = j + i*ny; a cell
 * holds 8 doubles (rho, rho*u, rho*v, rho*w, Bx, By, Bz, e) where Bx is the
 * x-face value at (i+1/2, j) and By the y-face value at (i, j+1/2).  Host
 * buffers passed through this ABI are in that array-of-structs layout
 * ("AoS": 8*nx*ny doubles).  Device storage is private to the library.
 *
 * Every function returns 0 on success and a negative MHD_ERR_* code on
 * failure; mhd_last_error() gives the message.  No exception crosses the ABI.
 * A handle owns one HIP stream and its device buffers; handles are not
 * internally synchronised (use one handle per host thread).
 *
 * Reference interface each entry point replaces (paths relative to the
 * reference root, P = ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_periodicXY,
 * S = ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_symmetricXY):
 * cited at each declaration.
 */
#ifndef MHD2D_H
#define MHD2D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHD2D_ABI_VERSION 1
#define MHD_GHOST 3

enum mhd_status {
    MHD_OK = 0,
    MHD_ERR_INVALID = -1,   /* bad argument or configuration */
    MHD_ERR_HIP = -2,       /* a HIP runtime call failed */
    MHD_ERR_NOMEM = -3,     /* device or host allocation failed */
    MHD_ERR_IO = -4,        /* file could not be read or written */
    MHD_ERR_STATE = -5      /* call not valid in the current state */
};

enum mhd_bc { MHD_BC_PERIODIC = 0, MHD_BC_SYMMETRIC = 1 };

typedef struct mhd_config {
    int nx, ny;          /* cells per axis including 3 ghost layers per side (P/const.hpp:10,15) */
    double dx, dy;       /* P/const.hpp:7,12 */
    double xmin, ymin;   /* lower corner of the interior (used by the initial-state setups) */
    double gamma;        /* gamma_mhd, P/const.hpp:18 */
    double cfl;          /* CFL, P/const.hpp:17 */
    double eps;          /* EPS, P/const.hpp:4 (1e-20 in every FP64 example) */
    double eta;          /* eta: uniform resistivity and the dt diffusion limit, P/const.hpp:25 */
    int bc_x, bc_y;      /* enum mhd_bc per axis: P/ (periodic x and y), S/ (symmetric x and y),
                            periodic x with symmetric y as ResistiveMHD_singleGPU_thrust/
                            lib_ResistiveMHD_2D_GPU_single/resistiveMHD_2D.cu:198-259 */
} mhd_config;

typedef struct mhd_handle mhd_handle;

/* Fills *cfg with the reference example defaults (gamma 5/3, CFL 0.7,
 * EPS 1e-20, eta 0, periodic x and y) for an interior of nx_int x ny_int
 * cells on [xmin, xmin+lx] x [ymin, ymin+ly].  Host only. */
int mhd_default_config(mhd_config *cfg, int nx_int, int ny_int, double xmin, double lx,
                       double ymin, double ly);

/* ResistiveMHD2D::ResistiveMHD2D() (P/resistiveMHD_2D.cu:11-22) together with
 * initializeDeviceConstants() (P/const.cu:4-24): allocates the device state
 * and copies the configuration.  The state starts all-zero, as the
 * reference's device_vector members do. */
int mhd_create(const mhd_config *cfg, mhd_handle **out);
int mhd_destroy(mhd_handle *h);
const char *mhd_last_error(const mhd_handle *h);   /* h may be NULL: last global error */
int mhd_get_config(const mhd_handle *h, mhd_config *cfg);

/* Replaces the user-written ResistiveMHD2D::initializeU body
 * (P/resistiveMHD_2D.hpp:28; e.g. EX/plasmoid/main_plasmoid.cu:85-156
 * writes U then applies the boundary): uploads an AoS host state of all
 * nx*ny cells as is.  The boundary is not applied; call mhd_apply_boundary. */
int mhd_set_state(mhd_handle *h, const double *host_aos);
/* ResistiveMHD2D::getU (P/resistiveMHD_2D.cu:431-434) as a host copy. */
int mhd_get_state(mhd_handle *h, double *host_aos);

/* Per-cell resistivity for the resistive flux (nx*ny doubles, cell (i, j)
 * at j + i*ny).  Replaces the example plugin getEta(x, y) evaluated at
 * (i*dx, j*dy) (EX/plasmoid/main_plasmoid.cu:159-172).  NULL clears it, after
 * which the uniform cfg.eta is used (EX/bhattacharjee/main_bhattacharjee.cu:131-139).
 * The dt diffusion limit always uses cfg.eta. */
int mhd_set_eta_field(mhd_handle *h, const double *host_eta);
int mhd_set_eta(mhd_handle *h, double eta);

/* Boundary::periodicBoundaryX2nd / Y2nd (P/boundary.cu:61-71,131-141) and
 * Boundary::symmetricBoundaryX2nd / Y2nd (S/boundary.cu:90-104,191-205)
 * applied to the current state.  axis 0 = x, 1 = y; kind = enum mhd_bc. */
int mhd_apply_boundary(mhd_handle *h, int axis, int kind);
/* Sets the boundary kinds oneStepRK2 applies after each stage (x first). */
int mhd_set_boundary(mhd_handle *h, int bc_x, int bc_y);

/* ResistiveMHD2D::calculateDt (P/resistiveMHD_2D.cu:277-291): CFL time step
 * of the current state over all nx*ny cells, with the eta diffusion limit. */
int mhd_calculate_dt(mhd_handle *h, double *dt);

/* ResistiveMHD2D::oneStepRK2 (P/resistiveMHD_2D.cu:133-195; S/ identical
 * but for the boundary): one RK2 step.  mhd_step_rk2 runs nsteps steps with
 * no host synchronisation between them. */
int mhd_one_step_rk2(mhd_handle *h);
int mhd_step_rk2(mhd_handle *h, int nsteps);

/* dt of the last step and the accumulated time (the examples' host loop
 * totalTime += dt, EX/plasmoid/main_plasmoid.cu:325). */
int mhd_get_dt(mhd_handle *h, double *dt);
int mhd_get_time(mhd_handle *h, double *t);
int mhd_set_time(mhd_handle *h, double t);

/* ResistiveMHD2D::checkCalculationIsCrashed (P/resistiveMHD_2D.cu:294-309):
 * *crashed = 1 if any cell's e is NaN after the last step. */
int mhd_check_crashed(mhd_handle *h, int *crashed);

/* ResistiveMHD2D::save (P/resistiveMHD_2D.cu:198-244): writes
 * {dir}/{name}_{step}.bin, raw little-endian FP64, all nx*ny cells in index
 * order k = j + i*ny, 8 doubles per cell, no header.  mhd_load reads such a
 * file back into the state (the reference has no reader). */
int mhd_save(mhd_handle *h, const char *dir, const char *name, int step);
int mhd_load(mhd_handle *h, const char *path);

/* Per-rank files of the MPI library's ResistiveMHD2D::save
 * (ResistiveMHD_multiGPU_cuda/ResistiveMHD2D_GPU/resistiveMHD2D.cu:457-486):
 * {dir}/{name}_{step}_{rank}.bin with the interior cells only (ghosts left
 * out), i-major, 8 doubles per cell.  mhd_load_subdomain reads such a file
 * into the interior and keeps the ghost cells; apply the boundary after. */
int mhd_save_subdomain(mhd_handle *h, const char *dir, const char *name, int step, int rank);
int mhd_load_subdomain(mhd_handle *h, const char *path);

int mhd_synchronize(mhd_handle *h);

/* Initial-state setups (host only, no GPU needed): fill an AoS buffer of
 * nx*ny cells, ghosts included, before the boundary is applied.  They replace
 * the examples' initializeU kernels (face B sampled at half positions as
 * EX/plasmoid/main_plasmoid.cu:97-126 does).
 *  - Orszag-Tang vortex: rho = gamma^2, p = gamma, v = (-sin y, sin x),
 *    B = (-sin y, sin 2x).
 *  - Harris sheet: Bx = tanh(y), rho = sech^2(y) + 0.2, p = rho/2, vy white
 *    noise of amplitude noise_amp from a counter hash of (seed, i, j).
 *  - Harris resistivity: eta0 + eta1 * exp(-(x^2 + y^2) / 4) per cell. */
int mhd_fill_orszag_tang(const mhd_config *cfg, double *host_aos);
int mhd_fill_harris(const mhd_config *cfg, unsigned long long seed, double noise_amp, double *host_aos);
int mhd_fill_harris_eta(const mhd_config *cfg, double eta0, double eta1, double *host_eta);

/* Decaying-turbulence state for the MPI library's subdomain case (no
 * reference example sets it up; it stands in for the user-written
 * initializeU of ResistiveMHD_multiGPU_cuda/ResistiveMHD2D_GPU/
 * resistiveMHD2D.hpp): rho = 1, p = 1, w = Bz = 0; (vx, vy) and the corner
 * vector potential Az are sums of cosines over the wave vectors
 * n = (nx, ny) != 0 with |n| <= kmax, k = 2*pi*n / L (L = interior width),
 * amplitudes 1/|n| (v) and 1/|n|^2 (Az), phases uniform from a counter hash
 * of (seed, mode, field).  B is the discrete curl of Az
 * (Bx = dAz/dy, By = -dAz/dx across each face), so the face divergence is
 * zero to round-off.  The amplitudes are scaled so that the interior rms of
 * |v| and of |B| (face values) are rms_v and rms_b, from the discrete
 * Parseval sums (exact for kmax < N/2). */
int mhd_fill_turbulence(const mhd_config *cfg, unsigned long long seed, int kmax, double rms_v, double rms_b,
                        double *host_aos);

/* Device-side helpers built on the setups: fill, upload, apply the
 * configured boundary (as initializeU does). */
int mhd_init_orszag_tang(mhd_handle *h);
int mhd_init_harris(mhd_handle *h, unsigned long long seed, double noise_amp, double eta0, double eta1);
int mhd_init_turbulence(mhd_handle *h, unsigned long long seed, int kmax, double rms_v, double rms_b);

/* Allocation-free step.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  nsteps RK2 steps of
 * mhd_step_rk2 (exact numerics, MHD_SCHEME_SINGLE, the boundaries of
 * cfg->bc_x / bc_y) of a state the caller owns on the current device,
 * started from its CFL dt, enqueued on `stream` (a hipStream_t; NULL: the
 * null stream) with no host synchronisation.  dev_state: 8 planes (rho,
 * rho u, rho v, rho w, Bx, By, Bz, e, face B as in the AoS layout), plane
 * v at dev_state + v * plane, cell (i, j) of a plane at i * pitch + j,
 * ghosts included; advanced in place.  dev_eta: the per-cell resistivity
 * in the layout of one plane, or NULL for the uniform cfg->eta.  pitch and
 * plane must be even with pitch >= ny and plane >= nx * pitch, and both
 * buffers 16-byte aligned.  dev_info (device, 3 doubles, or NULL): dt of
 * the last step, the time the steps advanced, 1.0 if an energy became NaN.
 * It allocates nothing: U*, the control block and the symmetric-x side row
 * live in a fixed scratch array in the library, 32 MiB on each GPU the
 * library runs on, held until the process ends.  The limit is plane <=
 * 2^19 doubles (about 724 x 724 cells with ghosts) and nx <= 8192; larger
 * grids return MHD_ERR_INVALID (use mhd_create / mhd_step_rk2).  Every call
 * on one GPU shares that GPU's scratch: calls that overlap on one GPU
 * (concurrent streams or host threads) race. */
int mhd_step_rk2_noalloc(const mhd_config *cfg, double *dev_state, long long pitch, long long plane,
                         const double *dev_eta, int nsteps, void *stream, double *dev_info);

/* Step scheme.
 *  - MHD_SCHEME_SINGLE (default): the single-GPU library's oneStepRK2
 *    (P/resistiveMHD_2D.cu:133-195, boundaries per cfg.bc_x / bc_y).
 *  - MHD_SCHEME_MULTI_SHARE: the MPI library's ResistiveMHD2D::oneStepRK2
 *    (ResistiveMHD_multiGPU_cuda/ResistiveMHD2D_GPU/resistiveMHD2D.cu:147-221)
 *    on one GPU's subdomain that is its own neighbour on all four sides, so
 *    sendrecv_U (ResistiveMHD_multiGPU_cuda/ResistiveMHD2D_GPU/mpi.cu:104-188)
 *    becomes an on-device periodic halo fill.  Differences from SINGLE: the
 *    face-centred resistive flux (flux_solver.cu:73-189 there), B advanced by
 *    the flux divergence from the face-smoothed field in both stages, the
 *    CT-contact corner EMF averaged over the two stages applied once to the
 *    step's initial face B (ct.cu:26-218 there), and dt over interior cells
 *    without the diffusion limit (resistiveMHD2D.cu:489-556 there); a step
 *    with dt < 0 counts as crashed.  Needs periodic x and y and no per-cell
 *    eta field (MHD_ERR_STATE otherwise); its resistivity is cfg.eta or the
 *    planes of mhd_set_eta_staggered.  Switching back to MHD_SCHEME_SINGLE
 *    drops those planes. */
enum mhd_scheme { MHD_SCHEME_SINGLE = 0, MHD_SCHEME_MULTI_SHARE = 1 };
int mhd_set_scheme(mhd_handle *h, int scheme);
int mhd_get_scheme(const mhd_handle *h, int *scheme);

/* MHD_SCHEME_MULTI_SHARE only: the MPI library's plug-in getEta(x, y)
 * (ResistiveMHD_multiGPU_cuda/ResistiveMHD2D_GPU/get_eta.hpp:6) sampled where
 * that library calls it, one nx*ny plane each (index j + i*ny, ghosts
 * included): x-faces at ((i + 1/2) dx, j dy) (flux_solver.cu:84,94 there),
 * y-faces at (i dx, (j + 1/2) dy) (flux_solver.cu:168,179) and cells at
 * (i dx, j dy) for the CT's cell EMF (ct.cu:41,56).  All three NULL: back to
 * the uniform cfg.eta.  The subdomain is its own periodic neighbour, so each
 * plane's ghost entries must equal their periodic images (as a periodic
 * getEta gives) and every entry must be >= 0: MHD_ERR_INVALID otherwise;
 * MHD_ERR_STATE under MHD_SCHEME_SINGLE. */
int mhd_set_eta_staggered(mhd_handle *h, const double *host_xface, const double *host_yface,
                          const double *host_cell);

/* Arithmetic of the step (no counterpart in the reference, whose single
 * code path is what MHD_NUMERICS_EXACT reproduces):
 *  - MHD_NUMERICS_EXACT (default): the reference's expressions operation for
 *    operation, with IEEE division and square root and no contraction;
 *    results are bitwise equal to an uncontracted IEEE double restatement of
 *    those expressions (the CPU oracle) on the same inputs.  A reference
 *    build whose compiler contracts a*b+c into FMAs (nvcc does by default)
 *    rounds differently, so it agrees with EXACT to rounding only.
 *  - MHD_NUMERICS_FAST: shared reciprocals, squared wave speeds, FMA
 *    contraction.  It differs from EXACT by rounding, but the reference's HLLD
 *    takes the physical flux of each star state, which jumps where a
 *    rotational wave speed crosses zero, so a rounding-level difference that
 *    flips a region seeds a larger local difference: FAST is not held to the
 *    1e-10 bar (tests check it at 1e-8 relative L2 on Orszag-Tang and small
 *    Harris runs).  On a 4096 x 2048 Harris sheet after 100 steps it is
 *    5e-3 relative in rho*u, exactly as far as the same restatement compiled
 *    with FMA contraction: the jumps, not FAST's rounding, set the size. */
enum mhd_numerics { MHD_NUMERICS_FAST = 0, MHD_NUMERICS_EXACT = 1 };
int mhd_set_numerics(mhd_handle *h, int mode);
int mhd_get_numerics(const mhd_handle *h, int *mode);

/* Diagnostic (no counterpart in the reference): the HLLD region of every
 * face of the current state under the arithmetic `mode`, from the stage
 * kernel's reconstruction.  host_rx[j + i*ny] is the face (i+1/2, j),
 * host_ry[j + i*ny] the face (i, j+1/2), each nx*ny bytes; bits 0-5 are the
 * indicators of P/hlld.cu:64-69 (outer-left, single-star left, double-star
 * left, outer-right, single-star right, double-star right), 0xFF marks faces
 * no update reads.  Counting faces whose bits differ between the two modes
 * measures where FAST's rounding changes the HLLD branch. */
int mhd_face_regions(mhd_handle *h, int mode, unsigned char *host_rx, unsigned char *host_ry);

/* Launch-shape control (segment length along x of one wave's sweep; 0 picks
 * the built-in default).  Exposed for the autotuner and the tests. */
int mhd_set_sweep_length(mhd_handle *h, int columns);
int mhd_get_sweep_length(const mhd_handle *h, int *columns);

/* Stage timing (no counterpart in the reference, which has no timers):
 * when enabled, each step records HIP events on the handle's stream around
 * the two stage kernels; mhd_profile_read synchronises and returns the summed
 * kernel time in ms of stage 1 and stage 2 launches since enabling, and the
 * number of steps timed. */
int mhd_profile_enable(mhd_handle *h, int on);
int mhd_profile_read(mhd_handle *h, double *ms_stage1, double *ms_stage2, int *steps);

/* ------------------------------------------------------------------ batch
 * Batched form (no counterpart in the reference, which runs one simulation
 * per process because dt is a global, P/const.hpp:20,44): nruns independent
 * runs, each with its own mhd_config (grid size, dx, dy, gamma, CFL, eta,
 * boundaries), state, optional eta field and dt, advanced together.  Each RK
 * stage is one launch over a work list of every run's (strip, segment)
 * items and each boundary pass one launch over all runs; steps replay a
 * captured hipGraph.  Per run the arithmetic is that of mhd_step_rk2 with
 * MHD_SCHEME_SINGLE: with the same numerics a run's state is bitwise the one
 * a single handle produces.  States use the same AoS layout as
 * mhd_set_state. */
typedef struct mhd_batch mhd_batch;
int mhd_batch_create(const mhd_config *cfgs, int nruns, mhd_batch **out);
int mhd_batch_destroy(mhd_batch *b);
const char *mhd_batch_last_error(const mhd_batch *b);   /* b may be NULL */
int mhd_batch_size(const mhd_batch *b, int *nruns);
int mhd_batch_set_state(mhd_batch *b, int run, const double *host_aos);
int mhd_batch_get_state(mhd_batch *b, int run, double *host_aos);
/* ResistiveMHD2D::save (P/resistiveMHD_2D.cu:198-244) of one run, in
 * mhd_save's layout ({dir}/{name}_{step}.bin), and the matching reader. */
int mhd_batch_save(mhd_batch *b, int run, const char *dir, const char *name, int step);
int mhd_batch_load(mhd_batch *b, int run, const char *path);
int mhd_batch_set_eta_field(mhd_batch *b, int run, const double *host_eta);   /* NULL: uniform cfg.eta */
/* the configured boundary of every run (x then y) */
int mhd_batch_apply_boundary(mhd_batch *b);
/* setups of mhd_fill_orszag_tang / mhd_fill_harris uploaded to one run
 * (the boundary is applied by mhd_batch_apply_boundary) */
int mhd_batch_init_orszag_tang(mhd_batch *b, int run);
int mhd_batch_init_harris(mhd_batch *b, int run, unsigned long long seed, double noise_amp);
/* nsteps RK2 steps of every run, no host synchronisation */
int mhd_batch_step(mhd_batch *b, int nsteps);
int mhd_batch_synchronize(mhd_batch *b);
int mhd_batch_get_dt(mhd_batch *b, int run, double *dt);
int mhd_batch_get_time(mhd_batch *b, int run, double *t);
int mhd_batch_check_crashed(mhd_batch *b, int run, int *crashed);
int mhd_batch_set_numerics(mhd_batch *b, int mode);
/* sweep length of the work items (0: chosen from the occupancy) */
int mhd_batch_set_sweep_length(mhd_batch *b, int columns);
int mhd_batch_get_sweep_length(const mhd_batch *b, int *columns);
/* replay steps from a captured hipGraph (default on) */
int mhd_batch_set_graph(mhd_batch *b, int on);
/* Launch shape of the batched stage kernels: 0 (the default) launches one
 * wavefront per work item; w > 0 launches w wavefronts per CU that loop
 * over the work list (a persistent launch). */
int mhd_batch_set_waves_per_cu(mhd_batch *b, int waves_per_cu);
int mhd_batch_get_waves_per_cu(const mhd_batch *b, int *waves_per_cu);
/* Packed last strips (no counterpart in the reference).  A wavefront
 * updates 60 interior rows of one column segment; a run's last strip of at
 * most 28 rows can instead be packed, with 64 / (rows + 4) segments side by
 * side in one wavefront, swept by a launch of its own after the full
 * strips'.  AUTO (the default) packs when the full strips alone would take
 * more than two rounds of resident wavefronts; results are bitwise the same
 * in every mode.  mhd_batch_get_packed_items: packed items in the current
 * work list (0: none). */
#define MHD_STRIP_PACK_AUTO 0
#define MHD_STRIP_PACK_ALWAYS 1
#define MHD_STRIP_PACK_NEVER 2
int mhd_batch_set_strip_packing(mhd_batch *b, int mode);
int mhd_batch_get_packed_items(const mhd_batch *b, int *items);
/* Autotuner: times 2*steps batch steps for each candidate sweep length on
 * the batch's current states, then for each launch shape (one wavefront
 * per item, or half or all of the resident wavefront slots per CU) at the
 * best length.  It keeps the fastest pair (the length written to *columns
 * if not NULL, the shape readable by mhd_batch_get_waves_per_cu) and
 * restores every state, time and dt, so the runs continue as if it had not
 * been called. */
int mhd_batch_autotune(mhd_batch *b, int steps, int *columns);

/* ---------------------------------------------------------------- adjoint
 * Discrete adjoint of the MHD_SCHEME_SINGLE step (no counterpart in the
 * reference, which has no differentiation), for a batch of runs as in
 * mhd_batch_*.  A rollout U_N = S_{N-1}(... S_0(BC(U0))) is run by
 * mhd_adjoint_forward, which records every step's dt and keeps a checkpoint
 * every `stride` steps (0: ceil(sqrt(nsteps)); MHD_ADJ_STRIDE_FIT: the largest
 * stride up to nsteps whose storage, 2 stride + ceil(nsteps / stride) states
 * of every run, fits in half of the free device memory plus what the object
 * already holds, so that with memory to spare the whole rollout stays
 * resident and the backward pass recomputes no segment; never below
 * ceil(sqrt(nsteps)); mhd_adjoint_get_stride reports the stride used).
 * mhd_adjoint_backward takes
 * the seed dL/dU_N of each run and returns dL/dU0 (of the state passed to
 * mhd_adjoint_set_state, before the boundary fill) and dL/deta (per cell,
 * ghosts included), with every step's dt held at its recorded value
 * ("frozen dt").  With frozen_dt = 1 the forward replays dt values set by
 * mhd_adjoint_set_dts instead of the CFL ones (for finite-difference checks
 * with the same dt sequence).  Exact numerics only; periodic or symmetric
 * boundaries in each direction (the symmetric-x fill's ghost Bx rebuild,
 * S/boundary.cu:56-88, is differentiated through the bottom boundary face
 * By that the CT leaves for it).  States use the AoS layout of
 * mhd_set_state. */
typedef struct mhd_adjoint mhd_adjoint;
int mhd_adjoint_create(const mhd_config *cfgs, int nruns, mhd_adjoint **out);
int mhd_adjoint_destroy(mhd_adjoint *a);
const char *mhd_adjoint_last_error(const mhd_adjoint *a);   /* a may be NULL */
int mhd_adjoint_set_eta_field(mhd_adjoint *a, int run, const double *host_eta);   /* NULL: uniform cfg.eta */
int mhd_adjoint_set_state(mhd_adjoint *a, int run, const double *host_aos);       /* U0 */
int mhd_adjoint_set_dts(mhd_adjoint *a, int run, const double *host_dts, int nsteps);
int mhd_adjoint_get_dts(mhd_adjoint *a, int run, double *host_dts, int nsteps);
#define MHD_ADJ_STRIDE_FIT (-1)
int mhd_adjoint_forward(mhd_adjoint *a, int nsteps, int stride, int frozen_dt);
int mhd_adjoint_get_stride(const mhd_adjoint *a, int *stride);   /* the last forward's checkpoint stride */
int mhd_adjoint_get_state(mhd_adjoint *a, int run, double *host_aos);   /* U_N (until the next forward or tangent) */
int mhd_adjoint_set_seed(mhd_adjoint *a, int run, const double *host_aos);   /* dL/dU_N */
int mhd_adjoint_backward(mhd_adjoint *a);
int mhd_adjoint_get_grad_state(mhd_adjoint *a, int run, double *host_aos);   /* dL/dU0 */
int mhd_adjoint_get_grad_eta(mhd_adjoint *a, int run, double *host_eta);     /* dL/deta */
/* Tangent-linear model of the same rollout (the FP64 reference the adjoint
 * is checked against): with the direction (dU0, deta) of each run (NULL: 0),
 * mhd_adjoint_tangent replays nsteps steps from U0 with the recorded dt and
 * carries dU along; mhd_adjoint_get_tangent returns dU_N.  For any seed W,
 * sum(W * dU_N) = sum(dL/dU0 * dU0) + sum(dL/deta * deta). */
int mhd_adjoint_set_direction(mhd_adjoint *a, int run, const double *host_dU0, const double *host_deta);
int mhd_adjoint_tangent(mhd_adjoint *a, int nsteps);
int mhd_adjoint_get_tangent(mhd_adjoint *a, int run, double *host_dUN);
/* Decision replay (no counterpart in the reference): the rollout of a
 * perturbed initial state P0 with the recorded dt and with every branch
 * decision of every step (minmod limiter, HLLD wave speeds and region, the
 * signs of Bx, min/max, and with mhd_adjoint_set_degenerate_tol > 0 the
 * degenerate-limit switch) taken from the rollout of U0.  P0 and the
 * perturbed resistivity are passed through mhd_adjoint_set_direction in
 * place of the direction: host_dU0 = P0 (AoS, before the boundary fill),
 * host_deta = the perturbed eta plane (for a run with an eta plane it is
 * required: NULL means eta = 0; a run on the uniform cfg.eta keeps it).
 * mhd_adjoint_get_tangent then returns P_N.  The map P0 -> P_N is smooth,
 * equals the rollout at P0 = U0, and its derivative there is the tangent's,
 * so central differences of a loss of P_N check the adjoint without the
 * jumps that branch switches put into differences of the rollout itself. */
int mhd_adjoint_replay(mhd_adjoint *a, int nsteps);
/* Whole-batch transfers (one copy and one synchronisation): every run's AoS
 * state (or eta plane, nx*ny) concatenated in run order.  The _device forms
 * take a device pointer on the handle's GPU (e.g. a torch tensor's storage)
 * and copy nothing across PCIe; the caller orders its own stream first. */
int mhd_adjoint_set_states(mhd_adjoint *a, const double *host_all);        /* all U0 */
int mhd_adjoint_get_states(mhd_adjoint *a, double *host_all);              /* all U_N */
int mhd_adjoint_set_seeds(mhd_adjoint *a, const double *host_all);         /* all dL/dU_N */
int mhd_adjoint_get_grad_states(mhd_adjoint *a, double *host_all);         /* all dL/dU0 */
int mhd_adjoint_get_grad_etas(mhd_adjoint *a, double *host_all);           /* all dL/deta */
int mhd_adjoint_set_states_device(mhd_adjoint *a, const double *dev_all);
int mhd_adjoint_set_eta_fields_device(mhd_adjoint *a, const double *dev_all);   /* every run's eta plane */
int mhd_adjoint_get_states_device(mhd_adjoint *a, double *dev_all);
int mhd_adjoint_set_seeds_device(mhd_adjoint *a, const double *dev_all);
int mhd_adjoint_get_grad_states_device(mhd_adjoint *a, double *dev_all);
int mhd_adjoint_get_grad_etas_device(mhd_adjoint *a, double *dev_all);
int mhd_adjoint_synchronize(mhd_adjoint *a);
/* Kernel form of the face adjoints (no counterpart in the reference): one
 * face per lane with atomics, or a sweep along x with LDS rings that
 * accumulates cell adjoints before they leave for HBM.  AUTO picks the sweep
 * for batches large enough to fill the GPU with it.  Results agree to the
 * rounding of the atomic sums. */
enum mhd_adj_face_kernel { MHD_ADJ_FACE_AUTO = 0, MHD_ADJ_FACE_PER_FACE = 1, MHD_ADJ_FACE_SWEEP = 2 };
int mhd_adjoint_set_face_kernel(mhd_adjoint *a, int mode);
/* Derivative rule at degenerate HLLD star states (no counterpart in the
 * reference, which has no derivatives).  The star-state quotients of
 * P/hlld.cu:332-338 and 345-351 divide by rho (S - u)(S - SM) - Bx^2 + EPS,
 * which is zero up to rounding where the transverse field vanishes and the
 * fast speed equals the Alfven speed; their derivatives there are of the
 * order of 1 / rounding noise.  With tol > 0 the tangent and the adjoint
 * take, on the faces where |rho (S - u)(S - SM) - Bx^2| < tol rho (S - u)^2,
 * the derivative of the degenerate limit (tangential star components equal
 * to the outer ones, Miyoshi & Kusano 2005).  The forward step is unchanged.
 * tol = 0 (the default) differentiates the step exactly.  tol in [0, 1). */
int mhd_adjoint_set_degenerate_tol(mhd_adjoint *a, double tol);
/* Sweep length of the forward (and recompute) stage kernels picked by the
 * batch autotuner (mhd_batch_autotune) on the initial states; a later
 * mhd_adjoint_forward starts from U0 as always. */
int mhd_adjoint_autotune(mhd_adjoint *a, int steps, int *columns);

#ifdef __cplusplus
}
#endif
#endif /* MHD2D_H */
