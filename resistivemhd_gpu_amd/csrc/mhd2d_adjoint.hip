// mhd2d_adjoint.hip -- discrete adjoint of the single-GPU library's RK2 step
// (ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_periodicXY/
// resistiveMHD_2D.cu:133-195) for a batch of runs, with dt frozen at the
// values the forward pass used, and sqrt(N) checkpointing (mhd_adjoint_* in
// include/mhd2d.h).
//
// Forward: the batched sweep kernels advance every run, recording each
// step's dt and keeping the state every `stride` steps.  Backward, segment by
// segment from the last: the segment's states are recomputed from its
// checkpoint with the recorded dt (bitwise the forward's), then each step is
// reversed.  The adjoint of one step, with lam the adjoint of U^{n+1}:
//
//   U^{n+1} = BC(U2), U2_nonB = (U + U* - dt dF(U*) - dt dG(U*)) / 2,
//   U2_B = B* ; U* = BC(U1), U1_nonB = U - dt dF(U) - dt dG(U),
//   U1_B = B* = B - dt curl E(F(U), G(U))
//
//   1. lam <- BC^T lam
//   2. stage-2 faces: lamF from lam, then the face adjoint on the centred U*
//      (flux_full_adj: MUSCL, HLLD, eta J) scattered to lamC and g_eta
//   3. lamBar = shift^T lamC + lam_nonB / 2
//   4. lamBar <- BC^T lamBar
//   5. lamB* = lamBar_B + lam_B ; lamE = -dt curl^T lamB*
//   6. stage-1 faces: lamF from lamBar and lamE, face adjoint on the centred U
//   7. lam = lamU + shift^T lamC  (the adjoint of U^n), lamU_nonB =
//      lam_nonB / 2 + lamBar_nonB, lamU_B = lamB*
//
// The face pieces' branch decisions (minmod, HLLD region, min/max) come
// from a primal recomputed with IEEE arithmetic on every branch input, so
// the adjoint differentiates the branch the forward step took.
// Contributions of faces to cells are summed with FP64 atomics (the
// gradient's last bits depend on their order).  Boundaries: periodic or
// symmetric in each direction.  The symmetric-x fill rebuilds the ghost Bx
// from By of the bottom boundary face row 2 as the stage's CT left it, before
// the y fill overwrites that row: the forward keeps the CT value there (and
// stage 2 copies it into U^{n+1}), and the CT adjoint, which covers every
// face with i, j >= 1, takes up the adjoint the x-fill transpose leaves in
// row 2.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/mhd2d.h"
#include "mhd2d_adjoint_face.h"
#include "mhd2d_batch_impl.h"
#include "mhd2d_kernels.h"

using mhd::Ctrl;
using mhd::Grid;
using mhd::Phys;

namespace {

// One run of the adjoint batch.
struct AdjDesc {
    Grid g;
    Phys ph;
    int bc_x, bc_y;
    int ntj;                      // 64-row tiles per column
    double *u, *us;               // U^n and U* (the forward batch's buffers)
    const double *eta;            // per-cell resistivity or nullptr (uniform ph.eta)
    double *lam, *lamBar, *lamC, *lamE, *geta;
    double *u0, *hist, *ckpt;     // initial state, segment states U^n, checkpoints
    double *hists;                // segment states U* (stage-1 results with their boundary)
    double *dts;                  // dt of every step of the rollout
    double dtol;                  // degenerate star-state threshold of the face derivatives (0: exact)
    long long aos;                // offset of this run's cells in the concatenated AoS transfer layout
    int swj;                      // sweep form: 64-row strips per column, face rows 2 + 64 s ..
    int tail0;                    // sweep form: first face row left to the per-face tail, 2 + 64 swj
    // tangent-linear model: direction (dU0, deta), tangent state dU, dU*,
    // tangent face fluxes dF, dG and corner EMF dE
    double *du0, *deta, *du, *dus, *dF, *dG, *dE;
};

// ---------------------------------------------------------------- kernels
// Work list: run r owns tiles [first[r], first[r+1]); a tile is 64
// consecutive rows of one column, one wavefront (blocks hold blockDim.x / 64
// tiles).  run[t] is the run of tile t (one load instead of a search over
// the runs).
struct TileMap {
    const long long *first;
    const int *run;
    int nruns;
};

__device__ __forceinline__ bool tile_cell(TileMap tm, int &r, int &i, int &j, const AdjDesc *runs)
{
    const long long t = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (t >= tm.first[tm.nruns]) return false;
    r = tm.run[t];
    const int lt = (int)(t - tm.first[r]);
    const int ntj = runs[r].ntj;
    i = lt / ntj;
    j = (lt - i * ntj) * 64 + (int)(threadIdx.x & 63);
    return true;
}

struct Soa {
    const double *u;
    int pitch;
    long long plane;
    __device__ double operator()(int v, int i, int j) const { return u[v * plane + (size_t)i * pitch + j]; }
};

struct EtaAcc {
    const double *e;
    int pitch;
    double uniform;
    __device__ double operator()(int i, int j) const { return e ? e[(size_t)i * pitch + j] : uniform; }
};

struct AtomicAdd {
    double *lamC, *geta;
    int pitch;
    long long plane;
    __device__ void operator()(int i, int j, int v, double b) const
    {
        if (v == 8) unsafeAtomicAdd(geta + (size_t)i * pitch + j, b);
        else unsafeAtomicAdd(lamC + v * plane + (size_t)i * pitch + j, b);
    }
};

enum Buf { kU0 = 0, kU = 1, kHist = 2, kCkpt = 3, kLam = 4, kLamBar = 5, kLamC = 7, kUs = 8, kHistS = 9 };

__device__ __forceinline__ double *buf_ptr(const AdjDesc &d, int which, int slot)
{
    const size_t s8 = (size_t)8 * d.g.plane;
    switch (which) {
    case kU0: return d.u0;
    case kU: return d.u;
    case kHist: return d.hist + slot * s8;
    case kCkpt: return d.ckpt + slot * s8;
    case kLam: return d.lam;
    case kLamBar: return d.lamBar;
    case kUs: return d.us;
    case kHistS: return d.hists + slot * s8;
    default: return d.lamC;
    }
}

__global__ __launch_bounds__(256) void copy_kernel(const AdjDesc *runs, TileMap tm, int src,
                                                   int src_slot, int dst, int dst_slot)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    if (j >= d.g.ny) return;
    const double *a = buf_ptr(d, src, src_slot);
    double *b = buf_ptr(d, dst, dst_slot);
    const size_t o = (size_t)i * d.g.pitch + j;
#pragma unroll
    for (int v = 0; v < 8; ++v) b[v * d.g.plane + o] = a[v * d.g.plane + o];
}

// Every run's state or adjoint plane set to or from the concatenated AoS
// transfer layout (8 doubles per cell, k = j + i*ny, runs back to back at
// AdjDesc::aos), one launch for the whole batch.
template <bool TO_AOS>
__global__ __launch_bounds__(256) void aos_all_kernel(const AdjDesc *runs, TileMap tm, int which,
                                                      double *aos)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    if (j >= d.g.ny) return;
    double *p = buf_ptr(d, which, 0) + (size_t)i * d.g.pitch + j;
    double *c = aos + d.aos + ((size_t)i * d.g.ny + j) * 8;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        if (TO_AOS) c[v] = p[v * d.g.plane];
        else p[v * d.g.plane] = c[v];
    }
}

// Every run's dL/deta plane to the concatenated layout (nx*ny per run, runs
// back to back at AdjDesc::aos / 8), one launch for the whole batch.
__global__ __launch_bounds__(256) void geta_all_kernel(const AdjDesc *runs, TileMap tm,
                                                       double *dst)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    if (j >= d.g.ny) return;
    dst[d.aos / 8 + (size_t)i * d.g.ny + j] = d.geta[(size_t)i * d.g.pitch + j];
}

// mask bits: 1 lamC, 2 geta
__global__ __launch_bounds__(256) void zero_kernel(const AdjDesc *runs, TileMap tm, int mask)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    if (j >= d.g.ny) return;
    const size_t o = (size_t)i * d.g.pitch + j;
    if (mask & 1) {
#pragma unroll
        for (int v = 0; v < 8; ++v) d.lamC[v * d.g.plane + o] = 0.0;
    }
    if (mask & 2) d.geta[o] = 0.0;
}

// Transpose of the y boundary fill (periodic: P/boundary.cu:74-141;
// symmetric: S/boundary.cu:107-205); blockIdx.y = run.  One ghost cell per
// lane, as the forward fill: lanes 6c .. 6c+5 of a wavefront take column
// col0 + c and ghost rows 0, 1, 2, ny-3, ny-2, ny-1 (ten columns per
// wavefront), each adding its ghost adjoint into the cell it was copied
// from and clearing it.  Source and ghost rows are disjoint (ny >= 12), so
// lanes never touch each other's cells except through the atomics on Bx.
// Symmetric y: the By rebuild chain of a column's bottom (top) ghosts is
// transposed by its q = 0 (q = 3) lane.
constexpr int kBctCols = 10;   // columns per wavefront

__global__ __launch_bounds__(256) void bct_y_kernel(const AdjDesc *runs, int which)
{
    const AdjDesc &d = runs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int i = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kBctCols + lane / 6;
    const int q = lane % 6;
    if (lane >= 6 * kBctCols || i >= d.g.nx) return;
    double *L = buf_ptr(d, which, 0);
    const long long pl = d.g.plane;
    const int ny = d.g.ny;
    const size_t row = (size_t)i * d.g.pitch;
    const int dst = q < 3 ? q : ny - 6 + q;   // 0, 1, 2, ny-3, ny-2, ny-1
    if (d.bc_y == MHD_BC_PERIODIC) {
        const int src = q < 3 ? ny - 6 + q : q;   // ny-6 .. ny-4, 3 .. 5
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            L[v * pl + row + src] += L[v * pl + row + dst];
            L[v * pl + row + dst] = 0.0;
        }
        return;
    }
    const int src = q < 3 ? 5 - q : ny - 1 - q;   // 5, 4, 3, ny-4, ny-5, ny-6
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        if (v == 5) continue;
        const double b = L[v * pl + row + dst];
        L[v * pl + row + dst] = 0.0;
        if (v == 4) unsafeAtomicAdd(&L[4 * pl + row + src], b);   // the By chains add to Bx too
        else L[v * pl + row + src] += b;
    }
    if (q != 0 && q != 3) return;
    double *by = L + 5 * pl + row;
    double *bx = L + 4 * pl;
    const size_t prev = (size_t)(i - 1) * d.g.pitch;
    const double c = d.ph.dy / d.ph.dx;
    if (q == 0) {
        // bottom: by2 = by3 + c d3, by1 = by2 + c d3, by0 = by1 + c d4
        double lb0 = by[0], lb1 = by[1], lb2 = by[2];
        by[0] = by[1] = by[2] = 0.0;
        lb1 += lb0;
        const double ld4 = c * lb0;
        lb2 += lb1;
        const double ld3 = c * lb1 + c * lb2;
        by[3] += lb2;
        if (i > 0) {
            // d_r = Bx(i, r) - Bx(i-1, r)
            unsafeAtomicAdd(&bx[row + 3], ld3);
            unsafeAtomicAdd(&bx[prev + 3], -ld3);
            unsafeAtomicAdd(&bx[row + 4], ld4);
            unsafeAtomicAdd(&bx[prev + 4], -ld4);
        }
    } else {
        // top: by(ny-3) = by(ny-4) - c dA, by(ny-2) = .. - c dB, by(ny-1) = .. - c dC
        double lt0 = by[ny - 3], lt1 = by[ny - 2], lt2 = by[ny - 1];
        by[ny - 3] = by[ny - 2] = by[ny - 1] = 0.0;
        lt1 += lt2;
        const double ldC = -c * lt2;
        lt0 += lt1;
        const double ldB = -c * lt1;
        by[ny - 4] += lt0;
        const double ldA = -c * lt0;
        if (i > 0) {
            const int rows[3] = {ny - 4, ny - 5, ny - 6};
            const double lds[3] = {ldA, ldB, ldC};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                unsafeAtomicAdd(&bx[row + rows[k]], lds[k]);
                unsafeAtomicAdd(&bx[prev + rows[k]], -lds[k]);
            }
        }
    }
}

// Transpose of the x fill (periodic: P/boundary.cu:4-71; symmetric:
// S/boundary.cu:4-104); blockIdx.y = run.  One ghost cell per lane, as the
// forward fill: wavefront w takes ghost column w % 6 (0, 1, 2, nx-3, nx-2,
// nx-1) and rows (w / 6) * 64 .. + 63.
// Symmetric x: the Bx rebuild of row j's left (right) ghost faces is
// transposed by its q = 0 (q = 3) lane.  The rebuild reads By of rows j and
// j-1 in columns 3, 4 (nx-4, nx-5, nx-6); at j = 3 that is the bottom
// boundary face By(i, 2) as the stage's CT left it (the y fill overwrites it
// afterwards), so its adjoint stays in row 2 for the CT adjoint to take up.
// By adds are atomic: neighbouring rows' chains and the mirror copies share
// those cells.
__global__ __launch_bounds__(256) void bct_x_kernel(const AdjDesc *runs, int which)
{
    const AdjDesc &d = runs[blockIdx.y];
    const int w = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int q = w % 6, j = (w / 6) * 64 + (int)(threadIdx.x & 63);
    if (j >= d.g.ny) return;
    double *L = buf_ptr(d, which, 0);
    const long long pl = d.g.plane;
    const int nx = d.g.nx;
    const size_t P = d.g.pitch;
    const size_t dst = (size_t)(q < 3 ? q : nx - 6 + q) * P + j;   // columns 0, 1, 2, nx-3, nx-2, nx-1
    if (d.bc_x == MHD_BC_PERIODIC) {
        const size_t src = (size_t)(q < 3 ? nx - 6 + q : q) * P + j;   // columns nx-6 .. nx-4, 3 .. 5
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            L[v * pl + src] += L[v * pl + dst];
            L[v * pl + dst] = 0.0;
        }
        return;
    }
    const size_t src = (size_t)(q < 3 ? 5 - q : nx - 1 - q) * P + j;   // columns 5, 4, 3, nx-4, nx-5, nx-6
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        if (v == 4) continue;
        const double b = L[v * pl + dst];
        L[v * pl + dst] = 0.0;
        if (v == 5) unsafeAtomicAdd(&L[5 * pl + src], b);
        else L[v * pl + src] += b;
    }
    if (q != 0 && q != 3) return;
    double *bx = L + 4 * pl;
    double *by = L + 5 * pl;
    const double k = d.ph.dx / d.ph.dy;
    if (q == 0) {
        // columns 2, 1, 0: b2 = bx3 + k d3, b1 = b2 + k d3, b0 = b1 + k d4,
        // d_c = By(c, j) - By(c, j-1) (row 0: b0 = b1 = b2 = bx3)
        double l0 = bx[j], l1 = bx[P + j], l2 = bx[2 * P + j];
        bx[j] = bx[P + j] = bx[2 * P + j] = 0.0;
        l1 += l0;
        l2 += l1;
        bx[3 * P + j] += l2;
        if (j > 0) {
            const double ld4 = k * l0, ld3 = k * l1 + k * l2;
            unsafeAtomicAdd(&by[3 * P + j], ld3);
            unsafeAtomicAdd(&by[3 * P + j - 1], -ld3);
            unsafeAtomicAdd(&by[4 * P + j], ld4);
            unsafeAtomicAdd(&by[4 * P + j - 1], -ld4);
        }
    } else {
        // columns nx-3, nx-2, nx-1: bA = bx(nx-4) - k dA, bB = bA - k dB,
        // bC = bB - k dC from columns nx-4, nx-5, nx-6
        const size_t cA = (size_t)(nx - 3) * P + j;
        double lA = bx[cA], lB = bx[cA + P], lC = bx[cA + 2 * P];
        bx[cA] = bx[cA + P] = bx[cA + 2 * P] = 0.0;
        lB += lC;
        lA += lB;
        bx[cA - P] += lA;
        if (j > 0) {
            const double lds[3] = {-k * lA, -k * lB, -k * lC};   // columns nx-4, nx-5, nx-6
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const size_t c = (size_t)(nx - 4 - m) * P + j;
                unsafeAtomicAdd(&by[c], lds[m]);
                unsafeAtomicAdd(&by[c - 1], -lds[m]);
            }
        }
    }
}

__device__ __forceinline__ bool non_b(int v) { return v != 4 && v != 5; }

// Face adjoints of one stage and direction: face (i, j) sits between cell
// (i, j) and (i+1, j) (DIR 0) or (i, j+1) (DIR 1).
// One wavefront per block: at ~500 registers a SIMD holds one such wave, and
// single-wave blocks let it take the next tile as soon as its own ends.
// slot >= 0: the primal state of this stage is history slot `slot` (U^n for
// stage 1, U* for stage 2), except stage 1 at slot 0 with ck >= 0, whose
// U^n is checkpoint ck (a segment starts from its checkpoint in place);
// slot < 0: the live buffers u / us.  DEGEN:
// the degenerate star-state rule is on (d.dtol > 0); without it the rule's
// test folds away at compile time.
// Adjoint of face (i, j) of direction DIR: its flux adjoint from the stage's
// cell adjoints, then the flux's input adjoints added to lamC / geta with
// global atomics.
template <int STAGE>
__device__ __forceinline__ const double *stage_prim(const AdjDesc &d, int slot, int ck)
{
    const size_t s8 = (size_t)8 * d.g.plane;
    if (slot < 0) return STAGE == 2 ? d.us : d.u;
    if (STAGE == 1 && slot == 0 && ck >= 0) return d.ckpt + (size_t)ck * s8;
    return (STAGE == 2 ? d.hists : d.hist) + (size_t)slot * s8;
}

template <int STAGE, int DIR, bool DEGEN>
__device__ __forceinline__ void face_adj_at(const AdjDesc &d, int i, int j, int step, int slot, int ck)
{
    const Grid &g = d.g;
    const double dt = d.dts[step];
    const double coef = (STAGE == 2 ? 0.5 : 1.0) * dt / (DIR == 0 ? d.ph.dx : d.ph.dy);
    const double *L = STAGE == 2 ? d.lam : d.lamBar;
    const size_t o = (size_t)i * g.pitch + j;
    const size_t on = DIR == 0 ? o + g.pitch : o + 1;
    double lamF[8];
    bool any = false;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        lamF[v] = non_b(v) ? coef * (L[v * g.plane + on] - L[v * g.plane + o]) : 0.0;
        any |= lamF[v] != 0.0;
    }
    if (STAGE == 1) {
        // E(i, j) = (G4(i, j) + G4(i+1, j) - F5(i, j) - F5(i, j+1)) / 4
        if (DIR == 0) lamF[5] = -0.25 * (d.lamE[o] + d.lamE[o - 1]);
        else lamF[4] = 0.25 * (d.lamE[o] + d.lamE[o - g.pitch]);
        any |= lamF[4] != 0.0 || lamF[5] != 0.0;
    }
    if (!any) return;
    const Soa S{stage_prim<STAGE>(d, slot, ck), g.pitch, g.plane};
    const EtaAcc E{d.eta, g.pitch, d.ph.eta};
    AtomicAdd add{d.lamC, d.geta, g.pitch, g.plane};
    const double par[4] = {d.ph.gamma - 1.0, d.ph.gamma, d.ph.eps, DEGEN ? d.dtol : 0.0};
    const double rpar[2] = {d.ph.dx, d.ph.dy};
    mhd_ad::flux_full_adj(S, E, DIR, i, j, par, rpar, lamF, add);
}

template <int STAGE, int DIR, bool DEGEN>
__global__ __launch_bounds__(64) void face_adj_kernel(const AdjDesc *runs, TileMap tm,
                                                       int step, int slot, int ck)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    const Grid &g = d.g;
    if (DIR == 0 ? (i < 2 || i > g.nx - 4 || j < 2 || j > g.ny - 3) : (i < 2 || i > g.nx - 3 || j < 2 || j > g.ny - 4))
        return;
    face_adj_at<STAGE, DIR, DEGEN>(d, i, j, step, slot, ck);
}

// Face rows a run's sweep strips leave over (rows tail0 .. ny-3 of the x
// faces, tail0 .. ny-4 of the y faces, every face column), one face per
// lane, both directions in one launch: blocks below xblocks take x faces,
// the rest y faces.  Per direction, run r owns lanes [tfirst[r],
// tfirst[r+1]), column-major over the tail rows.
__device__ __forceinline__ int tail_rows(const AdjDesc &d, int dir)
{
    return max(0, (dir == 0 ? d.g.ny - 2 : d.g.ny - 3) - d.tail0);
}

template <int STAGE, int DIR, bool DEGEN>
__device__ __forceinline__ void face_adj_tail(const AdjDesc *runs, const long long *tfirst, int nruns, long long t,
                                              int step, int slot, int ck)
{
    if (t >= tfirst[nruns]) return;
    int lo = 0, hi = nruns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tfirst[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const AdjDesc &d = runs[lo];
    const int rows = tail_rows(d, DIR);
    const long long k = t - tfirst[lo];
    face_adj_at<STAGE, DIR, DEGEN>(d, 2 + (int)(k / rows), d.tail0 + (int)(k % rows), step, slot, ck);
}

template <int STAGE, bool DEGEN>
__global__ __launch_bounds__(64) void face_adj_tail_kernel(const AdjDesc *runs, const long long *tfirst, int nruns,
                                                            unsigned xblocks, int step, int slot, int ck)
{
    if (blockIdx.x < xblocks)
        face_adj_tail<STAGE, 0, DEGEN>(runs, tfirst, nruns, (long long)blockIdx.x * 64 + threadIdx.x, step, slot, ck);
    else
        face_adj_tail<STAGE, 1, DEGEN>(runs, tfirst + nruns + 1, nruns,
                                       (long long)(blockIdx.x - xblocks) * 64 + threadIdx.x, step, slot, ck);
}

// ------------------------------------------------ face adjoints, sweep form
// A workgroup of two wavefronts takes one item: the 64 face rows j0 .. j0+63
// (j0 = 2 + 64 s) of one run and the face columns ia .. ib-1; the few rows
// past a run's last strip go to face_adj_tail_kernel.  Wave 0 computes the x faces and
// wave 1 the y faces of each column, in step.  Everything a column's faces
// read sits in LDS rings (rows j0-2 .. j0+65; state columns i-2 .. i+2,
// adjoint and resistivity columns as the stencils need).  The columns the
// next step adds arrive by direct-to-LDS loads while this step computes,
// straight into ring slots that hold one column more than a step reads, so
// the slot each load fills is one no wave reads at this step; the loop-top
// wait and barrier make them visible to the next step.  Cell and
// resistivity adjoints accumulate in an
// LDS ring of five columns and leave for HBM once per column when
// complete, instead of with 38 atomics per face.
constexpr int kSwRows = 68;   // LDS rows: global rows j0-2 .. j0+65
constexpr int kSwSt = 6;      // state columns i-2 .. i+2, and i+3 landing
constexpr int kSwRing = 3;    // adjoint / resistivity columns i, i+1 and i+2 landing; corner EMF i-1, i, i+1
constexpr int kSwAcc = 5;     // accumulator columns i-2 .. i+2

struct __attribute__((aligned(16))) SweepLds {   // every row 16-byte aligned
    double st[8][kSwSt][kSwRows];
    double lm[6][kSwRing][kSwRows];   // adjoint components 0, 1, 2, 3, 6, 7
    double le[kSwRing][kSwRows];      // corner-EMF adjoint (stage 1)
    double et[kSwRing][kSwRows];      // resistivity plane
    double acc[9][kSwAcc][kSwRows];   // cell adjoints 0..7, resistivity adjoint 8
};

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;

// s_waitcnt vmcnt(0) as the compiler's own instruction (so it knows the
// direct-to-LDS loads have landed): vmcnt 0, expcnt and lgkmcnt at maximum.
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Loads rows j0-2 .. j0+65 of one plane column (col_base = &plane[i * pitch],
// 64-byte aligned) into dst[kSwRows]: lane l < 34 moves rows 2l, 2l+1 with
// one 16-byte direct-to-LDS load (LDS address dst + 16 l).  Row pairs start
// at even rows, so every load is 16-byte aligned; a pair past the grid is
// clamped to the last even-started in-grid pair, which only changes rows
// >= ny - 1 (no valid face reads row ny - 1 or beyond; rows below 0 are
// likewise never read).
__device__ __forceinline__ void sweep_dma(const double *col_base, int j0, int ny, double *dst, int lane)
{
    if (lane < kSwRows / 2) {
        const int r0 = min(max(j0 - 2 + 2 * lane, 0), (ny - 2) & ~1);
        __builtin_amdgcn_global_load_lds((glb_void_t *)(col_base + r0), (lds_void_t *)dst, 16, 0, 0);
    }
}

// Accessors over the LDS rings.  Column indices are wave-uniform, so the
// ring slot (column mod ring length) is scalar arithmetic.
struct SweepState {
    const SweepLds *l;
    int j0;
    __device__ double operator()(int v, int ci, int cj) const { return l->st[v][ci % kSwSt][cj - j0 + 2]; }
};

struct SweepEta {
    const SweepLds *l;
    int j0;
    bool field;
    double uniform;
    __device__ double operator()(int ci, int cj) const { return field ? l->et[ci % kSwRing][cj - j0 + 2] : uniform; }
};

struct SweepAdd {
    SweepLds *l;
    int j0;
    __device__ void operator()(int ci, int cj, int v, double b) const
    {
        __hip_atomic_fetch_add(&l->acc[v][ci % kSwAcc][cj - j0 + 2], b, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
};
}  // namespace

template <>
struct mhd_ad::add_skips_zero<SweepAdd> {
    static constexpr bool value = false;
};

namespace {

// Item -> run, first row, face columns [ia, ib).
__device__ __forceinline__ bool sweep_item(const AdjDesc *runs, const long long *sfirst, int nruns, int cols,
                                           int &r, int &j0, int &ia, int &ib)
{
    const long long t = blockIdx.x;
    if (t >= sfirst[nruns]) return false;
    int lo = 0, hi = nruns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sfirst[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    r = lo;
    const long long lt = t - sfirst[lo];
    const int swj = runs[lo].swj;
    j0 = 2 + (int)(lt % swj) * 64;
    ia = 2 + (int)(lt / swj) * cols;
    ib = min(ia + cols, runs[lo].g.nx - 2);
    return true;
}

__device__ __forceinline__ int lm_index(int v) { return v < 4 ? v : v - 2; }

// The fields of a run the sweep reads, loaded once per workgroup (uniform).
struct SweepRun {
    int nx, ny, pitch;
    long long plane;
    const double *prim, *L, *lamE, *eta;
    double *lamC, *geta;
};

// Adds accumulator column m to HBM (global atomics: neighbouring items share
// the edge columns and rows) and clears it.
__device__ __forceinline__ void sweep_flush(SweepLds &l, const SweepRun &g, int m, int j0, int tid)
{
    const int sa = m % kSwAcc;
    // components 0..7: 16 threads each, rows tid & 15 + 16 k (five passes,
    // unrolled so that the five LDS reads issue together); the resistivity
    // adjoint: 68 threads
    const int v = tid >> 4;
    double *dst = g.lamC + v * g.plane + (size_t)m * g.pitch + (j0 - 2);
    double b[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int r = (tid & 15) + 16 * k;
        b[k] = r < kSwRows ? l.acc[v][sa][r] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int r = (tid & 15) + 16 * k;
        if (r < kSwRows) l.acc[v][sa][r] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int r = (tid & 15) + 16 * k;
        const int gj = j0 - 2 + r;
        if (b[k] != 0.0 && gj >= 0 && gj < g.ny) unsafeAtomicAdd(dst + r, b[k]);
    }
    if (tid < kSwRows) {
        const double b = l.acc[8][sa][tid];
        l.acc[8][sa][tid] = 0.0;
        const int gj = j0 - 2 + tid;
        if (b != 0.0 && gj >= 0 && gj < g.ny) unsafeAtomicAdd(g.geta + (size_t)m * g.pitch + gj, b);
    }
}

// Ring slot that plane p of column col occupies.
__device__ __forceinline__ double *sweep_slot(SweepLds &l, int p, int col)
{
    if (p < 8) return l.st[p][col % kSwSt];
    if (p < 14) return l.lm[p - 8][col % kSwRing];
    if (p == 14) return l.le[col % kSwRing];
    return l.et[col % kSwRing];
}

// One wave's share of the loads of a step, straight into their ring slots:
// wave 0 the 8 state planes of column cs, wave 1 the six adjoint planes and
// the resistivity of column cl and the corner-EMF adjoint of column ce.
template <int STAGE>
__device__ __forceinline__ void sweep_issue(SweepLds &l, const SweepRun &g, int wave, int cs, int cl, int ce, int j0,
                                            int lane)
{
    if (wave == 0) {
        const double *base = g.prim + (size_t)cs * g.pitch;
#pragma unroll
        for (int v = 0; v < 8; ++v) sweep_dma(base + v * g.plane, j0, g.ny, sweep_slot(l, v, cs), lane);
    } else {
        const double *base = g.L + (size_t)cl * g.pitch;
#pragma unroll
        for (int q = 0; q < 6; ++q)
            sweep_dma(base + (q < 4 ? q : q + 2) * g.plane, j0, g.ny, sweep_slot(l, 8 + q, cl), lane);
        if (STAGE == 1) sweep_dma(g.lamE + (size_t)ce * g.pitch, j0, g.ny, sweep_slot(l, 14, ce), lane);
        if (g.eta) sweep_dma(g.eta + (size_t)cl * g.pitch, j0, g.ny, sweep_slot(l, 15, cl), lane);
    }
}

template <int STAGE, bool DEGEN>
__global__ __launch_bounds__(128) void face_adj_sweep_kernel(const AdjDesc *runs, const long long *sfirst, int nruns,
                                                             int cols, int step, int slot, int ck)
{
    __shared__ SweepLds l;
    int r, j0, ia, ib;
    if (!sweep_item(runs, sfirst, nruns, cols, r, j0, ia, ib)) return;
    r = __builtin_amdgcn_readfirstlane(r);
    j0 = __builtin_amdgcn_readfirstlane(j0);
    ia = __builtin_amdgcn_readfirstlane(ia);
    ib = __builtin_amdgcn_readfirstlane(ib);
    const AdjDesc &d = runs[r];
    SweepRun g;
    g.nx = d.g.nx;
    g.ny = d.g.ny;
    g.pitch = d.g.pitch;
    g.plane = d.g.plane;
    g.prim = stage_prim<STAGE>(d, slot, ck);
    g.L = STAGE == 2 ? d.lam : d.lamBar;
    g.lamE = d.lamE;
    g.eta = d.eta;
    g.lamC = d.lamC;
    g.geta = d.geta;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // 0: x faces, 1: y faces
    const int lane = tid & 63;
    const int j = j0 + lane;
    const double dt = d.dts[step];
    const double coef = (STAGE == 2 ? 0.5 : 1.0) * dt / (wave == 0 ? d.ph.dx : d.ph.dy);
    const double par[4] = {d.ph.gamma - 1.0, d.ph.gamma, d.ph.eps, DEGEN ? d.dtol : 0.0};
    const double rpar[2] = {d.ph.dx, d.ph.dy};
    const double eta_uniform = d.ph.eta;

    // every column step ia reads, straight into the rings: state ia-2 ..
    // ia+2, adjoint and resistivity ia, ia+1, corner-EMF adjoint ia-1, ia;
    // the two waves take alternate planes
    for (int k = tid; k < 9 * kSwAcc * kSwRows; k += 128) (&l.acc[0][0][0])[k] = 0.0;
    if (wave == 0) {
        for (int c = ia - 2; c <= ia + 2; ++c)
            for (int v = 0; v < 8; ++v)
                sweep_dma(g.prim + v * g.plane + (size_t)c * g.pitch, j0, g.ny, l.st[v][c % kSwSt], lane);
    } else {
        for (int c = ia; c <= ia + 1; ++c) {
            for (int q = 0; q < 6; ++q)
                sweep_dma(g.L + (q < 4 ? q : q + 2) * g.plane + (size_t)c * g.pitch, j0, g.ny, l.lm[q][c % kSwRing],
                          lane);
            if (g.eta) sweep_dma(g.eta + (size_t)c * g.pitch, j0, g.ny, l.et[c % kSwRing], lane);
            if (STAGE == 1) sweep_dma(g.lamE + (size_t)(c - 1) * g.pitch, j0, g.ny, l.le[(c - 1) % kSwRing], lane);
        }
    }

    for (int i = ia; i < ib; ++i) {
        // the lane index, opaque to the compiler afresh each column: masks
        // derived from it (flush and copy bounds, face validity) are then
        // recomputed per column rather than hoisted out of the loop and held
        // in scalar registers across the face computation (which spilled
        // them to VGPR lanes)
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int lane = tid & 63;
        const int j = j0 + lane;
        wait_vm_all();       // this wave's loads for step i landed
        __syncthreads();     // and the other wave's; step i-1 done
        if (i - 2 >= ia - 1) sweep_flush(l, g, i - 2, j0, tid);
        if (i + 1 < ib) {
            // the columns step i+1 adds, into ring slots no wave reads at
            // step i: state i+3 (slot of i-3), adjoint and resistivity i+2
            // (slot of i-1), corner-EMF adjoint i+1 (slot of i-2)
            sweep_issue<STAGE>(l, g, wave, i + 3, i + 2, i + 1, j0, lane);
        }

        const bool valid = wave == 0 ? (i <= g.nx - 4 && j >= 2 && j <= g.ny - 3)
                                     : (i <= g.nx - 3 && j >= 2 && j <= g.ny - 4);
        if (valid) {
            const int ri = i % kSwRing, ri1 = (i + 1) % kSwRing;
            const int row = lane + 2;
            double lamF[8];
            bool any = false;
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                lamF[v] = 0.0;
                if (non_b(v)) {
                    const double *Lq = l.lm[lm_index(v)][0];
                    const double here = Lq[ri * kSwRows + row];
                    const double next = wave == 0 ? Lq[ri1 * kSwRows + row] : Lq[ri * kSwRows + row + 1];
                    lamF[v] = coef * (next - here);
                }
                any |= lamF[v] != 0.0;
            }
            if (STAGE == 1) {
                // E(i, j) = (G4(i, j) + G4(i+1, j) - F5(i, j) - F5(i, j+1)) / 4
                const double e0 = l.le[ri][row];
                if (wave == 0) lamF[5] = -0.25 * (e0 + l.le[ri][row - 1]);
                else lamF[4] = 0.25 * (e0 + l.le[(i - 1) % kSwRing][row]);
                any |= lamF[4] != 0.0 || lamF[5] != 0.0;
            }
            if (any) {
                SweepState S{&l, j0};
                SweepEta E{&l, j0, g.eta != nullptr, eta_uniform};
                SweepAdd add{&l, j0};
                mhd_ad::flux_full_adj(S, E, wave, i, j, par, rpar, lamF, add);
            }
        }
    }
    __syncthreads();
    for (int m = max(ia - 1, ib - 2); m <= ib + 1; ++m) sweep_flush(l, g, m, j0, tid);
}

// Adjoint of the face-to-centre averaging (shiftUToCenterForCT) at cell
// (i, j): component v of the stored state.
__device__ __forceinline__ double shift_t(const AdjDesc &d, int v, int i, int j)
{
    const Grid &g = d.g;
    const double *C = d.lamC + v * g.plane;
    const size_t o = (size_t)i * g.pitch + j;
    if (v == 4) return (i >= 1 ? 0.5 * C[o] : C[o]) + (i + 1 < g.nx ? 0.5 * C[o + g.pitch] : 0.0);
    if (v == 5) return (j >= 1 ? 0.5 * C[o] : C[o]) + (j + 1 < g.ny ? 0.5 * C[o + 1] : 0.0);
    return C[o];
}

// step 3: lamBar = shift^T lamC + lam_nonB / 2
__global__ __launch_bounds__(256) void cell2_kernel(const AdjDesc *runs, TileMap tm)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    if (j >= d.g.ny) return;
    const size_t o = (size_t)i * d.g.pitch + j;
    for (int v = 0; v < 8; ++v) {
        const double half = non_b(v) ? 0.5 * d.lam[v * d.g.plane + o] : 0.0;
        d.lamBar[v * d.g.plane + o] = shift_t(d, v, i, j) + half;
    }
}

// step 5: CT adjoint.  B* = B - dt curl E for cells i >= 1, j >= 1
// (P/ct.cu:44-59): bX* = bX - dt/dy (E(i,j) - E(i,j-1)),
// bY* = bY + dt/dx (E(i,j) - E(i-1,j)).
__device__ __forceinline__ double lam_bstar(const AdjDesc &d, int v, int i, int j)
{
    if (i < 1 || j < 1 || i >= d.g.nx || j >= d.g.ny) return 0.0;
    const size_t o = v * d.g.plane + (size_t)i * d.g.pitch + j;
    return d.lamBar[o] + d.lam[o];
}

// Also clears the cell's lamC for the stage-1 face adjoints that follow
// (nothing reads lamC between the stage-2 cell adjoint and them).
__global__ __launch_bounds__(256) void ct_adj_kernel(const AdjDesc *runs, TileMap tm, int step)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    const Grid &g = d.g;
    if (j >= g.ny) return;
    const size_t o = (size_t)i * g.pitch + j;
#pragma unroll
    for (int v = 0; v < 8; ++v) d.lamC[v * g.plane + o] = 0.0;
    const double dt = d.dts[step];
    const double cy = dt / d.ph.dy, cx = dt / d.ph.dx;
    double le = 0.0;
    if (i < g.nx - 1 && j < g.ny - 1)
        le = -cy * lam_bstar(d, 4, i, j) + cy * lam_bstar(d, 4, i, j + 1) + cx * lam_bstar(d, 5, i, j)
           - cx * lam_bstar(d, 5, i + 1, j);
    d.lamE[o] = le;
}

// step 7: lam = lamU + shift^T lamC, with the direct adjoint of U^n
// lamU = lam_nonB / 2 + lamBar_nonB (the stage-2 average and stage 1's
// identity term) and lamU_B = lam_B + lamBar_B on the CT-updated faces
// (i, j >= 1), 0 elsewhere (CT overwrites B from B^n there).  In place: no
// other thread reads this cell's lam (ct_adj, which reads neighbours, ran
// before).
__global__ __launch_bounds__(256) void cell1_kernel(const AdjDesc *runs, TileMap tm)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    if (j >= d.g.ny) return;
    const size_t o = (size_t)i * d.g.pitch + j;
    for (int v = 0; v < 8; ++v) {
        const size_t ov = v * d.g.plane + o;
        const double lu = non_b(v) ? 0.5 * d.lam[ov] + d.lamBar[ov] : lam_bstar(d, v, i, j);
        d.lam[ov] = lu + shift_t(d, v, i, j);
    }
}

// Sets every run's dt replay: dt_in = dts (or nullptr), dt_out = dts (or
// nullptr), and seq0 so that the next step reads/writes entry idx0.
__global__ void ctrl_seq_kernel(const AdjDesc *runs, Ctrl *ctrl, int nruns, int replay, int record, long long idx0)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    Ctrl &c = ctrl[r];
    c.dt_in = replay ? runs[r].dts : nullptr;
    c.dt_out = record ? runs[r].dts : nullptr;
    c.seq0 = c.steps - idx0;
}

__global__ void ctrl_clear_kernel(Ctrl *ctrl, int nruns)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    ctrl[r].dt_in = nullptr;
    ctrl[r].dt_out = nullptr;
    ctrl[r].time = 0.0;
}

__global__ __launch_bounds__(256) void pack_plane_kernel(const double *__restrict__ src, double *__restrict__ dst,
                                                         Grid g)
{
    const long long total = (long long)g.nx * g.ny;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(k / g.ny), j = (int)(k % g.ny);
        dst[k] = src[(size_t)i * g.pitch + j];
    }
}


// ------------------------------------------------- tangent-linear model
// dF, dG of the faces the step reads (zero elsewhere), from the primal state
// (U or U*) and its tangent, with the tangent of eta.
template <int DIR>
__global__ __launch_bounds__(256) void face_tlm_kernel(const AdjDesc *runs, TileMap tm,
                                                       int stage)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    const Grid &g = d.g;
    if (j >= g.ny) return;
    const size_t o = (size_t)i * g.pitch + j;
    double *out = DIR == 0 ? d.dF : d.dG;
    const bool inside = DIR == 0 ? (i >= 2 && i <= g.nx - 4 && j >= 2 && j <= g.ny - 3)
                                 : (i >= 2 && i <= g.nx - 3 && j >= 2 && j <= g.ny - 4);
    if (!inside) {
#pragma unroll
        for (int v = 0; v < 8; ++v) out[v * g.plane + o] = 0.0;
        return;
    }
    const Soa S{stage == 2 ? d.us : d.u, g.pitch, g.plane};
    const Soa dS{stage == 2 ? d.dus : d.du, g.pitch, g.plane};
    const EtaAcc E{d.eta, g.pitch, d.ph.eta};
    const EtaAcc dE{d.deta, g.pitch, 0.0};
    const double par[4] = {d.ph.gamma - 1.0, d.ph.gamma, d.ph.eps, d.dtol};
    const double rpar[2] = {d.ph.dx, d.ph.dy};
    double F[8], dF[8];
    mhd_ad::flux_full_tlm(S, dS, E, dE, DIR, i, j, par, rpar, F, dF);
#pragma unroll
    for (int v = 0; v < 8; ++v) out[v * g.plane + o] = dF[v];
}

// Decision replay of the face fluxes: the flux of the perturbed state P (the
// tangent's state buffers hold it) and perturbed resistivity (the direction's
// eta buffer; a run without an eta plane keeps its uniform eta), with every
// branch decision taken at the primal state.  The rest of the step is
// linear, so the tangent's EMF, stage and boundary kernels advance P
// exactly as the forward step advances a state.
template <int DIR>
__global__ __launch_bounds__(256) void face_rep_kernel(const AdjDesc *runs, TileMap tm,
                                                       int stage)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    const Grid &g = d.g;
    if (j >= g.ny) return;
    const size_t o = (size_t)i * g.pitch + j;
    double *out = DIR == 0 ? d.dF : d.dG;
    const bool inside = DIR == 0 ? (i >= 2 && i <= g.nx - 4 && j >= 2 && j <= g.ny - 3)
                                 : (i >= 2 && i <= g.nx - 3 && j >= 2 && j <= g.ny - 4);
    if (!inside) {
#pragma unroll
        for (int v = 0; v < 8; ++v) out[v * g.plane + o] = 0.0;
        return;
    }
    const Soa S{stage == 2 ? d.us : d.u, g.pitch, g.plane};
    const Soa SP{stage == 2 ? d.dus : d.du, g.pitch, g.plane};
    const EtaAcc E{d.eta, g.pitch, d.ph.eta};
    const EtaAcc EP{d.eta ? d.deta : nullptr, g.pitch, d.ph.eta};
    const double par[4] = {d.ph.gamma - 1.0, d.ph.gamma, d.ph.eps, d.dtol};
    const double rpar[2] = {d.ph.dx, d.ph.dy};
    double F[8], FP[8];
    mhd_ad::flux_full_rep(S, SP, E, EP, DIR, i, j, par, rpar, F, FP);
#pragma unroll
    for (int v = 0; v < 8; ++v) out[v * g.plane + o] = FP[v];
}

// Tangent of the corner EMF E(i, j) = (G4(i,j) + G4(i+1,j) - F5(i,j) - F5(i,j+1)) / 4.
__global__ __launch_bounds__(256) void emf_tlm_kernel(const AdjDesc *runs, TileMap tm)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    const Grid &g = d.g;
    if (j >= g.ny) return;
    const size_t o = (size_t)i * g.pitch + j;
    double e = 0.0;
    if (i < g.nx - 1 && j < g.ny - 1)
        e = 0.25 * (d.dG[4 * g.plane + o] + d.dG[4 * g.plane + o + g.pitch] - d.dF[5 * g.plane + o]
                    - d.dF[5 * g.plane + o + 1]);
    d.dE[o] = e;
}

// Tangent of a stage update of the interior cells (ghosts are refilled by
// the boundary): stage 1 writes dU* (B by CT), stage 2 writes dU (B = dB*).
__global__ __launch_bounds__(256) void stage_tlm_kernel(const AdjDesc *runs, TileMap tm,
                                                        int stage, int step)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    const Grid &g = d.g;
    if (i < 3 || i > g.nx - 4 || j < 2 || j > g.ny - 4) return;
    const size_t o = (size_t)i * g.pitch + j;
    const double dt = d.dts[step];
    const double dtdx = dt / d.ph.dx, dtdy = dt / d.ph.dy;
    if (j == 2) {
        // By of the bottom boundary face after the CT, which the symmetric-x
        // fill reads before the y fill overwrites it; stage 2 carries the
        // same value (U^n's face and the stage-1 EMF, both unchanged)
        const size_t q = 5 * g.plane + o;
        const double v = d.du[q] + dtdx * (d.dE[o] - d.dE[o - g.pitch]);
        if (stage == 1) d.dus[q] = v;
        else d.du[q] = v;
        return;
    }
    for (int v = 0; v < 8; ++v) {
        const size_t q = v * g.plane + o;
        if (v == 4 || v == 5) {
            if (stage == 1) {
                const double de = d.dE[o];
                d.dus[q] = v == 4 ? d.du[q] - dtdy * (de - d.dE[o - 1]) : d.du[q] + dtdx * (de - d.dE[o - g.pitch]);
            } else {
                d.du[q] = d.dus[q];   // B^{n+1} = B*
            }
            continue;
        }
        // the forward stages' operation order (P/resistiveMHD_2D.cu:67-82,
        // 103-126), so that the decision replay at the recorded initial
        // state is the forward rollout bit for bit
        const double fx = dtdx * (d.dF[q] - d.dF[q - g.pitch]), gy = dtdy * (d.dG[q] - d.dG[q - 1]);
        if (stage == 1) d.dus[q] = d.du[q] - fx - gy;
        else d.du[q] = 0.5 * (d.du[q] + d.dus[q] - fx - gy);
    }
}

// The boundary fill applied to a tangent state (it is linear): periodic or
// symmetric x, then periodic or symmetric y (the same copies and Bx / By
// rebuilds as the forward fill, in its operation order, so that the
// decision replay advances a state bit for bit as the forward step does).
__global__ __launch_bounds__(256) void bc_tlm_x_kernel(const AdjDesc *runs, int stage)
{
    const AdjDesc &d = runs[blockIdx.y];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d.g.ny) return;
    double *T = stage == 1 ? d.dus : d.du;
    const int nx = d.g.nx;
    const size_t P = d.g.pitch;
    const long long pl = d.g.plane;
    const int dst[6] = {0, 1, 2, nx - 3, nx - 2, nx - 1};
    if (d.bc_x == MHD_BC_PERIODIC) {
        const int src[6] = {nx - 6, nx - 5, nx - 4, 3, 4, 5};
        for (int v = 0; v < 8; ++v)
            for (int q = 0; q < 6; ++q) T[v * pl + dst[q] * P + j] = T[v * pl + src[q] * P + j];
        return;
    }
    const int src[6] = {5, 4, 3, nx - 4, nx - 5, nx - 6};
    for (int v = 0; v < 8; ++v) {
        if (v == 4) continue;
        for (int q = 0; q < 6; ++q) T[v * pl + dst[q] * P + j] = T[v * pl + src[q] * P + j];
    }
    const double *by = T + 5 * pl;
    double *bx = T + 4 * pl;
    const double dx = d.ph.dx, dy = d.ph.dy;
    if (j > 0) {
        const size_t c3 = 3 * P + j, cA = (size_t)(nx - 4) * P + j;
        const double d3 = by[c3] - by[c3 - 1];
        double b = bx[c3] + d3 / dy * dx;
        bx[2 * P + j] = b;
        b = b + d3 / dy * dx;
        bx[P + j] = b;
        bx[j] = b + (by[c3 + P] - by[c3 + P - 1]) / dy * dx;
        b = bx[cA] - (by[cA] - by[cA - 1]) / dy * dx;
        bx[cA + P] = b;
        b = b - (by[cA - P] - by[cA - P - 1]) / dy * dx;
        bx[cA + 2 * P] = b;
        bx[cA + 3 * P] = b - (by[cA - 2 * P] - by[cA - 2 * P - 1]) / dy * dx;
    } else {
        bx[0] = bx[P] = bx[2 * P] = bx[3 * P];
        const size_t cA = (size_t)(nx - 4) * P;
        bx[cA + P] = bx[cA + 2 * P] = bx[cA + 3 * P] = bx[cA];
    }
}

__global__ __launch_bounds__(256) void bc_tlm_y_kernel(const AdjDesc *runs, int stage)
{
    const AdjDesc &d = runs[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.g.nx) return;
    double *T = stage == 1 ? d.dus : d.du;
    const long long pl = d.g.plane;
    const int ny = d.g.ny;
    const size_t row = (size_t)i * d.g.pitch;
    const int dst[6] = {0, 1, 2, ny - 3, ny - 2, ny - 1};
    if (d.bc_y == MHD_BC_PERIODIC) {
        const int src[6] = {ny - 6, ny - 5, ny - 4, 3, 4, 5};
        for (int v = 0; v < 8; ++v)
            for (int q = 0; q < 6; ++q) T[v * pl + row + dst[q]] = T[v * pl + row + src[q]];
        return;
    }
    const int src[6] = {5, 4, 3, ny - 4, ny - 5, ny - 6};
    for (int v = 0; v < 8; ++v) {
        if (v == 5) continue;
        for (int q = 0; q < 6; ++q) T[v * pl + row + dst[q]] = T[v * pl + row + src[q]];
    }
    double *by = T + 5 * pl + row;
    const double *bx = T + 4 * pl;
    if (i > 0) {
        const size_t rm = (size_t)(i - 1) * d.g.pitch;
        const double d3 = bx[row + 3] - bx[rm + 3], d4 = bx[row + 4] - bx[rm + 4];
        const double dA = bx[row + ny - 4] - bx[rm + ny - 4], dB = bx[row + ny - 5] - bx[rm + ny - 5];
        const double dC = bx[row + ny - 6] - bx[rm + ny - 6];
        by[2] = by[3] + d3 / d.ph.dx * d.ph.dy;
        by[1] = by[2] + d3 / d.ph.dx * d.ph.dy;
        by[0] = by[1] + d4 / d.ph.dx * d.ph.dy;
        by[ny - 3] = by[ny - 4] - dA / d.ph.dx * d.ph.dy;
        by[ny - 2] = by[ny - 3] - dB / d.ph.dx * d.ph.dy;
        by[ny - 1] = by[ny - 2] - dC / d.ph.dx * d.ph.dy;
    } else {
        by[2] = by[1] = by[0] = by[3];
        by[ny - 3] = by[ny - 2] = by[ny - 1] = by[ny - 4];
    }
}

__global__ __launch_bounds__(256) void tlm_init_kernel(const AdjDesc *runs, TileMap tm)
{
    int r, i, j;
    if (!tile_cell(tm, r, i, j, runs)) return;
    const AdjDesc &d = runs[r];
    if (j >= d.g.ny) return;
    const size_t o = (size_t)i * d.g.pitch + j;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        d.du[v * d.g.plane + o] = d.du0[v * d.g.plane + o];
        d.dus[v * d.g.plane + o] = 0.0;
    }
}

thread_local std::string g_adj_error;

}  // namespace

struct mhd_adjoint {
    mhd_batch *fwd = nullptr;
    int nruns = 0;
    std::vector<AdjDesc> desc;
    AdjDesc *d_desc = nullptr;
    long long *d_first = nullptr;
    int *d_tile_run = nullptr;    // run of each tile
    TileMap tiles{};
    long long total_tiles = 0;
    // sweep form of the face adjoints: item prefix per run, columns per item
    long long *d_sfirst = nullptr;
    long long sweep_items = 0;
    // face rows the strips leave over: lane prefix per run, per direction
    long long *d_tfirst = nullptr;
    long long tail_faces[2] = {0, 0};
    int sweep_cols = 0;
    int face_kernel = MHD_ADJ_FACE_AUTO;
    bool face_kernel_auto_sweep = false;
    double dtol = 0.0;            // degenerate star-state threshold (AdjDesc::dtol of every run)
    double *pool = nullptr;       // lam, lamBar, lamC, lamE, geta, u0 per run
    double *hist = nullptr, *hists = nullptr, *ckpt = nullptr, *dts = nullptr;
    // per history slot k, three copies of the batch's run descriptors: A_k
    // (U = hist[k], U* = hists[k]), B_k (as A_k, stage 2 writing hist[k+1])
    // and C_k (as A_k, stage 2 writing the live state), so a step reads and
    // writes its states straight in the history
    mhd::RunDesc *d_slot_desc = nullptr;
    size_t slot_desc_cap = 0;   // descriptors
    bool last_in_hist = false;  // the forward pass left its last segment in the history
    size_t hist_cap = 0, hists_cap = 0, ckpt_cap = 0, dts_cap = 0;   // doubles
    int nsteps = 0, stride = 0, nckpt = 0;
    bool forward_done = false;
    double *tpool = nullptr;      // tangent buffers (allocated on first use)
    double *stage_all = nullptr;  // AoS staging of every run at once (allocated on first use)
    std::vector<size_t> aos_off;  // offset of each run's AoS state in stage_all
    std::string err;
};

namespace {

int afail(mhd_adjoint *a, int code, const std::string &msg)
{
    if (a) a->err = msg;
    g_adj_error = msg;
    return code;
}

#define AHIP(a, call)                                                                                       \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) return afail(a, MHD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

hipStream_t stream_of(const mhd_adjoint *a) { return a->fwd->stream; }

int upload_desc(mhd_adjoint *a)
{
    AHIP(a, hipMemcpyAsync(a->d_desc, a->desc.data(), sizeof(AdjDesc) * a->desc.size(), hipMemcpyHostToDevice,
                           stream_of(a)));
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    return MHD_OK;
}

dim3 tiles_grid(const mhd_adjoint *a) { return dim3((unsigned)((a->total_tiles + 3) / 4)); }
dim3 tiles_grid1(const mhd_adjoint *a) { return dim3((unsigned)a->total_tiles); }

// Sweep items: a 64-row strip of a run and `cols` face columns.  cols (8 to
// 64) minimises rounds of resident workgroups (two per CU) times the
// columns each item sweeps plus its start-up (about three columns' worth:
// loading five state columns, flushing four accumulator columns), so the
// last round is not left half empty.
// Face rows up to this many past a run's last full 64-row strip go to the
// per-face tail kernel instead of a mostly idle extra strip (x faces span
// ny - 4 rows, 64k + 2 for a 64k-row interior).
constexpr int kSwTailRows = 16;

int layout_sweep(mhd_adjoint *a)
{
    for (AdjDesc &d : a->desc) {
        const int rows = d.g.ny - 4;   // x-face rows 2 .. ny-3 (the y faces stop a row earlier)
        d.swj = rows / 64 + (rows % 64 > kSwTailRows ? 1 : 0);
        d.tail0 = 2 + 64 * d.swj;
    }
    std::vector<long long> tfirst(2 * (a->nruns + 1), 0);
    for (int r = 0; r < a->nruns; ++r) {
        const AdjDesc &d = a->desc[r];
        const long long tx = (long long)(d.g.nx - 5) * std::max(0, d.g.ny - 2 - d.tail0);
        const long long ty = (long long)(d.g.nx - 4) * std::max(0, d.g.ny - 3 - d.tail0);
        tfirst[r + 1] = tfirst[r] + tx;
        tfirst[a->nruns + 1 + r + 1] = tfirst[a->nruns + 1 + r] + ty;
    }
    a->tail_faces[0] = tfirst[a->nruns];
    a->tail_faces[1] = tfirst[2 * a->nruns + 1];
    AHIP(a, hipMalloc(&a->d_tfirst, sizeof(long long) * tfirst.size()));
    AHIP(a, hipMemcpyAsync(a->d_tfirst, tfirst.data(), sizeof(long long) * tfirst.size(), hipMemcpyHostToDevice,
                           stream_of(a)));
    long long strips_cols = 0;
    for (const AdjDesc &d : a->desc) strips_cols += (long long)d.swj * (d.g.nx - 4);
    const long long slots = 2LL * std::max(1, a->fwd->num_cu);
    const long long target = 16 * slots;
    double best_cost = 1e300;
    a->sweep_cols = 8;
    for (int cols = 8; cols <= 64; ++cols) {
        long long items = 0;
        for (const AdjDesc &d : a->desc) items += (long long)d.swj * ((d.g.nx - 4 + cols - 1) / cols);
        const double cost = (double)((items + slots - 1) / slots) * (cols + 3);
        if (cost < best_cost - 1e-9) {
            best_cost = cost;
            a->sweep_cols = cols;
        }
    }
    std::vector<long long> first(a->nruns + 1, 0);
    for (int r = 0; r < a->nruns; ++r) {
        const AdjDesc &d = a->desc[r];
        first[r + 1] = first[r] + (long long)d.swj * ((d.g.nx - 4 + a->sweep_cols - 1) / a->sweep_cols);
    }
    a->sweep_items = first[a->nruns];
    AHIP(a, hipMalloc(&a->d_sfirst, sizeof(long long) * (a->nruns + 1)));
    AHIP(a, hipMemcpyAsync(a->d_sfirst, first.data(), sizeof(long long) * (a->nruns + 1), hipMemcpyHostToDevice,
                           stream_of(a)));
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    // the per-face form wins where the sweep would leave the machine idle:
    // fewer strip-columns than ~16 rounds of 8-column items
    a->face_kernel_auto_sweep = strips_cols >= target * 8;
    return MHD_OK;
}

bool use_sweep(const mhd_adjoint *a)
{
    return a->face_kernel == MHD_ADJ_FACE_SWEEP || (a->face_kernel == MHD_ADJ_FACE_AUTO && a->face_kernel_auto_sweep);
}

template <int STAGE, bool DEGEN>
void launch_face_adj_t(const mhd_adjoint *a, int step, int slot, int ck, hipStream_t s)
{
    if (use_sweep(a)) {
        if (a->sweep_items)
            hipLaunchKernelGGL((face_adj_sweep_kernel<STAGE, DEGEN>), dim3((unsigned)a->sweep_items), dim3(128), 0, s,
                               a->d_desc, a->d_sfirst, a->nruns, a->sweep_cols, step, slot, ck);
        const unsigned xb = (unsigned)((a->tail_faces[0] + 63) / 64), yb = (unsigned)((a->tail_faces[1] + 63) / 64);
        if (xb + yb)
            hipLaunchKernelGGL((face_adj_tail_kernel<STAGE, DEGEN>), dim3(xb + yb), dim3(64), 0, s, a->d_desc,
                               a->d_tfirst, a->nruns, xb, step, slot, ck);
    } else {
        const dim3 tg1 = tiles_grid1(a), blk(64);
        hipLaunchKernelGGL((face_adj_kernel<STAGE, 0, DEGEN>), tg1, blk, 0, s, a->d_desc, a->tiles, step, slot, ck);
        hipLaunchKernelGGL((face_adj_kernel<STAGE, 1, DEGEN>), tg1, blk, 0, s, a->d_desc, a->tiles, step, slot, ck);
    }
}

// The face-adjoint launches of one step (stage s, both directions); ck: the
// checkpoint that holds slot 0's U^n.
int launch_face_adj(mhd_adjoint *a, int stage, int step, int slot, int ck)
{
    hipStream_t s = stream_of(a);
    const bool degen = a->dtol > 0.0;
    if (stage == 2) {
        if (degen) launch_face_adj_t<2, true>(a, step, slot, ck, s);
        else launch_face_adj_t<2, false>(a, step, slot, ck, s);
    } else {
        if (degen) launch_face_adj_t<1, true>(a, step, slot, ck, s);
        else launch_face_adj_t<1, false>(a, step, slot, ck, s);
    }
    AHIP(a, hipGetLastError());
    return MHD_OK;
}

int copy_all(mhd_adjoint *a, int src, int src_slot, int dst, int dst_slot)
{
    hipLaunchKernelGGL(copy_kernel, tiles_grid(a), dim3(256), 0, stream_of(a), a->d_desc, a->tiles, src,
                       src_slot, dst, dst_slot);
    AHIP(a, hipGetLastError());
    return MHD_OK;
}

int bc_transpose(mhd_adjoint *a, int which)
{
    const mhd_batch *b = a->fwd;
    hipLaunchKernelGGL(bct_y_kernel, dim3((b->max_nx + 4 * kBctCols - 1) / (4 * kBctCols), a->nruns), dim3(256), 0,
                       stream_of(a), a->d_desc, which);
    hipLaunchKernelGGL(bct_x_kernel, dim3((6 * ((b->max_ny + 63) / 64) + 3) / 4, a->nruns), dim3(256), 0,
                       stream_of(a), a->d_desc, which);
    AHIP(a, hipGetLastError());
    return MHD_OK;
}

int set_replay(mhd_adjoint *a, bool replay, bool record, long long idx0)
{
    hipLaunchKernelGGL(ctrl_seq_kernel, dim3((a->nruns + 63) / 64), dim3(64), 0, stream_of(a), a->d_desc,
                       a->fwd->ctrl, a->nruns, replay ? 1 : 0, record ? 1 : 0, idx0);
    AHIP(a, hipGetLastError());
    return MHD_OK;
}

// Grows a device buffer to hold at least n doubles (contents not kept).
int ensure(mhd_adjoint *a, double **p, size_t *cap, size_t n)
{
    if (*cap >= n) return MHD_OK;
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    AHIP(a, hipFree(*p));
    *p = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc(p, sizeof(double) * n);
    if (e != hipSuccess)
        return afail(a, e == hipErrorOutOfMemory ? MHD_ERR_NOMEM : MHD_ERR_HIP,
                     std::string("adjoint storage: ") + hipGetErrorString(e));
    *cap = n;
    return MHD_OK;
}

// Checkpoint stride: explicit (> 0), ceil(sqrt(nsteps)) (0), or for
// MHD_ADJ_STRIDE_FIT the largest stride up to nsteps whose history and
// checkpoint storage, (2 K + ceil(nsteps / K)) states of every run, fits in
// half of the device memory free now plus what this object already holds
// (ensure() frees it before it allocates more); at nsteps the whole rollout
// stays in the history and the backward pass recomputes nothing.
int choose_stride(const mhd_adjoint *a, int nsteps, int stride)
{
    const int kmin = std::max(1, (int)std::ceil(std::sqrt((double)nsteps)));
    if (stride > 0) return stride;
    if (stride == 0) return kmin;
    double state = 0.0;   // bytes of one state of every run
    for (const AdjDesc &d : a->desc) state += 8.0 * 8.0 * (double)d.g.plane;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return kmin;
    const double budget = 0.5 * (double)free_b + 8.0 * (double)(a->hist_cap + a->hists_cap + a->ckpt_cap);
    for (int K = nsteps; K > kmin; --K) {
        const int nck = (nsteps + K - 1) / K;
        if ((2.0 * K + nck) * state <= budget) return K;
    }
    return kmin;
}

int layout_storage(mhd_adjoint *a, int nsteps, int stride)
{
    const int K = choose_stride(a, nsteps, stride);
    const int nck = (nsteps + K - 1) / K;
    size_t h = 0, c = 0;
    for (const AdjDesc &d : a->desc) {
        h += (size_t)K * 8 * d.g.plane;
        c += (size_t)nck * 8 * d.g.plane;
    }
    int rc;
    if ((rc = ensure(a, &a->hist, &a->hist_cap, h))) return rc;
    if ((rc = ensure(a, &a->hists, &a->hists_cap, h))) return rc;
    if ((rc = ensure(a, &a->ckpt, &a->ckpt_cap, c))) return rc;
    size_t ho = 0, co = 0;
    for (AdjDesc &d : a->desc) {
        d.hist = a->hist + ho;
        d.hists = a->hists + ho;
        d.ckpt = a->ckpt + co;
        ho += (size_t)K * 8 * d.g.plane;
        co += (size_t)nck * 8 * d.g.plane;
    }
    a->stride = K;
    a->nckpt = nck;
    const mhd_batch *b = a->fwd;
    // K history slots, then nck checkpoint variants of slot 0 (U^n read from
    // the checkpoint in place)
    std::vector<mhd::RunDesc> sd((size_t)3 * (K + nck) * a->nruns);
    for (int k = 0; k < K + nck; ++k)
        for (int r = 0; r < a->nruns; ++r) {
            const AdjDesc &d = a->desc[r];
            const size_t s8 = (size_t)8 * d.g.plane;
            const int slot = k < K ? k : 0;
            mhd::RunDesc A = b->desc[r];
            A.u = k < K ? d.hist + slot * s8 : d.ckpt + (size_t)(k - K) * s8;
            A.us = d.hists + slot * s8;
            A.uo = nullptr;
            mhd::RunDesc B = A, C = A;
            B.uo = d.hist + (size_t)std::min(slot + 1, K - 1) * s8;
            C.uo = b->desc[r].u;
            sd[(size_t)(3 * k) * a->nruns + r] = A;
            sd[(size_t)(3 * k + 1) * a->nruns + r] = B;
            sd[(size_t)(3 * k + 2) * a->nruns + r] = C;
        }
    if (a->slot_desc_cap < sd.size()) {
        AHIP(a, hipStreamSynchronize(stream_of(a)));
        AHIP(a, hipFree(a->d_slot_desc));
        a->d_slot_desc = nullptr;
        a->slot_desc_cap = 0;
        AHIP(a, hipMalloc(&a->d_slot_desc, sizeof(mhd::RunDesc) * sd.size()));
        a->slot_desc_cap = sd.size();
    }
    AHIP(a, hipMemcpyAsync(a->d_slot_desc, sd.data(), sizeof(mhd::RunDesc) * sd.size(), hipMemcpyHostToDevice,
                           stream_of(a)));
    return upload_desc(a);   // synchronises: sd stays alive until the copy is done
}

const mhd::RunDesc *slot_desc(const mhd_adjoint *a, int k, int which)
{
    return a->d_slot_desc + (size_t)(3 * k + which) * a->nruns;
}

// Steps s0 .. s0+n-1 of the segment that starts from checkpoint ck, with
// every state written into the history: slot k gets U^{s0+k} (slot 0's is
// the checkpoint itself, read in place) and U* of that step.  last_to_live:
// the final step's U^{s0+n} goes to the live state (the forward pass's
// result); otherwise the final step stops after stage 1 (the backward pass
// needs its U* only).  Steps take dt as the control blocks say (record or
// replay).
int history_steps(mhd_adjoint *a, int n, bool last_to_live, int ck)
{
    mhd_batch *b = a->fwd;
    hipStream_t s = stream_of(a);
    for (int k = 0; k < n; ++k) {
        const int kd = k == 0 ? a->stride + ck : k;   // descriptor set of slot k
        const mhd::RunDesc *A = slot_desc(a, kd, 0);
        AHIP(a, mhd::launch_stage_batch(1, b->exact, A, b->d_item_run, b->total_items, b->xcds, b->items_per_xcd,
                                        b->blocks_per_xcd, b->cur, s, b->d_tail_run, b->tail_items));
        AHIP(a, mhd::launch_bc_batch(b->exact, A, b->nruns, b->max_nx, b->max_ny, true, false, 0, s));
        // stage 2 reads U^n and U* from slot k and writes U^{n+1} into slot
        // k+1 (or the live state); the boundary fill follows it there
        const mhd::RunDesc *B = nullptr, *after = nullptr;
        if (k + 1 < n) {
            B = slot_desc(a, kd, 1);
            after = slot_desc(a, k + 1, 0);
        } else if (last_to_live) {
            B = slot_desc(a, kd, 2);
            after = b->d_desc;
        }
        if (B) {
            AHIP(a, mhd::launch_stage_batch(2, b->exact, B, b->d_item_run, b->total_items, b->xcds, b->items_per_xcd,
                                            b->blocks_per_xcd, b->cur, s, b->d_tail_run, b->tail_items));
            AHIP(a, mhd::launch_bc_batch(b->exact, after, b->nruns, b->max_nx, b->max_ny, false, true, b->cur ^ 1,
                                         s));
            b->cur ^= 1;
        }
    }
    return MHD_OK;
}

int ensure_dts(mhd_adjoint *a, int nsteps)
{
    int rc;
    if (a->dts_cap < (size_t)nsteps * a->nruns) {
        // keep recorded values when growing
        std::vector<double> old;
        if (a->dts && a->dts_cap) {
            old.resize(a->dts_cap);
            AHIP(a, hipMemcpy(old.data(), a->dts, sizeof(double) * a->dts_cap, hipMemcpyDeviceToHost));
        }
        const size_t oldper = a->nruns ? a->dts_cap / a->nruns : 0;
        size_t cap = a->dts_cap;
        if ((rc = ensure(a, &a->dts, &cap, (size_t)nsteps * a->nruns))) return rc;
        a->dts_cap = cap;
        std::vector<double> fresh((size_t)nsteps * a->nruns, 0.0);
        for (int r = 0; r < a->nruns && oldper; ++r)
            std::copy(old.begin() + r * oldper, old.begin() + (r + 1) * oldper, fresh.begin() + (size_t)r * nsteps);
        AHIP(a, hipMemcpy(a->dts, fresh.data(), sizeof(double) * fresh.size(), hipMemcpyHostToDevice));
        for (int r = 0; r < a->nruns; ++r) a->desc[r].dts = a->dts + (size_t)r * nsteps;
        return upload_desc(a);
    }
    return MHD_OK;
}

int check_run(mhd_adjoint *a, int run)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    if (run < 0 || run >= a->nruns) return afail(a, MHD_ERR_INVALID, "run index out of range");
    return MHD_OK;
}

int per_run_steps(const mhd_adjoint *a) { return a->nruns ? (int)(a->dts_cap / a->nruns) : 0; }

}  // namespace

extern "C" {

const char *mhd_adjoint_last_error(const mhd_adjoint *a) { return a ? a->err.c_str() : g_adj_error.c_str(); }

int mhd_adjoint_create(const mhd_config *cfgs, int nruns, mhd_adjoint **out)
{
    if (!out) return afail(nullptr, MHD_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!cfgs || nruns < 1) return afail(nullptr, MHD_ERR_INVALID, "need at least one run configuration");
    mhd_adjoint *a = new (std::nothrow) mhd_adjoint();
    if (!a) return afail(nullptr, MHD_ERR_NOMEM, "host allocation failed");
    int rc = mhd_batch_create(cfgs, nruns, &a->fwd);
    if (rc) {
        const std::string msg = mhd_batch_last_error(nullptr);
        delete a;
        return afail(nullptr, rc, msg);
    }
    a->nruns = nruns;
    a->desc.resize(nruns);
    std::vector<long long> first(nruns + 1, 0);
    size_t total = 0;
    for (int r = 0; r < nruns; ++r) {
        const mhd::RunDesc &rd = a->fwd->desc[r];
        AdjDesc &d = a->desc[r];
        std::memset(&d, 0, sizeof d);
        d.g = rd.g;
        d.ph = rd.ph;
        d.bc_x = rd.bc_x;
        d.bc_y = rd.bc_y;
        d.ntj = (rd.g.ny + 63) / 64;
        d.u = rd.u;
        d.us = rd.us;
        first[r + 1] = first[r] + (long long)rd.g.nx * d.ntj;
        d.aos = r ? a->desc[r - 1].aos + 8LL * a->desc[r - 1].g.nx * a->desc[r - 1].g.ny : 0;
        total += (size_t)(3 * 8 + 2 + 8) * rd.g.plane;   // lam, lamBar, lamC; lamE, geta; u0
    }
    a->total_tiles = first[nruns];
    hipError_t e;
    if ((e = hipMalloc(&a->pool, sizeof(double) * total)) != hipSuccess ||
        (e = hipMalloc(&a->d_desc, sizeof(AdjDesc) * nruns)) != hipSuccess ||
        (e = hipMalloc(&a->d_first, sizeof(long long) * (nruns + 1))) != hipSuccess ||
        (e = hipMemsetAsync(a->pool, 0, sizeof(double) * total, a->fwd->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(a->d_first, first.data(), sizeof(long long) * (nruns + 1), hipMemcpyHostToDevice,
                            a->fwd->stream)) != hipSuccess ||
        (e = hipMalloc(&a->d_tile_run, sizeof(int) * (size_t)a->total_tiles)) != hipSuccess) {
        const std::string msg = std::string("mhd_adjoint_create: ") + hipGetErrorString(e);
        mhd_adjoint_destroy(a);
        return afail(nullptr, e == hipErrorOutOfMemory ? MHD_ERR_NOMEM : MHD_ERR_HIP, msg);
    }
    {
        std::vector<int> tile_run((size_t)a->total_tiles);
        for (int r = 0; r < nruns; ++r)
            std::fill(tile_run.begin() + first[r], tile_run.begin() + first[r + 1], r);
        if ((e = hipMemcpy(a->d_tile_run, tile_run.data(), sizeof(int) * tile_run.size(), hipMemcpyHostToDevice)) !=
            hipSuccess) {
            const std::string msg = std::string("mhd_adjoint_create: ") + hipGetErrorString(e);
            mhd_adjoint_destroy(a);
            return afail(nullptr, MHD_ERR_HIP, msg);
        }
        a->tiles = TileMap{a->d_first, a->d_tile_run, nruns};
    }
    size_t off = 0;
    for (AdjDesc &d : a->desc) {
        const size_t p8 = (size_t)8 * d.g.plane;
        d.lam = a->pool + off; off += p8;
        d.lamBar = a->pool + off; off += p8;
        d.lamC = a->pool + off; off += p8;
        d.lamE = a->pool + off; off += d.g.plane;
        d.geta = a->pool + off; off += d.g.plane;
        d.u0 = a->pool + off; off += p8;
    }
    if ((rc = layout_sweep(a)) || (rc = upload_desc(a))) {
        const std::string msg = a->err;
        mhd_adjoint_destroy(a);
        return afail(nullptr, rc, msg);
    }
    *out = a;
    return MHD_OK;
}

int mhd_adjoint_destroy(mhd_adjoint *a)
{
    if (!a) return MHD_OK;
    if (a->fwd) (void)hipStreamSynchronize(a->fwd->stream);
    (void)hipFree(a->pool);
    (void)hipFree(a->d_desc);
    (void)hipFree(a->d_first);
    (void)hipFree(a->d_tile_run);
    (void)hipFree(a->d_sfirst);
    (void)hipFree(a->d_tfirst);
    (void)hipFree(a->hist);
    (void)hipFree(a->hists);
    (void)hipFree(a->d_slot_desc);
    (void)hipFree(a->ckpt);
    (void)hipFree(a->dts);
    (void)hipFree(a->tpool);
    (void)hipFree(a->stage_all);
    mhd_batch_destroy(a->fwd);
    delete a;
    return MHD_OK;
}

int mhd_adjoint_autotune(mhd_adjoint *a, int steps, int *columns)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    int rc;
    // tune on the initial states (the next forward pass starts from U0 again)
    if ((rc = copy_all(a, kU0, 0, kU, 0))) return rc;
    if ((rc = mhd_batch_apply_boundary(a->fwd))) return afail(a, rc, a->fwd->err);
    if ((rc = mhd_batch_autotune(a->fwd, steps, columns))) return afail(a, rc, a->fwd->err);
    a->forward_done = false;
    return MHD_OK;
}

int mhd_adjoint_set_face_kernel(mhd_adjoint *a, int mode)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    if (mode != MHD_ADJ_FACE_AUTO && mode != MHD_ADJ_FACE_PER_FACE && mode != MHD_ADJ_FACE_SWEEP)
        return afail(a, MHD_ERR_INVALID, "mode must be MHD_ADJ_FACE_AUTO, _PER_FACE or _SWEEP");
    a->face_kernel = mode;
    return MHD_OK;
}

int mhd_adjoint_set_degenerate_tol(mhd_adjoint *a, double tol)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    if (!(tol >= 0.0 && tol < 1.0)) return afail(a, MHD_ERR_INVALID, "tol must be in [0, 1)");
    a->dtol = tol;
    for (AdjDesc &d : a->desc) d.dtol = tol;
    return upload_desc(a);
}

// The checkpoints and history no longer belong to the inputs: backward needs
// a new forward pass first.
static void invalidate_rollout(mhd_adjoint *a)
{
    a->forward_done = false;
    a->last_in_hist = false;
}

int mhd_adjoint_set_eta_field(mhd_adjoint *a, int run, const double *host_eta)
{
    int rc = check_run(a, run);
    if (rc) return rc;
    if ((rc = mhd_batch_set_eta_field(a->fwd, run, host_eta))) return afail(a, rc, a->fwd->err);
    a->desc[run].eta = a->fwd->desc[run].eta;
    invalidate_rollout(a);   // the recorded rollout used the previous field
    return upload_desc(a);
}

int mhd_adjoint_set_state(mhd_adjoint *a, int run, const double *host_aos)
{
    int rc = check_run(a, run);
    if (rc) return rc;
    if (!host_aos) return afail(a, MHD_ERR_INVALID, "NULL state");
    const AdjDesc &d = a->desc[run];
    const size_t bytes = sizeof(double) * 8 * (size_t)d.g.nx * d.g.ny;
    AHIP(a, hipMemcpyAsync(a->fwd->aos, host_aos, bytes, hipMemcpyHostToDevice, stream_of(a)));
    AHIP(a, mhd::launch_unpack(a->fwd->aos, d.u0, d.g, stream_of(a)));
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    invalidate_rollout(a);   // the recorded rollout started from the previous U0
    return MHD_OK;
}

int mhd_adjoint_set_dts(mhd_adjoint *a, int run, const double *host_dts, int nsteps)
{
    int rc = check_run(a, run);
    if (rc) return rc;
    if (!host_dts || nsteps < 1) return afail(a, MHD_ERR_INVALID, "need nsteps >= 1 dt values");
    if ((rc = ensure_dts(a, nsteps))) return rc;
    AHIP(a, hipMemcpy(a->desc[run].dts, host_dts, sizeof(double) * nsteps, hipMemcpyHostToDevice));
    invalidate_rollout(a);   // the recorded rollout's checkpoints used other dt values
    return MHD_OK;
}

int mhd_adjoint_get_dts(mhd_adjoint *a, int run, double *host_dts, int nsteps)
{
    int rc = check_run(a, run);
    if (rc) return rc;
    if (!host_dts || nsteps < 1 || nsteps > per_run_steps(a)) return afail(a, MHD_ERR_INVALID, "bad dt count");
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    AHIP(a, hipMemcpy(host_dts, a->desc[run].dts, sizeof(double) * nsteps, hipMemcpyDeviceToHost));
    return MHD_OK;
}

int mhd_adjoint_forward(mhd_adjoint *a, int nsteps, int stride, int frozen_dt)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    if (nsteps < 1) return afail(a, MHD_ERR_INVALID, "nsteps must be >= 1");
    if (stride < MHD_ADJ_STRIDE_FIT)
        return afail(a, MHD_ERR_INVALID, "stride must be > 0, 0 (ceil(sqrt(nsteps))) or MHD_ADJ_STRIDE_FIT");
    if (frozen_dt && per_run_steps(a) < nsteps) return afail(a, MHD_ERR_STATE, "frozen dt: set nsteps dt values first");
    int rc;
    if ((rc = ensure_dts(a, nsteps))) return rc;
    if ((rc = layout_storage(a, nsteps, stride))) return rc;
    mhd_batch *b = a->fwd;
    // U = BC(U0), time 0; dt recorded (or replayed) from entry 0
    if ((rc = copy_all(a, kU0, 0, kU, 0))) return rc;
    if ((rc = mhd_batch_apply_boundary(b))) return afail(a, rc, b->err);
    hipLaunchKernelGGL(ctrl_clear_kernel, dim3((a->nruns + 63) / 64), dim3(64), 0, stream_of(a), b->ctrl, a->nruns);
    if ((rc = set_replay(a, frozen_dt != 0, frozen_dt == 0, 0))) return rc;
    if ((rc = copy_all(a, kU, 0, kCkpt, 0))) return rc;
    for (int c = 0; c < a->nckpt; ++c) {
        const int n = std::min(a->stride, nsteps - c * a->stride);
        if (c + 1 < a->nckpt) {
            if ((rc = mhd_batch_step(b, n))) return afail(a, rc, b->err);
            if ((rc = copy_all(a, kU, 0, kCkpt, c + 1))) return rc;
        } else {
            // the last segment keeps its states in the history: the backward
            // pass starts there without recomputing it
            if ((rc = mhd_batch_refresh_dt(b))) return afail(a, rc, b->err);
            if ((rc = history_steps(a, n, true, c))) return rc;
        }
    }
    a->last_in_hist = true;
    if ((rc = set_replay(a, false, false, 0))) return rc;
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    a->nsteps = nsteps;
    a->forward_done = true;
    return MHD_OK;
}

int mhd_adjoint_get_stride(const mhd_adjoint *a, int *stride)
{
    if (!a || !stride) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint or stride");
    if (!a->forward_done) return afail(const_cast<mhd_adjoint *>(a), MHD_ERR_STATE, "run mhd_adjoint_forward first");
    *stride = a->stride;
    return MHD_OK;
}

int mhd_adjoint_get_state(mhd_adjoint *a, int run, double *host_aos)
{
    int rc = check_run(a, run);
    if (rc) return rc;
    if ((rc = mhd_batch_get_state(a->fwd, run, host_aos))) return afail(a, rc, a->fwd->err);
    return MHD_OK;
}

int mhd_adjoint_set_seed(mhd_adjoint *a, int run, const double *host_aos)
{
    int rc = check_run(a, run);
    if (rc) return rc;
    if (!host_aos) return afail(a, MHD_ERR_INVALID, "NULL seed");
    const AdjDesc &d = a->desc[run];
    const size_t bytes = sizeof(double) * 8 * (size_t)d.g.nx * d.g.ny;
    AHIP(a, hipMemcpyAsync(a->fwd->aos, host_aos, bytes, hipMemcpyHostToDevice, stream_of(a)));
    AHIP(a, mhd::launch_unpack(a->fwd->aos, d.lam, d.g, stream_of(a)));
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    return MHD_OK;
}

int mhd_adjoint_backward(mhd_adjoint *a)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    if (!a->forward_done) return afail(a, MHD_ERR_STATE, "run mhd_adjoint_forward first");
    mhd_batch *b = a->fwd;
    hipStream_t s = stream_of(a);
    const dim3 tg = tiles_grid(a), blk(256);
    int rc;
    hipLaunchKernelGGL(zero_kernel, tg, blk, 0, s, a->d_desc, a->tiles, 2);
    for (int c = a->nckpt - 1; c >= 0; --c) {
        const int s0 = c * a->stride, s1 = std::min(a->nsteps, s0 + a->stride);
        // recompute the segment with the recorded dt, each step's states
        // written straight into history slot st - s0: U^st and U* =
        // BC(stage 1 of U^st); U^{st+1} starts as a copy of U^st that stage 2
        // updates in place
        if (c == a->nckpt - 1 && a->last_in_hist) {
            a->last_in_hist = false;   // left there by the forward pass; later segments overwrite it
        } else {
            if ((rc = set_replay(a, true, false, s0))) return rc;
            if ((rc = history_steps(a, s1 - s0, false, c))) return rc;
        }
        for (int st = s1 - 1; st >= s0; --st) {
            const int slot = st - s0;
            if ((rc = bc_transpose(a, kLam))) return rc;
            hipLaunchKernelGGL(zero_kernel, tg, blk, 0, s, a->d_desc, a->tiles, 1);
            if ((rc = launch_face_adj(a, 2, st, slot, c))) return rc;
            hipLaunchKernelGGL(cell2_kernel, tg, blk, 0, s, a->d_desc, a->tiles);
            if ((rc = bc_transpose(a, kLamBar))) return rc;
            hipLaunchKernelGGL(ct_adj_kernel, tg, blk, 0, s, a->d_desc, a->tiles, st);   // and clears lamC
            if ((rc = launch_face_adj(a, 1, st, slot, c))) return rc;
            hipLaunchKernelGGL(cell1_kernel, tg, blk, 0, s, a->d_desc, a->tiles);
            AHIP(a, hipGetLastError());
        }
    }
    // the rollout started from BC(U0)
    if ((rc = bc_transpose(a, kLam))) return rc;
    if ((rc = set_replay(a, false, false, 0))) return rc;
    AHIP(a, hipStreamSynchronize(s));
    return MHD_OK;
}

int mhd_adjoint_get_grad_state(mhd_adjoint *a, int run, double *host_aos)
{
    int rc = check_run(a, run);
    if (rc) return rc;
    if (!host_aos) return afail(a, MHD_ERR_INVALID, "NULL output");
    const AdjDesc &d = a->desc[run];
    AHIP(a, mhd::launch_pack(d.lam, a->fwd->aos, d.g, stream_of(a)));
    AHIP(a, hipMemcpyAsync(host_aos, a->fwd->aos, sizeof(double) * 8 * (size_t)d.g.nx * d.g.ny,
                           hipMemcpyDeviceToHost, stream_of(a)));
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    return MHD_OK;
}

int mhd_adjoint_get_grad_eta(mhd_adjoint *a, int run, double *host_eta)
{
    int rc = check_run(a, run);
    if (rc) return rc;
    if (!host_eta) return afail(a, MHD_ERR_INVALID, "NULL output");
    const AdjDesc &d = a->desc[run];
    hipLaunchKernelGGL(pack_plane_kernel, dim3(1024), dim3(256), 0, stream_of(a), d.geta, a->fwd->aos, d.g);
    AHIP(a, hipGetLastError());
    AHIP(a, hipMemcpyAsync(host_eta, a->fwd->aos, sizeof(double) * (size_t)d.g.nx * d.g.ny, hipMemcpyDeviceToHost,
                           stream_of(a)));
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    return MHD_OK;
}


static int ensure_tangent(mhd_adjoint *a)
{
    if (a->tpool) return MHD_OK;
    size_t total = 0;
    for (const AdjDesc &d : a->desc) total += (size_t)(5 * 8 + 2) * d.g.plane;
    AHIP(a, hipMalloc(&a->tpool, sizeof(double) * total));
    AHIP(a, hipMemsetAsync(a->tpool, 0, sizeof(double) * total, stream_of(a)));
    size_t off = 0;
    for (AdjDesc &d : a->desc) {
        const size_t p8 = (size_t)8 * d.g.plane;
        d.du0 = a->tpool + off; off += p8;
        d.du = a->tpool + off; off += p8;
        d.dus = a->tpool + off; off += p8;
        d.dF = a->tpool + off; off += p8;
        d.dG = a->tpool + off; off += p8;
        d.dE = a->tpool + off; off += d.g.plane;
        d.deta = a->tpool + off; off += d.g.plane;
    }
    return upload_desc(a);
}

int mhd_adjoint_set_direction(mhd_adjoint *a, int run, const double *host_dU0, const double *host_deta)
{
    int rc = check_run(a, run);
    if (rc) return rc;
    if ((rc = ensure_tangent(a))) return rc;
    const AdjDesc &d = a->desc[run];
    const size_t cells = (size_t)d.g.nx * d.g.ny;
    if (host_dU0) {
        AHIP(a, hipMemcpyAsync(a->fwd->aos, host_dU0, sizeof(double) * 8 * cells, hipMemcpyHostToDevice, stream_of(a)));
        AHIP(a, mhd::launch_unpack(a->fwd->aos, d.du0, d.g, stream_of(a)));
    } else {
        AHIP(a, hipMemsetAsync(d.du0, 0, sizeof(double) * 8 * d.g.plane, stream_of(a)));
    }
    if (host_deta) {
        AHIP(a, hipMemcpyAsync(a->fwd->aos, host_deta, sizeof(double) * cells, hipMemcpyHostToDevice, stream_of(a)));
        AHIP(a, mhd::launch_scatter_plane(a->fwd->aos, d.deta, d.g, stream_of(a)));
    } else {
        AHIP(a, hipMemsetAsync(d.deta, 0, sizeof(double) * d.g.plane, stream_of(a)));
    }
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    return MHD_OK;
}

// The tangent-linear rollout (replay false) or the decision replay of a
// perturbed rollout (replay true): the same linear cell, CT and boundary
// kernels around the tangent or the replayed face fluxes.
static int run_tlm(mhd_adjoint *a, int nsteps, bool replay);

int mhd_adjoint_tangent(mhd_adjoint *a, int nsteps) { return run_tlm(a, nsteps, false); }

int mhd_adjoint_replay(mhd_adjoint *a, int nsteps) { return run_tlm(a, nsteps, true); }

static int run_tlm(mhd_adjoint *a, int nsteps, bool replay)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    if (nsteps < 1 || nsteps > per_run_steps(a)) return afail(a, MHD_ERR_STATE, "needs a recorded or set dt per step");
    int rc;
    if ((rc = ensure_tangent(a))) return rc;
    mhd_batch *b = a->fwd;
    hipStream_t s = stream_of(a);
    const dim3 tg = tiles_grid(a), blk(256);
    const dim3 gx((b->max_ny + 255) / 256, a->nruns), gy((b->max_nx + 255) / 256, a->nruns);
    if ((rc = copy_all(a, kU0, 0, kU, 0))) return rc;
    if ((rc = mhd_batch_apply_boundary(b))) return afail(a, rc, b->err);
    hipLaunchKernelGGL(tlm_init_kernel, tg, blk, 0, s, a->d_desc, a->tiles);
    hipLaunchKernelGGL(bc_tlm_x_kernel, gx, blk, 0, s, a->d_desc, 2);
    hipLaunchKernelGGL(bc_tlm_y_kernel, gy, blk, 0, s, a->d_desc, 2);
    if ((rc = set_replay(a, true, false, 0))) return rc;
    for (int st = 0; st < nsteps; ++st) {
        // tangents of stage 1 from (U, dU), primal U* = BC(stage 1 of U)
        if (replay) {
            hipLaunchKernelGGL(face_rep_kernel<0>, tg, blk, 0, s, a->d_desc, a->tiles, 1);
            hipLaunchKernelGGL(face_rep_kernel<1>, tg, blk, 0, s, a->d_desc, a->tiles, 1);
        } else {
            hipLaunchKernelGGL(face_tlm_kernel<0>, tg, blk, 0, s, a->d_desc, a->tiles, 1);
            hipLaunchKernelGGL(face_tlm_kernel<1>, tg, blk, 0, s, a->d_desc, a->tiles, 1);
        }
        hipLaunchKernelGGL(emf_tlm_kernel, tg, blk, 0, s, a->d_desc, a->tiles);
        hipLaunchKernelGGL(stage_tlm_kernel, tg, blk, 0, s, a->d_desc, a->tiles, 1, st);
        hipLaunchKernelGGL(bc_tlm_x_kernel, gx, blk, 0, s, a->d_desc, 1);
        hipLaunchKernelGGL(bc_tlm_y_kernel, gy, blk, 0, s, a->d_desc, 1);
        AHIP(a, mhd::launch_stage_batch(1, b->exact, b->d_desc, b->d_item_run, b->total_items, b->xcds,
                                        b->items_per_xcd, b->blocks_per_xcd, b->cur, s, b->d_tail_run, b->tail_items));
        AHIP(a, mhd::launch_bc_batch(b->exact, b->d_desc, b->nruns, b->max_nx, b->max_ny, true, false, 0, s));
        // tangents of stage 2 from (U*, dU*) (B keeps the stage-1 CT value)
        if (replay) {
            hipLaunchKernelGGL(face_rep_kernel<0>, tg, blk, 0, s, a->d_desc, a->tiles, 2);
            hipLaunchKernelGGL(face_rep_kernel<1>, tg, blk, 0, s, a->d_desc, a->tiles, 2);
        } else {
            hipLaunchKernelGGL(face_tlm_kernel<0>, tg, blk, 0, s, a->d_desc, a->tiles, 2);
            hipLaunchKernelGGL(face_tlm_kernel<1>, tg, blk, 0, s, a->d_desc, a->tiles, 2);
        }
        hipLaunchKernelGGL(stage_tlm_kernel, tg, blk, 0, s, a->d_desc, a->tiles, 2, st);
        hipLaunchKernelGGL(bc_tlm_x_kernel, gx, blk, 0, s, a->d_desc, 2);
        hipLaunchKernelGGL(bc_tlm_y_kernel, gy, blk, 0, s, a->d_desc, 2);
        AHIP(a, hipGetLastError());
        // primal stage 2 and its boundary: the state the next step starts from
        AHIP(a, mhd::launch_stage_batch(2, b->exact, b->d_desc, b->d_item_run, b->total_items, b->xcds,
                                        b->items_per_xcd, b->blocks_per_xcd, b->cur, s, b->d_tail_run, b->tail_items));
        AHIP(a, mhd::launch_bc_batch(b->exact, b->d_desc, b->nruns, b->max_nx, b->max_ny, false, true, b->cur ^ 1, s));
        b->cur ^= 1;
    }
    if ((rc = set_replay(a, false, false, 0))) return rc;
    AHIP(a, hipStreamSynchronize(s));
    return MHD_OK;
}

int mhd_adjoint_get_tangent(mhd_adjoint *a, int run, double *host_dUN)
{
    int rc = check_run(a, run);
    if (rc) return rc;
    if (!host_dUN || !a->tpool) return afail(a, MHD_ERR_STATE, "run mhd_adjoint_tangent first");
    const AdjDesc &d = a->desc[run];
    AHIP(a, mhd::launch_pack(d.du, a->fwd->aos, d.g, stream_of(a)));
    AHIP(a, hipMemcpyAsync(host_dUN, a->fwd->aos, sizeof(double) * 8 * (size_t)d.g.nx * d.g.ny,
                           hipMemcpyDeviceToHost, stream_of(a)));
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    return MHD_OK;
}


// ---- whole-batch transfers: one host copy and one synchronisation for
// every run (states concatenated in run order, AoS as mhd_set_state)
static int ensure_stage_all(mhd_adjoint *a)
{
    if (a->stage_all) return MHD_OK;
    a->aos_off.assign(a->nruns + 1, 0);
    for (int r = 0; r < a->nruns; ++r)
        a->aos_off[r + 1] = a->aos_off[r] + (size_t)8 * a->desc[r].g.nx * a->desc[r].g.ny;
    AHIP(a, hipMalloc(&a->stage_all, sizeof(double) * a->aos_off[a->nruns]));
    return MHD_OK;
}

// device: the buffer is device memory (for instance a torch tensor on the
// same GPU): no host copy; the library's stream waits for nothing else, so
// the caller synchronises its own stream before the call.
static int upload_all(mhd_adjoint *a, const double *src, int which, bool device)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    if (!src) return afail(a, MHD_ERR_INVALID, "NULL buffer");
    int rc;
    if ((rc = ensure_stage_all(a))) return rc;
    const double *stage = src;
    if (!device) {
        AHIP(a, hipMemcpyAsync(a->stage_all, src, sizeof(double) * a->aos_off[a->nruns], hipMemcpyHostToDevice,
                               stream_of(a)));
        stage = a->stage_all;
    }
    hipLaunchKernelGGL(aos_all_kernel<false>, tiles_grid(a), dim3(256), 0, stream_of(a), a->d_desc, a->tiles,
                       which, const_cast<double *>(stage));
    AHIP(a, hipGetLastError());
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    if (which == kU0) invalidate_rollout(a);   // new initial states: the recorded rollout is stale
    return MHD_OK;
}

static int download_all(mhd_adjoint *a, double *dst, int which, bool device)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    if (!dst) return afail(a, MHD_ERR_INVALID, "NULL buffer");
    int rc;
    if ((rc = ensure_stage_all(a))) return rc;
    double *stage = device ? dst : a->stage_all;
    hipLaunchKernelGGL(aos_all_kernel<true>, tiles_grid(a), dim3(256), 0, stream_of(a), a->d_desc, a->tiles,
                       which, stage);
    AHIP(a, hipGetLastError());
    if (!device)
        AHIP(a, hipMemcpyAsync(dst, a->stage_all, sizeof(double) * a->aos_off[a->nruns], hipMemcpyDeviceToHost,
                               stream_of(a)));
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    return MHD_OK;
}

int mhd_adjoint_set_states(mhd_adjoint *a, const double *host_all) { return upload_all(a, host_all, kU0, false); }
int mhd_adjoint_set_seeds(mhd_adjoint *a, const double *host_all) { return upload_all(a, host_all, kLam, false); }
int mhd_adjoint_get_states(mhd_adjoint *a, double *host_all) { return download_all(a, host_all, kU, false); }
int mhd_adjoint_get_grad_states(mhd_adjoint *a, double *host_all) { return download_all(a, host_all, kLam, false); }
int mhd_adjoint_set_states_device(mhd_adjoint *a, const double *dev_all) { return upload_all(a, dev_all, kU0, true); }
int mhd_adjoint_set_seeds_device(mhd_adjoint *a, const double *dev_all) { return upload_all(a, dev_all, kLam, true); }
int mhd_adjoint_get_states_device(mhd_adjoint *a, double *dev_all) { return download_all(a, dev_all, kU, true); }
int mhd_adjoint_get_grad_states_device(mhd_adjoint *a, double *dev_all)
{
    return download_all(a, dev_all, kLam, true);
}

// every run's dL/deta (nx*ny per run, concatenated), host or device
static int grad_eta_all(mhd_adjoint *a, double *dst, bool device)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    if (!dst) return afail(a, MHD_ERR_INVALID, "NULL buffer");
    int rc;
    if ((rc = ensure_stage_all(a))) return rc;
    double *stage = device ? dst : a->stage_all;
    const size_t off = a->aos_off[a->nruns] / 8;
    hipLaunchKernelGGL(geta_all_kernel, tiles_grid(a), dim3(256), 0, stream_of(a), a->d_desc, a->tiles,
                       stage);
    AHIP(a, hipGetLastError());
    if (!device)
        AHIP(a, hipMemcpyAsync(dst, a->stage_all, sizeof(double) * off, hipMemcpyDeviceToHost, stream_of(a)));
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    return MHD_OK;
}

int mhd_adjoint_set_eta_fields_device(mhd_adjoint *a, const double *dev_all)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    if (!dev_all) return afail(a, MHD_ERR_INVALID, "NULL buffer");
    mhd_batch *b = a->fwd;
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    size_t off = 0;
    for (int r = 0; r < a->nruns; ++r) {
        mhd::RunDesc &d = b->desc[r];
        if (!b->eta[r]) AHIP(a, hipMalloc(&b->eta[r], sizeof(double) * (size_t)d.g.plane));
        AHIP(a, mhd::launch_scatter_plane(dev_all + off, b->eta[r], d.g, stream_of(a)));
        d.eta = b->eta[r];
        a->desc[r].eta = d.eta;
        off += (size_t)d.g.nx * d.g.ny;
    }
    AHIP(a, hipMemcpyAsync(b->d_desc, b->desc.data(), sizeof(mhd::RunDesc) * (size_t)a->nruns, hipMemcpyHostToDevice,
                           stream_of(a)));
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    invalidate_rollout(a);   // the recorded rollout used the previous fields
    return upload_desc(a);
}

int mhd_adjoint_get_grad_etas(mhd_adjoint *a, double *host_all) { return grad_eta_all(a, host_all, false); }
int mhd_adjoint_get_grad_etas_device(mhd_adjoint *a, double *dev_all) { return grad_eta_all(a, dev_all, true); }

int mhd_adjoint_synchronize(mhd_adjoint *a)
{
    if (!a) return afail(nullptr, MHD_ERR_INVALID, "NULL adjoint");
    AHIP(a, hipStreamSynchronize(stream_of(a)));
    return MHD_OK;
}

}  // extern "C"
