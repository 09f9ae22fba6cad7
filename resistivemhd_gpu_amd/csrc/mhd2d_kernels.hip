// mhd2d_kernels.hip -- gfx950 kernels of the 2D resistive-MHD RK2 step.
//
// State layout in HBM: 8 FP64 planes (rho, rho*u, rho*v, rho*w, Bx, By, Bz,
// e), each nx rows (x, slow) of `pitch` doubles (y, fast), ghosts included.
//
// One RK stage is one launch of rk_stage_kernel: a wavefront owns a strip of
// kStripRows interior rows (y) and sweeps along x over a segment of columns.
// Lane l holds row j0-2+l.  Everything along x (MUSCL window, the x-face
// flux, EMF and flux of the previous column) stays in registers; y
// neighbours are exchanged with cross-lane shuffles.  Per column the wave
// reads one 64-row slice of the 8 planes and the resistivity, staged in LDS
// one column ahead by direct-to-LDS loads, computes the x-face HLLD +
// resistive flux, the y-face flux of the previous column, the corner EMF,
// and writes the updated cell two columns behind the load front.  The
// update's own operands come from LDS too: the cell's stage input (and in
// stage 2 U^n) is staged while the y-face flux is computed, and the
// previous column's face fluxes wait in per-lane LDS slots instead of
// registers held across both HLLD solves.  The next column's loads are
// issued after the x-face flux: a spill reload waits for every older vector
// memory load, so loads issued at the top of the column would be waited for
// a few hundred instructions later.
//
// Reference (paths relative to its root,
// P = ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_periodicXY,
// S = ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_symmetricXY,
// EX = ResistiveMHD_singleGPU_thrust/examples2D): oneStepRK2
// P/resistiveMHD_2D.cu:133-195.
#include <hip/hip_runtime.h>
#include "mhd2d_device.h"
#include "mhd2d_kernels.h"

namespace mhd {

#define PL(ptr, v) ((ptr) + (size_t)(v) * g.plane)

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;

// Rows of one staged column: global rows j0-3 .. j0+62 of a strip (66 rows,
// 33 16-byte row pairs; j0 - 3 = 60 sy is even, so every pair is 16-byte
// aligned).  Lane l's row j0-2+l sits at index l+1 and the row below it
// (the By face under the cell) at index l.
constexpr int kStageRows = 66;
constexpr int kStagePlanes = 9;   // 8 state planes and the resistivity plane

// x-frame primitive variables of one cell (P/calculate_half_Q.cu:4-38 with
// the centred B of shiftUToCenterForCT, P/resistiveMHD_2D.cu:313-368).
struct Cell {
    double rho, u, v, w, bxc, byc, bz, px, py;
};

template <bool EXACT>
__device__ __forceinline__ Cell make_cell(double rho, double mx, double my, double mz, double bxc,
                                          double byc, double bz, double e, const Phys &ph)
{
    Cell c;
    c.rho = rho;
    c.bxc = bxc;
    c.byc = byc;
    c.bz = bz;
    if (EXACT) {
        const SharedDen dr = div_prepare(mx, rho);
        c.u = div_by(mx, rho, dr);
        c.v = div_by(my, rho, dr);
        c.w = div_by(mz, rho, dr);
        // x sweep: GetBasicParamterFunctor on the centred state
        c.px = (ph.gamma - 1.0)
             * (e - 0.5 * (rho * (c.u * c.u + c.v * c.v + c.w * c.w))
             - 0.5 * (bxc * bxc + byc * byc + bz * bz));
        // y sweep: the same functor on the rotated state (P/hlld.cu:134-152),
        // whose sums run in the rotated order
        c.py = (ph.gamma - 1.0)
             * (e - 0.5 * (rho * (c.v * c.v + c.w * c.w + c.u * c.u))
             - 0.5 * (byc * byc + bz * bz + bxc * bxc));
    } else {
#pragma clang fp contract(fast)
        const double ir = rcp_fast(rho);
        c.u = mx * ir;
        c.v = my * ir;
        c.w = mz * ir;
        c.px = (ph.gamma - 1.0) * (e - 0.5 * rho * (c.u * c.u + c.v * c.v + c.w * c.w)
                                     - 0.5 * (bxc * bxc + byc * byc + bz * bz));
        c.py = c.px;
    }
    return c;
}

// Resident wavefronts per SIMD the register allocation targets (the exact
// arithmetic needs more live values than the fast one).
#ifndef MHD_WAVES_PER_EU_EXACT
#define MHD_WAVES_PER_EU_EXACT 2
#endif
#ifndef MHD_WAVES_PER_EU_FAST
#define MHD_WAVES_PER_EU_FAST 2
#endif

// SCHEME 0: the single-GPU library's step (P/resistiveMHD_2D.cu:133-195).
// SCHEME 1: the MPI library's step on one subdomain
// (ResistiveMHD_multiGPU_cuda/ResistiveMHD2D_GPU/resistiveMHD2D.cu:147-221,
// "M/" below): face-centred resistive flux (M/flux_solver.cu:73-189), B
// advanced by flux divergence from the face-smoothed field in both stages,
// no CT in the stages; the face fluxes the CT-contact EMF needs are stored
// in `aux` (planes F0, F5 at x-faces and G0, G4 at y-faces, owned by the
// cell on their low side) and the stage-2 state goes to uout != uold.
// SCHEME 2: SCHEME 1 with the MPI library's getEta sampled per index, etaf
// the x-face plane ((i + 1/2) dx, j dy), etaf + plane the y-face plane
// (i dx, (j + 1/2) dy) (M/flux_solver.cu:84,168); SCHEME 1 uses the
// uniform ph.eta (a separate instantiation, so the planes' loads and
// registers cost it nothing).
//
// stage_item is one work item (strip sy, segment sx) of one run; the
// single-run kernel and the batched work-list kernel both call it.
//
// TAIL: item of the packed last strip (tail_g = its rows + 4).  Lane
// q * tail_g + r holds row j0 - 2 + r of segment item * (64 / tail_g) + q:
// the loop runs over group 0's columns c, and group q's lanes work on
// column c + dcol, dcol = q * seg (loads clamped to the grid, stores only
// within the group's own segment).  Each group keeps two halo rows on each
// side, so the y-neighbour shifts that cross a group boundary only reach
// rows whose results are discarded, as at the edges of a full strip.
template <int STAGE, bool EXACT, int SCHEME, bool TAIL = false>
__device__ __forceinline__ void stage_item(int item, const double *__restrict__ uin, double *uout, const double *uold,
                                           const double *__restrict__ etaf, double *__restrict__ side,
                                           double *__restrict__ aux, Ctrl *__restrict__ ctrl, const Grid &g,
                                           const Phys &ph, int cur)
{
    constexpr bool MULTI = SCHEME >= 1;
    constexpr bool ETAM = SCHEME == 2;        // the MPI library's getEta planes
    constexpr bool ETA_PLANE = !MULTI || ETAM;   // etaf may be set (the uniform multi scheme never reads it)
    static_assert(!(TAIL && MULTI), "packed strips belong to the single-GPU scheme");
    const int lane = threadIdx.x;
    int sy, sx, row = lane, dcol = 0;
    bool group_ok = true;
    if (TAIL) {
        const int per = 64 / g.tail_g, q = lane / g.tail_g;
        sy = g.nstrips;
        sx = item * per;
        group_ok = q < per && sx + q < g.nsegs;
        row = q < per ? lane - q * g.tail_g : g.tail_g - 1;
        dcol = group_ok ? q * g.seg : 0;
    } else {
        sy = item % g.nstrips;
        sx = item / g.nstrips;
    }
    const int j0 = kGhost + sy * kStripRows;
    const int i0 = kGhost + sx * g.seg;
    // loop end (group 0's segment, in full when TAIL) and, in loop columns,
    // the end of this lane's own segment
    const int iend = TAIL ? i0 + g.seg : min(i0 + g.seg, g.nx - kGhost);
    const int send = TAIL ? (group_ok ? min(i0 + dcol + g.seg, g.nx - kGhost) - dcol : 0) : iend;
    const int jend = min(j0 + kStripRows, g.ny - kGhost);
    const int jr = j0 - 2 + row;               // row held by this lane
    const int jl = min(jr, g.ny - 1);          // clamped row for loads (rows past ny-1 are unused)
    const bool upd_row = (jr >= j0) && (jr < jend);
    const bool stale_row = (sy == 0) && (jr == kGhost - 1);
    const bool leader = lane == 0 && item == 0 && (!TAIL || g.nstrips == 0);

    // dt: ResistiveMHD2D::calculateDt (P/resistiveMHD_2D.cu:277-291), from
    // the per-cell minimum reduced at the end of the previous step.
    double dt;
    if (STAGE == 1) {
        const double m = ctrl->first_nan[cur] ? __builtin_nan("") : __longlong_as_double((long long)ctrl->acc[cur]);
        if (MULTI) {
            dt = m * ph.cfl;   // M/resistiveMHD2D.cu:541: no diffusion limit
        } else {
            const double lim = fmin(ph.dx * ph.dx / ph.eta, ph.dy * ph.dy / ph.eta);
            dt = fmin(m, lim) * ph.cfl;
        }
        const long long seq = ctrl->steps - ctrl->seq0;
        if (ctrl->dt_in) dt = ctrl->dt_in[seq];
        if (leader) {
            if (ctrl->dt_out) ctrl->dt_out[seq] = dt;
            ctrl->dt = dt;
            ctrl->acc[cur ^ 1] = kInfBits;
            ctrl->first_nan[cur ^ 1] = 0;
            ctrl->crashed = 0;
        }
    } else {
        dt = ctrl->dt;
        if (leader) {
            ctrl->time += dt;
            ctrl->steps += 1;
        }
    }
    const double dtdx = dt / ph.dx;
    const double dtdy = dt / ph.dy;
    const double two_dx = 2.0 * ph.dx, two_dy = 2.0 * ph.dy;
    const double inv_2dx = 1.0 / two_dx, inv_2dy = 1.0 / two_dy;
    const double inv_gm1 = 1.0 / (ph.gamma - 1.0);

    const double *Rh = PL(uin, 0), *Mx = PL(uin, 1), *My = PL(uin, 2), *Mz = PL(uin, 3);
    const double *Bx = PL(uin, 4), *By = PL(uin, 5), *Bz = PL(uin, 6), *En = PL(uin, 7);

    // Full strips stage each column in LDS one column ahead with direct
    // loads (global_load_lds), so a column's HBM latency overlaps the
    // previous column's arithmetic instead of stalling the wavefront.  One
    // slot serves: a column's slot is read only at the top of its iteration,
    // and the next column's loads are issued after those reads are done.
    // LDS per wavefront: 15,632 B in stage 1 and 18,800 B in stage 2, so
    // eight wavefronts (two per SIMD, the VGPR limit) fit a CU's 160 KiB.
    constexpr bool DMA = !TAIL;
    __shared__ __attribute__((aligned(16))) double slot[ETAM ? kStagePlanes + 1 : kStagePlanes][kStageRows];
    const int r0 = j0 - 3;
    const int rpair = min(r0 + 2 * lane, (g.ny - 2) & ~1);
    auto stage_col = [&](int col) {
        if (lane < kStageRows / 2) {
            // opaque, so the lane offset is not folded into one loop-invariant
            // 64-bit pointer per plane (VGPR pairs the loop would have to keep)
            size_t off = (size_t)col * g.pitch + rpair;
            asm volatile("" : "+v"(off));
#pragma unroll
            for (int v = 0; v < 8; ++v)
                __builtin_amdgcn_global_load_lds((glb_void_t *)(uin + (size_t)v * g.plane + off),
                                                 (lds_void_t *)slot[v], 16, 0, 0);
            if (ETA_PLANE && etaf)
                __builtin_amdgcn_global_load_lds((glb_void_t *)(etaf + off), (lds_void_t *)slot[8], 16, 0, 0);
            if (ETAM)
                __builtin_amdgcn_global_load_lds((glb_void_t *)(etaf + g.plane + off), (lds_void_t *)slot[ETAM ? 9 : 8],
                                                 16, 0, 0);
        }
    };
    // the stage input (and in stage 2 U^n) of the cell being updated,
    // staged before the y-face flux of the same column (not in stage 2 of
    // the multi-GPU scheme with getEta planes, whose extra staging plane
    // would take its LDS past eight wavefronts per CU)
    constexpr bool UPD_DMA = DMA && !(ETAM && STAGE == 2);
    __shared__ __attribute__((aligned(16))) double upd[UPD_DMA ? (STAGE == 2 ? 14 : 8) : 1][kStageRows];
    // the previous column's face fluxes (the components the update reads:
    // six of F and seven of G here, all eight of each in the MPI library's
    // step, whose B advances by flux divergence), one slot per lane, instead
    // of registers held across both HLLD solves
    constexpr bool FLDS = DMA;
    constexpr bool GLDS = DMA;
    __shared__ double fb[FLDS ? (MULTI ? 8 : 6) : 1][64];
    __shared__ double gb[GLDS ? (MULTI ? 8 : 7) : 1][64];
    auto stage_upd = [&](int col) {
        if (lane < kStageRows / 2) {
            // opaque, so the lane offset is not folded into one loop-invariant
            // 64-bit pointer per plane (VGPR pairs the loop would have to keep)
            size_t off = (size_t)col * g.pitch + rpair;
            asm volatile("" : "+v"(off));
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                // the multi-GPU scheme's stage 2 reads the smoothed face B
                // instead of the stage input's: planes 4 and 5 stage U^n's Bx
                // of the next column and By of this one for its U^n smoothing
                const double *src = uin + (size_t)v * g.plane + off;
                if (MULTI && STAGE == 2 && v == 4) src = uold + (size_t)4 * g.plane + off + g.pitch;
                if (MULTI && STAGE == 2 && v == 5) src = uold + (size_t)5 * g.plane + off;
                __builtin_amdgcn_global_load_lds((glb_void_t *)src, (lds_void_t *)upd[v], 16, 0, 0);
            }
#pragma unroll
            for (int k = 0; k < (STAGE == 2 ? 6 : 0); ++k) {
                const int v = k < 4 ? k : k + 2;
                __builtin_amdgcn_global_load_lds((glb_void_t *)(uold + (size_t)v * g.plane + off),
                                                 (lds_void_t *)upd[8 + k], 16, 0, 0);
            }
        }
    };
    if (DMA) stage_col(i0 - 2);

    double bx_prev = Bx[(size_t)(i0 - 3 + dcol) * g.pitch + jl];
    // multi-GPU scheme, stage 2 with the update staged: U^n's Bx of the two
    // columns before the next update's staged one (s-1 and s)
    constexpr bool OLDB_LDS = MULTI && STAGE == 2 && UPD_DMA;
    double obx_m = 0.0, obx_s = 0.0;
    if (OLDB_LDS) {
        obx_m = PL(uold, 4)[(size_t)(i0 - 1) * g.pitch + jl];
        obx_s = PL(uold, 4)[(size_t)i0 * g.pitch + jl];
    }
    Cell A{}, B{};                                   // cells c-2 and c-1
    Face7 Lx{};                                      // left state of face c-2+1/2
    Flux8 Fprev{}, Gprev{};                          // F(c-3+1/2), G(c-2)
    double Eprev = 0.0;                              // E at corner (c-3+1/2, j+1/2)
    double rJY = 0.0, rJZ = 0.0, rJYBZ = 0.0, rJZBY = 0.0;   // resistive terms of cell c-2
    double rTA = 0.0, rTB = 0.0;                     // M/: centred-difference terms of cell c-2
    double dtloc = __longlong_as_double((long long)kInfBits);
    int nanflag = 0;

    double eta_prev = 0.0;   // resistivity of cell c-1 (SCHEME 2: of x-face c-1 + 1/2)
    double eta_prev2 = 0.0;  // SCHEME 2: x-face c-2 + 1/2
    double etay_prev = 0.0;  // SCHEME 2: y-face (c-1, j + 1/2)

    for (int c = i0 - 2; c <= iend + 1; ++c) {
        double rho, mx, my, mz, bx, by, bz, e, bym, eta_c, etay_c = ph.eta;
        if (DMA) {
            // s_waitcnt vmcnt(0): this column's staging loads have landed
            __builtin_amdgcn_s_waitcnt(0x0F70);
            rho = slot[0][lane + 1]; mx = slot[1][lane + 1]; my = slot[2][lane + 1]; mz = slot[3][lane + 1];
            bx = slot[4][lane + 1]; by = slot[5][lane + 1]; bz = slot[6][lane + 1]; e = slot[7][lane + 1];
            bym = slot[5][lane];
            eta_c = ETA_PLANE && etaf ? slot[8][lane + 1] : ph.eta;
            if (ETAM) etay_c = slot[ETAM ? 9 : 8][lane + 1];
            // warm-up columns stage the next column at once; the others after
            // their x-face flux
            if (c < i0) {
                // s_waitcnt lgkmcnt(0): the slot's reads are done before the
                // next column's loads may land in it
                __builtin_amdgcn_s_waitcnt(0xC07F);
                stage_col(c + 1);
            }
        } else {
            const size_t o = (size_t)(TAIL ? min(c + dcol, g.nx - 1) : c) * g.pitch + jl;
            rho = Rh[o]; mx = Mx[o]; my = My[o]; mz = Mz[o];
            bx = Bx[o]; by = By[o]; bz = Bz[o]; e = En[o];
            bym = By[o - 1];
            eta_c = etaf ? etaf[o] : ph.eta;
        }
        const double bxc = 0.5 * (bx + bx_prev);
        bx_prev = bx;
        const double byc = 0.5 * (by + bym);
        const Cell C = make_cell<EXACT>(rho, mx, my, mz, bxc, byc, bz, e, ph);

        if (c >= i0) {
            // x-slopes of cell c-1 (P/muscl.cu:6-29,67-90): right state of face
            // c-2+1/2 and left state of face c-1+1/2.
            const double sr = minmod(B.rho - A.rho, C.rho - B.rho, ph.eps);
            const double su = minmod(B.u - A.u, C.u - B.u, ph.eps);
            const double sv = minmod(B.v - A.v, C.v - B.v, ph.eps);
            const double sw = minmod(B.w - A.w, C.w - B.w, ph.eps);
            const double sby = minmod(B.byc - A.byc, C.byc - B.byc, ph.eps);
            const double sbz = minmod(B.bz - A.bz, C.bz - B.bz, ph.eps);
            const double sp = minmod(B.px - A.px, C.px - B.px, ph.eps);

            // resistive terms of cell c-1 (EX/bhattacharjee/main_bhattacharjee.cu:142-241)
            const double bxc_up = lane_next(B.bxc), bxc_dn = lane_prev(B.bxc);
            const double bz_up = lane_next(B.bz), bz_dn = lane_prev(B.bz);
            const double eta = eta_prev;
            double jX, jY, jZ;
            if (EXACT) {
                jY = -(C.bz - A.bz) / two_dx;
                jZ = (C.byc - A.byc) / two_dx - (bxc_up - bxc_dn) / two_dy;
                jX = (bz_up - bz_dn) / two_dy;
            } else {
                jY = -(C.bz - A.bz) * inv_2dx;
                jZ = (C.byc - A.byc) * inv_2dx - (bxc_up - bxc_dn) * inv_2dy;
                jX = (bz_up - bz_dn) * inv_2dy;
            }
            const double eJX = eta * jX, eJY = eta * jY, eJZ = eta * jZ;
            const double eJYBZ = eJY * B.bz, eJZBY = eJZ * B.byc;
            const double eJZBX = eJZ * B.bxc, eJXBZ = eJX * B.bz;
            // M/flux_solver.cu:88-92,173-177: the face current averages the two
            // cells' centred differences, ((a_i - b_i) + a_i+1) - b_i+1
            double tA = 0.0, tB = 0.0;
            if (MULTI) {
                if (EXACT) {
                    tA = (C.byc - A.byc) / two_dx;
                    tB = (bxc_up - bxc_dn) / two_dy;
                } else {
                    tA = (C.byc - A.byc) * inv_2dx;
                    tB = (bxc_up - bxc_dn) * inv_2dy;
                }
            }

            Flux8 F{};
            if (c >= i0 + 1) {
                Face7 Rx;
                Rx.rho = B.rho - 0.5 * sr;
                Rx.u = B.u - 0.5 * su;
                Rx.v = B.v - 0.5 * sv;
                Rx.w = B.w - 0.5 * sw;
                Rx.by = B.byc - 0.5 * sby;
                Rx.bz = B.bz - 0.5 * sbz;
                Rx.p = B.px - 0.5 * sp;
                F = hlld<EXACT>(Lx, Rx, 0.5 * (A.bxc + B.bxc), ph, inv_gm1);
                if (MULTI) {
                    // face (c-2+1/2): cells A = c-2 and B = c-1 (M/flux_solver.cu:73-104)
                    const double fjY = EXACT ? -(B.bz - A.bz) / ph.dx : -(B.bz - A.bz) * (2.0 * inv_2dx);
                    const double fjZ = 0.5 * (((rTA - rTB) + tA) - tB);
                    const double feta = ETAM ? eta_prev2 : ph.eta;
                    const double fJY = feta * fjY, fJZ = feta * fjZ;
                    const double fJYBZ = fJY * 0.5 * (A.bz + B.bz);
                    const double fJZBY = fJZ * 0.5 * (A.byc + B.byc);
                    F.f5 -= fJZ;
                    F.f6 += fJY;
                    F.f7 += fJYBZ - fJZBY;
                } else {
                    F.f5 -= 0.5 * (rJZ + eJZ);
                    F.f6 += 0.5 * (rJY + eJY);
                    F.f7 += 0.5 * (rJYBZ + eJYBZ) - 0.5 * (rJZBY + eJZBY);
                }
            }
            Lx.rho = B.rho + 0.5 * sr;
            Lx.u = B.u + 0.5 * su;
            Lx.v = B.v + 0.5 * sv;
            Lx.w = B.w + 0.5 * sw;
            Lx.by = B.byc + 0.5 * sby;
            Lx.bz = B.bz + 0.5 * sbz;
            Lx.p = B.px + 0.5 * sp;
            rJY = eJY; rJZ = eJZ; rJYBZ = eJYBZ; rJZBY = eJZBY;
            rTA = tA; rTB = tB;

            if (DMA && c + 1 <= iend + 1) {
                __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0), as above
                stage_col(c + 1);
            }
            if (UPD_DMA && c >= i0 + 2 && c - 2 < send) stage_upd(c - 2);

            // y-face flux of column c-1 at faces j+1/2 (HLLD::calculateFluxG,
            // P/hlld.cu:209-233): rotated frame (u,v,w)->(v,w,u),
            // (Bx,By,Bz)->(By,Bz,Bx); neighbours along y come from other lanes.
            Flux8 G;
            {
                const double q0 = B.rho, q1 = B.v, q2 = B.w, q3 = B.u, q4 = B.bz, q5 = B.bxc, q6 = B.py;
                const double m0 = lane_prev(q0), m1 = lane_prev(q1), m2 = lane_prev(q2);
                const double m3 = lane_prev(q3), m4 = lane_prev(q4), m5 = lane_prev(q5);
                const double m6 = lane_prev(q6);
                const double p0 = lane_next(q0), p1 = lane_next(q1), p2 = lane_next(q2);
                const double p3 = lane_next(q3), p4 = lane_next(q4), p5 = lane_next(q5);
                const double p6 = lane_next(q6);
                const double t0 = minmod(q0 - m0, p0 - q0, ph.eps), t1 = minmod(q1 - m1, p1 - q1, ph.eps);
                const double t2 = minmod(q2 - m2, p2 - q2, ph.eps), t3 = minmod(q3 - m3, p3 - q3, ph.eps);
                const double t4 = minmod(q4 - m4, p4 - q4, ph.eps), t5 = minmod(q5 - m5, p5 - q5, ph.eps);
                const double t6 = minmod(q6 - m6, p6 - q6, ph.eps);
                Face7 Lg, Rg;
                Lg.rho = q0 + 0.5 * t0; Lg.u = q1 + 0.5 * t1; Lg.v = q2 + 0.5 * t2; Lg.w = q3 + 0.5 * t3;
                Lg.by = q4 + 0.5 * t4; Lg.bz = q5 + 0.5 * t5; Lg.p = q6 + 0.5 * t6;
                Rg.rho = lane_next(q0 - 0.5 * t0);
                Rg.u = lane_next(q1 - 0.5 * t1);
                Rg.v = lane_next(q2 - 0.5 * t2);
                Rg.w = lane_next(q3 - 0.5 * t3);
                Rg.by = lane_next(q4 - 0.5 * t4);
                Rg.bz = lane_next(q5 - 0.5 * t5);
                Rg.p = lane_next(q6 - 0.5 * t6);
                const double bn = 0.5 * (B.byc + lane_next(B.byc));
                const Flux8 Gr = hlld<EXACT>(Lg, Rg, bn, ph, inv_gm1);
                // rotate back (P/hlld.cu:170-194)
                G.f0 = Gr.f0; G.f1 = Gr.f3; G.f2 = Gr.f1; G.f3 = Gr.f2;
                G.f4 = Gr.f6; G.f5 = Gr.f4; G.f6 = Gr.f5; G.f7 = Gr.f7;
                if (MULTI) {
                    // face (c-1, j+1/2): cells j and j+1 (M/flux_solver.cu:158-189)
                    const double tA1 = lane_next(tA), tB1 = lane_next(tB);
                    const double gjX = EXACT ? (bz_up - B.bz) / ph.dy : (bz_up - B.bz) * (2.0 * inv_2dy);
                    const double gjZ = 0.5 * (((tA - tB) + tA1) - tB1);
                    const double geta = ETAM ? etay_prev : ph.eta;
                    const double gJX = geta * gjX, gJZ = geta * gjZ;
                    const double gJXBZ = gJX * 0.5 * (B.bz + bz_up);
                    const double gJZBX = gJZ * 0.5 * (B.bxc + bxc_up);
                    G.f4 += gJZ;
                    G.f6 -= gJX;
                    G.f7 += gJZBX - gJXBZ;
                } else {
                    const double eJZ1 = lane_next(eJZ), eJX1 = lane_next(eJX);
                    const double eJZBX1 = lane_next(eJZBX), eJXBZ1 = lane_next(eJXBZ);
                    G.f4 += 0.5 * (eJZ + eJZ1);
                    G.f6 -= 0.5 * (eJX + eJX1);
                    G.f7 += 0.5 * (eJZBX + eJZBX1) - 0.5 * (eJXBZ + eJXBZ1);
                }
            }

            // corner EMF at (c-2+1/2, j+1/2), P/ct.cu:22-41 (stage 1 only: the
            // stage-2 CT reuses stage-1 fluxes, P/resistiveMHD_2D.cu:167-168,190)
            double E = 0.0;
            if (!MULTI && STAGE == 1 && c >= i0 + 1) {
                const double f5up = lane_next(F.f5);
                const double gp4 = GLDS ? gb[4][lane] : Gprev.f4;
                E = 0.25 * (gp4 + G.f4 + (-1.0 * F.f5) + (-1.0 * f5up));
            }

            if (c >= i0 + 2) {
                const int s = c - 2;
                Flux8 Fp = Fprev, Gp = Gprev;
                if (FLDS && MULTI) {
                    Fp.f0 = fb[0][lane]; Fp.f1 = fb[1][lane]; Fp.f2 = fb[2][lane]; Fp.f3 = fb[3][lane];
                    Fp.f4 = fb[4][lane]; Fp.f5 = fb[5][lane]; Fp.f6 = fb[6][lane]; Fp.f7 = fb[7][lane];
                } else if (FLDS) {
                    Fp.f0 = fb[0][lane]; Fp.f1 = fb[1][lane]; Fp.f2 = fb[2][lane];
                    Fp.f3 = fb[3][lane]; Fp.f6 = fb[4][lane]; Fp.f7 = fb[5][lane];
                }
                if (GLDS && MULTI) {
                    Gp.f0 = gb[0][lane]; Gp.f1 = gb[1][lane]; Gp.f2 = gb[2][lane]; Gp.f3 = gb[3][lane];
                    Gp.f4 = gb[4][lane]; Gp.f5 = gb[5][lane]; Gp.f6 = gb[6][lane]; Gp.f7 = gb[7][lane];
                } else if (GLDS) {
                    Gp.f0 = gb[0][lane]; Gp.f1 = gb[1][lane]; Gp.f2 = gb[2][lane]; Gp.f3 = gb[3][lane];
                    Gp.f4 = gb[4][lane]; Gp.f6 = gb[5][lane]; Gp.f7 = gb[6][lane];
                }
                const double g0m = lane_prev(Gp.f0), g1m = lane_prev(Gp.f1);
                const double g2m = lane_prev(Gp.f2), g3m = lane_prev(Gp.f3);
                const double g6m = lane_prev(Gp.f6), g7m = lane_prev(Gp.f7);
                const double Em = lane_prev(E);
                if (MULTI) {
                    // face-smoothed B of the stage input (shift + back,
                    // M/resistiveMHD2D.cu:597-690): cells s = A and s+1 = B
                    const double g4m = lane_prev(Gp.f4), g5m = lane_prev(Gp.f5);
                    const double byc_up = lane_next(A.byc);
                    const double sbx = 0.5 * (A.bxc + B.bxc);
                    const double sby = 0.5 * (A.byc + byc_up);
                    double obx = 0.0, oby = 0.0;
                    if (STAGE == 2) {
                        // face-smoothed B of U^n at cell s
                        double bxs, bxm, bxp, bys, bym1;
                        if (OLDB_LDS) {
                            // s_waitcnt vmcnt(0): the staging loads of cell s have landed
                            __builtin_amdgcn_s_waitcnt(0x0F70);
                            const int r = lane + 1;
                            bxm = obx_m;
                            bxs = obx_s;
                            bxp = upd[4][r];
                            bys = upd[5][r];
                            bym1 = upd[5][r - 1];
                            obx_m = obx_s;
                            obx_s = bxp;
                        } else {
                            const size_t on = (size_t)s * g.pitch + jl;
                            bxs = PL(uold, 4)[on];
                            bxm = PL(uold, 4)[on - g.pitch];
                            bxp = PL(uold, 4)[on + g.pitch];
                            bys = PL(uold, 5)[on];
                            bym1 = PL(uold, 5)[on - 1];
                        }
                        const double byp = lane_next(bys);
                        const double cxs = 0.5 * (bxs + bxm), cxp = 0.5 * (bxp + bxs);
                        const double cys = 0.5 * (bys + bym1), cyp = 0.5 * (byp + bys);
                        obx = 0.5 * (cxs + cxp);
                        oby = 0.5 * (cys + cyp);
                    }
                    if (s < iend && upd_row) {
                        const size_t os = (size_t)s * g.pitch + jl;
                        double n0, n1, n2, n3, n4, n5, n6, n7;
                        // the stage input's (and U^n's) non-B components of cell s
                        double S0, S1, S2, S3, S6, S7, O0 = 0.0, O1 = 0.0, O2 = 0.0, O3 = 0.0, O6 = 0.0, O7 = 0.0;
                        if (UPD_DMA) {
                            // s_waitcnt vmcnt(0): the cell's staging loads have landed
                            __builtin_amdgcn_s_waitcnt(0x0F70);
                            const int r = lane + 1;
                            S0 = upd[0][r]; S1 = upd[1][r]; S2 = upd[2][r]; S3 = upd[3][r];
                            S6 = upd[6][r]; S7 = upd[7][r];
                            if (STAGE == 2) {
                                O0 = upd[8][r]; O1 = upd[9][r]; O2 = upd[10][r]; O3 = upd[11][r];
                                O6 = upd[12][r]; O7 = upd[13][r];
                            }
                        } else {
                            S0 = Rh[os]; S1 = Mx[os]; S2 = My[os]; S3 = Mz[os]; S6 = Bz[os]; S7 = En[os];
                            if (STAGE == 2) {
                                O0 = PL(uold, 0)[os]; O1 = PL(uold, 1)[os]; O2 = PL(uold, 2)[os];
                                O3 = PL(uold, 3)[os]; O6 = PL(uold, 6)[os]; O7 = PL(uold, 7)[os];
                            }
                        }
                        if (STAGE == 1) {
                            // oneStepFirst_kernel, M/resistiveMHD2D.cu:64-103
                            n0 = S0 - dtdx * (F.f0 - Fp.f0) - dtdy * (Gp.f0 - g0m);
                            n1 = S1 - dtdx * (F.f1 - Fp.f1) - dtdy * (Gp.f1 - g1m);
                            n2 = S2 - dtdx * (F.f2 - Fp.f2) - dtdy * (Gp.f2 - g2m);
                            n3 = S3 - dtdx * (F.f3 - Fp.f3) - dtdy * (Gp.f3 - g3m);
                            n4 = sbx - dtdx * (F.f4 - Fp.f4) - dtdy * (Gp.f4 - g4m);
                            n5 = sby - dtdx * (F.f5 - Fp.f5) - dtdy * (Gp.f5 - g5m);
                            n6 = S6 - dtdx * (F.f6 - Fp.f6) - dtdy * (Gp.f6 - g6m);
                            n7 = S7 - dtdx * (F.f7 - Fp.f7) - dtdy * (Gp.f7 - g7m);
                        } else {
                            // oneStepSecond_kernel, M/resistiveMHD2D.cu:106-144
                            n0 = 0.5 * (O0 + S0 - dtdx * (F.f0 - Fp.f0) - dtdy * (Gp.f0 - g0m));
                            n1 = 0.5 * (O1 + S1 - dtdx * (F.f1 - Fp.f1) - dtdy * (Gp.f1 - g1m));
                            n2 = 0.5 * (O2 + S2 - dtdx * (F.f2 - Fp.f2) - dtdy * (Gp.f2 - g2m));
                            n3 = 0.5 * (O3 + S3 - dtdx * (F.f3 - Fp.f3) - dtdy * (Gp.f3 - g3m));
                            n4 = 0.5 * (obx + sbx - dtdx * (F.f4 - Fp.f4) - dtdy * (Gp.f4 - g4m));
                            n5 = 0.5 * (oby + sby - dtdx * (F.f5 - Fp.f5) - dtdy * (Gp.f5 - g5m));
                            n6 = 0.5 * (O6 + S6 - dtdx * (F.f6 - Fp.f6) - dtdy * (Gp.f6 - g6m));
                            n7 = 0.5 * (O7 + S7 - dtdx * (F.f7 - Fp.f7) - dtdy * (Gp.f7 - g7m));
                        }
                        PL(uout, 0)[os] = n0; PL(uout, 1)[os] = n1; PL(uout, 2)[os] = n2; PL(uout, 3)[os] = n3;
                        PL(uout, 4)[os] = n4; PL(uout, 5)[os] = n5; PL(uout, 6)[os] = n6; PL(uout, 7)[os] = n7;
                        // face fluxes of the CT-contact EMF (M/ct.cu:58-63)
                        PL(aux, 0)[os] = F.f0;
                        PL(aux, 1)[os] = F.f5;
                        PL(aux, 2)[os] = Gp.f0;
                        PL(aux, 3)[os] = Gp.f4;
                    }
                } else if (s < send && (upd_row || stale_row)) {
                    const size_t os = (size_t)(s + dcol) * g.pitch + jl;
                    double n0, n1, n2, n3, n4, n5, n6, n7;
                    if (STAGE == 1) {
                        // oneStepFirstFunctor (P/resistiveMHD_2D.cu:53-86) and
                        // CT_kernel (P/ct.cu:44-59)
                        double U0, U1, U2, U3, U4, U5, U6, U7;
                        if (UPD_DMA) {
                            // s_waitcnt vmcnt(0): the cell's staging loads have landed
                            __builtin_amdgcn_s_waitcnt(0x0F70);
                            const int r = lane + 1;
                            U0 = upd[0][r]; U1 = upd[1][r]; U2 = upd[2][r]; U3 = upd[3][r];
                            U4 = upd[4][r]; U5 = upd[5][r]; U6 = upd[6][r]; U7 = upd[7][r];
                        } else {
                            U0 = Rh[os]; U1 = Mx[os]; U2 = My[os]; U3 = Mz[os];
                            U4 = Bx[os]; U5 = By[os]; U6 = Bz[os]; U7 = En[os];
                        }
                        n0 = U0 - dtdx * (F.f0 - Fp.f0) - dtdy * (Gp.f0 - g0m);
                        n1 = U1 - dtdx * (F.f1 - Fp.f1) - dtdy * (Gp.f1 - g1m);
                        n2 = U2 - dtdx * (F.f2 - Fp.f2) - dtdy * (Gp.f2 - g2m);
                        n3 = U3 - dtdx * (F.f3 - Fp.f3) - dtdy * (Gp.f3 - g3m);
                        n4 = U4 - dtdy * (E - Em);
                        n5 = U5 + dtdx * (E - Eprev);
                        n6 = U6 - dtdx * (F.f6 - Fp.f6) - dtdy * (Gp.f6 - g6m);
                        n7 = U7 - dtdx * (F.f7 - Fp.f7) - dtdy * (Gp.f7 - g7m);
                        if (upd_row) {
                            PL(uout, 0)[os] = n0; PL(uout, 1)[os] = n1; PL(uout, 2)[os] = n2; PL(uout, 3)[os] = n3;
                            PL(uout, 4)[os] = n4; PL(uout, 5)[os] = n5; PL(uout, 6)[os] = n6; PL(uout, 7)[os] = n7;
                        } else {
                            // CT value of the bottom boundary face, read by the
                            // symmetric-x boundary before the y boundary
                            // overwrites it (S/boundary.cu:56-75).
                            PL(uout, 5)[os] = n5;
                            side[s + dcol] = n5;
                        }
                    } else {
                        if (upd_row) {
                            // oneStepSecondFunctor (P/resistiveMHD_2D.cu:88-130);
                            // B is the stage-1 CT value (same EMF, same old B).
                            double S0, S1, S2, S3, S4, S5, S6, S7, O0, O1, O2, O3, O6, O7;
                            if (UPD_DMA) {
                                // s_waitcnt vmcnt(0): the cell's staging loads have landed
                                __builtin_amdgcn_s_waitcnt(0x0F70);
                                const int r = lane + 1;
                                S0 = upd[0][r]; S1 = upd[1][r]; S2 = upd[2][r]; S3 = upd[3][r];
                                S4 = upd[4][r]; S5 = upd[5][r]; S6 = upd[6][r]; S7 = upd[7][r];
                                O0 = upd[8][r]; O1 = upd[9][r]; O2 = upd[10][r]; O3 = upd[11][r];
                                O6 = upd[12][r]; O7 = upd[13][r];
                            } else {
                                S0 = Rh[os]; S1 = Mx[os]; S2 = My[os]; S3 = Mz[os];
                                S4 = Bx[os]; S5 = By[os]; S6 = Bz[os]; S7 = En[os];
                                O0 = PL(uold, 0)[os]; O1 = PL(uold, 1)[os]; O2 = PL(uold, 2)[os];
                                O3 = PL(uold, 3)[os]; O6 = PL(uold, 6)[os]; O7 = PL(uold, 7)[os];
                            }
                            n0 = 0.5 * (O0 + S0 - dtdx * (F.f0 - Fp.f0) - dtdy * (Gp.f0 - g0m));
                            n1 = 0.5 * (O1 + S1 - dtdx * (F.f1 - Fp.f1) - dtdy * (Gp.f1 - g1m));
                            n2 = 0.5 * (O2 + S2 - dtdx * (F.f2 - Fp.f2) - dtdy * (Gp.f2 - g2m));
                            n3 = 0.5 * (O3 + S3 - dtdx * (F.f3 - Fp.f3) - dtdy * (Gp.f3 - g3m));
                            n4 = S4;
                            n5 = S5;
                            n6 = 0.5 * (O6 + S6 - dtdx * (F.f6 - Fp.f6) - dtdy * (Gp.f6 - g6m));
                            n7 = 0.5 * (O7 + S7 - dtdx * (F.f7 - Fp.f7) - dtdy * (Gp.f7 - g7m));
                            PL(uout, 0)[os] = n0; PL(uout, 1)[os] = n1; PL(uout, 2)[os] = n2; PL(uout, 3)[os] = n3;
                            PL(uout, 4)[os] = n4; PL(uout, 5)[os] = n5; PL(uout, 6)[os] = n6; PL(uout, 7)[os] = n7;
                            // fused epilogue: per-cell dt of the new state for
                            // the next step, and the crash flag
                            const double d = cell_dt<EXACT>(n0, n1, n2, n3, n4, n5, n6, n7, ph);
                            if (d < dtloc) dtloc = d;
                            nanflag |= isnan(n7) ? 1 : 0;
                        } else {
                            PL(uout, 5)[os] = side[s + dcol];
                        }
                    }
                }
            }
            if (FLDS && MULTI) {
                fb[0][lane] = F.f0; fb[1][lane] = F.f1; fb[2][lane] = F.f2; fb[3][lane] = F.f3;
                fb[4][lane] = F.f4; fb[5][lane] = F.f5; fb[6][lane] = F.f6; fb[7][lane] = F.f7;
            } else if (FLDS) {
                fb[0][lane] = F.f0; fb[1][lane] = F.f1; fb[2][lane] = F.f2;
                fb[3][lane] = F.f3; fb[4][lane] = F.f6; fb[5][lane] = F.f7;
            } else {
                Fprev = F;
            }
            if (GLDS && MULTI) {
                gb[0][lane] = G.f0; gb[1][lane] = G.f1; gb[2][lane] = G.f2; gb[3][lane] = G.f3;
                gb[4][lane] = G.f4; gb[5][lane] = G.f5; gb[6][lane] = G.f6; gb[7][lane] = G.f7;
            } else if (GLDS) {
                gb[0][lane] = G.f0; gb[1][lane] = G.f1; gb[2][lane] = G.f2; gb[3][lane] = G.f3;
                gb[4][lane] = G.f4; gb[5][lane] = G.f6; gb[6][lane] = G.f7;
            } else {
                Gprev = G;
            }
            Eprev = E;
        }
        A = B;
        B = C;
        if (ETAM) {
            eta_prev2 = eta_prev;
            etay_prev = etay_c;
        }
        eta_prev = eta_c;
    }

    if (STAGE == 2 && !MULTI) {
        accumulate_dt(ctrl, cur ^ 1, dtloc);
        const unsigned long long any = __ballot(nanflag);
        if (lane == 0 && any) atomicOr(&ctrl->crashed, 1);
    }
}

// Workgroups are dealt to the XCDs round-robin (block b runs on XCD
// b % xcds).  Each XCD gets a contiguous range of items (strips fastest), so
// neighbouring strips and segments, which read the same halo lines, share
// that XCD's L2.
__device__ __forceinline__ int xcd_item(int xcds, int items_per_xcd)
{
    return (int)(blockIdx.x % xcds) * items_per_xcd + (int)(blockIdx.x / xcds);
}

template <int STAGE, bool EXACT, int SCHEME>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(EXACT ? MHD_WAVES_PER_EU_EXACT : MHD_WAVES_PER_EU_FAST)))
void rk_stage_kernel(const double *__restrict__ uin, double *uout, const double *uold,
                     const double *__restrict__ etaf, double *__restrict__ side, double *__restrict__ aux,
                     Ctrl *__restrict__ ctrl, Grid g, Phys ph, int cur)
{
    const int item = xcd_item(g.xcds, g.items_per_xcd);
    if (item >= g.nstrips * g.nsegs) return;
    stage_item<STAGE, EXACT, SCHEME>(item, uin, uout, uold, etaf, side, aux, ctrl, g, ph, cur);
}

// Batched form: one launch advances every run of a batch by one stage.  The
// work list holds the items of all runs back to back; item_run maps an item
// to its run, whose descriptor gives the grid, constants and buffers.  Each
// run keeps its own dt (its own control block).  XCD x owns the items
// [x * items_per_xcd, (x + 1) * items_per_xcd); its blocks_per_xcd
// wavefronts take them in turn.  blocks_per_xcd == items_per_xcd is one
// wavefront per item (PERSIST false: no loop, whose bookkeeping costs the
// exact kernel spill space); fewer makes the launch persistent (each
// wavefront loops over items).
template <int STAGE, bool EXACT>
__device__ __forceinline__ void batch_item(const RunDesc *__restrict__ runs, const int *__restrict__ item_run,
                                           int item, int cur)
{
    const RunDesc &d = runs[item_run[item]];
    const Grid g = d.g;
    const Phys ph = d.ph;
    if (STAGE == 1)
        stage_item<1, EXACT, 0>(item - d.first_item, d.u, d.us, d.u, d.eta, d.side, nullptr, d.ctrl, g, ph, cur);
    else
        stage_item<2, EXACT, 0>(item - d.first_item, d.us, d.uo ? d.uo : d.u, d.u, d.eta, d.side, nullptr, d.ctrl, g,
                                ph, cur);
}

// The packed last strips of a batch (runs whose grid has tail_g): a launch
// of its own after the full strips' launch, one wavefront per packed item,
// so the full-strip kernel keeps its register allocation.  tail_run maps a
// packed item to its run; the run's packed items start at first_tail.
template <int STAGE, bool EXACT>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(EXACT ? MHD_WAVES_PER_EU_EXACT : MHD_WAVES_PER_EU_FAST)))
void rk_stage_tail_batch_kernel(const RunDesc *__restrict__ runs, const int *__restrict__ tail_run, int total,
                                int cur)
{
    const int item = (int)blockIdx.x;
    if (item < total) tail_item<STAGE, EXACT>(runs, tail_run, item, cur);
}

template <int STAGE, bool EXACT>
__device__ __forceinline__ void tail_item(const RunDesc *__restrict__ runs, const int *__restrict__ tail_run, int item,
                                          int cur)
{
    const RunDesc &d = runs[tail_run[item]];
    const Grid g = d.g;
    const Phys ph = d.ph;
    if (STAGE == 1)
        stage_item<1, EXACT, 0, true>(item - d.first_tail, d.u, d.us, d.u, d.eta, d.side, nullptr, d.ctrl, g, ph,
                                      cur);
    else
        stage_item<2, EXACT, 0, true>(item - d.first_tail, d.us, d.uo ? d.uo : d.u, d.u, d.eta, d.side, nullptr,
                                      d.ctrl, g, ph, cur);
}

template <int STAGE, bool EXACT, bool PERSIST>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(EXACT ? MHD_WAVES_PER_EU_EXACT : MHD_WAVES_PER_EU_FAST)))
void rk_stage_batch_kernel(const RunDesc *__restrict__ runs, const int *__restrict__ item_run, int total_items,
                           int xcds, int items_per_xcd, int blocks_per_xcd, int cur, const int *__restrict__ tail_run,
                           int tail_items)
{
    if (!PERSIST) {
        // the packed items follow the full strips' blocks, so they fill the
        // launch's last round
        const int main_blocks = xcds * items_per_xcd;
        if ((int)blockIdx.x >= main_blocks) {
            if ((int)blockIdx.x - main_blocks < tail_items)
                tail_item<STAGE, EXACT>(runs, tail_run, (int)blockIdx.x - main_blocks, cur);
            return;
        }
        const int item = xcd_item(xcds, items_per_xcd);
        if (item < total_items) batch_item<STAGE, EXACT>(runs, item_run, item, cur);
        return;
    }
    const int x = (int)(blockIdx.x % xcds);
    for (int k = (int)(blockIdx.x / xcds); k < items_per_xcd; k += blocks_per_xcd) {
        const int item = x * items_per_xcd + k;
        if (item >= total_items) return;
        batch_item<STAGE, EXACT>(runs, item_run, item, cur);
    }
}

// ------------------------------------------------------------ boundaries
// Periodic x (P/boundary.cu:4-71) or symmetric x (S/boundary.cu:4-104).
// One ghost cell per lane: a wavefront takes one ghost column q (columns 0,
// 1, 2, nx-3, nx-2, nx-1) and 64 consecutive rows, so each plane's load and
// store is one coalesced run.  The symmetric bX rebuild reads By of the
// interior columns the mirror copied from (identical values), so one pass
// suffices.
template <int KIND, bool EXACT>
__device__ __forceinline__ void bc_x_cells(int q, int j, double *u, Ctrl *ctrl, int slot, const Grid &g,
                                           const Phys &ph)
{
    double dloc = __longlong_as_double((long long)kInfBits);
    if (j < g.ny) {
        const int nx = g.nx;
        const size_t pitch = g.pitch;
        const int dst = q < 3 ? q : nx - 6 + q;   // 0, 1, 2, nx-3, nx-2, nx-1
        const int src = KIND == 0 ? (q < 3 ? nx - 6 + q : q)           // nx-6 .. nx-4, 3 .. 5
                                  : (q < 3 ? 5 - q : nx - 1 - q);      // 5, 4, 3, nx-4, nx-5, nx-6
        double cell[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            if (KIND == 1 && v == 4) continue;
            cell[v] = PL(u, v)[(size_t)src * pitch + j];
            PL(u, v)[(size_t)dst * pitch + j] = cell[v];
        }
        if (KIND == 1) {
            // bX of the ghost columns from the zero divergence of the
            // mirrored By: columns 2, 1, 0 from columns 3, 3, 4 and columns
            // nx-3, nx-2, nx-1 from columns nx-4, nx-5, nx-6, each added in
            // turn
            const double *by = PL(u, 5);
            double *bx = PL(u, 4);
            double b;
            if (j > 0) {
                if (q < 3) {
                    const size_t c3 = (size_t)3 * pitch + j;
                    const double d3 = by[c3] - by[c3 - 1];
                    b = bx[c3] + d3 / ph.dy * ph.dx;                                       // column 2
                    if (q < 2) b = b + d3 / ph.dy * ph.dx;                                 // column 1
                    if (q < 1) b = b + (by[c3 + pitch] - by[c3 + pitch - 1]) / ph.dy * ph.dx;   // column 0
                } else {
                    const size_t cA = (size_t)(nx - 4) * pitch + j;
                    b = bx[cA] - (by[cA] - by[cA - 1]) / ph.dy * ph.dx;                    // column nx-3
                    if (q > 3) b = b - (by[cA - pitch] - by[cA - pitch - 1]) / ph.dy * ph.dx;           // nx-2
                    if (q > 4) b = b - (by[cA - 2 * pitch] - by[cA - 2 * pitch - 1]) / ph.dy * ph.dx;   // nx-1
                }
            } else {
                b = q < 3 ? bx[(size_t)3 * pitch] : bx[(size_t)(nx - 4) * pitch];
            }
            cell[4] = b;
            bx[(size_t)dst * pitch + j] = b;
        }
        if (ctrl && j >= kGhost && j < g.ny - kGhost) {
            const double d = cell_dt<EXACT>(cell[0], cell[1], cell[2], cell[3], cell[4], cell[5], cell[6], cell[7], ph);
            if (d < dloc) dloc = d;
        }
    }
    if (ctrl) accumulate_dt(ctrl, slot, dloc);
}

// Wavefront w of the x boundary: ghost column w % 6, rows (w / 6) * 64 ..
__device__ __forceinline__ int bcx_wave() { return (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); }

template <int KIND, bool EXACT>
__global__ __launch_bounds__(256) void bc_x_kernel(double *u, Ctrl *ctrl, int slot, Grid g, Phys ph)
{
    const int w = bcx_wave();
    bc_x_cells<KIND, EXACT>(w % 6, (w / 6) * 64 + (int)(threadIdx.x & 63), u, ctrl, slot, g, ph);
}

// Periodic y (P/boundary.cu:74-141) or symmetric y (S/boundary.cu:107-205),
// after the x boundary.  One ghost cell per lane: lanes 6c .. 6c+5 of a
// wavefront take column col0 + c, ghost rows 0, 1, 2 and ny-3, ny-2, ny-1
// (lanes 60..63 idle), so a load touches the two 24-byte ghost runs of ten
// neighbouring columns instead of 64 columns' worth of lines.
constexpr int kBcyCols = 10;   // columns per wavefront

template <int KIND, bool EXACT>
__device__ __forceinline__ void bc_y_cells(int col0, double *u, Ctrl *ctrl, int slot, const Grid &g,
                                           const Phys &ph)
{
    double dloc = __longlong_as_double((long long)kInfBits);
    const int lane = threadIdx.x & 63;
    const int q = lane % 6;
    const int i = col0 + lane / 6;
    if (lane < 6 * kBcyCols && i < g.nx) {
        const int ny = g.ny;
        const size_t row = (size_t)i * g.pitch;
        const int dst = q < 3 ? q : ny - 6 + q;   // 0, 1, 2, ny-3, ny-2, ny-1
        const int src = KIND == 0 ? (q < 3 ? ny - 6 + q : q)           // ny-6 .. ny-4, 3 .. 5
                                  : (q < 3 ? 5 - q : ny - 1 - q);      // 5, 4, 3, ny-4, ny-5, ny-6
        double cell[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            if (KIND == 1 && v == 5) continue;
            cell[v] = PL(u, v)[row + src];
            PL(u, v)[row + dst] = cell[v];
        }
        if (KIND == 1) {
            // By of the ghost rows from the zero divergence of the mirrored
            // Bx: rows 2, 1, 0 from rows 3, 3, 4 and rows ny-3, ny-2, ny-1
            // from rows ny-4, ny-5, ny-6, each added in turn
            const double *bx = PL(u, 4);
            double *by = PL(u, 5);
            double b;
            if (i > 0) {
                const size_t rm = (size_t)(i - 1) * g.pitch;
                if (q < 3) {
                    const double d3 = bx[row + 3] - bx[rm + 3];
                    b = by[row + 3] + d3 / ph.dx * ph.dy;                  // row 2
                    if (q < 2) b = b + d3 / ph.dx * ph.dy;                  // row 1
                    if (q < 1) b = b + (bx[row + 4] - bx[rm + 4]) / ph.dx * ph.dy;   // row 0
                } else {
                    const double dA = bx[row + ny - 4] - bx[rm + ny - 4];
                    b = by[row + ny - 4] - dA / ph.dx * ph.dy;             // row ny-3
                    if (q > 3) b = b - (bx[row + ny - 5] - bx[rm + ny - 5]) / ph.dx * ph.dy;   // row ny-2
                    if (q > 4) b = b - (bx[row + ny - 6] - bx[rm + ny - 6]) / ph.dx * ph.dy;   // row ny-1
                }
            } else {
                b = q < 3 ? by[row + 3] : by[row + ny - 4];
            }
            cell[5] = b;
            by[row + dst] = b;
        }
        if (ctrl) {
            const double d = cell_dt<EXACT>(cell[0], cell[1], cell[2], cell[3], cell[4], cell[5], cell[6], cell[7], ph);
            if (d < dloc) dloc = d;
            if (q == 0 && i == 0 && isnan(d)) ctrl->first_nan[slot] = 1;   // cell (0, 0)
        }
    }
    if (ctrl) accumulate_dt(ctrl, slot, dloc);
}

template <int KIND, bool EXACT>
__global__ __launch_bounds__(256) void bc_y_kernel(double *u, Ctrl *ctrl, int slot, Grid g, Phys ph)
{
    bc_y_cells<KIND, EXACT>((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kBcyCols, u, ctrl, slot, g,
                            ph);
}

// Batched boundaries: blockIdx.y is the run; which state (U or U*) and
// whether the dt of the ghost cells is reduced are the same for every run.
template <bool EXACT>
__global__ __launch_bounds__(256) void bc_x_batch_kernel(const RunDesc *__restrict__ runs, int stage_state,
                                                         int with_dt, int slot)
{
    const RunDesc &d = runs[blockIdx.y];
    const Grid g = d.g;
    const Phys ph = d.ph;
    if ((int)(blockIdx.x * (blockDim.x >> 6) / 6) * 64 >= g.ny) return;   // whole workgroup past this run's rows
    const int w = bcx_wave();
    const int q = w % 6, j = (w / 6) * 64 + (int)(threadIdx.x & 63);
    double *u = stage_state ? d.us : d.u;
    Ctrl *ctrl = with_dt ? d.ctrl : nullptr;
    if (d.bc_x == 0) bc_x_cells<0, EXACT>(q, j, u, ctrl, slot, g, ph);
    else bc_x_cells<1, EXACT>(q, j, u, ctrl, slot, g, ph);
}

template <bool EXACT>
__global__ __launch_bounds__(256) void bc_y_batch_kernel(const RunDesc *__restrict__ runs, int stage_state,
                                                         int with_dt, int slot)
{
    const RunDesc &d = runs[blockIdx.y];
    const Grid g = d.g;
    const Phys ph = d.ph;
    if ((int)(blockIdx.x * (blockDim.x >> 6)) * kBcyCols >= g.nx) return;   // whole workgroup past this run's columns
    const int col0 = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kBcyCols;
    double *u = stage_state ? d.us : d.u;
    Ctrl *ctrl = with_dt ? d.ctrl : nullptr;
    if (d.bc_y == 0) bc_y_cells<0, EXACT>(col0, u, ctrl, slot, g, ph);
    else bc_y_cells<1, EXACT>(col0, u, ctrl, slot, g, ph);
}

// Per-cell dt over all nx*ny cells (P/resistiveMHD_2D.cu:277-286) and the
// NaN scan of e, for a state set from outside the step.
// INTERIOR: the dt of the MPI library's calculateDt, interior cells only
// (M/resistiveMHD2D.cu:489-524); the NaN scan still covers every cell.
template <bool EXACT, bool INTERIOR>
__device__ __forceinline__ void dt_all_body(const double *__restrict__ u, Ctrl *ctrl, int slot, const Grid &g,
                                            const Phys &ph)
{
    double dloc = __longlong_as_double((long long)kInfBits);
    int nanflag = 0;
    // grid-stride over rows (blockIdx.x) and columns (threads): a few
    // thousand wavefronts each fold many cells before one atomic
    for (int i = blockIdx.x; i < g.nx; i += gridDim.x) {
        for (int j = threadIdx.x; j < g.ny; j += blockDim.x) {
            const size_t o = (size_t)i * g.pitch + j;
            const double e = PL(u, 7)[o];
            const double d = cell_dt<EXACT>(PL(u, 0)[o], PL(u, 1)[o], PL(u, 2)[o], PL(u, 3)[o], PL(u, 4)[o],
                                            PL(u, 5)[o], PL(u, 6)[o], e, ph);
            const bool inside = !INTERIOR || (i >= kGhost && i < g.nx - kGhost && j >= kGhost && j < g.ny - kGhost);
            if (inside && d < dloc) dloc = d;
            const int first = INTERIOR ? kGhost : 0;
            if (i == first && j == first && isnan(d)) ctrl->first_nan[slot] = 1;
            nanflag |= isnan(e) ? 1 : 0;   // checkCalculationIsCrashed, P/resistiveMHD_2D.cu:294-309
        }
    }
    accumulate_dt(ctrl, slot, dloc);
    const unsigned long long any = __ballot(nanflag);
    if ((threadIdx.x & 63) == 0 && any) atomicOr(&ctrl->crashed, 1);
}

template <bool EXACT, bool INTERIOR>
__global__ __launch_bounds__(256) void dt_all_kernel(const double *__restrict__ u, Ctrl *ctrl, int slot, Grid g,
                                                     Phys ph)
{
    dt_all_body<EXACT, INTERIOR>(u, ctrl, slot, g, ph);
}

// The same for every run of a batch (blockIdx.y = run).
template <bool EXACT>
__global__ __launch_bounds__(256) void dt_all_batch_kernel(const RunDesc *__restrict__ runs, int slot)
{
    const RunDesc &d = runs[blockIdx.y];
    const Grid g = d.g;
    const Phys ph = d.ph;
    dt_all_body<EXACT, false>(d.u, d.ctrl, slot, g, ph);
}

__global__ void ctrl_reset_batch_kernel(const RunDesc *__restrict__ runs, int nruns, int slot)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nruns) {
        Ctrl *c = runs[r].ctrl;
        c->acc[slot] = kInfBits;
        c->first_nan[slot] = 0;
        c->crashed = 0;
    }
}

__global__ void ctrl_reset_kernel(Ctrl *ctrl, int slot)
{
    if (threadIdx.x == 0) {
        ctrl->acc[slot] = kInfBits;
        ctrl->first_nan[slot] = 0;
        ctrl->crashed = 0;
    }
}

// ------------------------------------------------ MPI library's step (M/)
// The CT-contact EMF of M/ct.cu and the final CT, on one periodic subdomain.

// Columns per wavefront of the column-walking kernels below (smoothed
// cell EMF; corner EMF with the final CT).
#ifndef MHD_EC_COLUMNS
#define MHD_EC_COLUMNS 8
#endif
#ifndef MHD_CT_COLUMNS
#define MHD_CT_COLUMNS 32
#endif
constexpr int kEcSeg = MHD_EC_COLUMNS;
constexpr int kMSeg = MHD_CT_COLUMNS;

// Face value smoothed along its normal (shiftUToCenterForCT then
// backUToCenterHalfForCT, M/resistiveMHD2D.cu:597-690) from the face values
// before (m), at (c) and after (p) it.
__device__ __forceinline__ double smooth3(double m, double c, double p)
{
    const double c0 = 0.5 * (c + m);
    const double c1 = 0.5 * (p + c);
    return 0.5 * (c0 + c1);
}

// Cell-centred reference EMF u*bY - v*bX - eta*jZ of CT::setOldFlux2D /
// setNowFlux2D (M/ct.cu:26-65; the G4 array there is its negation) at
// cell (i, j) from its centred state and the face B around it: X0, X1 the
// x-face B of columns i-1 and i at rows j-1 .. j+1, Y0, Y1, Y2 the y-face B
// of columns i-1 .. i+1 at rows j-1 and j.
__device__ __forceinline__ double cell_emf(const double *u, const double *eta_c, size_t o, const double *X0,
                                           const double *X1, const double *Y0, const double *Y1, const double *Y2,
                                           const Grid &g, const Phys &ph)
{
    const double rho = PL(u, 0)[o];
    const double uu = PL(u, 1)[o] / rho;
    const double vv = PL(u, 2)[o] / rho;
    const double bX = 0.5 * (X1[1] + X0[1]);
    const double bY = 0.5 * (Y1[1] + Y1[0]);
    const double jZ = 0.25 * (
          (Y2[1] - Y1[1]) / ph.dx - (X1[2] - X1[1]) / ph.dy
        + (Y2[0] - Y1[0]) / ph.dx - (X1[1] - X1[0]) / ph.dy
        + (Y1[0] - Y0[0]) / ph.dx - (X0[1] - X0[0]) / ph.dy
        + (Y1[1] - Y0[1]) / ph.dx - (X0[2] - X0[1]) / ph.dy);
    const double eta = eta_c ? eta_c[o] : ph.eta;   // getEta(i dx, j dy), M/ct.cu:41,56
    return uu * bY - vv * bX - eta * jZ;
}

// Cell EMF of the state as stored, over the interior: one cell per lane,
// a workgroup per 256 rows of one column (neighbouring columns' workgroups
// run together, so the face values read around a cell come from L2).
__global__ __launch_bounds__(256) void m_ec_kernel(const double *__restrict__ u, const double *__restrict__ eta_c,
                                                   double *__restrict__ ec, Grid g, Phys ph)
{
    const int j = kGhost + blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.ny - kGhost) return;
    const size_t P = g.pitch;
    const double *bx = PL(u, 4), *by = PL(u, 5);
    for (int i = kGhost + blockIdx.y; i < g.nx - kGhost; i += gridDim.y) {
        const size_t o = (size_t)i * P + j;
        double X0[3], X1[3], Y0[2], Y1[2], Y2[2];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            X0[q] = bx[o - P - 1 + q];
            X1[q] = bx[o - 1 + q];
        }
        Y0[0] = by[o - P - 1]; Y0[1] = by[o - P];
        Y1[0] = by[o - 1];     Y1[1] = by[o];
        Y2[0] = by[o + P - 1]; Y2[1] = by[o + P];
        ec[o] = cell_emf(u, eta_c, o, X0, X1, Y0, Y1, Y2, g, ph);
    }
}

// Cell EMF of the face-smoothed state (stage 1 sees U^n after
// backUToCenterHalfForCT), over the interior.  Each smoothed face value
// takes three raw ones, so one lane per row walks kEcSeg columns and
// carries what it has formed: the smoothed X0, X1, Y0, Y1 and the raw bx of
// columns i and i+1 (the raw by of rows j-2 .. j+1 of column i+1 is read
// for Y2), so each face value is read once per lane.
__global__ __launch_bounds__(256) void m_ec_smooth_kernel(const double *__restrict__ u,
                                                          const double *__restrict__ eta_c, double *__restrict__ ec,
                                                          Grid g, Phys ph)
{
    const int jj = kGhost + blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = jj < g.ny - kGhost;
    const int j = on ? jj : kGhost;   // lanes past the interior read row kGhost's stencil and store nothing
    const int ib = kGhost + blockIdx.y * kEcSeg, ie = min(ib + kEcSeg, g.nx - kGhost);
    const size_t P = g.pitch;
    const double *bx = PL(u, 4), *by = PL(u, 5);
    auto raw = [&](int c, double *r) {
#pragma unroll
        for (int q = 0; q < 3; ++q) r[q] = bx[(size_t)c * P + j - 1 + q];
    };
    auto col_y = [&](int c, double *y) {
        const double *b = by + (size_t)c * P + j;
        y[0] = smooth3(b[-2], b[-1], b[0]);
        y[1] = smooth3(b[-1], b[0], b[1]);
    };
    double X0[3], X1[3], Y0[2], Y1[2], Y2[2], r0[3], r1[3];
    {
        double a[3], b[3];
        raw(ib - 2, a);
        raw(ib - 1, b);
        raw(ib, r0);
        raw(ib + 1, r1);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            X0[q] = smooth3(a[q], b[q], r0[q]);
            X1[q] = smooth3(b[q], r0[q], r1[q]);
        }
    }
    col_y(ib - 1, Y0);
    col_y(ib, Y1);
    for (int i = ib; i < ie; ++i) {
        col_y(i + 1, Y2);
        if (on) {
            const size_t o = (size_t)i * P + j;
            ec[o] = cell_emf(u, eta_c, o, X0, X1, Y0, Y1, Y2, g, ph);
        }
        if (i + 1 == ie) break;
        double r2[3];
        raw(i + 2, r2);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            X0[q] = X1[q];
            X1[q] = smooth3(r0[q], r1[q], r2[q]);
            r0[q] = r1[q];
            r1[q] = r2[q];
        }
        Y0[0] = Y1[0]; Y0[1] = Y1[1];
        Y1[0] = Y2[0]; Y1[1] = Y2[1];
    }
}

__device__ __forceinline__ double msign(double x) { return (double)((x > 0.0) - (x < 0.0)); }   // M/sign.cu

// Corner EMF of one flux set at (i+1/2, j+1/2) (M/ct.cu:147-169) from the
// x-face values F.f0, F.f5 of (i, j) and (i, j+1) (f0, f0N, f5, f5N), the
// y-face values G.f0, G.f4 of (i, j) and (i+1, j) (g0, g4, g0E, g4E) and
// the cell EMFs ec = F5 = -G4 of (i, j), (i, j+1), (i+1, j), (i+1, j+1).
struct CornerIn {
    double f0, f0N, f5, f5N, g0, g4, g0E, g4E, ec, ecN, ecE, ecNE;
};

__device__ __forceinline__ double contact_emf(const CornerIn &c)
{
    const double avg = 0.25 * (-c.f5 - c.f5N + c.g4 + c.g4E);
    const double sS = msign(c.f0), sN = msign(c.f0N);
    const double sW = msign(c.g0), sE = msign(c.g0E);
    const double eS = (1.0 + sS) * (c.g4 - (-c.ec)) + (1.0 - sS) * (c.g4E - (-c.ecE));
    const double eN = (1.0 + sN) * ((-c.ecN) - c.g4) + (1.0 - sN) * ((-c.ecNE) - c.g4E);
    const double eW = -(1.0 + sW) * (c.f5 - c.ec) - (1.0 - sW) * (c.f5N - c.ecN);
    const double eE = -(1.0 + sE) * (c.ecE - c.f5) - (1.0 - sE) * (c.ecNE - c.f5N);
    return avg + 0.125 * (eS - eN + eW - eE);
}

// Rows of the interior each wavefront of m_emf_ct_kernel updates.
constexpr int kMCtRows = 63;

// The corner EMF E = (E_old + E_now) / 2 at (i+1/2, j+1/2) (CT::divBClean,
// M/ct.cu:120-218) and the final CT from the step's initial face B
// (M/ct.cu:172-218) over the interior, then the per-cell dt and NaN flag of
// the new state (the next step's calculateDt, M/resistiveMHD2D.cu:489-524),
// in one pass: a wavefront takes kMCtRows rows and walks kMSeg columns.
// Lane L holds corner row jb - 1 + L; at column i it forms its corner's
// E(i, row) and updates cell (i, row) with E(i, row - 1) from the lane
// below (lane 0 only supplies that value) and E(i - 1, row) from the column
// before.  Column i+1's y-face values and cell EMFs, read for corner i,
// are carried to corner i+1, so each input is read once per lane.  The
// per-face and per-cell inputs are stored for the interior only; the ghost
// layer's values are those of the periodic image, which the self-exchange
// halo makes equal.
template <bool EXACT>
__global__ __launch_bounds__(256) void m_emf_ct_kernel(const double *__restrict__ aux_old,
                                                       const double *__restrict__ aux_now,
                                                       const double *__restrict__ ec_old,
                                                       const double *__restrict__ ec_now,
                                                       const double *__restrict__ uold, double *__restrict__ u,
                                                       Ctrl *__restrict__ ctrl, int slot, Grid g, Phys ph)
{
    const int lane = threadIdx.x & 63;
    const int jb = kGhost + kMCtRows * (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (jb >= g.ny - kGhost) return;   // whole wavefront past the interior
    const int lnx = g.nx - 2 * kGhost, lny = g.ny - 2 * kGhost;
    auto wrap = [](int a, int n, int l) { return a < kGhost ? a + l : (a >= n - kGhost ? a - l : a); };
    const bool on = jb - 1 + lane < g.ny - kGhost;   // corner rows kGhost-1 .. ny-kGhost-1
    const int j = on ? jb - 1 + lane : jb;           // lanes past the interior: a valid row, nothing stored
    const bool out = on && lane > 0;
    const int wj = wrap(j, g.ny, lny), wj1 = wrap(j + 1, g.ny, lny);
    const int ib = kGhost + blockIdx.y * kMSeg, ie = min(ib + kMSeg, g.nx - kGhost);
    const long long plane = g.plane;
    const double *aux[2] = {aux_old, aux_now}, *ecp[2] = {ec_old, ec_now};
    CornerIn c[2];
    {
        const size_t wi = (size_t)wrap(ib - 1, g.nx, lnx) * g.pitch;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            c[s].g0E = aux[s][2 * plane + wi + wj];
            c[s].g4E = aux[s][3 * plane + wi + wj];
            c[s].ecE = ecp[s][wi + wj];
            c[s].ecNE = ecp[s][wi + wj1];
        }
    }
    const double dt = ctrl->dt;
    double dloc = __longlong_as_double((long long)kInfBits);
    int nanflag = 0;
    double eW = 0.0;
    for (int i = ib - 1; i < ie; ++i) {
        const size_t wi = (size_t)wrap(i, g.nx, lnx) * g.pitch, wi1 = (size_t)wrap(i + 1, g.nx, lnx) * g.pitch;
        double es[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            CornerIn &q = c[s];
            q.g0 = q.g0E;
            q.g4 = q.g4E;
            q.ec = q.ecE;
            q.ecN = q.ecNE;
            q.f0 = aux[s][wi + wj];
            q.f0N = aux[s][wi + wj1];
            q.f5 = aux[s][plane + wi + wj];
            q.f5N = aux[s][plane + wi + wj1];
            q.g0E = aux[s][2 * plane + wi1 + wj];
            q.g4E = aux[s][3 * plane + wi1 + wj];
            q.ecE = ecp[s][wi1 + wj];
            q.ecNE = ecp[s][wi1 + wj1];
            es[s] = contact_emf(q);
        }
        const double e = 0.5 * (es[0] + es[1]);
        const double eS = lane_prev(e);   // every lane active: the loop is uniform
        if (i >= ib && out) {
            const size_t o = (size_t)i * g.pitch + j;
            const double nbx = PL(uold, 4)[o] - dt / ph.dy * (e - eS);
            const double nby = PL(uold, 5)[o] + dt / ph.dx * (e - eW);
            PL(u, 4)[o] = nbx;
            PL(u, 5)[o] = nby;
            const double en = PL(u, 7)[o];
            const double d = cell_dt<EXACT>(PL(u, 0)[o], PL(u, 1)[o], PL(u, 2)[o], PL(u, 3)[o], nbx, nby, PL(u, 6)[o],
                                            en, ph);
            if (d < dloc) dloc = d;
            if (i == kGhost && j == kGhost && isnan(d)) ctrl->first_nan[slot] = 1;
            nanflag |= isnan(en) ? 1 : 0;
        }
        eW = e;
    }
    accumulate_dt(ctrl, slot, dloc);
    const unsigned long long any = __ballot(nanflag);
    if (lane == 0 && any) atomicOr(&ctrl->crashed, 1);
}

// AoS (reference ConservationParameter order) <-> SoA planes.
__global__ __launch_bounds__(256) void unpack_kernel(const double *__restrict__ aos, double *__restrict__ u, Grid g)
{
    const long long total = (long long)g.nx * g.ny;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(k / g.ny), j = (int)(k % g.ny);
        const size_t o = (size_t)i * g.pitch + j;
#pragma unroll
        for (int v = 0; v < 8; ++v) PL(u, v)[o] = aos[(size_t)k * 8 + v];
    }
}

__global__ __launch_bounds__(256) void pack_kernel(const double *__restrict__ u, double *__restrict__ aos, Grid g)
{
    const long long total = (long long)g.nx * g.ny;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(k / g.ny), j = (int)(k % g.ny);
        const size_t o = (size_t)i * g.pitch + j;
#pragma unroll
        for (int v = 0; v < 8; ++v) aos[(size_t)k * 8 + v] = PL(u, v)[o];
    }
}

__global__ __launch_bounds__(256) void scatter_plane_kernel(const double *__restrict__ src, double *__restrict__ dst,
                                                            Grid g)
{
    const long long total = (long long)g.nx * g.ny;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(k / g.ny), j = (int)(k % g.ny);
        dst[(size_t)i * g.pitch + j] = src[k];
    }
}

// Diagnostic: region_bits of every face the update uses, in one arithmetic,
// from the stage kernel's reconstruction (cells with centred B, minmod
// slopes, the face's normal field the mean of the two cells' centred
// components).  rx[j + i*ny]: face (i+1/2, j); ry[j + i*ny]: face
// (i, j+1/2); 0xFF where no update reads the face.
template <bool EXACT>
__device__ __forceinline__ Cell cell_at(const double *u, const Grid &g, const Phys &ph, int i, int j)
{
    const size_t o = (size_t)i * g.pitch + j;
    const double bxc = 0.5 * (PL(u, 4)[o] + PL(u, 4)[o - g.pitch]);
    const double byc = 0.5 * (PL(u, 5)[o] + PL(u, 5)[o - 1]);
    return make_cell<EXACT>(PL(u, 0)[o], PL(u, 1)[o], PL(u, 2)[o], PL(u, 3)[o], bxc, byc, PL(u, 6)[o],
                            PL(u, 7)[o], ph);
}

// Left and right states of the face between cells a and b, with m before a
// and p after b, in the frame given by the field order.
__device__ __forceinline__ void face_states(const double (&m)[7], const double (&a)[7], const double (&b)[7],
                                            const double (&p)[7], double eps, Face7 &L, Face7 &R)
{
    double l[7], r[7];
#pragma unroll
    for (int v = 0; v < 7; ++v) {
        l[v] = a[v] + 0.5 * minmod(a[v] - m[v], b[v] - a[v], eps);
        r[v] = b[v] - 0.5 * minmod(b[v] - a[v], p[v] - b[v], eps);
    }
    L = Face7{l[0], l[1], l[2], l[3], l[4], l[5], l[6]};
    R = Face7{r[0], r[1], r[2], r[3], r[4], r[5], r[6]};
}

template <bool EXACT>
__global__ __launch_bounds__(256) void face_region_kernel(const double *__restrict__ u, Grid g, Phys ph,
                                                          unsigned char *__restrict__ rx,
                                                          unsigned char *__restrict__ ry)
{
    const double inv_gm1 = 1.0 / (ph.gamma - 1.0);
    const long long total = (long long)g.nx * g.ny;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(k / g.ny), j = (int)(k % g.ny);
        int cx = 0xFF, cy = 0xFF;
        if (i >= kGhost - 1 && i <= g.nx - kGhost - 1 && j >= kGhost && j < g.ny - kGhost) {
            Cell c[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) c[q] = cell_at<EXACT>(u, g, ph, i - 1 + q, j);
            double f[4][7];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double t[7] = {c[q].rho, c[q].u, c[q].v, c[q].w, c[q].byc, c[q].bz, c[q].px};
#pragma unroll
                for (int v = 0; v < 7; ++v) f[q][v] = t[v];
            }
            Face7 L, R;
            face_states(f[0], f[1], f[2], f[3], ph.eps, L, R);
            hlld<EXACT>(L, R, 0.5 * (c[1].bxc + c[2].bxc), ph, inv_gm1, &cx);
        }
        if (i >= kGhost && i < g.nx - kGhost && j >= kGhost - 1 && j <= g.ny - kGhost - 1) {
            Cell c[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) c[q] = cell_at<EXACT>(u, g, ph, i, j - 1 + q);
            double f[4][7];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                // rotated frame (u,v,w)->(v,w,u), (Bx,By,Bz)->(By,Bz,Bx)
                const double t[7] = {c[q].rho, c[q].v, c[q].w, c[q].u, c[q].bz, c[q].bxc, c[q].py};
#pragma unroll
                for (int v = 0; v < 7; ++v) f[q][v] = t[v];
            }
            Face7 L, R;
            face_states(f[0], f[1], f[2], f[3], ph.eps, L, R);
            hlld<EXACT>(L, R, 0.5 * (c[1].byc + c[2].byc), ph, inv_gm1, &cy);
        }
        rx[k] = (unsigned char)cx;
        ry[k] = (unsigned char)cy;
    }
}

// ------------------------------------------------------------- launchers
static int grid_for(long long n, int block)
{
    long long b = (n + block - 1) / block;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (int)b;
}

hipError_t launch_stage(int stage, bool exact, int scheme, const double *uin, double *uout, const double *uold,
                        const double *eta, double *side, double *aux, Ctrl *ctrl, const Grid &grid, const Phys &ph,
                        int cur, hipStream_t s)
{
    const Grid &g = grid;
    const int blocks = g.xcds * g.items_per_xcd;
#define MHD_LAUNCH_STAGE(ST, EX, SC)                                                                            \
    hipLaunchKernelGGL((rk_stage_kernel<ST, EX, SC>), dim3(blocks), dim3(64), 0, s, uin, uout, uold, eta, side, \
                       aux, ctrl, g, ph, cur)
#define MHD_LAUNCH_STAGE_EX(ST, SC) \
    do { if (exact) MHD_LAUNCH_STAGE(ST, true, SC); else MHD_LAUNCH_STAGE(ST, false, SC); } while (0)
    if (scheme == 0) {
        if (stage == 1) MHD_LAUNCH_STAGE_EX(1, 0); else MHD_LAUNCH_STAGE_EX(2, 0);
    } else if (!eta) {
        if (stage == 1) MHD_LAUNCH_STAGE_EX(1, 1); else MHD_LAUNCH_STAGE_EX(2, 1);
    } else {   // the multi-GPU scheme with getEta planes
        if (stage == 1) MHD_LAUNCH_STAGE_EX(1, 2); else MHD_LAUNCH_STAGE_EX(2, 2);
    }
#undef MHD_LAUNCH_STAGE_EX
#undef MHD_LAUNCH_STAGE
    return hipGetLastError();
}

hipError_t launch_stage_batch(int stage, bool exact, const RunDesc *runs, const int *item_run, int total_items,
                              int xcds, int items_per_xcd, int blocks_per_xcd, int cur, hipStream_t s,
                              const int *tail_run, int tail_items)
{
    const bool persist = blocks_per_xcd < items_per_xcd;
    const int blocks = xcds * blocks_per_xcd + (persist ? 0 : tail_items);
#define MHD_LAUNCH_BATCH(ST, EX)                                                                                 \
    do {                                                                                                         \
        if (persist)                                                                                             \
            hipLaunchKernelGGL((rk_stage_batch_kernel<ST, EX, true>), dim3(blocks), dim3(64), 0, s, runs, item_run, \
                               total_items, xcds, items_per_xcd, blocks_per_xcd, cur, tail_run, tail_items);     \
        else                                                                                                     \
            hipLaunchKernelGGL((rk_stage_batch_kernel<ST, EX, false>), dim3(blocks), dim3(64), 0, s, runs,        \
                               item_run, total_items, xcds, items_per_xcd, blocks_per_xcd, cur, tail_run,        \
                               tail_items);                                                                      \
    } while (0)
    if (stage == 1) {
        if (exact) MHD_LAUNCH_BATCH(1, true); else MHD_LAUNCH_BATCH(1, false);
    } else {
        if (exact) MHD_LAUNCH_BATCH(2, true); else MHD_LAUNCH_BATCH(2, false);
    }
#undef MHD_LAUNCH_BATCH
    if (persist && tail_items > 0) {
#define MHD_LAUNCH_TAIL(ST, EX)                                                                            \
    hipLaunchKernelGGL((rk_stage_tail_batch_kernel<ST, EX>), dim3(tail_items), dim3(64), 0, s, runs, tail_run, \
                       tail_items, cur)
        if (stage == 1) {
            if (exact) MHD_LAUNCH_TAIL(1, true); else MHD_LAUNCH_TAIL(1, false);
        } else {
            if (exact) MHD_LAUNCH_TAIL(2, true); else MHD_LAUNCH_TAIL(2, false);
        }
#undef MHD_LAUNCH_TAIL
    }
    return hipGetLastError();
}

hipError_t launch_bc_batch(bool exact, const RunDesc *runs, int nruns, int max_nx, int max_ny, bool stage_state,
                           bool with_dt, int slot, hipStream_t s)
{
    const dim3 gx((6 * ((max_ny + 63) / 64) + 3) / 4, nruns), gy((max_nx + 4 * kBcyCols - 1) / (4 * kBcyCols), nruns);
    const int st = stage_state ? 1 : 0, wd = with_dt ? 1 : 0;
    if (exact) {
        hipLaunchKernelGGL(bc_x_batch_kernel<true>, gx, dim3(256), 0, s, runs, st, wd, slot);
        hipLaunchKernelGGL(bc_y_batch_kernel<true>, gy, dim3(256), 0, s, runs, st, wd, slot);
    } else {
        hipLaunchKernelGGL(bc_x_batch_kernel<false>, gx, dim3(256), 0, s, runs, st, wd, slot);
        hipLaunchKernelGGL(bc_y_batch_kernel<false>, gy, dim3(256), 0, s, runs, st, wd, slot);
    }
    return hipGetLastError();
}

int stage_batch_waves_per_cu(bool exact)
{
    int a = 0, b = 0;
    if (exact) {
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, rk_stage_batch_kernel<1, true, false>, 64, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, rk_stage_batch_kernel<2, true, false>, 64, 0);
    } else {
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, rk_stage_batch_kernel<1, false, false>, 64, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, rk_stage_batch_kernel<2, false, false>, 64, 0);
    }
    const int w = a < b ? a : b;
    return w > 0 ? w : 1;
}

// 2D launch shape of the interior elementwise kernels: threads along y
// (contiguous), blocks along x folded over at most 16384 rows.
static dim3 interior_grid(const Grid &g, int extra)
{
    const int cols = g.ny - 2 * kGhost + extra;
    const int rows = g.nx - 2 * kGhost + extra;
    return dim3((cols + 255) / 256, rows < 16384 ? rows : 16384);
}

hipError_t launch_m_ec(bool smooth, const double *u, const double *eta_c, double *ec, const Grid &g, const Phys &ph,
                      hipStream_t s)
{
    if (smooth) {
        const dim3 grid((g.ny - 2 * kGhost + 255) / 256, (g.nx - 2 * kGhost + kEcSeg - 1) / kEcSeg);
        hipLaunchKernelGGL(m_ec_smooth_kernel, grid, dim3(256), 0, s, u, eta_c, ec, g, ph);
    } else {
        hipLaunchKernelGGL(m_ec_kernel, interior_grid(g, 0), dim3(256), 0, s, u, eta_c, ec, g, ph);
    }
    return hipGetLastError();
}

hipError_t launch_m_emf_ct(bool exact, const double *aux_old, const double *aux_now, const double *ec_old,
                           const double *ec_now, const double *uold, double *u, Ctrl *ctrl, int slot, const Grid &g,
                           const Phys &ph, hipStream_t s)
{
    const int waves = (g.ny - 2 * kGhost + kMCtRows - 1) / kMCtRows;
    const dim3 grid((waves + 3) / 4, (g.nx - 2 * kGhost + kMSeg - 1) / kMSeg);
    if (exact)
        hipLaunchKernelGGL((m_emf_ct_kernel<true>), grid, dim3(256), 0, s, aux_old, aux_now, ec_old, ec_now, uold, u,
                           ctrl, slot, g, ph);
    else
        hipLaunchKernelGGL((m_emf_ct_kernel<false>), grid, dim3(256), 0, s, aux_old, aux_now, ec_old, ec_now, uold, u,
                           ctrl, slot, g, ph);
    return hipGetLastError();
}

hipError_t launch_bc_x(int kind, bool exact, double *u, Ctrl *ctrl, int slot, const Grid &g, const Phys &ph,
                       hipStream_t s)
{
    const int blocks = (6 * ((g.ny + 63) / 64) + 3) / 4;   // one wavefront per ghost column and 64 rows
    if (kind == 0) {
        if (exact) hipLaunchKernelGGL((bc_x_kernel<0, true>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
        else hipLaunchKernelGGL((bc_x_kernel<0, false>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
    } else {
        if (exact) hipLaunchKernelGGL((bc_x_kernel<1, true>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
        else hipLaunchKernelGGL((bc_x_kernel<1, false>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
    }
    return hipGetLastError();
}

hipError_t launch_bc_y(int kind, bool exact, double *u, Ctrl *ctrl, int slot, const Grid &g, const Phys &ph,
                       hipStream_t s)
{
    const int blocks = (g.nx + 4 * kBcyCols - 1) / (4 * kBcyCols);   // four wavefronts of ten columns each
    if (kind == 0) {
        if (exact) hipLaunchKernelGGL((bc_y_kernel<0, true>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
        else hipLaunchKernelGGL((bc_y_kernel<0, false>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
    } else {
        if (exact) hipLaunchKernelGGL((bc_y_kernel<1, true>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
        else hipLaunchKernelGGL((bc_y_kernel<1, false>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
    }
    return hipGetLastError();
}

hipError_t launch_dt_all(bool exact, bool interior, const double *u, Ctrl *ctrl, int slot, const Grid &g,
                         const Phys &ph, hipStream_t s)
{
    hipLaunchKernelGGL(ctrl_reset_kernel, dim3(1), dim3(64), 0, s, ctrl, slot);
    const dim3 grid(g.nx < 1024 ? g.nx : 1024);
#define MHD_LAUNCH_DT(EX, IN) hipLaunchKernelGGL((dt_all_kernel<EX, IN>), grid, dim3(256), 0, s, u, ctrl, slot, g, ph)
    if (exact) { if (interior) MHD_LAUNCH_DT(true, true); else MHD_LAUNCH_DT(true, false); }
    else { if (interior) MHD_LAUNCH_DT(false, true); else MHD_LAUNCH_DT(false, false); }
#undef MHD_LAUNCH_DT
    return hipGetLastError();
}

hipError_t launch_dt_all_batch(bool exact, const RunDesc *runs, int nruns, int max_nx, int slot, hipStream_t s)
{
    hipLaunchKernelGGL(ctrl_reset_batch_kernel, dim3((nruns + 63) / 64), dim3(64), 0, s, runs, nruns, slot);
    // a few blocks per run: the batch as a whole fills the machine
    const dim3 grid(max_nx < 64 ? max_nx : 64, nruns);
    if (exact) hipLaunchKernelGGL((dt_all_batch_kernel<true>), grid, dim3(256), 0, s, runs, slot);
    else hipLaunchKernelGGL((dt_all_batch_kernel<false>), grid, dim3(256), 0, s, runs, slot);
    return hipGetLastError();
}

int stage_waves_per_cu(bool exact)
{
    int a = 0, b = 0;
    if (exact) {
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, rk_stage_kernel<1, true, 0>, 64, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, rk_stage_kernel<2, true, 0>, 64, 0);
    } else {
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, rk_stage_kernel<1, false, 0>, 64, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, rk_stage_kernel<2, false, 0>, 64, 0);
    }
    const int w = a < b ? a : b;
    return w > 0 ? w : 1;
}

hipError_t launch_unpack(const double *aos, double *u, const Grid &g, hipStream_t s)
{
    hipLaunchKernelGGL(unpack_kernel, dim3(grid_for((long long)g.nx * g.ny, 256)), dim3(256), 0, s, aos, u, g);
    return hipGetLastError();
}

hipError_t launch_pack(const double *u, double *aos, const Grid &g, hipStream_t s)
{
    hipLaunchKernelGGL(pack_kernel, dim3(grid_for((long long)g.nx * g.ny, 256)), dim3(256), 0, s, u, aos, g);
    return hipGetLastError();
}

hipError_t launch_face_regions(bool exact, const double *u, unsigned char *rx, unsigned char *ry, const Grid &g,
                               const Phys &ph, hipStream_t s)
{
    const dim3 grid(grid_for((long long)g.nx * g.ny, 256));
    if (exact) hipLaunchKernelGGL(face_region_kernel<true>, grid, dim3(256), 0, s, u, g, ph, rx, ry);
    else hipLaunchKernelGGL(face_region_kernel<false>, grid, dim3(256), 0, s, u, g, ph, rx, ry);
    return hipGetLastError();
}

hipError_t launch_scatter_plane(const double *src, double *dst, const Grid &g, hipStream_t s)
{
    hipLaunchKernelGGL(scatter_plane_kernel, dim3(grid_for((long long)g.nx * g.ny, 256)), dim3(256), 0, s, src, dst,
                       g);
    return hipGetLastError();
}

}  // namespace mhd
