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
// loads one 64-row slice of the 8 planes (coalesced), computes the x-face
// HLLD + resistive flux, the y-face flux of the previous column, the corner
// EMF, and writes the updated cell two columns behind the load front.
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

template <int STAGE, bool EXACT>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(EXACT ? MHD_WAVES_PER_EU_EXACT : MHD_WAVES_PER_EU_FAST)))
void rk_stage_kernel(const double *__restrict__ uin, double *uout,
                                                      const double *uold, const double *__restrict__ etaf,
                                                      double *__restrict__ side, Ctrl *__restrict__ ctrl,
                                                      Grid g, Phys ph, int cur)
{
    const int lane = threadIdx.x;
    // Workgroups are dealt to the XCDs round-robin (block b runs on XCD
    // b % xcds).  Each XCD gets a contiguous range of items (strips fastest),
    // so neighbouring strips and segments, which read the same halo lines,
    // share that XCD's L2.
    const int item = (blockIdx.x % g.xcds) * g.items_per_xcd + blockIdx.x / g.xcds;
    if (item >= g.nstrips * g.nsegs) return;
    const int sy = item % g.nstrips;
    const int sx = item / g.nstrips;
    const int j0 = kGhost + sy * kStripRows;
    const int i0 = kGhost + sx * g.seg;
    const int iend = min(i0 + g.seg, g.nx - kGhost);
    const int jend = min(j0 + kStripRows, g.ny - kGhost);
    const int jr = j0 - 2 + lane;              // row held by this lane
    const int jl = min(jr, g.ny - 1);          // clamped row for loads (rows past ny-1 are unused)
    const bool upd_row = (jr >= j0) && (jr < jend);
    const bool stale_row = (sy == 0) && (jr == kGhost - 1);

    // dt: ResistiveMHD2D::calculateDt (P/resistiveMHD_2D.cu:277-291), from
    // the per-cell minimum reduced at the end of the previous step.
    double dt;
    if (STAGE == 1) {
        const double m = __longlong_as_double((long long)ctrl->acc[cur]);
        const double lim = fmin(ph.dx * ph.dx / ph.eta, ph.dy * ph.dy / ph.eta);
        dt = fmin(m, lim) * ph.cfl;
        if (item == 0 && lane == 0) {
            ctrl->dt = dt;
            ctrl->acc[cur ^ 1] = kInfBits;
            ctrl->crashed = 0;
        }
    } else {
        dt = ctrl->dt;
        if (item == 0 && lane == 0) {
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

    double bx_prev = Bx[(size_t)(i0 - 3) * g.pitch + jl];
    Cell A{}, B{};                                   // cells c-2 and c-1
    Face7 Lx{};                                      // left state of face c-2+1/2
    Flux8 Fprev{}, Gprev{};                          // F(c-3+1/2), G(c-2)
    double Eprev = 0.0;                              // E at corner (c-3+1/2, j+1/2)
    double rJY = 0.0, rJZ = 0.0, rJYBZ = 0.0, rJZBY = 0.0;   // resistive terms of cell c-2
    double dtloc = __longlong_as_double((long long)kInfBits);
    int nanflag = 0;

    double eta_prev = 0.0;   // resistivity of cell c-1

    for (int c = i0 - 2; c <= iend + 1; ++c) {
        const size_t o = (size_t)c * g.pitch + jl;
        const double rho = Rh[o], mx = Mx[o], my = My[o], mz = Mz[o];
        const double bx = Bx[o], by = By[o], bz = Bz[o], e = En[o];
        const double bym = By[o - 1];
        const double eta_c = etaf ? etaf[o] : ph.eta;
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
            const double bxc_up = __shfl_down(B.bxc, 1, 64), bxc_dn = __shfl_up(B.bxc, 1, 64);
            const double bz_up = __shfl_down(B.bz, 1, 64), bz_dn = __shfl_up(B.bz, 1, 64);
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
                F.f5 -= 0.5 * (rJZ + eJZ);
                F.f6 += 0.5 * (rJY + eJY);
                F.f7 += 0.5 * (rJYBZ + eJYBZ) - 0.5 * (rJZBY + eJZBY);
            }
            Lx.rho = B.rho + 0.5 * sr;
            Lx.u = B.u + 0.5 * su;
            Lx.v = B.v + 0.5 * sv;
            Lx.w = B.w + 0.5 * sw;
            Lx.by = B.byc + 0.5 * sby;
            Lx.bz = B.bz + 0.5 * sbz;
            Lx.p = B.px + 0.5 * sp;
            rJY = eJY; rJZ = eJZ; rJYBZ = eJYBZ; rJZBY = eJZBY;

            // y-face flux of column c-1 at faces j+1/2 (HLLD::calculateFluxG,
            // P/hlld.cu:209-233): rotated frame (u,v,w)->(v,w,u),
            // (Bx,By,Bz)->(By,Bz,Bx); neighbours along y come from other lanes.
            Flux8 G;
            {
                const double q0 = B.rho, q1 = B.v, q2 = B.w, q3 = B.u, q4 = B.bz, q5 = B.bxc, q6 = B.py;
                const double m0 = __shfl_up(q0, 1, 64), m1 = __shfl_up(q1, 1, 64), m2 = __shfl_up(q2, 1, 64);
                const double m3 = __shfl_up(q3, 1, 64), m4 = __shfl_up(q4, 1, 64), m5 = __shfl_up(q5, 1, 64);
                const double m6 = __shfl_up(q6, 1, 64);
                const double p0 = __shfl_down(q0, 1, 64), p1 = __shfl_down(q1, 1, 64), p2 = __shfl_down(q2, 1, 64);
                const double p3 = __shfl_down(q3, 1, 64), p4 = __shfl_down(q4, 1, 64), p5 = __shfl_down(q5, 1, 64);
                const double p6 = __shfl_down(q6, 1, 64);
                const double t0 = minmod(q0 - m0, p0 - q0, ph.eps), t1 = minmod(q1 - m1, p1 - q1, ph.eps);
                const double t2 = minmod(q2 - m2, p2 - q2, ph.eps), t3 = minmod(q3 - m3, p3 - q3, ph.eps);
                const double t4 = minmod(q4 - m4, p4 - q4, ph.eps), t5 = minmod(q5 - m5, p5 - q5, ph.eps);
                const double t6 = minmod(q6 - m6, p6 - q6, ph.eps);
                Face7 Lg, Rg;
                Lg.rho = q0 + 0.5 * t0; Lg.u = q1 + 0.5 * t1; Lg.v = q2 + 0.5 * t2; Lg.w = q3 + 0.5 * t3;
                Lg.by = q4 + 0.5 * t4; Lg.bz = q5 + 0.5 * t5; Lg.p = q6 + 0.5 * t6;
                Rg.rho = __shfl_down(q0 - 0.5 * t0, 1, 64);
                Rg.u = __shfl_down(q1 - 0.5 * t1, 1, 64);
                Rg.v = __shfl_down(q2 - 0.5 * t2, 1, 64);
                Rg.w = __shfl_down(q3 - 0.5 * t3, 1, 64);
                Rg.by = __shfl_down(q4 - 0.5 * t4, 1, 64);
                Rg.bz = __shfl_down(q5 - 0.5 * t5, 1, 64);
                Rg.p = __shfl_down(q6 - 0.5 * t6, 1, 64);
                const double bn = 0.5 * (B.byc + __shfl_down(B.byc, 1, 64));
                const Flux8 Gr = hlld<EXACT>(Lg, Rg, bn, ph, inv_gm1);
                // rotate back (P/hlld.cu:170-194)
                G.f0 = Gr.f0; G.f1 = Gr.f3; G.f2 = Gr.f1; G.f3 = Gr.f2;
                G.f4 = Gr.f6; G.f5 = Gr.f4; G.f6 = Gr.f5; G.f7 = Gr.f7;
                const double eJZ1 = __shfl_down(eJZ, 1, 64), eJX1 = __shfl_down(eJX, 1, 64);
                const double eJZBX1 = __shfl_down(eJZBX, 1, 64), eJXBZ1 = __shfl_down(eJXBZ, 1, 64);
                G.f4 += 0.5 * (eJZ + eJZ1);
                G.f6 -= 0.5 * (eJX + eJX1);
                G.f7 += 0.5 * (eJZBX + eJZBX1) - 0.5 * (eJXBZ + eJXBZ1);
            }

            // corner EMF at (c-2+1/2, j+1/2), P/ct.cu:22-41 (stage 1 only: the
            // stage-2 CT reuses stage-1 fluxes, P/resistiveMHD_2D.cu:167-168,190)
            double E = 0.0;
            if (STAGE == 1 && c >= i0 + 1) {
                const double f5up = __shfl_down(F.f5, 1, 64);
                E = 0.25 * (Gprev.f4 + G.f4 + (-1.0 * F.f5) + (-1.0 * f5up));
            }

            if (c >= i0 + 2) {
                const int s = c - 2;
                const double g0m = __shfl_up(Gprev.f0, 1, 64), g1m = __shfl_up(Gprev.f1, 1, 64);
                const double g2m = __shfl_up(Gprev.f2, 1, 64), g3m = __shfl_up(Gprev.f3, 1, 64);
                const double g6m = __shfl_up(Gprev.f6, 1, 64), g7m = __shfl_up(Gprev.f7, 1, 64);
                const double Em = __shfl_up(E, 1, 64);
                if (s < iend && (upd_row || (STAGE == 1 && stale_row) || (STAGE == 2 && stale_row))) {
                    const size_t os = (size_t)s * g.pitch + jl;
                    double n0, n1, n2, n3, n4, n5, n6, n7;
                    if (STAGE == 1) {
                        // oneStepFirstFunctor (P/resistiveMHD_2D.cu:53-86) and
                        // CT_kernel (P/ct.cu:44-59)
                        const double U0 = Rh[os], U1 = Mx[os], U2 = My[os], U3 = Mz[os];
                        const double U4 = Bx[os], U5 = By[os], U6 = Bz[os], U7 = En[os];
                        n0 = U0 - dtdx * (F.f0 - Fprev.f0) - dtdy * (Gprev.f0 - g0m);
                        n1 = U1 - dtdx * (F.f1 - Fprev.f1) - dtdy * (Gprev.f1 - g1m);
                        n2 = U2 - dtdx * (F.f2 - Fprev.f2) - dtdy * (Gprev.f2 - g2m);
                        n3 = U3 - dtdx * (F.f3 - Fprev.f3) - dtdy * (Gprev.f3 - g3m);
                        n4 = U4 - dtdy * (E - Em);
                        n5 = U5 + dtdx * (E - Eprev);
                        n6 = U6 - dtdx * (F.f6 - Fprev.f6) - dtdy * (Gprev.f6 - g6m);
                        n7 = U7 - dtdx * (F.f7 - Fprev.f7) - dtdy * (Gprev.f7 - g7m);
                        if (upd_row) {
                            PL(uout, 0)[os] = n0; PL(uout, 1)[os] = n1; PL(uout, 2)[os] = n2; PL(uout, 3)[os] = n3;
                            PL(uout, 4)[os] = n4; PL(uout, 5)[os] = n5; PL(uout, 6)[os] = n6; PL(uout, 7)[os] = n7;
                        } else {
                            // CT value of the bottom boundary face, read by the
                            // symmetric-x boundary before the y boundary
                            // overwrites it (S/boundary.cu:56-75).
                            PL(uout, 5)[os] = n5;
                            side[s] = n5;
                        }
                    } else {
                        if (upd_row) {
                            // oneStepSecondFunctor (P/resistiveMHD_2D.cu:88-130);
                            // B is the stage-1 CT value (same EMF, same old B).
                            const double S0 = Rh[os], S1 = Mx[os], S2 = My[os], S3 = Mz[os];
                            const double S4 = Bx[os], S5 = By[os], S6 = Bz[os], S7 = En[os];
                            const double O0 = PL(uold, 0)[os], O1 = PL(uold, 1)[os], O2 = PL(uold, 2)[os];
                            const double O3 = PL(uold, 3)[os], O6 = PL(uold, 6)[os], O7 = PL(uold, 7)[os];
                            n0 = 0.5 * (O0 + S0 - dtdx * (F.f0 - Fprev.f0) - dtdy * (Gprev.f0 - g0m));
                            n1 = 0.5 * (O1 + S1 - dtdx * (F.f1 - Fprev.f1) - dtdy * (Gprev.f1 - g1m));
                            n2 = 0.5 * (O2 + S2 - dtdx * (F.f2 - Fprev.f2) - dtdy * (Gprev.f2 - g2m));
                            n3 = 0.5 * (O3 + S3 - dtdx * (F.f3 - Fprev.f3) - dtdy * (Gprev.f3 - g3m));
                            n4 = S4;
                            n5 = S5;
                            n6 = 0.5 * (O6 + S6 - dtdx * (F.f6 - Fprev.f6) - dtdy * (Gprev.f6 - g6m));
                            n7 = 0.5 * (O7 + S7 - dtdx * (F.f7 - Fprev.f7) - dtdy * (Gprev.f7 - g7m));
                            PL(uout, 0)[os] = n0; PL(uout, 1)[os] = n1; PL(uout, 2)[os] = n2; PL(uout, 3)[os] = n3;
                            PL(uout, 4)[os] = n4; PL(uout, 5)[os] = n5; PL(uout, 6)[os] = n6; PL(uout, 7)[os] = n7;
                            // fused epilogue: per-cell dt of the new state for
                            // the next step, and the crash flag
                            const double d = cell_dt<EXACT>(n0, n1, n2, n3, n4, n5, n6, n7, ph);
                            if (d < dtloc) dtloc = d;
                            nanflag |= isnan(n7) ? 1 : 0;
                        } else {
                            PL(uout, 5)[os] = side[s];
                        }
                    }
                }
            }
            Fprev = F;
            Gprev = G;
            Eprev = E;
        }
        A = B;
        B = C;
        eta_prev = eta_c;
    }

    if (STAGE == 2) {
        accumulate_dt(ctrl, cur ^ 1, dtloc);
        const unsigned long long any = __ballot(nanflag);
        if (lane == 0 && any) atomicOr(&ctrl->crashed, 1);
    }
}

// ------------------------------------------------------------ boundaries
// Periodic x (P/boundary.cu:4-71) or symmetric x (S/boundary.cu:4-104), one
// thread per row.  The symmetric bX rebuild reads By of the interior columns
// the mirror copied from (identical values), so one pass suffices.
template <int KIND, bool EXACT>
__global__ __launch_bounds__(256) void bc_x_kernel(double *u, Ctrl *ctrl, int slot, Grid g, Phys ph)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double dloc = __longlong_as_double((long long)kInfBits);
    if (j < g.ny) {
        const int nx = g.nx;
        const int dst[6] = {0, 1, 2, nx - 3, nx - 2, nx - 1};
        int src[6];
        if (KIND == 0) { src[0] = nx - 6; src[1] = nx - 5; src[2] = nx - 4; src[3] = 3; src[4] = 4; src[5] = 5; }
        else           { src[0] = 5; src[1] = 4; src[2] = 3; src[3] = nx - 4; src[4] = nx - 5; src[5] = nx - 6; }
        double cell[6][8];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                if (KIND == 1 && v == 4) continue;
                cell[q][v] = PL(u, v)[(size_t)src[q] * g.pitch + j];
                PL(u, v)[(size_t)dst[q] * g.pitch + j] = cell[q][v];
            }
        }
        if (KIND == 1) {
            const double *by = PL(u, 5);
            double *bx = PL(u, 4);
            if (j > 0) {
                // By of ghost columns 2, 1, 0 equals that of columns 3, 4, 5
                const double dby3 = by[(size_t)3 * g.pitch + j] - by[(size_t)3 * g.pitch + j - 1];
                const double dby4 = by[(size_t)4 * g.pitch + j] - by[(size_t)4 * g.pitch + j - 1];
                const double dby5 = by[(size_t)5 * g.pitch + j] - by[(size_t)5 * g.pitch + j - 1];
                cell[2][4] = bx[(size_t)3 * g.pitch + j] + dby3 / ph.dy * ph.dx;
                cell[1][4] = cell[2][4] + dby3 / ph.dy * ph.dx;
                cell[0][4] = cell[1][4] + dby4 / ph.dy * ph.dx;
                (void)dby5;
                const double dbyA = by[(size_t)(nx - 4) * g.pitch + j] - by[(size_t)(nx - 4) * g.pitch + j - 1];
                const double dbyB = by[(size_t)(nx - 5) * g.pitch + j] - by[(size_t)(nx - 5) * g.pitch + j - 1];
                const double dbyC = by[(size_t)(nx - 6) * g.pitch + j] - by[(size_t)(nx - 6) * g.pitch + j - 1];
                cell[3][4] = bx[(size_t)(nx - 4) * g.pitch + j] - dbyA / ph.dy * ph.dx;
                cell[4][4] = cell[3][4] - dbyB / ph.dy * ph.dx;
                cell[5][4] = cell[4][4] - dbyC / ph.dy * ph.dx;
            } else {
                cell[2][4] = bx[(size_t)3 * g.pitch + j];
                cell[1][4] = cell[2][4];
                cell[0][4] = cell[1][4];
                cell[3][4] = bx[(size_t)(nx - 4) * g.pitch + j];
                cell[4][4] = cell[3][4];
                cell[5][4] = cell[4][4];
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) bx[(size_t)dst[q] * g.pitch + j] = cell[q][4];
        }
        if (ctrl && j >= kGhost && j < g.ny - kGhost) {
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double d = cell_dt<EXACT>(cell[q][0], cell[q][1], cell[q][2], cell[q][3], cell[q][4], cell[q][5],
                                         cell[q][6], cell[q][7], ph);
                if (d < dloc) dloc = d;
            }
        }
    }
    if (ctrl) accumulate_dt(ctrl, slot, dloc);
}

// Periodic y (P/boundary.cu:74-141) or symmetric y (S/boundary.cu:107-205),
// one thread per column, after the x boundary.
template <int KIND, bool EXACT>
__global__ __launch_bounds__(256) void bc_y_kernel(double *u, Ctrl *ctrl, int slot, Grid g, Phys ph)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double dloc = __longlong_as_double((long long)kInfBits);
    if (i < g.nx) {
        const int ny = g.ny;
        const size_t row = (size_t)i * g.pitch;
        const int dst[6] = {0, 1, 2, ny - 3, ny - 2, ny - 1};
        int src[6];
        if (KIND == 0) { src[0] = ny - 6; src[1] = ny - 5; src[2] = ny - 4; src[3] = 3; src[4] = 4; src[5] = 5; }
        else           { src[0] = 5; src[1] = 4; src[2] = 3; src[3] = ny - 4; src[4] = ny - 5; src[5] = ny - 6; }
        double cell[6][8];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                if (KIND == 1 && v == 5) continue;
                cell[q][v] = PL(u, v)[row + src[q]];
                PL(u, v)[row + dst[q]] = cell[q][v];
            }
        }
        if (KIND == 1) {
            const double *bx = PL(u, 4);
            double *by = PL(u, 5);
            if (i > 0) {
                // Bx of ghost rows 2, 1, 0 equals that of rows 3, 4, 5 (and
                // rows ny-3.. that of ny-4..), column i-1 included
                const size_t rm = (size_t)(i - 1) * g.pitch;
                const double d3 = bx[row + 3] - bx[rm + 3];
                const double d4 = bx[row + 4] - bx[rm + 4];
                const double dA = bx[row + ny - 4] - bx[rm + ny - 4];
                const double dB = bx[row + ny - 5] - bx[rm + ny - 5];
                const double dC = bx[row + ny - 6] - bx[rm + ny - 6];
                cell[2][5] = by[row + 3] + d3 / ph.dx * ph.dy;
                cell[1][5] = cell[2][5] + d3 / ph.dx * ph.dy;
                cell[0][5] = cell[1][5] + d4 / ph.dx * ph.dy;
                cell[3][5] = by[row + ny - 4] - dA / ph.dx * ph.dy;
                cell[4][5] = cell[3][5] - dB / ph.dx * ph.dy;
                cell[5][5] = cell[4][5] - dC / ph.dx * ph.dy;
            } else {
                cell[2][5] = by[row + 3];
                cell[1][5] = cell[2][5];
                cell[0][5] = cell[1][5];
                cell[3][5] = by[row + ny - 4];
                cell[4][5] = cell[3][5];
                cell[5][5] = cell[4][5];
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) by[row + dst[q]] = cell[q][5];
        }
        if (ctrl) {
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double d = cell_dt<EXACT>(cell[q][0], cell[q][1], cell[q][2], cell[q][3], cell[q][4], cell[q][5],
                                         cell[q][6], cell[q][7], ph);
                if (d < dloc) dloc = d;
            }
        }
    }
    if (ctrl) accumulate_dt(ctrl, slot, dloc);
}

// Per-cell dt over all nx*ny cells (P/resistiveMHD_2D.cu:277-286) and the
// NaN scan of e, for a state set from outside the step.
template <bool EXACT>
__global__ __launch_bounds__(256) void dt_all_kernel(const double *__restrict__ u, Ctrl *ctrl, int slot, Grid g,
                                                     Phys ph)
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
            if (d < dloc) dloc = d;
            nanflag |= isnan(e) ? 1 : 0;   // checkCalculationIsCrashed, P/resistiveMHD_2D.cu:294-309
        }
    }
    accumulate_dt(ctrl, slot, dloc);
    const unsigned long long any = __ballot(nanflag);
    if ((threadIdx.x & 63) == 0 && any) atomicOr(&ctrl->crashed, 1);
}

__global__ void ctrl_reset_kernel(Ctrl *ctrl, int slot)
{
    if (threadIdx.x == 0) {
        ctrl->acc[slot] = kInfBits;
        ctrl->crashed = 0;
    }
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

// ------------------------------------------------------------- launchers
static int grid_for(long long n, int block)
{
    long long b = (n + block - 1) / block;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (int)b;
}

hipError_t launch_stage(int stage, bool exact, const double *uin, double *uout, const double *uold,
                        const double *eta, double *side, Ctrl *ctrl, const Grid &g, const Phys &ph, int cur,
                        hipStream_t s)
{
    const int blocks = g.xcds * g.items_per_xcd;
#define MHD_LAUNCH_STAGE(ST, EX)                                                                               \
    hipLaunchKernelGGL((rk_stage_kernel<ST, EX>), dim3(blocks), dim3(64), 0, s, uin, uout, uold, eta, side, ctrl, \
                       g, ph, cur)
    if (stage == 1) {
        if (exact) MHD_LAUNCH_STAGE(1, true); else MHD_LAUNCH_STAGE(1, false);
    } else {
        if (exact) MHD_LAUNCH_STAGE(2, true); else MHD_LAUNCH_STAGE(2, false);
    }
#undef MHD_LAUNCH_STAGE
    return hipGetLastError();
}

hipError_t launch_bc_x(int kind, bool exact, double *u, Ctrl *ctrl, int slot, const Grid &g, const Phys &ph,
                       hipStream_t s)
{
    const int blocks = (g.ny + 255) / 256;
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
    const int blocks = (g.nx + 255) / 256;
    if (kind == 0) {
        if (exact) hipLaunchKernelGGL((bc_y_kernel<0, true>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
        else hipLaunchKernelGGL((bc_y_kernel<0, false>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
    } else {
        if (exact) hipLaunchKernelGGL((bc_y_kernel<1, true>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
        else hipLaunchKernelGGL((bc_y_kernel<1, false>), dim3(blocks), dim3(256), 0, s, u, ctrl, slot, g, ph);
    }
    return hipGetLastError();
}

hipError_t launch_dt_all(bool exact, const double *u, Ctrl *ctrl, int slot, const Grid &g, const Phys &ph,
                         hipStream_t s)
{
    hipLaunchKernelGGL(ctrl_reset_kernel, dim3(1), dim3(64), 0, s, ctrl, slot);
    const dim3 grid(g.nx < 1024 ? g.nx : 1024);
    if (exact) hipLaunchKernelGGL((dt_all_kernel<true>), grid, dim3(256), 0, s, u, ctrl, slot, g, ph);
    else hipLaunchKernelGGL((dt_all_kernel<false>), grid, dim3(256), 0, s, u, ctrl, slot, g, ph);
    return hipGetLastError();
}

int stage_waves_per_cu(bool exact)
{
    int a = 0, b = 0;
    if (exact) {
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, rk_stage_kernel<1, true>, 64, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, rk_stage_kernel<2, true>, 64, 0);
    } else {
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, rk_stage_kernel<1, false>, 64, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, rk_stage_kernel<2, false>, 64, 0);
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

hipError_t launch_scatter_plane(const double *src, double *dst, const Grid &g, hipStream_t s)
{
    hipLaunchKernelGGL(scatter_plane_kernel, dim3(grid_for((long long)g.nx * g.ny, 256)), dim3(256), 0, s, src, dst,
                       g);
    return hipGetLastError();
}

}  // namespace mhd
