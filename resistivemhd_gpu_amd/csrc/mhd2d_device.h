// mhd2d_device.h -- device-side numerics of the resistive-MHD RK2 step for
// gfx950: primitive variables, minmod-limited MUSCL, the HLLD interface flux,
// the per-cell CFL time step and wavefront reductions.
//
// Arithmetic follows the reference's expressions operation for operation
// (paths relative to the reference root,
// P = ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_periodicXY), so
// that, built without floating-point contraction, the results round exactly
// as the reference's do.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mhd {

struct Phys {
    double gamma, dx, dy, eps, cfl, eta;
};

// Device-resident step control block: dt reduction slots, dt of the step in
// flight, accumulated time and the crash flag.
struct Ctrl {
    unsigned long long acc[2];   // running min of per-cell dt, as the bits of a non-negative double
    double dt;                   // dt of the current / last step
    double time;                 // accumulated time
    long long steps;             // completed steps
    int crashed;                 // NaN found in e after the last step
    int pad;
};

constexpr unsigned long long kInfBits = 0x7FF0000000000000ull;

struct Flux8 {
    double f0, f1, f2, f3, f4, f5, f6, f7;
};

// thrust::min / thrust::max (lhs, rhs): min picks rhs only if rhs < lhs,
// max picks rhs only if lhs < rhs.
__device__ __forceinline__ double tmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double tmax(double a, double b) { return (a < b) ? b : a; }

// P/minmod.hpp:6-16: sign(x) * max(min(|x|, sign(x)*y), EPS), sign as int.
__device__ __forceinline__ double minmod(double x, double y, double eps)
{
    int s = (x > 0) - (x < 0);
    return s * tmax(tmin(fabs(x), s * y), eps);
}

// P/hlld.cu:26-32
__device__ __forceinline__ double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }

// P/hlld.cu:448-467 getOneFlux
__device__ __forceinline__ Flux8 one_flux(double rho, double u, double v, double w,
                                          double bX, double bY, double bZ, double e, double pT)
{
    Flux8 f;
    f.f0 = rho * u;
    f.f1 = rho * u * u + pT - bX * bX;
    f.f2 = rho * u * v - bX * bY;
    f.f3 = rho * u * w - bX * bZ;
    f.f4 = 0.0;
    f.f5 = u * bY - v * bX;
    f.f6 = u * bZ - w * bX;
    f.f7 = (e + pT) * u - bX * (bX * u + bY * v + bZ * w);
    return f;
}

// Interface state in the normal frame: rho, u (normal), v, w, bY, bZ, p.
struct Face7 {
    double rho, u, v, w, by, bz, p;
};

// HLLD flux at one interface with normal field bX (both sides share it):
// parameters P/hlld.cu:256-433, candidates P/hlld.cu:469-549, selection as
// the sum of six indicator-weighted candidates P/hlld.cu:35-107.
__device__ __forceinline__ Flux8 hlld(const Face7 &Lq, const Face7 &Rq, double bX, const Phys &ph)
{
    const double gamma = ph.gamma, eps = ph.eps;
    const double rhoL = Lq.rho, uL = Lq.u, vL = Lq.v, wL = Lq.w, bXL = bX, bYL = Lq.by, bZL = Lq.bz, pL = Lq.p;
    const double rhoR = Rq.rho, uR = Rq.u, vR = Rq.v, wR = Rq.w, bXR = bX, bYR = Rq.by, bZR = Rq.bz, pR = Rq.p;

    const double eL = pL / (gamma - 1.0)
                    + 0.5 * rhoL * (uL * uL + vL * vL + wL * wL)
                    + 0.5 * (bXL * bXL + bYL * bYL + bZL * bZL);
    const double pTL = pL + 0.5 * (bXL * bXL + bYL * bYL + bZL * bZL);
    const double eR = pR / (gamma - 1.0)
                    + 0.5 * rhoR * (uR * uR + vR * vR + wR * wR)
                    + 0.5 * (bXR * bXR + bYR * bYR + bZR * bZR);
    const double pTR = pR + 0.5 * (bXR * bXR + bYR * bYR + bZR * bZR);

    const double csL = sqrt(gamma * pL / rhoL);
    const double caL = sqrt((bXL * bXL + bYL * bYL + bZL * bZL) / rhoL);
    const double vaL = sqrt(bXL * bXL / rhoL);
    const double cfL = sqrt(0.5 * (csL * csL + caL * caL
                     + sqrt((csL * csL + caL * caL) * (csL * csL + caL * caL)
                     - 4.0 * csL * csL * vaL * vaL)));
    const double csR = sqrt(gamma * pR / rhoR);
    const double caR = sqrt((bXR * bXR + bYR * bYR + bZR * bZR) / rhoR);
    const double vaR = sqrt(bXR * bXR / rhoR);
    const double cfR = sqrt(0.5 * (csR * csR + caR * caR
                     + sqrt((csR * csR + caR * caR) * (csR * csR + caR * caR)
                     - 4.0 * csR * csR * vaR * vaR)));

    double SL = tmin(uL, uR) - tmax(cfL, cfR);
    double SR = tmax(uL, uR) + tmax(cfL, cfR);
    SL = tmin(SL, 0.0);
    SR = tmax(SR, 0.0);

    const double SM = ((SR - uR) * rhoR * uR - (SL - uL) * rhoL * uL - pTR + pTL)
                    / ((SR - uR) * rhoR - (SL - uL) * rhoL + eps);
    const double pT1 = ((SR - uR) * rhoR * pTL - (SL - uL) * rhoL * pTR
                     + rhoL * rhoR * (SR - uR) * (SL - uL) * (uR - uL))
                     / ((SR - uR) * rhoR - (SL - uL) * rhoL + eps);

    const double rho1L = rhoL * (SL - uL) / (SL - SM + eps);
    const double u1L = SM;
    const double v1L = vL - bXL * bYL * (SM - uL) / (rhoL * (SL - uL) * (SL - SM) - bXL * bXL + eps);
    const double w1L = wL - bXL * bZL * (SM - uL) / (rhoL * (SL - uL) * (SL - SM) - bXL * bXL + eps);
    const double bX1L = bXL;
    const double bY1L = bYL * (rhoL * (SL - uL) * (SL - uL) - bXL * bXL)
                      / (rhoL * (SL - uL) * (SL - SM) - bXL * bXL + eps);
    const double bZ1L = bZL * (rhoL * (SL - uL) * (SL - uL) - bXL * bXL)
                      / (rhoL * (SL - uL) * (SL - SM) - bXL * bXL + eps);
    const double e1L = ((SL - uL) * eL - pTL * uL + pT1 * SM
                     + bXL * ((uL * bXL + vL * bYL + wL * bZL) - (u1L * bX1L + v1L * bY1L + w1L * bZ1L)))
                     / (SL - SM + eps);

    const double rho1R = rhoR * (SR - uR) / (SR - SM + eps);
    const double u1R = SM;
    const double v1R = vR - bXR * bYR * (SM - uR) / (rhoR * (SR - uR) * (SR - SM) - bXR * bXR + eps);
    const double w1R = wR - bXR * bZR * (SM - uR) / (rhoR * (SR - uR) * (SR - SM) - bXR * bXR + eps);
    const double bX1R = bXR;
    const double bY1R = bYR * (rhoR * (SR - uR) * (SR - uR) - bXR * bXR)
                      / (rhoR * (SR - uR) * (SR - SM) - bXR * bXR + eps);
    const double bZ1R = bZR * (rhoR * (SR - uR) * (SR - uR) - bXR * bXR)
                      / (rhoR * (SR - uR) * (SR - SM) - bXR * bXR + eps);
    const double e1R = ((SR - uR) * eR - pTR * uR + pT1 * SM
                     + bXR * ((uR * bXR + vR * bYR + wR * bZR) - (u1R * bX1R + v1R * bY1R + w1R * bZ1R)))
                     / (SR - SM + eps);

    const double S1L = SM - sqrt(bX1L * bX1L / rho1L);
    const double S1R = SM + sqrt(bX1R * bX1R / rho1R);

    // Region indicators (P/hlld.cu:64-69).  After the clamp SL <= 0 <= SR,
    // so the outer-left region is never selected.
    const bool in_oL = (SL > 0.0);
    const bool in_mL = (SL <= 0.0) && (0.0 < S1L);
    const bool in_iL = (S1L <= 0.0) && (0.0 < SM);
    const bool in_oR = (SR <= 0.0);
    const bool in_mR = (S1R <= 0.0) && (0.0 < SR);
    const bool in_iR = (SM <= 0.0) && (0.0 < S1R);

    const double sq1L = sqrt(rho1L), sq1R = sqrt(rho1R);
    const double sbx = sgn(bX1L);
    const double u2 = SM;
    const double v2 = (sq1L * v1L + sq1R * v1R + (bY1R - bY1L) * sbx) / (sq1L + sq1R + eps);
    const double w2 = (sq1L * w1L + sq1R * w1R + (bZ1R - bZ1L) * sbx) / (sq1L + sq1R + eps);
    const double bY2 = (sq1L * bY1R + sq1R * bY1L + sqrt(rho1L * rho1R) * (v1R - v1L) * sbx)
                     / (sq1L + sq1R + eps);
    const double bZ2 = (sq1L * bZ1R + sq1R * bZ1L + sqrt(rho1L * rho1R) * (w1R - w1L) * sbx)
                     / (sq1L + sq1R + eps);
    const double e2L = e1L - sq1L
                     * ((u1L * bX1L + v1L * bY1L + w1L * bZ1L) - (u2 * bXL + v2 * bY2 + w2 * bZ2))
                     * sgn(bXL);
    const double e2R = e1R + sq1R
                     * ((u1R * bX1R + v1R * bY1R + w1R * bZ1R) - (u2 * bXR + v2 * bY2 + w2 * bZ2))
                     * sgn(bXR);

    const Flux8 oL = one_flux(rhoL, uL, vL, wL, bXL, bYL, bZL, eL, pTL);
    const Flux8 mL = one_flux(rho1L, u1L, v1L, w1L, bXL, bY1L, bZ1L, e1L, pT1);
    const Flux8 iL = one_flux(rho1L, u2, v2, w2, bXL, bY2, bZ2, e2L, pT1);
    const Flux8 oR = one_flux(rhoR, uR, vR, wR, bXR, bYR, bZR, eR, pTR);
    const Flux8 mR = one_flux(rho1R, u1R, v1R, w1R, bXR, bY1R, bZ1R, e1R, pT1);
    const Flux8 iR = one_flux(rho1R, u2, v2, w2, bXR, bY2, bZ2, e2R, pT1);

    const double woL = (double)in_oL, wmL = (double)in_mL, wiL = (double)in_iL;
    const double woR = (double)in_oR, wmR = (double)in_mR, wiR = (double)in_iR;
#define MHD_SEL(c) (oL.c * woL + mL.c * wmL + iL.c * wiL + oR.c * woR + mR.c * wmR + iR.c * wiR)
    Flux8 f;
    f.f0 = MHD_SEL(f0); f.f1 = MHD_SEL(f1); f.f2 = MHD_SEL(f2); f.f3 = MHD_SEL(f3);
    f.f4 = MHD_SEL(f4); f.f5 = MHD_SEL(f5); f.f6 = MHD_SEL(f6); f.f7 = MHD_SEL(f7);
#undef MHD_SEL
    return f;
}

// calculateDtFunctor, P/resistiveMHD_2D.cu:247-274: the stored face B is
// used as if it were centred.
__device__ __forceinline__ double cell_dt(double rho, double mx, double my, double mz,
                                          double bX, double bY, double bZ, double e, const Phys &ph)
{
    const double u = mx / rho, v = my / rho, w = mz / rho;
    const double p = (ph.gamma - 1.0)
                   * (e - 0.5 * rho * (u * u + v * v + w * w)
                   - 0.5 * (bX * bX + bY * bY + bZ * bZ));
    const double cs = sqrt(ph.gamma * p / rho);
    const double ca = sqrt((bX * bX + bY * bY + bZ * bZ) / rho);
    const double maxSpeedX = fabs(u) + sqrt(cs * cs + ca * ca);
    const double maxSpeedY = fabs(v) + sqrt(cs * cs + ca * ca);
    return 1.0 / (maxSpeedX / ph.dx + maxSpeedY / ph.dy + ph.eps);
}

// Wavefront (64-lane) reductions.
__device__ __forceinline__ double wave_min(double x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double y = __shfl_xor(x, off, 64);
        x = (y < x) ? y : x;
    }
    return x;
}

__device__ __forceinline__ void accumulate_dt(Ctrl *ctrl, int slot, double local_min)
{
    // local_min is +inf when no valid cell contributed; NaN cells never lower it.
    double m = wave_min(local_min);
    if ((threadIdx.x & 63) == 0 && m < __longlong_as_double((long long)kInfBits))
        atomicMin(&ctrl->acc[slot], (unsigned long long)__double_as_longlong(m));
}

}  // namespace mhd
