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
    // per slot: the dt of the cell a sequential min_element visits first is
    // NaN, which then sticks (a NaN anywhere else is passed over)
    int first_nan[2];
    int pad;
    // dt replay (frozen dt): when dt_in is set, stage 1 of step `steps`
    // takes dt_in[steps - seq0] instead of the CFL dt; when dt_out is set,
    // it records the dt it used there.
    const double *dt_in;
    double *dt_out;
    long long seq0;
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

// IEEE f64 division a/b as the compiler expands it (v_div_scale of the
// denominator, reciprocal refined by two Newton steps, v_div_scale of the
// numerator, one correction, v_div_fmas, v_div_fixup), with the reciprocal of
// the scaled denominator computed once per denominator and reused.  A lane
// whose scaled denominator differs from the shared one (div_scale rescales
// only for extreme exponent gaps) recomputes it, so every quotient is
// bitwise the compiler's a/b.
struct SharedDen {
    double sb;   // v_div_scale(b, b, a0)
    double r;    // refined reciprocal of sb
};

__device__ __forceinline__ double div_recip_core(double sb)
{
    double r = __builtin_amdgcn_rcp(sb);
    double e = fma(-sb, r, 1.0);
    r = fma(r, e, r);
    e = fma(-sb, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ SharedDen div_prepare(double a0, double b)
{
    bool unused;
    SharedDen d;
    d.sb = __builtin_amdgcn_div_scale(a0, b, false, &unused);
    d.r = div_recip_core(d.sb);
    return d;
}

__device__ __forceinline__ double div_by(double a, double b, const SharedDen &d)
{
    bool unused, flag;
    const double sb = __builtin_amdgcn_div_scale(a, b, false, &unused);
    double r = d.r;
    const bool other = __double_as_longlong(sb) != __double_as_longlong(d.sb);
    if (__builtin_expect(__any(other), 0) && other) r = div_recip_core(sb);
    const double sa = __builtin_amdgcn_div_scale(a, b, true, &flag);
    double q = sa * r;
    const double e = fma(-sb, q, sa);
    q = __builtin_amdgcn_div_fmas(e, r, q, flag);
    return __builtin_amdgcn_div_fixup(q, b, a);
}

// The neighbouring lane's value, as __shfl_up(x, 1, 64) (lane_prev: lane
// l-1's, lane 0 keeps its own) and __shfl_down(x, 1, 64) (lane_next: lane
// l+1's, lane 63 keeps its own), by DPP wavefront shifts (wave_shr:1,
// wave_shl:1): two VALU moves instead of two LDS permutes and their wait.
__device__ __forceinline__ double dpp_shift(double x, int ctrl)
{
    const long long b = __double_as_longlong(x);
    const int lo = (int)(b & 0xffffffff), hi = (int)(b >> 32);
    int l2, h2;
    if (ctrl == 0x138) {
        l2 = __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xF, 0xF, false);
        h2 = __builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xF, 0xF, false);
    } else {
        l2 = __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xF, 0xF, false);
        h2 = __builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xF, 0xF, false);
    }
    return __longlong_as_double(((long long)h2 << 32) | (unsigned int)l2);
}
__device__ __forceinline__ double lane_prev(double x) { return dpp_shift(x, 0x138); }
__device__ __forceinline__ double lane_next(double x) { return dpp_shift(x, 0x130); }

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
// Region indicators of an interface as bits: outer-left, single-star left,
// double-star left, outer-right, single-star right, double-star right
// (P/hlld.cu:64-69).
__device__ __forceinline__ int region_bits(double SL, double S1L, double SM, double S1R, double SR)
{
    return (int)(SL > 0.0) | (int)((SL <= 0.0) && (0.0 < S1L)) << 1 | (int)((S1L <= 0.0) && (0.0 < SM)) << 2
         | (int)(SR <= 0.0) << 3 | (int)((S1R <= 0.0) && (0.0 < SR)) << 4 | (int)((SM <= 0.0) && (0.0 < S1R)) << 5;
}

// region: if not null, receives region_bits of the interface (a diagnostic;
// the stage kernels pass null and the store folds away).
__device__ __forceinline__ Flux8 hlld_exact(const Face7 &Lq, const Face7 &Rq, double bX, const Phys &ph,
                                            int *region = nullptr)
{
    const double gamma = ph.gamma, eps = ph.eps;
    const double rhoL = Lq.rho, uL = Lq.u, vL = Lq.v, wL = Lq.w, bXL = bX, bYL = Lq.by, bZL = Lq.bz, pL = Lq.p;
    const double rhoR = Rq.rho, uR = Rq.u, vR = Rq.v, wR = Rq.w, bXR = bX, bYR = Rq.by, bZR = Rq.bz, pR = Rq.p;

    // Divisions sharing a denominator reuse its reciprocal (div_by is the
    // compiler's IEEE division bit for bit).
    const double gm1 = gamma - 1.0;
    const SharedDen dg = div_prepare(pL, gm1);
    const double eL = div_by(pL, gm1, dg)
                    + 0.5 * rhoL * (uL * uL + vL * vL + wL * wL)
                    + 0.5 * (bXL * bXL + bYL * bYL + bZL * bZL);
    const double pTL = pL + 0.5 * (bXL * bXL + bYL * bYL + bZL * bZL);
    const double eR = div_by(pR, gm1, dg)
                    + 0.5 * rhoR * (uR * uR + vR * vR + wR * wR)
                    + 0.5 * (bXR * bXR + bYR * bYR + bZR * bZR);
    const double pTR = pR + 0.5 * (bXR * bXR + bYR * bYR + bZR * bZR);

    const SharedDen drL = div_prepare(gamma * pL, rhoL);
    const double csL = sqrt(div_by(gamma * pL, rhoL, drL));
    const double caL = sqrt(div_by(bXL * bXL + bYL * bYL + bZL * bZL, rhoL, drL));
    const double vaL = sqrt(div_by(bXL * bXL, rhoL, drL));
    const double cfL = sqrt(0.5 * (csL * csL + caL * caL
                     + sqrt((csL * csL + caL * caL) * (csL * csL + caL * caL)
                     - 4.0 * csL * csL * vaL * vaL)));
    const SharedDen drR = div_prepare(gamma * pR, rhoR);
    const double csR = sqrt(div_by(gamma * pR, rhoR, drR));
    const double caR = sqrt(div_by(bXR * bXR + bYR * bYR + bZR * bZR, rhoR, drR));
    const double vaR = sqrt(div_by(bXR * bXR, rhoR, drR));
    const double cfR = sqrt(0.5 * (csR * csR + caR * caR
                     + sqrt((csR * csR + caR * caR) * (csR * csR + caR * caR)
                     - 4.0 * csR * csR * vaR * vaR)));

    double SL = tmin(uL, uR) - tmax(cfL, cfR);
    double SR = tmax(uL, uR) + tmax(cfL, cfR);
    SL = tmin(SL, 0.0);
    SR = tmax(SR, 0.0);

    const double D0 = (SR - uR) * rhoR - (SL - uL) * rhoL + eps;
    const double nSM = (SR - uR) * rhoR * uR - (SL - uL) * rhoL * uL - pTR + pTL;
    const SharedDen d0 = div_prepare(nSM, D0);
    const double SM = div_by(nSM, D0, d0);
    const double pT1 = div_by((SR - uR) * rhoR * pTL - (SL - uL) * rhoL * pTR
                              + rhoL * rhoR * (SR - uR) * (SL - uL) * (uR - uL), D0, d0);

    const double DSL = SL - SM + eps;
    const double DL = rhoL * (SL - uL) * (SL - SM) - bXL * bXL + eps;
    const double nrho1L = rhoL * (SL - uL);
    const SharedDen dSL = div_prepare(nrho1L, DSL);
    const double n_v1L = bXL * bYL * (SM - uL);
    const SharedDen dL = div_prepare(n_v1L, DL);
    const double rho1L = div_by(nrho1L, DSL, dSL);
    const double u1L = SM;
    const double v1L = vL - div_by(n_v1L, DL, dL);
    const double w1L = wL - div_by(bXL * bZL * (SM - uL), DL, dL);
    const double bX1L = bXL;
    const double bY1L = div_by(bYL * (rhoL * (SL - uL) * (SL - uL) - bXL * bXL), DL, dL);
    const double bZ1L = div_by(bZL * (rhoL * (SL - uL) * (SL - uL) - bXL * bXL), DL, dL);
    const double e1L = div_by((SL - uL) * eL - pTL * uL + pT1 * SM
                              + bXL * ((uL * bXL + vL * bYL + wL * bZL) - (u1L * bX1L + v1L * bY1L + w1L * bZ1L)),
                              DSL, dSL);

    const double DSR = SR - SM + eps;
    const double DR = rhoR * (SR - uR) * (SR - SM) - bXR * bXR + eps;
    const double nrho1R = rhoR * (SR - uR);
    const SharedDen dSR = div_prepare(nrho1R, DSR);
    const double n_v1R = bXR * bYR * (SM - uR);
    const SharedDen dR = div_prepare(n_v1R, DR);
    const double rho1R = div_by(nrho1R, DSR, dSR);
    const double u1R = SM;
    const double v1R = vR - div_by(n_v1R, DR, dR);
    const double w1R = wR - div_by(bXR * bZR * (SM - uR), DR, dR);
    const double bX1R = bXR;
    const double bY1R = div_by(bYR * (rhoR * (SR - uR) * (SR - uR) - bXR * bXR), DR, dR);
    const double bZ1R = div_by(bZR * (rhoR * (SR - uR) * (SR - uR) - bXR * bXR), DR, dR);
    const double e1R = div_by((SR - uR) * eR - pTR * uR + pT1 * SM
                              + bXR * ((uR * bXR + vR * bYR + wR * bZR) - (u1R * bX1R + v1R * bY1R + w1R * bZ1R)),
                              DSR, dSR);

    const double S1L = SM - sqrt(bX1L * bX1L / rho1L);
    const double S1R = SM + sqrt(bX1R * bX1R / rho1R);

    // Region indicators (P/hlld.cu:64-69).
    const bool in_oL = (SL > 0.0);
    const bool in_mL = (SL <= 0.0) && (0.0 < S1L);
    const bool in_iL = (S1L <= 0.0) && (0.0 < SM);
    const bool in_oR = (SR <= 0.0);
    const bool in_mR = (S1R <= 0.0) && (0.0 < SR);
    const bool in_iR = (SM <= 0.0) && (0.0 < S1R);
    if (region) *region = region_bits(SL, S1L, SM, S1R, SR);

    const double sq1L = sqrt(rho1L), sq1R = sqrt(rho1R);
    const double sbx = sgn(bX1L);
    const double u2 = SM;
    const double D2 = sq1L + sq1R + eps;
    const double n_v2 = sq1L * v1L + sq1R * v1R + (bY1R - bY1L) * sbx;
    const SharedDen d2 = div_prepare(n_v2, D2);
    const double v2 = div_by(n_v2, D2, d2);
    const double w2 = div_by(sq1L * w1L + sq1R * w1R + (bZ1R - bZ1L) * sbx, D2, d2);
    const double bY2 = div_by(sq1L * bY1R + sq1R * bY1L + sqrt(rho1L * rho1R) * (v1R - v1L) * sbx, D2, d2);
    const double bZ2 = div_by(sq1L * bZ1R + sq1R * bZ1L + sqrt(rho1L * rho1R) * (w1R - w1L) * sbx, D2, d2);
    const double e2L = e1L - sq1L
                     * ((u1L * bX1L + v1L * bY1L + w1L * bZ1L) - (u2 * bXL + v2 * bY2 + w2 * bZ2))
                     * sgn(bXL);
    const double e2R = e1R + sq1R
                     * ((u1R * bX1R + v1R * bY1R + w1R * bZ1R) - (u2 * bXR + v2 * bY2 + w2 * bZ2))
                     * sgn(bXR);

    // The reference sums all six candidates times 0/1 weights.  When every
    // candidate is finite that sum equals the sum of the active candidates
    // alone.  After the clamp SL <= 0 <= SR and with S1L <= SM <= S1R, at most
    // one star region is active and the outer-right region (SR == 0) may be
    // active beside it, so one candidate is evaluated (plus the right state's
    // flux when SR == 0).  Lanes where that reasoning does not hold (a
    // non-finite star state, or an overlap the ordering excludes) take the
    // literal six-term sum below, so the result is the reference's in all cases.
    const int nmain = (int)in_mL + (int)in_iL + (int)in_iR + (int)in_mR;
    const double chk = rho1L + v1L + w1L + bY1L + bZ1L + e1L + rho1R + v1R + w1R + bY1R + bZ1R + e1R
                     + v2 + w2 + bY2 + bZ2 + e2L + e2R + SM + pT1 + eL + eR + pTL + pTR;
    const bool literal = in_oL || nmain > 1 || !(chk - chk == 0.0);

    Flux8 f;
    if (__builtin_expect(__any(literal), 0) && literal) {
        // P/hlld.cu:35-107: ((((oL*w + mL*w) + iL*w) + oR*w) + mR*w) + iR*w,
        // candidates evaluated one at a time (P/hlld.cu:537-542).
        const double woL = (double)in_oL, wmL = (double)in_mL, wiL = (double)in_iL;
        const double woR = (double)in_oR, wmR = (double)in_mR, wiR = (double)in_iR;
        Flux8 c = one_flux(rhoL, uL, vL, wL, bXL, bYL, bZL, eL, pTL);
        f.f0 = c.f0 * woL; f.f1 = c.f1 * woL; f.f2 = c.f2 * woL; f.f3 = c.f3 * woL;
        f.f4 = c.f4 * woL; f.f5 = c.f5 * woL; f.f6 = c.f6 * woL; f.f7 = c.f7 * woL;
#define MHD_ACC(C, W)                                                                       \
        f.f0 = f.f0 + C.f0 * W; f.f1 = f.f1 + C.f1 * W; f.f2 = f.f2 + C.f2 * W;             \
        f.f3 = f.f3 + C.f3 * W; f.f4 = f.f4 + C.f4 * W; f.f5 = f.f5 + C.f5 * W;             \
        f.f6 = f.f6 + C.f6 * W; f.f7 = f.f7 + C.f7 * W;
        c = one_flux(rho1L, u1L, v1L, w1L, bXL, bY1L, bZ1L, e1L, pT1);
        MHD_ACC(c, wmL)
        c = one_flux(rho1L, u2, v2, w2, bXL, bY2, bZ2, e2L, pT1);
        MHD_ACC(c, wiL)
        c = one_flux(rhoR, uR, vR, wR, bXR, bYR, bZR, eR, pTR);
        MHD_ACC(c, woR)
        c = one_flux(rho1R, u1R, v1R, w1R, bXR, bY1R, bZ1R, e1R, pT1);
        MHD_ACC(c, wmR)
        c = one_flux(rho1R, u2, v2, w2, bXR, bY2, bZ2, e2R, pT1);
        MHD_ACC(c, wiR)
#undef MHD_ACC
    } else {
        const bool left = in_mL || in_iL;
        const bool inner = in_iL || in_iR;
        const double rho = left ? rho1L : rho1R;
        const double v = inner ? v2 : (left ? v1L : v1R);
        const double w = inner ? w2 : (left ? w1L : w1R);
        const double by = inner ? bY2 : (left ? bY1L : bY1R);
        const double bz = inner ? bZ2 : (left ? bZ1L : bZ1R);
        const double e = inner ? (left ? e2L : e2R) : (left ? e1L : e1R);
        if (nmain == 1) {
            f = one_flux(rho, SM, v, w, bX, by, bz, e, pT1);
        } else {
            f.f0 = f.f1 = f.f2 = f.f3 = f.f4 = f.f5 = f.f6 = f.f7 = 0.0;
        }
        if (__builtin_expect(__any(in_oR), 0) && in_oR) {
            const Flux8 o = one_flux(rhoR, uR, vR, wR, bXR, bYR, bZR, eR, pTR);
            f.f0 = f.f0 + o.f0; f.f1 = f.f1 + o.f1; f.f2 = f.f2 + o.f2; f.f3 = f.f3 + o.f3;
            f.f4 = f.f4 + o.f4; f.f5 = f.f5 + o.f5; f.f6 = f.f6 + o.f6; f.f7 = f.f7 + o.f7;
        }
    }
    return f;
}

// calculateDtFunctor, P/resistiveMHD_2D.cu:247-274: the stored face B is
// used as if it were centred.
__device__ __forceinline__ double cell_dt_exact(double rho, double mx, double my, double mz,
                                                double bX, double bY, double bZ, double e, const Phys &ph)
{
    const SharedDen dr = div_prepare(mx, rho);
    const double u = div_by(mx, rho, dr), v = div_by(my, rho, dr), w = div_by(mz, rho, dr);
    const double p = (ph.gamma - 1.0)
                   * (e - 0.5 * rho * (u * u + v * v + w * w)
                   - 0.5 * (bX * bX + bY * bY + bZ * bZ));
    const double cs = sqrt(div_by(ph.gamma * p, rho, dr));
    const double ca = sqrt(div_by(bX * bX + bY * bY + bZ * bZ, rho, dr));
    const double maxSpeedX = fabs(u) + sqrt(cs * cs + ca * ca);
    const double maxSpeedY = fabs(v) + sqrt(cs * cs + ca * ca);
    return 1.0 / (maxSpeedX / ph.dx + maxSpeedY / ph.dy + ph.eps);
}

// ---------------------------------------------------------------------------
// Fast numerics: the same formulas with shared reciprocals (hardware
// reciprocal / reciprocal square root refined by Newton steps to within an
// ulp or two), squared wave speeds reused instead of square roots of
// squares, FMA contraction, and only the selected HLLD region's flux
// evaluated.  Results differ from the exact path by rounding only.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double rcp_fast(double x)
{
#pragma clang fp contract(fast)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// 1/sqrt(x) for x > 0
__device__ __forceinline__ double rsqrt_fast(double x)
{
#pragma clang fp contract(fast)
    double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * y;
    double g = x * y;
    double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    r = fma(-g, h, 0.5);
    h = fma(h, r, h);
    return 2.0 * h;
}

__device__ __forceinline__ double sqrt_fast(double x)
{
    return __builtin_sqrt(x);
}

__device__ __forceinline__ Flux8 hlld_fast(const Face7 &L, const Face7 &R, double bx, const Phys &ph,
                                           double inv_gm1, int *region = nullptr)
{
#pragma clang fp contract(fast)
    const double gamma = ph.gamma, eps = ph.eps;
    const double bx2 = bx * bx;

    const double irhoL = rcp_fast(L.rho), irhoR = rcp_fast(R.rho);
    const double B2L = bx2 + L.by * L.by + L.bz * L.bz;
    const double B2R = bx2 + R.by * R.by + R.bz * R.bz;
    const double eL = L.p * inv_gm1 + 0.5 * L.rho * (L.u * L.u + L.v * L.v + L.w * L.w) + 0.5 * B2L;
    const double eR = R.p * inv_gm1 + 0.5 * R.rho * (R.u * R.u + R.v * R.v + R.w * R.w) + 0.5 * B2R;
    const double pTL = L.p + 0.5 * B2L;
    const double pTR = R.p + 0.5 * B2R;

    const double cs2L = gamma * L.p * irhoL, ca2L = B2L * irhoL, va2L = bx2 * irhoL;
    const double cs2R = gamma * R.p * irhoR, ca2R = B2R * irhoR, va2R = bx2 * irhoR;
    const double aL = cs2L + ca2L, aR = cs2R + ca2R;
    const double cfL = sqrt_fast(0.5 * (aL + sqrt_fast(aL * aL - 4.0 * cs2L * va2L)));
    const double cfR = sqrt_fast(0.5 * (aR + sqrt_fast(aR * aR - 4.0 * cs2R * va2R)));

    const double cfm = tmax(cfL, cfR);
    const double SL = tmin(tmin(L.u, R.u) - cfm, 0.0);
    const double SR = tmax(tmax(L.u, R.u) + cfm, 0.0);

    const double dL = SL - L.u, dR = SR - R.u;
    const double rdL = dL * L.rho, rdR = dR * R.rho;
    const double iden = rcp_fast(rdR - rdL + eps);
    const double SM = (rdR * R.u - rdL * L.u - pTR + pTL) * iden;
    const double pT1 = (rdR * pTL - rdL * pTR + L.rho * R.rho * dR * dL * (R.u - L.u)) * iden;

    // single-star states
    const double iSLSM = rcp_fast(SL - SM + eps), iSRSM = rcp_fast(SR - SM + eps);
    const double rho1L = rdL * iSLSM, rho1R = rdR * iSRSM;
    const double iDL = rcp_fast(rdL * (SL - SM) - bx2 + eps);
    const double iDR = rcp_fast(rdR * (SR - SM) - bx2 + eps);
    const double tL = bx * (SM - L.u) * iDL, tR = bx * (SM - R.u) * iDR;
    const double v1L = L.v - L.by * tL, w1L = L.w - L.bz * tL;
    const double v1R = R.v - R.by * tR, w1R = R.w - R.bz * tR;
    const double fL = (rdL * dL - bx2) * iDL, fR = (rdR * dR - bx2) * iDR;
    const double by1L = L.by * fL, bz1L = L.bz * fL;
    const double by1R = R.by * fR, bz1R = R.bz * fR;
    const double vb1L = SM * bx + v1L * by1L + w1L * bz1L;
    const double vb1R = SM * bx + v1R * by1R + w1R * bz1R;
    const double e1L = (dL * eL - pTL * L.u + pT1 * SM + bx * ((L.u * bx + L.v * L.by + L.w * L.bz) - vb1L)) * iSLSM;
    const double e1R = (dR * eR - pTR * R.u + pT1 * SM + bx * ((R.u * bx + R.v * R.by + R.w * R.bz) - vb1R)) * iSRSM;

    const double isqL = rsqrt_fast(rho1L), isqR = rsqrt_fast(rho1R);
    const double sqL = rho1L * isqL, sqR = rho1R * isqR;
    const double abx = fabs(bx);
    const double S1L = SM - abx * isqL;
    const double S1R = SM + abx * isqR;
    if (region) *region = region_bits(SL, S1L, SM, S1R, SR);

    // double-star states
    const double s = sgn(bx);
    const double isum = rcp_fast(sqL + sqR + eps);
    const double v2 = (sqL * v1L + sqR * v1R + (by1R - by1L) * s) * isum;
    const double w2 = (sqL * w1L + sqR * w1R + (bz1R - bz1L) * s) * isum;
    const double sqLR = sqL * sqR;
    const double by2 = (sqL * by1R + sqR * by1L + sqLR * (v1R - v1L) * s) * isum;
    const double bz2 = (sqL * bz1R + sqR * bz1L + sqLR * (w1R - w1L) * s) * isum;
    const double vb2 = SM * bx + v2 * by2 + w2 * bz2;
    const double e2L = e1L - sqL * (vb1L - vb2) * s;
    const double e2R = e1R + sqR * (vb1R - vb2) * s;

    // Region of the interface (P/hlld.cu:64-69).  With SL <= 0 <= SR at most
    // one of the star regions holds; the outer-right region (SR == 0) may hold
    // together with one of them, and its flux is then added, as the
    // reference's indicator sum does.  A NaN speed poisons the flux, as a NaN
    // candidate does in the reference's sum.
    const bool left = SM > 0.0;
    const bool inner = left ? !(S1L > 0.0) : (S1R > 0.0);
    const bool main_on = left || (S1R > 0.0) || (SR > 0.0);
    const double rho = left ? rho1L : rho1R;
    const double v = inner ? v2 : (left ? v1L : v1R);
    const double w = inner ? w2 : (left ? w1L : w1R);
    const double by = inner ? by2 : (left ? by1L : by1R);
    const double bz = inner ? bz2 : (left ? bz1L : bz1R);
    const double e = inner ? (left ? e2L : e2R) : (left ? e1L : e1R);
    const double m = main_on ? 1.0 : 0.0;
    Flux8 f = one_flux(rho, SM, v, w, bx, by, bz, e, pT1);
    f.f0 *= m; f.f1 *= m; f.f2 *= m; f.f3 *= m; f.f5 *= m; f.f6 *= m; f.f7 *= m;
    if (SR <= 0.0) {
        const Flux8 o = one_flux(R.rho, R.u, R.v, R.w, bx, R.by, R.bz, eR, pTR);
        f.f0 += o.f0; f.f1 += o.f1; f.f2 += o.f2; f.f3 += o.f3; f.f5 += o.f5; f.f6 += o.f6; f.f7 += o.f7;
    }
    if (!(S1L == S1L) || !(S1R == S1R) || !(SM == SM)) {
        const double nan = __builtin_nan("");
        f.f0 = f.f1 = f.f2 = f.f3 = f.f5 = f.f6 = f.f7 = nan;
    }
    return f;
}

__device__ __forceinline__ double cell_dt_fast(double rho, double mx, double my, double mz,
                                               double bX, double bY, double bZ, double e, const Phys &ph)
{
#pragma clang fp contract(fast)
    const double ir = rcp_fast(rho);
    const double u = mx * ir, v = my * ir, w = mz * ir;
    const double B2 = bX * bX + bY * bY + bZ * bZ;
    const double p = (ph.gamma - 1.0) * (e - 0.5 * rho * (u * u + v * v + w * w) - 0.5 * B2);
    const double c = sqrt_fast((ph.gamma * p + B2) * ir);
    return rcp_fast((fabs(u) + c) / ph.dx + (fabs(v) + c) / ph.dy + ph.eps);
}

template <bool EXACT>
__device__ __forceinline__ Flux8 hlld(const Face7 &L, const Face7 &R, double bx, const Phys &ph, double inv_gm1,
                                      int *region = nullptr)
{
    if (EXACT) return hlld_exact(L, R, bx, ph, region);
    return hlld_fast(L, R, bx, ph, inv_gm1, region);
}

template <bool EXACT>
__device__ __forceinline__ double cell_dt(double rho, double mx, double my, double mz, double bX, double bY,
                                          double bZ, double e, const Phys &ph)
{
    if (EXACT) return cell_dt_exact(rho, mx, my, mz, bX, bY, bZ, e, ph);
    return cell_dt_fast(rho, mx, my, mz, bX, bY, bZ, e, ph);
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
