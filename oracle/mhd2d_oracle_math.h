/* mhd2d_oracle_math.h -- per-element arithmetic of the CPU oracle (test
 * infrastructure only): the reference's state, primitive and flux types,
 * minmod, the primitives, MUSCL, the HLLD parameters / candidates /
 * selection, the CFL dt of a cell and the two RK stage updates, each
 * restated from the cited reference lines (paths relative to the reference
 * root, P = ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_periodicXY).
 * Shared by the pass-by-pass restatement in mhd2d_oracle.c and the Thrust
 * OpenMP twin of the reference's pass structure (thrust_twin.cpp), so both
 * round every element the same way.  C99 and C++. */
#ifndef MHD2D_ORACLE_MATH_H
#define MHD2D_ORACLE_MATH_H
#include <math.h>

typedef struct { double rho, rhoU, rhoV, rhoW, bX, bY, bZ, e; } Cons;   /* P/conservation_parameter_struct.hpp:5-27 */
typedef struct { double rho, u, v, w, bX, bY, bZ, p; } Prim;            /* P/basic_parameter_struct.hpp:5-27 */
typedef struct { double f0, f1, f2, f3, f4, f5, f6, f7; } Flux;         /* P/flux_struct.hpp:5-27 */

/* thrust::min / thrust::max argument order (lhs, rhs): min returns rhs only
 * if rhs < lhs, max returns rhs only if lhs < rhs. */
static inline double tmin(double a, double b) { return (b < a) ? b : a; }
static inline double tmax(double a, double b) { return (a < b) ? b : a; }

/* ---------------------------------------------------------------- minmod */
/* P/minmod.hpp:6-16.  Pinned by the closed-form truth table (Q1) in tests. */
static inline double minmod(double x, double y, double eps)
{
    int sign_x = (x > 0) - (x < 0);
    double abs_x = fabs(x);
    return sign_x * tmax(tmin(abs_x, sign_x * y), eps);
}

/* --------------------------------------------------- primitive variables */
/* P/calculate_half_Q.cu:4-38 (GetBasicParamterFunctor).  Pinned by the
 * conservative->primitive round trip test. */
static inline Prim basic_param(const Cons *c, const Cons *cp1, double gamma)
{
    Prim q;
    double rho = c->rho;
    double u = c->rhoU / rho;
    double v = c->rhoV / rho;
    double w = c->rhoW / rho;
    double bX = c->bX, bY = c->bY, bZ = c->bZ, e = c->e;
    double p = (gamma - 1.0)
             * (e - 0.5 * (rho * (u * u + v * v + w * w))
             - 0.5 * (bX * bX + bY * bY + bZ * bZ));
    q.rho = rho; q.u = u; q.v = v; q.w = w;
    q.bX = 0.5 * (bX + cp1->bX);
    q.bY = bY; q.bZ = bZ; q.p = p;
    return q;
}

/* --------------------------------------------------------------- MUSCL */
/* P/muscl.cu:6-29 LeftParameterFunctor */
static inline Prim left_param(const Prim *m1, const Prim *q, const Prim *p1, double eps)
{
    Prim l;
    l.rho = q->rho + 0.5 * minmod(q->rho - m1->rho, p1->rho - q->rho, eps);
    l.u   = q->u   + 0.5 * minmod(q->u   - m1->u,   p1->u   - q->u,   eps);
    l.v   = q->v   + 0.5 * minmod(q->v   - m1->v,   p1->v   - q->v,   eps);
    l.w   = q->w   + 0.5 * minmod(q->w   - m1->w,   p1->w   - q->w,   eps);
    l.bY  = q->bY  + 0.5 * minmod(q->bY  - m1->bY,  p1->bY  - q->bY,  eps);
    l.bZ  = q->bZ  + 0.5 * minmod(q->bZ  - m1->bZ,  p1->bZ  - q->bZ,  eps);
    l.p   = q->p   + 0.5 * minmod(q->p   - m1->p,   p1->p   - q->p,   eps);
    l.bX  = q->bX;
    return l;
}

/* P/muscl.cu:67-90 RightParameterFunctor */
static inline Prim right_param(const Prim *q, const Prim *p1, const Prim *p2, double eps)
{
    Prim r;
    r.rho = p1->rho - 0.5 * minmod(p1->rho - q->rho, p2->rho - p1->rho, eps);
    r.u   = p1->u   - 0.5 * minmod(p1->u   - q->u,   p2->u   - p1->u,   eps);
    r.v   = p1->v   - 0.5 * minmod(p1->v   - q->v,   p2->v   - p1->v,   eps);
    r.w   = p1->w   - 0.5 * minmod(p1->w   - q->w,   p2->w   - p1->w,   eps);
    r.bY  = p1->bY  - 0.5 * minmod(p1->bY  - q->bY,  p2->bY  - p1->bY,  eps);
    r.bZ  = p1->bZ  - 0.5 * minmod(p1->bZ  - q->bZ,  p2->bZ  - p1->bZ,  eps);
    r.p   = p1->p   - 0.5 * minmod(p1->p   - q->p,   p2->p   - p1->p,   eps);
    r.bX  = q->bX;
    return r;
}

/* ---------------------------------------------------------------- HLLD */
/* P/hlld.cu:26-32 */
static inline double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }

/* P/hlld.cu:448-467 getOneFlux */
static inline Flux one_flux(double rho, double u, double v, double w,
                            double bX, double bY, double bZ, double e, double pT)
{
    Flux f;
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


/* The HLLD parameters of one interface (the reference's HLLDParameter,
 * P/hlld_parameter_struct.hpp:5-108): the outer states with their energy
 * and total pressure, the wave speeds, the single-star and double-star
 * states. */
typedef struct {
    double rhoL, uL, vL, wL, bXL, bYL, bZL, eL, pTL;
    double rhoR, uR, vR, wR, bXR, bYR, bZR, eR, pTR;
    double SL, SR, SM, S1L, S1R;
    double rho1L, u1L, v1L, w1L, bY1L, bZ1L, e1L, pT1L;
    double rho1R, u1R, v1R, w1R, bY1R, bZ1R, e1R, pT1R;
    double rho2L, rho2R, u2, v2, w2, bY2, bZ2, e2L, e2R, pT2L, pT2R;
} HlldParams;

/* The six candidate fluxes of one interface (P/hlld.cu:469-549):
 * outer, single-star and double-star states, left and right. */
typedef struct { Flux oL, mL, iL, oR, mR, iR; } HlldCandidates;

/* calculateHLLDParameterFunctor (P/hlld.cu:256-433), with the SL <= 0 <= SR
 * clamp at :315-318 and EPS in every denominator. */
static inline HlldParams hlld_params(const Prim *L, const Prim *R, double gamma, double eps)
{
    double rhoL = L->rho, uL = L->u, vL = L->v, wL = L->w, bXL = L->bX, bYL = L->bY, bZL = L->bZ, pL = L->p;
    double rhoR = R->rho, uR = R->u, vR = R->v, wR = R->w, bXR = R->bX, bYR = R->bY, bZR = R->bZ, pR = R->p;
    double eL = pL / (gamma - 1.0)
              + 0.5 * rhoL * (uL * uL + vL * vL + wL * wL)
              + 0.5 * (bXL * bXL + bYL * bYL + bZL * bZL);
    double pTL = pL + 0.5 * (bXL * bXL + bYL * bYL + bZL * bZL);
    double eR = pR / (gamma - 1.0)
              + 0.5 * rhoR * (uR * uR + vR * vR + wR * wR)
              + 0.5 * (bXR * bXR + bYR * bYR + bZR * bZR);
    double pTR = pR + 0.5 * (bXR * bXR + bYR * bYR + bZR * bZR);

    double csL = sqrt(gamma * pL / rhoL);
    double caL = sqrt((bXL * bXL + bYL * bYL + bZL * bZL) / rhoL);
    double vaL = sqrt(bXL * bXL / rhoL);
    double cfL = sqrt(0.5 * (csL * csL + caL * caL
               + sqrt((csL * csL + caL * caL) * (csL * csL + caL * caL)
               - 4.0 * csL * csL * vaL * vaL)));
    double csR = sqrt(gamma * pR / rhoR);
    double caR = sqrt((bXR * bXR + bYR * bYR + bZR * bZR) / rhoR);
    double vaR = sqrt(bXR * bXR / rhoR);
    double cfR = sqrt(0.5 * (csR * csR + caR * caR
               + sqrt((csR * csR + caR * caR) * (csR * csR + caR * caR)
               - 4.0 * csR * csR * vaR * vaR)));

    double SL = tmin(uL, uR) - tmax(cfL, cfR);
    double SR = tmax(uL, uR) + tmax(cfL, cfR);
    SL = tmin(SL, 0.0);
    SR = tmax(SR, 0.0);

    double SM = ((SR - uR) * rhoR * uR - (SL - uL) * rhoL * uL - pTR + pTL)
              / ((SR - uR) * rhoR - (SL - uL) * rhoL + eps);
    double pT1 = ((SR - uR) * rhoR * pTL - (SL - uL) * rhoL * pTR
               + rhoL * rhoR * (SR - uR) * (SL - uL) * (uR - uL))
               / ((SR - uR) * rhoR - (SL - uL) * rhoL + eps);
    double pT1L = pT1, pT1R = pT1;

    double rho1L = rhoL * (SL - uL) / (SL - SM + eps);
    double u1L = SM;
    double v1L = vL - bXL * bYL * (SM - uL) / (rhoL * (SL - uL) * (SL - SM) - bXL * bXL + eps);
    double w1L = wL - bXL * bZL * (SM - uL) / (rhoL * (SL - uL) * (SL - SM) - bXL * bXL + eps);
    double bX1L = bXL;
    double bY1L = bYL * (rhoL * (SL - uL) * (SL - uL) - bXL * bXL)
                / (rhoL * (SL - uL) * (SL - SM) - bXL * bXL + eps);
    double bZ1L = bZL * (rhoL * (SL - uL) * (SL - uL) - bXL * bXL)
                / (rhoL * (SL - uL) * (SL - SM) - bXL * bXL + eps);
    double e1L = ((SL - uL) * eL - pTL * uL + pT1 * SM
               + bXL * ((uL * bXL + vL * bYL + wL * bZL) - (u1L * bX1L + v1L * bY1L + w1L * bZ1L)))
               / (SL - SM + eps);

    double rho1R = rhoR * (SR - uR) / (SR - SM + eps);
    double u1R = SM;
    double v1R = vR - bXR * bYR * (SM - uR) / (rhoR * (SR - uR) * (SR - SM) - bXR * bXR + eps);
    double w1R = wR - bXR * bZR * (SM - uR) / (rhoR * (SR - uR) * (SR - SM) - bXR * bXR + eps);
    double bX1R = bXR;
    double bY1R = bYR * (rhoR * (SR - uR) * (SR - uR) - bXR * bXR)
                / (rhoR * (SR - uR) * (SR - SM) - bXR * bXR + eps);
    double bZ1R = bZR * (rhoR * (SR - uR) * (SR - uR) - bXR * bXR)
                / (rhoR * (SR - uR) * (SR - SM) - bXR * bXR + eps);
    double e1R = ((SR - uR) * eR - pTR * uR + pT1 * SM
               + bXR * ((uR * bXR + vR * bYR + wR * bZR) - (u1R * bX1R + v1R * bY1R + w1R * bZ1R)))
               / (SR - SM + eps);

    double S1L = SM - sqrt(bX1L * bX1L / rho1L);
    double S1R = SM + sqrt(bX1R * bX1R / rho1R);

    double rho2L = rho1L, rho2R = rho1R, u2 = SM;
    double v2 = (sqrt(rho1L) * v1L + sqrt(rho1R) * v1R + (bY1R - bY1L) * sgn(bX1L))
              / (sqrt(rho1L) + sqrt(rho1R) + eps);
    double w2 = (sqrt(rho1L) * w1L + sqrt(rho1R) * w1R + (bZ1R - bZ1L) * sgn(bX1L))
              / (sqrt(rho1L) + sqrt(rho1R) + eps);
    double bY2 = (sqrt(rho1L) * bY1R + sqrt(rho1R) * bY1L + sqrt(rho1L * rho1R) * (v1R - v1L) * sgn(bX1L))
               / (sqrt(rho1L) + sqrt(rho1R) + eps);
    double bZ2 = (sqrt(rho1L) * bZ1R + sqrt(rho1R) * bZ1L + sqrt(rho1L * rho1R) * (w1R - w1L) * sgn(bX1L))
               / (sqrt(rho1L) + sqrt(rho1R) + eps);
    double e2L = e1L - sqrt(rho1L)
               * ((u1L * bX1L + v1L * bY1L + w1L * bZ1L) - (u2 * bXL + v2 * bY2 + w2 * bZ2))
               * sgn(bXL);
    double e2R = e1R + sqrt(rho1R)
               * ((u1R * bX1R + v1R * bY1R + w1R * bZ1R) - (u2 * bXR + v2 * bY2 + w2 * bZ2))
               * sgn(bXR);
    double pT2L = pT1L, pT2R = pT1R;

    HlldParams P;
    P.rhoL = rhoL; P.uL = uL; P.vL = vL; P.wL = wL; P.bXL = bXL; P.bYL = bYL; P.bZL = bZL; P.eL = eL; P.pTL = pTL;
    P.rhoR = rhoR; P.uR = uR; P.vR = vR; P.wR = wR; P.bXR = bXR; P.bYR = bYR; P.bZR = bZR; P.eR = eR; P.pTR = pTR;
    P.SL = SL; P.SR = SR; P.SM = SM; P.S1L = S1L; P.S1R = S1R;
    P.rho1L = rho1L; P.u1L = u1L; P.v1L = v1L; P.w1L = w1L; P.bY1L = bY1L; P.bZ1L = bZ1L; P.e1L = e1L;
    P.pT1L = pT1L;
    P.rho1R = rho1R; P.u1R = u1R; P.v1R = v1R; P.w1R = w1R; P.bY1R = bY1R; P.bZ1R = bZ1R; P.e1R = e1R;
    P.pT1R = pT1R;
    P.rho2L = rho2L; P.rho2R = rho2R; P.u2 = u2; P.v2 = v2; P.w2 = w2; P.bY2 = bY2; P.bZ2 = bZ2;
    P.e2L = e2L; P.e2R = e2R; P.pT2L = pT2L; P.pT2R = pT2R;
    (void)bX1L; (void)bX1R;
    return P;
}

/* setFluxFunctor (P/hlld.cu:469-549): getOneFlux of each state. */
static inline HlldCandidates hlld_candidates(const HlldParams *P)
{
    HlldCandidates C;
    C.oL = one_flux(P->rhoL, P->uL, P->vL, P->wL, P->bXL, P->bYL, P->bZL, P->eL, P->pTL);
    C.mL = one_flux(P->rho1L, P->u1L, P->v1L, P->w1L, P->bXL, P->bY1L, P->bZ1L, P->e1L, P->pT1L);
    C.iL = one_flux(P->rho2L, P->u2, P->v2, P->w2, P->bXL, P->bY2, P->bZ2, P->e2L, P->pT2L);
    C.oR = one_flux(P->rhoR, P->uR, P->vR, P->wR, P->bXR, P->bYR, P->bZR, P->eR, P->pTR);
    C.mR = one_flux(P->rho1R, P->u1R, P->v1R, P->w1R, P->bXR, P->bY1R, P->bZ1R, P->e1R, P->pT1R);
    C.iR = one_flux(P->rho2R, P->u2, P->v2, P->w2, P->bXR, P->bY2, P->bZ2, P->e2R, P->pT2R);
    return C;
}

/* CalculateFluxFunctor (P/hlld.cu:35-107): the sum of the six candidates
 * weighted by the region indicators. */
static inline Flux hlld_select(const HlldParams *P, const HlldCandidates *C)
{
    double w_oL = (double)(P->SL > 0.0);
    double w_mL = (double)((P->SL <= 0.0) && (0.0 < P->S1L));
    double w_iL = (double)((P->S1L <= 0.0) && (0.0 < P->SM));
    double w_oR = (double)(P->SR <= 0.0);
    double w_mR = (double)((P->S1R <= 0.0) && (0.0 < P->SR));
    double w_iR = (double)((P->SM <= 0.0) && (0.0 < P->S1R));
#define SEL(c) (C->oL.c * w_oL + C->mL.c * w_mL + C->iL.c * w_iL + C->oR.c * w_oR + C->mR.c * w_mR + C->iR.c * w_iR)
    Flux f;
    f.f0 = SEL(f0); f.f1 = SEL(f1); f.f2 = SEL(f2); f.f3 = SEL(f3);
    f.f4 = SEL(f4); f.f5 = SEL(f5); f.f6 = SEL(f6); f.f7 = SEL(f7);
#undef SEL
    return f;
}

/* One HLLD interface flux from MUSCL states L and R.
 * Parameters: P/hlld.cu:256-433 (calculateHLLDParameterFunctor, with the
 * SL<=0<=SR clamp at :315-318 and EPS in every denominator).  Candidates:
 * P/hlld.cu:469-549 (setFluxFunctor).  Selection: P/hlld.cu:35-107 (sum of
 * the six indicator-weighted candidates).  The reference stores the
 * parameters and candidates in arrays between passes; every element is
 * computed from (dQLeft[k], dQRight[k]) alone, so computing them here in one
 * function gives the same values.  Pinned by consistency F(U,U) = F_phys(U),
 * x<->y rotation equivariance, the Bx=0 reduction and mirror symmetry. */
static inline Flux hlld_flux(const Prim *L, const Prim *R, double gamma, double eps)
{
    HlldParams P = hlld_params(L, R, gamma, eps);
    HlldCandidates C = hlld_candidates(&P);
    return hlld_select(&P, &C);
}

/* calculateDtFunctor: P/resistiveMHD_2D.cu:247-274 (stored face B is used as
 * if centred).  Pinned by the uniform-state closed form. */
static inline double cell_dt(const Cons *c, double gamma, double dx, double dy, double eps)
{
    double rho = c->rho;
    double u = c->rhoU / rho, v = c->rhoV / rho, w = c->rhoW / rho;
    double bX = c->bX, bY = c->bY, bZ = c->bZ, e = c->e;
    double p = (gamma - 1.0)
             * (e - 0.5 * rho * (u * u + v * v + w * w)
             - 0.5 * (bX * bX + bY * bY + bZ * bZ));
    double cs = sqrt(gamma * p / rho);
    double ca = sqrt((bX * bX + bY * bY + bZ * bZ) / rho);
    double maxSpeedX = fabs(u) + sqrt(cs * cs + ca * ca);
    double maxSpeedY = fabs(v) + sqrt(cs * cs + ca * ca);
    return 1.0 / (maxSpeedX / dx + maxSpeedY / dy + eps);
}

/* oneStepFirstFunctor: P/resistiveMHD_2D.cu:53-86 */
static inline Cons stage_first(const Cons *U, const Flux *F, const Flux *Fm, const Flux *G, const Flux *Gm,
                               double dt, double dx, double dy)
{
    Cons n;
    n.rho  = U->rho  - dt / dx * (F->f0 - Fm->f0) - dt / dy * (G->f0 - Gm->f0);
    n.rhoU = U->rhoU - dt / dx * (F->f1 - Fm->f1) - dt / dy * (G->f1 - Gm->f1);
    n.rhoV = U->rhoV - dt / dx * (F->f2 - Fm->f2) - dt / dy * (G->f2 - Gm->f2);
    n.rhoW = U->rhoW - dt / dx * (F->f3 - Fm->f3) - dt / dy * (G->f3 - Gm->f3);
    n.bX   = U->bX   - dt / dx * (F->f4 - Fm->f4) - dt / dy * (G->f4 - Gm->f4);
    n.bY   = U->bY   - dt / dx * (F->f5 - Fm->f5) - dt / dy * (G->f5 - Gm->f5);
    n.bZ   = U->bZ   - dt / dx * (F->f6 - Fm->f6) - dt / dy * (G->f6 - Gm->f6);
    n.e    = U->e    - dt / dx * (F->f7 - Fm->f7) - dt / dy * (G->f7 - Gm->f7);
    return n;
}

/* oneStepSecondFunctor: P/resistiveMHD_2D.cu:88-130 */
static inline Cons stage_second(const Cons *U, const Cons *B, const Flux *F, const Flux *Fm,
                                const Flux *G, const Flux *Gm, double dt, double dx, double dy)
{
    Cons n;
    n.rho  = 0.5 * (U->rho  + B->rho  - dt / dx * (F->f0 - Fm->f0) - dt / dy * (G->f0 - Gm->f0));
    n.rhoU = 0.5 * (U->rhoU + B->rhoU - dt / dx * (F->f1 - Fm->f1) - dt / dy * (G->f1 - Gm->f1));
    n.rhoV = 0.5 * (U->rhoV + B->rhoV - dt / dx * (F->f2 - Fm->f2) - dt / dy * (G->f2 - Gm->f2));
    n.rhoW = 0.5 * (U->rhoW + B->rhoW - dt / dx * (F->f3 - Fm->f3) - dt / dy * (G->f3 - Gm->f3));
    n.bX   = 0.5 * (U->bX   + B->bX   - dt / dx * (F->f4 - Fm->f4) - dt / dy * (G->f4 - Gm->f4));
    n.bY   = 0.5 * (U->bY   + B->bY   - dt / dx * (F->f5 - Fm->f5) - dt / dy * (G->f5 - Gm->f5));
    n.bZ   = 0.5 * (U->bZ   + B->bZ   - dt / dx * (F->f6 - Fm->f6) - dt / dy * (G->f6 - Gm->f6));
    n.e    = 0.5 * (U->e    + B->e    - dt / dx * (F->f7 - Fm->f7) - dt / dy * (G->f7 - Gm->f7));
    return n;
}


#endif
