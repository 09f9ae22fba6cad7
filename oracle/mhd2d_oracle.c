/*
 * mhd2d_oracle.c -- CPU restatement of the reference's single-GPU FP64
 * resistive-MHD RK2 step.  TEST INFRASTRUCTURE ONLY: this file is the
 * reference result that tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg compare the HIP library against.  It is never linked into, called by
 * or shipped as part of the product library.
 *
 * It follows the reference's pass structure literally: array-of-structs
 * state of nx*ny cells (3 ghost layers per side included in nx, ny), index
 * k = j + i*ny (y fastest), the same persistent intermediate arrays and the
 * same element ranges for every pass, so that values the reference leaves in
 * ghost cells (stale intermediates, wrap-around at column ends) are the same
 * here.  Build with -O2 -ffp-contract=off so every operation rounds exactly
 * as written.
 *
 * Reference files (paths relative to the reference root):
 *   P = ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_periodicXY
 *   S = ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_symmetricXY
 *   F32 = ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_single
 *   EX = ResistiveMHD_singleGPU_thrust/examples2D
 *
 * Pinning (which independent check pins each function) is listed in the
 * header comment of each function; tests/test_oracle.py holds the checks.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mhd2d_oracle.h"
#include "mhd2d_oracle_math.h"
#ifdef _OPENMP
#include <omp.h>
#endif

struct oracle_sim {
    int nx, ny, N;
    double dx, dy, gamma, cfl, eps, eta;
    int bc_x, bc_y;
    double dt, time;
    double dt_fixed;   /* > 0: every step uses this dt (tests of the flux and CT parts alone) */
    long steps;
    /* ResistiveMHD2D members, P/resistiveMHD_2D.hpp:11-23 */
    Cons *U, *UBar;
    Flux *fluxF, *fluxG;
    double *dtVector, *bXOld, *bYOld, *tmpVector;
    /* FluxSolver / HLLD members, P/flux_solver.hpp:9-10, P/hlld.hpp:7-23 */
    Flux *solverFlux, *hlldFlux;
    Prim *dQCenter, *dQLeft, *dQRight;
    Cons *tmpUForFluxG;
    /* CT members, P/ct.hpp:9-12 */
    Flux *oldFluxF, *oldFluxG;
    double *eZVector;
    /* face B of UBar right after the stage-1 CT (a snapshot for the tests;
     * the reference keeps no copy) */
    double *bXStage1, *bYStage1;
    /* per-cell resistivity sampled at cell (i, j): plays the role of the
     * example's getEta(i*dx, j*dy) (EX/plasmoid/main_plasmoid.cu:159-172) */
    double *etaField;
    /* the MPI library's getEta at its three sampling positions, per index
     * (ghosts included): x-face ((i + 0.5) dx, j dy) (M/flux_solver.cu:84,94),
     * y-face (i dx, (j + 0.5) dy) (M/flux_solver.cu:168,179) and cell
     * (i dx, j dy) (M/ct.cu:41,56); NULL: the uniform eta */
    double *etaXFaceM, *etaYFaceM, *etaCellM;
    /* scheme: ORACLE_SCHEME_SINGLE (the single-GPU library) or
     * ORACLE_SCHEME_MULTI_SHARE (the MPI library's step on one rank) */
    int scheme;
    /* multi-GPU library CT arrays, M/ct.cu:4-23 */
    double *oNF5, *oNG4, *oF5, *oG4, *oNF0, *oNG0;
    double *nNF5, *nNG4, *nF5, *nG4, *nNF0, *nNG0;
};

double oracle_minmod(double x, double y, double eps) { return minmod(x, y, eps); }

void oracle_hlld_flux(const double *left8, const double *right8, double gamma, double eps, double *flux8)
{
    Prim L, R;
    memcpy(&L, left8, sizeof L);
    memcpy(&R, right8, sizeof R);
    Flux f = hlld_flux(&L, &R, gamma, eps);
    memcpy(flux8, &f, sizeof f);
}

/* ------------------------------------------------------ flux assembly */
/* HLLD::setQX + calculateFluxF: P/hlld.cu:110-131,236-243 with the ranges
 * of P/calculate_half_Q.cu:41-53 and P/muscl.cu:32-42,88-102. */
static void hlld_flux_x(struct oracle_sim *s, const Cons *U)
{
    int N = s->N, ny = s->ny;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < N - ny; ++k)
        s->dQCenter[k] = basic_param(&U[k], &U[k + ny], s->gamma);
    #pragma omp parallel for schedule(static)
    for (int k = ny; k < N - ny; ++k)
        s->dQLeft[k] = left_param(&s->dQCenter[k - ny], &s->dQCenter[k], &s->dQCenter[k + ny], s->eps);
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < N - 2 * ny; ++k)
        s->dQRight[k] = right_param(&s->dQCenter[k], &s->dQCenter[k + ny], &s->dQCenter[k + 2 * ny], s->eps);
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < N; ++k)
        s->hlldFlux[k] = hlld_flux(&s->dQLeft[k], &s->dQRight[k], s->gamma, s->eps);
}

/* HLLD::calculateFluxG: P/hlld.cu:134-233.  The state is rotated
 * (u,v,w)->(v,w,u), (Bx,By,Bz)->(By,Bz,Bx), swept with stride 1 (which wraps
 * across columns at j = 0 and ny-1, as in the reference), and the flux is
 * rotated back. */
static void hlld_flux_y(struct oracle_sim *s, const Cons *U)
{
    int N = s->N;
    Cons *t = s->tmpUForFluxG;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < N; ++k) {                         /* P/hlld.cu:134-152 */
        t[k].rho = U[k].rho;
        t[k].rhoU = U[k].rhoV; t[k].rhoV = U[k].rhoW; t[k].rhoW = U[k].rhoU;
        t[k].bX = U[k].bY; t[k].bY = U[k].bZ; t[k].bZ = U[k].bX;
        t[k].e = U[k].e;
    }
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < N - 1; ++k)                       /* P/calculate_half_Q.cu:56-67 */
        s->dQCenter[k] = basic_param(&t[k], &t[k + 1], s->gamma);
    #pragma omp parallel for schedule(static)
    for (int k = 1; k < N - 1; ++k)                       /* P/muscl.cu:45-59 */
        s->dQLeft[k] = left_param(&s->dQCenter[k - 1], &s->dQCenter[k], &s->dQCenter[k + 1], s->eps);
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < N - 2; ++k)                       /* P/muscl.cu:106-126 */
        s->dQRight[k] = right_param(&s->dQCenter[k], &s->dQCenter[k + 1], &s->dQCenter[k + 2], s->eps);
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < N; ++k) {
        Flux f = hlld_flux(&s->dQLeft[k], &s->dQRight[k], s->gamma, s->eps);
        Flux g;                                          /* P/hlld.cu:170-194 */
        g.f0 = f.f0;
        g.f1 = f.f3; g.f2 = f.f1; g.f3 = f.f2;
        g.f4 = f.f6; g.f5 = f.f4; g.f6 = f.f5;
        g.f7 = f.f7;
        s->hlldFlux[k] = g;
    }
}

static inline double eta_at(const struct oracle_sim *s, int k)
{
    return s->etaField ? s->etaField[k] : s->eta;
}

/* Resistive eta*J flux added to F: EX/bhattacharjee/main_bhattacharjee.cu:142-182
 * (identical formula EX/plasmoid/main_plasmoid.cu:175-224), centred B.
 * Pinned by the linear-B closed form test. */
static void resistive_x(struct oracle_sim *s, const Cons *U, Flux *flux)
{
    int nx = s->nx, ny = s->ny;
    double dx = s->dx, dy = s->dy;
    #pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 2; ++i) {
        for (int j = 1; j < ny - 2; ++j) {
            int k = j + i * ny;
            double jY, jZ, eta;
            double etaJY, etaJYPlus1, etaJZ, etaJZPlus1;
            double etaJYBZ, etaJYBZPlus1, etaJZBY, etaJZBYPlus1;
            jY = -(U[j + (i + 1) * ny].bZ - U[j + (i - 1) * ny].bZ) / (2.0 * dx);
            jZ = (U[j + (i + 1) * ny].bY - U[j + (i - 1) * ny].bY) / (2.0 * dx)
               - (U[j + 1 + i * ny].bX - U[j - 1 + i * ny].bX) / (2.0 * dy);
            eta = eta_at(s, k);
            etaJY = eta * jY;
            etaJZ = eta * jZ;
            etaJYBZ = etaJY * U[j + i * ny].bZ;
            etaJZBY = etaJZ * U[j + i * ny].bY;
            jY = -(U[j + (i + 2) * ny].bZ - U[j + i * ny].bZ) / (2.0 * dx);
            jZ = (U[j + (i + 2) * ny].bY - U[j + i * ny].bY) / (2.0 * dx)
               - (U[j + 1 + (i + 1) * ny].bX - U[j - 1 + (i + 1) * ny].bX) / (2.0 * dy);
            eta = eta_at(s, k + ny);
            etaJYPlus1 = eta * jY;
            etaJZPlus1 = eta * jZ;
            etaJYBZPlus1 = etaJYPlus1 * U[j + (i + 1) * ny].bZ;
            etaJZBYPlus1 = etaJZPlus1 * U[j + (i + 1) * ny].bY;
            flux[k].f5 -= 0.5 * (etaJZ + etaJZPlus1);
            flux[k].f6 += 0.5 * (etaJY + etaJYPlus1);
            flux[k].f7 += 0.5 * (etaJYBZ + etaJYBZPlus1)
                        - 0.5 * (etaJZBY + etaJZBYPlus1);
        }
    }
}

/* Resistive eta*J flux added to G: EX/bhattacharjee/main_bhattacharjee.cu:201-241 */
static void resistive_y(struct oracle_sim *s, const Cons *U, Flux *flux)
{
    int nx = s->nx, ny = s->ny;
    double dx = s->dx, dy = s->dy;
    #pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 2; ++i) {
        for (int j = 1; j < ny - 2; ++j) {
            int k = j + i * ny;
            double jX, jZ, eta;
            double etaJX, etaJXPlus1, etaJZ, etaJZPlus1;
            double etaJZBX, etaJZBXPlus1, etaJXBZ, etaJXBZPlus1;
            jX = (U[j + 1 + i * ny].bZ - U[j - 1 + i * ny].bZ) / (2.0 * dy);
            jZ = (U[j + (i + 1) * ny].bY - U[j + (i - 1) * ny].bY) / (2.0 * dx)
               - (U[j + 1 + i * ny].bX - U[j - 1 + i * ny].bX) / (2.0 * dy);
            eta = eta_at(s, k);
            etaJX = eta * jX;
            etaJZ = eta * jZ;
            etaJXBZ = etaJX * U[j + i * ny].bZ;
            etaJZBX = etaJZ * U[j + i * ny].bX;
            jX = (U[j + 2 + i * ny].bZ - U[j + i * ny].bZ) / (2.0 * dy);
            jZ = (U[j + 1 + (i + 1) * ny].bY - U[j + 1 + (i - 1) * ny].bY) / (2.0 * dx)
               - (U[j + 2 + i * ny].bX - U[j + i * ny].bX) / (2.0 * dy);
            eta = eta_at(s, k + 1);
            etaJXPlus1 = eta * jX;
            etaJZPlus1 = eta * jZ;
            etaJXBZPlus1 = etaJXPlus1 * U[j + 1 + i * ny].bZ;
            etaJZBXPlus1 = etaJZPlus1 * U[j + 1 + i * ny].bX;
            flux[k].f4 += 0.5 * (etaJZ + etaJZPlus1);
            flux[k].f6 -= 0.5 * (etaJX + etaJXPlus1);
            flux[k].f7 += 0.5 * (etaJZBX + etaJZBXPlus1)
                        - 0.5 * (etaJXBZ + etaJXBZPlus1);
        }
    }
}

/* FluxSolver::getFluxF / getFluxG: P/flux_solver.cu:10-33 */
static void get_flux_f(struct oracle_sim *s, const Cons *U, Flux *out)
{
    hlld_flux_x(s, U);
    memcpy(s->solverFlux, s->hlldFlux, sizeof(Flux) * s->N);
    resistive_x(s, U, s->solverFlux);
    memcpy(out, s->solverFlux, sizeof(Flux) * s->N);
}

static void get_flux_g(struct oracle_sim *s, const Cons *U, Flux *out)
{
    hlld_flux_y(s, U);
    memcpy(s->solverFlux, s->hlldFlux, sizeof(Flux) * s->N);
    resistive_y(s, U, s->solverFlux);
    memcpy(out, s->solverFlux, sizeof(Flux) * s->N);
}

/* ------------------------------------------- face <-> centre averaging */
/* shiftUToCenterForCT: P/resistiveMHD_2D.cu:313-368 */
static void shift_to_center(struct oracle_sim *s, Cons *U)
{
    int nx = s->nx, ny = s->ny;
    double *tmp = s->tmpVector;
    for (int k = 0; k < s->N; ++k) tmp[k] = U[k].bX;
    #pragma omp parallel for schedule(static)
    for (int i = 1; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            U[j + i * ny].bX = 0.5 * (tmp[j + i * ny] + tmp[j + (i - 1) * ny]);
    for (int k = 0; k < s->N; ++k) tmp[k] = U[k].bY;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            U[j + i * ny].bY = 0.5 * (tmp[j + i * ny] + tmp[j - 1 + i * ny]);
}

/* backUToCenterHalfForCT: P/resistiveMHD_2D.cu:371-426 */
static void back_to_center_half(struct oracle_sim *s, Cons *U)
{
    int nx = s->nx, ny = s->ny;
    double *tmp = s->tmpVector;
    for (int k = 0; k < s->N; ++k) tmp[k] = U[k].bX;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nx - 1; ++i)
        for (int j = 0; j < ny; ++j)
            U[j + i * ny].bX = 0.5 * (tmp[j + i * ny] + tmp[j + (i + 1) * ny]);
    for (int k = 0; k < s->N; ++k) tmp[k] = U[k].bY;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny - 1; ++j)
            U[j + i * ny].bY = 0.5 * (tmp[j + i * ny] + tmp[j + 1 + i * ny]);
}

/* ------------------------------------------------------------------ CT */
/* CT::divBClean: P/ct.cu:22-85 (getEZVector_kernel, CT_kernel).  The EMF
 * array persists between calls; entries at i = nx-1 or j = ny-1 are never
 * written (they stay 0 as the reference's zero-initialised device_vector). */
static void div_b_clean(struct oracle_sim *s, const double *bXOld, const double *bYOld, Cons *U)
{
    int nx = s->nx, ny = s->ny;
    const Flux *F = s->oldFluxF, *G = s->oldFluxG;
    double *eZ = s->eZVector;
    double dt = s->dt;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < nx - 1; ++i) {
        for (int j = 0; j < ny - 1; ++j) {
            double eZG1 = G[j + i * ny].f4;
            double eZG2 = G[j + (i + 1) * ny].f4;
            double eZF1 = -1.0 * F[j + i * ny].f5;
            double eZF2 = -1.0 * F[j + 1 + i * ny].f5;
            eZ[j + i * ny] = 0.25 * (eZG1 + eZG2 + eZF1 + eZF2);
        }
    }
    #pragma omp parallel for schedule(static)
    for (int i = 1; i < nx; ++i) {
        for (int j = 1; j < ny; ++j) {
            int k = j + i * ny;
            U[k].bX = bXOld[k] - dt / s->dy * (eZ[k] - eZ[k - 1]);
            U[k].bY = bYOld[k] + dt / s->dx * (eZ[k] - eZ[k - ny]);
        }
    }
}

/* ------------------------------------------------------------ boundary */
static void copy_cell(Cons *dst, const Cons *src) { *dst = *src; }

/* P/boundary.cu:4-71 periodicBoundaryX2nd */
static void periodic_x(struct oracle_sim *s, Cons *U)
{
    int nx = s->nx, ny = s->ny;
    for (int j = 0; j < ny; ++j) {
        copy_cell(&U[j + 0 * ny], &U[j + (nx - 6) * ny]);
        copy_cell(&U[j + 1 * ny], &U[j + (nx - 5) * ny]);
        copy_cell(&U[j + 2 * ny], &U[j + (nx - 4) * ny]);
        copy_cell(&U[j + (nx - 3) * ny], &U[j + 3 * ny]);
        copy_cell(&U[j + (nx - 2) * ny], &U[j + 4 * ny]);
        copy_cell(&U[j + (nx - 1) * ny], &U[j + 5 * ny]);
    }
}

/* P/boundary.cu:74-141 periodicBoundaryY2nd */
static void periodic_y(struct oracle_sim *s, Cons *U)
{
    int nx = s->nx, ny = s->ny;
    for (int i = 0; i < nx; ++i) {
        copy_cell(&U[0 + i * ny], &U[ny - 6 + i * ny]);
        copy_cell(&U[1 + i * ny], &U[ny - 5 + i * ny]);
        copy_cell(&U[2 + i * ny], &U[ny - 4 + i * ny]);
        copy_cell(&U[ny - 3 + i * ny], &U[3 + i * ny]);
        copy_cell(&U[ny - 2 + i * ny], &U[4 + i * ny]);
        copy_cell(&U[ny - 1 + i * ny], &U[5 + i * ny]);
    }
}

static void mirror_no_bx(Cons *dst, const Cons *src)
{
    dst->rho = src->rho; dst->rhoU = src->rhoU; dst->rhoV = src->rhoV; dst->rhoW = src->rhoW;
    dst->bY = src->bY; dst->bZ = src->bZ; dst->e = src->e;
}

static void mirror_no_by(Cons *dst, const Cons *src)
{
    dst->rho = src->rho; dst->rhoU = src->rhoU; dst->rhoV = src->rhoV; dst->rhoW = src->rhoW;
    dst->bX = src->bX; dst->bZ = src->bZ; dst->e = src->e;
}

/* S/boundary.cu:4-104 symmetricBoundaryX2nd: mirror everything but bX (the
 * normal momentum is not flipped), then rebuild the ghost bX from div B = 0. */
static void symmetric_x(struct oracle_sim *s, Cons *U)
{
    int nx = s->nx, ny = s->ny;
    double dx = s->dx, dy = s->dy;
    for (int j = 0; j < ny; ++j) {
        mirror_no_bx(&U[j + 0 * ny], &U[j + 5 * ny]);
        mirror_no_bx(&U[j + 1 * ny], &U[j + 4 * ny]);
        mirror_no_bx(&U[j + 2 * ny], &U[j + 3 * ny]);
        mirror_no_bx(&U[j + (nx - 1) * ny], &U[j + (nx - 6) * ny]);
        mirror_no_bx(&U[j + (nx - 2) * ny], &U[j + (nx - 5) * ny]);
        mirror_no_bx(&U[j + (nx - 3) * ny], &U[j + (nx - 4) * ny]);
    }
    for (int j = 0; j < ny; ++j) {
        if (0 < j) {
            U[j + 2 * ny].bX = U[j + 3 * ny].bX + (U[j + 3 * ny].bY - U[j - 1 + 3 * ny].bY) / dy * dx;
            U[j + 1 * ny].bX = U[j + 2 * ny].bX + (U[j + 2 * ny].bY - U[j - 1 + 2 * ny].bY) / dy * dx;
            U[j + 0 * ny].bX = U[j + 1 * ny].bX + (U[j + 1 * ny].bY - U[j - 1 + 1 * ny].bY) / dy * dx;
            U[j + (nx - 3) * ny].bX = U[j + (nx - 4) * ny].bX
                - (U[j + (nx - 3) * ny].bY - U[j - 1 + (nx - 3) * ny].bY) / dy * dx;
            U[j + (nx - 2) * ny].bX = U[j + (nx - 3) * ny].bX
                - (U[j + (nx - 2) * ny].bY - U[j - 1 + (nx - 2) * ny].bY) / dy * dx;
            U[j + (nx - 1) * ny].bX = U[j + (nx - 2) * ny].bX
                - (U[j + (nx - 1) * ny].bY - U[j - 1 + (nx - 1) * ny].bY) / dy * dx;
        }
        if (j == 0) {
            U[j + 2 * ny].bX = U[j + 3 * ny].bX;
            U[j + 1 * ny].bX = U[j + 2 * ny].bX;
            U[j + 0 * ny].bX = U[j + 1 * ny].bX;
            U[j + (nx - 3) * ny].bX = U[j + (nx - 4) * ny].bX;
            U[j + (nx - 2) * ny].bX = U[j + (nx - 3) * ny].bX;
            U[j + (nx - 1) * ny].bX = U[j + (nx - 2) * ny].bX;
        }
    }
}

/* S/boundary.cu:107-205 symmetricBoundaryY2nd */
static void symmetric_y(struct oracle_sim *s, Cons *U)
{
    int nx = s->nx, ny = s->ny;
    double dx = s->dx, dy = s->dy;
    for (int i = 0; i < nx; ++i) {
        mirror_no_by(&U[0 + i * ny], &U[5 + i * ny]);
        mirror_no_by(&U[1 + i * ny], &U[4 + i * ny]);
        mirror_no_by(&U[2 + i * ny], &U[3 + i * ny]);
        mirror_no_by(&U[ny - 1 + i * ny], &U[ny - 6 + i * ny]);
        mirror_no_by(&U[ny - 2 + i * ny], &U[ny - 5 + i * ny]);
        mirror_no_by(&U[ny - 3 + i * ny], &U[ny - 4 + i * ny]);
    }
    for (int i = 0; i < nx; ++i) {
        if (0 < i) {
            U[2 + i * ny].bY = U[3 + i * ny].bY + (U[3 + i * ny].bX - U[3 + (i - 1) * ny].bX) / dx * dy;
            U[1 + i * ny].bY = U[2 + i * ny].bY + (U[2 + i * ny].bX - U[2 + (i - 1) * ny].bX) / dx * dy;
            U[0 + i * ny].bY = U[1 + i * ny].bY + (U[1 + i * ny].bX - U[1 + (i - 1) * ny].bX) / dx * dy;
            U[ny - 3 + i * ny].bY = U[ny - 4 + i * ny].bY
                - (U[ny - 3 + i * ny].bX - U[ny - 3 + (i - 1) * ny].bX) / dx * dy;
            U[ny - 2 + i * ny].bY = U[ny - 3 + i * ny].bY
                - (U[ny - 2 + i * ny].bX - U[ny - 2 + (i - 1) * ny].bX) / dx * dy;
            U[ny - 1 + i * ny].bY = U[ny - 2 + i * ny].bY
                - (U[ny - 1 + i * ny].bX - U[ny - 1 + (i - 1) * ny].bX) / dx * dy;
        }
        if (i == 0) {
            U[2 + i * ny].bY = U[3 + i * ny].bY;
            U[1 + i * ny].bY = U[2 + i * ny].bY;
            U[0 + i * ny].bY = U[1 + i * ny].bY;
            U[ny - 3 + i * ny].bY = U[ny - 4 + i * ny].bY;
            U[ny - 2 + i * ny].bY = U[ny - 3 + i * ny].bY;
            U[ny - 1 + i * ny].bY = U[ny - 2 + i * ny].bY;
        }
    }
}

/* Boundary order: x first, then y over full columns (which fills the
 * corners), as P/resistiveMHD_2D.cu:171-172 and, for periodic x with
 * symmetric y, F32/resistiveMHD_2D.cu:239-240,257-258. */
static void apply_bc(struct oracle_sim *s, Cons *U)
{
    if (s->bc_x == ORACLE_BC_PERIODIC) periodic_x(s, U); else symmetric_x(s, U);
    if (s->bc_y == ORACLE_BC_PERIODIC) periodic_y(s, U); else symmetric_y(s, U);
}

/* ------------------------------------------------------------------ dt */
/* ResistiveMHD2D::calculateDt: P/resistiveMHD_2D.cu:277-291, with
 * thrust::min_element's sequential semantics (first element, replaced only
 * by a strictly smaller one). */
static double calculate_dt(struct oracle_sim *s)
{
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < s->N; ++k)
        s->dtVector[k] = cell_dt(&s->U[k], s->gamma, s->dx, s->dy, s->eps);
    double m = s->dtVector[0];
    for (int k = 1; k < s->N; ++k)
        if (s->dtVector[k] < m) m = s->dtVector[k];
    double lim = fmin(s->dx * s->dx / s->eta, s->dy * s->dy / s->eta);
    s->dt = fmin(m, lim) * s->cfl;
    if (s->dt_fixed > 0.0) s->dt = s->dt_fixed;
    return s->dt;
}

/* ----------------------------------------------------------- RK stages */
/* ResistiveMHD2D::oneStepRK2: P/resistiveMHD_2D.cu:133-195 (S/ identical
 * but for the boundary calls).  Note the second CT call reuses the stage-1
 * fluxes saved by setOldFlux2D (:167-168, :190). */
static void one_step_rk2(struct oracle_sim *s)
{
    int N = s->N, ny = s->ny;
    Cons *U = s->U, *UBar = s->UBar;
    for (int k = 0; k < N; ++k) { s->bXOld[k] = U[k].bX; s->bYOld[k] = U[k].bY; }
    memcpy(UBar, U, sizeof(Cons) * N);

    calculate_dt(s);

    shift_to_center(s, U);
    get_flux_f(s, U, s->fluxF);
    get_flux_g(s, U, s->fluxG);
    back_to_center_half(s, U);

    #pragma omp parallel for schedule(static)
    for (int k = ny; k < N; ++k)
        UBar[k] = stage_first(&U[k], &s->fluxF[k], &s->fluxF[k - ny], &s->fluxG[k], &s->fluxG[k - 1],
                              s->dt, s->dx, s->dy);

    memcpy(s->oldFluxF, s->fluxF, sizeof(Flux) * N);     /* CT::setOldFlux2D, P/ct.cu:12-19 */
    memcpy(s->oldFluxG, s->fluxG, sizeof(Flux) * N);
    div_b_clean(s, s->bXOld, s->bYOld, UBar);
    for (int k = 0; k < N; ++k) { s->bXStage1[k] = UBar[k].bX; s->bYStage1[k] = UBar[k].bY; }   /* test snapshot */

    apply_bc(s, UBar);

    shift_to_center(s, UBar);
    get_flux_f(s, UBar, s->fluxF);
    get_flux_g(s, UBar, s->fluxG);
    back_to_center_half(s, UBar);

    #pragma omp parallel for schedule(static)
    for (int k = ny; k < N; ++k)
        U[k] = stage_second(&U[k], &UBar[k], &s->fluxF[k], &s->fluxF[k - ny], &s->fluxG[k], &s->fluxG[k - 1],
                            s->dt, s->dx, s->dy);

    div_b_clean(s, s->bXOld, s->bYOld, U);

    apply_bc(s, U);
    s->time += s->dt;
    s->steps += 1;
}

/* ================================================================ */
/* Multi-GPU library step on one rank (one GPU's share of the       */
/* decomposition).  M = ResistiveMHD_multiGPU_cuda/ResistiveMHD2D_GPU */
/* Same HLLD, MUSCL and primitives as above (expression for          */
/* expression; the M/ kernels only differ in their element ranges,   */
/* which matter for ghost cells the halo overwrites).                */
/* ================================================================ */

static inline int msign(double x) { return (x > 0.0) - (x < 0.0); }   /* M/sign.cu */

/* M/calculate_half_Q.cu:10-48 + M/muscl.cu (x: shift ny, y: shift 1),
 * with the M/ element ranges. */
static void m_hlld_sweep(struct oracle_sim *s, const Cons *U, int shift, Flux *out)
{
    int nx = s->nx, ny = s->ny;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx - 1; ++i)
        for (int j = 0; j < ny - 1; ++j) {
            int k = j + i * ny;
            s->dQCenter[k] = basic_param(&U[k], &U[k + shift], s->gamma);
        }
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j) {
            int k = j + i * ny;
            s->dQLeft[k] = left_param(&s->dQCenter[k - shift], &s->dQCenter[k], &s->dQCenter[k + shift], s->eps);
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx - 2; ++i)
        for (int j = 0; j < ny - 2; ++j) {
            int k = j + i * ny;
            s->dQRight[k] = right_param(&s->dQCenter[k], &s->dQCenter[k + shift], &s->dQCenter[k + 2 * shift], s->eps);
        }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < s->N; ++k)
        out[k] = hlld_flux(&s->dQLeft[k], &s->dQRight[k], s->gamma, s->eps);
}

/* M/flux_solver.cu:73-104 (getEta at the x-face position) */
static void m_resistive_x(struct oracle_sim *s, const Cons *U, Flux *flux)
{
    int nx = s->nx, ny = s->ny;
    double dx = s->dx, dy = s->dy;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 2; ++i)
        for (int j = 1; j < ny - 1; ++j) {
            int index = j + i * ny;
            double jY, jZ, eta, etaJY, etaJZ, etaJYBZ, etaJZBY;
            jY = -(U[index + ny].bZ - U[index].bZ) / dx;
            jZ = 0.5 * (
                 (U[index + ny].bY - U[index - ny].bY) / (2.0 * dx) - (U[index + 1].bX - U[index - 1].bX) / (2.0 * dy)
               + (U[index + 2 * ny].bY - U[index].bY) / (2.0 * dx) - (U[index + ny + 1].bX - U[index + ny - 1].bX) / (2.0 * dy)
            );
            eta = s->etaXFaceM ? s->etaXFaceM[index] : s->eta;
            etaJY = eta * jY;
            etaJZ = eta * jZ;
            etaJYBZ = etaJY * 0.5 * (U[index].bZ + U[index + ny].bZ);
            etaJZBY = etaJZ * 0.5 * (U[index].bY + U[index + ny].bY);
            flux[index].f5 -= etaJZ;
            flux[index].f6 += etaJY;
            flux[index].f7 += etaJYBZ - etaJZBY;
        }
}

/* M/flux_solver.cu:158-189 (getEta at the y-face position) */
static void m_resistive_y(struct oracle_sim *s, const Cons *U, Flux *flux)
{
    int nx = s->nx, ny = s->ny;
    double dx = s->dx, dy = s->dy;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 2; ++j) {
            int index = j + i * ny;
            double jX, jZ, eta, etaJX, etaJZ, etaJZBX, etaJXBZ;
            jX = (U[index + 1].bZ - U[index].bZ) / dy;
            jZ = 0.5 * (
                 (U[index + ny].bY - U[index - ny].bY) / (2.0 * dx) - (U[index + 1].bX - U[index - 1].bX) / (2.0 * dy)
               + (U[index + ny + 1].bY - U[index - ny + 1].bY) / (2.0 * dx) - (U[index + 2].bX - U[index].bX) / (2.0 * dy)
            );
            eta = s->etaYFaceM ? s->etaYFaceM[index] : s->eta;
            etaJX = eta * jX;
            etaJZ = eta * jZ;
            etaJXBZ = etaJX * 0.5 * (U[index].bZ + U[index + 1].bZ);
            etaJZBX = etaJZ * 0.5 * (U[index].bX + U[index + 1].bX);
            flux[index].f4 += etaJZ;
            flux[index].f6 -= etaJX;
            flux[index].f7 += etaJZBX - etaJXBZ;
        }
}

/* FluxSolver::getFluxF/G of M/flux_solver.cu:11-34 with M/hlld.cu:105-245 */
static void m_get_flux_f(struct oracle_sim *s, const Cons *U, Flux *out)
{
    m_hlld_sweep(s, U, s->ny, s->hlldFlux);
    memcpy(out, s->hlldFlux, sizeof(Flux) * s->N);
    m_resistive_x(s, U, out);
}

static void m_get_flux_g(struct oracle_sim *s, const Cons *U, Flux *out)
{
    int N = s->N;
    Cons *t = s->tmpUForFluxG;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < N; ++k) {                         /* M/hlld.cu:131-152 */
        t[k].rho = U[k].rho;
        t[k].rhoU = U[k].rhoV; t[k].rhoV = U[k].rhoW; t[k].rhoW = U[k].rhoU;
        t[k].bX = U[k].bY; t[k].bY = U[k].bZ; t[k].bZ = U[k].bX;
        t[k].e = U[k].e;
    }
    m_hlld_sweep(s, t, 1, s->hlldFlux);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < N; ++k) {                         /* M/hlld.cu:175-200 */
        Flux f = s->hlldFlux[k], g;
        g.f0 = f.f0; g.f1 = f.f3; g.f2 = f.f1; g.f3 = f.f2;
        g.f4 = f.f6; g.f5 = f.f4; g.f6 = f.f5; g.f7 = f.f7;
        out[k] = g;
    }
    m_resistive_y(s, U, out);
}

/* CT::setOldFlux2D / setNowFlux2D: M/ct.cu:26-65 */
static void m_set_flux(struct oracle_sim *s, const Flux *F, const Flux *G, const Cons *U,
                       double *NF5, double *NG4, double *F5, double *G4, double *NF0, double *NG0)
{
    int nx = s->nx, ny = s->ny;
    double dx = s->dx, dy = s->dy;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j) {
            int index = j + i * ny;
            double rho = U[index].rho;
            double u = U[index].rhoU / rho;
            double v = U[index].rhoV / rho;
            double bX = 0.5 * (U[index].bX + U[index - ny].bX);
            double bY = 0.5 * (U[index].bY + U[index - 1].bY);
            double jZ = 0.25 * (
                (U[index + ny].bY - U[index].bY) / dx - (U[index + 1].bX - U[index].bX) / dy
              + (U[index - 1 + ny].bY - U[index - 1].bY) / dx - (U[index].bX - U[index - 1].bX) / dy
              + (U[index - 1].bY - U[index - 1 - ny].bY) / dx - (U[index - ny].bX - U[index - ny - 1].bX) / dy
              + (U[index].bY - U[index - ny].bY) / dx - (U[index + 1 - ny].bX - U[index - ny].bX) / dy
            );
            double eta = s->etaCellM ? s->etaCellM[index] : s->eta;
            NF5[index] = F[index].f5;
            NG4[index] = G[index].f4;
            F5[index] = u * bY - v * bX - eta * jZ;
            G4[index] = -(u * bY - v * bX - eta * jZ);
            NF0[index] = F[index].f0;
            NG0[index] = G[index].f0;
        }
}

/* CT-contact corner EMF of one flux set: M/ct.cu:147-169 */
static inline double m_emf(int index, int ny, const double *NF5, const double *NG4, const double *F5,
                           const double *G4, const double *NF0, const double *NG0)
{
    double avg = 0.25 * (
        - NF5[index] - NF5[index + 1]
        + NG4[index] + NG4[index + ny]
    );
    double eS = (1.0 + msign(NF0[index])) * (NG4[index] - G4[index])
              + (1.0 - msign(NF0[index])) * (NG4[index + ny] - G4[index + ny]);
    double eN = (1.0 + msign(NF0[index + 1])) * (G4[index + 1] - NG4[index])
              + (1.0 - msign(NF0[index + 1])) * (G4[index + 1 + ny] - NG4[index + ny]);
    double eW = -(1.0 + msign(NG0[index])) * (NF5[index] - F5[index])
               -(1.0 - msign(NG0[index])) * (NF5[index + 1] - F5[index + 1]);
    double eE = -(1.0 + msign(NG0[index + ny])) * (F5[index + ny] - NF5[index])
               -(1.0 - msign(NG0[index + ny])) * (F5[index + 1 + ny] - NF5[index + 1]);
    return avg + 0.125 * (eS - eN + eW - eE);
}

/* CT::divBClean: M/ct.cu:120-218 */
static void m_div_b_clean(struct oracle_sim *s, Cons *U)
{
    int nx = s->nx, ny = s->ny;
    double *eZ = s->eZVector;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx - 1; ++i)
        for (int j = 0; j < ny - 1; ++j) {
            int index = j + i * ny;
            double eOld = m_emf(index, ny, s->oNF5, s->oNG4, s->oF5, s->oG4, s->oNF0, s->oNG0);
            double eNow = m_emf(index, ny, s->nNF5, s->nNG4, s->nF5, s->nG4, s->nNF0, s->nNG0);
            eZ[index] = 0.5 * (eOld + eNow);
        }
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j) {
            int k = j + i * ny;
            U[k].bX = s->bXOld[k] - s->dt / s->dy * (eZ[k] - eZ[k - 1]);
            U[k].bY = s->bYOld[k] + s->dt / s->dx * (eZ[k] - eZ[k - ny]);
        }
}

/* sendrecv_U on one rank (M/mpi.cu:104-188): the rank is its own left,
 * right, up and down neighbour.  x: interior rows of the 3 edge columns;
 * then y: the 3 edge rows of every column. */
static void m_halo(struct oracle_sim *s, Cons *U)
{
    int nx = s->nx, ny = s->ny, b = 3, lnx = nx - 6, lny = ny - 6;
    for (int i = 0; i < b; ++i)
        for (int j = 0; j < lny; ++j) {
            Cons r = U[j + b + (lnx + i) * ny];        /* sent right, received from the left */
            Cons l = U[j + b + (b + i) * ny];          /* sent left, received from the right */
            U[j + b + i * ny] = r;
            U[j + b + (lnx + b + i) * ny] = l;
        }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < b; ++j) {
            Cons d = U[j + lny + i * ny];
            Cons u = U[j + b + i * ny];
            U[j + i * ny] = d;
            U[j + lny + b + i * ny] = u;
        }
}

/* calculateDt: M/resistiveMHD2D.cu:489-556 -- interior cells only, no
 * diffusion limit. */
static double m_calculate_dt(struct oracle_sim *s)
{
    int nx = s->nx, ny = s->ny;
    double m = cell_dt(&s->U[3 + 3 * ny], s->gamma, s->dx, s->dy, s->eps);
    for (int i = 3; i < nx - 3; ++i)
        for (int j = 3; j < ny - 3; ++j) {
            double d = cell_dt(&s->U[j + i * ny], s->gamma, s->dx, s->dy, s->eps);
            if (d < m) m = d;
        }
    s->dt = m * s->cfl;
    if (s->dt_fixed > 0.0) s->dt = s->dt_fixed;
    return s->dt;
}

/* oneStepFirst_kernel / oneStepSecond_kernel ranges: 0 < i, 0 < j */
static void m_one_step_rk2(struct oracle_sim *s)
{
    int N = s->N, nx = s->nx, ny = s->ny;
    Cons *U = s->U, *UBar = s->UBar;
    for (int k = 0; k < N; ++k) { s->bXOld[k] = U[k].bX; s->bYOld[k] = U[k].bY; }
    memcpy(UBar, U, sizeof(Cons) * N);
    m_calculate_dt(s);

    shift_to_center(s, U);
    m_get_flux_f(s, U, s->fluxF);
    m_get_flux_g(s, U, s->fluxG);
    back_to_center_half(s, U);
    m_set_flux(s, s->fluxF, s->fluxG, U, s->oNF5, s->oNG4, s->oF5, s->oG4, s->oNF0, s->oNG0);
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j) {
            int k = j + i * ny;
            UBar[k] = stage_first(&U[k], &s->fluxF[k], &s->fluxF[k - ny], &s->fluxG[k], &s->fluxG[k - 1],
                                  s->dt, s->dx, s->dy);
        }
    m_halo(s, UBar);

    shift_to_center(s, UBar);
    m_get_flux_f(s, UBar, s->fluxF);
    m_get_flux_g(s, UBar, s->fluxG);
    back_to_center_half(s, UBar);
#pragma omp parallel for schedule(static)
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j) {
            int k = j + i * ny;
            U[k] = stage_second(&U[k], &UBar[k], &s->fluxF[k], &s->fluxF[k - ny], &s->fluxG[k], &s->fluxG[k - 1],
                                s->dt, s->dx, s->dy);
        }
    m_halo(s, U);
    m_set_flux(s, s->fluxF, s->fluxG, U, s->nNF5, s->nNG4, s->nF5, s->nG4, s->nNF0, s->nNG0);
    m_div_b_clean(s, U);
    m_halo(s, U);
    s->time += s->dt;
    s->steps += 1;
}

/* -------------------------------------------------------------- C API */
struct oracle_sim *oracle_create(int nx, int ny, double dx, double dy, double gamma, double cfl,
                                 double eps, double eta, int bc_x, int bc_y)
{
    if (nx < 12 || ny < 12) return NULL;
    struct oracle_sim *s = (struct oracle_sim *)calloc(1, sizeof *s);
    if (!s) return NULL;
    s->nx = nx; s->ny = ny; s->N = nx * ny;
    s->dx = dx; s->dy = dy; s->gamma = gamma; s->cfl = cfl; s->eps = eps; s->eta = eta;
    s->bc_x = bc_x; s->bc_y = bc_y;
    size_t N = (size_t)s->N;
    /* zero-initialised, as thrust::device_vector of these structs */
    s->U = calloc(N, sizeof(Cons)); s->UBar = calloc(N, sizeof(Cons));
    s->fluxF = calloc(N, sizeof(Flux)); s->fluxG = calloc(N, sizeof(Flux));
    s->dtVector = calloc(N, sizeof(double)); s->bXOld = calloc(N, sizeof(double));
    s->bYOld = calloc(N, sizeof(double)); s->tmpVector = calloc(N, sizeof(double));
    s->solverFlux = calloc(N, sizeof(Flux)); s->hlldFlux = calloc(N, sizeof(Flux));
    s->dQCenter = calloc(N, sizeof(Prim)); s->dQLeft = calloc(N, sizeof(Prim)); s->dQRight = calloc(N, sizeof(Prim));
    s->tmpUForFluxG = calloc(N, sizeof(Cons));
    s->oldFluxF = calloc(N, sizeof(Flux)); s->oldFluxG = calloc(N, sizeof(Flux));
    s->eZVector = calloc(N, sizeof(double));
    s->bXStage1 = calloc(N, sizeof(double)); s->bYStage1 = calloc(N, sizeof(double));
    s->oNF5 = calloc(N, sizeof(double)); s->oNG4 = calloc(N, sizeof(double)); s->oF5 = calloc(N, sizeof(double));
    s->oG4 = calloc(N, sizeof(double)); s->oNF0 = calloc(N, sizeof(double)); s->oNG0 = calloc(N, sizeof(double));
    s->nNF5 = calloc(N, sizeof(double)); s->nNG4 = calloc(N, sizeof(double)); s->nF5 = calloc(N, sizeof(double));
    s->nG4 = calloc(N, sizeof(double)); s->nNF0 = calloc(N, sizeof(double)); s->nNG0 = calloc(N, sizeof(double));
    if (!s->oNF5 || !s->oNG4 || !s->oF5 || !s->oG4 || !s->oNF0 || !s->oNG0 ||
        !s->nNF5 || !s->nNG4 || !s->nF5 || !s->nG4 || !s->nNF0 || !s->nNG0) {
        oracle_destroy(s);
        return NULL;
    }
    if (!s->U || !s->UBar || !s->fluxF || !s->fluxG || !s->dtVector || !s->bXOld || !s->bYOld ||
        !s->tmpVector || !s->solverFlux || !s->hlldFlux || !s->dQCenter || !s->dQLeft ||
        !s->dQRight || !s->tmpUForFluxG || !s->oldFluxF || !s->oldFluxG || !s->eZVector) {
        oracle_destroy(s);
        return NULL;
    }
    return s;
}

void oracle_destroy(struct oracle_sim *s)
{
    if (!s) return;
    free(s->U); free(s->UBar); free(s->fluxF); free(s->fluxG); free(s->dtVector); free(s->bXOld);
    free(s->bYOld); free(s->tmpVector); free(s->solverFlux); free(s->hlldFlux); free(s->dQCenter);
    free(s->dQLeft); free(s->dQRight); free(s->tmpUForFluxG); free(s->oldFluxF); free(s->oldFluxG);
    free(s->eZVector); free(s->etaField); free(s->bXStage1); free(s->bYStage1);
    free(s->etaXFaceM); free(s->etaYFaceM); free(s->etaCellM);
    free(s->oNF5); free(s->oNG4); free(s->oF5); free(s->oG4); free(s->oNF0); free(s->oNG0);
    free(s->nNF5); free(s->nNG4); free(s->nF5); free(s->nG4); free(s->nNF0); free(s->nNG0);
    free(s);
}

void oracle_set_state(struct oracle_sim *s, const double *aos) { memcpy(s->U, aos, sizeof(Cons) * s->N); }
void oracle_get_state(const struct oracle_sim *s, double *aos) { memcpy(aos, s->U, sizeof(Cons) * s->N); }

/* Face B (bX, bY planes of nx*ny) of UBar right after the stage-1 CT of the
 * last single-GPU step (P/resistiveMHD_2D.cu:167-168).  Pins Q7: the second
 * CT (:190) reuses the stage-1 EMF and the old face B, so the interior face
 * B of U^{n+1} must equal this bit for bit. */
void oracle_get_stage1_face_b(const struct oracle_sim *s, double *bx, double *by)
{
    memcpy(bx, s->bXStage1, sizeof(double) * s->N);
    memcpy(by, s->bYStage1, sizeof(double) * s->N);
}

/* HLLD region indicators (P/hlld.cu:64-69) of every face of the current
 * state, as bits: outer-left, single-star left, double-star left,
 * outer-right, single-star right, double-star right.  The faces are those
 * the next stage solves: the state is centred (shift_to_center) and
 * reconstructed as in hlld_flux_x / hlld_flux_y.  rx[k]: face between
 * cells k and k+ny; ry[k]: face between k and k+1 (0 where the reference's
 * ranges leave a face unset).  Test hook for the numerics comparison. */
static unsigned char region_bits(const HlldParams *P)
{
    return (unsigned char)((P->SL > 0.0) | ((P->SL <= 0.0) && (0.0 < P->S1L)) << 1
                           | ((P->S1L <= 0.0) && (0.0 < P->SM)) << 2 | (P->SR <= 0.0) << 3
                           | ((P->S1R <= 0.0) && (0.0 < P->SR)) << 4 | ((P->SM <= 0.0) && (0.0 < P->S1R)) << 5);
}

void oracle_face_regions(struct oracle_sim *s, unsigned char *rx, unsigned char *ry)
{
    int N = s->N, ny = s->ny;
    Cons *c = (Cons *)malloc(sizeof(Cons) * N);
    memcpy(c, s->U, sizeof(Cons) * N);
    shift_to_center(s, c);
    memset(rx, 0, N);
    memset(ry, 0, N);
    for (int k = 0; k < N - ny; ++k)
        s->dQCenter[k] = basic_param(&c[k], &c[k + ny], s->gamma);
    for (int k = ny; k < N - ny; ++k)
        s->dQLeft[k] = left_param(&s->dQCenter[k - ny], &s->dQCenter[k], &s->dQCenter[k + ny], s->eps);
    for (int k = 0; k < N - 2 * ny; ++k)
        s->dQRight[k] = right_param(&s->dQCenter[k], &s->dQCenter[k + ny], &s->dQCenter[k + 2 * ny], s->eps);
    for (int k = ny; k < N - 2 * ny; ++k) {
        HlldParams P = hlld_params(&s->dQLeft[k], &s->dQRight[k], s->gamma, s->eps);
        rx[k] = region_bits(&P);
    }
    Cons *t = s->tmpUForFluxG;
    for (int k = 0; k < N; ++k) {
        t[k].rho = c[k].rho;
        t[k].rhoU = c[k].rhoV; t[k].rhoV = c[k].rhoW; t[k].rhoW = c[k].rhoU;
        t[k].bX = c[k].bY; t[k].bY = c[k].bZ; t[k].bZ = c[k].bX;
        t[k].e = c[k].e;
    }
    for (int k = 0; k < N - 1; ++k)
        s->dQCenter[k] = basic_param(&t[k], &t[k + 1], s->gamma);
    for (int k = 1; k < N - 1; ++k)
        s->dQLeft[k] = left_param(&s->dQCenter[k - 1], &s->dQCenter[k], &s->dQCenter[k + 1], s->eps);
    for (int k = 0; k < N - 2; ++k)
        s->dQRight[k] = right_param(&s->dQCenter[k], &s->dQCenter[k + 1], &s->dQCenter[k + 2], s->eps);
    for (int k = 1; k < N - 2; ++k) {
        HlldParams P = hlld_params(&s->dQLeft[k], &s->dQRight[k], s->gamma, s->eps);
        ry[k] = region_bits(&P);
    }
    free(c);
}

/* The CT-contact corner EMF of one flux set (m_emf) on a 2 x 2 patch of
 * cells (i, j), (i, j+1), (i+1, j), (i+1, j+1), each array in that order:
 * face fluxes NF5 (x-face F.f5), NG4 (y-face G.f4), NF0, NG0 (mass fluxes)
 * and the cell terms F5, G4 of M/ct.cu:26-65.  Test hook for the upwind
 * pin against Gardiner & Stone (2005). */
double oracle_ct_contact_emf(const double *NF5, const double *NG4, const double *F5, const double *G4,
                             const double *NF0, const double *NG0)
{
    return m_emf(0, 2, NF5, NG4, F5, G4, NF0, NG0);
}

int oracle_set_eta_field(struct oracle_sim *s, const double *eta)
{
    if (!eta) { free(s->etaField); s->etaField = NULL; return 0; }
    if (!s->etaField) s->etaField = (double *)malloc(sizeof(double) * s->N);
    if (!s->etaField) return -1;
    memcpy(s->etaField, eta, sizeof(double) * s->N);
    return 0;
}

/* The MPI library's getEta sampled at x-faces, y-faces and cells (all three
 * planes, or all NULL for the uniform eta). */
int oracle_set_eta_staggered(struct oracle_sim *s, const double *xface, const double *yface, const double *cell)
{
    if (!xface && !yface && !cell) {
        free(s->etaXFaceM); free(s->etaYFaceM); free(s->etaCellM);
        s->etaXFaceM = s->etaYFaceM = s->etaCellM = NULL;
        return 0;
    }
    if (!xface || !yface || !cell) return -1;
    double **dst[3] = {&s->etaXFaceM, &s->etaYFaceM, &s->etaCellM};
    const double *src[3] = {xface, yface, cell};
    for (int q = 0; q < 3; ++q) {
        if (!*dst[q]) *dst[q] = (double *)malloc(sizeof(double) * s->N);
        if (!*dst[q]) return -1;
        memcpy(*dst[q], src[q], sizeof(double) * s->N);
    }
    return 0;
}

void oracle_apply_boundary(struct oracle_sim *s) { apply_bc(s, s->U); }

void oracle_apply_boundary_axis(struct oracle_sim *s, int axis, int kind)
{
    if (axis == 0) { if (kind == ORACLE_BC_PERIODIC) periodic_x(s, s->U); else symmetric_x(s, s->U); }
    else           { if (kind == ORACLE_BC_PERIODIC) periodic_y(s, s->U); else symmetric_y(s, s->U); }
}

double oracle_calculate_dt(struct oracle_sim *s)
{
    return s->scheme == ORACLE_SCHEME_MULTI_SHARE ? m_calculate_dt(s) : calculate_dt(s);
}

void oracle_apply_halo(struct oracle_sim *s) { m_halo(s, s->U); }

void oracle_step(struct oracle_sim *s, int nsteps)
{
    for (int n = 0; n < nsteps; ++n) {
        if (s->scheme == ORACLE_SCHEME_MULTI_SHARE) m_one_step_rk2(s);
        else one_step_rk2(s);
    }
}

int oracle_set_scheme(struct oracle_sim *s, int scheme)
{
    if (scheme == ORACLE_SCHEME_MULTI_SHARE && (s->etaField || s->bc_x != ORACLE_BC_PERIODIC || s->bc_y != ORACLE_BC_PERIODIC))
        return -1;
    s->scheme = scheme;
    return 0;
}

/* checkCalculationIsCrashed of M/resistiveMHD2D.cu:567-579 also flags dt < 0 */
int oracle_check_crashed_multi(const struct oracle_sim *s)
{
    return oracle_check_crashed(s) || s->dt < 0;
}

double oracle_get_dt(const struct oracle_sim *s) { return s->dt; }
void oracle_set_fixed_dt(struct oracle_sim *s, double dt) { s->dt_fixed = dt; }
double oracle_get_time(const struct oracle_sim *s) { return s->time; }

/* checkCalculationIsCrashed: P/resistiveMHD_2D.cu:294-309 */
int oracle_check_crashed(const struct oracle_sim *s)
{
    for (int k = 0; k < s->N; ++k) if (isnan(s->U[k].e)) return 1;
    return 0;
}

/* Fluxes of the last stage-1 call (the ones CT used), for white-box tests. */
void oracle_get_stage1_fluxes(const struct oracle_sim *s, double *F, double *G)
{
    memcpy(F, s->oldFluxF, sizeof(Flux) * s->N);
    memcpy(G, s->oldFluxG, sizeof(Flux) * s->N);
}

/* Stand-alone flux evaluation on a given state (shift, F, G, back), for the
 * HLLD/resistive white-box tests.  Leaves U unchanged except for the face
 * smoothing that shift+back applies, exactly as the reference does. */
void oracle_compute_fluxes(struct oracle_sim *s, double *aos_state, double *F, double *G)
{
    Cons *U = (Cons *)aos_state;
    shift_to_center(s, U);
    if (s->scheme == ORACLE_SCHEME_MULTI_SHARE) {
        m_get_flux_f(s, U, (Flux *)F);
        m_get_flux_g(s, U, (Flux *)G);
    } else {
        get_flux_f(s, U, (Flux *)F);
        get_flux_g(s, U, (Flux *)G);
    }
    back_to_center_half(s, U);
}

double oracle_cell_dt(const double *cell8, double gamma, double dx, double dy, double eps)
{
    Cons c;
    memcpy(&c, cell8, sizeof c);
    return cell_dt(&c, gamma, dx, dy, eps);
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
