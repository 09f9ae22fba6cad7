// thrust_twin.cpp -- the reference's single-GPU RK2 step restated with its
// own pass structure on rocThrust's OpenMP backend: the CPU baseline the
// benchmark times (test and measurement infrastructure, not part of the
// product library).  Built host-only with
//   hipcc --offload-host-only -fopenmp -DTHRUST_DEVICE_SYSTEM=THRUST_DEVICE_SYSTEM_OMP
// (oracle/Makefile), so every thrust::device_vector lives in host memory
// and every transform / for_each / reduction runs on OpenMP threads.
//
// What is kept from the reference (paths relative to its root,
// P = ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_periodicXY,
// S = ..._symmetricXY, EX = ResistiveMHD_singleGPU_thrust/examples2D):
//   - array-of-structs state, primitives and fluxes of nx*ny cells;
//   - one pass per reference transform or kernel of oneStepRK2
//     (P/resistiveMHD_2D.cu:133-195): the face-field copies, dt
//     (transform + min_element, :277-291), the centring of B (:313-426),
//     per direction the primitives, left and right MUSCL states, the HLLD
//     parameters as an array of structs (P/hlld.cu:256-444), the six
//     candidate fluxes (:448-565) and their selection (:35-131), the y
//     sweep's state and flux rotations (:134-233), the flux solver's copies
//     and the example's resistive kernels (EX/bhattacharjee/
//     main_bhattacharjee.cu:131-257), the stage transforms, CT
//     (P/ct.cu:12-85), the boundary kernels (P/boundary.cu, S/boundary.cu)
//     and the NaN check (P/resistiveMHD_2D.cu:294-309) after every step;
//   - the element ranges of every pass.
// The per-element arithmetic is the oracle's (mhd2d_oracle_math.h), so the
// twin's states equal the oracle's bit for bit (tests/test_thrust_twin.py).
#include <thrust/copy.h>
#include <thrust/device_vector.h>
#include <thrust/extrema.h>
#include <thrust/for_each.h>
#include <thrust/iterator/counting_iterator.h>
#include <thrust/transform.h>
#include <thrust/transform_reduce.h>

#include <cmath>
#include <cstring>
#include <new>
#include <omp.h>

#include "mhd2d_oracle_math.h"

namespace {

using thrust::raw_pointer_cast;
typedef thrust::counting_iterator<long> Count;

enum { BC_PERIODIC = 0, BC_SYMMETRIC = 1 };

struct Twin {
    int nx, ny;
    long N;
    double dx, dy, gamma, cfl, eps, eta;
    int bc_x, bc_y;
    double dt = 0.0, time = 0.0;
    thrust::device_vector<Cons> U, UBar, tmpU;
    thrust::device_vector<Flux> fluxF, fluxG, solverFlux, hlldFlux, oldFluxF, oldFluxG;
    thrust::device_vector<Prim> dQCenter, dQLeft, dQRight;
    thrust::device_vector<HlldParams> params;
    thrust::device_vector<HlldCandidates> cands;
    thrust::device_vector<double> dtVector, bXOld, bYOld, tmpVector, eZVector, etaField;
    bool has_eta = false;

    Twin(int nx_, int ny_) : nx(nx_), ny(ny_), N((long)nx_ * ny_)
    {
        const Cons c0{};
        const Flux f0{};
        const Prim p0{};
        U.assign(N, c0); UBar.assign(N, c0); tmpU.assign(N, c0);
        fluxF.assign(N, f0); fluxG.assign(N, f0); solverFlux.assign(N, f0); hlldFlux.assign(N, f0);
        oldFluxF.assign(N, f0); oldFluxG.assign(N, f0);
        dQCenter.assign(N, p0); dQLeft.assign(N, p0); dQRight.assign(N, p0);
        params.resize(N);
        cands.resize(N);
        dtVector.assign(N, 0.0); bXOld.assign(N, 0.0); bYOld.assign(N, 0.0); tmpVector.assign(N, 0.0);
        eZVector.assign(N, 0.0);
    }
};

template <class F>
void each(long lo, long hi, F f)
{
    if (hi > lo) thrust::for_each(Count(lo), Count(hi), f);
}

// 2D index space [i0, i1) x [j0, j1) flattened, i slow
template <class F>
void each2(int i0, int i1, int j0, int j1, F f)
{
    if (i1 <= i0 || j1 <= j0) return;
    const long nj = j1 - j0;
    thrust::for_each(Count(0), Count((long)(i1 - i0) * nj), [=](long t) { f(i0 + (int)(t / nj), j0 + (int)(t % nj)); });
}

// ------------------------------------------------------------- flux passes
// HLLD::calculateFluxF / calculateFluxG with the reference's ranges:
// primitives of k < N - s, left states of [s, N - s), right states of
// k < N - 2 s (s = ny along x, 1 along y), then parameters, candidates and
// selection over every element.
void hlld_sweep(Twin &t, const Cons *Ud, long s)
{
    const long N = t.N;
    const double gamma = t.gamma, eps = t.eps;
    Prim *C = raw_pointer_cast(t.dQCenter.data());
    Prim *L = raw_pointer_cast(t.dQLeft.data());
    Prim *R = raw_pointer_cast(t.dQRight.data());
    HlldParams *P = raw_pointer_cast(t.params.data());
    HlldCandidates *Q = raw_pointer_cast(t.cands.data());
    Flux *H = raw_pointer_cast(t.hlldFlux.data());
    thrust::transform(Count(0), Count(N - s), t.dQCenter.begin(),
                      [=](long k) { return basic_param(&Ud[k], &Ud[k + s], gamma); });
    thrust::transform(Count(s), Count(N - s), t.dQLeft.begin() + s,
                      [=](long k) { return left_param(&C[k - s], &C[k], &C[k + s], eps); });
    thrust::transform(Count(0), Count(N - 2 * s), t.dQRight.begin(),
                      [=](long k) { return right_param(&C[k], &C[k + s], &C[k + 2 * s], eps); });
    thrust::transform(Count(0), Count(N), t.params.begin(),
                      [=](long k) { return hlld_params(&L[k], &R[k], gamma, eps); });
    thrust::transform(Count(0), Count(N), t.cands.begin(), [=](long k) { return hlld_candidates(&P[k]); });
    thrust::transform(Count(0), Count(N), t.hlldFlux.begin(), [=](long k) { return hlld_select(&P[k], &Q[k]); });
    (void)H;
}

double eta_of(const double *e, double eta, long k) { return e ? e[k] : eta; }

// EX/bhattacharjee/main_bhattacharjee.cu:142-182 (centred B)
void resistive_x(Twin &t, const Cons *U, Flux *flux)
{
    const int ny = t.ny;
    const double dx = t.dx, dy = t.dy, eta0 = t.eta;
    const double *ef = t.has_eta ? raw_pointer_cast(t.etaField.data()) : nullptr;
    each2(1, t.nx - 2, 1, t.ny - 2, [=](int i, int j) {
        const long k = j + (long)i * ny;
        double jY = -(U[j + (i + 1) * (long)ny].bZ - U[j + (i - 1) * (long)ny].bZ) / (2.0 * dx);
        double jZ = (U[j + (i + 1) * (long)ny].bY - U[j + (i - 1) * (long)ny].bY) / (2.0 * dx)
                  - (U[j + 1 + i * (long)ny].bX - U[j - 1 + i * (long)ny].bX) / (2.0 * dy);
        double eta = eta_of(ef, eta0, k);
        const double etaJY = eta * jY, etaJZ = eta * jZ;
        const double etaJYBZ = etaJY * U[k].bZ, etaJZBY = etaJZ * U[k].bY;
        jY = -(U[j + (i + 2) * (long)ny].bZ - U[k].bZ) / (2.0 * dx);
        jZ = (U[j + (i + 2) * (long)ny].bY - U[k].bY) / (2.0 * dx)
           - (U[j + 1 + (i + 1) * (long)ny].bX - U[j - 1 + (i + 1) * (long)ny].bX) / (2.0 * dy);
        eta = eta_of(ef, eta0, k + ny);
        const double etaJY1 = eta * jY, etaJZ1 = eta * jZ;
        const double etaJYBZ1 = etaJY1 * U[k + ny].bZ, etaJZBY1 = etaJZ1 * U[k + ny].bY;
        flux[k].f5 -= 0.5 * (etaJZ + etaJZ1);
        flux[k].f6 += 0.5 * (etaJY + etaJY1);
        flux[k].f7 += 0.5 * (etaJYBZ + etaJYBZ1) - 0.5 * (etaJZBY + etaJZBY1);
    });
}

// EX/bhattacharjee/main_bhattacharjee.cu:201-241
void resistive_y(Twin &t, const Cons *U, Flux *flux)
{
    const int ny = t.ny;
    const double dx = t.dx, dy = t.dy, eta0 = t.eta;
    const double *ef = t.has_eta ? raw_pointer_cast(t.etaField.data()) : nullptr;
    each2(1, t.nx - 2, 1, t.ny - 2, [=](int i, int j) {
        const long k = j + (long)i * ny;
        double jX = (U[j + 1 + i * (long)ny].bZ - U[j - 1 + i * (long)ny].bZ) / (2.0 * dy);
        double jZ = (U[j + (i + 1) * (long)ny].bY - U[j + (i - 1) * (long)ny].bY) / (2.0 * dx)
                  - (U[j + 1 + i * (long)ny].bX - U[j - 1 + i * (long)ny].bX) / (2.0 * dy);
        double eta = eta_of(ef, eta0, k);
        const double etaJX = eta * jX, etaJZ = eta * jZ;
        const double etaJXBZ = etaJX * U[k].bZ, etaJZBX = etaJZ * U[k].bX;
        jX = (U[j + 2 + i * (long)ny].bZ - U[k].bZ) / (2.0 * dy);
        jZ = (U[j + 1 + (i + 1) * (long)ny].bY - U[j + 1 + (i - 1) * (long)ny].bY) / (2.0 * dx)
           - (U[j + 2 + i * (long)ny].bX - U[k].bX) / (2.0 * dy);
        eta = eta_of(ef, eta0, k + 1);
        const double etaJX1 = eta * jX, etaJZ1 = eta * jZ;
        const double etaJXBZ1 = etaJX1 * U[k + 1].bZ, etaJZBX1 = etaJZ1 * U[k + 1].bX;
        flux[k].f4 += 0.5 * (etaJZ + etaJZ1);
        flux[k].f6 -= 0.5 * (etaJX + etaJX1);
        flux[k].f7 += 0.5 * (etaJZBX + etaJZBX1) - 0.5 * (etaJXBZ + etaJXBZ1);
    });
}

// FluxSolver::getFluxF / getFluxG (P/flux_solver.cu:10-33)
void get_flux_f(Twin &t, thrust::device_vector<Cons> &U, thrust::device_vector<Flux> &out)
{
    hlld_sweep(t, raw_pointer_cast(U.data()), t.ny);
    thrust::copy(t.hlldFlux.begin(), t.hlldFlux.end(), t.solverFlux.begin());
    resistive_x(t, raw_pointer_cast(U.data()), raw_pointer_cast(t.solverFlux.data()));
    thrust::copy(t.solverFlux.begin(), t.solverFlux.end(), out.begin());
}

void get_flux_g(Twin &t, thrust::device_vector<Cons> &U, thrust::device_vector<Flux> &out)
{
    // rotation (u, v, w) -> (v, w, u), (Bx, By, Bz) -> (By, Bz, Bx) (P/hlld.cu:134-152)
    thrust::transform(U.begin(), U.end(), t.tmpU.begin(), [](const Cons &c) {
        Cons r;
        r.rho = c.rho; r.rhoU = c.rhoV; r.rhoV = c.rhoW; r.rhoW = c.rhoU;
        r.bX = c.bY; r.bY = c.bZ; r.bZ = c.bX; r.e = c.e;
        return r;
    });
    hlld_sweep(t, raw_pointer_cast(t.tmpU.data()), 1);
    // and back (P/hlld.cu:170-194)
    thrust::transform(t.hlldFlux.begin(), t.hlldFlux.end(), t.hlldFlux.begin(), [](const Flux &f) {
        Flux g;
        g.f0 = f.f0; g.f1 = f.f3; g.f2 = f.f1; g.f3 = f.f2;
        g.f4 = f.f6; g.f5 = f.f4; g.f6 = f.f5; g.f7 = f.f7;
        return g;
    });
    thrust::copy(t.hlldFlux.begin(), t.hlldFlux.end(), t.solverFlux.begin());
    resistive_y(t, raw_pointer_cast(U.data()), raw_pointer_cast(t.solverFlux.data()));
    thrust::copy(t.solverFlux.begin(), t.solverFlux.end(), out.begin());
}

// ------------------------------------------------- centring of face B
// shiftUToCenterForCT (P/resistiveMHD_2D.cu:313-368)
void shift_to_center(Twin &t, thrust::device_vector<Cons> &Uv)
{
    Cons *U = raw_pointer_cast(Uv.data());
    double *tmp = raw_pointer_cast(t.tmpVector.data());
    const int ny = t.ny;
    thrust::transform(Uv.begin(), Uv.end(), t.tmpVector.begin(), [](const Cons &c) { return c.bX; });
    each2(1, t.nx, 0, ny, [=](int i, int j) { U[j + (long)i * ny].bX = 0.5 * (tmp[j + (long)i * ny] + tmp[j + (long)(i - 1) * ny]); });
    thrust::transform(Uv.begin(), Uv.end(), t.tmpVector.begin(), [](const Cons &c) { return c.bY; });
    each2(0, t.nx, 1, ny, [=](int i, int j) { U[j + (long)i * ny].bY = 0.5 * (tmp[j + (long)i * ny] + tmp[j - 1 + (long)i * ny]); });
}

// backUToCenterHalfForCT (P/resistiveMHD_2D.cu:371-426)
void back_to_center_half(Twin &t, thrust::device_vector<Cons> &Uv)
{
    Cons *U = raw_pointer_cast(Uv.data());
    double *tmp = raw_pointer_cast(t.tmpVector.data());
    const int ny = t.ny;
    thrust::transform(Uv.begin(), Uv.end(), t.tmpVector.begin(), [](const Cons &c) { return c.bX; });
    each2(0, t.nx - 1, 0, ny, [=](int i, int j) { U[j + (long)i * ny].bX = 0.5 * (tmp[j + (long)i * ny] + tmp[j + (long)(i + 1) * ny]); });
    thrust::transform(Uv.begin(), Uv.end(), t.tmpVector.begin(), [](const Cons &c) { return c.bY; });
    each2(0, t.nx, 0, ny - 1, [=](int i, int j) { U[j + (long)i * ny].bY = 0.5 * (tmp[j + (long)i * ny] + tmp[j + 1 + (long)i * ny]); });
}

// ------------------------------------------------------------------- CT
// CT::divBClean (P/ct.cu:22-85) with the fluxes saved by setOldFlux2D
void div_b_clean(Twin &t, thrust::device_vector<Cons> &Uv)
{
    const Flux *F = raw_pointer_cast(t.oldFluxF.data()), *G = raw_pointer_cast(t.oldFluxG.data());
    double *eZ = raw_pointer_cast(t.eZVector.data());
    const double *bxo = raw_pointer_cast(t.bXOld.data()), *byo = raw_pointer_cast(t.bYOld.data());
    Cons *U = raw_pointer_cast(Uv.data());
    const int ny = t.ny;
    const double dt = t.dt, dx = t.dx, dy = t.dy;
    each2(0, t.nx - 1, 0, ny - 1, [=](int i, int j) {
        const long k = j + (long)i * ny;
        const double eZG1 = G[k].f4, eZG2 = G[k + ny].f4;
        const double eZF1 = -1.0 * F[k].f5, eZF2 = -1.0 * F[k + 1].f5;
        eZ[k] = 0.25 * (eZG1 + eZG2 + eZF1 + eZF2);
    });
    each2(1, t.nx, 1, ny, [=](int i, int j) {
        const long k = j + (long)i * ny;
        U[k].bX = bxo[k] - dt / dy * (eZ[k] - eZ[k - 1]);
        U[k].bY = byo[k] + dt / dx * (eZ[k] - eZ[k - ny]);
    });
}

// ------------------------------------------------------------- boundary
// P/boundary.cu:4-141 and S/boundary.cu:4-205: one element per column
// (x) or row (y), as the reference's kernels
void boundary(Twin &t, thrust::device_vector<Cons> &Uv)
{
    Cons *U = raw_pointer_cast(Uv.data());
    const int nx = t.nx, ny = t.ny;
    const double dx = t.dx, dy = t.dy;
    auto at = [=](int i, int j) -> Cons & { return U[j + (long)i * ny]; };
    if (t.bc_x == BC_PERIODIC) {
        each(0, ny, [=](long j) {
            at(0, j) = at(nx - 6, j); at(1, j) = at(nx - 5, j); at(2, j) = at(nx - 4, j);
            at(nx - 3, j) = at(3, j); at(nx - 2, j) = at(4, j); at(nx - 1, j) = at(5, j);
        });
    } else {
        each(0, ny, [=](long j) {
            const int src[6] = {5, 4, 3, nx - 6, nx - 5, nx - 4}, dst[6] = {0, 1, 2, nx - 1, nx - 2, nx - 3};
            for (int q = 0; q < 6; ++q) {
                Cons &d = at(dst[q], j);
                const Cons &s = at(src[q], j);
                d.rho = s.rho; d.rhoU = s.rhoU; d.rhoV = s.rhoV; d.rhoW = s.rhoW; d.bY = s.bY; d.bZ = s.bZ; d.e = s.e;
            }
        });
        each(0, ny, [=](long j) {
            if (0 < j) {
                at(2, j).bX = at(3, j).bX + (at(3, j).bY - at(3, j - 1).bY) / dy * dx;
                at(1, j).bX = at(2, j).bX + (at(2, j).bY - at(2, j - 1).bY) / dy * dx;
                at(0, j).bX = at(1, j).bX + (at(1, j).bY - at(1, j - 1).bY) / dy * dx;
                at(nx - 3, j).bX = at(nx - 4, j).bX - (at(nx - 3, j).bY - at(nx - 3, j - 1).bY) / dy * dx;
                at(nx - 2, j).bX = at(nx - 3, j).bX - (at(nx - 2, j).bY - at(nx - 2, j - 1).bY) / dy * dx;
                at(nx - 1, j).bX = at(nx - 2, j).bX - (at(nx - 1, j).bY - at(nx - 1, j - 1).bY) / dy * dx;
            } else {
                at(2, j).bX = at(3, j).bX; at(1, j).bX = at(2, j).bX; at(0, j).bX = at(1, j).bX;
                at(nx - 3, j).bX = at(nx - 4, j).bX; at(nx - 2, j).bX = at(nx - 3, j).bX;
                at(nx - 1, j).bX = at(nx - 2, j).bX;
            }
        });
    }
    if (t.bc_y == BC_PERIODIC) {
        each(0, nx, [=](long i) {
            at(i, 0) = at(i, ny - 6); at(i, 1) = at(i, ny - 5); at(i, 2) = at(i, ny - 4);
            at(i, ny - 3) = at(i, 3); at(i, ny - 2) = at(i, 4); at(i, ny - 1) = at(i, 5);
        });
    } else {
        each(0, nx, [=](long i) {
            const int src[6] = {5, 4, 3, ny - 6, ny - 5, ny - 4}, dst[6] = {0, 1, 2, ny - 1, ny - 2, ny - 3};
            for (int q = 0; q < 6; ++q) {
                Cons &d = at(i, dst[q]);
                const Cons &s = at(i, src[q]);
                d.rho = s.rho; d.rhoU = s.rhoU; d.rhoV = s.rhoV; d.rhoW = s.rhoW; d.bX = s.bX; d.bZ = s.bZ; d.e = s.e;
            }
        });
        each(0, nx, [=](long i) {
            if (0 < i) {
                at(i, 2).bY = at(i, 3).bY + (at(i, 3).bX - at(i - 1, 3).bX) / dx * dy;
                at(i, 1).bY = at(i, 2).bY + (at(i, 2).bX - at(i - 1, 2).bX) / dx * dy;
                at(i, 0).bY = at(i, 1).bY + (at(i, 1).bX - at(i - 1, 1).bX) / dx * dy;
                at(i, ny - 3).bY = at(i, ny - 4).bY - (at(i, ny - 3).bX - at(i - 1, ny - 3).bX) / dx * dy;
                at(i, ny - 2).bY = at(i, ny - 3).bY - (at(i, ny - 2).bX - at(i - 1, ny - 2).bX) / dx * dy;
                at(i, ny - 1).bY = at(i, ny - 2).bY - (at(i, ny - 1).bX - at(i - 1, ny - 1).bX) / dx * dy;
            } else {
                at(i, 2).bY = at(i, 3).bY; at(i, 1).bY = at(i, 2).bY; at(i, 0).bY = at(i, 1).bY;
                at(i, ny - 3).bY = at(i, ny - 4).bY; at(i, ny - 2).bY = at(i, ny - 3).bY;
                at(i, ny - 1).bY = at(i, ny - 2).bY;
            }
        });
    }
}

// ------------------------------------------------------------------- dt
// ResistiveMHD2D::calculateDt (P/resistiveMHD_2D.cu:277-291)
void calculate_dt(Twin &t)
{
    const double gamma = t.gamma, dx = t.dx, dy = t.dy, eps = t.eps;
    thrust::transform(t.U.begin(), t.U.end(), t.dtVector.begin(),
                      [=](const Cons &c) { return cell_dt(&c, gamma, dx, dy, eps); });
    const double m = *thrust::min_element(t.dtVector.begin(), t.dtVector.end());
    const double lim = std::fmin(t.dx * t.dx / t.eta, t.dy * t.dy / t.eta);
    t.dt = std::fmin(m, lim) * t.cfl;
}

// ResistiveMHD2D::oneStepRK2 (P/resistiveMHD_2D.cu:133-195)
void one_step(Twin &t)
{
    const long N = t.N, ny = t.ny;
    thrust::transform(t.U.begin(), t.U.end(), t.bXOld.begin(), [](const Cons &c) { return c.bX; });
    thrust::transform(t.U.begin(), t.U.end(), t.bYOld.begin(), [](const Cons &c) { return c.bY; });
    thrust::copy(t.U.begin(), t.U.end(), t.UBar.begin());
    calculate_dt(t);

    shift_to_center(t, t.U);
    get_flux_f(t, t.U, t.fluxF);
    get_flux_g(t, t.U, t.fluxG);
    back_to_center_half(t, t.U);
    {
        const Cons *U = raw_pointer_cast(t.U.data());
        const Flux *F = raw_pointer_cast(t.fluxF.data()), *G = raw_pointer_cast(t.fluxG.data());
        const double dt = t.dt, dx = t.dx, dy = t.dy;
        thrust::transform(Count(ny), Count(N), t.UBar.begin() + ny, [=](long k) {
            return stage_first(&U[k], &F[k], &F[k - ny], &G[k], &G[k - 1], dt, dx, dy);
        });
    }
    thrust::copy(t.fluxF.begin(), t.fluxF.end(), t.oldFluxF.begin());   // CT::setOldFlux2D
    thrust::copy(t.fluxG.begin(), t.fluxG.end(), t.oldFluxG.begin());
    div_b_clean(t, t.UBar);
    boundary(t, t.UBar);

    shift_to_center(t, t.UBar);
    get_flux_f(t, t.UBar, t.fluxF);
    get_flux_g(t, t.UBar, t.fluxG);
    back_to_center_half(t, t.UBar);
    {
        Cons *U = raw_pointer_cast(t.U.data());
        const Cons *B = raw_pointer_cast(t.UBar.data());
        const Flux *F = raw_pointer_cast(t.fluxF.data()), *G = raw_pointer_cast(t.fluxG.data());
        const double dt = t.dt, dx = t.dx, dy = t.dy;
        thrust::transform(Count(ny), Count(N), t.U.begin() + ny, [=](long k) {
            return stage_second(&U[k], &B[k], &F[k], &F[k - ny], &G[k], &G[k - 1], dt, dx, dy);
        });
    }
    div_b_clean(t, t.U);
    boundary(t, t.U);
    t.time += t.dt;
}

// ResistiveMHD2D::checkCalculationIsCrashed (P/resistiveMHD_2D.cu:294-309)
bool crashed(Twin &t)
{
    return thrust::transform_reduce(t.U.begin(), t.U.end(), [](const Cons &c) { return (bool)std::isnan(c.e); },
                                    false, thrust::logical_or<bool>());
}

}  // namespace

extern "C" {

void *twin_create(int nx, int ny, double dx, double dy, double gamma, double cfl, double eps, double eta, int bc_x,
                  int bc_y)
{
    if (nx < 12 || ny < 12) return nullptr;
    Twin *t = new (std::nothrow) Twin(nx, ny);
    if (!t) return nullptr;
    t->dx = dx; t->dy = dy; t->gamma = gamma; t->cfl = cfl; t->eps = eps; t->eta = eta;
    t->bc_x = bc_x; t->bc_y = bc_y;
    return t;
}

void twin_destroy(void *p) { delete static_cast<Twin *>(p); }

void twin_set_state(void *p, const double *aos)
{
    Twin *t = static_cast<Twin *>(p);
    thrust::copy(reinterpret_cast<const Cons *>(aos), reinterpret_cast<const Cons *>(aos) + t->N, t->U.begin());
}

void twin_get_state(void *p, double *aos)
{
    Twin *t = static_cast<Twin *>(p);
    thrust::copy(t->U.begin(), t->U.end(), reinterpret_cast<Cons *>(aos));
}

void twin_set_eta_field(void *p, const double *eta)
{
    Twin *t = static_cast<Twin *>(p);
    if (!eta) {
        t->has_eta = false;
        return;
    }
    t->etaField.assign(eta, eta + t->N);
    t->has_eta = true;
}

void twin_apply_boundary(void *p) { boundary(*static_cast<Twin *>(p), static_cast<Twin *>(p)->U); }

// nsteps steps of the examples' time loop: oneStepRK2, then the crash check
// (EX/plasmoid/main_plasmoid.cu:305-326).  Returns 1 if a step crashed.
int twin_step(void *p, int nsteps)
{
    Twin *t = static_cast<Twin *>(p);
    for (int n = 0; n < nsteps; ++n) {
        one_step(*t);
        if (crashed(*t)) return 1;
    }
    return 0;
}

double twin_get_dt(void *p) { return static_cast<Twin *>(p)->dt; }
double twin_get_time(void *p) { return static_cast<Twin *>(p)->time; }
void twin_set_threads(int n) { omp_set_num_threads(n); }
int twin_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
