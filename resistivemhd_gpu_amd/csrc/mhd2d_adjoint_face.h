// mhd2d_adjoint_face.h -- assembly of the per-face flux of the RK step from
// the cell states, and its tangent-linear and adjoint, shared by the HIP
// adjoint kernels and host code.  The generated pieces (mhd2d_face_ad.h)
// take the centred states of the four cells of a face's MUSCL stencil and
// the resistive stencil; this header gathers those inputs from a state
// accessor, rotates the y sweep into the x frame and back
// (ResistiveMHD_singleGPU_thrust/lib_ResistiveMHD_2D_GPU_periodicXY/
// hlld.cu:134-233), adds the resistive terms (flux_solver.cu:10-33) and
// scatters the input adjoints back to cells.
//
// A state accessor S provides `double operator()(int v, int i, int j)`
// (component v of cell (i, j), face-centred B as stored); an eta accessor E
// provides `double operator()(int i, int j)`.
#pragma once
#include "mhd2d_face_ad.h"

// Loops over stencil slots are unrolled so that accessors see constant
// offsets from the face (a ring-buffer accessor indexes small arrays by them).
#if defined(__clang__)
#define MHD_UNROLL _Pragma("unroll")
#elif defined(__GNUC__)
#define MHD_UNROLL _Pragma("GCC unroll 16")
#else
#define MHD_UNROLL
#endif

namespace mhd_ad {

// Centred component (shiftUToCenterForCT, resistiveMHD_2D.cu:313-368):
// Bx averaged over i-1/2, i+1/2 (for i >= 1), By over j-1/2, j+1/2 (j >= 1).
template <class S>
MHD_HD inline double centred(const S &U, int v, int i, int j)
{
    if (v == 4) return i >= 1 ? 0.5 * (U(4, i, j) + U(4, i - 1, j)) : U(4, i, j);
    if (v == 5) return j >= 1 ? 0.5 * (U(5, i, j) + U(5, i, j - 1)) : U(5, i, j);
    return U(v, i, j);
}

// Component of the unrotated state that slot v of the y sweep's rotated
// state holds: (rho, rho v, rho w, rho u, By, Bz, Bx, e) (hlld.cu:134-152).
MHD_HD inline int rot_y(int v)
{
    constexpr int r[8] = {0, 2, 3, 1, 5, 6, 4, 7};
    return r[v];
}

// Cell of stencil slot c (0..3 = A..D) of face (i, j): x faces sit between
// (i, j) and (i+1, j), y faces between (i, j) and (i, j+1).
MHD_HD inline void stencil_cell(int dir, int i, int j, int c, int &ci, int &cj)
{
    ci = dir == 0 ? i - 1 + c : i;
    cj = dir == 0 ? j : j - 1 + c;
}

template <class S>
MHD_HD inline void gather_face(const S &U, int dir, int i, int j, double *in)
{
    MHD_UNROLL
    for (int c = 0; c < 4; ++c) {
        int ci, cj;
        stencil_cell(dir, i, j, c, ci, cj);
        MHD_UNROLL
        for (int v = 0; v < 8; ++v) in[8 * c + v] = centred(U, dir == 0 ? v : rot_y(v), ci, cj);
    }
}

// Resistive stencil of face (i, j): see resistive_x.ssa / resistive_y.ssa.
template <class S, class E>
MHD_HD inline void gather_resist(const S &U, const E &eta, int dir, int i, int j, double *in)
{
    if (dir == 0) {
        MHD_UNROLL
        for (int c = 0; c < 4; ++c) {
            in[c] = centred(U, 6, i - 1 + c, j);
            in[4 + c] = centred(U, 5, i - 1 + c, j);
        }
        in[8] = centred(U, 4, i, j - 1);
        in[9] = centred(U, 4, i, j + 1);
        in[10] = centred(U, 4, i + 1, j - 1);
        in[11] = centred(U, 4, i + 1, j + 1);
        in[12] = eta(i, j);
        in[13] = eta(i + 1, j);
    } else {
        MHD_UNROLL
        for (int c = 0; c < 4; ++c) {
            in[c] = centred(U, 6, i, j - 1 + c);
            in[4 + c] = centred(U, 4, i, j - 1 + c);
        }
        in[8] = centred(U, 5, i - 1, j);
        in[9] = centred(U, 5, i + 1, j);
        in[10] = centred(U, 5, i - 1, j + 1);
        in[11] = centred(U, 5, i + 1, j + 1);
        in[12] = eta(i, j);
        in[13] = eta(i, j + 1);
    }
}

// Cell and component each resistive input comes from (eta: v = 8).
MHD_HD inline void resist_slot(int dir, int i, int j, int s, int &ci, int &cj, int &v)
{
    if (dir == 0) {
        if (s < 8) { ci = i - 1 + (s & 3); cj = j; v = s < 4 ? 6 : 5; return; }
        if (s < 12) { ci = s < 10 ? i : i + 1; cj = (s & 1) ? j + 1 : j - 1; v = 4; return; }
        ci = s == 12 ? i : i + 1; cj = j; v = 8;
    } else {
        if (s < 8) { ci = i; cj = j - 1 + (s & 3); v = s < 4 ? 6 : 4; return; }
        if (s < 12) { ci = (s & 1) ? i + 1 : i - 1; cj = s < 10 ? j : j + 1; v = 5; return; }
        ci = i; cj = s == 12 ? j : j + 1; v = 8;
    }
}

// Whether an adder skips zero contributions: a global atomic costs a memory
// transaction, so a zero one is worth a test; an LDS add is cheaper than
// the test and the exec-mask change around it.
template <class Add>
struct add_skips_zero {
    static constexpr bool value = true;
};

// Scatter of one face's input adjoints: in_b (4 cells x 8 slots, rotated
// for y faces) and the resistive stencil's remaining slots 8..13.
template <class Add>
MHD_HD inline void scatter_face(int dir, int i, int j, const double *in_b, const double *rin_b, Add &add)
{
    MHD_UNROLL
    for (int c = 0; c < 4; ++c) {
        int ci, cj;
        stencil_cell(dir, i, j, c, ci, cj);
        MHD_UNROLL
        for (int v = 0; v < 8; ++v) {
            const double b = in_b[8 * c + v];
            if (!add_skips_zero<Add>::value || b != 0.0) add(ci, cj, dir == 0 ? v : rot_y(v), b);
        }
    }
    MHD_UNROLL
    for (int s = 8; s < 14; ++s) {
        if (add_skips_zero<Add>::value && rin_b[s] == 0.0) continue;
        int ci, cj, v;
        resist_slot(dir, i, j, s, ci, cj, v);
        add(ci, cj, v, rin_b[s]);
    }
}

// Full flux (conserved order f0..f7) of face (i, j): HLLD of the MUSCL
// states plus the resistive terms.  par = {gamma - 1, gamma, eps},
// rpar = {dx, dy}.
template <class S, class E>
MHD_HD inline void flux_full(const S &U, const E &eta, int dir, int i, int j, const double *par, const double *rpar,
                             double *F)
{
    double in[32], ff[7], rin[14], d[3];
    gather_face(U, dir, i, j, in);
    face_flux(in, par, ff);
    gather_resist(U, eta, dir, i, j, rin);
    if (dir == 0) {
        resist_x(rin, rpar, d);
        F[0] = ff[0]; F[1] = ff[1]; F[2] = ff[2]; F[3] = ff[3]; F[4] = 0.0;
        F[5] = ff[4] - d[0];
        F[6] = ff[5] + d[1];
        F[7] = ff[6] + d[2];
    } else {
        resist_y(rin, rpar, d);
        // rotate back (hlld.cu:170-194), then the resistive terms
        F[0] = ff[0]; F[1] = ff[3]; F[2] = ff[1]; F[3] = ff[2];
        F[4] = ff[5] + d[0];
        F[5] = 0.0;
        F[6] = ff[4] - d[1];
        F[7] = ff[6] + d[2];
    }
}

// Adjoint of flux_full: for the flux adjoint lamF (conserved order), calls
// add(ci, cj, v, value) for every centred-state input (v = 0..7) and
// add(ci, cj, 8, value) for every eta input.  The direction-dependent
// gathers and scatters sit in branches on dir around one copy of the
// generated face adjoint, so a caller with a run-time (wave-uniform) dir
// compiles each with constant stencil offsets.
template <class S, class E, class Add>
MHD_HD inline void flux_full_adj(const S &U, const E &eta, int dir, int i, int j, const double *par,
                                 const double *rpar, const double *lamF, Add &add)
{
    double in[32], in_b[32], ff_b[7], rin[14], rin_b[14], d_b[3];
    if (dir == 0) {
        gather_face(U, 0, i, j, in);
        gather_resist(U, eta, 0, i, j, rin);
        ff_b[0] = lamF[0]; ff_b[1] = lamF[1]; ff_b[2] = lamF[2]; ff_b[3] = lamF[3];
        ff_b[4] = lamF[5]; ff_b[5] = lamF[6]; ff_b[6] = lamF[7];
        d_b[0] = -lamF[5]; d_b[1] = lamF[6]; d_b[2] = lamF[7];
        resist_x_adj(rin, rpar, d_b, rin_b);
    } else {
        gather_face(U, 1, i, j, in);
        gather_resist(U, eta, 1, i, j, rin);
        ff_b[0] = lamF[0]; ff_b[1] = lamF[2]; ff_b[2] = lamF[3]; ff_b[3] = lamF[1];
        ff_b[4] = lamF[6]; ff_b[5] = lamF[4]; ff_b[6] = lamF[7];
        d_b[0] = lamF[4]; d_b[1] = -lamF[6]; d_b[2] = lamF[7];
        resist_y_adj(rin, rpar, d_b, rin_b);
    }
    face_flux_adj(in, par, ff_b, in_b);
    // the resistive terms' Bz and By (x faces) / Bz and Bx (y faces) inputs
    // belong to the same four cells: fold them in before scattering
    if (dir == 0) {
        MHD_UNROLL
        for (int c = 0; c < 4; ++c) {
            in_b[8 * c + 6] += rin_b[c];
            in_b[8 * c + 5] += rin_b[4 + c];
        }
        scatter_face(0, i, j, in_b, rin_b, add);
    } else {
        MHD_UNROLL
        for (int c = 0; c < 4; ++c) {
            in_b[8 * c + 5] += rin_b[c];       // rotated slot of Bz
            in_b[8 * c + 6] += rin_b[4 + c];   // rotated slot of Bx
        }
        scatter_face(1, i, j, in_b, rin_b, add);
    }
}

// Tangent of flux_full for a direction (dU of the stored state, deta).
template <class S, class E>
MHD_HD inline void flux_full_tlm(const S &U, const S &dU, const E &eta, const E &deta, int dir, int i, int j,
                                 const double *par, const double *rpar, double *F, double *dF)
{
    double in[32], in_d[32], ff[7], ff_d[7], rin[14], rin_d[14], d[3], d_d[3];
    gather_face(U, dir, i, j, in);
    gather_face(dU, dir, i, j, in_d);
    face_flux_tlm(in, in_d, par, ff, ff_d);
    gather_resist(U, eta, dir, i, j, rin);
    gather_resist(dU, deta, dir, i, j, rin_d);
    if (dir == 0) {
        resist_x_tlm(rin, rin_d, rpar, d, d_d);
        F[0] = ff[0]; F[1] = ff[1]; F[2] = ff[2]; F[3] = ff[3]; F[4] = 0.0;
        F[5] = ff[4] - d[0]; F[6] = ff[5] + d[1]; F[7] = ff[6] + d[2];
        dF[0] = ff_d[0]; dF[1] = ff_d[1]; dF[2] = ff_d[2]; dF[3] = ff_d[3]; dF[4] = 0.0;
        dF[5] = ff_d[4] - d_d[0]; dF[6] = ff_d[5] + d_d[1]; dF[7] = ff_d[6] + d_d[2];
    } else {
        resist_y_tlm(rin, rin_d, rpar, d, d_d);
        F[0] = ff[0]; F[1] = ff[3]; F[2] = ff[1]; F[3] = ff[2];
        F[4] = ff[5] + d[0]; F[5] = 0.0; F[6] = ff[4] - d[1]; F[7] = ff[6] + d[2];
        dF[0] = ff_d[0]; dF[1] = ff_d[3]; dF[2] = ff_d[1]; dF[3] = ff_d[2];
        dF[4] = ff_d[5] + d_d[0]; dF[5] = 0.0; dF[6] = ff_d[4] - d_d[1]; dF[7] = ff_d[6] + d_d[2];
    }
}

// Decision replay of flux_full: FP is the flux of the perturbed state P
// (and perturbed resistivity etaP) with every branch decision (minmod,
// HLLD speeds and region, signs, the degenerate-limit switch) taken at the
// base state U; F is the flux of U.  Central differences of the rollout
// built from it check the tangent and the adjoint without the jumps of the
// branches themselves.
template <class S, class E>
MHD_HD inline void flux_full_rep(const S &U, const S &P, const E &eta, const E &etaP, int dir, int i, int j,
                                 const double *par, const double *rpar, double *F, double *FP)
{
    double in[32], in_p[32], ff[7], ff_p[7], rin[14], rin_p[14], d[3], d_p[3];
    gather_face(U, dir, i, j, in);
    gather_face(P, dir, i, j, in_p);
    face_flux_rep(in, in_p, par, ff, ff_p);
    gather_resist(U, eta, dir, i, j, rin);
    gather_resist(P, etaP, dir, i, j, rin_p);
    if (dir == 0) {
        resist_x_rep(rin, rin_p, rpar, d, d_p);
        F[0] = ff[0]; F[1] = ff[1]; F[2] = ff[2]; F[3] = ff[3]; F[4] = 0.0;
        F[5] = ff[4] - d[0]; F[6] = ff[5] + d[1]; F[7] = ff[6] + d[2];
        FP[0] = ff_p[0]; FP[1] = ff_p[1]; FP[2] = ff_p[2]; FP[3] = ff_p[3]; FP[4] = 0.0;
        FP[5] = ff_p[4] - d_p[0]; FP[6] = ff_p[5] + d_p[1]; FP[7] = ff_p[6] + d_p[2];
    } else {
        resist_y_rep(rin, rin_p, rpar, d, d_p);
        F[0] = ff[0]; F[1] = ff[3]; F[2] = ff[1]; F[3] = ff[2];
        F[4] = ff[5] + d[0]; F[5] = 0.0; F[6] = ff[4] - d[1]; F[7] = ff[6] + d[2];
        FP[0] = ff_p[0]; FP[1] = ff_p[3]; FP[2] = ff_p[1]; FP[3] = ff_p[2];
        FP[4] = ff_p[5] + d_p[0]; FP[5] = 0.0; FP[6] = ff_p[4] - d_p[1]; FP[7] = ff_p[6] + d_p[2];
    }
}

}  // namespace mhd_ad
