// Host build of the flux assembly and its generated tangent/adjoint
// (resistivemhd_gpu_amd/csrc/mhd2d_adjoint_face.h), for the CPU tests:
// compiled with g++ at test time and called through ctypes.
#include "mhd2d_adjoint_face.h"

namespace {
struct Aos {
    const double *u;
    int ny;
    double operator()(int v, int i, int j) const { return u[((long)i * ny + j) * 8 + v]; }
};
struct Eta {
    const double *e;
    int ny;
    double uniform;
    double operator()(int i, int j) const { return e ? e[(long)i * ny + j] : uniform; }
};
struct AddGrid {
    double *lam;    // (nx, ny, 8)
    double *geta;   // (nx, ny)
    int ny;
    void operator()(int i, int j, int v, double b)
    {
        if (v == 8) geta[(long)i * ny + j] += b;
        else lam[((long)i * ny + j) * 8 + v] += b;
    }
};
}  // namespace

extern "C" {

void h_face_flux(const double *in, const double *par, double *out) { mhd_ad::face_flux(in, par, out); }
void h_face_flux_tlm(const double *in, const double *in_d, const double *par, double *out, double *out_d)
{
    mhd_ad::face_flux_tlm(in, in_d, par, out, out_d);
}
void h_face_flux_adj(const double *in, const double *par, const double *out_b, double *in_b)
{
    mhd_ad::face_flux_adj(in, par, out_b, in_b);
}
void h_resist_tlm(int dir, const double *in, const double *in_d, const double *par, double *out, double *out_d)
{
    if (dir == 0) mhd_ad::resist_x_tlm(in, in_d, par, out, out_d);
    else mhd_ad::resist_y_tlm(in, in_d, par, out, out_d);
}
void h_resist_adj(int dir, const double *in, const double *par, const double *out_b, double *in_b)
{
    if (dir == 0) mhd_ad::resist_x_adj(in, par, out_b, in_b);
    else mhd_ad::resist_y_adj(in, par, out_b, in_b);
}

// Flux of face (i, j) of direction dir of an AoS state (nx, ny, 8).
void h_flux_full(const double *U, const double *eta, double eta0, int nx, int ny, int dir, int i, int j,
                 const double *par, const double *rpar, double *F)
{
    (void)nx;
    const Aos S{U, ny};
    const Eta E{eta, ny, eta0};
    mhd_ad::flux_full(S, E, dir, i, j, par, rpar, F);
}

void h_flux_full_tlm(const double *U, const double *dU, const double *eta, const double *deta, double eta0,
                     int nx, int ny, int dir, int i, int j, const double *par, const double *rpar, double *F,
                     double *dF)
{
    (void)nx;
    const Aos S{U, ny}, dS{dU, ny};
    const Eta E{eta, ny, eta0}, dE{deta, ny, 0.0};
    mhd_ad::flux_full_tlm(S, dS, E, dE, dir, i, j, par, rpar, F, dF);
}

void h_face_flux_rep(const double *in, const double *in_p, const double *par, double *out, double *out_p)
{
    mhd_ad::face_flux_rep(in, in_p, par, out, out_p);
}

void h_flux_full_rep(const double *U, const double *P, const double *eta, const double *etaP, double eta0, int nx,
                     int ny, int dir, int i, int j, const double *par, const double *rpar, double *F, double *FP)
{
    (void)nx;
    const Aos S{U, ny}, SP{P, ny};
    const Eta E{eta, ny, eta0}, EP{etaP, ny, eta0};
    mhd_ad::flux_full_rep(S, SP, E, EP, dir, i, j, par, rpar, F, FP);
}

// Adjoint of the face flux: adds into lamUc (centred-state adjoint, AoS)
// and geta.
void h_flux_full_adj(const double *U, const double *eta, double eta0, int nx, int ny, int dir, int i, int j,
                     const double *par, const double *rpar, const double *lamF, double *lamUc, double *geta)
{
    (void)nx;
    const Aos S{U, ny};
    const Eta E{eta, ny, eta0};
    AddGrid add{lamUc, geta, ny};
    mhd_ad::flux_full_adj(S, E, dir, i, j, par, rpar, lamF, add);
}

}  // extern "C"
