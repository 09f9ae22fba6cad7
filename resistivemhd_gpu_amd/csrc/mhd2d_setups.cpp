// mhd2d_setups.cpp -- initial states of the benchmark and parity workloads,
// computed on the host into the reference's array-of-structs layout
// (8 doubles per cell, k = j + i*ny, ghosts included).  They stand in for
// the examples' user-written initializeU kernels; face B is sampled at the
// face positions and the energy uses the cell-centred field, as
// ResistiveMHD_singleGPU_thrust/examples2D/plasmoid/main_plasmoid.cu:85-126
// does.  Cell i's centre is x = xmin + (i - 3 + 1/2) dx.  In both setups the
// x-face field depends on y only and the y-face field on x only, so the face
// samples equal the centre samples and B is discretely divergence-free.
#include <cmath>
#include <cstdint>
#include "../../include/mhd2d.h"

namespace {

inline double xc(const mhd_config *c, int i) { return c->xmin + (i - MHD_GHOST + 0.5) * c->dx; }
inline double yc(const mhd_config *c, int j) { return c->ymin + (j - MHD_GHOST + 0.5) * c->dy; }

// splitmix64 finaliser of a (seed, i, j) counter: reproducible noise that
// does not depend on the order cells are visited in.
inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double uniform01(uint64_t seed, int i, int j)
{
    const uint64_t key = (seed << 42) ^ ((uint64_t)(uint32_t)i << 21) ^ (uint64_t)(uint32_t)j;
    return (double)(mix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

inline void put(double *aos, long k, double rho, double u, double v, double w, double bx, double by,
                double bz, double bxc, double byc, double p, double gamma)
{
    double *c = aos + 8 * k;
    c[0] = rho;
    c[1] = rho * u;
    c[2] = rho * v;
    c[3] = rho * w;
    c[4] = bx;
    c[5] = by;
    c[6] = bz;
    c[7] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (bxc * bxc + byc * byc + bz * bz);
}

bool valid(const mhd_config *c) { return c && c->nx >= 4 * MHD_GHOST && c->ny >= 4 * MHD_GHOST; }

}  // namespace

extern "C" int mhd_fill_orszag_tang(const mhd_config *cfg, double *aos)
{
    if (!valid(cfg) || !aos) return MHD_ERR_INVALID;
    const double g = cfg->gamma;
    for (int i = 0; i < cfg->nx; ++i) {
        for (int j = 0; j < cfg->ny; ++j) {
            const double x = xc(cfg, i), y = yc(cfg, j);
            const double u = -std::sin(y), v = std::sin(x);
            const double bxc = -std::sin(y), byc = std::sin(2.0 * x);
            const double bx = -std::sin(y);                 // x-face (i+1/2, j)
            const double by = std::sin(2.0 * x);            // y-face (i, j+1/2)
            put(aos, (long)j + (long)i * cfg->ny, g * g, u, v, 0.0, bx, by, 0.0, bxc, byc, g, g);
        }
    }
    return MHD_OK;
}

extern "C" int mhd_fill_harris(const mhd_config *cfg, unsigned long long seed, double noise_amp, double *aos)
{
    if (!valid(cfg) || !aos) return MHD_ERR_INVALID;
    const double g = cfg->gamma;
    for (int i = 0; i < cfg->nx; ++i) {
        for (int j = 0; j < cfg->ny; ++j) {
            const double y = yc(cfg, j);
            const double ch = std::cosh(y);
            const double rho = 1.0 / (ch * ch) + 0.2;
            const double p = 0.5 * rho;
            const double vy = noise_amp * (2.0 * uniform01(seed, i, j) - 1.0);
            const double bx = std::tanh(y);
            put(aos, (long)j + (long)i * cfg->ny, rho, 0.0, vy, 0.0, bx, 0.0, 0.0, bx, 0.0, p, g);
        }
    }
    return MHD_OK;
}

extern "C" int mhd_fill_harris_eta(const mhd_config *cfg, double eta0, double eta1, double *eta)
{
    if (!valid(cfg) || !eta) return MHD_ERR_INVALID;
    for (int i = 0; i < cfg->nx; ++i) {
        for (int j = 0; j < cfg->ny; ++j) {
            const double x = xc(cfg, i), y = yc(cfg, j);
            eta[(long)j + (long)i * cfg->ny] = eta0 + eta1 * std::exp(-(x * x + y * y) / 4.0);
        }
    }
    return MHD_OK;
}

extern "C" int mhd_default_config(mhd_config *cfg, int nx_int, int ny_int, double xmin, double lx, double ymin,
                                  double ly)
{
    if (!cfg || nx_int < 2 * MHD_GHOST || ny_int < 2 * MHD_GHOST || !(lx > 0.0) || !(ly > 0.0))
        return MHD_ERR_INVALID;
    cfg->nx = nx_int + 2 * MHD_GHOST;
    cfg->ny = ny_int + 2 * MHD_GHOST;
    cfg->dx = lx / nx_int;
    cfg->dy = ly / ny_int;
    cfg->xmin = xmin;
    cfg->ymin = ymin;
    cfg->gamma = 5.0 / 3.0;
    cfg->cfl = 0.7;
    cfg->eps = 1e-20;
    cfg->eta = 0.0;
    cfg->bc_x = MHD_BC_PERIODIC;
    cfg->bc_y = MHD_BC_PERIODIC;
    return MHD_OK;
}
