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
#include <thread>
#include <vector>
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

// One Fourier mode of the turbulence setup: wave numbers (per unit length),
// amplitude and phase of vx, vy and Az.
struct Mode {
    int nx, ny;
    double kx, ky, av, aa;
    double ph[3];
};

// Field value sum_n a_n cos(kx x + ky y + phi_n) as sum over nx of
// cos(kx x) Re P_nx(y) - sin(kx x) Im P_nx(y), P_nx(y) = sum_ny a e^{i(phi + ky y)}.
struct Separable {
    int nkx;                   // nx = 0..kmax
    std::vector<double> pr, pi;   // [nkx][rows]
};

Separable row_tables(const std::vector<Mode> &modes, int kmax, int field, double amp_scale, const double *y,
                     int rows)
{
    Separable t;
    t.nkx = kmax + 1;
    t.pr.assign((size_t)t.nkx * rows, 0.0);
    t.pi.assign((size_t)t.nkx * rows, 0.0);
    for (const Mode &m : modes) {
        const double a = (field == 2 ? m.aa : m.av) * amp_scale;
        for (int r = 0; r < rows; ++r) {
            const double arg = m.ky * y[r] + m.ph[field];
            t.pr[(size_t)m.nx * rows + r] += a * std::cos(arg);
            t.pi[(size_t)m.nx * rows + r] += a * std::sin(arg);
        }
    }
    return t;
}

inline double eval(const Separable &t, int rows, int r, const double *cx, const double *sx)
{
    double f = 0.0;
    for (int q = 0; q < t.nkx; ++q) f += cx[q] * t.pr[(size_t)q * rows + r] - sx[q] * t.pi[(size_t)q * rows + r];
    return f;
}

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

extern "C" int mhd_fill_turbulence(const mhd_config *cfg, unsigned long long seed, int kmax, double rms_v,
                                   double rms_b, double *aos)
{
    if (!valid(cfg) || !aos || kmax < 1 || !(rms_v >= 0.0) || !(rms_b >= 0.0)) return MHD_ERR_INVALID;
    const int nxi = cfg->nx - 2 * MHD_GHOST, nyi = cfg->ny - 2 * MHD_GHOST;
    if (2 * kmax >= nxi || 2 * kmax >= nyi) return MHD_ERR_INVALID;   // modes must be resolved
    const double two_pi = 6.283185307179586;
    const double lx = cfg->dx * nxi, ly = cfg->dy * nyi;
    // modes of the half plane (nx > 0, or nx = 0 and ny > 0) so that no two
    // are equal or opposite
    std::vector<Mode> modes;
    for (int a = 0; a <= kmax; ++a)
        for (int b = -kmax; b <= kmax; ++b) {
            if ((a == 0 && b <= 0) || a * a + b * b > kmax * kmax) continue;
            Mode m;
            m.nx = a;
            m.ny = b;
            m.kx = two_pi * a / lx;
            m.ky = two_pi * b / ly;
            const double n2 = (double)(a * a + b * b);
            m.av = 1.0 / std::sqrt(n2);
            m.aa = 1.0 / n2;
            const uint64_t id = (uint64_t)modes.size();
            for (int f = 0; f < 3; ++f) m.ph[f] = two_pi * uniform01(seed, (int)id, f + 1);
            modes.push_back(m);
        }
    // interior rms from the discrete Parseval sums: mean cos^2 = 1/2 per mode
    // and field; a face difference of Az multiplies a mode by 2 sin(k d / 2) / d
    double sv2 = 0.0, sb2 = 0.0;
    for (const Mode &m : modes) {
        sv2 += m.av * m.av;   // (vx^2 + vy^2) / 2 per mode
        const double dyf = 2.0 * std::sin(0.5 * m.ky * cfg->dy) / cfg->dy;
        const double dxf = 2.0 * std::sin(0.5 * m.kx * cfg->dx) / cfg->dx;
        sb2 += 0.5 * m.aa * m.aa * (dxf * dxf + dyf * dyf);
    }
    const double scale_v = sv2 > 0.0 ? rms_v / std::sqrt(sv2) : 0.0;
    const double scale_a = sb2 > 0.0 ? rms_b / std::sqrt(sb2) : 0.0;

    const int nx = cfg->nx, ny = cfg->ny;
    std::vector<double> ycen(ny), ycor(ny + 1);
    for (int j = 0; j < ny; ++j) ycen[j] = yc(cfg, j);
    for (int j = -1; j < ny; ++j) ycor[j + 1] = cfg->ymin + (j - MHD_GHOST + 1.0) * cfg->dy;   // corner (., j+1/2)
    const Separable tvx = row_tables(modes, kmax, 0, scale_v, ycen.data(), ny);
    const Separable tvy = row_tables(modes, kmax, 1, scale_v, ycen.data(), ny);
    const Separable taz = row_tables(modes, kmax, 2, scale_a, ycor.data(), ny + 1);
    const int nkx = kmax + 1;
    auto col_tables = [&](double x, double *c, double *sn) {
        for (int q = 0; q < nkx; ++q) {
            const double kx = two_pi * q / lx;
            c[q] = std::cos(kx * x);
            sn[q] = std::sin(kx * x);
        }
    };
    const double g = cfg->gamma;
    // columns are independent given Az on the corner column to their left
    auto work = [&](int i0, int i1) {
        std::vector<double> cc(nkx), sc(nkx), ca(nkx), sa(nkx);
        std::vector<double> a_left(ny + 1), a_right(ny + 1), bx_left(ny);
        auto corner_column = [&](int i, double *a) {   // Az at (i+1/2, j+1/2), j = -1..ny-1
            col_tables(cfg->xmin + (i - MHD_GHOST + 1.0) * cfg->dx, ca.data(), sa.data());
            for (int r = 0; r <= ny; ++r) a[r] = eval(taz, ny + 1, r, ca.data(), sa.data());
        };
        corner_column(i0 - 2, a_left.data());
        corner_column(i0 - 1, a_right.data());
        for (int r = 0; r < ny; ++r) bx_left[r] = (a_right[r + 1] - a_right[r]) / cfg->dy;
        a_left.swap(a_right);
        for (int i = i0; i < i1; ++i) {
            corner_column(i, a_right.data());
            col_tables(xc(cfg, i), cc.data(), sc.data());
            for (int j = 0; j < ny; ++j) {
                const double u = eval(tvx, ny, j, cc.data(), sc.data());
                const double v = eval(tvy, ny, j, cc.data(), sc.data());
                const double bx = (a_right[j + 1] - a_right[j]) / cfg->dy;       // x-face (i+1/2, j)
                const double by = -(a_right[j + 1] - a_left[j + 1]) / cfg->dx;   // y-face (i, j+1/2)
                const double bym = -(a_right[j] - a_left[j]) / cfg->dx;          // y-face (i, j-1/2)
                const double bxc = 0.5 * (bx + bx_left[j]);
                const double byc = 0.5 * (by + bym);
                put(aos, (long)j + (long)i * ny, 1.0, u, v, 0.0, bx, by, 0.0, bxc, byc, 1.0, g);
                bx_left[j] = bx;
            }
            a_left.swap(a_right);
        }
    };
    unsigned hw = std::thread::hardware_concurrency();
    int nthreads = (int)(hw == 0 ? 1 : (hw > 16 ? 16 : hw));
    if ((long long)nx * ny < (1LL << 18)) nthreads = 1;
    std::vector<std::thread> pool;
    const int chunk = (nx + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        const int a = t * chunk, b = (t + 1) * chunk < nx ? (t + 1) * chunk : nx;
        if (a >= b) break;
        if (t == nthreads - 1 || b == nx) work(a, b);
        else pool.emplace_back(work, a, b);
    }
    for (std::thread &th : pool) th.join();
    return MHD_OK;
}
