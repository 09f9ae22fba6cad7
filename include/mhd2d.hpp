// mhd2d.hpp -- C++ wrapper over the C ABI in mhd2d.h with the method names
// of the reference's single-GPU class (ResistiveMHD_singleGPU_thrust/
// lib_ResistiveMHD_2D_GPU_periodicXY/resistiveMHD_2D.hpp:9-52) and its
// boundary setters (ResistiveMHD_singleGPU_thrust/
// lib_ResistiveMHD_2D_GPU_symmetricXY/boundary.hpp:6-30), so that an example
// main written against the reference keeps its time loop.  Header-only; link
// with libmhd2d.so.  Errors throw mhd2d::Error (the C ABI itself never
// throws).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "mhd2d.h"

namespace mhd2d {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const mhd_handle *h)
{
    if (rc != MHD_OK) throw Error(rc, std::string("libmhd2d: ") + mhd_last_error(h));
}

class ResistiveMHD2D {
public:
    // ResistiveMHD2D() + initializeDeviceConstants(): the configuration
    // replaces the reference's link-time globals (const.hpp).
    explicit ResistiveMHD2D(const mhd_config &cfg) : cfg_(cfg)
    {
        check(mhd_create(&cfg_, &h_), nullptr);
    }
    ~ResistiveMHD2D() { mhd_destroy(h_); }
    ResistiveMHD2D(const ResistiveMHD2D &) = delete;
    ResistiveMHD2D &operator=(const ResistiveMHD2D &) = delete;

    // initializeU: the user's initial state (nx*ny cells, 8 doubles each,
    // index j + i*ny, ghosts included), then the configured boundary.
    void initializeU(const std::vector<double> &U)
    {
        setU(U);
        check(mhd_apply_boundary(h_, 0, cfg_.bc_x), h_);
        check(mhd_apply_boundary(h_, 1, cfg_.bc_y), h_);
    }
    void setU(const std::vector<double> &U)
    {
        if (U.size() != size()) throw Error(MHD_ERR_INVALID, "state size != nx*ny*8");
        check(mhd_set_state(h_, U.data()), h_);
    }
    std::vector<double> getU()
    {
        std::vector<double> U(size());
        check(mhd_get_state(h_, U.data()), h_);
        return U;
    }

    void oneStepRK2() { check(mhd_one_step_rk2(h_), h_); }
    void steps(int n) { check(mhd_step_rk2(h_, n), h_); }

    // calculateDt sets the member dt, as the reference sets its global dt.
    void calculateDt() { check(mhd_calculate_dt(h_, &dt), h_); }
    // dt of the last step (the reference's global dt after oneStepRK2)
    double lastDt()
    {
        double d;
        check(mhd_get_dt(h_, &d), h_);
        return d;
    }
    double totalTime()
    {
        double t;
        check(mhd_get_time(h_, &t), h_);
        return t;
    }

    bool checkCalculationIsCrashed()
    {
        int c = 0;
        check(mhd_check_crashed(h_, &c), h_);
        return c != 0;
    }

    void save(const std::string &directoryname, const std::string &filenameWithoutStep, int step)
    {
        check(mhd_save(h_, directoryname.c_str(), filenameWithoutStep.c_str(), step), h_);
    }
    void load(const std::string &path) { check(mhd_load(h_, path.c_str()), h_); }

    // The MPI library's per-rank save (ResistiveMHD_multiGPU_cuda/
    // ResistiveMHD2D_GPU/resistiveMHD2D.cu:457-486): interior cells only.
    void saveSubdomain(const std::string &directoryname, const std::string &filenameWithoutStep, int step, int rank)
    {
        check(mhd_save_subdomain(h_, directoryname.c_str(), filenameWithoutStep.c_str(), step, rank), h_);
    }
    void loadSubdomain(const std::string &path) { check(mhd_load_subdomain(h_, path.c_str()), h_); }

    // MHD_SCHEME_SINGLE (default) or MHD_SCHEME_MULTI_SHARE: the MPI
    // library's step on one rank, periodic halo filled on the device.
    void setScheme(int scheme) { check(mhd_set_scheme(h_, scheme), h_); }

    // Boundary::periodicBoundaryX2nd, ... applied to the current state
    void periodicBoundaryX2nd() { check(mhd_apply_boundary(h_, 0, MHD_BC_PERIODIC), h_); }
    void periodicBoundaryY2nd() { check(mhd_apply_boundary(h_, 1, MHD_BC_PERIODIC), h_); }
    void symmetricBoundaryX2nd() { check(mhd_apply_boundary(h_, 0, MHD_BC_SYMMETRIC), h_); }
    void symmetricBoundaryY2nd() { check(mhd_apply_boundary(h_, 1, MHD_BC_SYMMETRIC), h_); }

    // getEta plug-in replacement: per-cell resistivity (nx*ny), or uniform
    void setEtaField(const std::vector<double> &eta)
    {
        if (eta.size() != (size_t)cfg_.nx * cfg_.ny) throw Error(MHD_ERR_INVALID, "eta size != nx*ny");
        check(mhd_set_eta_field(h_, eta.data()), h_);
    }
    void setUniformEta(double eta) { check(mhd_set_eta_field(h_, nullptr), h_); check(mhd_set_eta(h_, eta), h_); cfg_.eta = eta; }
    // the MPI library's getEta plug-in under MHD_SCHEME_MULTI_SHARE: sampled
    // at x-faces, y-faces and cells (nx*ny each, periodic ghost entries)
    void setEtaStaggered(const std::vector<double> &xface, const std::vector<double> &yface,
                         const std::vector<double> &cell)
    {
        const size_t n = (size_t)cfg_.nx * cfg_.ny;
        if (xface.size() != n || yface.size() != n || cell.size() != n) throw Error(MHD_ERR_INVALID, "eta size != nx*ny");
        check(mhd_set_eta_staggered(h_, xface.data(), yface.data(), cell.data()), h_);
    }

    mhd_handle *handle() { return h_; }
    const mhd_config &config() const { return cfg_; }
    size_t size() const { return (size_t)cfg_.nx * cfg_.ny * 8; }

    double dt = 0.0;

private:
    mhd_config cfg_;
    mhd_handle *h_ = nullptr;
};

}  // namespace mhd2d
