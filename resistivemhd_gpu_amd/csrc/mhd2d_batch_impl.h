// mhd2d_batch_impl.h -- internal state of a batch (mhd_batch of
// include/mhd2d.h), shared by the batch and adjoint sources.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

#include "../../include/mhd2d.h"
#include "mhd2d_kernels.h"

struct mhd_batch {
    int nruns = 0;
    std::vector<mhd_config> cfg;
    std::vector<mhd::RunDesc> desc;          // host copy of the descriptors
    std::vector<double *> eta;          // per-run resistivity planes (nullptr: uniform)
    mhd::RunDesc *d_desc = nullptr;
    int *d_item_run = nullptr;
    int item_cap = 0;
    int *d_tail_run = nullptr;   // packed last strips: item -> run
    int tail_cap = 0, tail_items = 0;
    int strip_packing = MHD_STRIP_PACK_AUTO;
    double *pool = nullptr;             // U, U* and the side row of every run
    mhd::Ctrl *ctrl = nullptr;               // one control block per run
    double *aos = nullptr;              // staging for host transfers (largest run)
    hipStream_t stream = nullptr;
    int device = 0, num_cu = 0;
    bool exact = true;
    int seg_request = 0, seg = 0;
    int total_items = 0, xcds = 1, items_per_xcd = 0;
    // stage launch shape: waves_per_cu resident wavefronts per CU looping
    // over the work list, or 0 for one wavefront per item
    int waves_per_cu = 0, blocks_per_xcd = 0;
    int max_nx = 0, max_ny = 0;
    int cur = 0;
    bool dt_valid = false;
    bool use_graph = true;
    hipGraphExec_t graph[2] = {nullptr, nullptr};   // captured steps, by the control slot they start from
    std::string err;
};

// Per-cell dt of every run's current state into its control slot (done
// lazily before the first step after a state change).
int mhd_batch_refresh_dt(mhd_batch *b);
