// pw_solver.h -- the solver handle shared by the C ABI (pw_api.cpp) and the
// shot server (pw_server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pw_device.h"

struct pw_solver {
    pw_config cfg;
    int prec = PW_FP32;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    pw::DevBuffers b{};
    void* pool = nullptr;
    size_t pool_bytes = 0;
    size_t state_bytes_total = 0;  // bytes of the state + memory-variable region (zeroed on reset)
    // medium
    bool has_medium = false;
    int em = -23, epml = 0;
    float inv_scale_pml = 1.0f;
    double vp_max = 0.0, courant = 0.0;
    bool layered = false;
    pw_layered spec{};
    // comp16 / bf16 dynamic scale (Half/main.cu:800-807, 1326-1364)
    float dyn_scale = 1.0f, stress_base = 1.0f, vel_base = 1.0f, source_scale = 1.0f;
    // source
    bool has_source = false;
    unsigned src_off = 0;
    double f0 = 0.0, w[4] = {1, 1, 1, 0};
    // receivers
    int nrec = 0;
    bool ref_receivers = false;  // set by pw_set_receivers_reference: the five reference channels
    int* d_rfield = nullptr;
    unsigned* d_roff = nullptr;
    float* d_seis = nullptr;
    // snapshots
    int nsnap = 0, snap_y = 0;
    std::vector<int> snap_step, snap_field;
    float* d_snaps = nullptr;
    // full-volume snapshots (pw_set_volume_snapshots): decoded on the compute
    // stream into one of two device staging volumes, then copied to pinned host
    // memory on copy_stream while the steps go on
    int nvol = 0;
    std::vector<int> vol_step, vol_field;
    std::vector<float*> vol_host;        // hipHostMalloc'ed, one volume each
    std::vector<hipEvent_t> vol_done;    // recorded on copy_stream after each copy
    std::vector<char> vol_taken;         // the snapshot's step has run
    float* d_vstage[2] = {nullptr, nullptr};
    hipEvent_t ev_decoded[2] = {nullptr, nullptr}, ev_copied[2] = {nullptr, nullptr};
    int vol_next = 0;
    hipStream_t copy_stream = nullptr;
    int step = 0;
    // per-kernel timing (pw_set_profiling)
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;     // created lazily, reused
    std::vector<hipEvent_t> ev_pending;  // (start, stop) pairs around step-kernel launches
    std::vector<char> ev_kind;           // per pair: 0 velocity, 1 stress
    long long prof_steps = 0;
};

namespace pw {

// Creates a solver (pw_create).
pw_status solver_create(const pw_config* cfg, pw_solver** out);
// Device bytes pw_create allocates for this configuration (state, memory
// variables, coefficients, tables), without the receiver and snapshot buffers.
size_t solver_pool_bytes(const pw_config& c);
// Installs a medium from device-resident FP64 base fields: field k of cell
// (x, y, z) of the interior at d_base[k] + x + y*row_stride + z*plane_stride.
// d_res: kReduceSlots doubles of device memory for the medium's maxima, or null
// to allocate them for the call.
pw_status install_medium(pw_solver* s, const double* const* d_base, int nxi, int nyi, int nzi, long long row_stride,
                         long long plane_stride, double* d_res = nullptr);

// What one step does besides the step kernels (Half/main.cu:1326-1377 for the
// scaled forms): the dynamic rescale ratio applied first (0 = none), the source
// amplitude, and the factors that turn stored stress / velocity into physical
// units for the receivers.  *dyn carries the dynamic scale from step to step.
struct StepPlan {
    float ratio;
    double amp;
    float s_stress, s_vel;
};
StepPlan plan_step(const pw_solver* s, int it, float* dyn);

}  // namespace pw
