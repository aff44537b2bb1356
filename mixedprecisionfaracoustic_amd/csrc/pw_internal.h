// pw_internal.h -- shared declarations of the porowave library (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/porowave.h"

namespace pw {

// Thread-local error text behind pw_last_error().
void set_error(const std::string& msg);
pw_status fail(pw_status st, const std::string& msg);

#define PW_HIP_TRY(expr)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::pw::fail(e_ == hipErrorOutOfMemory ? PW_ERR_OUT_OF_MEMORY : PW_ERR_HIP, \
                              std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

// CPML memory-variable slots per damped (cell, axis): 4 terms in the velocity
// kernel and 4 in the stress kernel (slots 0-7), one variable each; comp16
// keeps the reference Half program's pair per term (Half/main.cu:201-240): psi
// in slots 0-7 and the previous derivative in slots 8-15.
constexpr int kMemPerAxis = 16;

// Grid geometry shared by every kernel.  Arrays are padded SoA: row pitch is
// nx rounded up to 64 elements, plane = pitch * ny.
struct Geom {
    int nx, ny, nz;
    int pitch;            // elements per row
    long long plane;      // elements per z-plane
    long long cells;      // pitch * ny * nz (allocation size per array)
    int lo[3], hi[3];     // CPML slab widths per axis (cells with a nonzero profile)
    long long slab_cells[3];
};

// 1-D CPML tables per axis, (c, e) for the full-point and half-point profiles:
// c = dt*d/2, e = exp(-d dt), as used by the single-variable recursion
//   psi = phi - c*der;  der += psi;  phi = e*(psi - c*der).
struct CpmlTables {
    std::vector<double> c_full[3], e_full[3], c_half[3], e_half[3];
};

// Host-side precompute helpers (pw_medium.cpp).
void cpml_profile_reference_f32(int n, int np, float h, float dt, float rc, float v0, float* d, float* d2,
                                float* e, float* e2);
void cpml_profile_f64(int n, int np, double h, double dt, double rc, double v0, double* d, double* d2,
                      double* e, double* e2);
float ricker_f32(int it, float dt, float f0);
double ricker_f64(int it, double dt, double f0);

}  // namespace pw
