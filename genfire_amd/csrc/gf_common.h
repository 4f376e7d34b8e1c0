// gf_common.h -- shared definitions for the GENFIRE MI355X library (libgenfire.so).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/genfire.h"

namespace gf {

// Error carried from a failing HIP call or argument check up to the C-ABI, which
// stores the message in the context and returns the status code.
struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void raise(int code, const std::string& msg);

#define GF_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            ::gf::raise(GENFIRE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                             " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define GF_CHECK(cond, code, msg)                    \
    do {                                             \
        if (!(cond)) ::gf::raise((code), (msg));     \
    } while (0)

#define GF_LAUNCH_CHECK() GF_HIP(hipGetLastError())

// Device buffer owned by a job; freed on destruction.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        if (n == bytes && p) return;
        release();
        if (n) {
            hipError_t e = hipMalloc(&p, n);
            if (e != hipSuccess)
                raise(GENFIRE_ERR_NOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed: " +
                                             hipGetErrorString(e));
        }
        bytes = n;
    }
    // at least n bytes: keeps a larger buffer (no free and allocate when a
    // buffer alternates between two uses of different sizes)
    void ensure(size_t n) {
        if (p && bytes >= n) return;
        release();
        alloc(n);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Per-projection constants of the gridding index pass, computed on the host in
// FP32 with no contraction (see gf_host_math.cpp) and read by the kernels.
struct PlaneConst {
    float nv[3];    // normVECs(p,:) = rotMATs(:,3,p), single
    float nh[3];    // nv / norm(nv), single
    float nn;       // (nv1*nv1 + nv2*nv2) + nv3*nv3, single
    float lu[9];    // LU factors of rotMATs(:,:,p), column-major, single
    int piv[3];     // LAPACK-style row interchanges (0-based)
    int axis;       // dominant axis of nh (enumeration direction)
    double nhd[3];  // nh as double, for the candidate interval bound
};

// Half-grid geometry of fillInFourierGrid_DFT.m:122-136.
struct HalfGrid {
    int lo[3];
    int len[3];
    int64_t total() const { return (int64_t)len[0] * len[1] * len[2]; }
};

HalfGrid half_grid(int n1, int n2, int centro);

// Host math shared by the C-ABI (gf_host_math.cpp).
void rotation_f32(const double v1[3], const double v2[3], const double v3[3], double phi,
                  double theta, double psi, float Rf[9]);
void plane_setup(const float R[9], PlaneConst* pc);
void zyz_matrix(double phi, double theta, double psi, double R[9]);
void rotation_f64(const double v1[3], const double v2[3], const double v3[3], double phi, double theta, double psi,
                  double R[9]);

}  // namespace gf
