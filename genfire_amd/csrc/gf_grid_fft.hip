// gf_grid_fft.hip -- FFT gridding (griddingMethod = 1): fillInFourierGrid.m:27-168
// and src/hermitianSymmetrize.m:22-65.
//
// 1. k_fg_load: every projection zero-padded to d x d (My_paddzero) and put in
//    FFT order (ifftshift), as complex; then two C2C passes of the batched 2D FFT
//    (my_fft = fftshift(fftn(ifftshift(x))), fillInFourierGrid.m:60; the
//    fftshift is folded into the frequency -> FFT index mapping below).
// 2. k_fg_scatter, one launch per projection, voxel-centric: the reference walks
//    the rotated lattice points and appends (voxel, value, distance) to a list
//    (:126-141) that accumarray then sums (:157-158).  Here each thread owns one
//    column of voxels through the plane's slab and finds, for each voxel, the
//    lattice points whose rotated coordinates match it under the reference's own
//    test (round(m) + shift == voxel, |shift| <= shiftMax, distance <= cutoff),
//    re-evaluating that test with the same single-precision operations.  Every
//    voxel is owned by one thread per launch and the launches run in projection
//    order, so the FP64 sums need no atomics and are the same on every run.
// 3. k_fg_finalize: measuredK = sum(w v)/sum(w) rounded to single (:159-160),
//    hermitianSymmetrize's average of each voxel with its conjugate mate (the
//    padded odd-size flip pairs k with -k; k = -d/2 of an even edge meets the
//    zero padding), the single cast of runGridding, written in FFT order.
#include "gf_grid.h"

namespace gf {
namespace {

__device__ __forceinline__ int signed_freq(int s, int n) { return s < (n + 1) / 2 ? s : s - n; }

// Projection p's slice: real projection (Dim1 x Dim2, column-major) padded into
// d x d around the centre pixel floor(d/2) and ifftshifted, imaginary part zero.
__global__ void k_fg_load(const float* __restrict__ proj, int dim1, int dim2, int P, int d,
                          float2* __restrict__ F)
{
    const int64_t plane = (int64_t)d * d, total = plane * P;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int s1 = (int)(t % d), s2 = (int)((t / d) % d);
        const int64_t p = t / plane;
        const int a = signed_freq(s1, d) + dim1 / 2;
        const int b = signed_freq(s2, d) + dim2 / 2;
        float v = 0.0f;
        if (a >= 0 && a < dim1 && b >= 0 && b < dim2) v = proj[(int64_t)a + (int64_t)dim1 * (b + (int64_t)dim2 * p)];
        F[t] = make_float2(v, 0.0f);
    }
}

__global__ __launch_bounds__(256) void k_fg_scatter(FftPlane pl, int d, int shift_max, float rho, double cutoff,
                                                    const float2* __restrict__ F, double2* __restrict__ acc,
                                                    double* __restrict__ sw, uint32_t* __restrict__ hits)
{
    const int c0 = d / 2;                  // centre pixel, 0-based (nc - 1)
    const int kmin = -c0, kmax = d - 1 - c0;
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= (int64_t)d * d) return;
    const int A = pl.axis, B = A == 0 ? 1 : 0, C = A == 2 ? 1 : 2;
    int v[3];
    v[B] = (int)(col % d) + kmin;
    v[C] = (int)(col / d) + kmin;
    // slab |n . v| <= rho along the dominant axis of the normal
    const float na = pl.nrm[A];
    const float t = pl.nrm[B] * (float)v[B] + pl.nrm[C] * (float)v[C];
    float e0 = (-rho - t) / na, e1 = (rho - t) / na;
    if (e0 > e1) { const float q = e0; e0 = e1; e1 = q; }
    const int lo = max(kmin, (int)ceilf(e0)), hi = min(kmax, (int)floorf(e1));
    const float cut_f = rho;
    for (int va = lo; va <= hi; ++va) {
        v[A] = va;
        const float vx = (float)v[0], vy = (float)v[1], vz = (float)v[2];
        // in-plane coordinates of the voxel (candidate window only)
        const float q1 = pl.rot[0] * vx + pl.rot[1] * vy + pl.rot[2] * vz;
        const float q2 = pl.rot[3] * vx + pl.rot[4] * vy + pl.rot[5] * vz;
        const int i_lo = max(kmin, (int)ceilf(q1 - cut_f)), i_hi = min(kmax, (int)floorf(q1 + cut_f));
        const int j_lo = max(kmin, (int)ceilf(q2 - cut_f)), j_hi = min(kmax, (int)floorf(q2 + cut_f));
        double ax = 0.0, ay = 0.0, aw = 0.0;
        uint32_t n = 0;
        // lattice points in list order: ky (columns) outer, kx (rows) inner
        for (int ky = j_lo; ky <= j_hi; ++ky) {
            for (int kx = i_lo; kx <= i_hi; ++kx) {
                const float fx = (float)kx, fy = (float)ky;
                // rotkCoords = R' * [kx; ky; 0] in single (:87)
                float m[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) m[i] = __fadd_rn(__fmul_rn(pl.rot[i], fx), __fmul_rn(pl.rot[i + 3], fy));
                // tmp = round(measured) + shift must be this voxel, |shift| <= shiftMax (:126-131)
                bool ok = true;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const int s = v[i] - (int)roundf(m[i]);
                    ok = ok && s >= -shift_max && s <= shift_max;
                }
                if (!ok) continue;
                const float dx = __fsub_rn(m[0], vx), dy = __fsub_rn(m[1], vy), dz = __fsub_rn(m[2], vz);
                const float dist = __fsqrt_rn(
                    __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz)));   // :132
                if (!((double)dist <= cutoff)) continue;                                             // :136
                const float2 f = F[(int64_t)(kx < 0 ? kx + d : kx) + (int64_t)d * (ky < 0 ? ky + d : ky)];
                if (f.x == -999.0f && f.y == 0.0f) continue;                                          // :108-113
                const float w = __fdiv_rn(1.0f, __fadd_rn(dist, 1e-5f));                              // :153-155
                ax += (double)__fmul_rn(f.x, w);
                ay += (double)__fmul_rn(f.y, w);
                aw += (double)w;
                ++n;
            }
        }
        if (n) {
            const int64_t vi = (int64_t)(v[0] - kmin) + (int64_t)d * ((v[1] - kmin) + (int64_t)d * (v[2] - kmin));
            double2 a = acc[vi];
            a.x += ax;
            a.y += ay;
            acc[vi] = a;
            sw[vi] += aw;
            hits[vi] += n;
        }
    }
}

// measuredK(v) = sum(w v) / sum(w) where sum(w) > 0, as single (:157-160)
__device__ __forceinline__ float2 fg_value(const double2* acc, const double* sw, int64_t i)
{
    const double w = sw[i];
    if (!(w > 0.0)) return make_float2(0.0f, 0.0f);
    const double2 a = acc[i];
    return make_float2((float)(a.x / w), (float)(a.y / w));
}

__global__ void k_fg_finalize(int d, const double2* __restrict__ acc, const double* __restrict__ sw,
                              float2* __restrict__ M)
{
    const int64_t total = (int64_t)d * d * d;
    const int c0 = d / 2;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int k1 = signed_freq((int)(s % d), d), k2 = signed_freq((int)((s / d) % d), d),
                  k3 = signed_freq((int)(s / ((int64_t)d * d)), d);
        const int64_t vi = (int64_t)(k1 + c0) + (int64_t)d * ((k2 + c0) + (int64_t)d * (k3 + c0));
        const float2 a = fg_value(acc, sw, vi);
        float2 b = make_float2(0.0f, 0.0f);
        const int m1 = c0 - k1, m2 = c0 - k2, m3 = c0 - k3;   // centred index of -k
        if (m1 < d && m2 < d && m3 < d) b = fg_value(acc, sw, (int64_t)m1 + (int64_t)d * (m2 + (int64_t)d * m3));
        const int cnt = (a.x != 0.0f || a.y != 0.0f) + (b.x != 0.0f || b.y != 0.0f);   // :43-45
        double re = (double)a.x + (double)b.x, im = (double)a.y - (double)b.y;           // :41
        if (cnt) {                                                                       // :47
            re /= (double)cnt;
            im /= (double)cnt;
        }
        M[s] = make_float2((float)re, (float)im);
    }
}

int fg_blocks(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    return (int)(b < 65536 * 4 ? (b < 1 ? 1 : b) : 65536 * 4);
}

}  // namespace

void fftgrid_load(const float* d_proj, int dim1, int dim2, int P, int d, float2* d_F, hipStream_t st)
{
    const int64_t total = (int64_t)d * d * P;
    k_fg_load<<<fg_blocks(total, 256), 256, 0, st>>>(d_proj, dim1, dim2, P, d, d_F);
    GF_LAUNCH_CHECK();
}

void fftgrid_scatter(const FftPlane& pl, int d, int shift_max, float rho, double cutoff, const float2* d_Fp,
                     double2* d_acc, double* d_sw, uint32_t* d_hits, hipStream_t st)
{
    const int64_t cols = (int64_t)d * d;
    k_fg_scatter<<<(unsigned)((cols + 255) / 256), 256, 0, st>>>(pl, d, shift_max, rho, cutoff, d_Fp, d_acc, d_sw,
                                                                 d_hits);
    GF_LAUNCH_CHECK();
}

void fftgrid_finalize(int d, const double2* d_acc, const double* d_sw, float2* d_M, hipStream_t st)
{
    const int64_t total = (int64_t)d * d * d;
    k_fg_finalize<<<fg_blocks(total, 256), 256, 0, st>>>(d, d_acc, d_sw, d_M);
    GF_LAUNCH_CHECK();
}

}  // namespace gf
