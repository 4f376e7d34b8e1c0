// gf_project.hip -- real-space projector and R-factor sums
// (src/calculate3Dprojection_RealSpaceinterp.m:1-57, reconstruct.m:168-185,
// src/Tian_calc_Rfactor_realspace.m:1-24).
//
// Proj(i,j) = sum_k Vol(R [X_i; Y_j; Z_k]) with centred coordinates
// X = i - floor(d1/2) (..RealSpaceinterp.m:18-20) and trilinear interpolation with
// extrapolation value 0: a sample is zero unless every coordinate lies in the
// closed interval [0, d-1] of its axis.  One lane marches one ray (one output
// pixel) along z; a 16x16 pixel tile keeps neighbouring rays on neighbouring
// voxels so the eight-voxel gathers hit L1/L2.  The R-factor sums of each tile
// are written as per-tile partials and reduced in a fixed order.
#include "gf_grid.h"

namespace gf {

namespace {

constexpr int kTile = 16;

__device__ __forceinline__ bool lin_axis(float f, int n, int* i0, float* t)
{
    if (!(f >= 0.0f) || f > (float)(n - 1)) return false;
    int i = (int)floorf(f);
    if (i > n - 2) i = n - 2;
    *i0 = i;
    *t = f - (float)i;
    return true;
}

__global__ __launch_bounds__(kTile* kTile) void k_project(ProjArgs a)
{
    const int p = blockIdx.y;
    const int tiles1 = (a.m1 + kTile - 1) / kTile;
    const int ti = blockIdx.x % tiles1, tj = blockIdx.x / tiles1;
    const int i = ti * kTile + (threadIdx.x % kTile);
    const int j = tj * kTile + (threadIdx.x / kTile);
    const bool valid = i < a.m1 && j < a.m2;
    const float* R = a.R + 9 * p;
    const int d1 = a.d1, d2 = a.d2, d3 = a.d3;
    // pixel (i,j) of the measured window sits at (i + s1, j + s2) of the d1 x d2
    // back projection (My_stripzero of the full projection, Tian...:16)
    const int s1 = d1 / 2 - a.m1 / 2, s2 = d2 / 2 - a.m2 / 2;
    const float X = (float)(i + s1 - d1 / 2), Y = (float)(j + s2 - d2 / 2);
    const float c1 = (float)(d1 / 2), c2 = (float)(d2 / 2), c3 = (float)(d3 / 2);
    const float bx = R[0] * X + R[3] * Y + c1;
    const float by = R[1] * X + R[4] * Y + c2;
    const float bz = R[2] * X + R[5] * Y + c3;
    float acc = 0.0f;
    if (valid) {
        for (int k = 0; k < d3; ++k) {
            const float Z = (float)(k - d3 / 2);
            const float fx = bx + R[6] * Z, fy = by + R[7] * Z, fz = bz + R[8] * Z;
            int a0, b0, e0;
            float ta, tb, te;
            if (!lin_axis(fx, d1, &a0, &ta) || !lin_axis(fy, d2, &b0, &tb) || !lin_axis(fz, d3, &e0, &te))
                continue;
            const float* v = a.vol + (int64_t)a0 + (int64_t)d1 * (b0 + (int64_t)d2 * e0);
            const int64_t sy = d1, sz = (int64_t)d1 * d2;
            const float v000 = v[0], v100 = v[1], v010 = v[sy], v110 = v[sy + 1];
            const float v001 = v[sz], v101 = v[sz + 1], v011 = v[sz + sy], v111 = v[sz + sy + 1];
            const float x00 = v000 + ta * (v100 - v000), x10 = v010 + ta * (v110 - v010);
            const float x01 = v001 + ta * (v101 - v001), x11 = v011 + ta * (v111 - v011);
            const float y0 = x00 + tb * (x10 - x00), y1 = x01 + tb * (x11 - x01);
            acc += y0 + te * (y1 - y0);
        }
    }
    double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
    if (valid) {
        const int64_t o = (int64_t)i + (int64_t)a.m1 * (j + (int64_t)a.m2 * p);
        if (a.out) a.out[o] = acc;
        if (a.meas) {
            const double m = (double)a.meas[o];
            const double b = (double)acc;
            const double dlt = a.post_hoc ? fabs(fabs(b) - fabs(m)) : fabs(b - m);
            q0 = dlt;
            q1 = (b - m) * (b - m);
            q2 = fabs(m);
            q3 = m * m;
        }
    }
    if (!a.sums) return;
    __shared__ double red[4][kTile * kTile / 64];
    for (int off = 32; off > 0; off >>= 1) {
        q0 += __shfl_down(q0, off, 64);
        q1 += __shfl_down(q1, off, 64);
        q2 += __shfl_down(q2, off, 64);
        q3 += __shfl_down(q3, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = q0; red[1][w] = q1; red[2][w] = q2; red[3][w] = q3; }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int k = 0; k < kTile * kTile / 64; ++k) t += red[threadIdx.x][k];
        a.sums[((int64_t)p * 4 + threadIdx.x) * gridDim.x + blockIdx.x] = t;
    }
}

__global__ void k_project_reduce(const double* sums, int P, int nblk, double* out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * 4) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += sums[(int64_t)t * nblk + b];
    out[t] = s;
}

}  // namespace

int project_blocks(int m1, int m2)
{
    return ((m1 + kTile - 1) / kTile) * ((m2 + kTile - 1) / kTile);
}

void project(const ProjArgs& a, hipStream_t st)
{
    GF_CHECK(a.d1 >= 2 && a.d2 >= 2 && a.d3 >= 2, GENFIRE_ERR_ARG, "projector needs every volume dimension >= 2");
    GF_CHECK(a.m1 <= a.d1 && a.m2 <= a.d2, GENFIRE_ERR_ARG, "measured projection larger than the volume");
    dim3 grid(project_blocks(a.m1, a.m2), a.P);
    k_project<<<grid, kTile * kTile, 0, st>>>(a);
    GF_LAUNCH_CHECK();
}

void project_reduce(const double* d_sums, int P, int nblk, double* d_out4, hipStream_t st)
{
    const int n = P * 4;
    k_project_reduce<<<(n + 255) / 256, 256, 0, st>>>(d_sums, P, nblk, d_out4);
    GF_LAUNCH_CHECK();
}

}  // namespace gf
