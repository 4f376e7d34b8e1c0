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


__device__ __forceinline__ bool lin_axis(float f, int n, int* i0, float* t)
{
    if (!(f >= 0.0f) || f > (float)(n - 1)) return false;
    int i = (int)floorf(f);
    if (i > n - 2) i = n - 2;
    *i0 = i;
    *t = f - (float)i;
    return true;
}

// Unaligned pair load: two adjacent floats of the volume in one 8-byte access.
typedef float pair_t __attribute__((ext_vector_type(2), aligned(4)));
// Two pairs of the z-pair copies in one 16-byte access (8-byte aligned).
typedef float quad_t __attribute__((ext_vector_type(4), aligned(8)));

// With more than 8 projections the workgroups of one projection run on one XCD:
// the hardware deals workgroups to the 8 XCDs round-robin, so physical block b
// goes to XCD b % 8.  XCD x takes the contiguous block of per = ceil(P / 8)
// projections starting at x * per, in groups of G neighbouring projections whose
// tiles are interleaved (slot = (g * tiles + tile) * G + m for member m of the
// XCD's group g).  Neighbouring projections of a tilt
// series send nearly the same rays through the model, so the G workgroups of
// one tile read the same voxels at about the same time and share the XCD's L2
// fills: each projection reads the whole model (twice the voxels in the z-pair
// copies), which would otherwise stream from HBM or the MALL once per projection.
// With at most 8 projections (per = 1) every XCD instead takes a contiguous
// eighth of each projection's tiles, so a single projection uses the whole chip.
constexpr int kXcd = 8;
constexpr int kProjGroup = 6;   // G, tilt256's 91 projections: 12 per XCD, two groups
constexpr int kLanes = 64;   // pixels of one wave along the fast pixel axis
constexpr int kRows = 4;     // waves per workgroup along the other pixel axis
constexpr float kRowAlign = 0.9f;   // lane rows when a pairing's |R| reaches this, else patches

// A wave's 64 lanes are 64 neighbouring pixels along the pixel axis (i or j,
// R[11]) whose direction in the volume is closest to the fast axis of one of the
// model's forms (x- or y-fastest, R[9]): the largest of |R00|, |R11|, |R10|,
// |R01|.  For tilts about y the pixel-y direction
// is the volume's y axis, so each gather of a wave reads a few contiguous runs
// instead of one cache line per lane.
//
// Rotations that map the volume's y axis onto itself (R[10] set: tilts about y)
// take a second form when d2 is a multiple of 4: every ray of a pixel column i
// then samples the same (x, z) cell with the same weights at each step, its y
// being exactly j + s2, so one lane marches the four rays of an aligned group
// of four y and reads each corner of the cell for all four with one 16-byte
// load -- a quarter of the gather instructions of the pair form.  Each ray's
// interpolation is the pair form's: with ty = 0 its y-lerp returns the lower
// value, and the ray at y = d2-1 (cell d2-2, ty = 1) takes the two values from
// the same load.  A workgroup covers 16 pixel columns x 64 y; the tile slots
// past the ones it needs only write zero partials.
__device__ __forceinline__ void project_ytilt(const ProjArgs& a, const float* R, int p, int tile,
                                              double& q0, double& q1, double& q2, double& q3)
{
    const int d1 = a.d1, d2 = a.d2, d3 = a.d3;
    const int s1 = d1 / 2 - a.m1 / 2, s2 = d2 / 2 - a.m2 / 2;
    const int g0 = s2 / 4, ng = (s2 + a.m2 + 3) / 4 - g0;   // groups of four y covering the window
    const int yb = (ng + 15) / 16;
    const int w = (int)threadIdx.x / kLanes, l = (int)threadIdx.x % kLanes;
    const int i = (tile / yb) * 16 + 4 * w + l / 16;
    const int y0 = 4 * (g0 + (tile % yb) * 16 + l % 16);
    if (i >= a.m1 || y0 >= s2 + a.m2) return;
    const float X = (float)(i + s1 - d1 / 2), Y = (float)(y0 - d2 / 2);
    const float c1 = (float)(d1 / 2), c3 = (float)(d3 / 2);
    const float bx = R[0] * X + R[3] * Y + c1;      // R[3] = R[5] = 0: the same for the four rays
    const float bz = R[2] * X + R[5] * Y + c3;
    const float rx = R[6], rz = R[8];
    float klo = 0.0f, khi = (float)(d3 - 1);
    const float bb[2] = {bx, bz}, rr[2] = {rx, rz};
    const float hi[2] = {(float)(d1 - 1), (float)(d3 - 1)};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        if (rr[q] != 0.0f) {
            const float t0 = (0.0f - bb[q]) / rr[q] + c3, t1 = (hi[q] - bb[q]) / rr[q] + c3;
            klo = fmaxf(klo, fminf(t0, t1) - 1.0f);
            khi = fminf(khi, fmaxf(t0, t1) + 1.0f);
        } else if (!(bb[q] >= 0.0f && bb[q] <= hi[q])) {
            khi = -1.0f;
        }
    }
    const int k0 = (int)fminf((float)d3, fmaxf(0.0f, ceilf(klo)));
    const int k1 = (int)fmaxf(-1.0f, fminf((float)(d3 - 1), floorf(khi)));
    const bool top = y0 + 3 == d2 - 1;      // the fourth ray sits in cell d2-2 with ty = 1
    const float* vol = a.vol_y + y0;        // index y + d2*(x + d1*z)
    const int sx = d2, sz = d1 * d2;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = k0; k <= k1; ++k) {
        const float Z = (float)(k - d3 / 2);
        const float fx = bx + rx * Z, fz = bz + rz * Z;
        int a0, e0;
        float ta, te;
        if (!lin_axis(fx, d1, &a0, &ta) || !lin_axis(fz, d3, &e0, &te)) continue;
        const float* v = vol + d2 * (a0 + d1 * e0);
        const float4 A = *reinterpret_cast<const float4*>(v);             // (a0,   y0..y0+3, e0)
        const float4 B = *reinterpret_cast<const float4*>(v + sx);        // (a0+1, ., e0)
        const float4 C = *reinterpret_cast<const float4*>(v + sz);        // (a0,   ., e0+1)
        const float4 D = *reinterpret_cast<const float4*>(v + sz + sx);   // (a0+1, ., e0+1)
        const float a4[4] = {A.x, A.y, A.z, A.w}, b4[4] = {B.x, B.y, B.z, B.w};
        const float c4[4] = {C.x, C.y, C.z, C.w}, d4[4] = {D.x, D.y, D.z, D.w};
        float L[4], U[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            L[r] = a4[r] + ta * (b4[r] - a4[r]);
            U[r] = c4[r] + ta * (d4[r] - c4[r]);
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) acc[r] += L[r] + te * (U[r] - L[r]);
        const float ty = 1.0f;
        const float l3 = top ? L[2] + ty * (L[3] - L[2]) : L[3];
        const float u3 = top ? U[2] + ty * (U[3] - U[2]) : U[3];
        acc[3] += l3 + te * (u3 - l3);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = y0 + r - s2;
        if (j < 0 || j >= a.m2) continue;
        const int64_t o = (int64_t)i + (int64_t)a.m1 * (j + (int64_t)a.m2 * p);
        if (a.out) a.out[o] = acc[r];
        if (a.meas) {
            const double m = (double)a.meas[o];
            const double b = (double)acc[r];
            q0 += a.post_hoc ? fabs(fabs(b) - fabs(m)) : fabs(b - m);
            q1 += (b - m) * (b - m);
            q2 += fabs(m);
            q3 += m * m;
        }
    }
}

__device__ __forceinline__ int ytilt_tiles(const ProjArgs& a)
{
    const int s2 = a.d2 / 2 - a.m2 / 2, g0 = s2 / 4, ng = (s2 + a.m2 + 3) / 4 - g0;
    return ((a.m1 + 15) / 16) * ((ng + 15) / 16);
}

__device__ __forceinline__ void tile_sums(const ProjArgs& a, int p, int tile, int tiles, double q0, double q1,
                                          double q2, double q3);

__global__ __launch_bounds__(kLanes* kRows) void k_project(ProjArgs a, int tiles, int per, int group)
{
    const int xcd = blockIdx.x % kXcd, slot = blockIdx.x / kXcd;
    int p, tile;
    if (per == 1) {   // at most 8 projections: XCD x takes the x-th eighth of every projection's tiles
        const int tc = (tiles + kXcd - 1) / kXcd;
        p = slot / tc;
        tile = xcd * tc + slot % tc;
        if (tile >= tiles || p >= a.P) return;
    } else {
        const int k = (slot / (group * tiles)) * group + slot % group;   // projection within the XCD's block
        p = xcd * per + k;
        tile = (slot / group) % tiles;
        if (k >= per || p >= a.P) return;
    }
    const float* R = a.R + kRStride * p;
    if (R[10] != 0.0f && a.d2 % 4 == 0 && ytilt_tiles(a) <= tiles) {
        double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
        if (tile < ytilt_tiles(a)) project_ytilt(a, R, p, tile, q0, q1, q2, q3);
        tile_sums(a, p, tile, tiles, q0, q1, q2, q3);
        return;
    }
    const bool yfast = R[9] != 0.0f;     // the model's y-fastest forms
    int i, j;
    if (R[11] == 2.0f) {                  // 8 x 8 pixel patch per wave, 16 x 16 per workgroup
        const int tp1 = (a.m1 + 15) / 16, w = (int)threadIdx.x / kLanes, l = (int)threadIdx.x % kLanes;
        i = (tile % tp1) * 16 + (w & 1) * 8 + (l & 7);
        j = (tile / tp1) * 16 + (w >> 1) * 8 + (l >> 3);
    } else if (R[11] != 0.0f) {           // lanes along pixel j
        const int tiles2 = (a.m2 + kLanes - 1) / kLanes;
        i = (tile / tiles2) * kRows + (int)(threadIdx.x / kLanes);
        j = (tile % tiles2) * kLanes + (int)(threadIdx.x % kLanes);
    } else {
        const int tiles1 = (a.m1 + kLanes - 1) / kLanes;
        i = (tile % tiles1) * kLanes + (int)(threadIdx.x % kLanes);
        j = (tile / tiles1) * kRows + (int)(threadIdx.x / kLanes);
    }
    const bool valid = i < a.m1 && j < a.m2;
    const int d1 = a.d1, d2 = a.d2, d3 = a.d3;
    // pixel (i,j) of the measured window sits at (i + s1, j + s2) of the d1 x d2
    // back projection (My_stripzero of the full projection, Tian...:16)
    const int s1 = d1 / 2 - a.m1 / 2, s2 = d2 / 2 - a.m2 / 2;
    const float X = (float)(i + s1 - d1 / 2), Y = (float)(j + s2 - d2 / 2);
    const float c1 = (float)(d1 / 2), c2 = (float)(d2 / 2), c3 = (float)(d3 / 2);
    const float bx = R[0] * X + R[3] * Y + c1;
    const float by = R[1] * X + R[4] * Y + c2;
    const float bz = R[2] * X + R[5] * Y + c3;
    const float rx = R[6], ry = R[7], rz = R[8];
    float acc = 0.0f;
    if (valid) {
        // steps whose sample can lie inside the box, from the three slab
        // intervals widened by one step; the exact test stays per step
        float klo = 0.0f, khi = (float)(d3 - 1);
        const float bb[3] = {bx, by, bz}, rr[3] = {rx, ry, rz};
        const float hi[3] = {(float)(d1 - 1), (float)(d2 - 1), (float)(d3 - 1)};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (rr[q] != 0.0f) {
                const float t0 = (0.0f - bb[q]) / rr[q] + c3, t1 = (hi[q] - bb[q]) / rr[q] + c3;
                klo = fmaxf(klo, fminf(t0, t1) - 1.0f);
                khi = fminf(khi, fmaxf(t0, t1) + 1.0f);
            } else if (!(bb[q] >= 0.0f && bb[q] <= hi[q])) {
                khi = -1.0f;
            }
        }
        const int k0 = (int)fminf((float)d3, fmaxf(0.0f, ceilf(klo)));
        const int k1 = (int)fmaxf(-1.0f, fminf((float)(d3 - 1), floorf(khi)));
        if (yfast && a.vol_z) {
            // z-pair copies: the cell's four voxels at x = a0 are one 16-byte
            // load, (b0, e0) (b0, e0+1) (b0+1, e0) (b0+1, e0+1), from the copy
            // whose pairs start at e0's parity; the same at x = a0 + 1
            const int sx = 2 * d2;
            for (int k = k0; k <= k1; ++k) {
                const float Z = (float)(k - d3 / 2);
                const float fx = bx + rx * Z, fy = by + ry * Z, fz = bz + rz * Z;
                int a0, b0, e0;
                float ta, tb, te;
                if (!lin_axis(fx, d1, &a0, &ta) || !lin_axis(fy, d2, &b0, &tb) || !lin_axis(fz, d3, &e0, &te))
                    continue;
                const int odd = e0 & 1;
                const float* v = a.vol_z + (odd ? a.zodd : 0) + 2 * (b0 + d2 * (a0 + d1 * (e0 >> 1)));
                const quad_t q0 = *reinterpret_cast<const quad_t*>(v);        // x = a0
                const quad_t q1 = *reinterpret_cast<const quad_t*>(v + sx);   // x = a0 + 1
                const float x00 = q0.x + ta * (q1.x - q0.x), x10 = q0.z + ta * (q1.z - q0.z);
                const float x01 = q0.y + ta * (q1.y - q0.y), x11 = q0.w + ta * (q1.w - q0.w);
                const float y0 = x00 + tb * (x10 - x00), y1 = x01 + tb * (x11 - x01);
                acc += y0 + te * (y1 - y0);
            }
        } else if (yfast) {
            const float* vol = a.vol_y;           // index y + d2*(x + d1*z)
            const int sx = d2, sz = d1 * d2;
            for (int k = k0; k <= k1; ++k) {
                const float Z = (float)(k - d3 / 2);
                const float fx = bx + rx * Z, fy = by + ry * Z, fz = bz + rz * Z;
                int a0, b0, e0;
                float ta, tb, te;
                if (!lin_axis(fx, d1, &a0, &ta) || !lin_axis(fy, d2, &b0, &tb) || !lin_axis(fz, d3, &e0, &te))
                    continue;
                const float* v = vol + (b0 + d2 * (a0 + d1 * e0));
                const pair_t q00 = *reinterpret_cast<const pair_t*>(v);            // (a0,   b0..b0+1, e0)
                const pair_t q10 = *reinterpret_cast<const pair_t*>(v + sx);       // (a0+1, ., e0)
                const pair_t q01 = *reinterpret_cast<const pair_t*>(v + sz);       // (a0,   ., e0+1)
                const pair_t q11 = *reinterpret_cast<const pair_t*>(v + sz + sx);  // (a0+1, ., e0+1)
                // same interpolation order as the x-fastest branch: along x, y, then z
                const float x00 = q00.x + ta * (q10.x - q00.x), x10 = q00.y + ta * (q10.y - q00.y);
                const float x01 = q01.x + ta * (q11.x - q01.x), x11 = q01.y + ta * (q11.y - q01.y);
                const float y0 = x00 + tb * (x10 - x00), y1 = x01 + tb * (x11 - x01);
                acc += y0 + te * (y1 - y0);
            }
        } else if (a.vol_zx) {
            // z-pair copies of the x-fastest model: the cell's four voxels at
            // y = b0 are one 16-byte load, (a0, e0) (a0, e0+1) (a0+1, e0)
            // (a0+1, e0+1); the same at y = b0 + 1
            const int sy = 2 * d1;
            for (int k = k0; k <= k1; ++k) {
                const float Z = (float)(k - d3 / 2);
                const float fx = bx + rx * Z, fy = by + ry * Z, fz = bz + rz * Z;
                int a0, b0, e0;
                float ta, tb, te;
                if (!lin_axis(fx, d1, &a0, &ta) || !lin_axis(fy, d2, &b0, &tb) || !lin_axis(fz, d3, &e0, &te))
                    continue;
                const int odd = e0 & 1;
                const float* v = a.vol_zx + (odd ? a.zodd : 0) + 2 * (a0 + d1 * (b0 + d2 * (e0 >> 1)));
                const quad_t q0 = *reinterpret_cast<const quad_t*>(v);        // y = b0
                const quad_t q1 = *reinterpret_cast<const quad_t*>(v + sy);   // y = b0 + 1
                const float x00 = q0.x + ta * (q0.z - q0.x), x10 = q1.x + ta * (q1.z - q1.x);
                const float x01 = q0.y + ta * (q0.w - q0.y), x11 = q1.y + ta * (q1.w - q1.y);
                const float y0 = x00 + tb * (x10 - x00), y1 = x01 + tb * (x11 - x01);
                acc += y0 + te * (y1 - y0);
            }
        } else {
            const float* vol = a.vol;             // index x + d1*(y + d2*z)
            const int sy = d1, sz = d1 * d2;
            for (int k = k0; k <= k1; ++k) {
                const float Z = (float)(k - d3 / 2);
                const float fx = bx + rx * Z, fy = by + ry * Z, fz = bz + rz * Z;
                int a0, b0, e0;
                float ta, tb, te;
                if (!lin_axis(fx, d1, &a0, &ta) || !lin_axis(fy, d2, &b0, &tb) || !lin_axis(fz, d3, &e0, &te))
                    continue;
                const float* v = vol + (a0 + d1 * (b0 + d2 * e0));
                const pair_t q00 = *reinterpret_cast<const pair_t*>(v);
                const pair_t q10 = *reinterpret_cast<const pair_t*>(v + sy);
                const pair_t q01 = *reinterpret_cast<const pair_t*>(v + sz);
                const pair_t q11 = *reinterpret_cast<const pair_t*>(v + sz + sy);
                const float x00 = q00.x + ta * (q00.y - q00.x), x10 = q10.x + ta * (q10.y - q10.x);
                const float x01 = q01.x + ta * (q01.y - q01.x), x11 = q11.x + ta * (q11.y - q11.x);
                const float y0 = x00 + tb * (x10 - x00), y1 = x01 + tb * (x11 - x01);
                acc += y0 + te * (y1 - y0);
            }
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
    tile_sums(a, p, tile, tiles, q0, q1, q2, q3);
}

// The R-factor partials of one workgroup: wave sums, then a fixed-order sum of
// the four waves into slot (p, c, tile).
__device__ __forceinline__ void tile_sums(const ProjArgs& a, int p, int tile, int tiles, double q0, double q1,
                                          double q2, double q3)
{
    if (!a.sums) return;
    __shared__ double red[4][kRows];
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
        for (int k = 0; k < kRows; ++k) t += red[threadIdx.x][k];
        a.sums[((int64_t)p * 4 + threadIdx.x) * tiles + tile] = t;
    }
}

// y-fastest copy of an x-fastest volume (dst[y + d2*(x + d1*z)] = src[x + d1*(y + d2*z)])
// through a 32 x 32 LDS tile per z-slice.
__global__ __launch_bounds__(256) void k_yfast(const float* __restrict__ src, float* __restrict__ dst, int d1, int d2,
                                               int d3)
{
    __shared__ float t[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32, z = blockIdx.z;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    const int64_t zoff = (int64_t)d1 * d2 * z;
    for (int r = ty; r < 32; r += 8) {
        const int x = bx + tx, y = by + r;
        if (x < d1 && y < d2) t[r][tx] = src[zoff + x + (int64_t)d1 * y];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int y = by + tx, x = bx + r;
        if (x < d1 && y < d2) dst[zoff + y + (int64_t)d2 * x] = t[tx][r];
    }
}

// z-pair copy of a y-fastest volume: pair m of copy `odd` holds z = 2m + odd + s,
// s = 0, 1, interleaved per voxel: dst[s + 2*(y + d2*(x + d1*m))]; z >= d3 is 0.
__global__ __launch_bounds__(256) void k_zpair(const float* __restrict__ src, float* __restrict__ dst, int d1, int d2,
                                               int d3, int odd, int64_t total)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(t & 1);
        const int64_t yx = (t >> 1) % ((int64_t)d1 * d2), m = (t >> 1) / ((int64_t)d1 * d2);
        const int64_t z = 2 * m + odd + s;
        dst[t] = z < d3 ? src[yx + (int64_t)d1 * d2 * z] : 0.0f;
    }
}

__device__ __forceinline__ int wrap_n(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// z-pair copies of the x-fastest crop: pair m of copy `odd` holds z = 2m + odd + s
// at dst[s + 2 (x + d1 (y + d2 m))], straight from the FFT-order volume.
__global__ __launch_bounds__(256) void k_crop_zpairs_x(const float* __restrict__ src, int n1, int n2, int n3,
                                                       float* __restrict__ dst, int d1, int d2, int d3,
                                                       int64_t per)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 2 * per;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = t % per, odd = t / per;
        const int s = (int)(o & 1);
        const int64_t xy = (o >> 1) % ((int64_t)d1 * d2), m = (o >> 1) / ((int64_t)d1 * d2);
        const int x = (int)(xy % d1), y = (int)(xy / d1);
        const int z = (int)(2 * m + odd + s);
        dst[t] = z < d3 ? src[(int64_t)wrap_n(x - d1 / 2, n1) +
                              (int64_t)n1 * (wrap_n(y - d2 / 2, n2) + (int64_t)n2 * wrap_n(z - d3 / 2, n3))]
                        : 0.0f;
    }
}

// One wave per (projection, sum): lane sums over a stride of 64, then the wave,
// in a fixed order.
__global__ __launch_bounds__(256) void k_project_reduce(const double* sums, int P, int nblk, double* out)
{
    const int w = (int)(blockIdx.x * blockDim.x + threadIdx.x) / 64, l = (int)threadIdx.x % 64;
    if (w >= P * 4) return;
    double s = 0.0;
    for (int b = l; b < nblk; b += 64) s += sums[(int64_t)w * nblk + b];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (l == 0) out[w] = s;
}

// The z-pair copies (as k_zpair) straight from the FFT-order iterate: model
// voxel (x, y, z) is src[wrap(x - d1/2) + n1 (wrap(y - d2/2) + n2 wrap(z - d3/2))]
// (the centred crop of fftshift).  A workgroup takes a 32 x 32 (x, y) tile of
// pair m: it stages slices 2m, 2m+1, 2m+2 through LDS (coalesced along x) and
// writes pair m of both copies (coalesced along y).
__global__ __launch_bounds__(256) void k_crop_zpairs(const float* __restrict__ src, int n1, int n2, int n3,
                                                     float* __restrict__ dst, int d1, int d2, int d3,
                                                     int64_t zodd)
{
    __shared__ float t[3][32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32, m = blockIdx.z;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int z = 2 * m + q;
        for (int r = ty; r < 32; r += 8) {
            const int x = bx + tx, y = by + r;
            float v = 0.0f;
            if (x < d1 && y < d2 && z < d3)
                v = src[(int64_t)wrap_n(x - d1 / 2, n1) +
                        (int64_t)n1 * (wrap_n(y - d2 / 2, n2) + (int64_t)n2 * wrap_n(z - d3 / 2, n3))];
            t[q][r][tx] = v;
        }
    }
    __syncthreads();
    const int l = threadIdx.x % 64, yy = l >> 1, sp = l & 1;
    const int y = by + yy;
    for (int xr = threadIdx.x / 64; xr < 32; xr += 4) {
        const int x = bx + xr;
        if (x >= d1 || y >= d2) continue;
        const int64_t o = sp + 2 * (y + (int64_t)d2 * (x + (int64_t)d1 * m));
        dst[o] = t[sp][yy][xr];             // even copy: z = 2m + sp
        dst[zodd + o] = t[1 + sp][yy][xr];  // odd copy: z = 2m + 1 + sp
    }
}

}  // namespace

int project_blocks(int m1, int m2)
{
    // tiles are kLanes x kRows pixels, oriented either way, or 16 x 16 patches:
    // the largest count
    const int a = ((m1 + kLanes - 1) / kLanes) * ((m2 + kRows - 1) / kRows);
    const int b = ((m2 + kLanes - 1) / kLanes) * ((m1 + kRows - 1) / kRows);
    const int c = ((m1 + 15) / 16) * ((m2 + 15) / 16);
    return std::max(a, std::max(b, c));
}

void project_orientation(float* R)
{
    // lanes along pixel i (column 0 of R in the volume) or j (column 1), the
    // model x- or y-fastest: the pairing whose lane direction is closest to the
    // layout's fast axis, |R[3c + r]| largest; ties keep (j, y) then (i, x), so a
    // tilt series about y (theta = 0 included) needs no x-fastest form.  When no
    // pairing is within about 25 degrees (|R| < 0.9) a row of 64 lanes drifts
    // across tens of cache lines per gather, and a wave takes an 8 x 8 pixel
    // patch instead (R[11] = 2) on the layout its two pixel axes favour.
    const float jy = fabsf(R[4]), ix = fabsf(R[0]), iy = fabsf(R[1]), jx = fabsf(R[3]);
    float lanes = 1.0f;
    bool layout_y = true;
    float best = jy;
    if (ix > best) { best = ix; lanes = 0.0f; layout_y = false; }
    if (iy > best) { best = iy; lanes = 0.0f; layout_y = true; }
    if (jx > best) { best = jx; lanes = 1.0f; layout_y = false; }
    if (best < kRowAlign) {
        lanes = 2.0f;
        layout_y = iy + jy >= ix + jx;
    }
    R[9] = layout_y ? 1.0f : 0.0f;
    R[11] = lanes;
}

bool project_y_exact(const float* R)
{
    return R[1] == 0.0f && R[3] == 0.0f && R[4] == 1.0f && R[5] == 0.0f && R[7] == 0.0f;
}

void volume_yfast(const float* d_src, float* d_dst, int d1, int d2, int d3, hipStream_t st)
{
    dim3 grid((d1 + 31) / 32, (d2 + 31) / 32, d3);
    k_yfast<<<grid, 256, 0, st>>>(d_src, d_dst, d1, d2, d3);
    GF_LAUNCH_CHECK();
}

int64_t zpair_floats(int d1, int d2, int d3) { return 2 * (int64_t)d1 * d2 * (d3 / 2 + 1); }

void volume_zpairs(const float* d_vol_y, float* d_dst, int d1, int d2, int d3, hipStream_t st)
{
    const int64_t total = zpair_floats(d1, d2, d3);
    for (int odd = 0; odd < 2; ++odd) {
        const int64_t blocks = std::min<int64_t>((total + 255) / 256, 1 << 20);
        k_zpair<<<(unsigned)blocks, 256, 0, st>>>(d_vol_y, d_dst + odd * total, d1, d2, d3, odd, total);
        GF_LAUNCH_CHECK();
    }
}

void crop_zpairs_x(const float* d_src, int n1, int n2, int n3, float* d_dst, int d1, int d2, int d3, hipStream_t st)
{
    const int64_t per = zpair_floats(d1, d2, d3);
    const int64_t blocks = std::min<int64_t>((2 * per + 255) / 256, 1 << 20);
    k_crop_zpairs_x<<<(unsigned)blocks, 256, 0, st>>>(d_src, n1, n2, n3, d_dst, d1, d2, d3, per);
    GF_LAUNCH_CHECK();
}

void crop_zpairs(const float* d_src, int n1, int n2, int n3, float* d_dst, int d1, int d2, int d3, hipStream_t st)
{
    dim3 grid((d1 + 31) / 32, (d2 + 31) / 32, d3 / 2 + 1);
    k_crop_zpairs<<<grid, 256, 0, st>>>(d_src, n1, n2, n3, d_dst, d1, d2, d3, zpair_floats(d1, d2, d3));
    GF_LAUNCH_CHECK();
}

void project(const ProjArgs& a, hipStream_t st)
{
    GF_CHECK(a.P >= 1, GENFIRE_ERR_ARG, "projector needs at least one projection");
    GF_CHECK(a.d1 >= 2 && a.d2 >= 2 && a.d3 >= 2, GENFIRE_ERR_ARG, "projector needs every volume dimension >= 2");
    GF_CHECK(a.m1 <= a.d1 && a.m2 <= a.d2, GENFIRE_ERR_ARG, "measured projection larger than the volume");
    GF_CHECK((int64_t)a.d1 * a.d2 * a.d3 < (1ll << 31), GENFIRE_ERR_UNSUPPORTED, "projector volume above 2^31 voxels");
    const int tiles = project_blocks(a.m1, a.m2);
    const int per = (a.P + kXcd - 1) / kXcd, group = std::min(kProjGroup, per);
    const int groups = (per + group - 1) / group;
    const int64_t blocks = per == 1 ? (int64_t)kXcd * a.P * ((tiles + kXcd - 1) / kXcd)
                                    : (int64_t)kXcd * groups * group * tiles;
    // a dispatch counts its x extent in work items, in 32 bits
    GF_CHECK(blocks * kLanes * kRows < (1ll << 32), GENFIRE_ERR_UNSUPPORTED,
             "projector launch above 2^32 work items: too many projections of this size");
    k_project<<<(unsigned)blocks, kLanes * kRows, 0, st>>>(a, tiles, per, group);
    GF_LAUNCH_CHECK();
}

void project_reduce(const double* d_sums, int P, int nblk, double* d_out4, hipStream_t st)
{
    const int n = P * 4;
    k_project_reduce<<<(n + 3) / 4, 256, 0, st>>>(d_sums, P, nblk, d_out4);
    GF_LAUNCH_CHECK();
}

}  // namespace gf
