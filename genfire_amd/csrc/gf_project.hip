// gf_project.hip -- real-space projector and R-factor sums
// (src/calculate3Dprojection_RealSpaceinterp.m:1-57, reconstruct.m:168-185,
// src/Tian_calc_Rfactor_realspace.m:1-24).
//
// Proj(i,j) = sum_k Vol(R [X_i; Y_j; Z_k]) with centred coordinates
// X = i - floor(d1/2) (..RealSpaceinterp.m:18-20) and trilinear interpolation with
// extrapolation value 0: a sample is zero unless every coordinate lies in the
// closed interval [0, d-1] of its axis.  One lane marches one ray (one output
// pixel) along z, summing its samples in ascending k.  The R-factor sums of
// each tile are written as per-tile partials and reduced in a fixed order.
//
// The model is read in a z-duplicated form: element (x, y, z) of 8 bytes holds
// the voxel (x, y, z) and its neighbour (x, y, z + 1), laid out y-fastest
// (index y + (d2+1) (x + (d1+1) z)) or x-fastest (x + (d1+1) (y + (d2+1) z)),
// with a zero row past the last x and y and a zero neighbour past the last z.
// A sample's eight voxels are then two 16-byte loads: the four (y, y+1) x
// (z, z+1) values at x = a0 and at a0 + 1 (y-fastest form), or the four
// (x, x+1) x (z, z+1) values at y = b0 and b0 + 1 (x-fastest form), with no
// clamp at the top face: a coordinate equal to d - 1 takes cell d - 1 with
// weight 0 on the zero row.
#include <algorithm>
#include <cstdlib>

#include "gf_grid.h"

namespace gf {

namespace {

constexpr int kXcd = 8;
constexpr int kProjGroup = 6;       // four-ray form: neighbouring projections whose tiles interleave
constexpr int kRayGroup = 12;       // the same for one ray per lane
constexpr int kLanes = 64;          // pixels of one wave along the fast pixel axis (four-ray form)
constexpr int kRows = 4;            // waves per workgroup along the other pixel axis
constexpr int kRun = 32;            // one ray per lane: a wave's run along the fast pixel axis
constexpr int kRunRows = 2;         // ... and its rows along the other (kRun x kRunRows = 64 lanes)
constexpr float kRowAlign = 0.9f;   // lane rows when a pairing's |R| reaches this, else patches
// z-duplicated forms above kSlabMinForm are projected in slabs of at most
// kSlabForm bytes (project_slabs), half the 256 MiB last-level cache
constexpr int64_t kSlabMinForm = 128ll << 20;
constexpr int64_t kSlabForm = 128ll << 20;

// Two (voxel, z + 1 neighbour) pairs in one 16-byte access (8-byte aligned).
typedef float quad_t __attribute__((ext_vector_type(4), aligned(8)));

__device__ __forceinline__ bool lin_axis(float f, int n, int* i0, float* t)
{
    if (!(f >= 0.0f) || f > (float)(n - 1)) return false;
    int i = (int)floorf(f);
    if (i > n - 2) i = n - 2;
    *i0 = i;
    *t = f - (float)i;
    return true;
}

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

// Every projection maps y exactly onto y (the four-ray form).  With more than 8
// projections the workgroups of one projection run on one XCD: the hardware
// deals workgroups to the 8 XCDs round-robin, so physical block b goes to XCD
// b % 8.  XCD x takes the contiguous block of per = ceil(P / 8) projections
// starting at x * per, in groups of G neighbouring projections whose tiles are
// interleaved (slot = (g * tiles + tile) * G + m for member m of the XCD's
// group g), so that the G workgroups of one tile read nearly the same voxels
// at about the same time and share the XCD's L2 fills.  With at most 8
// projections (per = 1) every XCD instead takes a contiguous eighth of each
// projection's tiles.
__global__ __launch_bounds__(kLanes* kRows) void k_project_ytilt(ProjArgs a, int tiles, int per, int group)
{
    const int xcd = blockIdx.x % kXcd, slot = blockIdx.x / kXcd;
    int p, tile;
    if (per == 1) {
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
    double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
    if (tile < ytilt_tiles(a)) project_ytilt(a, a.R + kRStride * p, p, tile, q0, q1, q2, q3);
    tile_sums(a, p, tile, tiles, q0, q1, q2, q3);
}

// ---------------------------------------------------------------------------
// One ray per lane over the z-duplicated model (every other rotation)
// ---------------------------------------------------------------------------

struct Ray {
    float bx, by, bz;   // sample coordinates at Z = 0
    float rx, ry, rz;   // their step per unit of Z (the third column of R)
};

// The sample coordinates of step k are fma(r, Z, b) with Z = k - floor(d3/2),
// the same expression in the range search and in the march.
__device__ __forceinline__ bool ray_inside(const Ray& r, float Z, float hx, float hy, float hz)
{
    const float fx = __builtin_fmaf(r.rx, Z, r.bx);
    const float fy = __builtin_fmaf(r.ry, Z, r.by);
    const float fz = __builtin_fmaf(r.rz, Z, r.bz);
    return fx >= 0.0f && fx <= hx && fy >= 0.0f && fy <= hy && fz >= 0.0f && fz <= hz;
}

// Steps [k0, k1] whose sample lies inside the box.  Each coordinate is a
// correctly rounded fma of a linear function of Z, hence monotonic in k, so the
// inside steps form one interval: estimate it from the three slab intervals,
// then move each end to the exact boundary with the inside test itself.
__device__ __forceinline__ void ray_range(const Ray& r, int d1, int d2, int d3, int* k0, int* k1)
{
    const int c3 = d3 / 2;
    const float hx = (float)(d1 - 1), hy = (float)(d2 - 1), hz = (float)(d3 - 1);
    float klo = 0.0f, khi = hz;
    const float bb[3] = {r.bx, r.by, r.bz}, rr[3] = {r.rx, r.ry, r.rz}, hi[3] = {hx, hy, hz};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (rr[q] != 0.0f) {
            const float t0 = (0.0f - bb[q]) / rr[q] + (float)c3, t1 = (hi[q] - bb[q]) / rr[q] + (float)c3;
            klo = fmaxf(klo, fminf(t0, t1) - 1.0f);
            khi = fminf(khi, fmaxf(t0, t1) + 1.0f);
        } else if (!(bb[q] >= 0.0f && bb[q] <= hi[q])) {
            khi = -1.0f;
        }
    }
    int a = (int)fminf((float)d3, fmaxf(0.0f, ceilf(klo)));
    int b = (int)fmaxf(-1.0f, fminf((float)(d3 - 1), floorf(khi)));
    if (a <= b) {
        while (a > 0 && ray_inside(r, (float)(a - 1 - c3), hx, hy, hz)) --a;
        while (b < d3 - 1 && ray_inside(r, (float)(b + 1 - c3), hx, hy, hz)) ++b;
        while (a <= b && !ray_inside(r, (float)(a - c3), hx, hy, hz)) ++a;
        while (b >= a && !ray_inside(r, (float)(b - c3), hx, hy, hz)) --b;
    }
    *k0 = a;
    *k1 = b;
}

// Pixel of this lane in tile `tile` of a projection whose lane arrangement is
// R[11]: 1, a wave takes 32 consecutive j in two neighbouring rows i (lanes
// 0-31 row i, 32-63 row i + 1), a tile 32 j x 8 i; 0, the same with i and j
// exchanged; 2, 8 x 8 patches, 16 x 16 per tile.  The two rows of a wave
// sample cells about one voxel apart, so a row's second corner load (x = a0 +
// 1 in the y-fastest form) meets the lines the other row's first load brought
// into L1 at the same step: for 91 projections of a 256^3 model this takes the
// projector from 2.37 to 2.15 ms per evaluation against one row of 64 j (16 j x
// 4 i: 2.39 ms).
__device__ __forceinline__ void ray_pixel(const ProjArgs& a, float form, int tile, int* i, int* j)
{
    const int w = (int)threadIdx.x / kLanes, l = (int)threadIdx.x % kLanes;
    if (form == 2.0f) {
        const int tp1 = (a.m1 + 15) / 16;
        *i = (tile % tp1) * 16 + (w & 1) * 8 + (l & 7);
        *j = (tile / tp1) * 16 + (w >> 1) * 8 + (l >> 3);
    } else if (form != 0.0f) {
        const int tiles2 = (a.m2 + kRun - 1) / kRun;
        *i = (tile / tiles2) * (kRows * kRunRows) + w * kRunRows + l / kRun;
        *j = (tile % tiles2) * kRun + l % kRun;
    } else {
        const int tiles1 = (a.m1 + kRun - 1) / kRun;
        *i = (tile % tiles1) * kRun + l % kRun;
        *j = (tile / tiles1) * (kRows * kRunRows) + w * kRunRows + l / kRun;
    }
}

// Pixel (i, j) of the measured window sits at (i + s1, j + s2) of the d1 x d2
// back projection (My_stripzero of the full projection, Tian...:16).
__device__ __forceinline__ Ray make_ray(const ProjArgs& a, const float* R, int i, int j)
{
    const int s1 = a.d1 / 2 - a.m1 / 2, s2 = a.d2 / 2 - a.m2 / 2;
    const float X = (float)(i + s1 - a.d1 / 2), Y = (float)(j + s2 - a.d2 / 2);
    Ray r;
    r.bx = __builtin_fmaf(R[0], X, __builtin_fmaf(R[3], Y, (float)(a.d1 / 2)));
    r.by = __builtin_fmaf(R[1], X, __builtin_fmaf(R[4], Y, (float)(a.d2 / 2)));
    r.bz = __builtin_fmaf(R[2], X, __builtin_fmaf(R[5], Y, (float)(a.d3 / 2)));
    r.rx = R[6];
    r.ry = R[7];
    r.rz = R[8];
    return r;
}

// Byte offset of cell (a0, b0, e0)'s element: index c + s1 r + s12 e0 with
// (c, r) = (b0, a0), s1 = d2 + 1 in the y-fastest form and (a0, b0), d1 + 1 in
// the x-fastest one, s12 = (d1 + 1)(d2 + 1); in 32 bits (24-bit products)
// unless the form exceeds 4 GiB (WIDE).
template <bool WIDE>
__device__ __forceinline__ auto zdup_offset(int c, int r, int e0, uint32_t s1, uint32_t s12)
{
    if constexpr (WIDE) {
        return ((uint64_t)s12 * (uint64_t)e0 + (uint64_t)s1 * (uint64_t)r + (uint64_t)c) * 8;
    } else {
        const uint32_t i = __umul24(s12, (uint32_t)e0) + (__umul24(s1, (uint32_t)r) + (uint32_t)c);
        return i << 3;
    }
}

// March the ray from Z to Z1 (Z = k - floor(d3/2) counts whole steps, exact in
// FP32), adding its samples to acc in ascending k.  v0 is the form's base, v1
// the next x row (YF) or y row.
template <bool YF, bool WIDE>
__device__ __forceinline__ float march(const char* __restrict__ v0, const char* __restrict__ v1, uint32_t s1,
                                       uint32_t s12, const Ray& r, float Z, float Z1, float acc)
{
    for (; Z <= Z1; Z += 1.0f) {
        const float fx = __builtin_fmaf(r.rx, Z, r.bx), fy = __builtin_fmaf(r.ry, Z, r.by);
        const float fz = __builtin_fmaf(r.rz, Z, r.bz);
        const int a0 = (int)fx, b0 = (int)fy, e0 = (int)fz;   // f >= 0: truncation is floor
        const float ta = __builtin_amdgcn_fractf(fx), tb = __builtin_amdgcn_fractf(fy);
        const float te = __builtin_amdgcn_fractf(fz);
        const auto off = YF ? zdup_offset<WIDE>(b0, a0, e0, s1, s12) : zdup_offset<WIDE>(a0, b0, e0, s1, s12);
        const quad_t q0 = *reinterpret_cast<const quad_t*>(v0 + off);
        const quad_t q1 = *reinterpret_cast<const quad_t*>(v1 + off);
        float x00, x10, x01, x11;   // x-interpolated values at (y, z) = (b0, e0), (b0+1, e0), (b0, e0+1), (b0+1, e0+1)
        if constexpr (YF) {         // q = (b0, e0) (b0, e0+1) (b0+1, e0) (b0+1, e0+1) at x = a0 and a0 + 1
            x00 = q0.x + ta * (q1.x - q0.x);
            x10 = q0.z + ta * (q1.z - q0.z);
            x01 = q0.y + ta * (q1.y - q0.y);
            x11 = q0.w + ta * (q1.w - q0.w);
        } else {                    // q = (a0, e0) (a0, e0+1) (a0+1, e0) (a0+1, e0+1) at y = b0 and b0 + 1
            x00 = q0.x + ta * (q0.z - q0.x);
            x10 = q1.x + ta * (q1.z - q1.x);
            x01 = q0.y + ta * (q0.w - q0.y);
            x11 = q1.y + ta * (q1.w - q1.y);
        }
        const float y0 = x00 + tb * (x10 - x00), y1 = x01 + tb * (x11 - x01);
        acc += y0 + te * (y1 - y0);
    }
    return acc;
}

// Cubic convolution (interp3 'cubic', calculate3Dprojection_RealSpaceinterp.m:
// 45,51): Keys' kernel with a = -1/2 on the 4 x 4 x 4 nodes around the cell
// min(floor(f), d - 2), read from the padded form (whose boundary nodes follow
// Keys' rule), one 16-byte load per (x, z) node pair: the four y nodes.  The
// weights of t in [0, 1] for the nodes i - 1 .. i + 2 are
// ((-t + 2) t - 1) t / 2, ((3t - 5) t t + 2) / 2, ((-3t + 4) t + 1) t / 2,
// (t - 1) t t / 2.  Offsets in 64 bits: the form passes 4 GiB at d = 1024.
typedef float quad4_t __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ void keys_w(float t, float w[4])
{
    w[0] = ((-t + 2.0f) * t - 1.0f) * t * 0.5f;
    w[1] = ((3.0f * t - 5.0f) * t * t + 2.0f) * 0.5f;
    w[2] = ((-3.0f * t + 4.0f) * t + 1.0f) * t * 0.5f;
    w[3] = (t - 1.0f) * t * t * 0.5f;
}

__device__ __forceinline__ float march_cubic(const float* __restrict__ form, int d1, int d2, int d3, const Ray& r,
                                             float Z, float Z1)
{
    const int64_t py = d2 + 2, pxy = (int64_t)(d1 + 2) * py;
    float acc = 0.0f;
    for (; Z <= Z1; Z += 1.0f) {
        const float fx = __builtin_fmaf(r.rx, Z, r.bx), fy = __builtin_fmaf(r.ry, Z, r.by);
        const float fz = __builtin_fmaf(r.rz, Z, r.bz);
        const int a0 = min((int)fx, d1 - 2), b0 = min((int)fy, d2 - 2), e0 = min((int)fz, d3 - 2);
        float wx[4], wy[4], wz[4];
        keys_w(fx - (float)a0, wx);
        keys_w(fy - (float)b0, wy);
        keys_w(fz - (float)e0, wz);
        // padded node (a0 - 1, b0 - 1, e0 - 1) sits at (a0, b0, e0)
        const float* base = form + ((int64_t)b0 + py * (int64_t)a0 + pxy * (int64_t)e0);
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float sc = 0.0f;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const quad4_t q = *reinterpret_cast<const quad4_t*>(base + py * a + pxy * c);
                sc += wx[a] * (wy[0] * q.x + wy[1] * q.y + wy[2] * q.z + wy[3] * q.w);
            }
            s += wz[c] * sc;
        }
        acc += s;
    }
    return acc;
}

// Slab sweep (models whose z-duplicated form passes 128 MiB): the form is cut
// into nslab slabs of slab_t z planes and every projection's rays cross one
// slab per launch, so all projections share each slab's fills in the 256 MiB
// last-level cache instead of each streaming the whole form (91 projections of
// a 512^3 model, a 1.08 GB form: 26.6 -> 19.2 ms per evaluation in 9 slabs; of
// a 256^3 model, 135 MB: 2.22 -> 2.16 ms in 2).  A ray's running sum
// waits in acc[pixel] between launches; slabs are taken in the ray's own k
// order (ascending z when rz >= 0, descending otherwise), so the sum adds the
// same samples in the same order as one march and is bit for bit the same.
//
// The first k in [k0, k1 + 1] from which pred holds (pred monotonic, false
// then true), starting from the estimate e and checked exactly; a wrong
// estimate falls back to bisection.
template <class Pred>
__device__ __forceinline__ int first_true(int k0, int k1, float e, Pred pred)
{
    int r = (int)fminf((float)(k1 + 1), fmaxf((float)k0, e));
    if ((r == k0 || !pred(r - 1)) && (r == k1 + 1 || pred(r))) return r;
    int lo = k0, hi = k1 + 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pred(mid)) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Narrows the ray's steps [k0, k1] to those whose sample's z lies in
// [zlo, zhi): fz(k) = fma(rz, Z, bz) is monotonic in k.
__device__ __forceinline__ void slab_range(const Ray& r, int c3, float zlo, float zhi, int* k0, int* k1)
{
    const int a = *k0, b = *k1;
    if (a > b) return;
    auto fz = [&](int k) { return __builtin_fmaf(r.rz, (float)(k - c3), r.bz); };
    if (r.rz == 0.0f) {
        if (!(r.bz >= zlo && r.bz < zhi)) *k1 = a - 1;
        return;
    }
    const float inv = 1.0f / r.rz, c = (float)c3;
    int lo, hi;
    if (r.rz > 0.0f) {
        lo = first_true(a, b, ceilf((zlo - r.bz) * inv + c), [&](int k) { return fz(k) >= zlo; });
        hi = first_true(a, b, ceilf((zhi - r.bz) * inv + c), [&](int k) { return fz(k) >= zhi; });
    } else {
        lo = first_true(a, b, floorf((zhi - r.bz) * inv + c) + 1.0f, [&](int k) { return fz(k) < zhi; });
        hi = first_true(a, b, floorf((zlo - r.bz) * inv + c) + 1.0f, [&](int k) { return fz(k) < zlo; });
    }
    *k0 = lo;
    *k1 = hi - 1;
}

// Every projection of a launch that is not all exact tilts about y: one ray
// per lane on the z-duplicated forms, a wave's lanes in two runs of 32 along
// the pixel axis R[11] picks (or 8 x 8 patches), and the block mapping of the
// four-ray form:
// XCD x takes a contiguous block of ceil(P/8) projections and interleaves the
// tiles of groups of G neighbouring ones, so that the G workgroups of one tile
// share the XCD's L2 fills.  With at most 8 projections each XCD takes an
// eighth of every projection's tiles.
template <bool WIDE, bool CUBIC, bool SLAB>
__global__ __launch_bounds__(kLanes* kRows) void k_project(ProjArgs a, int tiles, int per, int group, int slab,
                                                           int nslab, int slab_t)
{
    static_assert(!(CUBIC && SLAB), "the slab sweep is for the trilinear forms");
    const int xcd = blockIdx.x % kXcd, slot = blockIdx.x / kXcd;
    int p, tile;
    if (per == 1) {
        const int tc = (tiles + kXcd - 1) / kXcd;
        p = slot / tc;
        tile = xcd * tc + slot % tc;
        if (tile >= tiles || p >= a.P) return;
    } else {
        const int k = (slot / (group * tiles)) * group + slot % group;
        p = xcd * per + k;
        tile = (slot / group) % tiles;
        if (k >= per || p >= a.P) return;
    }
    const int d1 = a.d1, d2 = a.d2, d3 = a.d3, c3 = d3 / 2;
    const uint32_t sx = d1 + 1, sy = d2 + 1, sxy = sx * sy;
    const float* R = a.R + kRStride * p;
    int i, j;
    ray_pixel(a, R[11], tile, &i, &j);
    float acc = 0.0f;
    const bool valid = i < a.m1 && j < a.m2;
    const int64_t o = (int64_t)i + (int64_t)a.m1 * (j + (int64_t)a.m2 * p);
    if (valid) {
        const Ray r = make_ray(a, R, i, j);
        int k0, k1;
        ray_range(r, d1, d2, d3, &k0, &k1);
        if constexpr (SLAB) {
            const int s = r.rz >= 0.0f ? slab : nslab - 1 - slab;
            slab_range(r, c3, (float)(s * slab_t), (float)min(d3, (s + 1) * slab_t), &k0, &k1);
            if (slab > 0) acc = a.acc[o];
        }
        const float Z0 = (float)(k0 - c3), Z1 = (float)(k1 - c3);
        if constexpr (CUBIC) {
            acc = march_cubic(a.vcubic, d1, d2, d3, r, Z0, Z1);
        } else if (R[9] != 0.0f) {
            const char* v0 = reinterpret_cast<const char*>(a.zy);
            acc = march<true, WIDE>(v0, v0 + 8 * (size_t)sy, sy, sxy, r, Z0, Z1, acc);
        } else {
            const char* v0 = reinterpret_cast<const char*>(a.zx);
            acc = march<false, WIDE>(v0, v0 + 8 * (size_t)sx, sx, sxy, r, Z0, Z1, acc);
        }
    }
    if constexpr (SLAB) {
        if (slab < nslab - 1) {   // uniform: the sums follow the last slab
            if (valid) a.acc[o] = acc;
            return;
        }
    }
    double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
    if (valid) {
        if (a.out) a.out[o] = acc;
        if (a.meas) {
            const double m = (double)a.meas[o];
            const double b = (double)acc;
            q0 = a.post_hoc ? fabs(fabs(b) - fabs(m)) : fabs(b - m);
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
    __syncthreads();   // red is reused by the next call of the same workgroup
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

// z-duplicated forms of the model (see the top of this file): element
// (x, y, z) = {vol(x, y, z), vol(x, y, z + 1)}, zero where x = d1, y = d2 or
// z + 1 = d3.  The model is read either from a plain column-major d1 x d2 x d3
// volume (n = d, no shift) or as the centred crop of an FFT-order n1 x n2 x n3
// volume (voxel (x, y, z) at (wrap(x - d1/2), wrap(y - d2/2), wrap(z - d3/2))).
__device__ __forceinline__ int wrap_n(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

struct ModelSrc {
    const float* src;
    int n1, n2, n3;   // source extents
    int o1, o2, o3;   // shift of voxel (x, y, z): source (x - o1, y - o2, z - o3) wrapped
    int d1, d2, d3;   // model extents
    __device__ __forceinline__ float at(int x, int y, int z) const
    {
        if (x >= d1 || y >= d2 || z >= d3) return 0.0f;
        return src[(int64_t)wrap_n(x - o1, n1) + (int64_t)n1 * (wrap_n(y - o2, n2) + (int64_t)n2 * wrap_n(z - o3, n3))];
    }
};

// y-fastest form: a workgroup transposes a 32 x 32 (x, y) tile of slices z and
// z + 1 through LDS (read along x, written along y with the pairs interleaved).
__global__ __launch_bounds__(256) void k_zdup_y(ModelSrc m, float* __restrict__ dst)
{
    __shared__ float t[2][32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32, z = blockIdx.z;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
#pragma unroll
    for (int q = 0; q < 2; ++q)
        for (int r = ty; r < 32; r += 8) t[q][r][tx] = m.at(bx + tx, by + r, z + q);
    __syncthreads();
    const int X1 = m.d1 + 1, Y2 = m.d2 + 1;
    const int l = threadIdx.x % 64, yy = l >> 1, sp = l & 1;
    const int y = by + yy;
    for (int xr = threadIdx.x / 64; xr < 32; xr += 4) {
        const int x = bx + xr;
        if (x >= X1 || y >= Y2) continue;
        dst[sp + 2 * ((int64_t)y + (int64_t)Y2 * (x + (int64_t)X1 * z))] = t[sp][yy][xr];
    }
}

// x-fastest form: one thread per element, reading along x.
__global__ __launch_bounds__(256) void k_zdup_x(ModelSrc m, float2* __restrict__ dst, int64_t total)
{
    const int X1 = m.d1 + 1, Y1 = m.d2 + 1;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(e % X1);
        const int64_t yz = e / X1;
        const int y = (int)(yz % Y1), z = (int)(yz / Y1);
        dst[e] = make_float2(m.at(x, y, z), m.at(x, y, z + 1));
    }
}

// The cubic projector's padded form (see gf_grid.h): padded node (x, y, z),
// each in [-1, d], is the model node, or on an axis where it lies outside Keys'
// boundary combination 3 f(0) - 3 f(1) + f(2) (and its mirror at the top),
// applied on every such axis: a sum of up to 27 model nodes.  One thread per
// padded node, y fastest.
__device__ __forceinline__ int keys_taps(int i, int n, int idx[3], float w[3])
{
    if (i < 0) { idx[0] = 0; idx[1] = 1; idx[2] = 2; w[0] = 3.0f; w[1] = -3.0f; w[2] = 1.0f; return 3; }
    if (i >= n) { idx[0] = n - 1; idx[1] = n - 2; idx[2] = n - 3; w[0] = 3.0f; w[1] = -3.0f; w[2] = 1.0f; return 3; }
    idx[0] = i; w[0] = 1.0f;
    return 1;
}

__global__ __launch_bounds__(256) void k_cubic_form(ModelSrc m, float* __restrict__ dst, int64_t total)
{
    const int py = m.d2 + 2, px = m.d1 + 2;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int y = (int)(e % py) - 1;
        const int64_t xz = e / py;
        const int x = (int)(xz % px) - 1, z = (int)(xz / px) - 1;
        int ix[3], iy[3], iz[3];
        float wx[3], wy[3], wz[3];
        const int nx = keys_taps(x, m.d1, ix, wx), ny = keys_taps(y, m.d2, iy, wy), nz = keys_taps(z, m.d3, iz, wz);
        float v = 0.0f;
        for (int c = 0; c < nz; ++c)
            for (int b = 0; b < ny; ++b)
                for (int a = 0; a < nx; ++a) v += wz[c] * wy[b] * wx[a] * m.at(ix[a], iy[b], iz[c]);
        dst[e] = v;
    }
}

void cubic_launch(const ModelSrc& m, float* d_dst, hipStream_t st)
{
    GF_CHECK(m.d1 >= 4 && m.d2 >= 4 && m.d3 >= 4, GENFIRE_ERR_ARG,
             "cubic projector: interp3 'cubic' needs at least 4 nodes along every axis");
    const int64_t total = cubic_form_floats(m.d1, m.d2, m.d3);
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 1 << 20);
    k_cubic_form<<<(unsigned)blocks, 256, 0, st>>>(m, d_dst, total);
    GF_LAUNCH_CHECK();
}

void zdup_launch(const ModelSrc& m, float* d_dst, bool yfast, hipStream_t st)
{
    if (yfast) {
        dim3 grid((m.d1 + 1 + 31) / 32, (m.d2 + 1 + 31) / 32, m.d3);
        k_zdup_y<<<grid, 256, 0, st>>>(m, d_dst);
    } else {
        const int64_t total = (int64_t)(m.d1 + 1) * (m.d2 + 1) * m.d3;
        const int64_t blocks = std::min<int64_t>((total + 255) / 256, 1 << 20);
        k_zdup_x<<<(unsigned)blocks, 256, 0, st>>>(m, reinterpret_cast<float2*>(d_dst), total);
    }
    GF_LAUNCH_CHECK();
}

}  // namespace

int project_blocks(int m1, int m2)
{
    // tiles are kRun x (kRows kRunRows) pixels of the one-ray forms, oriented
    // either way, or 16 x 16 patches (which also bound the four-ray form's 16
    // columns x 16 groups of four y): the largest count
    constexpr int tr = kRows * kRunRows;
    const int a = ((m1 + kRun - 1) / kRun) * ((m2 + tr - 1) / tr);
    const int b = ((m2 + kRun - 1) / kRun) * ((m1 + tr - 1) / tr);
    const int c = ((m1 + 15) / 16) * ((m2 + 15) / 16);
    return std::max(a, std::max(b, c));
}

void project_orientation(float* R)
{
    // lanes along pixel i (column 0 of R in the volume) or j (column 1), the
    // model x- or y-fastest: the pairing whose lane direction is closest to the
    // layout's fast axis, |R[3c + r]| largest; ties keep (j, y) then (i, x), so a
    // tilt series about y (theta = 0 included) needs no x-fastest form.  When no
    // pairing is within about 25 degrees (|R| < 0.9) a wave's runs of 32 lanes
    // drift across tens of cache lines per gather, and a wave takes an 8 x 8 pixel
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


int64_t zdup_floats(int d1, int d2, int d3) { return 2 * (int64_t)(d1 + 1) * (d2 + 1) * d3; }

int project_slabs(int d1, int d2, int d3)
{
    // GENFIRE_PROJECTOR_SLABS=n sets the slab count (1: one pass) for any model
    // size: the tests compare the sweep with one pass on small models
    if (const char* e = std::getenv("GENFIRE_PROJECTOR_SLABS")) {
        const int n = std::atoi(e);
        if (n >= 1) return std::min(n, d3);
    }
    const int64_t form = 4 * zdup_floats(d1, d2, d3);
    if (form <= kSlabMinForm) return 1;
    return (int)std::min<int64_t>(d3, (form + kSlabForm - 1) / kSlabForm);
}

int64_t cubic_form_floats(int d1, int d2, int d3) { return (int64_t)(d1 + 2) * (d2 + 2) * (d3 + 2); }

void volume_cubic_form(const float* d_vol, int d1, int d2, int d3, float* d_dst, hipStream_t st)
{
    cubic_launch(ModelSrc{d_vol, d1, d2, d3, 0, 0, 0, d1, d2, d3}, d_dst, st);
}

void crop_cubic_form(const float* d_src, int n1, int n2, int n3, float* d_dst, int d1, int d2, int d3,
                     hipStream_t st)
{
    cubic_launch(ModelSrc{d_src, n1, n2, n3, d1 / 2, d2 / 2, d3 / 2, d1, d2, d3}, d_dst, st);
}

void volume_zdup(const float* d_vol, int d1, int d2, int d3, float* d_dst, bool yfast, hipStream_t st)
{
    zdup_launch(ModelSrc{d_vol, d1, d2, d3, 0, 0, 0, d1, d2, d3}, d_dst, yfast, st);
}

void crop_zdup(const float* d_src, int n1, int n2, int n3, float* d_dst, int d1, int d2, int d3, bool yfast,
               hipStream_t st)
{
    zdup_launch(ModelSrc{d_src, n1, n2, n3, d1 / 2, d2 / 2, d3 / 2, d1, d2, d3}, d_dst, yfast, st);
}

ProjForms project_forms(const float* Rm, int P, int d2, bool cubic)
{
    ProjForms f;
    if (cubic) {
        f.cubic = true;
        return f;
    }
    bool all_exact = P > 0;
    for (int p = 0; p < P; ++p) {
        const float* r = Rm + (size_t)kRStride * p;
        all_exact = all_exact && r[10] != 0.0f;
        f.zy = f.zy || r[9] != 0.0f;
        f.zx = f.zx || r[9] == 0.0f;
    }
    f.four_ray = all_exact && d2 % 4 == 0;
    if (f.four_ray) f.zy = f.zx = false;
    return f;
}

namespace {

void launch_ray(const ProjArgs& a, int tiles, hipStream_t st)
{
    const int per = (a.P + kXcd - 1) / kXcd, group = std::min(kRayGroup, per);
    const int groups = (per + group - 1) / group;
    const int64_t blocks = per == 1 ? (int64_t)kXcd * a.P * ((tiles + kXcd - 1) / kXcd)
                                    : (int64_t)kXcd * groups * group * tiles;
    // a dispatch counts its x extent in work items, in 32 bits
    GF_CHECK(blocks * kLanes * kRows < (1ll << 32), GENFIRE_ERR_UNSUPPORTED,
             "projector launch above 2^32 work items: too many projections of this size");
    // 32-bit offsets from 24-bit products: the form below 4 GiB and the plane
    // size (d1 + 1)(d2 + 1) below 2^24
    const bool wide = 4 * zdup_floats(a.d1, a.d2, a.d3) > (int64_t)UINT32_MAX ||
                      (int64_t)(a.d1 + 1) * (a.d2 + 1) >= (1ll << 24);
    const dim3 g((unsigned)blocks), b(kLanes * kRows);
    const int nslab = a.acc && !a.vcubic ? project_slabs(a.d1, a.d2, a.d3) : 1;
    if (nslab > 1) {
        const int t = (a.d3 + nslab - 1) / nslab, ns = (a.d3 + t - 1) / t;
        for (int s = 0; s < ns; ++s) {
            if (wide) k_project<true, false, true><<<g, b, 0, st>>>(a, tiles, per, group, s, ns, t);
            else k_project<false, false, true><<<g, b, 0, st>>>(a, tiles, per, group, s, ns, t);
            GF_LAUNCH_CHECK();
        }
        return;
    }
    if (a.vcubic) k_project<true, true, false><<<g, b, 0, st>>>(a, tiles, per, group, 0, 1, a.d3);
    else if (wide) k_project<true, false, false><<<g, b, 0, st>>>(a, tiles, per, group, 0, 1, a.d3);
    else k_project<false, false, false><<<g, b, 0, st>>>(a, tiles, per, group, 0, 1, a.d3);
    GF_LAUNCH_CHECK();
}

}  // namespace

void project(const ProjArgs& a, hipStream_t st)
{
    GF_CHECK(a.P >= 1, GENFIRE_ERR_ARG, "projector needs at least one projection");
    GF_CHECK(a.d1 >= 2 && a.d2 >= 2 && a.d3 >= 2, GENFIRE_ERR_ARG, "projector needs every volume dimension >= 2");
    GF_CHECK(a.m1 >= 1 && a.m2 >= 1 && a.m1 <= a.d1 && a.m2 <= a.d2, GENFIRE_ERR_ARG,
             "measured projection empty or larger than the volume");
    GF_CHECK((int64_t)a.d1 * a.d2 * a.d3 < (1ll << 31), GENFIRE_ERR_UNSUPPORTED, "projector volume above 2^31 voxels");
    const int tiles = project_blocks(a.m1, a.m2);
    if (a.four_ray && !a.vcubic) {
        GF_CHECK(a.vol_y && a.d2 % 4 == 0, GENFIRE_ERR_ARG, "four-ray projector form needs the y-fastest model");
        const int per = (a.P + kXcd - 1) / kXcd, group = std::min(kProjGroup, per);
        const int groups = (per + group - 1) / group;
        const int64_t blocks = per == 1 ? (int64_t)kXcd * a.P * ((tiles + kXcd - 1) / kXcd)
                                        : (int64_t)kXcd * groups * group * tiles;
        // a dispatch counts its x extent in work items, in 32 bits
        GF_CHECK(blocks * kLanes * kRows < (1ll << 32), GENFIRE_ERR_UNSUPPORTED,
                 "projector launch above 2^32 work items: too many projections of this size");
        k_project_ytilt<<<(unsigned)blocks, kLanes * kRows, 0, st>>>(a, tiles, per, group);
        GF_LAUNCH_CHECK();
        return;
    }
    if (a.vcubic)
        GF_CHECK(a.d1 >= 4 && a.d2 >= 4 && a.d3 >= 4, GENFIRE_ERR_ARG,
                 "cubic projector: interp3 'cubic' needs at least 4 nodes along every axis");
    else
        GF_CHECK(a.zy || a.zx, GENFIRE_ERR_ARG, "projector needs a z-duplicated model form");
    launch_ray(a, tiles, st);
}

void project_reduce(const double* d_sums, int P, int nblk, double* d_out4, hipStream_t st)
{
    const int n = P * 4;
    k_project_reduce<<<(n + 3) / 4, 256, 0, st>>>(d_sums, P, nblk, d_out4);
    GF_LAUNCH_CHECK();
}

}  // namespace gf
