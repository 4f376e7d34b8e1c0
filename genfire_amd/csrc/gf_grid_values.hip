// gf_grid_values.hip -- DFT values, weighted average and centrosymmetric fill of
// fillInFourierGrid_DFT.m:236-300, plus the centred/FFT-order layout helpers.
#include "gf_grid.h"

namespace gf {

namespace {

constexpr int kVT = 256;  // points per workgroup of the value kernel (one per lane)
constexpr int kCB = 16;   // projection columns staged in LDS per step

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// exp(-2 pi i k t) for an integer k and an FP32 t.  The phase in turns, k*t, is
// exact in FP64 (|k| < 2^12, t has 24 mantissa bits); its fractional part is
// exact; only the final FP32 sincospi of the reduced turn rounds, so the
// argument error is 2^-25 turn whatever the size of k*t.
__device__ __forceinline__ float2 cis_turns(double k, float t)
{
    const double ph = k * (double)t;
    const float r = (float)(ph - rint(ph));
    float s, c;
    sincospif(2.0f * r, &s, &c);
    return make_float2(c, -s);
}

// F_j = sum_b exp(-2 pi i k2o[b] v_j) * sum_a exp(-2 pi i k1o[a] u_j) P[a,b]
// (the separable form of the exponent K10*u + K20*v at :243).  One lane per
// point, 256 points of one projection per workgroup.  The inner sum runs over
// 16-row blocks: exp(-2 pi i (16h + l - c1) u) = E_hi(h) * E_lo(l) with both
// factors from cis_turns, so every factor carries 2^-25-turn phase accuracy.
// Projection columns stream through LDS 16 at a time and are read as broadcasts.
__global__ __launch_bounds__(kVT) void k_dft_fp32(const float* __restrict__ proj_all, int dim1,
                                                  int dim2, const float* __restrict__ un,
                                                  const float* __restrict__ vn,
                                                  const PointTile* __restrict__ tiles,
                                                  float2* __restrict__ F)
{
    extern __shared__ float sP[];  // (16*nh1) rows x kCB columns
    const PointTile tl = tiles[blockIdx.x];
    const int i = threadIdx.x;
    const bool valid = i < tl.count;
    const int64_t j = tl.begin + (valid ? i : 0);
    const float t1 = valid ? un[j] : 0.0f;
    const float t2 = valid ? vn[j] : 0.0f;
    const float* P = proj_all + (int64_t)tl.proj * dim1 * dim2;
    const int c1 = dim1 / 2, c2 = dim2 / 2;
    const int nh1 = (dim1 + 15) / 16;
    const int rows = nh1 * 16;

    float2 elo1[16], elo2[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) {
        elo1[l] = cis_turns((double)l, t1);
        elo2[l] = cis_turns((double)l, t2);
    }
    float Fr = 0.0f, Fi = 0.0f;
    for (int b0 = 0; b0 < dim2; b0 += kCB) {
        __syncthreads();
        for (int idx = threadIdx.x; idx < rows * kCB; idx += kVT) {
            const int a = idx % rows, bb = idx / rows;
            sP[a * kCB + bb] = (a < dim1 && b0 + bb < dim2) ? P[a + (int64_t)dim1 * (b0 + bb)] : 0.0f;
        }
        __syncthreads();
        float gr[kCB], gi[kCB];
#pragma unroll
        for (int bb = 0; bb < kCB; ++bb) { gr[bb] = 0.0f; gi[bb] = 0.0f; }
        for (int h = 0; h < nh1; ++h) {
            const float2 ehi = cis_turns((double)(16 * h - c1), t1);
#pragma unroll
            for (int l = 0; l < 16; ++l) {
                const float2 e = cmul(ehi, elo1[l]);
                const float4* row = reinterpret_cast<const float4*>(sP + (16 * h + l) * kCB);
#pragma unroll
                for (int q = 0; q < kCB / 4; ++q) {
                    const float4 r = row[q];
                    gr[4 * q + 0] += e.x * r.x; gi[4 * q + 0] += e.y * r.x;
                    gr[4 * q + 1] += e.x * r.y; gi[4 * q + 1] += e.y * r.y;
                    gr[4 * q + 2] += e.x * r.z; gi[4 * q + 2] += e.y * r.z;
                    gr[4 * q + 3] += e.x * r.w; gi[4 * q + 3] += e.y * r.w;
                }
            }
        }
        const float2 e2hi = cis_turns((double)(b0 - c2), t2);
#pragma unroll
        for (int bb = 0; bb < kCB; ++bb) {
            const float2 e = cmul(e2hi, elo2[bb]);
            Fr += gr[bb] * e.x - gi[bb] * e.y;
            Fi += gr[bb] * e.y + gi[bb] * e.x;
        }
    }
    if (valid) F[j] = make_float2(Fr, Fi);
}

// :271-289 for one projection (each half-grid voxel appears at most once per
// projection, and projections are launched in order, so the sums are formed in
// the reference's list order without atomics):
//   exact (D < 1e-7): value overwrite, flag set (last projection wins, :291);
//   else w = 1/D in single (:275), accumulated in FP64.
__global__ void k_make_tiles(const int64_t* __restrict__ proj_off, const int64_t* __restrict__ tile_off, int P,
                             int tile_pts, PointTile* __restrict__ tiles, int ntiles)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    int lo = 0, hi = P - 1;   // the projection p with tile_off[p] <= t < tile_off[p + 1]
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (tile_off[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const int64_t b = proj_off[lo] + (int64_t)(t - tile_off[lo]) * tile_pts;
    PointTile tl;
    tl.begin = b;
    tl.count = (int32_t)min((int64_t)tile_pts, proj_off[lo + 1] - b);
    tl.proj = lo;
    tiles[t] = tl;
}

__global__ void k_scatter(const uint32_t* __restrict__ ind, const float* __restrict__ dist,
                          const float2* __restrict__ F, int64_t begin, int64_t count,
                          double2* __restrict__ acc, double* __restrict__ sw,
                          uint32_t* __restrict__ hits, float2* __restrict__ exact)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int64_t i = begin + t;
    const uint32_t h = ind[i];
    const float d = dist[i];
    const float2 f = F[i];
    const uint32_t hh = hits[h];
    if ((double)d < 1e-7) {
        exact[h] = f;
        hits[h] = (hh + 1u) | kExactBit;
    } else {
        // acc and sw are not cleared: a voxel's first point (no hits yet) writes
        // them.  After an exact point they may hold stale sums, which the
        // finalize ignores (the exact value wins).
        const float w = 1.0f / d;
        const double wd = (double)w;
        const bool first = hh == 0u;
        double2 a = first ? make_double2(0.0, 0.0) : acc[h];
        const double s = first ? 0.0 : sw[h];
        a.x += wd * (double)f.x;
        a.y += wd * (double)f.y;
        acc[h] = a;
        sw[h] = s + wd;
        hits[h] = hh + 1u;
    }
}

__device__ __forceinline__ int signed_freq(int s, int n) { return s < (n + 1) / 2 ? s : s - n; }

// FS(index_interp) = sum(wF)/sum(w); FS(index_k) = exact (:287-291); single cast
// (GENFIRE_Reconstructor.m:182); CentroSymmetricity.m: k1 <= 0 from FS(k),
// k1 > 0 from conj(FS(-k)) (all four parity branches reduce to this rule).
// Written directly in FFT order: voxel s holds frequency k = s (mod n).
// One workgroup per kFinRows rows (y) of one z plane of the FFT-order grid, a
// thread per x position (strided by 256): no division; a voxel with no hits
// (most of the grid) reads only its hit count.
constexpr int kFinRows = 8;
__global__ __launch_bounds__(256) void k_finalize(HalfGrid hg, int n1, int n2, int centro,
                                                  const double2* __restrict__ acc, const double* __restrict__ sw,
                                                  const uint32_t* __restrict__ hits, const float2* __restrict__ exact,
                                                  float2* __restrict__ M)
{
    const int s3 = blockIdx.y;
    const int y0 = blockIdx.x * kFinRows, y1 = min(n2, y0 + kFinRows);
    for (int s2 = y0; s2 < y1; ++s2)
        for (int s1 = threadIdx.x; s1 < n1; s1 += 256) {
            int k1 = signed_freq(s1, n1), k2 = signed_freq(s2, n2), k3 = signed_freq(s3, n1);
            bool cj = false;
            if (centro && k1 > 0) { k1 = -k1; k2 = -k2; k3 = -k3; cj = true; }
            const int64_t h = (int64_t)(k1 - hg.lo[0]) +
                              (int64_t)hg.len[0] * ((k2 - hg.lo[1]) + (int64_t)hg.len[1] * (k3 - hg.lo[2]));
            float2 v = make_float2(0.0f, 0.0f);
            const uint32_t hh = hits[h];
            if (hh & kExactBit) {
                v = exact[h];
            } else if (hh != 0) {   // hits > 0: positive weights were summed
                const double w = sw[h];
                if (w > 0.0) {
                    const double2 a = acc[h];
                    v = make_float2((float)(a.x / w), (float)(a.y / w));
                }
            }
            if (cj) v.y = -v.y;
            M[(int64_t)s1 + (int64_t)n1 * (s2 + (int64_t)n2 * s3)] = v;
        }
}

__global__ void k_shift(const float2* __restrict__ in, float2* __restrict__ out, int n1, int n2,
                        int n3, int to_centred)
{
    const int64_t total = (int64_t)n1 * n2 * n3;
    const int o1 = to_centred ? (n1 + 1) / 2 : n1 / 2;
    const int o2 = to_centred ? (n2 + 1) / 2 : n2 / 2;
    const int o3 = to_centred ? (n3 + 1) / 2 : n3 / 2;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int c1 = (int)(t % n1), c2 = (int)((t / n1) % n2), c3 = (int)(t / ((int64_t)n1 * n2));
        const int64_t src = (int64_t)((c1 + o1) % n1) +
                            (int64_t)n1 * (((c2 + o2) % n2) + (int64_t)n2 * ((c3 + o3) % n3));
        out[t] = in[src];
    }
}

__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

__global__ void k_crop(const void* __restrict__ src, int src_complex, int n1, int n2, int n3,
                       float* __restrict__ dst, int d1, int d2, int d3)
{
    const int64_t total = (int64_t)d1 * d2 * d3;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(t % d1), b = (int)((t / d1) % d2), c = (int)(t / ((int64_t)d1 * d2));
        const int s1 = wrap(a - d1 / 2, n1), s2 = wrap(b - d2 / 2, n2), s3 = wrap(c - d3 / 2, n3);
        const int64_t si = (int64_t)s1 + (int64_t)n1 * (s2 + (int64_t)n2 * s3);
        dst[t] = src_complex ? static_cast<const float2*>(src)[si].x : static_cast<const float*>(src)[si];
    }
}

template <class T>
__global__ void k_pad(const T* __restrict__ src, int d1, int d2, int d3, T* __restrict__ dst, int n1,
                      int n2, int n3)
{
    const int64_t total = (int64_t)n1 * n2 * n3;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int s1 = (int)(s % n1), s2 = (int)((s / n1) % n2), s3 = (int)(s / ((int64_t)n1 * n2));
        // the centred offset of FFT index s is its signed frequency; box index = k + floor(d/2)
        const int a = signed_freq(s1, n1) + d1 / 2;
        const int b = signed_freq(s2, n2) + d2 / 2;
        const int c = signed_freq(s3, n3) + d3 / 2;
        T v = T(0);
        if (a >= 0 && a < d1 && b >= 0 && b < d2 && c >= 0 && c < d3)
            v = src[(int64_t)a + (int64_t)d1 * (b + (int64_t)d2 * c)];
        dst[s] = v;
    }
}

// The same pad for bytes (the support mask): one workgroup per (row segment of
// 1024, y, z), four consecutive x per thread written as one word, no division.
__global__ __launch_bounds__(256) void k_pad_u8(const uint8_t* __restrict__ src, int d1, int d2, int d3,
                                                uint8_t* __restrict__ dst, int n1, int n2, int n3)
{
    const int s2 = blockIdx.y, s3 = blockIdx.z, s1 = (blockIdx.x * 256 + threadIdx.x) * 4;
    if (s1 >= n1) return;
    const int b = signed_freq(s2, n2) + d2 / 2, c = signed_freq(s3, n3) + d3 / 2;
    const bool bc = b >= 0 && b < d2 && c >= 0 && c < d3;
    const uint8_t* row = bc ? src + (int64_t)d1 * (b + (int64_t)d2 * c) : src;
    uint8_t v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = s1 + q < n1 ? signed_freq(s1 + q, n1) + d1 / 2 : -1;
        v[q] = (bc && a >= 0 && a < d1) ? row[a] : (uint8_t)0;
    }
    uint8_t* out = dst + s1 + (int64_t)n1 * (s2 + (int64_t)n2 * s3);
    if (s1 + 4 <= n1 && (n1 & 3) == 0) {
        *reinterpret_cast<uint32_t*>(out) = (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) |
                                            ((uint32_t)v[3] << 24);
    } else {
        for (int q = 0; q < 4 && s1 + q < n1; ++q) out[q] = v[q];
    }
}

__global__ void k_expand_f16(const uint16_t* __restrict__ h, const float* __restrict__ scale, int64_t per_proj,
                             int64_t total, float* __restrict__ out)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x)
        out[t] = (float)__builtin_bit_cast(_Float16, h[t]) * scale[t / per_proj];
}

int grid_blocks(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    return (int)(b < 65536 * 4 ? (b < 1 ? 1 : b) : 65536 * 4);
}

}  // namespace

void grid_values_fp32(const float* d_proj, int dim1, int dim2, const float* d_un, const float* d_vn,
                      const PointTile* d_tiles, int ntiles, float2* d_F, hipStream_t st)
{
    if (ntiles <= 0) return;
    const size_t lds = (size_t)((dim1 + 15) / 16) * 16 * kCB * sizeof(float);
    GF_CHECK(lds <= 64 * 1024, GENFIRE_ERR_UNSUPPORTED, "projection Dim1 too large for the DFT kernel");
    k_dft_fp32<<<ntiles, kVT, lds, st>>>(d_proj, dim1, dim2, d_un, d_vn, d_tiles, d_F);
    GF_LAUNCH_CHECK();
}

void grid_make_tiles(const int64_t* d_proj_off, const int64_t* d_tile_off, int P, int tile_pts, PointTile* d_tiles,
                     int ntiles, hipStream_t st)
{
    if (ntiles <= 0) return;
    k_make_tiles<<<(ntiles + 255) / 256, 256, 0, st>>>(d_proj_off, d_tile_off, P, tile_pts, d_tiles, ntiles);
    GF_LAUNCH_CHECK();
}

void grid_scatter(const uint32_t* d_ind, const float* d_dist, const float2* d_F, int64_t begin,
                  int64_t count, double2* d_acc, double* d_sw, uint32_t* d_hits, float2* d_exact,
                  hipStream_t st)
{
    if (count <= 0) return;
    const int threads = 256;
    const int blocks = (int)((count + threads - 1) / threads);
    k_scatter<<<blocks, threads, 0, st>>>(d_ind, d_dist, d_F, begin, count, d_acc, d_sw, d_hits, d_exact);
    GF_LAUNCH_CHECK();
}

void grid_finalize(const HalfGrid& hg, int n1, int n2, int centro, const double2* d_acc,
                   const double* d_sw, const uint32_t* d_hits, const float2* d_exact, float2* d_M,
                   hipStream_t st)
{
    GF_CHECK(n2 <= 65535 * kFinRows && n1 <= 65535, GENFIRE_ERR_UNSUPPORTED, "grid edge above 65535");
    const dim3 grid((n2 + kFinRows - 1) / kFinRows, n1);
    k_finalize<<<grid, 256, 0, st>>>(hg, n1, n2, centro, d_acc, d_sw, d_hits, d_exact, d_M);
    GF_LAUNCH_CHECK();
}

void expand_f16(const uint16_t* d_h, const float* d_scale, int64_t per_proj, int P, float* d_out, hipStream_t st)
{
    const int64_t total = per_proj * P;
    k_expand_f16<<<grid_blocks(total, 256), 256, 0, st>>>(d_h, d_scale, per_proj, total, d_out);
    GF_LAUNCH_CHECK();
}

void shift_complex(const float2* d_in, float2* d_out, int n1, int n2, int n3, int to_centred,
                   hipStream_t st)
{
    const int64_t total = (int64_t)n1 * n2 * n3;
    k_shift<<<grid_blocks(total, 256), 256, 0, st>>>(d_in, d_out, n1, n2, n3, to_centred);
    GF_LAUNCH_CHECK();
}

void crop_centre(const void* d_src, int src_complex, int n1, int n2, int n3, float* d_dst, int d1,
                 int d2, int d3, hipStream_t st)
{
    const int64_t total = (int64_t)d1 * d2 * d3;
    k_crop<<<grid_blocks(total, 256), 256, 0, st>>>(d_src, src_complex, n1, n2, n3, d_dst, d1, d2, d3);
    GF_LAUNCH_CHECK();
}

void pad_centre(const float* d_src, int d1, int d2, int d3, float* d_dst, int n1, int n2, int n3,
                hipStream_t st)
{
    const int64_t total = (int64_t)n1 * n2 * n3;
    k_pad<float><<<grid_blocks(total, 256), 256, 0, st>>>(d_src, d1, d2, d3, d_dst, n1, n2, n3);
    GF_LAUNCH_CHECK();
}

void pad_centre_u8(const uint8_t* d_src, int d1, int d2, int d3, uint8_t* d_dst, int n1, int n2,
                   int n3, hipStream_t st)
{
    GF_CHECK(n2 <= 65535 && n3 <= 65535, GENFIRE_ERR_UNSUPPORTED, "grid edge above 65535");
    const dim3 grid((n1 + 1023) / 1024, n2, n3);
    k_pad_u8<<<grid, 256, 0, st>>>(d_src, d1, d2, d3, d_dst, n1, n2, n3);
    GF_LAUNCH_CHECK();
}


// ===========================================================================
// MFMA form of the DFT values (fillInFourierGrid_DFT.m:243 as a complex GEMM).
// G[j, b] = sum_a E1[j, a] P[a, b] is a real GEMM with the E1 real and imaginary
// parts as two A operands sharing the B operand P; F_j = sum_b E2[j, b] G[j, b].
// A wave owns 16 points (the 16 rows of v_mfma_f32_16x16x4f32); lane l supplies
// A[l%16][l/16] = E1(point l%16, a) built from three LDS tables of FP64-reduced
// phases, a = 128i + 16j + k: E1 = E1hh(i) * E1hm(j) * E1lo(k), the first two
// multiplied once per 16-row step; B[l/16][l%16] = P[a][b] from LDS; and it
// holds G[4(l/16)+r][16nt + l%16].  The E2 contraction runs on the accumulators
// (E2 split the same way, b = 128i + 16j + k) and ends in a 16-lane reduction.
// P is pre-transposed to row-major [a][b] (zero-padded) once per job.
// The three-level tables keep a workgroup's LDS small enough for three
// workgroups per CU up to 512 x 512 projections, and the register budget is
// held to three waves per SIMD.
// ===========================================================================
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kMP = 64;    // points per workgroup (4 waves x 16)
constexpr int kKC = 16;    // projection rows staged per step (two buffers)

__global__ void k_transpose_proj(const float* __restrict__ proj, int dim1, int dim2, int P, int Kp, int Np,
                                 float* __restrict__ Pt)
{
    const int64_t total = (int64_t)P * Kp * Np;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int b = (int)(t % Np), a = (int)((t / Np) % Kp), p = (int)(t / ((int64_t)Np * Kp));
        Pt[t] = (a < dim1 && b < dim2) ? proj[a + (int64_t)dim1 * (b + (int64_t)dim2 * p)] : 0.0f;
    }
}

// Rows of the three-level phase tables for H = Kp/16 (or Np/16) 16-row blocks.
__host__ __device__ constexpr int hh_rows(int H) { return (H + 7) / 8; }

// The six phase tables of a 64-point tile, contiguous from t1hh:
// [R1] E1(128 i - c1), [8] E1(16 j), [16] E1(k), then the same three for E2,
// each row 64 points wide.
__device__ __forceinline__ void fill_phase_tables(float2* t1hh, int R1, int R2, int c1, int c2,
                                                  const float* __restrict__ un, const float* __restrict__ vn,
                                                  const PointTile& tl, int tid)
{
    const int rows1 = R1 + 24;
    for (int idx = tid; idx < kMP * (rows1 + R2 + 24); idx += 256) {
        const int pt = idx % kMP, row = idx / kMP;
        const bool ok = pt < tl.count;
        const bool first = row < rows1;
        const float t = ok ? (first ? un[tl.begin + pt] : vn[tl.begin + pt]) : 0.0f;
        const int r = first ? row : row - rows1, R = first ? R1 : R2, c = first ? c1 : c2;
        const int k = r < R ? 128 * r - c : (r < R + 8 ? 16 * (r - R) : r - R - 8);
        t1hh[(first ? 0 : rows1 * kMP) + r * kMP + pt] = cis_turns((double)k, t);
    }
}

template <int NCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void k_dft_mfma(const float* __restrict__ Pt, int Kp, int Np, int dim1, int dim2,
                const float* __restrict__ un, const float* __restrict__ vn,
                const PointTile* __restrict__ tiles, float2* __restrict__ F)
{
    constexpr int STR = NCH + 16;           // LDS row stride: 4 rows of 16 lanes on distinct banks
    constexpr int NTT = NCH / 16;
    static_assert(NTT <= 8 && 8 % NTT == 0, "a column chunk stays inside one 128-column block");
    extern __shared__ float smem[];
    const int H1 = Kp / 16, H2 = Np / 16;
    const int R1 = hh_rows(H1), R2 = hh_rows(H2);
    float* Pl = smem;                                             // 2 x kKC x STR
    float2* t1hh = reinterpret_cast<float2*>(Pl + 2 * kKC * STR); // [R1][64]: a = 128i
    float2* t1hm = t1hh + R1 * kMP;                               // [8][64]:  16j
    float2* t1lo = t1hm + 8 * kMP;                                // [16][64]: k
    float2* t2hh = t1lo + 16 * kMP;                               // [R2][64]
    float2* t2hm = t2hh + R2 * kMP;                               // [8][64]
    float2* t2lo = t2hm + 8 * kMP;                                // [16][64]
    const PointTile tl = tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const int c1 = dim1 / 2, c2 = dim2 / 2;
    fill_phase_tables(t1hh, R1, R2, c1, c2, un, vn, tl, tid);
    const float* Pp = Pt + (int64_t)tl.proj * Kp * Np;
    const int w = tid / 64, l = tid % 64, li = l % 16, lk = l / 16;
    const int pa = 16 * w + li;
    float Fr[4] = {0.0f, 0.0f, 0.0f, 0.0f}, Fi[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // Projection rows are staged through two LDS buffers: the next chunk's
    // global loads are in flight while the MFMAs consume the current one.
    constexpr int SPT = kKC * NCH / 256;   // staged values per thread and chunk
    static_assert(kKC * NCH % 256 == 0, "staging split");
    const int nkc = Kp / kKC;              // Kp is a multiple of kKC
    for (int nc = 0; nc < Np; nc += NCH) {
        v4f acr[NTT], aci[NTT];
#pragma unroll
        for (int nt = 0; nt < NTT; ++nt) { acr[nt] = v4f{0, 0, 0, 0}; aci[nt] = v4f{0, 0, 0, 0}; }
        float pre[SPT];
        auto stage_load = [&](int kc) {
#pragma unroll
            for (int q = 0; q < SPT; ++q) {
                const int idx = tid + 256 * q, k = idx / NCH, c = idx % NCH;
                pre[q] = (nc + c < Np) ? Pp[(int64_t)(kc + k) * Np + nc + c] : 0.0f;
            }
        };
        stage_load(0);
        for (int ic = 0; ic < nkc; ++ic) {   // rows 16 ic .. 16 ic + 15 (kKC = 16)
            float* Pb = Pl + (ic & 1) * (kKC * STR);
#pragma unroll
            for (int q = 0; q < SPT; ++q) {
                const int idx = tid + 256 * q, k = idx / NCH, c = idx % NCH;
                Pb[k * STR + c] = pre[q];
            }
            __syncthreads();
            if (ic + 1 < nkc) stage_load((ic + 1) * kKC);
            const float2 e1h = cmul(t1hh[(ic >> 3) * kMP + pa], t1hm[(ic & 7) * kMP + pa]);
#pragma unroll
            for (int kk = 0; kk < kKC / 4; ++kk) {
                const float2 e1 = cmul(e1h, t1lo[(4 * kk + lk) * kMP + pa]);
                const float* prow = Pb + (4 * kk + lk) * STR + li;
#pragma unroll
                for (int nt = 0; nt < NTT; ++nt) {
                    const float b = prow[16 * nt];
                    acr[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(e1.x, b, acr[nt], 0, 0, 0);
                    aci[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(e1.y, b, aci[nt], 0, 0, 0);
                }
            }
        }
        __syncthreads();   // the buffers are refilled by the next column chunk
        // F += sum_b E2(pt, b) G(pt, b): lane holds columns nc + 16nt + li of points 16w + 4lk + r
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int pt = 16 * w + 4 * lk + r;
            float hr = 0.0f, hi = 0.0f;
#pragma unroll
            for (int nt = 0; nt < NTT; ++nt) {
                const int m = nc / 16 + nt;
                if (m < H2) {
                    const float2 e = t2hm[(m & 7) * kMP + pt];
                    const float gr = acr[nt][r], gi = aci[nt][r];
                    hr += gr * e.x - gi * e.y;
                    hi += gr * e.y + gi * e.x;
                }
            }
            const float2 el = cmul(t2hh[(nc >> 7) * kMP + pt], t2lo[li * kMP + pt]);
            float fr = hr * el.x - hi * el.y, fi = hr * el.y + hi * el.x;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                fr += __shfl_xor(fr, off, 64);
                fi += __shfl_xor(fi, off, 64);
            }
            Fr[r] += fr;
            Fi[r] += fi;
        }
    }
    if (li == 0)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int pt = 16 * w + 4 * lk + r;
            if (pt < tl.count) F[tl.begin + pt] = make_float2(Fr[r], Fi[r]);
        }
}

}  // namespace

int mfma_tile_points() { return kMP; }

void grid_prep_proj(const float* d_proj, int dim1, int dim2, int P, float* d_Pt, hipStream_t st)
{
    const int Kp = (dim1 + kKC - 1) / kKC * kKC, Np = (dim2 + 15) / 16 * 16;
    const int64_t total = (int64_t)P * Kp * Np;
    k_transpose_proj<<<grid_blocks(total, 256), 256, 0, st>>>(d_proj, dim1, dim2, P, Kp, Np, d_Pt);
    GF_LAUNCH_CHECK();
}

size_t grid_prep_floats(int dim1, int dim2, int P)
{
    return (size_t)P * ((dim1 + kKC - 1) / kKC * kKC) * ((dim2 + 15) / 16 * 16);
}

void grid_values_mfma_fp32(const float* d_Pt, int dim1, int dim2, const float* d_un, const float* d_vn,
                           const PointTile* d_tiles, int ntiles, float2* d_F, hipStream_t st)
{
    if (ntiles <= 0) return;
    const int Kp = (dim1 + kKC - 1) / kKC * kKC, Np = (dim2 + 15) / 16 * 16;
    const int H1 = Kp / 16, H2 = Np / 16;
    auto lds = [&](int nch) {
        return sizeof(float) * 2 * kKC * (nch + 16) + sizeof(float2) * kMP * (hh_rows(H1) + hh_rows(H2) + 48);
    };
    if (Np >= 128) {
        const size_t b = lds(128);
        GF_CHECK(b <= 160 * 1024, GENFIRE_ERR_UNSUPPORTED, "projection too large for the MFMA DFT kernel");
        if (b > 65536)
            GF_HIP(hipFuncSetAttribute((const void*)k_dft_mfma<128>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b));
        k_dft_mfma<128><<<ntiles, 256, b, st>>>(d_Pt, Kp, Np, dim1, dim2, d_un, d_vn, d_tiles, d_F);
    } else if (Np >= 64) {
        k_dft_mfma<64><<<ntiles, 256, lds(64), st>>>(d_Pt, Kp, Np, dim1, dim2, d_un, d_vn, d_tiles, d_F);
    } else {
        k_dft_mfma<16><<<ntiles, 256, lds(16), st>>>(d_Pt, Kp, Np, dim1, dim2, d_un, d_vn, d_tiles, d_F);
    }
    GF_LAUNCH_CHECK();
}

// ===========================================================================
// Reduced-precision MFMA forms of the same GEMM (v_mfma_f32_16x16x32_{bf16,f16},
// FP32 accumulation).  Lane l supplies A[l%16][8(l/16)+q] = E1(point l%16,
// a0+q) for q = 0..7 and B[8(l/16)+q][l%16] = P[a0+q][b]: eight consecutive
// projection rows, which the split kernel stores row-contiguous per column
// ([p][b][a], a padded to a multiple of 64) so a lane reads them as one
// 16-byte LDS access.
//   kSplitBf16: x = hi + lo with hi = x rounded to bf16 and lo = the exact
//     residual rounded to bf16 (|x - hi - lo| <= 2^-17 |x|);
//     A.B = Ah.Bh + Ah.Bl + Al.Bh (the dropped Al.Bl is below 2^-16 |A||B|).
//   kFp16: P / s_p in FP16 with s_p = max|P| of projection p, E1 in FP16,
//     one product; the sum is multiplied by s_p.
// E1 = E1hi(a >> 4) * E1lo(a & 15): a lane's eight rows share one E1hi entry
// and always use the same eight E1lo entries, which it keeps in registers.
// ===========================================================================
namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef unsigned short v8u __attribute__((ext_vector_type(8)));

constexpr int kLpKC = 64;            // projection rows per LDS stage (two k-steps of 32)
constexpr int kLpSK = kLpKC + 8;     // LDS column stride in 16-bit elements (144 bytes)

// hi = x rounded to nearest bf16, lo = the exact residual x - hi rounded to
// nearest bf16: |lo| <= 2^-8 |x| and |x - hi - lo| <= 2^-17 |x|
__device__ __forceinline__ void split_bf16(float x, unsigned short& hi, unsigned short& lo)
{
    const __bf16 h = (__bf16)x;
    hi = __builtin_bit_cast(unsigned short, h);
    lo = __builtin_bit_cast(unsigned short, (__bf16)(x - (float)h));
}

__global__ void k_proj_maxabs(const float* __restrict__ proj, int64_t n, float* __restrict__ scale)
{
    const float* src = proj + n * blockIdx.x;
    float m = 0.0f;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) m = fmaxf(m, fabsf(src[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
    __shared__ float red[16];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, red[w]);
        scale[blockIdx.x] = m > 0.0f ? m : 1.0f;
    }
}

template <int MODE>
__global__ void k_split_proj(const float* __restrict__ proj, int dim1, int dim2, int P, int Kp, int Np,
                             const float* __restrict__ scale, unsigned short* __restrict__ hi,
                             unsigned short* __restrict__ lo)
{
    const int64_t total = (int64_t)P * Np * Kp;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(t % Kp), b = (int)((t / Kp) % Np), p = (int)(t / ((int64_t)Kp * Np));
        const float x = (a < dim1 && b < dim2) ? proj[a + (int64_t)dim1 * (b + (int64_t)dim2 * p)] : 0.0f;
        if (MODE == GENFIRE_PREC_SPLIT_BF16) {
            unsigned short h, l;
            split_bf16(x, h, l);
            hi[t] = h;
            lo[t] = l;
        } else {
            hi[t] = __builtin_bit_cast(unsigned short, (_Float16)(x / scale[p]));
        }
    }
}

template <int MODE, int NCH>
__global__ __launch_bounds__(256) void k_dft_mfma_lp(const unsigned short* __restrict__ Phi,
                                                     const unsigned short* __restrict__ Plo,
                                                     const float* __restrict__ scale, int Kp, int Np, int dim1,
                                                     int dim2, const float* __restrict__ un,
                                                     const float* __restrict__ vn, const PointTile* __restrict__ tiles,
                                                     float2* __restrict__ F)
{
    constexpr bool SPLIT = MODE == GENFIRE_PREC_SPLIT_BF16;
    constexpr int NTT = NCH / 16;
    constexpr int NB = SPLIT ? 2 : 1;       // staged operand arrays (hi, lo)
    static_assert(NTT <= 8 && 8 % NTT == 0, "a column chunk stays inside one 128-column block");
    extern __shared__ float4 smem4[];
    unsigned short* Bl = reinterpret_cast<unsigned short*>(smem4);   // [NB][NCH][kLpSK]
    const int H1 = Kp / 16, H2 = Np / 16;
    const int R1 = hh_rows(H1), R2 = hh_rows(H2);
    float2* t1hh = reinterpret_cast<float2*>(Bl + NB * NCH * kLpSK);  // [R1][64]: a = 128i
    float2* t1hm = t1hh + R1 * kMP;                                   // [8][64]:  16j
    float2* t1lo = t1hm + 8 * kMP;                                    // [16][64]: k
    float2* t2hh = t1lo + 16 * kMP;                                   // [R2][64]
    float2* t2hm = t2hh + R2 * kMP;                                   // [8][64]
    float2* t2lo = t2hm + 8 * kMP;                                    // [16][64]
    const PointTile tl = tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const int c1 = dim1 / 2, c2 = dim2 / 2;
    fill_phase_tables(t1hh, R1, R2, c1, c2, un, vn, tl, tid);
    __syncthreads();
    const int w = tid / 64, l = tid % 64, li = l % 16, lk = l / 16;
    const int pa = 16 * w + li;
    float2 elo[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) elo[q] = t1lo[(8 * (lk & 1) + q) * kMP + pa];
    const int64_t pbase = (int64_t)tl.proj * Np * Kp;
    float Fr[4] = {0.0f, 0.0f, 0.0f, 0.0f}, Fi[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int nc = 0; nc < Np; nc += NCH) {
        v4f acr[NTT], aci[NTT];
#pragma unroll
        for (int nt = 0; nt < NTT; ++nt) { acr[nt] = v4f{0, 0, 0, 0}; aci[nt] = v4f{0, 0, 0, 0}; }
        for (int kc = 0; kc < Kp; kc += kLpKC) {
            __syncthreads();
            // stage columns nc..nc+NCH-1, rows kc..kc+63 (16-byte units)
            for (int idx = tid; idx < NB * NCH * (kLpKC / 8); idx += 256) {
                const int u = idx % (kLpKC / 8), c = (idx / (kLpKC / 8)) % NCH, h = idx / (NCH * (kLpKC / 8));
                const unsigned short* src = h ? Plo : Phi;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (nc + c < Np) v = *reinterpret_cast<const uint4*>(src + pbase + (int64_t)(nc + c) * Kp + kc + 8 * u);
                *reinterpret_cast<uint4*>(Bl + (h * NCH + c) * kLpSK + 8 * u) = v;
            }
            __syncthreads();
#pragma unroll
            for (int ks = 0; ks < kLpKC / 32; ++ks) {
                const int a0 = kc + 32 * ks + 8 * lk;          // this lane's rows a0..a0+7
                const float2 ehi = cmul(t1hh[(a0 >> 7) * kMP + pa], t1hm[((a0 >> 4) & 7) * kMP + pa]);
                v8u are, aim, are2, aim2;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const float2 e = cmul(ehi, elo[q]);
                    if (SPLIT) {
                        unsigned short h, lo;
                        split_bf16(e.x, h, lo);
                        are[q] = h;
                        are2[q] = lo;
                        split_bf16(e.y, h, lo);
                        aim[q] = h;
                        aim2[q] = lo;
                    } else {
                        are[q] = __builtin_bit_cast(unsigned short, (_Float16)e.x);
                        aim[q] = __builtin_bit_cast(unsigned short, (_Float16)e.y);
                    }
                }
#pragma unroll
                for (int nt = 0; nt < NTT; ++nt) {
                    const int col = 16 * nt + li, koff = 32 * ks + 8 * lk;
                    const v8u bh = *reinterpret_cast<const v8u*>(Bl + col * kLpSK + koff);
                    if (SPLIT) {
                        const v8u bl = *reinterpret_cast<const v8u*>(Bl + (NCH + col) * kLpSK + koff);
                        const v8bf Ah = __builtin_bit_cast(v8bf, are), Al = __builtin_bit_cast(v8bf, are2);
                        const v8bf Ih = __builtin_bit_cast(v8bf, aim), Il = __builtin_bit_cast(v8bf, aim2);
                        const v8bf Bh = __builtin_bit_cast(v8bf, bh), Bo = __builtin_bit_cast(v8bf, bl);
                        acr[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, acr[nt], 0, 0, 0);
                        acr[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bo, acr[nt], 0, 0, 0);
                        acr[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, acr[nt], 0, 0, 0);
                        aci[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Il, Bh, aci[nt], 0, 0, 0);
                        aci[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ih, Bo, aci[nt], 0, 0, 0);
                        aci[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ih, Bh, aci[nt], 0, 0, 0);
                    } else {
                        const v8h B = __builtin_bit_cast(v8h, bh);
                        acr[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8h, are), B, acr[nt], 0, 0, 0);
                        aci[nt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8h, aim), B, aci[nt], 0, 0, 0);
                    }
                }
            }
        }
        // F += sum_b E2(pt, b) G(pt, b), as in the FP32 form
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int pt = 16 * w + 4 * lk + r;
            float hr = 0.0f, hi = 0.0f;
#pragma unroll
            for (int nt = 0; nt < NTT; ++nt) {
                const int m = nc / 16 + nt;
                if (m < H2) {
                    const float2 e = t2hm[(m & 7) * kMP + pt];
                    const float gr = acr[nt][r], gi = aci[nt][r];
                    hr += gr * e.x - gi * e.y;
                    hi += gr * e.y + gi * e.x;
                }
            }
            const float2 el = cmul(t2hh[(nc >> 7) * kMP + pt], t2lo[li * kMP + pt]);
            float fr = hr * el.x - hi * el.y, fi = hr * el.y + hi * el.x;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                fr += __shfl_xor(fr, off, 64);
                fi += __shfl_xor(fi, off, 64);
            }
            Fr[r] += fr;
            Fi[r] += fi;
        }
    }
    const float sc = SPLIT ? 1.0f : scale[tl.proj];
    if (li == 0)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int pt = 16 * w + 4 * lk + r;
            if (pt < tl.count) F[tl.begin + pt] = make_float2(Fr[r] * sc, Fi[r] * sc);
        }
}

int lp_rows(int dim1) { return (dim1 + kLpKC - 1) / kLpKC * kLpKC; }
int lp_cols(int dim2) { return (dim2 + 15) / 16 * 16; }

template <int MODE, int NCH> size_t lp_lds(int Kp, int Np)
{
    constexpr int NB = MODE == GENFIRE_PREC_SPLIT_BF16 ? 2 : 1;
    return sizeof(unsigned short) * NB * NCH * kLpSK + sizeof(float2) * kMP * (hh_rows(Kp / 16) + hh_rows(Np / 16) + 48);
}

template <int MODE, int NCH>
void launch_lp(const unsigned short* hi, const unsigned short* lo, const float* scale, int dim1, int dim2,
               const float* un, const float* vn, const PointTile* tiles, int ntiles, float2* F, hipStream_t st)
{
    const int Kp = lp_rows(dim1), Np = lp_cols(dim2);
    const size_t b = lp_lds<MODE, NCH>(Kp, Np);
    GF_CHECK(b <= 160 * 1024, GENFIRE_ERR_UNSUPPORTED, "projection too large for the MFMA DFT kernel");
    if (b > 65536)
        GF_HIP(hipFuncSetAttribute((const void*)k_dft_mfma_lp<MODE, NCH>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)b));
    k_dft_mfma_lp<MODE, NCH><<<ntiles, 256, b, st>>>(hi, lo, scale, Kp, Np, dim1, dim2, un, vn, tiles, F);
    GF_LAUNCH_CHECK();
}

template <int MODE>
void launch_lp_n(const unsigned short* hi, const unsigned short* lo, const float* scale, int dim1, int dim2,
                 const float* un, const float* vn, const PointTile* tiles, int ntiles, float2* F, hipStream_t st)
{
    // 64-column chunks: twice the waves per SIMD of 128-column ones (fewer
    // accumulators), which measured 6% faster at 256^2 projections
    const int Np = lp_cols(dim2);
    if (Np >= 64) launch_lp<MODE, 64>(hi, lo, scale, dim1, dim2, un, vn, tiles, ntiles, F, st);
    else launch_lp<MODE, 16>(hi, lo, scale, dim1, dim2, un, vn, tiles, ntiles, F, st);
}

}  // namespace

size_t grid_prep_lp_elems(int dim1, int dim2, int P) { return (size_t)P * lp_rows(dim1) * lp_cols(dim2); }

void grid_prep_lp(int mode, const float* d_proj, int dim1, int dim2, int P, unsigned short* d_hi,
                  unsigned short* d_lo, float* d_scale, hipStream_t st)
{
    const int Kp = lp_rows(dim1), Np = lp_cols(dim2);
    const int64_t total = (int64_t)P * Kp * Np;
    if (mode == GENFIRE_PREC_SPLIT_BF16) {
        k_split_proj<GENFIRE_PREC_SPLIT_BF16><<<grid_blocks(total, 256), 256, 0, st>>>(d_proj, dim1, dim2, P, Kp, Np,
                                                                                     d_scale, d_hi, d_lo);
    } else {
        GF_CHECK(mode == GENFIRE_PREC_FP16, GENFIRE_ERR_ARG, "unknown gridding precision");
        k_proj_maxabs<<<P, 1024, 0, st>>>(d_proj, (int64_t)dim1 * dim2, d_scale);
        GF_LAUNCH_CHECK();
        k_split_proj<GENFIRE_PREC_FP16><<<grid_blocks(total, 256), 256, 0, st>>>(d_proj, dim1, dim2, P, Kp, Np,
                                                                               d_scale, d_hi, d_lo);
    }
    GF_LAUNCH_CHECK();
}

void grid_values_mfma_lp(int mode, const unsigned short* d_hi, const unsigned short* d_lo, const float* d_scale,
                         int dim1, int dim2, const float* d_un, const float* d_vn, const PointTile* d_tiles,
                         int ntiles, float2* d_F, hipStream_t st)
{
    if (ntiles <= 0) return;
    if (mode == GENFIRE_PREC_SPLIT_BF16)
        launch_lp_n<GENFIRE_PREC_SPLIT_BF16>(d_hi, d_lo, d_scale, dim1, dim2, d_un, d_vn, d_tiles, ntiles, d_F, st);
    else
        launch_lp_n<GENFIRE_PREC_FP16>(d_hi, d_lo, d_scale, dim1, dim2, d_un, d_vn, d_tiles, ntiles, d_F, st);
}

}  // namespace gf
