// gf_grid_index.hip -- gridding index pass of fillInFourierGrid_DFT.m:175-192.
//
// The reference scans the whole half grid once per projection and keeps the
// points with D = |K.nh| < cutoff whose in-plane coordinates pass
// |u| <= n1/2, |v| <= n2/2.  Here each projection enumerates only the thin slab
// around its plane: along the dominant axis a of the unit normal, every column
// (k_b, k_c) of the other two axes holds the candidates
//     |s + nh_a k_a| < cutoff,  s = nh_b k_b + nh_c k_c,
// an interval of about 2*cutoff/|nh_a| + 1 <= 3.5 values, widened by a margin
// (2e-3 in D) larger than the FP32 rounding of D for |k| < 4096.  Each
// candidate then runs the exact FP32 test, so the kept set is bit-identical to
// the full scan.  This file is compiled with -ffp-contract=off: every FP32
// expression rounds once per operation, in the reference's order.
//
// Three launches per gridding: count (per column-block totals), one-workgroup
// exclusive scan, emit (block-local scan + global offset).  Output order is
// (projection, column block, column, k_a ascending) -- deterministic.
#pragma clang fp contract(off)

#include "gf_grid.h"

namespace gf {

namespace {

constexpr int kThreads = 256;
constexpr double kMarginD = 2e-3;

struct IndexArgs {
    const PlaneConst* planes;
    HalfGrid hg;
    double cutoff;
    float h1, h2;  // n1/2, n2/2 as single (exact for n < 2^24)
    int nblk;      // column blocks per projection (max over projections)
};

// D, closest point and in-plane test for one grid point, in the reference's
// FP32 operation order (distancePointsPlane_YY :333, closestpoint :347-348,
// mldivide :182 as LU solve, :190).
__device__ __forceinline__ bool point_test(const PlaneConst& pc, float K1, float K2, float K3,
                                           double cutoff, float h1, float h2, float* D, float* u,
                                           float* v)
{
#pragma clang fp contract(off)
    const float t = (K1 * pc.nh[0] + K2 * pc.nh[1]) + K3 * pc.nh[2];
    const float d = fabsf(t);
    if (!((double)d < cutoff)) return false;
    const float s = (K1 * pc.nv[0] + K2 * pc.nv[1]) + K3 * pc.nv[2];
    const float vv = (0.0f - s) / pc.nn;
    float b[3];
    b[0] = K1 + vv * pc.nv[0];
    b[1] = K2 + vv * pc.nv[1];
    b[2] = K3 + vv * pc.nv[2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int p = pc.piv[j];
        if (p != j) {
            const float tmp = b[j];
            b[j] = (p == 1) ? b[1] : b[2];
            if (p == 1) b[1] = tmp; else b[2] = tmp;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (b[k] != 0.0f)
#pragma unroll
            for (int i = k + 1; i < 3; ++i) b[i] = b[i] - b[k] * pc.lu[i + 3 * k];
#pragma unroll
    for (int k = 2; k >= 0; --k)
        if (b[k] != 0.0f) {
            b[k] = b[k] / pc.lu[k + 3 * k];
#pragma unroll
            for (int i = 0; i < k; ++i) b[i] = b[i] - b[k] * pc.lu[i + 3 * k];
        }
    if (fabsf(b[0]) <= h1 && fabsf(b[1]) <= h2) {
        *D = d;
        *u = b[0];
        *v = b[1];
        return true;
    }
    return false;
}

struct Column {
    int a, b, c;       // dominant axis and the two column axes (b < c)
    int kb, kc;        // column coordinates
    int ka_lo, ka_hi;  // candidate interval along a (inclusive, clipped)
};

__device__ __forceinline__ bool make_column(const PlaneConst& pc, const HalfGrid& hg, double cutoff,
                                            int64_t col, Column* out)
{
    const int a = pc.axis;
    const int b = (a == 0) ? 1 : 0;
    const int c = (a == 2) ? 1 : 2;
    const int64_t ncol = (int64_t)hg.len[b] * hg.len[c];
    if (col >= ncol) return false;
    out->a = a; out->b = b; out->c = c;
    out->kb = hg.lo[b] + (int)(col % hg.len[b]);
    out->kc = hg.lo[c] + (int)(col / hg.len[b]);
    const double na = pc.nhd[a];
    const double s = pc.nhd[b] * out->kb + pc.nhd[c] * out->kc;
    const double e1 = (-cutoff - kMarginD - s) / na;
    const double e2 = (cutoff + kMarginD - s) / na;
    const double lo = fmin(e1, e2), hi = fmax(e1, e2);
    int ilo = (int)ceil(lo), ihi = (int)floor(hi);
    ilo = max(ilo, hg.lo[a]);
    ihi = min(ihi, hg.lo[a] + hg.len[a] - 1);
    out->ka_lo = ilo;
    out->ka_hi = ihi;
    return true;
}

__device__ __forceinline__ void coords(const Column& col, int ka, float K[3])
{
    // selects, not a store at a runtime index (which puts the array in scratch)
#pragma unroll
    for (int q = 0; q < 3; ++q) K[q] = (float)(col.a == q ? ka : (col.b == q ? col.kb : col.kc));
}

__global__ __launch_bounds__(kThreads) void k_index_count(IndexArgs args, uint32_t* blk_counts)
{
    const int p = blockIdx.y;
    const PlaneConst& pc = args.planes[p];   // read in place: a private copy indexed by axis sits in scratch
    const int64_t col = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    uint32_t cnt = 0;
    Column cl;
    if (make_column(pc, args.hg, args.cutoff, col, &cl)) {
        for (int ka = cl.ka_lo; ka <= cl.ka_hi; ++ka) {
            float K[3], D, u, v;
            coords(cl, ka, K);
            cnt += point_test(pc, K[0], K[1], K[2], args.cutoff, args.h1, args.h2, &D, &u, &v) ? 1u : 0u;
        }
    }
    __shared__ uint32_t red[kThreads / 64];
    // wave reduction then across waves
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kThreads / 64; ++w) t += red[w];
        blk_counts[(int64_t)p * args.nblk + blockIdx.x] = t;
    }
}

// Exclusive scan of n counts (int64 offsets) in three launches: each 1024-thread
// workgroup sums a chunk of kScanChunk counts, one workgroup scans the chunk
// sums, each workgroup rescans its chunk from its carry.  Integer sums, so the
// result does not depend on the split.
constexpr int kScanPer = 8;                       // counts per thread per chunk
constexpr int kScanChunk = 1024 * kScanPer;

// inclusive scan of one value per thread over the 1024-thread workgroup
__device__ __forceinline__ int64_t block_scan_incl(int64_t v, int64_t* wsum)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const int64_t u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    int64_t pre = 0;
    for (int q = 0; q < w; ++q) pre += wsum[q];
    __syncthreads();
    return v + pre;
}

__global__ __launch_bounds__(1024) void k_scan_sums(const uint32_t* __restrict__ counts, int64_t n,
                                                    int64_t* __restrict__ csum)
{
    __shared__ int64_t wsum[16];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
    int64_t v = 0;
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) v += base + q < n ? counts[base + q] : 0u;
    const int64_t inc = block_scan_incl(v, wsum);
    if (threadIdx.x == 1023) csum[blockIdx.x] = inc;
}

__global__ __launch_bounds__(1024) void k_scan_carry(int64_t* __restrict__ csum, int nchunk, int64_t* total)
{
    __shared__ int64_t wsum[16];
    int64_t run = 0;
    for (int b0 = 0; b0 < nchunk; b0 += 1024) {
        const int b = b0 + (int)threadIdx.x;
        const int64_t v = b < nchunk ? csum[b] : 0;
        const int64_t inc = block_scan_incl(v, wsum);
        if (b < nchunk) csum[b] = run + inc - v;   // exclusive carry of chunk b
        int64_t all = 0;                             // this round's total: the wave sums
        for (int q = 0; q < 16; ++q) all += wsum[q];
        __syncthreads();
        run += all;
    }
    if (threadIdx.x == 0) *total = run;
}

__global__ __launch_bounds__(1024) void k_scan_write(const uint32_t* __restrict__ counts, int64_t n,
                                                     const int64_t* __restrict__ carry, int64_t* __restrict__ offsets)
{
    __shared__ int64_t wsum[16];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
    uint32_t c[kScanPer];
    int64_t v = 0;
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) {
        c[q] = base + q < n ? counts[base + q] : 0u;
        v += c[q];
    }
    int64_t run = carry[blockIdx.x] + block_scan_incl(v, wsum) - v;
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) {
        if (base + q < n) offsets[base + q] = run;
        run += c[q];
    }
}

__global__ __launch_bounds__(kThreads) void k_index_emit(IndexArgs args, const int64_t* blk_offsets,
                                                         int n1, int n2, uint32_t* ind, float* dist,
                                                         float* un, float* vn)
{
    const int p = blockIdx.y;
    const PlaneConst& pc = args.planes[p];
    const int64_t col = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    Column cl;
    const bool have = make_column(pc, args.hg, args.cutoff, col, &cl);
    uint32_t cnt = 0;
    if (have)
        for (int ka = cl.ka_lo; ka <= cl.ka_hi; ++ka) {
            float K[3], D, u, v;
            coords(cl, ka, K);
            cnt += point_test(pc, K[0], K[1], K[2], args.cutoff, args.h1, args.h2, &D, &u, &v) ? 1u : 0u;
        }
    // block-exclusive scan of cnt
    __shared__ uint32_t sc[kThreads];
    sc[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? sc[threadIdx.x - off] : 0u;
        __syncthreads();
        sc[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t pos = blk_offsets[(int64_t)p * args.nblk + blockIdx.x] + (sc[threadIdx.x] - cnt);
    if (!have || cnt == 0) return;
    const HalfGrid& hg = args.hg;
    const float fn1 = (float)n1, fn2 = (float)n2;
    for (int ka = cl.ka_lo; ka <= cl.ka_hi; ++ka) {
        float K[3], D, u, v;
        coords(cl, ka, K);
        if (!point_test(pc, K[0], K[1], K[2], args.cutoff, args.h1, args.h2, &D, &u, &v)) continue;
        const int i1 = (int)K[0] - hg.lo[0], i2 = (int)K[1] - hg.lo[1], i3 = (int)K[2] - hg.lo[2];
        ind[pos] = (uint32_t)(i1 + (int64_t)hg.len[0] * (i2 + (int64_t)hg.len[1] * i3));
        dist[pos] = D;
        un[pos] = u / fn1;   // G_CP_plane(1,:)/n1, single (:237)
        vn[pos] = v / fn2;   // G_CP_plane(2,:)/n2, single (:238)
        ++pos;
    }
}

}  // namespace

void grid_index_count(const PlaneConst* d_planes, int P, const HalfGrid& hg, double cutoff, int n1,
                      int n2, int nblk, uint32_t* d_blk_counts, hipStream_t st)
{
    IndexArgs a{d_planes, hg, cutoff, (float)n1 / 2.0f, (float)n2 / 2.0f, nblk};
    dim3 grid(nblk, P);
    k_index_count<<<grid, kThreads, 0, st>>>(a, d_blk_counts);
    GF_LAUNCH_CHECK();
}

size_t grid_index_scan_scratch(int64_t n)
{
    return sizeof(int64_t) * (size_t)std::max<int64_t>((n + kScanChunk - 1) / kScanChunk, 1);
}

void grid_index_scan(const uint32_t* d_counts, int64_t n, int64_t* d_offsets, int64_t* d_total,
                     int64_t* d_scratch, hipStream_t st)
{
    const int nchunk = (int)std::max<int64_t>((n + kScanChunk - 1) / kScanChunk, 1);
    int64_t* csum = d_scratch;   // the chunk sums, then their exclusive carries
    k_scan_sums<<<nchunk, 1024, 0, st>>>(d_counts, n, csum);
    GF_LAUNCH_CHECK();
    k_scan_carry<<<1, 1024, 0, st>>>(csum, nchunk, d_total);
    GF_LAUNCH_CHECK();
    k_scan_write<<<nchunk, 1024, 0, st>>>(d_counts, n, csum, d_offsets);
    GF_LAUNCH_CHECK();
}

void grid_index_emit(const PlaneConst* d_planes, int P, const HalfGrid& hg, double cutoff, int n1,
                     int n2, int nblk, const int64_t* d_blk_offsets, uint32_t* ind, float* dist,
                     float* un, float* vn, hipStream_t st)
{
    IndexArgs a{d_planes, hg, cutoff, (float)n1 / 2.0f, (float)n2 / 2.0f, nblk};
    dim3 grid(nblk, P);
    k_index_emit<<<grid, kThreads, 0, st>>>(a, d_blk_offsets, n1, n2, ind, dist, un, vn);
    GF_LAUNCH_CHECK();
}

int index_threads() { return kThreads; }

}  // namespace gf
