// gf_fft.hip -- 3D FFT along one axis at a time, and the GENFIRE / DR iteration
// passes built on it (reconstruct.m:126-165).
//
// Each workgroup owns a tile of TP pencils of length N along the transformed
// axis.  The tile is staged in LDS ([pencil][element]); a mixed-radix (8,4,2,3,5)
// Stockham autosort FFT runs in place in LDS, every lane owning whole butterflies
// of a stage and a barrier separating a stage's reads from its writes.  Global
// loads and stores are coalesced: along axis 0 a tile is TP*N contiguous
// elements; along axes 1 and 2 a tile is TP consecutive x-positions (TP*8 bytes
// contiguous per row), so TP >= 16 gives whole 128-byte lines.
//
// The iteration is five passes over the spectrum (plus a tiny reduction):
//   first : x -> P_R(x) (or 2 P_R(x) - x for DR) -> complex -> FFT axis 0
//   plain : FFT axis 1
//   fused : FFT axis 2 -> errK partials -> k(C) = M(C) -> IFFT axis 2
//   plain : IFFT axis 1
//   last  : IFFT axis 0 -> real part -> x (GENFIRE) or x + P_F - P_R(x) (DR)
#include <cmath>
#include <mutex>
#include <unordered_map>

#include "gf_fft_core.h"
#include "gf_grid.h"

namespace gf {

namespace {

using namespace fftc;

struct Tile {
    int64_t base, ps, es;
    int valid;
};

// Row c of the rows a pass visits along an outer axis of length n: every row
// (b == 0), or only the b rows of the centred box in FFT order, [0, b - b/2)
// and [n - b/2, n).
__host__ __device__ __forceinline__ int64_t box_row(int64_t c, int n, int b)
{
    return (b == 0 || c < b - b / 2) ? c : c + (n - b);
}
__host__ __device__ constexpr int rows_of(int n, int b) { return b == 0 ? n : b; }

// Tile t of TP pencils along `axis`: axis 0 pencils are contiguous rows; axes 1
// and 2 take TP consecutive x-positions at one (z) or (y) coordinate.  yb, zb
// restrict the y and z rows visited to the centred box (0: all); for axis 0 the
// host only restricts y when TP divides both box segments, so a tile's pencils
// stay consecutive.
__device__ __forceinline__ Tile tile_of(int axis, int n1, int n2, int n3, int TP, int64_t t, int yb = 0,
                                        int zb = 0)
{
    Tile g;
    if (axis == 0) {
        const int ny = rows_of(n2, yb);
        const int64_t np = (int64_t)ny * rows_of(n3, zb);
        const int64_t first = t * TP;
        const int64_t y = box_row(first % ny, n2, yb), z = box_row(first / ny, n3, zb);
        g.base = (y + (int64_t)n2 * z) * n1; g.ps = n1; g.es = 1;
        g.valid = (int)min((int64_t)TP, np - first);
    } else {
        const int tpr = (n1 + TP - 1) / TP;
        const int64_t outer = t / tpr;
        const int i1 = (int)(t % tpr) * TP;
        g.ps = 1;
        if (axis == 1) { g.base = i1 + (int64_t)n1 * n2 * box_row(outer, n3, zb); g.es = n1; }
        else { g.base = i1 + (int64_t)n1 * box_row(outer, n2, yb); g.es = (int64_t)n1 * n2; }
        g.valid = min(TP, n1 - i1);
    }
    return g;
}

__host__ __device__ constexpr int64_t ntiles(int axis, int n1, int n2, int n3, int TP, int yb = 0, int zb = 0)
{
    return axis == 0 ? ((int64_t)rows_of(n2, yb) * rows_of(n3, zb) + TP - 1) / TP
                     : (int64_t)((n1 + TP - 1) / TP) * (axis == 1 ? rows_of(n3, zb) : rows_of(n2, yb));
}

enum Mode { kPlain = 0, kFirst = 1, kFused = 2, kLast = 3 };

struct PassArgs {
    IterArgs it;
    float2* data;        // plain pass data (in place)
    const float2* src;   // plain pass input when not data (out of place), else null
    const float2* tw;    // exp(-2 pi i m / N), m < N
    int axis;
    int yb, zb;          // plain passes: visit only the y / z rows of the centred box (0: all)
};

__device__ __forceinline__ float pr_apply(float v, uint8_t s, int positivity, int support)
{
    if (positivity && v < 0.0f) v = 0.0f;   // x(x<0) = 0   (reconstruct.m:127)
    if (support) v = v * (float)s;          // x .* S       (:130)
    return v;
}

// Element load of a pass (mode-specific): complex spectrum, or the iterate through
// P_R (GENFIRE: P_R(x); DR: 2 P_R(x) - x) with the optional best-copy of P_R(x).
template <int MODE>
__device__ __forceinline__ float2 pass_load(const PassArgs& a, float2* K, int64_t addr)
{
    if (MODE == kFirst) {
        const IterArgs& it = a.it;
        const float xv = it.x[addr];
        const float r = pr_apply(xv, it.S[addr], it.positivity, it.support);
        if (it.best) it.best[addr] = r;
        return make_float2(it.dr ? (2.0f * r - xv) : r, 0.0f);
    }
    if (MODE == kPlain && a.src) return a.src[addr];
    return K[addr];
}

// Element store of a pass: complex (scaled), or real(ifftn) folded into the
// iterate (GENFIRE: x = P_F; DR: x = x + P_F - P_R(x)).
template <int MODE>
__device__ __forceinline__ void pass_store(const PassArgs& a, float2* K, int64_t addr, float2 val, float scale)
{
    if (MODE == kLast) {
        const IterArgs& it = a.it;
        const float pf = val.x * scale;                 // real(ifftn(k)) (:165)
        if (it.pf_out) it.pf_out[addr] = pf;
        if (it.dr) {
            const float xv = it.x[addr];
            const float r = pr_apply(xv, it.S[addr], it.positivity, it.support);
            it.x[addr] = (xv + pf) - r;
        } else {
            it.x[addr] = pf;
        }
    } else {
        K[addr] = scale != 1.0f ? cscale(val, scale) : val;
    }
}

template <bool SWZ>
__device__ __forceinline__ double csr_replace(float2* lds, int N, const IterArgs& it, int x0, int y,
                                              int valid, int tid, int nthreads)
{
    const int64_t r0 = (int64_t)x0 + (int64_t)it.n1 * y;
    const int64_t beg = it.csr_off[r0], end = it.csr_off[r0 + valid];
    double err = 0.0;
    for (int64_t e = beg + tid; e < end; e += nthreads) {
        const uint32_t m = it.csr_meta[e];
        const int z = (int)(m & 0x7FFu), x = (int)((m >> 11) & 0x7FFu);
        float2* slot = lds + (x - x0) * N + (SWZ ? swz(z, x - x0) : z);
        const float2 kv = *slot, mv = it.csr_val[e];
        err += (double)fabsf(sqrtf(kv.x * kv.x + kv.y * kv.y) - sqrtf(mv.x * mv.x + mv.y * mv.y));
        if (m >> 31) *slot = mv;
    }
    return err;
}

__device__ __forceinline__ void wave_partial(double err, double* out, int64_t slot)
{
    for (int off = 32; off > 0; off >>= 1) err += __shfl_down(err, off, 64);
    if ((threadIdx.x & 63) == 0) out[slot] = err;
}

// ============================================================================
// Generic LDS kernel (any length with factors 2,3,5): every stage in LDS.
// ============================================================================
constexpr int tile_pencils(int N) { return N <= 256 ? 32 : (N <= 512 ? 16 : 8); }

// One Stockham stage of radix R on every pencil of the tile; LS = product of the
// radices already applied.  Butterfly j of a pencil reads x[j + r*N/R], applies
// the twiddles exp(-+2 pi i r (j mod LS) / (LS R)), and writes
// y[(j - j mod LS) R + j mod LS + q LS].
template <int N, int TP, int R, int LS, bool INV>
__device__ __forceinline__ void fft_stage(float2* lds, const float2* __restrict__ tw)
{
    constexpr int M = N / R;
    constexpr int NBF = TP * M;
    constexpr int NB = (NBF + kFT - 1) / kFT;
    float2 v[NB][R];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int b = threadIdx.x + i * kFT;
        if (NBF % kFT == 0 || b < NBF) {
            const int p = b / M, j = b % M, k = j % LS;
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = lds[p * N + j + r * M];
            if (LS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    float2 w = tw[r * k * (N / (LS * R))];
                    if (INV) w.y = -w.y;
                    v[i][r] = cmul(v[i][r], w);
                }
            }
            Dft<R, INV>::run(v[i]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int b = threadIdx.x + i * kFT;
        if (NBF % kFT == 0 || b < NBF) {
            const int p = b / M, j = b % M, k = j % LS;
            const int base = p * N + (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) lds[base + q * LS] = v[i][q];
        }
    }
    __syncthreads();
}

template <int N, int TP, bool INV, int LS, int REM> struct Stages {
    static __device__ __forceinline__ void run(float2* lds, const float2* tw)
    {
        constexpr int R = pick_radix(REM);
        fft_stage<N, TP, R, LS, INV>(lds, tw);
        Stages<N, TP, INV, LS * R, REM / R>::run(lds, tw);
    }
};
template <int N, int TP, bool INV, int LS> struct Stages<N, TP, INV, LS, 1> {
    static __device__ __forceinline__ void run(float2*, const float2*) {}
};

template <int N, int MODE, bool INV>
__global__ __launch_bounds__(kFT) void k_fft(PassArgs a)
{
    constexpr int TP = tile_pencils(N);
    extern __shared__ float2 lds[];
    const IterArgs& it = a.it;
    const Tile g = tile_of(a.axis, it.n1, it.n2, it.n3, TP, blockIdx.x, a.yb, a.zb);
    float2* __restrict__ K = (MODE == kPlain) ? a.data : it.K;

    for (int idx = threadIdx.x; idx < TP * N; idx += kFT) {
        const int p = a.axis == 0 ? idx / N : idx % TP;
        const int e = a.axis == 0 ? idx % N : idx / TP;
        float2 val = make_float2(0.0f, 0.0f);
        if (p < g.valid) val = pass_load<MODE>(a, K, g.base + p * g.ps + e * g.es);
        lds[p * N + e] = val;
    }
    __syncthreads();

    Stages<N, TP, INV, 1, N>::run(lds, a.tw);

    if (MODE == kFused) {
        const int tpr = (it.n1 + TP - 1) / TP;
        const int x0 = (int)(blockIdx.x % tpr) * TP, y = (int)(blockIdx.x / tpr);
        const double err = csr_replace<false>(lds, N, it, x0, y, g.valid, threadIdx.x, kFT);
        wave_partial(err, it.err_part, (int64_t)blockIdx.x * (kFT / 64) + (threadIdx.x >> 6));
        __syncthreads();
        Stages<N, TP, true, 1, N>::run(lds, a.tw);
    }

    const float scale = (INV || MODE == kFused) ? 1.0f / (float)N : 1.0f;
    for (int idx = threadIdx.x; idx < TP * N; idx += kFT) {
        const int p = a.axis == 0 ? idx / N : idx % TP;
        const int e = a.axis == 0 ? idx % N : idx / TP;
        if (p >= g.valid) continue;
        pass_store<MODE>(a, K, g.base + p * g.ps + e * g.es, lds[p * N + e], scale);
    }
}

template <int N, int MODE, bool INV>
__global__ __launch_bounds__(reg_threads<N>()) void k_fft_reg(PassArgs a)
{
    using PL = RegPlan<N>;
    constexpr int E = PL::E, T = N / E, TP = PL::TP, R0 = PL::R0, RL = reg_last_radix<N>();
    constexpr int NT = reg_threads<N>();
    extern __shared__ float2 lds[];
    const IterArgs& it = a.it;
    const int t = threadIdx.x;
    const int j = a.axis == 0 ? t % T : t / TP;
    const int p = a.axis == 0 ? t / T : t % TP;
    const Tile g = tile_of(a.axis, it.n1, it.n2, it.n3, TP, blockIdx.x, a.yb, a.zb);
    const bool valid = p < g.valid;
    float2* __restrict__ K = (MODE == kPlain) ? a.data : it.K;
    const int64_t pbase = g.base + p * g.ps;
    float2* twt = lds + TP * N;   // stage twiddle table past the tile
    reg_tw_fill<N>(twt, a.tw, t, NT);

    float2 v[E];
#pragma unroll
    for (int i = 0; i < E / R0; ++i)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int e = j + T * i + r * (N / R0);
            v[i * R0 + r] = valid ? pass_load<MODE>(a, K, pbase + e * g.es) : make_float2(0.0f, 0.0f);
        }

    reg_fft<N, INV>(v, lds, p * N, j, p, twt);

    if (MODE == kFused) {
        // natural order into LDS, replace, back into the first-stage layout
        reg_store_x<N, E, RL>(lds, p * N, v, j, p);
        __syncthreads();
        const int tpr = (it.n1 + TP - 1) / TP;
        const int x0 = (int)(blockIdx.x % tpr) * TP, y = (int)(blockIdx.x / tpr);
        const double err = csr_replace<true>(lds, N, it, x0, y, g.valid, t, NT);
        wave_partial(err, it.err_part, (int64_t)blockIdx.x * (NT / 64) + (t >> 6));
        __syncthreads();
        reg_read_x<N, E, R0>(lds, p * N, v, j, p);
        __syncthreads();
        reg_fft<N, true>(v, lds, p * N, j, p, twt);
    }

    if (!valid) return;
    const float scale = (INV || MODE == kFused) ? 1.0f / (float)N : 1.0f;
#pragma unroll
    for (int i = 0; i < E / RL; ++i)
#pragma unroll
        for (int q = 0; q < RL; ++q) {
            const int e = j + T * i + q * (N / RL);
            pass_store<MODE>(a, K, pbase + e * g.es, v[i * RL + q], scale);
        }
}

#define GF_FFT_SIZES(X) \
    X(4) X(8) X(16) X(32) X(45) X(48) X(64) X(75) X(96) X(128) X(160) X(192) X(256) X(320) X(384) X(512) X(640) X(768) X(1024)

// Workgroups and per-wave partial slots of a pass launch.
struct LaunchShape {
    int64_t blocks;
    int threads;
    size_t lds;
};

template <int N> constexpr int pass_pencils()
{
    if constexpr (RegPlan<N>::ok) return RegPlan<N>::TP;
    else return tile_pencils(N);
}

template <int N> LaunchShape shape_n(int axis, int n1, int n2, int n3, int yb = 0, int zb = 0)
{
    if constexpr (RegPlan<N>::ok) {
        return {ntiles(axis, n1, n2, n3, RegPlan<N>::TP, yb, zb), reg_threads<N>(),
                ((size_t)RegPlan<N>::TP * N + reg_tw_size<N>()) * sizeof(float2)};
    } else {
        return {ntiles(axis, n1, n2, n3, tile_pencils(N), yb, zb), kFT, (size_t)tile_pencils(N) * N * sizeof(float2)};
    }
}

int pass_pencils_of(int n)
{
    switch (n) {
#define GF_CASE(NN) case NN: return pass_pencils<NN>();
        GF_FFT_SIZES(GF_CASE)
#undef GF_CASE
        default:
            raise(GENFIRE_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " is not supported");
    }
}

LaunchShape shape_of(int n, int axis, int n1, int n2, int n3)
{
    switch (n) {
#define GF_CASE(NN) case NN: return shape_n<NN>(axis, n1, n2, n3);
        GF_FFT_SIZES(GF_CASE)
#undef GF_CASE
        default:
            raise(GENFIRE_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " is not supported");
    }
}

template <int N, int MODE, bool INV>
void launch_n(const PassArgs& a, hipStream_t st)
{
    const LaunchShape sh = shape_n<N>(a.axis, a.it.n1, a.it.n2, a.it.n3, a.yb, a.zb);
    if constexpr (RegPlan<N>::ok) {
        if (sh.lds > 65536)
            GF_HIP(hipFuncSetAttribute((const void*)k_fft_reg<N, MODE, INV>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.lds));
        k_fft_reg<N, MODE, INV><<<(unsigned)sh.blocks, sh.threads, sh.lds, st>>>(a);
    }
    else
        k_fft<N, MODE, INV><<<(unsigned)sh.blocks, sh.threads, sh.lds, st>>>(a);
    GF_LAUNCH_CHECK();
}

template <int MODE, bool INV>
void launch(int n, const PassArgs& a, hipStream_t st)
{
    switch (n) {
#define GF_CASE(NN) case NN: launch_n<NN, MODE, INV>(a, st); return;
        GF_FFT_SIZES(GF_CASE)
#undef GF_CASE
        default:
            raise(GENFIRE_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) +
                                               " is not supported (lengths with factors 2, 3, 5 from the built list)");
    }
}

int axis_len(const IterArgs& it, int axis) { return axis == 0 ? it.n1 : (axis == 1 ? it.n2 : it.n3); }

// errK[it] = sum(part) / den, one workgroup, fixed summation order.
// out[0], or out[*ctr] with *ctr advanced by one when ctr is set (the iteration
// index kept on the device, so that one captured launch serves every iteration)
__global__ __launch_bounds__(1024) void k_err_reduce(const double* part, int64_t n, double den, float* out, int* ctr)
{
    __shared__ double s[1024];
    double t = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) t += part[i];
    s[threadIdx.x] = t;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (ctr) {
            out[*ctr] = (float)(s[0] / den);
            *ctr += 1;
        } else {
            *out = (float)(s[0] / den);
        }
    }
}

__device__ __forceinline__ double kspace_coord(int c, int n)
{
    // make_Kspace_indices.m:19-31: even n -> (c - n/2)/(n/2); odd n -> (c - h)/h, h=(n-1)/2
    if (n == 1) return 0.0;
    if (n % 2 == 0) return (double)(c - n / 2) / (double)(n / 2);
    return (double)(c - (n - 1) / 2) / (double)((n - 1) / 2);
}

__device__ __forceinline__ bool measured_at(const float2* M, const uint8_t* mask, int64_t s)
{
    const float2 m = M[s];
    return (m.x != 0.0f || m.y != 0.0f) && (mask == nullptr || mask[s] != 0);
}

// number of measured voxels (M ~= 0 & mask, reconstruct.m:99) per z-pencil (x, y)
__global__ void k_csr_count(const float2* __restrict__ M, const uint8_t* __restrict__ mask, int n1, int n2,
                            int n3, uint32_t* __restrict__ cnt)
{
    const int64_t np = (int64_t)n1 * n2;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
        uint32_t c = 0;
        for (int z = 0; z < n3; ++z) c += measured_at(M, mask, q + np * z) ? 1u : 0u;
        cnt[q] = c;
    }
}

__global__ void k_csr_fill(const float2* __restrict__ M, const uint8_t* __restrict__ mask, int n1, int n2,
                           int n3, const int64_t* __restrict__ off, uint32_t* __restrict__ meta,
                           float2* __restrict__ val)
{
    const int64_t np = (int64_t)n1 * n2;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
        int64_t o = off[q];
        const uint32_t x = (uint32_t)(q % n1);
        for (int z = 0; z < n3; ++z) {
            const int64_t s = q + np * z;
            if (!measured_at(M, mask, s)) continue;
            meta[o] = (uint32_t)z | (x << 11);
            val[o] = M[s];
            ++o;
        }
    }
}

// enforced = (1 - Q) > delay (reconstruct.m:42,120), Q from the centred index
// exactly as make_Kspace_indices computes it in FP64.
__global__ void k_csr_enforce(int n1, int n2, int n3, const int64_t* __restrict__ off,
                              uint32_t* __restrict__ meta, double delay)
{
    const int64_t np = (int64_t)n1 * n2;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
        const int s1 = (int)(q % n1), s2 = (int)(q / n1);
        const double v1 = kspace_coord((s1 + n1 / 2) % n1, n1);
        const double v2 = kspace_coord((s2 + n2 / 2) % n2, n2);
        for (int64_t e = off[q]; e < off[q + 1]; ++e) {
            uint32_t m = meta[e] & 0x7FFFFFFFu;
            const int z = (int)(m & 0x7FFu);
            const double v3 = kspace_coord((z + n3 / 2) % n3, n3);
            const double Q = sqrt((v2 * v2 + v1 * v1) + v3 * v3);
            if (1.0 - Q > delay) m |= 0x80000000u;
            meta[e] = m;
        }
    }
}

__global__ __launch_bounds__(256) void k_den_part(const float2* __restrict__ val, int64_t n,
                                                  double* __restrict__ part)
{
    double t = 0.0;
    for (int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x; s < n; s += (int64_t)gridDim.x * 256) {
        const float2 m = val[s];
        t += (double)sqrtf(m.x * m.x + m.y * m.y);
    }
    __shared__ double r[256];
    r[threadIdx.x] = t;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) r[threadIdx.x] += r[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = r[0];
}

__global__ void k_sum_parts(const double* part, int n, double* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < n; ++i) t += part[i];
        *out = t;
    }
}

int pencil_blocks(int64_t np) { return (int)std::min<int64_t>((np + 255) / 256, 4096); }

}  // namespace

// ---- twiddle tables, one per length, built once per device ----
namespace {
struct TwKey {
    int dev, n;
    bool operator==(const TwKey& o) const { return dev == o.dev && n == o.n; }
};
struct TwHash {
    size_t operator()(const TwKey& k) const { return (size_t)k.dev * 131071u + (size_t)k.n; }
};
std::mutex g_tw_mu;
std::unordered_map<TwKey, float2*, TwHash> g_tw;
}  // namespace

const float2* fft_twiddles(int n)
{
    int dev = 0;
    GF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto itr = g_tw.find({dev, n});
    if (itr != g_tw.end()) return itr->second;
    std::vector<float2> h(n);
    for (int m = 0; m < n; ++m) {
        const double ang = -2.0 * M_PI * (double)m / (double)n;
        h[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    float2* d = nullptr;
    GF_HIP(hipMalloc(&d, sizeof(float2) * n));
    GF_HIP(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    g_tw[{dev, n}] = d;
    return d;
}


bool fft_supported(int n)
{
    switch (n) {
#define GF_CASE(NN) case NN: return true;
        GF_FFT_SIZES(GF_CASE)
#undef GF_CASE
        default: return false;
    }
}

void fft_axis(float2* d, int n1, int n2, int n3, int axis, int inverse, hipStream_t st)
{
    PassArgs a{};
    a.it.n1 = n1; a.it.n2 = n2; a.it.n3 = n3;
    a.data = d;
    a.axis = axis;
    const int n = axis_len(a.it, axis);
    if (n == 1) return;
    a.tw = fft_twiddles(n);
    if (inverse) launch<kPlain, true>(n, a, st);
    else launch<kPlain, false>(n, a, st);
}

void fft_axis_from(const float2* src, float2* d, int n1, int n2, int n3, int axis, int inverse, hipStream_t st)
{
    PassArgs a{};
    a.it.n1 = n1; a.it.n2 = n2; a.it.n3 = n3;
    a.data = d;
    a.src = src;
    a.axis = axis;
    const int n = axis_len(a.it, axis);
    GF_CHECK(n > 1, GENFIRE_ERR_ARG, "out-of-place pass along an axis of length 1");
    a.tw = fft_twiddles(n);
    if (inverse) launch<kPlain, true>(n, a, st);
    else launch<kPlain, false>(n, a, st);
}

void fft_axis_rows(float2* d, int n1, int n2, int n3, int axis, int inverse, int yb, int zb, hipStream_t st)
{
    GF_CHECK(yb >= 0 && yb <= n2 && zb >= 0 && zb <= n3 && (axis != 1 || yb == 0) && (axis != 2 || zb == 0),
             GENFIRE_ERR_ARG, "fft_axis_rows: box rows along the transformed axis or beyond the grid");
    PassArgs a{};
    a.it.n1 = n1; a.it.n2 = n2; a.it.n3 = n3;
    a.data = d;
    a.axis = axis;
    const int n = axis_len(a.it, axis);
    if (n == 1) return;
    yb = yb == n2 ? 0 : yb;
    zb = zb == n3 ? 0 : zb;
    if (axis == 0 && (yb > 0 || zb > 0)) {
        // a tile's TP pencils must stay consecutive in memory: TP divides the
        // rows per z plane and both y segments of the box; else every row
        const int tp = pass_pencils_of(n);
        const bool ok = (yb > 0 ? ((yb - yb / 2) % tp == 0 && (yb / 2) % tp == 0) : n2 % tp == 0);
        if (!ok) yb = zb = 0;
    }
    a.yb = yb;
    a.zb = zb;
    a.tw = fft_twiddles(n);
    if (inverse) launch<kPlain, true>(n, a, st);
    else launch<kPlain, false>(n, a, st);
}

void fft3(float2* d, int n1, int n2, int n3, int inverse, hipStream_t st)
{
    for (int ax = 0; ax < 3; ++ax) fft_axis(d, n1, n2, n3, ax, inverse, st);
}

void iter_first(const IterArgs& it, hipStream_t st)
{
    PassArgs a{};
    a.it = it;
    a.axis = 0;
    a.tw = fft_twiddles(it.n1);
    launch<kFirst, false>(it.n1, a, st);
}

int64_t iter_fused_partials(const IterArgs& it)
{
    const LaunchShape sh = shape_of(it.n3, 2, it.n1, it.n2, it.n3);
    return sh.blocks * (sh.threads / 64);
}

void iter_fused(const IterArgs& it, hipStream_t st)
{
    PassArgs a{};
    a.it = it;
    a.axis = 2;
    a.tw = fft_twiddles(it.n3);
    launch<kFused, false>(it.n3, a, st);
}

void iter_last(const IterArgs& it, hipStream_t st)
{
    PassArgs a{};
    a.it = it;
    a.axis = 0;
    a.tw = fft_twiddles(it.n1);
    launch<kLast, true>(it.n1, a, st);
}

void iter_errk_reduce(const double* d_part, int64_t n, double den, float* d_errK_it, hipStream_t st)
{
    k_err_reduce<<<1, 1024, 0, st>>>(d_part, n, den, d_errK_it, nullptr);
    GF_LAUNCH_CHECK();
}

void iter_errk_reduce_next(const double* d_part, int64_t n, double den, float* d_errK, int* d_ctr, hipStream_t st)
{
    k_err_reduce<<<1, 1024, 0, st>>>(d_part, n, den, d_errK, d_ctr);
    GF_LAUNCH_CHECK();
}

void csr_count(const float2* d_M, const uint8_t* d_mask, int n1, int n2, int n3, uint32_t* d_cnt, hipStream_t st)
{
    k_csr_count<<<pencil_blocks((int64_t)n1 * n2), 256, 0, st>>>(d_M, d_mask, n1, n2, n3, d_cnt);
    GF_LAUNCH_CHECK();
}

void csr_fill(const float2* d_M, const uint8_t* d_mask, int n1, int n2, int n3, const int64_t* d_off,
              uint32_t* d_meta, float2* d_val, hipStream_t st)
{
    k_csr_fill<<<pencil_blocks((int64_t)n1 * n2), 256, 0, st>>>(d_M, d_mask, n1, n2, n3, d_off, d_meta, d_val);
    GF_LAUNCH_CHECK();
}

void csr_enforce(int n1, int n2, int n3, const int64_t* d_off, uint32_t* d_meta, double delay, hipStream_t st)
{
    k_csr_enforce<<<pencil_blocks((int64_t)n1 * n2), 256, 0, st>>>(n1, n2, n3, d_off, d_meta, delay);
    GF_LAUNCH_CHECK();
}

void csr_den(const float2* d_val, int64_t n, double* d_tmp, int ntmp, double* d_out, hipStream_t st)
{
    k_den_part<<<ntmp, 256, 0, st>>>(d_val, n, d_tmp);
    GF_LAUNCH_CHECK();
    k_sum_parts<<<1, 64, 0, st>>>(d_tmp, ntmp, d_out);
    GF_LAUNCH_CHECK();
}

}  // namespace gf
