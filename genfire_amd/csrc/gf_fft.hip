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

#include "gf_grid.h"

namespace gf {

namespace {

constexpr int kFT = 256;

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (-i) for the forward transform, a * (+i) for the inverse.
template <bool INV> __device__ __forceinline__ float2 mul_mi(float2 a)
{
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, bool INV> struct Dft;

template <bool INV> struct Dft<2, INV> {
    static __device__ __forceinline__ void run(float2* v)
    {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV> struct Dft<4, INV> {
    static __device__ __forceinline__ void run(float2* v)
    {
        const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const float2 t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <bool INV> struct Dft<8, INV> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, INV>::run(e);
        Dft<4, INV>::run(o);
        const float c = 0.70710678118654752440f;
        // o[q] *= exp(-+ i pi q / 4)
        o[1] = INV ? make_float2(c * (o[1].x - o[1].y), c * (o[1].x + o[1].y))
                   : make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));
        o[2] = mul_mi<INV>(o[2]);
        o[3] = INV ? make_float2(-c * (o[3].x + o[3].y), c * (o[3].x - o[3].y))
                   : make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = cadd(e[q], o[q]);
            v[q + 4] = csub(e[q], o[q]);
        }
    }
};

template <bool INV> struct Dft<3, INV> {
    static __device__ __forceinline__ void run(float2* v)
    {
        const float s3 = 0.86602540378443864676f;
        const float2 t = cadd(v[1], v[2]);
        const float2 d = csub(v[1], v[2]);
        const float2 a = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
        const float2 b = cscale(mul_mi<INV>(d), s3);  // -i*s3*d forward
        v[0] = cadd(v[0], t);
        v[1] = cadd(a, b);
        v[2] = csub(a, b);
    }
};

template <bool INV> struct Dft<5, INV> {
    static __device__ __forceinline__ void run(float2* v)
    {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        const float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        const float2 d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
        const float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        const float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        const float2 b1 = make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
        const float2 b2 = make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
        const float2 ib1 = mul_mi<INV>(b1), ib2 = mul_mi<INV>(b2);  // -i*b forward
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(a1, ib1);
        v[4] = csub(a1, ib1);
        v[2] = cadd(a2, ib2);
        v[3] = csub(a2, ib2);
    }
};

constexpr int pick_radix(int rem)
{
    return rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 3 == 0 ? 3 : rem % 5 == 0 ? 5 : rem;
}

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

struct Tile {
    int64_t base, ps, es;
    int valid;
};

__device__ __forceinline__ Tile tile_of(int axis, int n1, int n2, int n3, int TP, int64_t t)
{
    Tile g;
    if (axis == 0) {
        const int64_t np = (int64_t)n2 * n3;
        const int64_t first = t * TP;
        g.base = first * n1; g.ps = n1; g.es = 1;
        g.valid = (int)min((int64_t)TP, np - first);
    } else {
        const int tpr = (n1 + TP - 1) / TP;
        const int64_t outer = t / tpr;
        const int i1 = (int)(t % tpr) * TP;
        g.ps = 1;
        if (axis == 1) { g.base = i1 + (int64_t)n1 * n2 * outer; g.es = n1; }
        else { g.base = i1 + (int64_t)n1 * outer; g.es = (int64_t)n1 * n2; }
        g.valid = min(TP, n1 - i1);
    }
    return g;
}

enum Mode { kPlain = 0, kFirst = 1, kFused = 2, kLast = 3 };

struct PassArgs {
    IterArgs it;
    float2* data;        // plain pass data (in place)
    const float2* tw;    // exp(-2 pi i m / N), m < N
    int axis;
};

__device__ __forceinline__ float pr_apply(float v, uint8_t s, int positivity, int support)
{
    if (positivity && v < 0.0f) v = 0.0f;   // x(x<0) = 0   (reconstruct.m:127)
    if (support) v = v * (float)s;          // x .* S       (:130)
    return v;
}

template <int N, int MODE, bool INV>
__global__ __launch_bounds__(kFT) void k_fft(PassArgs a)
{
    constexpr int TP = tile_pencils(N);
    extern __shared__ float2 lds[];
    const IterArgs& it = a.it;
    const int n1 = it.n1, n2 = it.n2, n3 = it.n3;
    const Tile g = tile_of(a.axis, n1, n2, n3, TP, blockIdx.x);
    float2* __restrict__ K = (MODE == kPlain) ? a.data : it.K;

    // ---- load ----
    for (int idx = threadIdx.x; idx < TP * N; idx += kFT) {
        const int p = a.axis == 0 ? idx / N : idx % TP;
        const int e = a.axis == 0 ? idx % N : idx / TP;
        float2 val = make_float2(0.0f, 0.0f);
        if (p < g.valid) {
            const int64_t addr = g.base + p * g.ps + e * g.es;
            if (MODE == kFirst) {
                const float xv = it.x[addr];
                const float r = pr_apply(xv, it.S[addr], it.positivity, it.support);
                if (it.best) it.best[addr] = r;
                val.x = it.dr ? (2.0f * r - xv) : r;
            } else {
                val = K[addr];
            }
        }
        lds[p * N + e] = val;
    }
    __syncthreads();

    Stages<N, TP, INV, 1, N>::run(lds, a.tw);

    if (MODE == kFused) {
        // errK numerator (reconstruct.m:135) and Fourier replacement (:163)
        double err = 0.0;
        for (int idx = threadIdx.x; idx < TP * N; idx += kFT) {
            const int p = idx % TP, e = idx / TP;
            if (p >= g.valid) continue;
            const int64_t addr = g.base + p * g.ps + e * g.es;
            const uint8_t f = it.flags[addr];
            if (f) {
                const float2 m = it.M[addr];
                const float2 kv = lds[p * N + e];
                if (f & kFlagMeasured)
                    err += (double)fabsf(sqrtf(kv.x * kv.x + kv.y * kv.y) - sqrtf(m.x * m.x + m.y * m.y));
                if (f & kFlagEnforced) lds[p * N + e] = m;
            }
        }
        // one partial per wave: no LDS beyond the tile (which fills 64 KiB at N=512)
        for (int off = 32; off > 0; off >>= 1) err += __shfl_down(err, off, 64);
        if ((threadIdx.x & 63) == 0) it.err_part[(int64_t)blockIdx.x * (kFT / 64) + (threadIdx.x >> 6)] = err;
        __syncthreads();
        Stages<N, TP, true, 1, N>::run(lds, a.tw);
    }

    // ---- store ----
    const float scale = (INV || MODE == kFused) ? 1.0f / (float)N : 1.0f;
    for (int idx = threadIdx.x; idx < TP * N; idx += kFT) {
        const int p = a.axis == 0 ? idx / N : idx % TP;
        const int e = a.axis == 0 ? idx % N : idx / TP;
        if (p >= g.valid) continue;
        const int64_t addr = g.base + p * g.ps + e * g.es;
        float2 val = lds[p * N + e];
        if (MODE == kLast) {
            const float pf = val.x * scale;           // real(ifftn(k)) (:165)
            if (it.pf_out) it.pf_out[addr] = pf;
            if (it.dr) {
                const float xv = it.x[addr];
                const float r = pr_apply(xv, it.S[addr], it.positivity, it.support);
                it.x[addr] = (xv + pf) - r;
            } else {
                it.x[addr] = pf;
            }
        } else {
            if (scale != 1.0f) val = cscale(val, scale);
            K[addr] = val;
        }
    }
}

// ---- twiddle tables, one per length, built once per device ----
struct TwKey {
    int dev, n;
    bool operator==(const TwKey& o) const { return dev == o.dev && n == o.n; }
};
struct TwHash {
    size_t operator()(const TwKey& k) const { return (size_t)k.dev * 131071u + (size_t)k.n; }
};
std::mutex g_tw_mu;
std::unordered_map<TwKey, float2*, TwHash> g_tw;

const float2* twiddles(int n)
{
    int dev = 0;
    GF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto itr = g_tw.find({dev, n});
    if (itr != g_tw.end()) return itr->second;
    std::vector<float2> h(n);
    for (int m = 0; m < n; ++m) {
        // exp(-2 pi i m / n) from an exactly reduced FP64 angle
        const double ang = -2.0 * M_PI * (double)m / (double)n;
        h[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    float2* d = nullptr;
    GF_HIP(hipMalloc(&d, sizeof(float2) * n));
    GF_HIP(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    g_tw[{dev, n}] = d;
    return d;
}

int64_t num_tiles(int axis, int n1, int n2, int n3, int TP)
{
    if (axis == 0) return ((int64_t)n2 * n3 + TP - 1) / TP;
    const int64_t tpr = (n1 + TP - 1) / TP;
    return tpr * (axis == 1 ? n3 : n2);
}

template <int N, int MODE, bool INV>
void launch_n(const PassArgs& a, hipStream_t st)
{
    constexpr int TP = tile_pencils(N);
    const int64_t tiles = num_tiles(a.axis, a.it.n1, a.it.n2, a.it.n3, TP);
    const size_t lds = (size_t)TP * N * sizeof(float2);
    k_fft<N, MODE, INV><<<(unsigned)tiles, kFT, lds, st>>>(a);
    GF_LAUNCH_CHECK();
}

#define GF_FFT_SIZES(X) \
    X(4) X(8) X(16) X(32) X(45) X(48) X(64) X(75) X(96) X(128) X(160) X(192) X(256) X(320) X(384) X(512) X(640) X(768) X(1024)

template <int MODE, bool INV>
void launch(int n, const PassArgs& a, hipStream_t st)
{
    switch (n) {
#define GF_CASE(NN) case NN: launch_n<NN, MODE, INV>(a, st); return;
        GF_FFT_SIZES(GF_CASE)
#undef GF_CASE
        default:
            raise(GENFIRE_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) +
                                               " is not supported (supported: 16..1024 with factors 2,3,5 in the built list)");
    }
}

int axis_len(const IterArgs& it, int axis) { return axis == 0 ? it.n1 : (axis == 1 ? it.n2 : it.n3); }

// errK[it] = sum(part) / den, one workgroup, fixed summation order.
__global__ __launch_bounds__(256) void k_err_reduce(const double* part, int n, double den, float* out)
{
    __shared__ double s[256];
    double t = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) t += part[i];
    s[threadIdx.x] = t;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (float)(s[0] / den);
}

__device__ __forceinline__ double kspace_coord(int c, int n)
{
    // make_Kspace_indices.m:19-31: even n -> (c - n/2)/(n/2); odd n -> (c - h)/h, h=(n-1)/2
    if (n == 1) return 0.0;
    if (n % 2 == 0) return (double)(c - n / 2) / (double)(n / 2);
    return (double)(c - (n - 1) / 2) / (double)((n - 1) / 2);
}

// flags per FFT-order voxel: measured = (M ~= 0) & mask (reconstruct.m:99);
// enforced = measured & (1 - Q) > delay (:42,:120), Q from the centred index
// exactly as make_Kspace_indices computes it in FP64.
__global__ void k_flags(const float2* __restrict__ M, const uint8_t* __restrict__ mask, int n1, int n2,
                        int n3, double delay, uint8_t* __restrict__ flags)
{
    const int64_t total = (int64_t)n1 * n2 * n3;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total;
         s += (int64_t)gridDim.x * blockDim.x) {
        const float2 m = M[s];
        const bool meas = (m.x != 0.0f || m.y != 0.0f) && (mask == nullptr || mask[s] != 0);
        uint8_t f = 0;
        if (meas) {
            const int s1 = (int)(s % n1), s2 = (int)((s / n1) % n2), s3 = (int)(s / ((int64_t)n1 * n2));
            const int c1 = (s1 + n1 / 2) % n1, c2 = (s2 + n2 / 2) % n2, c3 = (s3 + n3 / 2) % n3;
            const double v1 = kspace_coord(c1, n1), v2 = kspace_coord(c2, n2), v3 = kspace_coord(c3, n3);
            const double q = sqrt((v2 * v2 + v1 * v1) + v3 * v3);
            f = kFlagMeasured;
            if (1.0 - q > delay) f |= kFlagEnforced;
        }
        flags[s] = f;
    }
}

__global__ __launch_bounds__(256) void k_den_part(const float2* __restrict__ M,
                                                  const uint8_t* __restrict__ flags, int64_t n,
                                                  double* __restrict__ part)
{
    double t = 0.0;
    for (int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x; s < n; s += (int64_t)gridDim.x * 256)
        if (flags[s] & kFlagMeasured) {
            const float2 m = M[s];
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

}  // namespace

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
    a.tw = twiddles(n);
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
    a.tw = twiddles(it.n1);
    launch<kFirst, false>(it.n1, a, st);
}

int iter_fused_blocks(const IterArgs& it)
{
    int TP = 8;
    switch (it.n3) {
#define GF_CASE(NN) case NN: TP = tile_pencils(NN); break;
        GF_FFT_SIZES(GF_CASE)
#undef GF_CASE
        default: break;
    }
    return (int)num_tiles(2, it.n1, it.n2, it.n3, TP) * (kFT / 64);
}

void iter_fused(const IterArgs& it, hipStream_t st)
{
    PassArgs a{};
    a.it = it;
    a.axis = 2;
    a.tw = twiddles(it.n3);
    launch<kFused, false>(it.n3, a, st);
}

void iter_last(const IterArgs& it, hipStream_t st)
{
    PassArgs a{};
    a.it = it;
    a.axis = 0;
    a.tw = twiddles(it.n1);
    launch<kLast, true>(it.n1, a, st);
}

void iter_errk_reduce(const double* d_part, int n, double den, float* d_errK_it, hipStream_t st)
{
    k_err_reduce<<<1, 256, 0, st>>>(d_part, n, den, d_errK_it);
    GF_LAUNCH_CHECK();
}

void iter_flags(const float2* d_M, const uint8_t* d_mask, int n1, int n2, int n3, double delay,
                uint8_t* d_flags, hipStream_t st)
{
    const int64_t total = (int64_t)n1 * n2 * n3;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
    k_flags<<<blocks, 256, 0, st>>>(d_M, d_mask, n1, n2, n3, delay, d_flags);
    GF_LAUNCH_CHECK();
}

void iter_den(const float2* d_M, const uint8_t* d_flags, int64_t n, double* d_tmp, int ntmp,
              double* d_out, hipStream_t st)
{
    k_den_part<<<ntmp, 256, 0, st>>>(d_M, d_flags, n, d_tmp);
    GF_LAUNCH_CHECK();
    k_sum_parts<<<1, 64, 0, st>>>(d_tmp, ntmp, d_out);
    GF_LAUNCH_CHECK();
}

}  // namespace gf
