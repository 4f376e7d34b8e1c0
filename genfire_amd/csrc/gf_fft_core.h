// gf_fft_core.h -- device building blocks of the FFT kernels: complex helpers,
// radix-2/3/4/5/6/8/10/16 butterflies and the register-resident Stockham plans
// for lengths 64..1024.
#pragma once

#include <hip/hip_runtime.h>

namespace gf {
namespace fftc {

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

// Composite radices as P x Q (input n = Q*a + b, output k = c + P*d):
// X[c + P d] = sum_b W_Q^(bd) [W_N^(bc) sum_a x[Q a + b] W_P^(ac)].
template <int P, int Q, bool INV>
__device__ __forceinline__ void dft_pq(float2* v, const float2 (&w)[P][Q])
{
    float2 t[P][Q];
#pragma unroll
    for (int b = 0; b < Q; ++b) {
        float2 q[P];
#pragma unroll
        for (int a = 0; a < P; ++a) q[a] = v[Q * a + b];
        Dft<P, INV>::run(q);
#pragma unroll
        for (int c = 0; c < P; ++c) t[c][b] = (c == 0 || b == 0) ? q[c] : cmul(q[c], w[c][b]);
    }
#pragma unroll
    for (int c = 0; c < P; ++c) {
        float2 q[Q];
#pragma unroll
        for (int b = 0; b < Q; ++b) q[b] = t[c][b];
        Dft<Q, INV>::run(q);
#pragma unroll
        for (int d = 0; d < Q; ++d) v[c + P * d] = q[d];
    }
}

// 6 = 2 x 3, twiddles W_6^b (b = 1, 2) = cos(pi b/3) -+ i sin(pi b/3)
template <bool INV> struct Dft<6, INV> {
    static __device__ __forceinline__ void run(float2* v)
    {
        const float s3 = 0.86602540378443864676f;
        const float sg = INV ? 1.0f : -1.0f;
        const float2 w[2][3] = {{{1.0f, 0.0f}, {1.0f, 0.0f}, {1.0f, 0.0f}},
                                {{1.0f, 0.0f}, {0.5f, sg * s3}, {-0.5f, sg * s3}}};
        dft_pq<2, 3, INV>(v, w);
    }
};

// 10 = 2 x 5, twiddles W_10^b = cos(pi b/5) -+ i sin(pi b/5), b = 1..4
template <bool INV> struct Dft<10, INV> {
    static __device__ __forceinline__ void run(float2* v)
    {
        const float c1 = 0.80901699437494742410f, s1 = 0.58778525229247312917f;
        const float c2 = 0.30901699437494742410f, s2 = 0.95105651629515357212f;
        const float sg = INV ? 1.0f : -1.0f;
        const float2 w[2][5] = {{{1.0f, 0.0f}, {1.0f, 0.0f}, {1.0f, 0.0f}, {1.0f, 0.0f}, {1.0f, 0.0f}},
                                {{1.0f, 0.0f}, {c1, sg * s1}, {c2, sg * s2}, {-c2, sg * s2}, {-c1, sg * s1}}};
        dft_pq<2, 5, INV>(v, w);
    }
};

constexpr int pick_radix(int rem)
{
    return rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 3 == 0 ? 3 : rem % 5 == 0 ? 5 : rem;
}

// 16 = 4 x 4: X[c + 4b] = sum_a W4^(ab) W16^(ac) sum_m x[a + 4m] W4^(mc).
template <bool INV> struct Dft<16, INV> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 t[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            float2 q[4] = {v[a], v[a + 4], v[a + 8], v[a + 12]};
            Dft<4, INV>::run(q);
#pragma unroll
            for (int c = 0; c < 4; ++c) t[a][c] = q[c];
        }
        // W16^(ac), a, c in 1..3 (forward sign -)
        const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
        const float c2 = 0.70710678118654752440f;
        const float2 w1 = make_float2(c1, INV ? s1 : -s1);
        const float2 w2 = make_float2(c2, INV ? c2 : -c2);
        const float2 w3 = make_float2(s1, INV ? c1 : -c1);
        const float2 w6 = make_float2(-c2, INV ? c2 : -c2);
        const float2 w9 = make_float2(-c1, INV ? -s1 : s1);
        t[1][1] = cmul(t[1][1], w1); t[1][2] = cmul(t[1][2], w2); t[1][3] = cmul(t[1][3], w3);
        t[2][1] = cmul(t[2][1], w2); t[2][2] = mul_mi<INV>(t[2][2]); t[2][3] = cmul(t[2][3], w6);
        t[3][1] = cmul(t[3][1], w3); t[3][2] = cmul(t[3][2], w6); t[3][3] = cmul(t[3][3], w9);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float2 q[4] = {t[0][c], t[1][c], t[2][c], t[3][c]};
            Dft<4, INV>::run(q);
#pragma unroll
            for (int b = 0; b < 4; ++b) v[c + 4 * b] = q[b];
        }
    }
};

// LDS column swizzle of the register kernel: element `col` of pencil `p` lives at
// col ^ (((col >> 3) ^ p) & 15).  The 16 lanes an LDS access serves together then
// fall on 16 distinct bank pairs for every pattern used here: 16 consecutive
// columns of one pencil, the radix-8 Stockham writes 8b+q and 64m+k+8q, and one
// column of 16 different pencils (rows N*8 bytes apart would otherwise share a bank).
__device__ __forceinline__ int swz(int col, int p) { return col ^ (((col >> 3) ^ p) & 15); }

// The swizzle is linear over xor: swz(col, p) = swz_lin(col) ^ (p & 15).  So
// for an element whose index splits into a lane part and a compile-time part
// with no bits in common, e = lane_e | ce, the position is a per-lane value xor
// a constant, swz_lin(lane_e) ^ (p & 15) ^ swz_lin(ce).  With the lane value
// below 2^LB, the constant's bits from LB up add instead (an immediate offset of
// the LDS instruction) and its low bits take one xor, shared by every access
// with the same low bits.  The *_x forms below address LDS this way; they take
// the kernel's dynamic LDS base and the row's element offset from it (a
// multiple of 2^LB) and place every element exactly where swz does.
__host__ __device__ constexpr int swz_lin(int c) { return c ^ ((c >> 3) & 15); }
constexpr bool is_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }
constexpr int ilog2(int x)
{
    int r = 0;
    while ((1 << r) < x) ++r;
    return r;
}
constexpr int lane_bits(int bound) { return ilog2(bound) < 4 ? 4 : ilog2(bound); }

// Byte offset of the lane part: row offset rowoff (elements) plus the lane value.
__device__ __forceinline__ uint32_t lane_bytes(int rowoff, int lane_e, int p)
{
    return 8u * (uint32_t)(rowoff + (swz_lin(lane_e) ^ (p & 15)));
}

template <int LB>
__device__ __forceinline__ float2& lds_at(float2* lds, uint32_t lb, int ce)
{
    const int K = swz_lin(ce), klo = K & ((1 << LB) - 1), khi = K - klo;
    return *reinterpret_cast<float2*>(reinterpret_cast<char*>(lds) + ((lb ^ (uint32_t)(8 * klo)) + (uint32_t)(8 * khi)));
}

// errK numerator (reconstruct.m:135) and Fourier replacement (:163) on the
// natural-order spectrum of the tile's z-pencils [x0, x0+valid) at row y.  The
// measured voxels come from the per-pencil list (meta = z | x<<11 | enforced<<31).
// ============================================================================
// Register kernel (lengths 64..1024 with factors 2, 3, 5 from the plans below).  A pencil is owned by
// T = N/E lanes holding E elements each; every radix stage is computed in
// registers and only the two exchanges between the three stages go through
// LDS.  The first stage's inputs and the last stage's outputs are the elements
// j + T*i + r*N/R of the lane's pencil, so global loads and stores go straight
// from and to registers: along axis 0 consecutive lanes take consecutive
// elements, along axes 1 and 2 consecutive lanes take consecutive pencils
// (TP * 8 bytes contiguous per row).
// ============================================================================
template <int N> struct RegPlan { static constexpr bool ok = false; };
// Power-of-two lengths: radix 8/16 stages.  Lengths with a factor 3 or 5: the
// odd (or 6/10) radix last, so that every earlier stage's span LS stays a power
// of two, and E = lcm of the radices.
template <> struct RegPlan<64> { static constexpr bool ok = true; static constexpr int E = 8, R0 = 8, R1 = 8, R2 = 1, R3 = 1, TP = 32; };
template <> struct RegPlan<96> { static constexpr bool ok = true; static constexpr int E = 12, R0 = 4, R1 = 4, R2 = 6, R3 = 1, TP = 16; };
template <> struct RegPlan<128> { static constexpr bool ok = true; static constexpr int E = 8, R0 = 8, R1 = 8, R2 = 2, R3 = 1, TP = 16; };
template <> struct RegPlan<160> { static constexpr bool ok = true; static constexpr int E = 20, R0 = 4, R1 = 4, R2 = 10, R3 = 1, TP = 16; };
template <> struct RegPlan<192> { static constexpr bool ok = true; static constexpr int E = 12, R0 = 4, R1 = 4, R2 = 4, R3 = 3, TP = 16; };
template <> struct RegPlan<256> { static constexpr bool ok = true; static constexpr int E = 8, R0 = 8, R1 = 8, R2 = 4, R3 = 1, TP = 16; };
template <> struct RegPlan<320> { static constexpr bool ok = true; static constexpr int E = 20, R0 = 4, R1 = 4, R2 = 4, R3 = 5, TP = 16; };
template <> struct RegPlan<384> { static constexpr bool ok = true; static constexpr int E = 12, R0 = 4, R1 = 4, R2 = 4, R3 = 6, TP = 16; };
template <> struct RegPlan<512> { static constexpr bool ok = true; static constexpr int E = 8, R0 = 8, R1 = 8, R2 = 8, R3 = 1, TP = 16; };
template <> struct RegPlan<640> { static constexpr bool ok = true; static constexpr int E = 20, R0 = 4, R1 = 4, R2 = 4, R3 = 10, TP = 8; };
template <> struct RegPlan<768> { static constexpr bool ok = true; static constexpr int E = 24, R0 = 8, R1 = 8, R2 = 4, R3 = 3, TP = 8; };
template <> struct RegPlan<1024> { static constexpr bool ok = true; static constexpr int E = 16, R0 = 16, R1 = 16, R2 = 4, R3 = 1, TP = 8; };

template <int N> constexpr int reg_threads() { return RegPlan<N>::TP * (N / RegPlan<N>::E); }
template <int N> constexpr int reg_last_radix()
{
    return RegPlan<N>::R3 > 1 ? RegPlan<N>::R3 : RegPlan<N>::R2 > 1 ? RegPlan<N>::R2 : RegPlan<N>::R1;
}

// Twiddles of the register FFT, stage by stage: the stage of span LS and radix
// R (every stage after the first) holds (R-1) x LS entries, entry (r-1)*LS + k
// = exp(-2 pi i r k / (LS R)), so the lanes of a wave read consecutive entries
// (an LDS copy of it has no bank conflicts, unlike the strided r*k reads of the
// plain N-entry table).
template <int N> constexpr int reg_tw_size()
{
    using PL = RegPlan<N>;
    int sz = (PL::R1 - 1) * PL::R0;
    if (PL::R2 > 1) sz += (PL::R2 - 1) * PL::R0 * PL::R1;
    if (PL::R3 > 1) sz += (PL::R3 - 1) * PL::R0 * PL::R1 * PL::R2;
    return sz;
}

// Fills the stage table from the plain table tw[m] = exp(-2 pi i m / N); lanes
// t of nt share the work.
template <int N>
__device__ __forceinline__ void reg_tw_fill(float2* dst, const float2* __restrict__ tw, int t, int nt)
{
    using PL = RegPlan<N>;
    int off = 0;
    auto fill = [&](int LS, int R) {
        const int n = (R - 1) * LS;
        for (int i = t; i < n; i += nt) {
            const int r = 1 + i / LS, k = i % LS;
            dst[off + i] = tw[r * k * (N / (LS * R))];
        }
        off += n;
    };
    fill(PL::R0, PL::R1);
    if (PL::R2 > 1) fill(PL::R0 * PL::R1, PL::R2);
    if (PL::R3 > 1) fill(PL::R0 * PL::R1 * PL::R2, PL::R3);
}

template <int N, int E, int R, int LS, bool INV>
__device__ __forceinline__ void reg_stage(float2 (&v)[E], int j, const float2* __restrict__ tws)
{
    constexpr int T = N / E;
#pragma unroll
    for (int i = 0; i < E / R; ++i) {
        if (LS > 1) {
            const int k = (j + T * i) & (LS - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tws[(r - 1) * LS + k];
                if (INV) w.y = -w.y;
                v[i * R + r] = cmul(v[i * R + r], w);
            }
        }
        Dft<R, INV>::run(&v[i * R]);
    }
}

template <int N, int E, int R, int LS>
__device__ __forceinline__ void reg_write(float2* row, const float2 (&v)[E], int j, int p)
{
    constexpr int T = N / E;
#pragma unroll
    for (int i = 0; i < E / R; ++i) {
        const int b = j + T * i;
        const int k = b & (LS - 1);
        const int base = (b - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) row[swz(base + q * LS, p)] = v[i * R + q];
    }
}

template <int N, int E, int R>
__device__ __forceinline__ void reg_read(const float2* row, float2 (&v)[E], int j, int p)
{
    constexpr int T = N / E;
#pragma unroll
    for (int i = 0; i < E / R; ++i)
#pragma unroll
        for (int r = 0; r < R; ++r) v[i * R + r] = row[swz(j + T * i + r * (N / R), p)];
}

// reg_write and reg_read on the row at element offset rowoff from the dynamic
// LDS base lds (rowoff = p N for pencil p).
//   read: e = j | (T i + r N/R), T a power of two.
//   write, R and LS powers of two with LS <= T: b = j + T i, k = j & (LS-1),
//     e = (R (j & ~(LS-1)) + k) | (R T i + q LS);
//   write with b = j + T i < LS always (T E/R <= LS): e = j | (T i + q LS).
// Other stages (radix 3, 5, 6, 10) take reg_write's swz addressing.
template <int N, int E, int R, int LS>
__device__ __forceinline__ void reg_write_x(float2* lds, int rowoff, const float2 (&v)[E], int j, int p)
{
    constexpr int T = N / E;
    if constexpr (is_pow2(R) && is_pow2(LS) && is_pow2(T) && LS <= T) {
        constexpr int LB = lane_bits(R * T);
        static_assert(N % (1 << LB) == 0, "row offsets p N are multiples of 2^LB");
        const uint32_t lb = lane_bytes(rowoff, R * (j & ~(LS - 1)) + (j & (LS - 1)), p);
#pragma unroll
        for (int i = 0; i < E / R; ++i)
#pragma unroll
            for (int q = 0; q < R; ++q) lds_at<LB>(lds, lb, R * T * i + q * LS) = v[i * R + q];
    } else if constexpr (is_pow2(LS) && is_pow2(T) && T * (E / R) <= LS) {
        constexpr int LB = lane_bits(T);
        static_assert(N % (1 << LB) == 0, "row offsets p N are multiples of 2^LB");
        const uint32_t lb = lane_bytes(rowoff, j, p);
#pragma unroll
        for (int i = 0; i < E / R; ++i)
#pragma unroll
            for (int q = 0; q < R; ++q) lds_at<LB>(lds, lb, T * i + q * LS) = v[i * R + q];
    } else {
        reg_write<N, E, R, LS>(lds + rowoff, v, j, p);
    }
}

template <int N, int E, int R>
__device__ __forceinline__ void reg_read_x(float2* lds, int rowoff, float2 (&v)[E], int j, int p)
{
    constexpr int T = N / E, LB = lane_bits(T);
    static_assert(is_pow2(T) && N % (1 << LB) == 0, "lane part j, row offsets p N multiples of 2^LB");
    const uint32_t lb = lane_bytes(rowoff, j, p);
#pragma unroll
    for (int i = 0; i < E / R; ++i)
#pragma unroll
        for (int r = 0; r < R; ++r) v[i * R + r] = lds_at<LB>(lds, lb, T * i + r * (N / R));
}

// The natural-order store of a transform's outputs, element j + T i + q N/RL
// from register i RL + q (the output slots of the last stage of radix RL).
template <int N, int E, int RL>
__device__ __forceinline__ void reg_store_x(float2* lds, int rowoff, const float2 (&v)[E], int j, int p)
{
    constexpr int T = N / E, LB = lane_bits(T);
    static_assert(is_pow2(T) && N % (1 << LB) == 0, "lane part j, row offsets p N multiples of 2^LB");
    const uint32_t lb = lane_bytes(rowoff, j, p);
#pragma unroll
    for (int i = 0; i < E / RL; ++i)
#pragma unroll
        for (int q = 0; q < RL; ++q) lds_at<LB>(lds, lb, T * i + q * (N / RL)) = v[i * RL + q];
}

// In-register Stockham FFT of the pencils of a tile: pencil p's row lies at
// element offset rowoff = p N from the dynamic LDS base lds; twt is the stage
// table of reg_tw_fill (normally in LDS).
template <int N, bool INV>
__device__ __forceinline__ void reg_fft(float2 (&v)[RegPlan<N>::E], float2* lds, int rowoff, int j, int p,
                                        const float2* twt)
{
    using PL = RegPlan<N>;
    constexpr int E = PL::E, R0 = PL::R0, R1 = PL::R1, R2 = PL::R2, R3 = PL::R3;
    static_assert(R0 * R1 * R2 * R3 == N && E % R0 == 0 && E % R1 == 0 && E % R2 == 0 && E % R3 == 0,
                  "register FFT plan");
    static_assert(((R0 * R1) & (R0 * R1 - 1)) == 0 && (R3 == 1 || ((R0 * R1 * R2) & (R0 * R1 * R2 - 1)) == 0),
                  "the span of every stage but the last must be a power of two");
    reg_stage<N, E, R0, 1, INV>(v, j, twt);
    reg_write_x<N, E, R0, 1>(lds, rowoff, v, j, p);
    __syncthreads();
    reg_read_x<N, E, R1>(lds, rowoff, v, j, p);
    __syncthreads();
    reg_stage<N, E, R1, R0, INV>(v, j, twt);
    if constexpr (R2 > 1) {
        reg_write_x<N, E, R1, R0>(lds, rowoff, v, j, p);
        __syncthreads();
        reg_read_x<N, E, R2>(lds, rowoff, v, j, p);
        __syncthreads();
        reg_stage<N, E, R2, R0 * R1, INV>(v, j, twt + (R1 - 1) * R0);
    }
    if constexpr (R3 > 1) {
        reg_write_x<N, E, R2, R0 * R1>(lds, rowoff, v, j, p);
        __syncthreads();
        reg_read_x<N, E, R3>(lds, rowoff, v, j, p);
        __syncthreads();
        reg_stage<N, E, R3, R0 * R1 * R2, INV>(v, j, twt + (R1 - 1) * R0 + (R2 - 1) * R0 * R1);
    }
}

// Orders the LDS accesses of one wave: a pencil whose T lanes lie in one wave
// (T divides 64) exchanges its elements through LDS with no workgroup barrier.
// A wave's LDS instructions execute in order; the fences keep the compiler from
// moving its accesses across this point.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// reg_fft for pencils that each lie in one wave: the stage exchanges are
// ordered per wave, with no workgroup barrier.
template <int N, bool INV>
__device__ __forceinline__ void reg_fft_wave(float2 (&v)[RegPlan<N>::E], float2* lds, int rowoff, int j, int p,
                                             const float2* twt)
{
    using PL = RegPlan<N>;
    constexpr int E = PL::E, R0 = PL::R0, R1 = PL::R1, R2 = PL::R2, R3 = PL::R3;
    static_assert(64 % (N / E) == 0, "a pencil's lanes must lie in one wave");
    reg_stage<N, E, R0, 1, INV>(v, j, twt);
    reg_write_x<N, E, R0, 1>(lds, rowoff, v, j, p);
    wave_sync();
    reg_read_x<N, E, R1>(lds, rowoff, v, j, p);
    wave_sync();
    reg_stage<N, E, R1, R0, INV>(v, j, twt);
    if constexpr (R2 > 1) {
        reg_write_x<N, E, R1, R0>(lds, rowoff, v, j, p);
        wave_sync();
        reg_read_x<N, E, R2>(lds, rowoff, v, j, p);
        wave_sync();
        reg_stage<N, E, R2, R0 * R1, INV>(v, j, twt + (R1 - 1) * R0);
    }
    if constexpr (R3 > 1) {
        reg_write_x<N, E, R2, R0 * R1>(lds, rowoff, v, j, p);
        wave_sync();
        reg_read_x<N, E, R3>(lds, rowoff, v, j, p);
        wave_sync();
        reg_stage<N, E, R3, R0 * R1 * R2, INV>(v, j, twt + (R1 - 1) * R0 + (R2 - 1) * R0 * R1);
    }
}

}  // namespace fftc
}  // namespace gf
