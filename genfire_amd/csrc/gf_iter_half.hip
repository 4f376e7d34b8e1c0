// gf_iter_half.hip -- the GENFIRE / DR iteration on the half spectrum
// (reconstruct.m:126-165 with the object real).
//
// The iterate is real, so its spectrum is kept for kx in [0, n1/2] only (h =
// n1/2+1 planes).  real(ifftn(k')) of the replaced full spectrum k' equals the
// C2R transform of the half spectrum when
//   * for 0 < kx < n1/2 the stored value is the Hermitian part
//     (k'(s) + conj(k'(mate)))/2 of the voxel s and its unstored mate -s, which
//     is M(s) where both are measured with M(-s) = conj(M(s)) (all interior
//     voxels of a CentroSymmetricity grid), the average of the two measured
//     values where they differ (the +-n/2 lines the reference fills from its
//     extended half grid), and (M + k)/2 where only one of the pair is measured;
//   * the kx = 0 and kx = n1/2 planes hold both s and -s, are replaced
//     voxel by voxel, and contribute only their real part after the y and z
//     inverse transforms (the C2R takes Re of its first and last bins).
// errK's numerator counts |k(-s)| = |k(s)| for the mate, so each stored voxel
// carries up to two terms.  The rules are exact for any measuredK and mask.
//
// Five passes, each with one contiguous side and one side of 128-byte rows at a
// 4-KiB stride (no 2-MiB strides):
//   P1 x: real rows -> P_R / DR reflection -> R2C (complex FFT of n1/2 + split)
//         -> layout A  [y][kx][z]   (A[y + n2*(kx + h*z)])
//   P2 y: A pencils (contiguous) -> FFT -> layout B  [kx][z/16][ky][z%16]
//         (B[(z & 15) + 16*(ky + n2*((z >> 4) + (n3/16)*kx))]: a y tile's 16 z
//         of every ky are one contiguous block, a z pencil 128-byte runs)
//   P3 z: B pencils (contiguous) -> FFT -> errK + replacement -> IFFT -> B
//   P4 y: B (strided) -> IFFT -> A
//   P5 x: A (strided) -> Re of kx = 0, n1/2 -> C2R -> x (GENFIRE) or DR update
// Between two iterations P5 and the next P1 run as one pass (k_half_last_first).
//
// Box pruning.  With the support constraint on, P_R(x) vanishes outside the
// centred Dim1 x Dim2 x Dim1 box, and only the box part of P_F's output is ever
// used again (the next P_R, the monitor model, the returned P_R(x)).  So every
// pass skips real-space rows outside the box: P1/P5 only touch rows (y, z) in
// the box, P2 reads only y in the box and only z in the box, P3 reads and writes
// only z in the box, P4 writes only y in the box.  Frequencies are never pruned.
// GENFIRE's transform input P_R(x) lies in the box, so its passes compute the
// same numbers as without pruning.  DR's input 2 P_R(x) - x does not: there the
// iterate is carried in the form x_{k+1} = P_R(x_k) + d_k, d_k = real(ifftn(D_k))
// with D_k = k' - k the replacement delta (nonzero only on the measured-voxel
// lists), which follows from x_{k+1} = x_k + P_F(y_k) - P_R(x_k) and P_F(y_k) =
// y_k + d_k.  Then fftn(2 P_R(x_{k+1}) - x_{k+1}) = fftn(2 P_R(x_{k+1}) -
// P_R(x_k)) - fftn(d_k): a box-supported transform minus the Hermitian part D'_k
// of D_k (D_k itself off the kx = 0, n1/2 planes, where C2R then R2C is the
// identity; (D(s) + conj D(-s))/2 on them).  The z pass subtracts D' at the list
// entries, forms the new D there, and transforms D, not k', back.
#include <algorithm>
#include <cmath>

#include "gf_fft_core.h"
#include "gf_grid.h"

namespace gf {

namespace {

using namespace fftc;

constexpr int kHP = 16;  // pencils per workgroup: 16 x 8 bytes = one 128-byte row
constexpr int kZP = 8;   // pencils per workgroup of the fused z pass (contiguous pencils)

template <int N> constexpr int half_threads() { return kHP * (N / RegPlan<N>::E); }
// Rows per workgroup of the x passes (P1/P5: real rows of the M-point complex
// transform) and their workgroup size.
template <int M> constexpr int xrows() { return kHP; }
template <int M> constexpr int x_threads() { return xrows<M>() * (M / RegPlan<M>::E); }
template <int N> constexpr int zfused_threads() { return kZP * (N / RegPlan<N>::E); }

__device__ __forceinline__ float prh(float v, uint8_t s, int positivity, int support)
{
    if (positivity && v < 0.0f) v = 0.0f;   // x(x<0) = 0   (reconstruct.m:127)
    if (support) v = v * (float)s;          // x .* S       (:130)
    return v;
}

__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// FFT-order index s of an axis of length n lies in the centred box of extent b
// (signed offsets -b/2 .. b-b/2-1): s < b - b/2 or s >= n - b/2.
__host__ __device__ __forceinline__ bool inb(int s, int n, int b) { return s < b - b / 2 || s >= n - b / 2; }

// The tiles of ts consecutive indices along an axis of length n that meet the
// box: [0, lo_end) and [hi_begin, n/ts); all n/ts when the two runs meet.
struct BoxTiles {
    int lo_end, hi_begin, count;
};
__host__ __device__ __forceinline__ BoxTiles box_tiles(int n, int b, int ts)
{
    const int nt = n / ts, lo_end = (b - b / 2 + ts - 1) / ts, hi_begin = (n - b / 2) / ts;
    if (lo_end >= hi_begin) return {nt, nt, nt};
    return {lo_end, hi_begin, lo_end + (nt - hi_begin)};
}
__host__ __device__ __forceinline__ int box_tile(const BoxTiles& t, int i)
{
    return i < t.lo_end ? i : t.hi_begin + (i - t.lo_end);
}

// Tile of an x pass (P1/P5): one z in the box, 16 consecutive y rows that meet
// it; cy = the launch's y-tile count (the largest of its jobs).
template <int M>
__device__ __forceinline__ bool x_tile(const HalfIterArgs& a, int cy, int& i3, int& i2_0)
{
    constexpr int XR = xrows<M>();
    const BoxTiles ty = box_tiles(a.n2, a.by, XR), tz = box_tiles(a.n3, a.bz, 1);
    const int zi = (int)blockIdx.x / cy, yt = (int)blockIdx.x % cy;
    if (zi >= tz.count || yt >= ty.count) return false;
    i3 = box_tile(tz, zi);
    i2_0 = box_tile(ty, yt) * XR;
    return true;
}

// Tile of a y pass (P2/P4): one kx, 16 consecutive z that meet the box.
__device__ __forceinline__ bool y_tile(const HalfIterArgs& a, int cz, int& kx, int& i3_0)
{
    const BoxTiles tz = box_tiles(a.n3, a.bz, kHP);
    kx = (int)blockIdx.x / cz;
    const int zt = (int)blockIdx.x % cz;
    if (zt >= tz.count) return false;
    i3_0 = box_tile(tz, zt) * kHP;
    return true;
}

// Layout B of the y and z passes: plane kx holds n3/16 blocks of 16 z, each
// block [ky][z % 16].  A y tile (one kx, 16 z from a multiple of 16) writes
// and reads one contiguous block of n2 x 16 elements; a z pencil reads 16
// contiguous z per block (128 bytes), and the 8 pencils of a z-pass tile
// (consecutive ky) 1 KiB runs.  The z lengths with a register plan are
// multiples of 16.
static_assert(kHP == 16, "layout B blocks are a y tile's 16 z");
__device__ __forceinline__ float2* b_block(const HalfIterArgs& a, int kx, int z0)
{
    return a.B + (int64_t)a.n2 * a.n3 * kx + (int64_t)16 * a.n2 * (z0 >> 4);
}
// element (ky, z) of a plane of B
__device__ __forceinline__ int b_off(int ky, int z, int n2) { return (z & 15) + 16 * (ky + n2 * (z >> 4)); }

// Which elements of the real pair (xc, xc+1) of a row lie in the box (bit 0, 1).
__device__ __forceinline__ int pair_in(bool row_in, int xc, int n, int b)
{
    return row_in ? ((inb(xc, n, b) ? 1 : 0) | (inb(xc + 1, n, b) ? 2 : 0)) : 0;
}

// ---------------------------------------------------------------------------
// P1: P_R and the real-to-complex x transform.  N = n1, complex length M = N/2.
// Tile: 16 consecutive y rows at one z.  Lane (p, j), p = t / T, owns the
// element pairs (2m, 2m+1), m = j + T*s (s = 0..E-1), of row p as
// z[m] = x[2m] + i x[2m+1].  The forward FFT takes element s in register slot
// i*R0 + r with s = i + r*E/R0; the inverse FFT leaves element s in slot
// i*RL + q with s = i + q*E/RL (RL its last radix).
// ---------------------------------------------------------------------------
template <int M> __device__ __forceinline__ constexpr int in_slot(int s)
{
    return (s % (RegPlan<M>::E / RegPlan<M>::R0)) * RegPlan<M>::R0 + s / (RegPlan<M>::E / RegPlan<M>::R0);
}
template <int M> __device__ __forceinline__ constexpr int out_slot(int s)
{
    constexpr int RL = reg_last_radix<M>();
    return (s % (RegPlan<M>::E / RL)) * RL + s / (RegPlan<M>::E / RL);
}

// A zero the compiler cannot see through.  An LDS table address built with it
// keeps one transform's twiddle loads from being reused, and so kept live, by
// the next transform of the same kernel.
__device__ __forceinline__ int opaque_zero()
{
    int z = 0;
    asm volatile("" : "+v"(z));
    return z;
}

// Copies a table into LDS (the real split's twiddles).  The transforms read
// their twiddles from LDS stage tables (reg_tw_fill): as global loads they kept
// the texture address unit busy for most of a pass.  The first barrier inside
// the transform orders the copy.
__device__ __forceinline__ void tw_to_lds(float2* dst, const float2* __restrict__ src, int n, int t, int nt)
{
    for (int i = t; i < n; i += nt) dst[i] = src[i];
}

// threadIdx.x behind an empty asm statement.  Each phase of a fused pass takes
// its lane indices from this, so the compiler recomputes the phase's addresses
// instead of keeping those of an earlier phase live across the whole kernel.
__device__ __forceinline__ int phase_tid()
{
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Forward M-point FFT of the tile's rows (v in input slots; a row's lanes lie in
// one wave and the twiddle tables are filled and ordered by the caller), the R2C split
//   X[k] = (Z[k] + conj Z[M-k])/2 + W_N^k (Z[k] - conj Z[M-k])/(2i), k = 0..M,
// and the transposed store to layout A.
template <int N>
__device__ __forceinline__ void r2c_store(const HalfIterArgs& a, float2* lds, float2 (&v)[RegPlan<N / 2>::E],
                                          int i2_0, int i3, const float2* __restrict__ twM,
                                          const float2* __restrict__ twN)
{
    constexpr int M = N / 2;
    using PL = RegPlan<M>;
    constexpr int E = PL::E, T = M / E, RL = reg_last_radix<M>();
    const int t = phase_tid(), j = t % T, p = t / T;
    const int n2 = a.n2, h = M + 1;
    reg_fft_wave<M, false>(v, lds, p * M, j, p, twM);
    reg_store_x<M, E, RL>(lds, p * M, v, j, p);
    __syncthreads();
    constexpr int XR = xrows<M>();
    const int pp = t % XR, c = t / XR;
    const float2* rp = lds + pp * M;
    float2* out = a.A + (i2_0 + pp) + (int64_t)n2 * h * i3;
    if (!inb(i2_0 + pp, n2, a.by)) return;   // a row outside the box is zero and is never read
    auto split = [&](int kx) {
        const float2 zk = rp[swz(kx == M ? 0 : kx, pp)];          // Z[k mod M]
        const float2 zm = conjf2(rp[swz(kx == 0 ? 0 : M - kx, pp)]);  // conj Z[(M - k) mod M]
        const float2 ev = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y + zm.y));
        const float2 od = cscale(mul_mi<false>(csub(zk, zm)), 0.5f);
        return cadd(ev, cmul(twN[kx], od));
    };
    float2 X[M / T];
#pragma unroll
    for (int q = 0; q < M / T; ++q) X[q] = split(c + q * T);
#pragma unroll
    for (int q = 0; q < M / T; ++q) out[(int64_t)n2 * (c + q * T)] = X[q];
    if (c == 0) out[(int64_t)n2 * M] = split(M);   // Nyquist bin
}

// The value P1 transforms for the real pair xv: P_R(x) (GENFIRE) or the
// reflection 2 P_R(x) - x (DR); P_R(x) also goes to best when that is set.
__device__ __forceinline__ float2 p1_value(const HalfIterArgs& a, float2 xv, uint16_t s2, int64_t ad, bool wr = true)
{
    const float ra = prh(xv.x, (uint8_t)(s2 & 0xFF), a.positivity, a.support);
    const float rb = prh(xv.y, (uint8_t)(s2 >> 8), a.positivity, a.support);
    if (a.best && wr) *reinterpret_cast<float2*>(a.best + ad) = make_float2(ra, rb);
    return a.dr ? make_float2(2.0f * ra - xv.x, 2.0f * rb - xv.y) : make_float2(ra, rb);
}

// Loads of the real pair at ad restricted to the box: x (when needed) and the
// support bytes; elements outside the box read as zero.
__device__ __forceinline__ void load_pair(const HalfIterArgs& a, int64_t ad, int in, bool need_x, float2& xv,
                                          uint16_t& s2)
{
    xv = make_float2(0.0f, 0.0f);
    s2 = 0;
    if (!in) return;
    if (need_x) {
        xv = *reinterpret_cast<const float2*>(a.x + ad);
        if (!(in & 1)) xv.x = 0.0f;
        if (!(in & 2)) xv.y = 0.0f;
    }
    s2 = *reinterpret_cast<const uint16_t*>(a.S + ad);
}

template <int N>
__global__ __launch_bounds__(x_threads<N / 2>()) void k_half_first(HalfBatch hb, int cy,
                                                                      const float2* __restrict__ twM,
                                                                      const float2* __restrict__ twN)
{
    const HalfIterArgs& a = hb.job[blockIdx.y];   // the job of this launch row
    constexpr int M = N / 2;
    using PL = RegPlan<M>;
    constexpr int E = PL::E, T = M / E;
    extern __shared__ float2 lds[];
    const int t = threadIdx.x, j = t % T, p = t / T;
    const int n2 = a.n2;
    int i3, i2_0;
    if (!x_tile<M>(a, cy, i3, i2_0)) return;
    float2* twlM = lds + xrows<M>() * (M + 1);
    float2* twlN = twlM + reg_tw_size<M>();
    reg_tw_fill<M>(twlM, twM, t, x_threads<M>());
    tw_to_lds(twlN, twN, M + 1, t, x_threads<M>());
    const int64_t row = (int64_t)N * ((i2_0 + p) + (int64_t)n2 * i3);
    const bool row_in = inb(i2_0 + p, n2, a.by);

    // All loads are issued before any store (best may be written), so they are
    // in flight together.
    float2 xv[E];
    uint16_t s2[E];
    int in[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        in[s] = pair_in(row_in, 2 * (j + T * s), N, a.bx);
        load_pair(a, row + 2 * (j + T * s), in[s], true, xv[s], s2[s]);
    }
    float2 v[E];
#pragma unroll
    for (int s = 0; s < E; ++s) v[in_slot<M>(s)] = p1_value(a, xv[s], s2[s], row + 2 * (j + T * s), in[s] != 0);
    __syncthreads();   // the twiddle tables
    r2c_store<N>(a, lds, v, i2_0, i3, twlM, twlN);
}

// ---------------------------------------------------------------------------
// P2: y forward, A -> B.  Tile: one kx, 16 consecutive z.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(half_threads<N>()) void k_half_y_fwd(HalfBatch hb, int cz, const float2* __restrict__ tw)
{
    const HalfIterArgs& a = hb.job[blockIdx.y];   // the job of this launch row
    using PL = RegPlan<N>;
    constexpr int E = PL::E, T = N / E, R0 = PL::R0, RL = reg_last_radix<N>();
    extern __shared__ float2 lds[];
    const int t = threadIdx.x, j = t % T, p = t / T;
    const int n2 = N, n3 = a.n3, h = a.n1 / 2 + 1;
    int kx, i3_0;
    if (!y_tile(a, cz, kx, i3_0)) return;
    float2* twl = lds + kHP * N;
    reg_tw_fill<N>(twl, tw, t, half_threads<N>());
    const float2* src = a.A + (int64_t)n2 * (kx + (int64_t)h * (i3_0 + p));
    const bool z_in = inb(i3_0 + p, n3, a.bz);
    float2 v[E];
#pragma unroll
    for (int i = 0; i < E / R0; ++i)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int y = j + T * i + r * (N / R0);
            v[i * R0 + r] = (z_in && inb(y, N, a.by)) ? src[y] : make_float2(0.0f, 0.0f);
        }
    __syncthreads();   // the twiddle table; the pencils' transforms are wave-local
    reg_fft_wave<N, false>(v, lds, p * N, j, p, twl);
    reg_store_x<N, E, RL>(lds, p * N, v, j, p);
    __syncthreads();
    const int pp = t % kHP, c = t / kHP;
    float2* dst = b_block(a, kx, i3_0) + pp;   // the tile's contiguous block of B
    float2 w[N / T];
    reg_read_x<N, E, E>(lds, pp * N, w, c, pp);   // column c of pencil pp: elements c + q T
#pragma unroll
    for (int q = 0; q < N / T; ++q) dst[kHP * (c + q * T)] = w[q];
}

// ---------------------------------------------------------------------------
// P3: z forward, errK, replacement, z inverse, in place on B.  Tile: one kx,
// 8 consecutive ky (8 contiguous pencils; the kernel holds two transforms' worth
// of registers, so smaller workgroups keep three of them resident per CU).
// ---------------------------------------------------------------------------
// The box is exactly half of each z pencil (bz = N/2, the usual oversampling
// ratio of 2) and the plan's first and last radices are at least 4: then the
// rows that lie outside the box are whole radix slots, known at compile time,
// and the compiler drops the zero inputs' arithmetic in the first forward stage
// and the unused outputs' in the last inverse stage.
template <int N> constexpr bool zhalf_ok() { return RegPlan<N>::R0 >= 4 && reg_last_radix<N>() >= 4 && RegPlan<N>::R0 % 4 == 0 && reg_last_radix<N>() % 4 == 0; }
// slot r of a radix-R stage covers z in [r N/R, (r+1) N/R): inside the half box
// [0, N/4) u [3N/4, N) when r < R/4 or r >= 3R/4
template <int R> __device__ __forceinline__ constexpr bool half_slot(int r) { return r < R / 4 || r >= 3 * R / 4; }

// P3: the z pencils of layout B, contiguous.  A workgroup takes kZP pencils (one
// kx, kZP consecutive ky); the T lanes of a pencil lie in one wave (T divides 64),
// so each pencil's transforms, its list entries (errK terms, replacement, DR
// delta) and its LDS row are handled by its own lanes with wave-ordered LDS
// exchanges: the only workgroup barrier orders the shared twiddle table.  The
// errK partials are one per wave, err_part[blockIdx.x * W + w].
template <int N, bool HALF>
__global__ __launch_bounds__(zfused_threads<N>()) void k_half_z_fused(HalfBatch hb, const float2* __restrict__ tw)
{
    const HalfIterArgs& a = hb.job[blockIdx.y];   // the job of this launch row
    using PL = RegPlan<N>;
    constexpr int E = PL::E, T = N / E, R0 = PL::R0, RL = reg_last_radix<N>();
    constexpr int NT = zfused_threads<N>(), W = NT / 64;
    static_assert(64 % T == 0 && NT % 64 == 0, "a pencil's lanes lie in one wave");
    extern __shared__ float2 lds[];
    const int t = threadIdx.x, j = t % T, p = t / T;
    const int n2 = a.n2, n3 = N;
    const int tpr = n2 / kZP;
    const int kx = blockIdx.x / tpr, ky_0 = (blockIdx.x % tpr) * kZP;
    // LDS: the tile, then the twiddle table
    float2* twl = lds + kZP * N;
    reg_tw_fill<N>(twl, tw, t, NT);
    const float2* bk = a.B + (int64_t)n2 * n3 * kx;   // plane kx of B
    float2 v[E];
#pragma unroll
    for (int i = 0; i < E / R0; ++i)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int z = j + T * i + r * (N / R0);
            v[i * R0 + r] = (HALF ? half_slot<R0>(r) : inb(z, N, a.bz)) ? bk[b_off(ky_0 + p, z, n2)]
                                                                        : make_float2(0.0f, 0.0f);
        }
    // The pencil's list range and each lane's first kPre entries (e0, e0 + T)
    // are loaded before the forward transform, which hides their latency; most
    // pencils have fewer entries than 2 T, and later rounds load as they go.
    const int64_t rp = (int64_t)(ky_0 + p) + (int64_t)n2 * kx;
    const int beg = (int)a.csr_off[rp], end = (int)a.csr_off[rp + 1];   // < 2^31 entries
    const int e0 = beg + j;
    constexpr int kPre = 2;
    uint32_t mq[kPre];
    float2 mgq[kPre], Vq[kPre], dq[kPre], dnq[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        mq[u] = 0;
        mgq[u] = Vq[u] = dq[u] = dnq[u] = make_float2(0.0f, 0.0f);
        const int e = e0 + u * T;
        if (e < end) {
            mq[u] = a.csr_meta[e];
            mgq[u] = a.csr_mag[e];
            Vq[u] = a.csr_val[e];
            if (a.dmode) dq[u] = a.csr_d[e];
        }
    }
    __syncthreads();   // the twiddle table
    float2* rowl = lds + p * N;
    reg_fft_wave<N, false>(v, lds, p * N, j, p, twl);
    reg_store_x<N, E, RL>(lds, p * N, v, j, p);
    wave_sync();
    // DR on the box: the transform of the box part minus D' of the last
    // iteration is the transform of the iterate's reflection.  The Rfree terms
    // read it from LDS; otherwise the errK loop subtracts in registers.
    if (a.dmode && a.rf_off) {
        auto sub_d = [&](uint32_t m, float2 dv) {
            float2* slot = rowl + swz((int)(m & 0x7FFu), p);
            const float2 kv = *slot;
            *slot = make_float2(kv.x - dv.x, kv.y - dv.y);
        };
#pragma unroll
        for (int u = 0; u < kPre; ++u)
            if (e0 + u * T < end) sub_d(mq[u], dq[u]);
        for (int e = e0 + kPre * T; e < end; e += T) sub_d(a.csr_meta[e], a.csr_d[e]);
        wave_sync();
    }
    // Rfree terms (reconstruct.m:137-147) on the transform before replacement
    if (a.rf_off) {
        const int rb = (int)a.rf_off[rp], re = (int)a.rf_off[rp + 1];
        for (int e = rb + j; e < re; e += T) {
            const uint32_t m = a.rf_meta[e];
            const float2 kv = rowl[swz((int)(m & 0x7FFu), p)];
            const float2 V = a.rf_val[e];
            const float dx = kv.x - V.x, dy = kv.y - V.y;
            a.rf_terms[e] = sqrtf(dx * dx + dy * dy);
        }
        wave_sync();
    }
    // errK numerator (:135) and replacement (:163) from the half-spectrum lists;
    // the numerator is reduced here, per wave, so that it is not live through the
    // inverse transform
    {
        float err = 0.0f;   // at most a few entries per lane
        // one entry: returns the new delta D = k' - k (DR)
        auto entry = [&](int e, uint32_t m, float2 mg, float2 V, float2 dv) {
            float2* slot = rowl + swz((int)(m & 0x7FFu), p);
            float2 kv = *slot;
            if (a.dmode && !a.rf_off) kv = make_float2(kv.x - dv.x, kv.y - dv.y);
            const float ak = sqrtf(kv.x * kv.x + kv.y * kv.y);
            if (m & kHalfTermA) err += fabsf(ak - mg.x);
            if (m & kHalfTermB) err += fabsf(ak - mg.y);
            float2 kn = kv;
            if (m & kHalfEnforced)
                kn = (m & kHalfAverage) ? make_float2(0.5f * (V.x + kv.x), 0.5f * (V.y + kv.y)) : V;
            float2 dn = make_float2(0.0f, 0.0f);
            if (a.dmode) {
                dn = make_float2(kn.x - kv.x, kn.y - kv.y);
                a.csr_d[e] = dn;
            } else if (m & kHalfEnforced) {
                *slot = kn;
            }
            return dn;
        };
#pragma unroll
        for (int u = 0; u < kPre; ++u) {
            const int e = e0 + u * T;
            if (e < end) dnq[u] = entry(e, mq[u], mgq[u], Vq[u], dq[u]);
        }
        for (int e = e0 + kPre * T; e < end; e += T) {
            float2 dv = make_float2(0.0f, 0.0f);
            if (a.dmode && !a.rf_off) dv = a.csr_d[e];
            entry(e, a.csr_meta[e], a.csr_mag[e], a.csr_val[e], dv);
        }
        double de = (double)err;
        for (int off = 32; off > 0; off >>= 1) de += __shfl_down(de, off, 64);
        if ((t & 63) == 0) a.err_part[(int64_t)blockIdx.x * W + (t >> 6)] = de;
    }
    // DR on the box: the inverse transform takes the sparse D
    if (a.dmode) {
        for (int i = j; i < N; i += T) rowl[i] = make_float2(0.0f, 0.0f);
        wave_sync();
#pragma unroll
        for (int u = 0; u < kPre; ++u)
            if (e0 + u * T < end) rowl[swz((int)(mq[u] & 0x7FFu), p)] = dnq[u];
        for (int e = e0 + kPre * T; e < end; e += T) rowl[swz((int)(a.csr_meta[e] & 0x7FFu), p)] = a.csr_d[e];
    }
    wave_sync();
    // the inverse transform takes fresh lane indices and its own address of the
    // twiddle table (opaque_zero) so that nothing of the forward one stays live
    const int t2 = phase_tid(), j2 = t2 % T, p2 = t2 / T;
    reg_read_x<N, E, R0>(lds, p2 * N, v, j2, p2);
    wave_sync();
    reg_fft_wave<N, true>(v, lds, p2 * N, j2, p2, lds + (kZP * N + opaque_zero()));
    const float sc = 1.0f / (float)N;
    float2* bk2 = a.B + (int64_t)n2 * n3 * kx;
#pragma unroll
    for (int i = 0; i < E / RL; ++i)
#pragma unroll
        for (int q = 0; q < RL; ++q) {
            const int z = j2 + T * i + q * (N / RL);
            if (HALF ? half_slot<RL>(q) : inb(z, N, a.bz)) bk2[b_off(ky_0 + p2, z, n2)] = cscale(v[i * RL + q], sc);
        }
}

// ---------------------------------------------------------------------------
// P4: y inverse, B -> A.  Tile: one kx, 16 consecutive z (strided read).
// ---------------------------------------------------------------------------
// errK of the iteration, stage 1 (y-inverse launch): chunk c of the z pass's
// partials summed into err_chunk[c], lane sums over a fixed stride, wave sums,
// then the waves in order.
__device__ void errk_chunks(const HalfIterArgs& a, double* red)
{
    const int t = threadIdx.x, nt = blockDim.x;
    const int64_t len = (a.err_n + kErrChunks - 1) / kErrChunks;
    for (int c = blockIdx.x; c < kErrChunks; c += gridDim.x) {
        const int64_t b = (int64_t)c * len, e = min(a.err_n, b + len);
        double s = 0.0;
        for (int64_t i = b + t; i < e; i += nt) s += a.err_part[i];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
        if ((t & 63) == 0) red[t >> 6] = s;
        __syncthreads();
        if (t == 0) {
            double sum = 0.0;
            for (int w = 0; w < nt / 64; ++w) sum += red[w];
            a.err_chunk[c] = sum;
        }
        __syncthreads();   // red is the tile's LDS
    }
}

// errK of the iteration, stage 2 (the last x pass's launch, one lane).
__device__ __forceinline__ void errk_final(const HalfIterArgs& a)
{
    if (a.err_stage != 2 || blockIdx.x != 0 || threadIdx.x != 0) return;
    double sum = 0.0;
    for (int c = 0; c < kErrChunks; ++c) sum += a.err_chunk[c];
    a.errk[*a.errctr] = (float)(sum / a.err_den);
    *a.errctr += 1;
}

// D on the kx = 0 and n1/2 planes -> its Hermitian part (the mate of entry e on
// the same plane is csr_mate; each pair is handled from its lower entry).
// Workgroups b of nb take the entries in a grid stride.
__device__ void dherm_part(const HalfIterArgs& a, int b, int nb)
{
    const int64_t n0 = a.csr_off[a.n2];
    const int64_t mb = a.csr_off[(int64_t)(a.n1 / 2) * a.n2], me = a.csr_off[(int64_t)(a.n1 / 2 + 1) * a.n2];
    const int64_t cnt = n0 + (me - mb);
    for (int64_t i = (int64_t)b * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)nb * blockDim.x) {
        const int64_t e = i < n0 ? i : mb + (i - n0);
        const int64_t m = a.csr_mate[i];
        if (m < e) continue;   // the pair is handled from its lower entry (or has no mate)
        const float2 de = a.csr_d[e];
        if (m == e) {
            a.csr_d[e] = make_float2(de.x, 0.0f);
            continue;
        }
        const float2 dm = a.csr_d[m];
        a.csr_d[e] = make_float2(0.5f * (de.x + dm.x), 0.5f * (de.y - dm.y));
        a.csr_d[m] = make_float2(0.5f * (dm.x + de.x), 0.5f * (dm.y - de.y));
    }
}

template <int N>
__global__ __launch_bounds__(half_threads<N>()) void k_half_y_inv(HalfBatch hb, int cz, const float2* __restrict__ tw)
{
    const HalfIterArgs& a = hb.job[blockIdx.y];   // the job of this launch row
    using PL = RegPlan<N>;
    constexpr int E = PL::E, T = N / E, R0 = PL::R0, RL = reg_last_radix<N>();
    extern __shared__ float2 lds[];
    // the iteration's errK chunks and D's Hermitian part: both need the whole z
    // pass, which this launch follows, and nothing before the next z pass reads them
    if (a.err_stage == 1) errk_chunks(a, reinterpret_cast<double*>(lds));
    if (a.dherm) {
        const int nb = min(16, (int)gridDim.x);
        if ((int)blockIdx.x < nb) dherm_part(a, (int)blockIdx.x, nb);
    }
    const int t = threadIdx.x, j = t % T, p = t / T;
    const int n2 = N, n3 = a.n3, h = a.n1 / 2 + 1;
    int kx, i3_0;
    if (!y_tile(a, cz, kx, i3_0)) return;
    float2* twl = lds + kHP * N;
    reg_tw_fill<N>(twl, tw, t, half_threads<N>());
    {
        const int pp = t % kHP, c = t / kHP;
        const float2* src = b_block(a, kx, i3_0) + pp;
        const bool z_in = inb(i3_0 + pp, n3, a.bz);
        float2 w[N / T];
#pragma unroll
        for (int q = 0; q < N / T; ++q) w[q] = z_in ? src[kHP * (c + q * T)] : make_float2(0.0f, 0.0f);
        reg_store_x<N, E, E>(lds, pp * N, w, c, pp);   // column c of pencil pp: elements c + q T
    }
    __syncthreads();
    const int t2 = phase_tid(), j2 = t2 % T, p2 = t2 / T;
    float2 v[E];
    reg_read_x<N, E, R0>(lds, p2 * N, v, j2, p2);
    wave_sync();
    reg_fft_wave<N, true>(v, lds, p2 * N, j2, p2, twl);
    float2* dst = a.A + (int64_t)n2 * (kx + (int64_t)h * (i3_0 + p2));
    const float sc = 1.0f / (float)N;
    if (!inb(i3_0 + p2, n3, a.bz)) return;
#pragma unroll
    for (int i = 0; i < E / RL; ++i)
#pragma unroll
        for (int q = 0; q < RL; ++q) {
            const int y = j2 + T * i + q * (N / RL);
            if (inb(y, N, a.by)) dst[y] = cscale(v[i * RL + q], sc);
        }
}

// ---------------------------------------------------------------------------
// P5: Re of the kx = 0 and kx = M bins, complex-to-real x transform, and the
// iterate update.  Tile as P1.  The C2R split is
//   Z[k] = (X[k] + conj X[M-k])/2 + i W_N^-k (X[k] - conj X[M-k])/2,
//   z = IFFT_M(Z) / M,  x[2m] = Re z[m], x[2m+1] = Im z[m].
// Leaves the unscaled z in output slots of v.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void load_c2r(const HalfIterArgs& a, float2* lds, float2 (&v)[RegPlan<N / 2>::E], int i2_0,
                                         int i3, const float2* __restrict__ twM, const float2* __restrict__ twN)
{
    constexpr int M = N / 2;
    using PL = RegPlan<M>;
    constexpr int E = PL::E, T = M / E, R0 = PL::R0;
    constexpr int XR = xrows<M>();
    float2* nyq = lds + XR * M;    // X[M] of each pencil
    const int t = phase_tid(), j = t % T, p = t / T;
    const int n2 = a.n2, h = M + 1;
    {
        const int pp = t % XR, c = t / XR;
        const float2* src = a.A + (i2_0 + pp) + (int64_t)n2 * h * i3;
        const bool y_in = inb(i2_0 + pp, n2, a.by);   // rows outside the box are not formed
        float2 w[M / T];
#pragma unroll
        for (int q = 0; q < M / T; ++q) w[q] = y_in ? src[(int64_t)n2 * (c + q * T)] : make_float2(0.0f, 0.0f);
        if (c == 0) w[0].y = 0.0f;                 // Re of the kx = 0 bin
        reg_store_x<M, E, E>(lds, pp * M, w, c, pp);   // column c of pencil pp: elements c + q T
        if (c == 0) nyq[pp] = make_float2(y_in ? src[(int64_t)n2 * M].x : 0.0f, 0.0f);   // Re of kx = M
    }
    __syncthreads();
    float2* rowl = lds + p * M;
    constexpr int LB = lane_bits(T);
    const uint32_t lbj = lane_bytes(p * M, j, p);
#pragma unroll
    for (int i = 0; i < E / R0; ++i)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int e = j + T * i + r * (M / R0);
            const float2 xk = lds_at<LB>(lds, lbj, T * i + r * (M / R0));   // element e
            const float2 xm = conjf2(e == 0 ? nyq[p] : rowl[swz(M - e, p)]);
            const float2 ev = make_float2(0.5f * (xk.x + xm.x), 0.5f * (xk.y + xm.y));
            const float2 od = cscale(cmul(csub(xk, xm), conjf2(twN[e])), 0.5f);
            v[i * R0 + r] = make_float2(ev.x - od.y, ev.y + od.x);   // ev + i*od
        }
    wave_sync();   // the row's lanes lie in one wave
    reg_fft_wave<M, true>(v, lds, p * M, j, p, twM);
}

// The iterate update of the real pair at ad from pf = real(ifftn(k')) (:165):
// x = pf (GENFIRE) or x = x + pf - P_R(x) (DR).  Returns the new pair; stores
// it when store_x is set.
__device__ __forceinline__ float2 p5_update(const HalfIterArgs& a, float2 pf, float2 xv, uint16_t s2, int64_t ad,
                                           bool store_x)
{
    if (a.pf_out) *reinterpret_cast<float2*>(a.pf_out + ad) = pf;
    float2 xn = pf;
    if (a.dr) {
        const float ra = prh(xv.x, (uint8_t)(s2 & 0xFF), a.positivity, a.support);
        const float rb = prh(xv.y, (uint8_t)(s2 >> 8), a.positivity, a.support);
        xn = make_float2((xv.x + pf.x) - ra, (xv.y + pf.y) - rb);
    }
    if (store_x) *reinterpret_cast<float2*>(a.x + ad) = xn;
    return xn;
}

// DR on the box, one real pair: d = real(ifftn(D)) of the last replacement,
// r = P_R(x_k), x_{k+1} = r + d, pf = P_F(y_k) = (2 r - x_k) + d.  Returns
// x_{k+1}; r goes to *rk.
__device__ __forceinline__ float2 p5_dmode(const HalfIterArgs& a, float2 d, float2 xv, uint16_t s2, int64_t ad,
                                          float2* rk)
{
    const float ra = prh(xv.x, (uint8_t)(s2 & 0xFF), a.positivity, a.support);
    const float rb = prh(xv.y, (uint8_t)(s2 >> 8), a.positivity, a.support);
    if (a.pf_out)
        *reinterpret_cast<float2*>(a.pf_out + ad) = make_float2((2.0f * ra - xv.x) + d.x, (2.0f * rb - xv.y) + d.y);
    const float2 xn = make_float2(ra + d.x, rb + d.y);
    *reinterpret_cast<float2*>(a.x + ad) = xn;
    *rk = make_float2(ra, rb);
    return xn;
}

template <int N>
__global__ __launch_bounds__(x_threads<N / 2>()) void k_half_last(HalfBatch hb, int cy,
                                                                     const float2* __restrict__ twM,
                                                                     const float2* __restrict__ twN)
{
    const HalfIterArgs& a = hb.job[blockIdx.y];   // the job of this launch row
    constexpr int M = N / 2;
    using PL = RegPlan<M>;
    constexpr int E = PL::E, T = M / E;
    extern __shared__ float2 lds[];
    const int n2 = a.n2;
    int i3, i2_0;
    errk_final(a);
    if (!x_tile<M>(a, cy, i3, i2_0)) return;
    float2* twlM = lds + xrows<M>() * (M + 1);
    float2* twlN = twlM + reg_tw_size<M>();
    reg_tw_fill<M>(twlM, twM, (int)threadIdx.x, x_threads<M>());
    tw_to_lds(twlN, twN, M + 1, (int)threadIdx.x, x_threads<M>());
    float2 v[E];
    load_c2r<N>(a, lds, v, i2_0, i3, twlM, twlN);
    const int t = phase_tid(), j = t % T, p = t / T;
    const float sc = 1.0f / (float)M;
    const int64_t row = (int64_t)N * ((i2_0 + p) + (int64_t)n2 * i3);
    const bool row_in = inb(i2_0 + p, n2, a.by);
    float2 xv[E];
    uint16_t s2[E];
    int in[E];
#pragma unroll
    for (int s = 0; s < E; ++s) {
        in[s] = pair_in(row_in, 2 * (j + T * s), N, a.bx);
        load_pair(a, row + 2 * (j + T * s), a.dr ? in[s] : 0, true, xv[s], s2[s]);
    }
#pragma unroll
    for (int s = 0; s < E; ++s) {
        if (!in[s]) continue;
        const float2 d = cscale(v[out_slot<M>(s)], sc);
        if (a.dmode) {
            float2 rk;
            p5_dmode(a, d, xv[s], s2[s], row + 2 * (j + T * s), &rk);
        } else {
            p5_update(a, d, xv[s], s2[s], row + 2 * (j + T * s), true);
        }
    }
}

// P5 of one iteration fused with P1 of the next: the C2R output updates the
// iterate in registers, which is transformed again without a round trip of x
// through memory.  GENFIRE's x is the C2R output itself and is not stored
// (pf_out carries it where the R-factor needs it); DR stores x.
//
// The forward transform reads the LDS twiddle tables through addresses built
// with opaque_zero(): with the same addresses the compiler keeps the inverse
// transform's twiddle loads live for reuse by the forward one, which doubles the
// register count and halves occupancy.
template <int N>
__global__ __launch_bounds__(x_threads<N / 2>()) void k_half_last_first(HalfBatch hb, int cy,
                                                                           const float2* __restrict__ twM,
                                                                           const float2* __restrict__ twN)
{
    const HalfIterArgs& a = hb.job[blockIdx.y];   // the job of this launch row
    constexpr int M = N / 2;
    using PL = RegPlan<M>;
    constexpr int E = PL::E, T = M / E;
    extern __shared__ float2 lds[];
    const int n2 = a.n2;
    int i3, i2_0;
    errk_final(a);
    if (!x_tile<M>(a, cy, i3, i2_0)) return;
    float2* twlM = lds + xrows<M>() * (M + 1);
    float2* twlN = twlM + reg_tw_size<M>();
    reg_tw_fill<M>(twlM, twM, (int)threadIdx.x, x_threads<M>());
    tw_to_lds(twlN, twN, M + 1, (int)threadIdx.x, x_threads<M>());
    float2 v[E];
    load_c2r<N>(a, lds, v, i2_0, i3, twlM, twlN);
    const float sc = 1.0f / (float)M;
    const int tu = phase_tid(), ju = tu % T, pu = tu / T;
    const int64_t row = (int64_t)N * ((i2_0 + pu) + (int64_t)n2 * i3) + 2 * ju;
    const bool row_in = inb(i2_0 + pu, n2, a.by);
    // The update runs in two halves of E/2 pairs, each loading its iterate and
    // support pairs just before use (the scheduling barrier keeps the second
    // half's loads from being hoisted above the first half's arithmetic):
    // 56 VGPRs instead of 70 at N = 512, four workgroups per CU instead of
    // three, 110.5 -> 98.2 us per launch on 512^3 with the 256^3 box.
    float2 w[E];
    constexpr int H = E / 2;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        float2 xv[H];
        uint16_t s2[H];
        int in[H];
#pragma unroll
        for (int q = 0; q < H; ++q) {
            const int s = hh * H + q;
            in[q] = pair_in(row_in, 2 * (ju + T * s), N, a.bx);
            load_pair(a, row + 2 * T * s, in[q], a.dr != 0, xv[q], s2[q]);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < H; ++q) {
            const int s = hh * H + q;
            const int64_t ad = row + 2 * T * s;
            float2 u = make_float2(0.0f, 0.0f);
            if (in[q]) {
                const float2 d = cscale(v[out_slot<M>(s)], sc);
                if (a.dmode) {
                    // next transform input: 2 P_R(x_{k+1}) - P_R(x_k)
                    float2 rk;
                    const float2 xn = p5_dmode(a, d, xv[q], s2[q], ad, &rk);
                    const float2 rn = make_float2(prh(xn.x, (uint8_t)(s2[q] & 0xFF), a.positivity, a.support),
                                                  prh(xn.y, (uint8_t)(s2[q] >> 8), a.positivity, a.support));
                    if (a.best) *reinterpret_cast<float2*>(a.best + ad) = rn;
                    u = make_float2(2.0f * rn.x - rk.x, 2.0f * rn.y - rk.y);
                } else {
                    const float2 xn = p5_update(a, d, xv[q], s2[q], ad, a.dr != 0);
                    u = p1_value(a, xn, s2[q], ad);
                }
            }
            w[in_slot<M>(s)] = u;
        }
    }
#pragma unroll
    for (int s = 0; s < E; ++s) v[s] = w[s];
    const int z0 = opaque_zero();
    r2c_store<N>(a, lds, v, i2_0, i3, lds + (xrows<M>() * (M + 1) + z0),
                 lds + (xrows<M>() * (M + 1) + reg_tw_size<M>() + z0));
}

// ---------------------------------------------------------------------------
// Half-spectrum measured-voxel lists (pencil id ky + n2*kx, z ascending).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool hmeas(const float2* M, const uint8_t* mask, int64_t s)
{
    const float2 m = M[s];
    return (m.x != 0.0f || m.y != 0.0f) && (mask == nullptr || mask[s] != 0);
}

struct HalfVox {
    int64_t s, mate;
    bool self_paired;
};

__device__ __forceinline__ HalfVox half_vox(int kx, int ky, int z, int n1, int n2, int n3)
{
    HalfVox v;
    v.s = (int64_t)kx + (int64_t)n1 * (ky + (int64_t)n2 * z);
    const int mx = (n1 - kx) % n1, my = (n2 - ky) % n2, mz = (n3 - z) % n3;
    v.mate = (int64_t)mx + (int64_t)n1 * (my + (int64_t)n2 * mz);
    v.self_paired = (kx == 0) || (2 * kx == n1);
    return v;
}

// The lists are built per (pencil, z chunk): kHcsrChunks chunks of a pencil's z
// range, counted, scanned in (pencil, chunk) order and filled, so a pencil's
// entries stay in z order while four times as many lanes walk the grid.
constexpr int kHcsrChunks = 4;

__global__ void k_hcsr_count(const float2* __restrict__ M, const uint8_t* __restrict__ mask, int n1, int n2,
                             int n3, uint32_t* __restrict__ cnt)
{
    const int h = n1 / 2 + 1;
    const int64_t np = (int64_t)h * n2;
    // lanes take consecutive kx (contiguous voxels of M and of the mates); the
    // count goes to slot (ky + n2*kx) * kHcsrChunks + chunk
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < np * kHcsrChunks;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int kx = (int)(q % h), ky = (int)((q / h) % n2), ch = (int)(q / np);
        const int z0 = (int)((int64_t)n3 * ch / kHcsrChunks), z1 = (int)((int64_t)n3 * (ch + 1) / kHcsrChunks);
        uint32_t c = 0;
        for (int z = z0; z < z1; ++z) {
            const HalfVox hv = half_vox(kx, ky, z, n1, n2, n3);
            const bool A = hmeas(M, mask, hv.s);
            const bool Bm = hmeas(M, mask, hv.mate);   // on the kx = 0, n1/2 planes: the in-plane mate
            c += (A || Bm) ? 1u : 0u;
        }
        cnt[((int64_t)ky + (int64_t)n2 * kx) * kHcsrChunks + ch] = c;
    }
}

// csr_off[p] = the first chunk offset of pencil p (csr_off[np], the total, is
// written by the scan)
__global__ void k_hcsr_pencil_off(const int64_t* __restrict__ coff, int64_t np, int64_t* __restrict__ off)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x)
        off[p] = coff[p * kHcsrChunks];
}

__global__ void k_hcsr_fill(const float2* __restrict__ M, const uint8_t* __restrict__ mask, int n1, int n2,
                            int n3, const int64_t* __restrict__ coff, uint32_t* __restrict__ meta,
                            float2* __restrict__ val, float2* __restrict__ mag)
{
    const int h = n1 / 2 + 1;
    const int64_t np = (int64_t)h * n2;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < np * kHcsrChunks;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int kx = (int)(q % h), ky = (int)((q / h) % n2), ch = (int)(q / np);   // as in k_hcsr_count
        const int z0 = (int)((int64_t)n3 * ch / kHcsrChunks), z1 = (int)((int64_t)n3 * (ch + 1) / kHcsrChunks);
        int64_t o = coff[((int64_t)ky + (int64_t)n2 * kx) * kHcsrChunks + ch];
        for (int z = z0; z < z1; ++z) {
            const HalfVox hv = half_vox(kx, ky, z, n1, n2, n3);
            const bool A = hmeas(M, mask, hv.s);
            const bool Bm = hmeas(M, mask, hv.mate);
            const bool B = !hv.self_paired && Bm;
            if (!A && !Bm) continue;
            if (!A && hv.self_paired) {
                // listed without terms or replacement: it carries the mate's
                // share of the replacement delta on these planes
                meta[o] = (uint32_t)z | ((uint32_t)ky << 11) | kHalfPassive;
                val[o] = make_float2(0.0f, 0.0f);
                mag[o] = make_float2(0.0f, 0.0f);
                ++o;
                continue;
            }
            const float2 ms = M[hv.s];
            const float2 mm = hv.self_paired ? make_float2(0.0f, 0.0f) : conjf2(M[hv.mate]);
            uint32_t f = (uint32_t)z | ((uint32_t)ky << 11);
            float2 V;
            if (A && B) V = make_float2(0.5f * (ms.x + mm.x), 0.5f * (ms.y + mm.y));
            else V = A ? ms : mm;
            if (A) f |= kHalfTermA;
            if (B) f |= kHalfTermB;
            if (!hv.self_paired && A != B) f |= kHalfAverage;   // the mate is stored on the kx = 0, n1/2 planes
            meta[o] = f;
            val[o] = V;
            mag[o] = make_float2(A ? sqrtf(ms.x * ms.x + ms.y * ms.y) : 0.0f,
                                 B ? sqrtf(mm.x * mm.x + mm.y * mm.y) : 0.0f);
            ++o;
        }
    }
}

__device__ __forceinline__ double kcoord(int c, int n)
{
    // make_Kspace_indices.m:19-31
    if (n == 1) return 0.0;
    if (n % 2 == 0) return (double)(c - n / 2) / (double)(n / 2);
    return (double)(c - (n - 1) / 2) / (double)((n - 1) / 2);
}

// enforced = (1 - Q) > delay for the voxel (its mate has the same Q).
__global__ void k_hcsr_enforce(int n1, int n2, int n3, const int64_t* __restrict__ off,
                               uint32_t* __restrict__ meta, double delay)
{
    const int h = n1 / 2 + 1;
    const int64_t np = (int64_t)h * n2;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
        const int kx = (int)(q / n2), ky = (int)(q % n2);
        const double v1 = kcoord((kx + n1 / 2) % n1, n1);
        const double v2 = kcoord((ky + n2 / 2) % n2, n2);
        for (int64_t e = off[q]; e < off[q + 1]; ++e) {
            uint32_t m = meta[e] & ~kHalfEnforced;
            if (m & kHalfPassive) continue;
            const int z = (int)(m & 0x7FFu);
            const double v3 = kcoord((z + n3 / 2) % n3, n3);
            const double Q = sqrt((v2 * v2 + v1 * v1) + v3 * v3);
            if (1.0 - Q > delay) m |= kHalfEnforced;
            meta[e] = m;
        }
    }
}

__global__ __launch_bounds__(256) void k_hcsr_den(const float2* __restrict__ mag, const uint32_t* __restrict__ meta,
                                                  int64_t n, double* __restrict__ part)
{
    double t = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        const uint32_t m = meta[e];
        const float2 g = mag[e];
        if (m & kHalfTermA) t += (double)g.x;
        if (m & kHalfTermB) t += (double)g.y;
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

__global__ __launch_bounds__(256) void k_rfree_reduce(const float* __restrict__ terms,
                                                     const int32_t* __restrict__ order,
                                                     const int32_t* __restrict__ shell_off,
                                                     const double* __restrict__ den, int nbins,
                                                     float* __restrict__ out, const int* __restrict__ ctr)
{
    __shared__ double red[256];
    const int sh = blockIdx.x;
    double t = 0.0;
    for (int e = shell_off[sh] + (int)threadIdx.x; e < shell_off[sh + 1]; e += 256) t += (double)terms[order[e]];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[sh + (int64_t)nbins * (*ctr)] = den[sh] > 0.0 ? (float)(red[0] / den[sh]) : -1.0f;
}

// D -> D' on the kx = 0 and n1/2 planes: (D(s) + conj D(-s))/2 per mate pair,
// Re D(s) for a voxel that is its own mate.  Plane entries are the list's
// first run (pencils kx = 0) and its last (kx = n1/2); mate[i] is the entry of
// plane entry i's in-plane mate.
__global__ void k_hcsr_mates(const int64_t* __restrict__ off, const uint32_t* __restrict__ meta, int n1, int n2,
                             int n3, int64_t n0, int64_t mb, int64_t cnt, int64_t* __restrict__ mate)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i < n0 ? i : mb + (i - n0);
        const int kx = i < n0 ? 0 : n1 / 2;
        const uint32_t m = meta[e];
        const int z = (int)(m & 0x7FFu), ky = (int)((m >> 11) & 0x7FFu);
        const int mky = (n2 - ky) % n2, mz = (n3 - z) % n3;
        const int64_t q = (int64_t)mky + (int64_t)n2 * kx;
        int64_t lo = off[q], hi = off[q + 1], found = -1;
        while (lo < hi) {   // entries of a pencil are z-ascending
            const int64_t mid = (lo + hi) / 2;
            const int zm = (int)(meta[mid] & 0x7FFu);
            if (zm == mz) {
                found = mid;
                break;
            }
            if (zm < mz) lo = mid + 1;
            else hi = mid;
        }
        mate[i] = found;
    }
}

__global__ void k_hsum(const double* part, int n, double* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += part[i];
        *out = s;
    }
}

int pblocks(int64_t np) { return (int)std::min<int64_t>((np + 255) / 256, 4096); }

// Dynamic LDS above 64 KiB (the 1024-point passes) must be requested per kernel.
template <class K> void set_lds(K kern, size_t bytes)
{
    if (bytes > 65536)
        GF_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// Launch extents of the pruned passes: the largest tile counts of the batch.
struct XTiles {
    int cy, cz;
};
template <int M> XTiles x_tiles(const HalfBatch& b)
{
    XTiles r{0, 0};
    for (int i = 0; i < b.count; ++i) {
        const HalfIterArgs& a = b.job[i];
        r.cy = std::max(r.cy, box_tiles(a.n2, a.by, xrows<M>()).count);
        r.cz = std::max(r.cz, box_tiles(a.n3, a.bz, 1).count);
    }
    return r;
}
int y_tiles(const HalfBatch& b)
{
    int cz = 0;
    for (int i = 0; i < b.count; ++i) cz = std::max(cz, box_tiles(b.job[i].n3, b.job[i].bz, kHP).count);
    return cz;
}

// LDS of an x pass: the tile and the Nyquist bins, then the twiddle tables of
// the N/2-point transform and of the real split (N/2 + 1 entries).
template <int N> constexpr size_t x_pass_lds()
{
    return ((size_t)xrows<N / 2>() * (N / 2 + 1) + reg_tw_size<N / 2>() + N / 2 + 1) * sizeof(float2);
}

template <int N> void launch_first(const HalfBatch& b, hipStream_t st)
{
    const size_t lds = x_pass_lds<N>();
    set_lds(k_half_first<N>, lds);
    const XTiles xt = x_tiles<N / 2>(b);
    k_half_first<N><<<dim3(xt.cy * xt.cz, b.count), x_threads<N / 2>(), lds, st>>>(b, xt.cy, fft_twiddles(N / 2),
                                                                                   fft_twiddles(N));
    GF_LAUNCH_CHECK();
}
template <int N> void launch_last(const HalfBatch& b, hipStream_t st)
{
    const size_t lds = x_pass_lds<N>();
    set_lds(k_half_last<N>, lds);
    const XTiles xt = x_tiles<N / 2>(b);
    k_half_last<N><<<dim3(xt.cy * xt.cz, b.count), x_threads<N / 2>(), lds, st>>>(b, xt.cy, fft_twiddles(N / 2),
                                                                                  fft_twiddles(N));
    GF_LAUNCH_CHECK();
}
template <int N> void launch_last_first(const HalfBatch& b, hipStream_t st)
{
    const size_t lds = x_pass_lds<N>();
    set_lds(k_half_last_first<N>, lds);
    const float2* twM = fft_twiddles(N / 2);
    const float2* twN = fft_twiddles(N);
    const XTiles xt = x_tiles<N / 2>(b);
    k_half_last_first<N><<<dim3(xt.cy * xt.cz, b.count), x_threads<N / 2>(), lds, st>>>(b, xt.cy, twM, twN);
    GF_LAUNCH_CHECK();
}
template <int N> void launch_y(const HalfBatch& b, int inverse, hipStream_t st)
{
    const HalfIterArgs& a = b.job[0];
    const size_t lds = ((size_t)kHP * N + reg_tw_size<N>()) * sizeof(float2);   // tile + twiddle table
    const int cz = y_tiles(b);
    const dim3 blocks((a.n1 / 2 + 1) * cz, b.count);
    if (inverse) {
        set_lds(k_half_y_inv<N>, lds);
        k_half_y_inv<N><<<blocks, half_threads<N>(), lds, st>>>(b, cz, fft_twiddles(N));
    } else {
        set_lds(k_half_y_fwd<N>, lds);
        k_half_y_fwd<N><<<blocks, half_threads<N>(), lds, st>>>(b, cz, fft_twiddles(N));
    }
    GF_LAUNCH_CHECK();
}
template <int N> void launch_z(const HalfBatch& b, hipStream_t st)
{
    const HalfIterArgs& a = b.job[0];
    // tile + twiddle table
    const size_t lds = (size_t)kZP * N * sizeof(float2) + (size_t)reg_tw_size<N>() * sizeof(float2);
    const dim3 grid((a.n1 / 2 + 1) * (a.n2 / kZP), b.count);
    bool half = zhalf_ok<N>();
    for (int i = 0; i < b.count; ++i) half = half && b.job[i].bz == N / 2;
    if (half) {
        set_lds(k_half_z_fused<N, true>, lds);
        k_half_z_fused<N, true><<<grid, zfused_threads<N>(), lds, st>>>(b, fft_twiddles(N));
    } else {
        set_lds(k_half_z_fused<N, false>, lds);
        k_half_z_fused<N, false><<<grid, zfused_threads<N>(), lds, st>>>(b, fft_twiddles(N));
    }
    GF_LAUNCH_CHECK();
}

}  // namespace

bool half_supported(int n1, int n2, int n3)
{
    // lengths with a register plan: 64, 96, 128, 160, 192, 256, 320, 384, 512, 640, 768, 1024
    auto reg = [](int n) {
        switch (n) {
            case 64: case 96: case 128: case 160: case 192: case 256: case 320: case 384: case 512: case 640:
            case 768: case 1024: return true;
            default: return false;
        }
    };
    return n1 % 2 == 0 && n1 >= 128 && reg(n1 / 2) && reg(n2) && reg(n3);
}



#define GF_HALF_SWITCH(n, CALL)                                               \
    switch (n) {                                                              \
        case 64: CALL(64); break;                                             \
        case 96: CALL(96); break;                                             \
        case 128: CALL(128); break;                                           \
        case 160: CALL(160); break;                                           \
        case 192: CALL(192); break;                                           \
        case 256: CALL(256); break;                                           \
        case 320: CALL(320); break;                                           \
        case 384: CALL(384); break;                                           \
        case 512: CALL(512); break;                                           \
        case 640: CALL(640); break;                                           \
        case 768: CALL(768); break;                                           \
        case 1024: CALL(1024); break;                                         \
        default: raise(GENFIRE_ERR_UNSUPPORTED, "half-spectrum length not supported"); \
    }



namespace {
template <int N> int z_waves() { return zfused_threads<N>() / 64; }
}  // namespace

// errK partials of the z pass: one per wave of each workgroup
int64_t half_partials(int n1, int n2, int n3)
{
    int w = 0;
#define C(NN) w = z_waves<NN>()
    GF_HALF_SWITCH(n3, C)
#undef C
    return (int64_t)(n1 / 2 + 1) * (n2 / kZP) * w;
}

// x lengths n1: the complex transform runs at n1/2
#define GF_HALF_SWITCH_X(n, CALL)                                             \
    switch (n) {                                                              \
        case 128: CALL(128); break;                                           \
        case 192: CALL(192); break;                                           \
        case 256: CALL(256); break;                                           \
        case 320: CALL(320); break;                                           \
        case 384: CALL(384); break;                                           \
        case 512: CALL(512); break;                                           \
        case 640: CALL(640); break;                                           \
        case 768: CALL(768); break;                                           \
        case 1024: CALL(1024); break;                                         \
        default: raise(GENFIRE_ERR_UNSUPPORTED, "half-spectrum x length not supported"); \
    }

namespace {

HalfBatch one(const HalfIterArgs& a)
{
    HalfBatch b{};
    b.job[0] = a;
    b.count = 1;
    return b;
}

void check_batch(const HalfBatch& b)
{
    GF_CHECK(b.count >= 1 && b.count <= kHalfBatchMax, GENFIRE_ERR_ARG, "half-spectrum batch of 1..8 jobs");
    for (int i = 1; i < b.count; ++i)
        GF_CHECK(b.job[i].n1 == b.job[0].n1 && b.job[i].n2 == b.job[0].n2 && b.job[i].n3 == b.job[0].n3,
                 GENFIRE_ERR_ARG, "a half-spectrum batch needs jobs of one grid shape");
}

}  // namespace

void half_first(const HalfBatch& b, hipStream_t st)
{
    check_batch(b);
#define C(NN) launch_first<NN>(b, st)
    GF_HALF_SWITCH_X(b.job[0].n1, C)
#undef C
}
void half_y(const HalfBatch& b, int inverse, hipStream_t st)
{
    check_batch(b);
#define C(NN) launch_y<NN>(b, inverse, st)
    GF_HALF_SWITCH(b.job[0].n2, C)
#undef C
}
void half_z(const HalfBatch& b, hipStream_t st)
{
    check_batch(b);
#define C(NN) launch_z<NN>(b, st)
    GF_HALF_SWITCH(b.job[0].n3, C)
#undef C
}
void half_last(const HalfBatch& b, hipStream_t st)
{
    check_batch(b);
#define C(NN) launch_last<NN>(b, st)
    GF_HALF_SWITCH_X(b.job[0].n1, C)
#undef C
}
void half_last_first(const HalfBatch& b, hipStream_t st)
{
    check_batch(b);
#define C(NN) launch_last_first<NN>(b, st)
    GF_HALF_SWITCH_X(b.job[0].n1, C)
#undef C
}

void hcsr_mates(const int64_t* d_off, const uint32_t* d_meta, int n1, int n2, int n3, int64_t n_plane0,
                int64_t plane_m_begin, int64_t n_plane, int64_t* d_mate, hipStream_t st)
{
    if (n_plane <= 0) return;
    k_hcsr_mates<<<pblocks(n_plane), 256, 0, st>>>(d_off, d_meta, n1, n2, n3, n_plane0, plane_m_begin, n_plane,
                                                   d_mate);
    GF_LAUNCH_CHECK();
}

void half_first(const HalfIterArgs& a, hipStream_t st) { half_first(one(a), st); }
void half_y(const HalfIterArgs& a, int inverse, hipStream_t st) { half_y(one(a), inverse, st); }
void half_z(const HalfIterArgs& a, hipStream_t st) { half_z(one(a), st); }
void half_last(const HalfIterArgs& a, hipStream_t st) { half_last(one(a), st); }
void half_last_first(const HalfIterArgs& a, hipStream_t st) { half_last_first(one(a), st); }

void rfree_reduce(const float* d_terms, const int32_t* d_order, const int32_t* d_shell_off, const double* d_den,
                  int nbins, float* d_Rfree, const int* d_ctr, hipStream_t st)
{
    k_rfree_reduce<<<nbins, 256, 0, st>>>(d_terms, d_order, d_shell_off, d_den, nbins, d_Rfree, d_ctr);
    GF_LAUNCH_CHECK();
}

int hcsr_chunks() { return kHcsrChunks; }
void hcsr_count(const float2* d_M, const uint8_t* d_mask, int n1, int n2, int n3, uint32_t* d_cnt, hipStream_t st)
{
    k_hcsr_count<<<pblocks((int64_t)(n1 / 2 + 1) * n2 * kHcsrChunks), 256, 0, st>>>(d_M, d_mask, n1, n2, n3, d_cnt);
    GF_LAUNCH_CHECK();
}
void hcsr_pencil_off(const int64_t* d_coff, int64_t np, int64_t* d_off, hipStream_t st)
{
    k_hcsr_pencil_off<<<pblocks(np), 256, 0, st>>>(d_coff, np, d_off);
    GF_LAUNCH_CHECK();
}
void hcsr_fill(const float2* d_M, const uint8_t* d_mask, int n1, int n2, int n3, const int64_t* d_coff,
               uint32_t* d_meta, float2* d_val, float2* d_mag, hipStream_t st)
{
    k_hcsr_fill<<<pblocks((int64_t)(n1 / 2 + 1) * n2 * kHcsrChunks), 256, 0, st>>>(d_M, d_mask, n1, n2, n3, d_coff,
                                                                                     d_meta, d_val, d_mag);
    GF_LAUNCH_CHECK();
}
void hcsr_enforce(int n1, int n2, int n3, const int64_t* d_off, uint32_t* d_meta, double delay, hipStream_t st)
{
    k_hcsr_enforce<<<pblocks((int64_t)(n1 / 2 + 1) * n2), 256, 0, st>>>(n1, n2, n3, d_off, d_meta, delay);
    GF_LAUNCH_CHECK();
}
void hcsr_den(const float2* d_mag, const uint32_t* d_meta, int64_t n, double* d_tmp, int ntmp, double* d_out,
              hipStream_t st)
{
    k_hcsr_den<<<ntmp, 256, 0, st>>>(d_mag, d_meta, n, d_tmp);
    GF_LAUNCH_CHECK();
    k_hsum<<<1, 64, 0, st>>>(d_tmp, ntmp, d_out);
    GF_LAUNCH_CHECK();
}

}  // namespace gf
