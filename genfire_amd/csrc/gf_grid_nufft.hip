// gf_grid_nufft.hip -- the DFT values of fillInFourierGrid_DFT.m:236-248,
//   F_j = sum_{a,b} P[a,b] exp(-2 pi i ((a - c1) u_j + (b - c2) v_j)),
// as a type-2 non-uniform FFT instead of the direct sum.  One projection:
//   1. b_k = P_k / (phi^1(k1) phi^2(k2)) for the modes k = (a - c1, b - c2),
//      placed at (k1 mod M1, k2 mod M2) of an M1 x M2 grid (M >= 2 Dim);
//   2. G = forward 2D FFT of that grid (the register FFT passes);
//   3. F_j = sum over the w x w fine-grid nodes l nearest (M1 u_j, M2 v_j) of
//      G_l phi1(M1 u_j - l1) phi2(M2 v_j - l2).
// phi is the exponential-of-semicircle kernel phi(z) = exp(beta (sqrt(1 -
// (2z/w)^2) - 1)) on |z| < w/2 (Barnett, Magland and af Klinteberg, "A parallel
// non-uniform fast Fourier transform library based on an exponential of
// semicircle kernel", SIAM J. Sci. Comput. 41 (2019)), w = 8 and beta = 0.97 pi
// w (1 - 1/(2 sigma)) with sigma = M / Dim; phi^(k) = integral of phi(z)
// cos(2 pi k z / M), by Gauss-Legendre quadrature in FP64 on the host.  With
// these parameters the values agree with the direct sum to about 1e-7 relative:
// the same scale as FP32 rounding of the direct sum, and 1e-2 of the 1e-5 bar
// on gridded values.  The grid is at most a few MB per projection and stays in
// L2 while its points are interpolated (points are grouped by projection).
#include <cmath>
#include <vector>

#include "gf_grid.h"

namespace gf {

namespace {

constexpr int kNW = 8;           // kernel width in fine-grid nodes
constexpr int kNufftTile = 4;    // point tiles (64 points each) per workgroup

__global__ __launch_bounds__(256) void k_nufft_load(const float* __restrict__ proj, int d1, int d2, int P, int M1,
                                                    int M2, const float* __restrict__ inv1,
                                                    const float* __restrict__ inv2, float2* __restrict__ fine)
{
    const int c1 = d1 / 2, c2 = d2 / 2;
    const int64_t total = (int64_t)M1 * M2 * P;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int l1 = (int)(t % M1), l2 = (int)((t / M1) % M2), p = (int)(t / ((int64_t)M1 * M2));
        // mode k at grid index k mod M: k >= 0 at l < Dim - c, k < 0 at l >= M - c
        const int a = l1 < d1 - c1 ? l1 + c1 : (l1 >= M1 - c1 ? l1 - M1 + c1 : -1);
        const int b = l2 < d2 - c2 ? l2 + c2 : (l2 >= M2 - c2 ? l2 - M2 + c2 : -1);
        float v = 0.0f;
        if (a >= 0 && b >= 0) v = proj[a + (int64_t)d1 * (b + (int64_t)d2 * p)] * inv1[a] * inv2[b];
        fine[t] = make_float2(v, 0.0f);
    }
}

__device__ __forceinline__ float es_kernel(float z, float beta)
{
    // z in units of w/2; zero outside (-1, 1]
    const float s = 1.0f - z * z;
    return s >= 0.0f ? expf(beta * (sqrtf(s) - 1.0f)) : 0.0f;
}

__global__ __launch_bounds__(64 * kNufftTile) void k_nufft_interp(const float2* __restrict__ G, int M1, int M2,
                                                                  const float* __restrict__ un,
                                                                  const float* __restrict__ vn,
                                                                  const PointTile* __restrict__ tiles, int ntiles,
                                                                  float beta1, float beta2, float2* __restrict__ F)
{
    const int ti = (int)blockIdx.x * kNufftTile + (int)threadIdx.x / 64;
    if (ti >= ntiles) return;
    const PointTile tl = tiles[ti];
    const int i = (int)threadIdx.x % 64;
    if (i >= tl.count) return;
    const int64_t j = tl.begin + i;
    const float s1 = (float)M1 * un[j], s2 = (float)M2 * vn[j];
    const int l1 = (int)ceilf(s1 - 0.5f * kNW), l2 = (int)ceilf(s2 - 0.5f * kNW);
    float w1[kNW], w2[kNW];
    int r1[kNW];
#pragma unroll
    for (int q = 0; q < kNW; ++q) {
        w1[q] = es_kernel((s1 - (float)(l1 + q)) * (2.0f / kNW), beta1);
        w2[q] = es_kernel((s2 - (float)(l2 + q)) * (2.0f / kNW), beta2);
        int m = (l1 + q) % M1;
        r1[q] = m < 0 ? m + M1 : m;
    }
    const float2* Gp = G + (int64_t)M1 * M2 * tl.proj;
    float fr = 0.0f, fi = 0.0f;
#pragma unroll
    for (int u = 0; u < kNW; ++u) {
        int m2 = (l2 + u) % M2;
        m2 = m2 < 0 ? m2 + M2 : m2;
        const float2* row = Gp + (int64_t)M1 * m2;
        float rr = 0.0f, ri = 0.0f;
#pragma unroll
        for (int q = 0; q < kNW; ++q) {
            const float2 g = row[r1[q]];
            rr += w1[q] * g.x;
            ri += w1[q] * g.y;
        }
        fr += w2[u] * rr;
        fi += w2[u] * ri;
    }
    F[j] = make_float2(fr, fi);
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n).
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

}  // namespace

int nufft_kernel_width() { return kNW; }

int nufft_fine_len(int dim)
{
    for (int m = std::max(2 * dim, 2 * kNW); m <= 4 * std::max(dim, kNW); ++m)
        if (fft_supported(m)) return m;
    return 0;
}

double nufft_beta(int dim, int M)
{
    const double sigma = (double)M / (double)dim, pi = 3.14159265358979323846;
    return 0.97 * pi * kNW * (1.0 - 1.0 / (2.0 * sigma));
}

// 1 / phi^(a - dim/2) for the mode index a = 0..dim-1:
// phi^(k) = w * integral_0^{pi/2} exp(beta (cos t - 1)) cos(pi k w sin t / M) cos t dt
// (z = (w/2) sin t removes the kernel's square-root edge), 64-node Gauss-Legendre.
std::vector<float> nufft_deconv(int dim, int M)
{
    // the 64-node rule, computed once per process (its Newton iterations cost
    // about 0.4 ms of host time)
    struct Rule {
        std::vector<double> x, w;
        Rule() { gauss_legendre(64, x, w); }
    };
    static const Rule rule;
    const std::vector<double>& x = rule.x;
    const std::vector<double>& wq = rule.w;
    const double beta = nufft_beta(dim, M), pi = 3.14159265358979323846, h = 0.25 * pi;
    std::vector<float> inv(dim);
    for (int a = 0; a < dim; ++a) {
        const int k = a - dim / 2;
        double s = 0.0;
        for (int q = 0; q < 64; ++q) {
            const double t = h * (x[q] + 1.0);   // [0, pi/2]
            s += wq[q] * std::exp(beta * (std::cos(t) - 1.0)) * std::cos(pi * k * kNW * std::sin(t) / M) * std::cos(t);
        }
        inv[a] = (float)(1.0 / (kNW * h * s));
    }
    return inv;
}

void nufft_load(const float* d_proj, int d1, int d2, int P, int M1, int M2, const float* d_inv1,
                const float* d_inv2, float2* d_fine, hipStream_t st)
{
    const int64_t total = (int64_t)M1 * M2 * P;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 1 << 20);
    k_nufft_load<<<(unsigned)blocks, 256, 0, st>>>(d_proj, d1, d2, P, M1, M2, d_inv1, d_inv2, d_fine);
    GF_LAUNCH_CHECK();
}

void nufft_interp(const float2* d_G, int M1, int M2, double beta1, double beta2, const float* d_un,
                  const float* d_vn, const PointTile* d_tiles, int ntiles, float2* d_F, hipStream_t st)
{
    if (ntiles <= 0) return;
    const int blocks = (ntiles + kNufftTile - 1) / kNufftTile;
    k_nufft_interp<<<blocks, 64 * kNufftTile, 0, st>>>(d_G, M1, M2, d_un, d_vn, d_tiles, ntiles, (float)beta1,
                                                      (float)beta2, d_F);
    GF_LAUNCH_CHECK();
}

}  // namespace gf
