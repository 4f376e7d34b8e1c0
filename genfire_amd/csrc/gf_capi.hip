// gf_capi.hip -- the extern "C" boundary of libgenfire (include/genfire.h) and the
// host orchestration of one reconstruction job on one HIP stream.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>

#include "gf_grid.h"

struct genfire_ctx {
    int device = 0;
    std::mutex mu;
    std::string err;
};

namespace gf {

[[noreturn]] void raise(int code, const std::string& msg) { throw Error{code, msg}; }

namespace {

// MATLAB round: half away from zero.
double mround(double x) { return x < 0 ? -std::floor(-x + 0.5) : std::floor(x + 0.5); }

// round(linspace(1, N, m)) (reconstruct.m:103) with MATLAB's linspace
// d1 + (0:n1).*(d2-d1)./n1 and the exact endpoint.
std::vector<int> cutoff_schedule(int N, int m)
{
    std::vector<int> out(m);
    if (m == 1) { out[0] = N; return out; }
    const double d1 = 1.0, d2 = (double)N;
    const int n1 = m - 1;
    for (int k = 0; k <= n1; ++k) {
        double y = d1 + ((double)k * (d2 - d1)) / (double)n1;
        if (k == n1) y = d2;
        if (d1 == d2) y = d1;
        out[k] = (int)mround(y);
    }
    return out;
}

// MATLAB colon a:d:b with MathWorks' symmetric evaluation (colonop).
std::vector<double> colon(double a, double d, double b)
{
    const double tol = 2.0 * 2.220446049250313e-16 * std::max(std::fabs(a), std::fabs(b));
    const double sig = d < 0 ? -1.0 : 1.0;
    int n = (int)mround((b - a) / d);
    if (sig * (a + n * d - b) > tol) --n;
    std::vector<double> v;
    if (n < 0) return v;
    double c = a + n * d;
    if (sig * (c - b) > -tol) c = b;
    v.assign(n + 1, 0.0);
    for (int k = 0; k <= n / 2; ++k) {
        v[k] = a + k * d;
        v[n - k] = c - k * d;
    }
    if (n % 2 == 0) v[n / 2] = (a + c) / 2;
    return v;
}

// constraintEnforcementDelayIndicators, GENFIRE_Reconstructor.m:156-165.
std::vector<double> delay_indicators(int mode)
{
    std::vector<double> v;
    if (mode == 1) {
        v = colon(0.95, -0.1, -0.15);
        v.insert(v.end(), {-10.0, -10.0, -10.0});
        auto up = colon(-0.15, 0.1, 0.95);
        v.insert(v.end(), up.begin(), up.end());
    } else if (mode == 2) {
        v = colon(0.95, -0.1, -0.15);
        v.insert(v.end(), {-10.0, -10.0});
    } else if (mode == 3) {
        v.assign(5, -999.0);
    } else {
        raise(GENFIRE_ERR_ARG, "GENFIRE: ERROR! constraintEnforcementMode value " + std::to_string(mode) +
                                   " not understood");
    }
    return v;
}

struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    void init()
    {
        if (!a) {
            GF_HIP(hipEventCreate(&a));
            GF_HIP(hipEventCreate(&b));
        }
    }
    ~EventTimer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Per-kernel event pairs for the live roofline measurement.
struct KStats {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[GENFIRE_KSTAT_COUNT];
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    int64_t launches[GENFIRE_KSTAT_COUNT] = {};
    double ms[GENFIRE_KSTAT_COUNT] = {};
    std::pair<hipEvent_t, hipEvent_t> take()
    {
        if (!pool.empty()) {
            auto e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t a, b;
        GF_HIP(hipEventCreate(&a));
        GF_HIP(hipEventCreate(&b));
        return {a, b};
    }
    void collect()
    {
        for (int k = 0; k < GENFIRE_KSTAT_COUNT; ++k) {
            for (auto& e : pending[k]) {
                float t = 0.0f;
                GF_HIP(hipEventSynchronize(e.second));
                GF_HIP(hipEventElapsedTime(&t, e.first, e.second));
                ms[k] += t;
                launches[k] += 1;
                pool.push_back(e);
            }
            pending[k].clear();
        }
    }
    void reset()
    {
        for (int k = 0; k < GENFIRE_KSTAT_COUNT; ++k) { launches[k] = 0; ms[k] = 0.0; }
    }
    ~KStats()
    {
        for (auto& e : pool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        for (auto& v : pending)
            for (auto& e : v) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    }
};

}  // namespace
}  // namespace gf

using namespace gf;

struct genfire_job {
    genfire_ctx* ctx = nullptr;
    genfire_params prm{};
    int dim1 = 0, dim2 = 0, P = 0, n1 = 0, n2 = 0;
    HalfGrid hg{};
    std::vector<double> delays;
    hipStream_t st = nullptr;

    // inputs
    DevBuf proj, support_box, init_box;
    std::vector<double> angles;  // P x 3 column-major
    bool have_inputs = false, have_support = false, have_init = false;

    // gridding
    DevBuf planes, blk_counts, blk_offsets, total, ind, dist, un, vn, F, tiles, Pt, Phi, Plo, Pscale;
    DevBuf acc, sw, hits, exact;
    DevBuf M, mask, recifft;
    std::vector<int64_t> proj_off;
    int64_t npts = 0;
    bool gridded = false, have_mask = false;

    // iteration
    DevBuf x, K, S, best, pfbuf, err_part, errK, den_tmp, den, recon;
    DevBuf csr_cnt, csr_off, csr_meta, csr_val, csr_mag;
    int64_t nmeas = 0;
    DevBuf vol, vol_y, Rmat, rsums, rout;
    bool rmat_any_y = false;
    std::vector<float> errK_h, errR_h;
    std::vector<double> Rarr_h, Rarr2_h;
    bool reconstructed = false;

    // timing
    double t_grid = 0, t_iter = 0, t_rf = 0, t_total = 0;
    bool prof = false;
    KStats ks;

    // async
    std::thread worker;
    std::atomic<int> async_state{0};  // 0 idle, 1 running, 2 done
    int async_status = 0;
    std::string async_err;

    ~genfire_job()
    {
        if (worker.joinable()) worker.join();
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

void set_err(genfire_ctx* ctx, const std::string& m)
{
    if (!ctx) return;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->err = m;
}

template <class F> int guarded(genfire_ctx* ctx, F&& f)
{
    try {
        f();
        return GENFIRE_OK;
    } catch (const Error& e) {
        set_err(ctx, e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_err(ctx, "host allocation failed");
        return GENFIRE_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_err(ctx, e.what());
        return GENFIRE_ERR_ARG;
    }
}

void use_device(genfire_ctx* ctx) { GF_HIP(hipSetDevice(ctx->device)); }

int64_t full_size(const genfire_job* j) { return (int64_t)j->n1 * j->n2 * j->n1; }
int64_t box_size(const genfire_job* j) { return (int64_t)j->dim1 * j->dim2 * j->dim1; }

void kstat_begin(genfire_job* j, int k, std::pair<hipEvent_t, hipEvent_t>* ev)
{
    if (!j->prof) return;
    *ev = j->ks.take();
    GF_HIP(hipEventRecord(ev->first, j->st));
    (void)k;
}

void kstat_end(genfire_job* j, int k, std::pair<hipEvent_t, hipEvent_t>& ev)
{
    if (!j->prof) return;
    GF_HIP(hipEventRecord(ev.second, j->st));
    j->ks.pending[k].push_back(ev);
}

// ZYZ matrices of the projector (calculate3Dprojection_RealSpaceinterp.m:13-15),
// kRStride floats per projection with the y-fastest flag at [9].
std::vector<float> proj_matrices(const double* angles, int P, bool* any_y)
{
    std::vector<float> Rm((size_t)kRStride * P, 0.0f);
    *any_y = false;
    for (int p = 0; p < P; ++p) {
        double R[9];
        zyz_matrix(angles[p], angles[P + p], angles[2 * P + p], R);
        float* r = Rm.data() + (size_t)kRStride * p;
        for (int q = 0; q < 9; ++q) r[q] = (float)R[q];
        if (project_wants_yfast(r)) { r[9] = 1.0f; *any_y = true; }
    }
    return Rm;
}

// ---------------------------------------------------------------------------
// fillInFourierGrid_DFT + CentroSymmetricity + runGridding post-processing
// ---------------------------------------------------------------------------
void do_grid(genfire_job* j)
{
    GF_CHECK(j->have_inputs, GENFIRE_ERR_STATE, "GENFIRE: inputs not set (call genfire_job_set_inputs)");
    const int P = j->P, n1 = j->n1, n2 = j->n2;
    const int centro = j->prm.DFT_CentroSymmetricity ? 1 : 0;
    hipStream_t st = j->st;
    EventTimer tm;
    tm.init();
    GF_HIP(hipEventRecord(tm.a, st));

    // rotation matrices and plane constants (fillInFourierGrid_DFT.m:100-118)
    std::vector<PlaneConst> pcs(P);
    int64_t max_cols = 1;
    for (int p = 0; p < P; ++p) {
        float Rf[9];
        rotation_f32(j->prm.vector1, j->prm.vector2, j->prm.vector3, j->angles[p], j->angles[P + p],
                     j->angles[2 * P + p], Rf);
        plane_setup(Rf, &pcs[p]);
        const int a = pcs[p].axis, b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
        max_cols = std::max<int64_t>(max_cols, (int64_t)j->hg.len[b] * j->hg.len[c]);
    }
    j->planes.alloc(sizeof(PlaneConst) * P);
    GF_HIP(hipMemcpyAsync(j->planes.p, pcs.data(), sizeof(PlaneConst) * P, hipMemcpyHostToDevice, st));

    const int nthr = index_threads();
    const int nblk = (int)((max_cols + nthr - 1) / nthr);
    const int64_t nb_total = (int64_t)nblk * P;
    j->blk_counts.alloc(sizeof(uint32_t) * nb_total);
    j->blk_offsets.alloc(sizeof(int64_t) * nb_total);
    j->total.alloc(sizeof(int64_t));
    const double cutoff = j->prm.interpolationCutoffDistance;
    grid_index_count(j->planes.as<PlaneConst>(), P, j->hg, cutoff, n1, n2, nblk,
                     j->blk_counts.as<uint32_t>(), st);
    grid_index_scan(j->blk_counts.as<uint32_t>(), nb_total, j->blk_offsets.as<int64_t>(),
                    j->total.as<int64_t>(), st);
    std::vector<int64_t> offs(nb_total);
    int64_t npts = 0;
    GF_HIP(hipMemcpyAsync(offs.data(), j->blk_offsets.p, sizeof(int64_t) * nb_total, hipMemcpyDeviceToHost, st));
    GF_HIP(hipMemcpyAsync(&npts, j->total.p, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    GF_HIP(hipStreamSynchronize(st));
    j->npts = npts;
    j->proj_off.assign(P + 1, npts);
    for (int p = 0; p < P; ++p) j->proj_off[p] = offs[(int64_t)p * nblk];

    const int64_t cap = std::max<int64_t>(npts, 1);
    j->ind.alloc(sizeof(uint32_t) * cap);
    j->dist.alloc(sizeof(float) * cap);
    j->un.alloc(sizeof(float) * cap);
    j->vn.alloc(sizeof(float) * cap);
    j->F.alloc(sizeof(float2) * cap);
    grid_index_emit(j->planes.as<PlaneConst>(), P, j->hg, cutoff, n1, n2, nblk, j->blk_offsets.as<int64_t>(),
                    j->ind.as<uint32_t>(), j->dist.as<float>(), j->un.as<float>(), j->vn.as<float>(), st);

    // DFT values (:236-248).  Default: the MFMA GEMM form on 64-point tiles in
    // the job's precision; with FP32, GENFIRE_GRID_KERNEL=valu selects the VALU
    // kernel (256-point tiles).
    const int prec = j->prm.precision;
    const char* gk = std::getenv("GENFIRE_GRID_KERNEL");
    const bool use_valu = prec == GENFIRE_PREC_FP32 && gk && std::string(gk) == "valu";
    const int tile_pts = use_valu ? 256 : mfma_tile_points();
    std::vector<PointTile> tl;
    for (int p = 0; p < P; ++p)
        for (int64_t b = j->proj_off[p]; b < j->proj_off[p + 1]; b += tile_pts)
            tl.push_back({b, (int32_t)std::min<int64_t>(tile_pts, j->proj_off[p + 1] - b), p});
    if (!tl.empty()) {
        j->tiles.alloc(sizeof(PointTile) * tl.size());
        GF_HIP(hipMemcpyAsync(j->tiles.p, tl.data(), sizeof(PointTile) * tl.size(), hipMemcpyHostToDevice, st));
        if (prec != GENFIRE_PREC_FP32) {
            const size_t ne = grid_prep_lp_elems(j->dim1, j->dim2, P);
            j->Phi.alloc(sizeof(uint16_t) * ne);
            if (prec == GENFIRE_PREC_SPLIT_BF16) j->Plo.alloc(sizeof(uint16_t) * ne);
            j->Pscale.alloc(sizeof(float) * P);
            grid_prep_lp(prec, j->proj.as<float>(), j->dim1, j->dim2, P, j->Phi.as<uint16_t>(),
                         j->Plo.as<uint16_t>(), j->Pscale.as<float>(), st);
        } else if (!use_valu) {
            j->Pt.alloc(sizeof(float) * grid_prep_floats(j->dim1, j->dim2, P));
            grid_prep_proj(j->proj.as<float>(), j->dim1, j->dim2, P, j->Pt.as<float>(), st);
        }
        std::pair<hipEvent_t, hipEvent_t> ev;
        kstat_begin(j, GENFIRE_KSTAT_GRID_VALUES, &ev);
        if (prec != GENFIRE_PREC_FP32)
            grid_values_mfma_lp(prec, j->Phi.as<uint16_t>(), j->Plo.as<uint16_t>(), j->Pscale.as<float>(), j->dim1,
                                j->dim2, j->un.as<float>(), j->vn.as<float>(), j->tiles.as<PointTile>(),
                                (int)tl.size(), j->F.as<float2>(), st);
        else if (use_valu)
            grid_values_fp32(j->proj.as<float>(), j->dim1, j->dim2, j->un.as<float>(), j->vn.as<float>(),
                             j->tiles.as<PointTile>(), (int)tl.size(), j->F.as<float2>(), st);
        else
            grid_values_mfma_fp32(j->Pt.as<float>(), j->dim1, j->dim2, j->un.as<float>(), j->vn.as<float>(),
                                  j->tiles.as<PointTile>(), (int)tl.size(), j->F.as<float2>(), st);
        kstat_end(j, GENFIRE_KSTAT_GRID_VALUES, ev);
    }

    // weighted average (:267-291), projection by projection in list order
    const int64_t ht = j->hg.total();
    j->acc.alloc(sizeof(double2) * ht);
    j->sw.alloc(sizeof(double) * ht);
    j->hits.alloc(sizeof(uint32_t) * ht);
    j->exact.alloc(sizeof(float2) * ht);
    GF_HIP(hipMemsetAsync(j->acc.p, 0, j->acc.bytes, st));
    GF_HIP(hipMemsetAsync(j->sw.p, 0, j->sw.bytes, st));
    GF_HIP(hipMemsetAsync(j->hits.p, 0, j->hits.bytes, st));
    for (int p = 0; p < P; ++p)
        grid_scatter(j->ind.as<uint32_t>(), j->dist.as<float>(), j->F.as<float2>(), j->proj_off[p],
                     j->proj_off[p + 1] - j->proj_off[p], j->acc.as<double2>(), j->sw.as<double>(),
                     j->hits.as<uint32_t>(), j->exact.as<float2>(), st);

    // CentroSymmetricity + single cast, straight into FFT order
    j->M.alloc(sizeof(float2) * full_size(j));
    grid_finalize(j->hg, n1, n2, centro, j->acc.as<double2>(), j->sw.as<double>(), j->hits.as<uint32_t>(),
                  j->exact.as<float2>(), j->M.as<float2>(), st);
    j->have_mask = false;
    j->mask.release();

    // recIFFT = crop(real(fftshift(ifftn(ifftshift(measuredK))))) (:304)
    if (j->prm.compute_recIFFT) {
        j->K.alloc(sizeof(float2) * full_size(j));
        GF_HIP(hipMemcpyAsync(j->K.p, j->M.p, j->M.bytes, hipMemcpyDeviceToDevice, st));
        fft3(j->K.as<float2>(), n1, n2, n1, 1, st);
        j->recifft.alloc(sizeof(float) * box_size(j));
        crop_centre(j->K.p, 1, n1, n2, n1, j->recifft.as<float>(), j->dim1, j->dim2, j->dim1, st);
    }
    GF_HIP(hipEventRecord(tm.b, st));
    GF_HIP(hipEventSynchronize(tm.b));
    float ms = 0.0f;
    GF_HIP(hipEventElapsedTime(&ms, tm.a, tm.b));
    j->t_grid = ms;
    // the FP64 accumulators and the GEMM operands are only needed for the values above
    j->acc.release();
    j->sw.release();
    j->exact.release();
    j->Pt.release();
    j->Phi.release();
    j->Plo.release();
    j->gridded = true;
    j->reconstructed = false;
}

// ---------------------------------------------------------------------------
// reconstruct.m:28-194 (GENFIRE) and the Douglas-Rachford variant
// ---------------------------------------------------------------------------
void run_rfactor(genfire_job* j, const float* d_model_fft, int slot);

void do_reconstruct(genfire_job* j)
{
    GF_CHECK(j->gridded, GENFIRE_ERR_STATE, "GENFIRE: measuredK not available (call genfire_job_grid first)");
    const int N = j->prm.numIterations;
    const int n1 = j->n1, n2 = j->n2, n3 = j->n1;
    const int64_t nf = full_size(j);
    hipStream_t st = j->st;
    const bool dr = j->prm.algorithm == GENFIRE_ALG_DR;
    const bool monitor = j->prm.monitor_R != 0;
    const int L = std::max(1, j->prm.monitorR_loopLength);

    const bool half = half_supported(n1, n2, n3);
    const int64_t hn = (int64_t)(n1 / 2 + 1) * n2 * n3;   // half-spectrum size
    j->x.alloc(sizeof(float) * nf);
    j->K.alloc(sizeof(float2) * (half ? std::max<int64_t>(2 * hn, nf) : nf));
    j->S.alloc(nf);
    j->best.alloc(sizeof(float) * nf);
    if (monitor && (dr || half)) j->pfbuf.alloc(sizeof(float) * nf);
    IterArgs ia{};
    ia.n1 = n1; ia.n2 = n2; ia.n3 = n3;
    GF_CHECK(n1 <= 2048 && n3 <= 2048, GENFIRE_ERR_UNSUPPORTED, "GENFIRE: grid edge above 2048");
    const int64_t nparts = half ? half_partials(n1, n2, n3) : iter_fused_partials(ia);
    j->err_part.alloc(sizeof(double) * nparts);
    j->errK.alloc(sizeof(float) * std::max(N, 1));
    const int ntmp = 1024;
    j->den_tmp.alloc(sizeof(double) * ntmp);
    j->den.alloc(sizeof(double));

    EventTimer all, loop;
    all.init();
    loop.init();
    GF_HIP(hipEventRecord(all.a, st));

    // padded support (reconstruct.m:34) and initial object (:76-80), ifftshifted (:109-110)
    if (j->have_support) {
        pad_centre_u8(j->support_box.as<uint8_t>(), j->dim1, j->dim2, j->dim1, j->S.as<uint8_t>(), n1, n2, n3, st);
    } else {
        // Support = ones(Dim1,Dim2,Dim1) (GENFIRE_Reconstructor.m:146-148), padded
        DevBuf ones;
        ones.alloc(box_size(j));
        GF_HIP(hipMemsetAsync(ones.p, 1, ones.bytes, st));
        pad_centre_u8(ones.as<uint8_t>(), j->dim1, j->dim2, j->dim1, j->S.as<uint8_t>(), n1, n2, n3, st);
        GF_HIP(hipStreamSynchronize(st));
    }
    if (j->have_init)
        pad_centre(j->init_box.as<float>(), j->dim1, j->dim2, j->dim1, j->x.as<float>(), n1, n2, n3, st);
    else
        GF_HIP(hipMemsetAsync(j->x.p, 0, j->x.bytes, st));

    const std::vector<int> sched = cutoff_schedule(N, (int)j->delays.size());
    GF_CHECK(N >= 1 && sched.back() == N, GENFIRE_ERR_ARG, "GENFIRE: numIterations must be >= 1");

    // measured voxels (errInd, reconstruct.m:99) as per-pencil lists
    const uint8_t* mask = j->have_mask ? j->mask.as<uint8_t>() : nullptr;
    const int64_t npen = half ? (int64_t)(n1 / 2 + 1) * n2 : (int64_t)n1 * n2;
    j->csr_cnt.alloc(sizeof(uint32_t) * npen);
    j->csr_off.alloc(sizeof(int64_t) * (npen + 1));
    if (half) hcsr_count(j->M.as<float2>(), mask, n1, n2, n3, j->csr_cnt.as<uint32_t>(), st);
    else csr_count(j->M.as<float2>(), mask, n1, n2, n3, j->csr_cnt.as<uint32_t>(), st);
    grid_index_scan(j->csr_cnt.as<uint32_t>(), npen, j->csr_off.as<int64_t>(), j->csr_off.as<int64_t>() + npen, st);
    GF_HIP(hipMemcpyAsync(&j->nmeas, j->csr_off.as<int64_t>() + npen, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    GF_HIP(hipStreamSynchronize(st));
    const int64_t ne = std::max<int64_t>(j->nmeas, 1);
    j->csr_meta.alloc(sizeof(uint32_t) * ne);
    j->csr_val.alloc(sizeof(float2) * ne);
    if (half) {
        j->csr_mag.alloc(sizeof(float2) * ne);
        hcsr_fill(j->M.as<float2>(), mask, n1, n2, n3, j->csr_off.as<int64_t>(), j->csr_meta.as<uint32_t>(),
                  j->csr_val.as<float2>(), j->csr_mag.as<float2>(), st);
        hcsr_den(j->csr_mag.as<float2>(), j->csr_meta.as<uint32_t>(), j->nmeas, j->den_tmp.as<double>(), ntmp,
                 j->den.as<double>(), st);
    } else {
        csr_fill(j->M.as<float2>(), mask, n1, n2, n3, j->csr_off.as<int64_t>(), j->csr_meta.as<uint32_t>(),
                 j->csr_val.as<float2>(), st);
        csr_den(j->csr_val.as<float2>(), j->nmeas, j->den_tmp.as<double>(), ntmp, j->den.as<double>(), st);
    }
    // errK denominator: sum |M(errInd)| (reconstruct.m:135)
    double den = 0.0;
    GF_HIP(hipMemcpyAsync(&den, j->den.p, sizeof(double), hipMemcpyDeviceToHost, st));
    GF_HIP(hipStreamSynchronize(st));

    const int nR = monitor ? N / L : 0;
    j->errR_h.assign(nR, 0.0f);
    j->Rarr_h.assign((size_t)nR * j->P, 0.0);
    j->Rarr2_h.assign((size_t)nR * j->P, 0.0);
    j->t_rf = 0.0;

    ia.x = j->x.as<float>();
    ia.K = j->K.as<float2>();
    ia.S = j->S.as<uint8_t>();
    ia.csr_off = j->csr_off.as<int64_t>();
    ia.csr_meta = j->csr_meta.as<uint32_t>();
    ia.csr_val = j->csr_val.as<float2>();
    ia.err_part = j->err_part.as<double>();
    ia.positivity = j->prm.constraintPositivity ? 1 : 0;
    ia.support = j->prm.constraintSupport ? 1 : 0;
    ia.dr = dr ? 1 : 0;
    HalfIterArgs ha{};
    ha.x = ia.x;
    ha.A = j->K.as<float2>();
    ha.B = j->K.as<float2>() + hn;
    ha.S = ia.S;
    ha.csr_off = ia.csr_off;
    ha.csr_meta = ia.csr_meta;
    ha.csr_val = ia.csr_val;
    ha.csr_mag = j->csr_mag.as<float2>();
    ha.err_part = ia.err_part;
    ha.n1 = n1; ha.n2 = n2; ha.n3 = n3;
    ha.positivity = ia.positivity;
    ha.support = ia.support;
    ha.dr = ia.dr;

    GF_HIP(hipEventRecord(loop.a, st));
    size_t cur = 0;
    for (int it = 1; it <= N; ++it) {
        if (cur < sched.size() && it == sched[cur]) {          // :118-124
            size_t last = cur;
            for (size_t q = cur; q < sched.size(); ++q)
                if (sched[q] == it) last = q;
            if (half) hcsr_enforce(n1, n2, n3, j->csr_off.as<int64_t>(), j->csr_meta.as<uint32_t>(), j->delays[last], st);
            else csr_enforce(n1, n2, n3, j->csr_off.as<int64_t>(), j->csr_meta.as<uint32_t>(), j->delays[last], st);
            cur = last + 1;
        }
        // The schedule always ends with a change at it == N, which resets bestErr
        // (:123), so the returned reconstruction is P_R(x) of the last iteration:
        // it is written there and nowhere else.
        ia.best = (it == N) ? j->best.as<float>() : nullptr;
        const bool mon_it = monitor && (it % L == 0);
        ia.pf_out = (mon_it && dr) ? j->pfbuf.as<float>() : nullptr;
        std::pair<hipEvent_t, hipEvent_t> e_it, e_pass, e_fused;
        kstat_begin(j, GENFIRE_KSTAT_ITERATION, &e_it);
        if (half) {
            // P1 runs on its own only in the first iteration; later ones get it
            // from the fused P5 + P1 pass that ends the iteration before.
            if (it == 1) {
                ha.best = ia.best;
                ha.pf_out = nullptr;
                half_first(ha, st);
            }
            kstat_begin(j, GENFIRE_KSTAT_FFT_PASS, &e_pass);
            half_y(ha, 0, st);
            kstat_end(j, GENFIRE_KSTAT_FFT_PASS, e_pass);
            kstat_begin(j, GENFIRE_KSTAT_FFT_FUSED, &e_fused);
            half_z(ha, st);
            kstat_end(j, GENFIRE_KSTAT_FFT_FUSED, e_fused);
            half_y(ha, 1, st);
            ha.pf_out = mon_it ? j->pfbuf.as<float>() : nullptr;
            if (it < N) {
                ha.best = (it + 1 == N) ? j->best.as<float>() : nullptr;
                half_last_first(ha, st);
            } else {
                ha.best = nullptr;
                half_last(ha, st);
            }
        } else {
            iter_first(ia, st);
            kstat_begin(j, GENFIRE_KSTAT_FFT_PASS, &e_pass);
            fft_axis(j->K.as<float2>(), n1, n2, n3, 1, 0, st);
            kstat_end(j, GENFIRE_KSTAT_FFT_PASS, e_pass);
            kstat_begin(j, GENFIRE_KSTAT_FFT_FUSED, &e_fused);
            iter_fused(ia, st);
            kstat_end(j, GENFIRE_KSTAT_FFT_FUSED, e_fused);
            fft_axis(j->K.as<float2>(), n1, n2, n3, 1, 1, st);
            iter_last(ia, st);
        }
        iter_errk_reduce(j->err_part.as<double>(), nparts, den, j->errK.as<float>() + (it - 1), st);
        kstat_end(j, GENFIRE_KSTAT_ITERATION, e_it);
        if (mon_it) {
            const float* model = (dr || half) ? j->pfbuf.as<float>() : j->x.as<float>();
            run_rfactor(j, model, it / L - 1);
        }
    }
    GF_HIP(hipEventRecord(loop.b, st));

    // reconstruction = My_stripzero(fftshift(best), [Dim1 Dim2 Dim1]) (:192)
    j->recon.alloc(sizeof(float) * box_size(j));
    crop_centre(j->best.p, 0, n1, n2, n3, j->recon.as<float>(), j->dim1, j->dim2, j->dim1, st);
    j->errK_h.assign(N, 0.0f);
    GF_HIP(hipMemcpyAsync(j->errK_h.data(), j->errK.p, sizeof(float) * N, hipMemcpyDeviceToHost, st));
    GF_HIP(hipEventRecord(all.b, st));
    GF_HIP(hipEventSynchronize(all.b));
    float ms = 0.0f;
    GF_HIP(hipEventElapsedTime(&ms, loop.a, loop.b));
    j->t_iter = ms - j->t_rf;
    GF_HIP(hipEventElapsedTime(&ms, all.a, all.b));
    j->t_total = j->t_grid + ms;
    if (j->prof) j->ks.collect();
    j->reconstructed = true;
}

// R-factor monitor (reconstruct.m:168-185): model = crop(fftshift(x)) then, for
// every projection, R1 = sum|back-pj|/sum|pj| and R2 = ||back-pj||_F/||pj||_F.
void run_rfactor(genfire_job* j, const float* d_model_fft, int slot)
{
    hipStream_t st = j->st;
    const int P = j->P;
    EventTimer tm;
    tm.init();
    GF_HIP(hipEventRecord(tm.a, st));
    j->vol.alloc(sizeof(float) * box_size(j));
    crop_centre(d_model_fft, 0, j->n1, j->n2, j->n1, j->vol.as<float>(), j->dim1, j->dim2, j->dim1, st);
    if (!j->Rmat.p) {
        std::vector<float> Rm = proj_matrices(j->angles.data(), P, &j->rmat_any_y);
        j->Rmat.alloc(sizeof(float) * Rm.size());
        GF_HIP(hipMemcpyAsync(j->Rmat.p, Rm.data(), sizeof(float) * Rm.size(), hipMemcpyHostToDevice, st));
        GF_HIP(hipStreamSynchronize(st));
    }
    if (j->rmat_any_y) {
        j->vol_y.alloc(sizeof(float) * box_size(j));
        volume_yfast(j->vol.as<float>(), j->vol_y.as<float>(), j->dim1, j->dim2, j->dim1, st);
    }
    const int nblk = project_blocks(j->dim1, j->dim2);
    j->rsums.alloc(sizeof(double) * (size_t)P * 4 * nblk);
    j->rout.alloc(sizeof(double) * (size_t)P * 4);
    ProjArgs pa{};
    pa.vol = j->vol.as<float>();
    pa.vol_y = j->rmat_any_y ? j->vol_y.as<float>() : nullptr;
    pa.d1 = j->dim1; pa.d2 = j->dim2; pa.d3 = j->dim1;
    pa.R = j->Rmat.as<float>();
    pa.P = P;
    pa.out = nullptr;
    pa.meas = j->proj.as<float>();
    pa.m1 = j->dim1; pa.m2 = j->dim2;
    pa.sums = j->rsums.as<double>();
    pa.post_hoc = 0;
    std::pair<hipEvent_t, hipEvent_t> ev;
    kstat_begin(j, GENFIRE_KSTAT_PROJECT, &ev);
    project(pa, st);
    kstat_end(j, GENFIRE_KSTAT_PROJECT, ev);
    project_reduce(j->rsums.as<double>(), P, nblk, j->rout.as<double>(), st);
    std::vector<double> h((size_t)P * 4);
    GF_HIP(hipMemcpyAsync(h.data(), j->rout.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    GF_HIP(hipEventRecord(tm.b, st));
    GF_HIP(hipEventSynchronize(tm.b));
    float ms = 0.0f;
    GF_HIP(hipEventElapsedTime(&ms, tm.a, tm.b));
    j->t_rf += ms;
    double mean = 0.0;
    for (int p = 0; p < P; ++p) {
        const double r1 = h[4 * p + 0] / h[4 * p + 2];
        const double r2 = std::sqrt(h[4 * p + 1]) / std::sqrt(h[4 * p + 3]);
        j->Rarr_h[(size_t)slot * P + p] = r1;
        j->Rarr2_h[(size_t)slot * P + p] = r2;
        mean += r1;
    }
    j->errR_h[slot] = (float)(mean / P);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int genfire_version(void) { return 100; }

int genfire_params_default(genfire_params* p)
{
    if (!p) return GENFIRE_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    p->numIterations = 50;
    p->oversamplingRatio = 3.0;
    p->interpolationCutoffDistance = 0.7;
    p->constraintEnforcementMode = 1;
    p->constraintPositivity = 1;
    p->constraintSupport = 1;
    p->DFT_CentroSymmetricity = 1;
    p->vector1[2] = 1.0;
    p->vector2[1] = 1.0;
    p->vector3[0] = 1.0;
    p->monitor_R = 0;
    p->monitorR_loopLength = 10;
    p->algorithm = GENFIRE_ALG_GENFIRE;
    p->precision = GENFIRE_PREC_FP32;
    p->compute_recIFFT = 1;
    return GENFIRE_OK;
}

int genfire_ctx_create(int32_t device, genfire_ctx** out)
{
    if (!out) return GENFIRE_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GENFIRE_ERR_HIP;
    if (device < 0 || device >= n) return GENFIRE_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return GENFIRE_ERR_HIP;
    auto* c = new (std::nothrow) genfire_ctx();
    if (!c) return GENFIRE_ERR_NOMEM;
    c->device = device;
    *out = c;
    return GENFIRE_OK;
}

int genfire_ctx_destroy(genfire_ctx* ctx)
{
    delete ctx;
    return GENFIRE_OK;
}

const char* genfire_last_error(const genfire_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : "no context";
}

int genfire_job_create(genfire_ctx* ctx, const genfire_params* p, int32_t dim1, int32_t dim2,
                       int32_t num_projs, genfire_job** out)
{
    if (!ctx || !p || !out) return GENFIRE_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        GF_CHECK(dim1 >= 2 && dim2 >= 2 && num_projs >= 1, GENFIRE_ERR_ARG,
                 "GENFIRE: projections must be at least 2 x 2 and P >= 1");
        GF_CHECK(p->numIterations >= 1, GENFIRE_ERR_ARG, "GENFIRE: numIterations must be >= 1");
        GF_CHECK(p->oversamplingRatio > 0, GENFIRE_ERR_ARG, "GENFIRE: oversamplingRatio must be positive");
        GF_CHECK(p->algorithm == GENFIRE_ALG_GENFIRE || p->algorithm == GENFIRE_ALG_DR, GENFIRE_ERR_ARG,
                 "GENFIRE: unknown algorithm");
        GF_CHECK(p->precision == GENFIRE_PREC_FP32 || p->precision == GENFIRE_PREC_SPLIT_BF16 ||
                     p->precision == GENFIRE_PREC_FP16,
                 GENFIRE_ERR_ARG, "GENFIRE: unknown gridding precision");
        use_device(ctx);
        auto* j = new genfire_job();
        std::unique_ptr<genfire_job> guard(j);
        j->ctx = ctx;
        j->prm = *p;
        j->dim1 = dim1;
        j->dim2 = dim2;
        j->P = num_projs;
        j->n1 = (int)mround(dim1 * p->oversamplingRatio);   // :168
        j->n2 = (int)mround(dim2 * p->oversamplingRatio);   // :169
        GF_CHECK(j->n1 >= dim1 && j->n2 >= dim2, GENFIRE_ERR_ARG, "GENFIRE: oversampled size smaller than projection");
        GF_CHECK(fft_supported(j->n1) && fft_supported(j->n2), GENFIRE_ERR_UNSUPPORTED,
                 "GENFIRE: oversampled grid " + std::to_string(j->n1) + "x" + std::to_string(j->n2) +
                     " has an FFT length this build does not support");
        j->hg = half_grid(j->n1, j->n2, p->DFT_CentroSymmetricity ? 1 : 0);
        j->delays = delay_indicators(p->constraintEnforcementMode);
        GF_HIP(hipStreamCreateWithFlags(&j->st, hipStreamNonBlocking));
        *out = guard.release();
    });
}

int genfire_job_destroy(genfire_job* job)
{
    if (!job) return GENFIRE_OK;
    (void)hipSetDevice(job->ctx->device);
    delete job;
    return GENFIRE_OK;
}

int genfire_job_sizes(const genfire_job* job, int32_t* n1, int32_t* n2, int64_t half_len[3])
{
    if (!job) return GENFIRE_ERR_ARG;
    if (n1) *n1 = job->n1;
    if (n2) *n2 = job->n2;
    if (half_len)
        for (int i = 0; i < 3; ++i) half_len[i] = job->hg.len[i];
    return GENFIRE_OK;
}

int genfire_job_set_inputs(genfire_job* j, const float* projections, const double* angles,
                           const uint8_t* support, const float* initial_object)
{
    if (!j) return GENFIRE_ERR_ARG;
    if (!projections || !angles) {
        set_err(j->ctx, "GENFIRE: projections and angles are required");
        return GENFIRE_ERR_ARG;
    }
    return guarded(j->ctx, [&] {
        use_device(j->ctx);
        const size_t np = (size_t)j->dim1 * j->dim2 * j->P;
        j->proj.alloc(sizeof(float) * np);
        GF_HIP(hipMemcpyAsync(j->proj.p, projections, sizeof(float) * np, hipMemcpyHostToDevice, j->st));
        j->angles.assign(angles, angles + 3 * (size_t)j->P);
        j->have_support = support != nullptr;
        if (support) {
            j->support_box.alloc(box_size(j));
            GF_HIP(hipMemcpyAsync(j->support_box.p, support, box_size(j), hipMemcpyHostToDevice, j->st));
        }
        j->have_init = initial_object != nullptr;
        if (initial_object) {
            j->init_box.alloc(sizeof(float) * box_size(j));
            GF_HIP(hipMemcpyAsync(j->init_box.p, initial_object, sizeof(float) * box_size(j),
                                  hipMemcpyHostToDevice, j->st));
        }
        GF_HIP(hipStreamSynchronize(j->st));
        j->Rmat.release();
        j->have_inputs = true;
        j->gridded = false;
        j->reconstructed = false;
    });
}

int genfire_job_grid(genfire_job* j)
{
    if (!j) return GENFIRE_ERR_ARG;
    return guarded(j->ctx, [&] {
        use_device(j->ctx);
        do_grid(j);
        if (j->prof) j->ks.collect();
    });
}

int genfire_job_grid_count(const genfire_job* j, int64_t* n)
{
    if (!j || !n) return GENFIRE_ERR_ARG;
    if (!j->gridded) return GENFIRE_ERR_STATE;
    *n = j->npts;
    return GENFIRE_OK;
}

int genfire_job_get_points(genfire_job* j, int64_t* ind, int32_t* proj, float* dist, float* un, float* vn,
                           float* values)
{
    if (!j) return GENFIRE_ERR_ARG;
    return guarded(j->ctx, [&] {
        GF_CHECK(j->gridded, GENFIRE_ERR_STATE, "GENFIRE: not gridded");
        use_device(j->ctx);
        const int64_t n = j->npts;
        if (n == 0) return;
        if (ind) {
            std::vector<uint32_t> t(n);
            GF_HIP(hipMemcpy(t.data(), j->ind.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) ind[i] = t[i];
        }
        if (proj)
            for (int p = 0; p < j->P; ++p)
                for (int64_t i = j->proj_off[p]; i < j->proj_off[p + 1]; ++i) proj[i] = p;
        if (dist) GF_HIP(hipMemcpy(dist, j->dist.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        if (un) GF_HIP(hipMemcpy(un, j->un.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        if (vn) GF_HIP(hipMemcpy(vn, j->vn.p, sizeof(float) * n, hipMemcpyDeviceToHost));
        if (values) GF_HIP(hipMemcpy(values, j->F.p, sizeof(float2) * n, hipMemcpyDeviceToHost));
    });
}

int genfire_job_get_measuredK(genfire_job* j, float* out)
{
    if (!j || !out) return GENFIRE_ERR_ARG;
    return guarded(j->ctx, [&] {
        GF_CHECK(j->gridded, GENFIRE_ERR_STATE, "GENFIRE: not gridded");
        use_device(j->ctx);
        DevBuf tmp;
        tmp.alloc(sizeof(float2) * full_size(j));
        shift_complex(j->M.as<float2>(), tmp.as<float2>(), j->n1, j->n2, j->n1, 1, j->st);
        GF_HIP(hipMemcpyAsync(out, tmp.p, tmp.bytes, hipMemcpyDeviceToHost, j->st));
        GF_HIP(hipStreamSynchronize(j->st));
    });
}

int genfire_job_get_hits(genfire_job* j, int32_t* out)
{
    if (!j || !out) return GENFIRE_ERR_ARG;
    return guarded(j->ctx, [&] {
        GF_CHECK(j->gridded && j->hits.p, GENFIRE_ERR_STATE, "GENFIRE: not gridded");
        use_device(j->ctx);
        const int64_t n = j->hg.total();
        std::vector<uint32_t> t(n);
        GF_HIP(hipMemcpy(t.data(), j->hits.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) out[i] = (int32_t)(t[i] & ~kExactBit);
    });
}

int genfire_job_get_recIFFT(genfire_job* j, float* out)
{
    if (!j || !out) return GENFIRE_ERR_ARG;
    return guarded(j->ctx, [&] {
        GF_CHECK(j->gridded && j->recifft.p, GENFIRE_ERR_STATE, "GENFIRE: recIFFT not computed");
        use_device(j->ctx);
        GF_HIP(hipMemcpy(out, j->recifft.p, j->recifft.bytes, hipMemcpyDeviceToHost));
    });
}

int genfire_job_set_measuredK(genfire_job* j, const float* mk, const uint8_t* mask)
{
    if (!j || !mk) return GENFIRE_ERR_ARG;
    return guarded(j->ctx, [&] {
        use_device(j->ctx);
        const int64_t nf = full_size(j);
        DevBuf tmp;
        tmp.alloc(sizeof(float2) * nf);
        GF_HIP(hipMemcpyAsync(tmp.p, mk, tmp.bytes, hipMemcpyHostToDevice, j->st));
        j->M.alloc(sizeof(float2) * nf);
        shift_complex(tmp.as<float2>(), j->M.as<float2>(), j->n1, j->n2, j->n1, 0, j->st);
        j->have_mask = mask != nullptr;
        if (mask) {
            // measuredK_mask in FFT order: voxel s holds the centred voxel (s + floor(n/2)) mod n
            const int a1 = j->n1, a2 = j->n2, a3 = j->n1;
            std::vector<uint8_t> m8(nf);
            for (int64_t s = 0; s < nf; ++s) {
                const int s1 = (int)(s % a1), s2 = (int)((s / a1) % a2), s3 = (int)(s / ((int64_t)a1 * a2));
                const int64_t c = (int64_t)((s1 + a1 / 2) % a1) +
                                  (int64_t)a1 * (((s2 + a2 / 2) % a2) + (int64_t)a2 * ((s3 + a3 / 2) % a3));
                m8[s] = mask[c] != 0;
            }
            j->mask.alloc(nf);
            GF_HIP(hipMemcpyAsync(j->mask.p, m8.data(), nf, hipMemcpyHostToDevice, j->st));
            GF_HIP(hipStreamSynchronize(j->st));
        }
        GF_HIP(hipStreamSynchronize(j->st));
        j->gridded = true;
        j->reconstructed = false;
    });
}

int genfire_job_reconstruct(genfire_job* j)
{
    if (!j) return GENFIRE_ERR_ARG;
    return guarded(j->ctx, [&] {
        use_device(j->ctx);
        do_reconstruct(j);
    });
}

int genfire_job_get_reconstruction(genfire_job* j, float* out)
{
    if (!j || !out) return GENFIRE_ERR_ARG;
    return guarded(j->ctx, [&] {
        GF_CHECK(j->reconstructed, GENFIRE_ERR_STATE, "GENFIRE: not reconstructed");
        use_device(j->ctx);
        GF_HIP(hipMemcpy(out, j->recon.p, j->recon.bytes, hipMemcpyDeviceToHost));
    });
}

int genfire_job_get_errK(genfire_job* j, float* out)
{
    if (!j || !out) return GENFIRE_ERR_ARG;
    if (!j->reconstructed) return GENFIRE_ERR_STATE;
    std::memcpy(out, j->errK_h.data(), sizeof(float) * j->errK_h.size());
    return GENFIRE_OK;
}

int genfire_job_get_errR(genfire_job* j, float* errR, double* Rarr, double* Rarr2)
{
    if (!j) return GENFIRE_ERR_ARG;
    if (!j->reconstructed) return GENFIRE_ERR_STATE;
    const size_t nR = j->errR_h.size();
    if (errR) std::memcpy(errR, j->errR_h.data(), sizeof(float) * nR);
    // stored slot-major (slot*P + p): exactly the column-major P x nR layout
    if (Rarr) std::memcpy(Rarr, j->Rarr_h.data(), sizeof(double) * j->Rarr_h.size());
    if (Rarr2) std::memcpy(Rarr2, j->Rarr2_h.data(), sizeof(double) * j->Rarr2_h.size());
    return GENFIRE_OK;
}

int genfire_job_submit(genfire_job* j)
{
    if (!j) return GENFIRE_ERR_ARG;
    if (j->async_state.load() == 1) {
        set_err(j->ctx, "GENFIRE: job already running");
        return GENFIRE_ERR_STATE;
    }
    if (j->worker.joinable()) j->worker.join();
    j->async_state = 1;
    j->worker = std::thread([j] {
        int rc = genfire_job_grid(j);
        if (rc == GENFIRE_OK) rc = genfire_job_reconstruct(j);
        j->async_status = rc;
        j->async_state = 2;
    });
    return GENFIRE_OK;
}

int genfire_job_poll(genfire_job* j, int32_t* done)
{
    if (!j || !done) return GENFIRE_ERR_ARG;
    *done = j->async_state.load() == 2 ? 1 : 0;
    return GENFIRE_OK;
}

int genfire_job_wait(genfire_job* j)
{
    if (!j) return GENFIRE_ERR_ARG;
    if (j->worker.joinable()) j->worker.join();
    if (j->async_state.load() == 0) return GENFIRE_OK;
    j->async_state = 0;
    return j->async_status;
}

int genfire_job_timing(const genfire_job* j, double ms[4])
{
    if (!j || !ms) return GENFIRE_ERR_ARG;
    ms[0] = j->t_grid;
    ms[1] = j->t_iter;
    ms[2] = j->t_rf;
    ms[3] = j->t_total;
    return GENFIRE_OK;
}

int genfire_job_set_profiling(genfire_job* j, int32_t on)
{
    if (!j) return GENFIRE_ERR_ARG;
    j->prof = on != 0;
    j->ks.reset();
    return GENFIRE_OK;
}

int genfire_job_kernel_stats(const genfire_job* j, int32_t k, int64_t* launches, double* total_ms)
{
    if (!j || k < 0 || k >= GENFIRE_KSTAT_COUNT) return GENFIRE_ERR_ARG;
    if (launches) *launches = j->ks.launches[k];
    if (total_ms) *total_ms = j->ks.ms[k];
    return GENFIRE_OK;
}

int genfire_project(genfire_ctx* ctx, const float* vol, int32_t d1, int32_t d2, int32_t d3, double phi,
                    double theta, double psi, float* proj)
{
    if (!ctx || !vol || !proj) return GENFIRE_ERR_ARG;
    return guarded(ctx, [&] {
        use_device(ctx);
        GF_CHECK(d1 >= 2 && d2 >= 2 && d3 >= 2, GENFIRE_ERR_ARG, "GENFIRE: volume dimensions must be >= 2");
        hipStream_t st;
        GF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        DevBuf v, vy, r, o;
        const size_t nv = (size_t)d1 * d2 * d3;
        v.alloc(sizeof(float) * nv);
        o.alloc(sizeof(float) * (size_t)d1 * d2);
        const double ang[3] = {phi, theta, psi};
        bool any_y = false;
        std::vector<float> Rf = proj_matrices(ang, 1, &any_y);
        r.alloc(sizeof(float) * Rf.size());
        GF_HIP(hipMemcpyAsync(v.p, vol, sizeof(float) * nv, hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(r.p, Rf.data(), sizeof(float) * Rf.size(), hipMemcpyHostToDevice, st));
        if (any_y) {
            vy.alloc(sizeof(float) * nv);
            volume_yfast(v.as<float>(), vy.as<float>(), d1, d2, d3, st);
        }
        ProjArgs pa{};
        pa.vol = v.as<float>();
        pa.vol_y = vy.as<float>();
        pa.d1 = d1; pa.d2 = d2; pa.d3 = d3;
        pa.R = r.as<float>();
        pa.P = 1;
        pa.out = o.as<float>();
        pa.m1 = d1; pa.m2 = d2;
        project(pa, st);
        GF_HIP(hipMemcpyAsync(proj, o.p, o.bytes, hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        GF_HIP(hipStreamDestroy(st));
    });
}

int genfire_rfactor(genfire_ctx* ctx, const float* model, int32_t d1, int32_t d2, int32_t d3,
                    const float* projections, int32_t p1, int32_t p2, int32_t P, const double* angles,
                    double* Rtotal, double* Rarr, float* simu)
{
    if (!ctx || !model || !projections || !angles || P < 1) return GENFIRE_ERR_ARG;
    return guarded(ctx, [&] {
        use_device(ctx);
        GF_CHECK(p1 <= d1 && p2 <= d2, GENFIRE_ERR_ARG, "GENFIRE: projection larger than the model");
        hipStream_t st;
        GF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        DevBuf v, vy, r, o, m, s, s4;
        const size_t nv = (size_t)d1 * d2 * d3, np = (size_t)p1 * p2 * P;
        v.alloc(sizeof(float) * nv);
        m.alloc(sizeof(float) * np);
        o.alloc(sizeof(float) * np);
        bool any_y = false;
        std::vector<float> Rf = proj_matrices(angles, P, &any_y);
        r.alloc(sizeof(float) * Rf.size());
        GF_HIP(hipMemcpyAsync(v.p, model, sizeof(float) * nv, hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(m.p, projections, sizeof(float) * np, hipMemcpyHostToDevice, st));
        GF_HIP(hipMemcpyAsync(r.p, Rf.data(), sizeof(float) * Rf.size(), hipMemcpyHostToDevice, st));
        if (any_y) {
            vy.alloc(sizeof(float) * nv);
            volume_yfast(v.as<float>(), vy.as<float>(), d1, d2, d3, st);
        }
        const int nblk = project_blocks(p1, p2);
        s.alloc(sizeof(double) * (size_t)P * 4 * nblk);
        s4.alloc(sizeof(double) * (size_t)P * 4);
        ProjArgs pa{};
        pa.vol = v.as<float>();
        pa.vol_y = vy.as<float>();
        pa.d1 = d1; pa.d2 = d2; pa.d3 = d3;
        pa.R = r.as<float>();
        pa.P = P;
        pa.out = o.as<float>();
        pa.meas = m.as<float>();
        pa.m1 = p1; pa.m2 = p2;
        pa.sums = s.as<double>();
        pa.post_hoc = 1;
        project(pa, st);
        project_reduce(s.as<double>(), P, nblk, s4.as<double>(), st);
        std::vector<double> h((size_t)P * 4);
        GF_HIP(hipMemcpyAsync(h.data(), s4.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
        if (simu) GF_HIP(hipMemcpyAsync(simu, o.p, o.bytes, hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        GF_HIP(hipStreamDestroy(st));
        double mean = 0.0;
        for (int p = 0; p < P; ++p) {
            const double rr = h[4 * p] / h[4 * p + 2];
            if (Rarr) Rarr[p] = rr;
            mean += rr;
        }
        if (Rtotal) *Rtotal = mean / P;
    });
}

int genfire_fftn(genfire_ctx* ctx, float* data, int32_t n1, int32_t n2, int32_t n3, int32_t inverse)
{
    if (!ctx || !data) return GENFIRE_ERR_ARG;
    return guarded(ctx, [&] {
        use_device(ctx);
        GF_CHECK(n1 >= 1 && n2 >= 1 && n3 >= 1, GENFIRE_ERR_ARG, "GENFIRE: bad FFT size");
        GF_CHECK((n1 == 1 || fft_supported(n1)) && (n2 == 1 || fft_supported(n2)) && (n3 == 1 || fft_supported(n3)),
                 GENFIRE_ERR_UNSUPPORTED, "GENFIRE: FFT length not supported");
        hipStream_t st;
        GF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        DevBuf d;
        const size_t n = (size_t)n1 * n2 * n3;
        d.alloc(sizeof(float2) * n);
        GF_HIP(hipMemcpyAsync(d.p, data, sizeof(float2) * n, hipMemcpyHostToDevice, st));
        fft3(d.as<float2>(), n1, n2, n3, inverse ? 1 : 0, st);
        GF_HIP(hipMemcpyAsync(data, d.p, sizeof(float2) * n, hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        GF_HIP(hipStreamDestroy(st));
    });
}

}  // extern "C"
