// gf_job.hip -- the phases of one reconstruction job: gridding
// (fillInFourierGrid_DFT.m:85-310 + CentroSymmetricity.m:2-72 + runGridding,
// GENFIRE_Reconstructor.m:172-186) and the iteration (reconstruct.m:28-194 and
// the Douglas-Rachford variant), each on the job's HIP stream.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>

#include "gf_job.h"

namespace gf {

[[noreturn]] void raise(int code, const std::string& msg) { throw Error{code, msg}; }

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

namespace {

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

}  // namespace

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

std::pair<hipEvent_t, hipEvent_t> KStats::take()
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

void KStats::collect()
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

void KStats::reset()
{
    for (int k = 0; k < GENFIRE_KSTAT_COUNT; ++k) { launches[k] = 0; ms[k] = 0.0; }
}

KStats::~KStats()
{
    for (auto& e : pool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    for (auto& v : pending)
        for (auto& e : v) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
}

void set_err(genfire_ctx* ctx, const std::string& m)
{
    if (!ctx) return;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->err = m;
}

void use_device(genfire_ctx* ctx) { GF_HIP(hipSetDevice(ctx->device)); }

void kstat_begin(genfire_job* j, int k, std::pair<hipEvent_t, hipEvent_t>* ev)
{
    if (!j->prof) return;
    *ev = j->ks.take();
    GF_HIP(hipEventRecord(ev->first, j->run_st));
    (void)k;
}

void kstat_end(genfire_job* j, int k, std::pair<hipEvent_t, hipEvent_t>& ev)
{
    if (!j->prof) return;
    GF_HIP(hipEventRecord(ev.second, j->run_st));
    j->ks.pending[k].push_back(ev);
}

// ZYZ matrices of the projector (calculate3Dprojection_RealSpaceinterp.m:13-15),
// kRStride floats per projection with the y-fastest flag at [9], the exact-y
// flag at [10] and the lane-axis flag at [11], into Rm (P x kRStride floats).
void proj_matrices_into(const double* angles, int P, const double* vectors, float* Rm)
{
    for (int p = 0; p < P; ++p) {
        double R[9];
        if (vectors)   // src/calculate3Dprojection_RealSpaceinterp_general.m:3-7
            rotation_f64(vectors, vectors + 3, vectors + 6, angles[p], angles[P + p], angles[2 * P + p], R);
        else
            zyz_matrix(angles[p], angles[P + p], angles[2 * P + p], R);
        float* r = Rm + (size_t)kRStride * p;
        for (int q = 0; q < kRStride; ++q) r[q] = q < 9 ? (float)R[q] : 0.0f;
        if (project_y_exact(r)) {
            r[9] = 1.0f; r[10] = 1.0f; r[11] = 1.0f;
        } else {
            project_orientation(r);
        }
    }
}

std::vector<float> proj_matrices(const double* angles, int P, bool* any_y, const double* vectors)
{
    std::vector<float> Rm((size_t)kRStride * P, 0.0f);
    proj_matrices_into(angles, P, vectors, Rm.data());
    *any_y = false;
    for (int p = 0; p < P; ++p) *any_y = *any_y || Rm[(size_t)kRStride * p + 9] != 0.0f;
    return Rm;
}

// ---------------------------------------------------------------------------
// fillInFourierGrid_DFT + CentroSymmetricity + runGridding post-processing
// ---------------------------------------------------------------------------
namespace {

// fillInFourierGrid_DFT.m:85-291 + CentroSymmetricity: measuredK into j->M.
void grid_dft_body(genfire_job* j, hipStream_t st)
{
    const int P = j->P, n1 = j->n1, n2 = j->n2;
    const int centro = j->prm.DFT_CentroSymmetricity ? 1 : 0;

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
    j->scan_tmp.ensure(grid_index_scan_scratch(nb_total));
    grid_index_scan(j->blk_counts.as<uint32_t>(), nb_total, j->blk_offsets.as<int64_t>(),
                    j->total.as<int64_t>(), j->scan_tmp.as<int64_t>(), st);
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

    // DFT values (:236-248).  FP32: the type-2 non-uniform FFT where the padded
    // grid has a supported FFT length (both edges up to 512), else the direct
    // sum as an MFMA GEMM; GENFIRE_GRID_KERNEL=mfma or =valu selects the direct
    // sum on MFMA or on the VALU (256-point tiles).  The reduced precisions
    // always run the direct sum on MFMA.  64-point tiles otherwise.
    const int prec = j->prm.precision;
    const char* gk = std::getenv("GENFIRE_GRID_KERNEL");
    const std::string gks = gk ? gk : "";
    const bool use_valu = prec == GENFIRE_PREC_FP32 && gks == "valu";
    const int nu1 = nufft_fine_len(j->dim1), nu2 = nufft_fine_len(j->dim2);
    const bool use_nufft = prec == GENFIRE_PREC_FP32 && !use_valu && gks != "mfma" && nu1 > 0 && nu2 > 0;
    const int tile_pts = use_valu ? 256 : mfma_tile_points();
    // tiles of tile_pts points within one projection, generated on the device
    // from the projection offsets and the first tile of each projection (kept
    // in the job: the asynchronous copy may read them after this scope)
    std::vector<int64_t>& tm = j->tile_meta_h;
    tm.assign(2 * (size_t)(P + 1), 0);
    for (int p = 0; p <= P; ++p) tm[p] = j->proj_off[p];
    for (int p = 0; p < P; ++p)
        tm[P + 1 + p + 1] = tm[P + 1 + p] + (j->proj_off[p + 1] - j->proj_off[p] + tile_pts - 1) / tile_pts;
    const int64_t ntiles = tm[2 * (size_t)P + 1];
    GF_CHECK(ntiles < ((int64_t)1 << 31), GENFIRE_ERR_UNSUPPORTED, "gridding: too many point tiles");
    if (ntiles > 0) {
        j->tile_meta.alloc(sizeof(int64_t) * tm.size());
        GF_HIP(hipMemcpyAsync(j->tile_meta.p, tm.data(), sizeof(int64_t) * tm.size(), hipMemcpyHostToDevice, st));
        j->tiles.alloc(sizeof(PointTile) * ntiles);
        grid_make_tiles(j->tile_meta.as<int64_t>(), j->tile_meta.as<int64_t>() + P + 1, P, tile_pts,
                        j->tiles.as<PointTile>(), (int)ntiles, st);
        if (prec != GENFIRE_PREC_FP32) {
            const size_t ne = grid_prep_lp_elems(j->dim1, j->dim2, P);
            j->Phi.alloc(sizeof(uint16_t) * ne);
            if (prec == GENFIRE_PREC_SPLIT_BF16) j->Plo.alloc(sizeof(uint16_t) * ne);
            j->Pscale.alloc(sizeof(float) * P);
            grid_prep_lp(prec, j->proj.as<float>(), j->dim1, j->dim2, P, j->Phi.as<uint16_t>(),
                         j->Plo.as<uint16_t>(), j->Pscale.as<float>(), st);
        } else if (use_nufft) {
            // the deconvolution factors depend on the sizes only: computed and
            // uploaded at the job's first gridding (kept in the job, as the
            // tiles are, for the asynchronous copy)
            if (!j->nu_inv.p) {
                j->nu_inv_h = nufft_deconv(j->dim1, nu1);
                const std::vector<float> inv2 = nufft_deconv(j->dim2, nu2);
                j->nu_inv_h.insert(j->nu_inv_h.end(), inv2.begin(), inv2.end());
                j->nu_inv.alloc(sizeof(float) * j->nu_inv_h.size());
                GF_HIP(hipMemcpyAsync(j->nu_inv.p, j->nu_inv_h.data(), sizeof(float) * j->nu_inv_h.size(),
                                      hipMemcpyHostToDevice, st));
            }
            j->Pt.alloc(sizeof(float2) * (size_t)nu1 * nu2 * P);   // the padded spectra
        } else if (!use_valu) {
            j->Pt.alloc(sizeof(float) * grid_prep_floats(j->dim1, j->dim2, P));
            grid_prep_proj(j->proj.as<float>(), j->dim1, j->dim2, P, j->Pt.as<float>(), st);
        }
        std::pair<hipEvent_t, hipEvent_t> ev;
        kstat_begin(j, GENFIRE_KSTAT_GRID_VALUES, &ev);
        if (use_nufft) {
            float2* G = j->Pt.as<float2>();
            nufft_load(j->proj.as<float>(), j->dim1, j->dim2, P, nu1, nu2, j->nu_inv.as<float>(),
                       j->nu_inv.as<float>() + j->dim1, G, st);
            fft_axis(G, nu1, nu2, P, 0, 0, st);
            fft_axis(G, nu1, nu2, P, 1, 0, st);
            nufft_interp(G, nu1, nu2, nufft_beta(j->dim1, nu1), nufft_beta(j->dim2, nu2), j->un.as<float>(),
                         j->vn.as<float>(), j->tiles.as<PointTile>(), (int)ntiles, j->F.as<float2>(), st);
        } else if (prec != GENFIRE_PREC_FP32)
            grid_values_mfma_lp(prec, j->Phi.as<uint16_t>(), j->Plo.as<uint16_t>(), j->Pscale.as<float>(), j->dim1,
                                j->dim2, j->un.as<float>(), j->vn.as<float>(), j->tiles.as<PointTile>(),
                                (int)ntiles, j->F.as<float2>(), st);
        else if (use_valu)
            grid_values_fp32(j->proj.as<float>(), j->dim1, j->dim2, j->un.as<float>(), j->vn.as<float>(),
                             j->tiles.as<PointTile>(), (int)ntiles, j->F.as<float2>(), st);
        else
            grid_values_mfma_fp32(j->Pt.as<float>(), j->dim1, j->dim2, j->un.as<float>(), j->vn.as<float>(),
                                  j->tiles.as<PointTile>(), (int)ntiles, j->F.as<float2>(), st);
        kstat_end(j, GENFIRE_KSTAT_GRID_VALUES, ev);
    }

    // weighted average (:267-291), projection by projection in list order
    const int64_t ht = j->hg.total();
    j->acc.alloc(sizeof(double2) * ht);
    j->sw.alloc(sizeof(double) * ht);
    j->hits.alloc(sizeof(uint32_t) * ht);
    j->exact.alloc(sizeof(float2) * ht);
    GF_HIP(hipMemsetAsync(j->hits.p, 0, j->hits.bytes, st));   // acc and sw are written by each voxel's first point
    for (int p = 0; p < P; ++p)
        grid_scatter(j->ind.as<uint32_t>(), j->dist.as<float>(), j->F.as<float2>(), j->proj_off[p],
                     j->proj_off[p + 1] - j->proj_off[p], j->acc.as<double2>(), j->sw.as<double>(),
                     j->hits.as<uint32_t>(), j->exact.as<float2>(), st);

    // CentroSymmetricity + single cast, straight into FFT order
    j->M.alloc(sizeof(float2) * full_size(j));
    grid_finalize(j->hg, n1, n2, centro, j->acc.as<double2>(), j->sw.as<double>(), j->hits.as<uint32_t>(),
                  j->exact.as<float2>(), j->M.as<float2>(), st);
}

// fillInFourierGrid.m:27-163 (FFT gridding) + hermitianSymmetrize: measuredK into
// j->M, hits over the centred d^3 grid.
void grid_fft_body(genfire_job* j, hipStream_t st)
{
    const int P = j->P, d = j->n1;
    const int64_t plane = (int64_t)d * d, vox = plane * d;
    DevBuf spec;
    spec.alloc(sizeof(float2) * plane * P);
    fftgrid_load(j->proj.as<float>(), j->dim1, j->dim2, P, d, spec.as<float2>(), st);   // :59-61
    fft_axis(spec.as<float2>(), d, d, P, 0, 0, st);
    fft_axis(spec.as<float2>(), d, d, P, 1, 0, st);

    j->acc.alloc(sizeof(double2) * vox);
    j->sw.alloc(sizeof(double) * vox);
    j->hits.alloc(sizeof(uint32_t) * vox);
    GF_HIP(hipMemsetAsync(j->acc.p, 0, j->acc.bytes, st));
    GF_HIP(hipMemsetAsync(j->sw.p, 0, j->sw.bytes, st));
    GF_HIP(hipMemsetAsync(j->hits.p, 0, j->hits.bytes, st));

    // shiftMax (:117-121); a matched point lies within the cutoff and, for
    // shiftMax = 0, inside the voxel's unit cube (radius sqrt(3)/2)
    const double cutoff = j->prm.interpolationCutoffDistance;
    const int shift_max = j->prm.allowMultipleGridMatches ? (int)mround(cutoff) : 0;
    const double radius = std::min(cutoff, (shift_max + 0.5) * std::sqrt(3.0));
    const float rho = (float)(radius + 0.02 + 2e-5 * d);
    for (int p = 0; p < P; ++p) {
        double R[9];
        zyz_matrix(j->angles[p], j->angles[P + p], j->angles[2 * P + p], R);   // :80-84 (R' of the projector's)
        FftPlane pl{};
        for (int i = 0; i < 6; ++i) pl.rot[i] = (float)R[i];
        for (int i = 0; i < 3; ++i) pl.nrm[i] = (float)R[6 + i];
        pl.axis = 0;
        for (int i = 1; i < 3; ++i)
            if (std::fabs(pl.nrm[i]) > std::fabs(pl.nrm[pl.axis])) pl.axis = i;
        fftgrid_scatter(pl, d, shift_max, rho, cutoff, spec.as<float2>() + plane * p, j->acc.as<double2>(),
                        j->sw.as<double>(), j->hits.as<uint32_t>(), st);
    }
    j->M.alloc(sizeof(float2) * full_size(j));
    fftgrid_finalize(d, j->acc.as<double2>(), j->sw.as<double>(), j->M.as<float2>(), st);   // :157-163
    j->npts = 0;
    GF_HIP(hipStreamSynchronize(st));   // spec is freed on return
}

}  // namespace

void do_grid(genfire_job* j)
{
    GF_CHECK(j->have_inputs, GENFIRE_ERR_STATE, "GENFIRE: inputs not set (call genfire_job_set_inputs)");
    const int n1 = j->n1, n2 = j->n2;
    hipStream_t st = j->st;
    EventTimer tm;
    tm.init();
    GF_HIP(hipEventRecord(tm.a, st));
    if (j->prm.griddingMethod == 1)
        grid_fft_body(j, st);
    else
        grid_dft_body(j, st);
    j->have_mask = false;
    j->mask.release();

    // recIFFT = crop(real(fftshift(ifftn(ifftshift(measuredK))))) (:304)
    if (j->prm.compute_recIFFT) {
        // z, then y on the rows with z in the Dim1 box, then x on the rows with y
        // and z in the box: the crop reads nothing else
        j->K.ensure(sizeof(float2) * full_size(j));   // the iteration's spectrum buffer, when it exists
        fft_axis_from(j->M.as<float2>(), j->K.as<float2>(), n1, n2, n1, 2, 1, st);   // M stays as gridded
        fft_axis_rows(j->K.as<float2>(), n1, n2, n1, 1, 1, 0, j->dim1, st);
        fft_axis_rows(j->K.as<float2>(), n1, n2, n1, 0, 1, j->dim2, j->dim1, st);
        j->recifft.alloc(sizeof(float) * box_size(j));
        crop_centre(j->K.p, 1, n1, n2, n1, j->recifft.as<float>(), j->dim1, j->dim2, j->dim1, st);
    }
    GF_HIP(hipEventRecord(tm.b, st));
    GF_HIP(hipEventSynchronize(tm.b));
    float ms = 0.0f;
    GF_HIP(hipEventElapsedTime(&ms, tm.a, tm.b));
    j->t_grid = ms;
    // The FP64 accumulators and the value operands stay allocated for the
    // job's next gridding (a request server frees them before the iteration
    // set-up, release_grid_scratch).
    j->gridded = true;
    j->reconstructed = false;
}

// ---------------------------------------------------------------------------
// reconstruct.m:28-194 (GENFIRE) and the Douglas-Rachford variant
// ---------------------------------------------------------------------------
namespace {

// R-factor monitor (reconstruct.m:168-185): model = crop(fftshift(x)) then, for
// every projection, the sums of R1 = sum|back-pj|/sum|pj| and
// R2 = ||back-pj||_F/||pj||_F, kept on the device until recon_end.
void run_rfactor(genfire_job* j, const float* d_model_fft, int slot)
{
    hipStream_t st = j->run_st;
    const int P = j->P;
    EventTimer& tm = *j->rs.rf[slot];
    GF_HIP(hipEventRecord(tm.a, st));
    // the model's forms the projections read: the y-fastest crop (every
    // projection an exact tilt about y), else the z-duplicated forms
    const ProjForms& f = j->rforms;
    if (f.four_ray) {
        crop_centre(d_model_fft, 0, j->n1, j->n2, j->n1, j->vol.as<float>(), j->dim1, j->dim2, j->dim1, st);
        volume_yfast(j->vol.as<float>(), j->vol_y.as<float>(), j->dim1, j->dim2, j->dim1, st);
    }
    if (f.zy) crop_zdup(d_model_fft, j->n1, j->n2, j->n1, j->zy.as<float>(), j->dim1, j->dim2, j->dim1, true, st);
    if (f.zx) crop_zdup(d_model_fft, j->n1, j->n2, j->n1, j->zx.as<float>(), j->dim1, j->dim2, j->dim1, false, st);
    if (f.cubic) crop_cubic_form(d_model_fft, j->n1, j->n2, j->n1, j->vcubic.as<float>(), j->dim1, j->dim2, j->dim1, st);
    const int nblk = project_blocks(j->dim1, j->dim2);
    ProjArgs pa{};
    pa.four_ray = f.four_ray ? 1 : 0;
    pa.vol_y = f.four_ray ? j->vol_y.as<float>() : nullptr;
    pa.zy = f.zy ? j->zy.as<float>() : nullptr;
    pa.zx = f.zx ? j->zx.as<float>() : nullptr;
    pa.vcubic = f.cubic ? j->vcubic.as<float>() : nullptr;
    pa.d1 = j->dim1; pa.d2 = j->dim2; pa.d3 = j->dim1;
    pa.R = j->Rmat.as<float>();
    pa.P = P;
    pa.out = nullptr;
    pa.meas = j->proj.as<float>();
    pa.m1 = j->dim1; pa.m2 = j->dim2;
    pa.sums = j->rsums.as<double>();
    pa.post_hoc = 0;
    pa.acc = j->racc.as<float>();   // set when the model's forms take a slab sweep
    std::pair<hipEvent_t, hipEvent_t> ev;
    kstat_begin(j, GENFIRE_KSTAT_PROJECT, &ev);
    project(pa, st);
    kstat_end(j, GENFIRE_KSTAT_PROJECT, ev);
    project_reduce(j->rsums.as<double>(), P, nblk, j->rout.as<double>() + (size_t)slot * P * 4, st);
    GF_HIP(hipEventRecord(tm.b, st));
}

// The build's seeded permutation for the withheld points (the reference's
// randperm after rng(RandomSeed), reconstruct.m:52-55, is MATLAB's generator):
// candidates are ordered by splitmix64(seed ^ splitmix64(index)), index the
// centred column-major linear index, ties by index.
uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// reconstruct.m:46-69: per resolution shell [edge_s - 0.05, edge_s+1) of
// Q = make_Kspace_indices, withhold floor(count * percent) (at least one) of
// the measured voxels in the permuted order from measuredK_mask, and keep their
// values.  Builds the device mask and the per-half-pencil evaluation lists.
void rfree_select(genfire_job* j)
{
    hipStream_t st = j->run_st;
    const int n1 = j->n1, n2 = j->n2, n3 = j->n1;
    const int64_t nf = full_size(j);
    const int nb = j->prm.numBinsRfree;
    const double pct = j->prm.percentValuesForRfree;
    GF_CHECK(nb >= 1 && nb <= 1024, GENFIRE_ERR_ARG, "GENFIRE: numBinsRfree must be 1..1024");
    GF_CHECK(pct > 0.0 && pct <= 1.0, GENFIRE_ERR_ARG, "GENFIRE: percentValuesForRfree must be in (0, 1]");
    std::vector<float2> M(nf);
    std::vector<uint8_t> mask(nf, 1);
    GF_HIP(hipMemcpyAsync(M.data(), j->M.p, sizeof(float2) * nf, hipMemcpyDeviceToHost, st));
    if (j->have_mask) GF_HIP(hipMemcpyAsync(mask.data(), j->mask.p, nf, hipMemcpyDeviceToHost, st));
    GF_HIP(hipStreamSynchronize(st));
    // shell edges: linspace(0, 1, nb + 1) (MATLAB's, endpoint exact)
    std::vector<double> edge(nb + 1);
    for (int k = 0; k <= nb; ++k) edge[k] = (k == nb) ? 1.0 : 0.0 + ((double)k * (1.0 - 0.0)) / (double)nb;
    const double res_range = -0.05;
    auto coord = [](int c, int n) {   // make_Kspace_indices.m vector entry of centred index c
        if (n == 1) return 0.0;
        if (n % 2 == 0) return (double)(c - n / 2) / (double)(n / 2);
        return (double)(c - (n - 1) / 2) / (double)((n - 1) / 2);
    };
    const uint64_t seed = splitmix64((uint64_t)j->prm.RandomSeed);
    struct Cand {
        uint64_t key;
        int64_t c, f;
    };
    std::vector<std::vector<Cand>> shells(nb);
    for (int64_t f = 0; f < nf; ++f) {
        const float2 m = M[f];
        if (m.x == 0.0f && m.y == 0.0f) continue;   // measuredK ~= 0 (:50)
        const int x = (int)(f % n1), y = (int)((f / n1) % n2), z = (int)(f / ((int64_t)n1 * n2));
        const int c1 = (x + n1 / 2) % n1, c2 = (y + n2 / 2) % n2, c3 = (z + n3 / 2) % n3;   // FFT order -> centred
        const double v1 = coord(c1, n1), v2 = coord(c2, n2), v3 = coord(c3, n3);
        const double Q = std::sqrt((v2 * v2 + v1 * v1) + v3 * v3);
        const int64_t c = c1 + (int64_t)n1 * (c2 + (int64_t)n2 * c3);
        for (int sh = 0; sh < nb; ++sh)
            if (Q >= edge[sh] + res_range && Q < edge[sh + 1])
                shells[sh].push_back({splitmix64(seed ^ splitmix64((uint64_t)c)), c, f});
    }
    struct Ent {
        int64_t pencil;
        uint32_t meta;
        float2 val;
        int shell;
    };
    std::vector<Ent> ents;
    std::vector<double> den(nb, 0.0);
    for (int sh = 0; sh < nb; ++sh) {
        auto& v = shells[sh];
        if (v.empty()) continue;
        size_t cut = (size_t)std::floor((double)v.size() * pct);
        if (cut == 0) cut = 1;
        std::partial_sort(v.begin(), v.begin() + cut, v.end(), [](const Cand& a, const Cand& b) {
            return a.key != b.key ? a.key < b.key : a.c < b.c;
        });
        for (size_t q = 0; q < cut; ++q) {
            const int64_t f = v[q].f;
            mask[f] = 0;   // measuredK_mask(R_freeInd) = false (:66)
            float2 val = M[f];
            den[sh] += std::sqrt((double)val.x * val.x + (double)val.y * val.y);
            int x = (int)(f % n1), y = (int)((f / n1) % n2), z = (int)(f / ((int64_t)n1 * n2));
            if (2 * x > n1) {   // stored as its mate: k(s) = conj(k(-s))
                x = (n1 - x) % n1;
                y = (n2 - y) % n2;
                z = (n3 - z) % n3;
                val.y = -val.y;
            }
            ents.push_back({(int64_t)y + (int64_t)n2 * x, (uint32_t)z | ((uint32_t)y << 11), val, sh});
        }
    }
    // per half pencil lists (stable counting order), and the shell order for the sums
    const int64_t npen = (int64_t)(n1 / 2 + 1) * n2;
    std::vector<int64_t> off(npen + 1, 0);
    for (const Ent& e : ents) off[e.pencil + 1]++;
    for (int64_t q = 0; q < npen; ++q) off[q + 1] += off[q];
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    const size_t ne = std::max<size_t>(ents.size(), 1);
    std::vector<uint32_t> meta(ne, 0);
    std::vector<float2> val(ne, make_float2(0.0f, 0.0f));
    std::vector<int32_t> slot_of(ents.size());
    for (size_t q = 0; q < ents.size(); ++q) {
        const int64_t sl = pos[ents[q].pencil]++;
        meta[sl] = ents[q].meta;
        val[sl] = ents[q].val;
        slot_of[q] = (int32_t)sl;
    }
    std::vector<int32_t> shell_off(nb + 1, 0), order(ne, 0);
    for (const Ent& e : ents) shell_off[e.shell + 1]++;
    for (int sh = 0; sh < nb; ++sh) shell_off[sh + 1] += shell_off[sh];
    {
        std::vector<int32_t> sp(shell_off.begin(), shell_off.end() - 1);
        for (size_t q = 0; q < ents.size(); ++q) order[sp[ents[q].shell]++] = slot_of[q];
    }
    j->rf_bins = nb;
    j->rf_off.alloc(sizeof(int64_t) * (npen + 1));
    j->rf_meta.alloc(sizeof(uint32_t) * ne);
    j->rf_val.alloc(sizeof(float2) * ne);
    j->rf_terms.alloc(sizeof(float) * ne);
    j->rf_order.alloc(sizeof(int32_t) * ne);
    j->rf_shell_off.alloc(sizeof(int32_t) * (nb + 1));
    j->rf_den.alloc(sizeof(double) * nb);
    j->Rfree.alloc(sizeof(float) * (size_t)nb * std::max(j->prm.numIterations, 1));
    if (!j->mask.p) j->mask.alloc(nf);
    GF_HIP(hipMemcpyAsync(j->mask.p, mask.data(), nf, hipMemcpyHostToDevice, st));
    GF_HIP(hipMemcpyAsync(j->rf_off.p, off.data(), sizeof(int64_t) * (npen + 1), hipMemcpyHostToDevice, st));
    GF_HIP(hipMemcpyAsync(j->rf_meta.p, meta.data(), sizeof(uint32_t) * ne, hipMemcpyHostToDevice, st));
    GF_HIP(hipMemcpyAsync(j->rf_val.p, val.data(), sizeof(float2) * ne, hipMemcpyHostToDevice, st));
    GF_HIP(hipMemcpyAsync(j->rf_order.p, order.data(), sizeof(int32_t) * ne, hipMemcpyHostToDevice, st));
    GF_HIP(hipMemcpyAsync(j->rf_shell_off.p, shell_off.data(), sizeof(int32_t) * (nb + 1), hipMemcpyHostToDevice,
                          st));
    GF_HIP(hipMemcpyAsync(j->rf_den.p, den.data(), sizeof(double) * nb, hipMemcpyHostToDevice, st));
    GF_HIP(hipStreamSynchronize(st));
    j->have_mask = true;
}

// save_temp snapshots (reconstruct.m:149-159): the candidate P_R(x) of
// iteration it is written by the pass that forms it into cand[it % 2] (FFT
// order, box rows); once errK(it) is reduced it replaces the best box when
// errK(it) < bestErr (reset to +inf at every enforcement stage, :123).
bool snap_on(const genfire_job* j) { return j->snap_L > 0 && j->snap_fn != nullptr; }
float* snap_cand(const genfire_job* j, int it) { return const_cast<float*>(j->cand[it & 1].as<float>()); }

__global__ void k_best_reset(double* best_err) { *best_err = INFINITY; }

__global__ void k_best_decide(const float* errK, int slot, double* best_err, int* flag)
{
    const double e = (double)errK[slot];
    const int better = e < *best_err;
    if (better) *best_err = e;
    *flag = better;
}

// best_box = crop of cand when *flag (the centred Dim box of the FFT-order grid)
__global__ void k_best_crop(const int* flag, const float* __restrict__ cand, int n1, int n2, int n3,
                            float* __restrict__ box, int d1, int d2, int d3)
{
    if (!*flag) return;
    const int64_t total = (int64_t)d1 * d2 * d3;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(t % d1), b = (int)((t / d1) % d2), c = (int)(t / ((int64_t)d1 * d2));
        int s1 = a - d1 / 2, s2 = b - d2 / 2, s3 = c - d3 / 2;
        s1 += s1 < 0 ? n1 : 0;
        s2 += s2 < 0 ? n2 : 0;
        s3 += s3 < 0 ? n3 : 0;
        box[t] = cand[(int64_t)s1 + (int64_t)n1 * (s2 + (int64_t)n2 * s3)];
    }
}

void snap_update(genfire_job* j, int it)
{
    hipStream_t st = j->run_st;
    k_best_decide<<<1, 1, 0, st>>>(j->errK.as<float>(), it - 1, j->best_err.as<double>(), j->best_flag.as<int>());
    GF_LAUNCH_CHECK();
    const int64_t total = box_size(j);
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 8192);
    k_best_crop<<<blocks, 256, 0, st>>>(j->best_flag.as<int>(), snap_cand(j, it), j->n1, j->n2, j->n1,
                                        j->best_box.as<float>(), j->dim1, j->dim2, j->dim1);
    GF_LAUNCH_CHECK();
    if (it % j->snap_L == 0) {
        j->snap_h.resize((size_t)total);
        GF_HIP(hipMemcpyAsync(j->snap_h.data(), j->best_box.p, sizeof(float) * total, hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        j->snap_fn(j->snap_user, it, j->snap_h.data());
    }
}

HalfIterArgs p5_args(const genfire_job* j, int it)
{
    const ReconState& r = j->rs;
    HalfIterArgs ha = r.ha;
    const bool mon_it = r.monitor && (it % r.L == 0);
    ha.pf_out = mon_it ? const_cast<float*>(j->pfbuf.as<float>()) : nullptr;
    ha.best = snap_on(j) ? snap_cand(j, it + 1)
                         : (it + 1 == r.N) ? const_cast<float*>(j->best.as<float>()) : nullptr;
    // the iteration's errK from the y-inverse launch's chunk sums
    ha.errk = const_cast<float*>(j->errK.as<float>());
    ha.errctr = const_cast<int*>(j->errctr.as<int>());
    ha.err_chunk = const_cast<double*>(j->err_chunk.as<double>());
    ha.err_den = r.den;
    ha.err_stage = 2;
    return ha;
}

}  // namespace

void recon_begin(genfire_job* j)
{
    GF_CHECK(j->gridded, GENFIRE_ERR_STATE, "GENFIRE: measuredK not available (call genfire_job_grid first)");
    ReconState& r = j->rs;
    r.N = j->prm.numIterations;
    const int N = r.N;
    const int n1 = j->n1, n2 = j->n2, n3 = j->n1;
    const int64_t nf = full_size(j);
    if (!j->run_st) j->run_st = j->st;
    hipStream_t st = j->run_st;
    r.dr = j->prm.algorithm == GENFIRE_ALG_DR;
    r.monitor = j->prm.monitor_R != 0;
    r.L = std::max(1, j->prm.monitorR_loopLength);
    const bool dr = r.dr, monitor = r.monitor;

    // GENFIRE_HALF=0 runs the full-spectrum (complex) passes on a grid the
    // half-spectrum passes also handle (the two forms compared in the tests).
    const char* hf = std::getenv("GENFIRE_HALF");
    r.half = half_supported(n1, n2, n3) && !(hf && std::string(hf) == "0");
    const bool half = r.half;
    r.hn = (int64_t)(n1 / 2 + 1) * n2 * n3;   // half-spectrum size
    const int64_t hn = r.hn;
    r.all.init();
    r.loop.init();
    GF_HIP(hipEventRecord(r.all.a, st));
    j->x.alloc(sizeof(float) * nf);
    j->K.ensure(sizeof(float2) * (half ? std::max<int64_t>(2 * hn, nf) : nf));
    j->S.alloc(nf);
    j->best.alloc(sizeof(float) * nf);
    if (monitor && (dr || half)) j->pfbuf.alloc(sizeof(float) * nf);
    IterArgs& ia = r.ia;
    ia = IterArgs{};
    ia.n1 = n1; ia.n2 = n2; ia.n3 = n3;
    GF_CHECK(n1 <= 2048 && n3 <= 2048, GENFIRE_ERR_UNSUPPORTED, "GENFIRE: grid edge above 2048");
    r.nparts = half ? half_partials(n1, n2, n3) : iter_fused_partials(ia);
    j->err_part.alloc(sizeof(double) * r.nparts);
    j->err_chunk.alloc(sizeof(double) * kErrChunks);
    j->errK.alloc(sizeof(float) * std::max(N, 1));
    j->errctr.alloc(sizeof(int));
    GF_HIP(hipMemsetAsync(j->errctr.p, 0, sizeof(int), st));
    const int ntmp = 1024;
    j->den_tmp.alloc(sizeof(double) * ntmp);
    j->den.alloc(sizeof(double));

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

    r.sched = cutoff_schedule(N, (int)j->delays.size());
    r.cur = 0;
    GF_CHECK(N >= 1 && r.sched.back() == N, GENFIRE_ERR_ARG, "GENFIRE: numIterations must be >= 1");

    j->rf_bins = 0;
    if (j->prm.calculate_Rfree) {
        GF_CHECK(half, GENFIRE_ERR_UNSUPPORTED, "GENFIRE: Rfree needs a grid on the half-spectrum path");
        rfree_select(j);
    }

    // measured voxels (errInd, reconstruct.m:99) as per-pencil lists
    const uint8_t* mask = j->have_mask ? j->mask.as<uint8_t>() : nullptr;
    const int64_t npen = half ? (int64_t)(n1 / 2 + 1) * n2 : (int64_t)n1 * n2;
    const int64_t nseg = half ? npen * hcsr_chunks() : npen;   // counted segments
    j->csr_cnt.alloc(sizeof(uint32_t) * nseg);
    j->csr_off.alloc(sizeof(int64_t) * (npen + 1));
    j->scan_tmp.ensure(grid_index_scan_scratch(nseg));
    if (half) {
        j->csr_coff.alloc(sizeof(int64_t) * nseg);
        hcsr_count(j->M.as<float2>(), mask, n1, n2, n3, j->csr_cnt.as<uint32_t>(), st);
        grid_index_scan(j->csr_cnt.as<uint32_t>(), nseg, j->csr_coff.as<int64_t>(), j->csr_off.as<int64_t>() + npen,
                        j->scan_tmp.as<int64_t>(), st);
        hcsr_pencil_off(j->csr_coff.as<int64_t>(), npen, j->csr_off.as<int64_t>(), st);
    } else {
        csr_count(j->M.as<float2>(), mask, n1, n2, n3, j->csr_cnt.as<uint32_t>(), st);
        grid_index_scan(j->csr_cnt.as<uint32_t>(), npen, j->csr_off.as<int64_t>(), j->csr_off.as<int64_t>() + npen,
                        j->scan_tmp.as<int64_t>(), st);
    }
    GF_HIP(hipMemcpyAsync(&j->nmeas, j->csr_off.as<int64_t>() + npen, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    GF_HIP(hipStreamSynchronize(st));
    const int64_t ne = std::max<int64_t>(j->nmeas, 1);
    j->csr_meta.alloc(sizeof(uint32_t) * ne);
    j->csr_val.alloc(sizeof(float2) * ne);
    if (half) {
        j->csr_mag.alloc(sizeof(float2) * ne);
        hcsr_fill(j->M.as<float2>(), mask, n1, n2, n3, j->csr_coff.as<int64_t>(), j->csr_meta.as<uint32_t>(),
                  j->csr_val.as<float2>(), j->csr_mag.as<float2>(), st);
        j->csr_coff.release();
        hcsr_den(j->csr_mag.as<float2>(), j->csr_meta.as<uint32_t>(), j->nmeas, j->den_tmp.as<double>(), ntmp,
                 j->den.as<double>(), st);
    } else {
        csr_fill(j->M.as<float2>(), mask, n1, n2, n3, j->csr_off.as<int64_t>(), j->csr_meta.as<uint32_t>(),
                 j->csr_val.as<float2>(), st);
        csr_den(j->csr_val.as<float2>(), j->nmeas, j->den_tmp.as<double>(), ntmp, j->den.as<double>(), st);
    }
    // errK denominator: sum |M(errInd)| (reconstruct.m:135)
    GF_HIP(hipMemcpyAsync(&r.den, j->den.p, sizeof(double), hipMemcpyDeviceToHost, st));
    GF_HIP(hipStreamSynchronize(st));

    // R-factor monitor buffers and projection matrices
    r.nR = monitor ? N / r.L : 0;
    j->errR_h.assign(r.nR, 0.0f);
    j->Rarr_h.assign((size_t)r.nR * j->P, 0.0);
    j->Rarr2_h.assign((size_t)r.nR * j->P, 0.0);
    j->t_rf = 0.0;
    while ((int)r.rf.size() < r.nR) {
        r.rf.emplace_back(new EventTimer());
        r.rf.back()->init();
    }
    if (r.nR > 0) {
        bool any_y = false;
        std::vector<float> Rm = proj_matrices(j->angles.data(), j->P, &any_y);
        j->rforms = project_forms(Rm.data(), j->P, j->dim2, j->prm.projectorInterpolation == GENFIRE_INTERP_CUBIC);
        j->Rmat.alloc(sizeof(float) * Rm.size());
        GF_HIP(hipMemcpyAsync(j->Rmat.p, Rm.data(), sizeof(float) * Rm.size(), hipMemcpyHostToDevice, st));
        const ProjForms& f = j->rforms;
        if (f.four_ray) {
            j->vol.alloc(sizeof(float) * box_size(j));
            j->vol_y.alloc(sizeof(float) * box_size(j));
        }
        if (f.zy) j->zy.alloc(sizeof(float) * zdup_floats(j->dim1, j->dim2, j->dim1));
        if (f.zx) j->zx.alloc(sizeof(float) * zdup_floats(j->dim1, j->dim2, j->dim1));
        if (f.cubic) j->vcubic.alloc(sizeof(float) * cubic_form_floats(j->dim1, j->dim2, j->dim1));
        j->rsums.alloc(sizeof(double) * (size_t)j->P * 4 * project_blocks(j->dim1, j->dim2));
        if ((f.zy || f.zx) && project_slabs(j->dim1, j->dim2, j->dim1) > 1)
            j->racc.alloc(sizeof(float) * (size_t)j->P * j->dim1 * j->dim2);
        j->rout.alloc(sizeof(double) * (size_t)r.nR * j->P * 4);
        GF_HIP(hipStreamSynchronize(st));
    }

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
    HalfIterArgs& ha = r.ha;
    ha = HalfIterArgs{};
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
    // Box pruning needs P_R(x) to vanish outside the Dim box (support on);
    // GENFIRE_PRUNE=0 turns it off (every row transformed, as for support off).
    const char* pe = std::getenv("GENFIRE_PRUNE");
    const bool box = half && ia.support && !(pe && std::string(pe) == "0");
    ha.bx = box ? j->dim1 : n1;
    ha.by = box ? j->dim2 : n2;
    ha.bz = box ? j->dim1 : n3;
    ha.dmode = (box && dr) ? 1 : 0;
    if (ha.dmode) {
        // replacement deltas start at zero (D'_0 = 0: the first transform is of
        // 2 P_R(x_1) - x_1 with x_1 in the box), and the in-plane mates of the
        // kx = 0 and n1/2 plane entries
        j->csr_d.alloc(sizeof(float2) * std::max<int64_t>(j->nmeas, 1));
        GF_HIP(hipMemsetAsync(j->csr_d.p, 0, j->csr_d.bytes, st));
        int64_t o[3];
        const int64_t* off = j->csr_off.as<int64_t>();
        GF_HIP(hipMemcpyAsync(&o[0], off + n2, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        GF_HIP(hipMemcpyAsync(&o[1], off + (int64_t)(n1 / 2) * n2, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        GF_HIP(hipMemcpyAsync(&o[2], off + (int64_t)(n1 / 2 + 1) * n2, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        const int64_t npl = o[0] + (o[2] - o[1]);
        j->csr_mate.alloc(sizeof(int64_t) * std::max<int64_t>(npl, 1));
        hcsr_mates(off, j->csr_meta.as<uint32_t>(), n1, n2, n3, o[0], o[1], npl, j->csr_mate.as<int64_t>(), st);
        ha.csr_d = j->csr_d.as<float2>();
        ha.csr_mate = j->csr_mate.as<int64_t>();
    } else {
        j->csr_d.release();
        j->csr_mate.release();
    }
    if (j->rf_bins > 0) {
        ha.rf_off = j->rf_off.as<int64_t>();
        ha.rf_meta = j->rf_meta.as<uint32_t>();
        ha.rf_val = j->rf_val.as<float2>();
        ha.rf_terms = j->rf_terms.as<float>();
    }
    GF_HIP(hipEventRecord(r.loop.a, st));
    r.begun = true;
}

void recon_pre(genfire_job* j, int it)
{
    ReconState& r = j->rs;
    hipStream_t st = j->run_st;
    const int n1 = j->n1, n2 = j->n2, n3 = j->n1;
    if (r.cur < r.sched.size() && it == r.sched[r.cur]) {          // :118-124
        size_t last = r.cur;
        for (size_t q = r.cur; q < r.sched.size(); ++q)
            if (r.sched[q] == it) last = q;
        if (r.half) hcsr_enforce(n1, n2, n3, j->csr_off.as<int64_t>(), j->csr_meta.as<uint32_t>(), j->delays[last], st);
        else csr_enforce(n1, n2, n3, j->csr_off.as<int64_t>(), j->csr_meta.as<uint32_t>(), j->delays[last], st);
        r.cur = last + 1;
    }
    // The schedule always ends with a change at it == N, which resets bestErr
    // (:123), so the returned reconstruction is P_R(x) of the last iteration:
    // it is written there and nowhere else.  P1 runs on its own only in the
    // first iteration; later ones get it from the fused P5 + P1 pass that ends
    // the iteration before.
    if (r.half && it == 1) {
        HalfIterArgs ha = r.ha;
        ha.best = snap_on(j) ? snap_cand(j, 1) : (r.N == 1) ? j->best.as<float>() : nullptr;
        ha.pf_out = nullptr;
        half_first(ha, st);
    }
}

// Rfree of this iteration from the z pass's terms (slot = the device-side
// iteration index, advanced later by the errK reduction)
void rfree_step(genfire_job* j, hipStream_t st)
{
    if (j->rf_bins <= 0) return;
    rfree_reduce(j->rf_terms.as<float>(), j->rf_order.as<int32_t>(), j->rf_shell_off.as<int32_t>(),
                 j->rf_den.as<double>(), j->rf_bins, j->Rfree.as<float>(), j->errctr.as<int>(), st);
}

// The args of an iteration's y-inverse launch: it also reduces errK and forms
// D's Hermitian part (DR on the box).
HalfIterArgs y_inv_args(const genfire_job* j)
{
    const ReconState& r = j->rs;
    HalfIterArgs a = r.ha;
    a.err_chunk = const_cast<double*>(j->err_chunk.as<double>());
    a.err_n = r.nparts;
    a.err_stage = 1;
    a.dherm = a.dmode;
    return a;
}

void recon_body(genfire_job* j, int it)
{
    ReconState& r = j->rs;
    hipStream_t st = j->run_st;
    const int n1 = j->n1, n2 = j->n2, n3 = j->n1;
    std::pair<hipEvent_t, hipEvent_t> e_pass, e_fused;
    if (r.half) {
        kstat_begin(j, GENFIRE_KSTAT_FFT_PASS, &e_pass);
        half_y(r.ha, 0, st);
        kstat_end(j, GENFIRE_KSTAT_FFT_PASS, e_pass);
        kstat_begin(j, GENFIRE_KSTAT_FFT_FUSED, &e_fused);
        half_z(r.ha, st);
        kstat_end(j, GENFIRE_KSTAT_FFT_FUSED, e_fused);
        rfree_step(j, st);
        half_y(y_inv_args(j), 1, st);
        if (it < r.N) half_last_first(p5_args(j, it), st);
        else half_last(p5_args(j, it), st);
    } else {
        IterArgs ia = r.ia;
        const bool mon_it = r.monitor && (it % r.L == 0);
        ia.best = snap_on(j) ? snap_cand(j, it) : (it == r.N) ? j->best.as<float>() : nullptr;
        ia.pf_out = (mon_it && r.dr) ? j->pfbuf.as<float>() : nullptr;
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
}

void half_body_batch(genfire_job* const* jobs, const int* its, int count, hipStream_t st)
{
    for (int c0 = 0; c0 < count; c0 += kHalfBatchMax) {
        const int c1 = std::min(count, c0 + kHalfBatchMax);
        HalfBatch all{}, inv{}, cont{}, fin{};
        for (int i = c0; i < c1; ++i) {
            const genfire_job* j = jobs[i];
            GF_CHECK(j->rs.half, GENFIRE_ERR_ARG, "half_body_batch: job is not on the half-spectrum path");
            all.job[all.count++] = j->rs.ha;
            inv.job[inv.count++] = y_inv_args(j);
            if (its[i] < j->rs.N) cont.job[cont.count++] = p5_args(j, its[i]);
            else fin.job[fin.count++] = p5_args(j, its[i]);
        }
        half_y(all, 0, st);
        half_z(all, st);
        for (int i = c0; i < c1; ++i) rfree_step(jobs[i], st);
        half_y(inv, 1, st);
        if (cont.count) half_last_first(cont, st);
        if (fin.count) half_last(fin, st);
    }
}

void recon_post(genfire_job* j, int it)
{
    ReconState& r = j->rs;
    if (!r.half)   // the half-spectrum y-inverse pass reduced errK already
        iter_errk_reduce_next(j->err_part.as<double>(), r.nparts, r.den, j->errK.as<float>(), j->errctr.as<int>(),
                              j->run_st);
    if (r.monitor && (it % r.L == 0) && it / r.L - 1 < r.nR) {
        const float* model = (r.dr || r.half) ? j->pfbuf.as<float>() : j->x.as<float>();
        run_rfactor(j, model, it / r.L - 1);
    }
}

void recon_end(genfire_job* j)
{
    ReconState& r = j->rs;
    hipStream_t st = j->run_st;
    const int N = r.N;
    GF_HIP(hipEventRecord(r.loop.b, st));
    // reconstruction = My_stripzero(fftshift(best), [Dim1 Dim2 Dim1]) (:192)
    j->recon.alloc(sizeof(float) * box_size(j));
    if (snap_on(j))   // the best box kept for the snapshots: P_R(x) of the last iteration (a stage at N)
        GF_HIP(hipMemcpyAsync(j->recon.p, j->best_box.p, sizeof(float) * box_size(j), hipMemcpyDeviceToDevice, st));
    else
        crop_centre(j->best.p, 0, j->n1, j->n2, j->n1, j->recon.as<float>(), j->dim1, j->dim2, j->dim1, st);
    j->errK_h.assign(N, 0.0f);
    GF_HIP(hipMemcpyAsync(j->errK_h.data(), j->errK.p, sizeof(float) * N, hipMemcpyDeviceToHost, st));
    j->Rfree_h.assign((size_t)j->rf_bins * N, -1.0f);
    if (j->rf_bins > 0)
        GF_HIP(hipMemcpyAsync(j->Rfree_h.data(), j->Rfree.p, sizeof(float) * j->Rfree_h.size(),
                              hipMemcpyDeviceToHost, st));
    std::vector<double> h((size_t)r.nR * j->P * 4);
    if (r.nR > 0) GF_HIP(hipMemcpyAsync(h.data(), j->rout.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    GF_HIP(hipEventRecord(r.all.b, st));
    GF_HIP(hipEventSynchronize(r.all.b));
    const int P = j->P;
    j->t_rf = 0.0;
    for (int slot = 0; slot < r.nR; ++slot) {
        double mean = 0.0;
        const double* s = h.data() + (size_t)slot * P * 4;
        for (int p = 0; p < P; ++p) {
            const double r1 = s[4 * p + 0] / s[4 * p + 2];
            const double r2 = std::sqrt(s[4 * p + 1]) / std::sqrt(s[4 * p + 3]);
            j->Rarr_h[(size_t)slot * P + p] = r1;
            j->Rarr2_h[(size_t)slot * P + p] = r2;
            mean += r1;
        }
        j->errR_h[slot] = (float)(mean / P);
        float ms = 0.0f;
        GF_HIP(hipEventElapsedTime(&ms, r.rf[slot]->a, r.rf[slot]->b));
        j->t_rf += ms;
    }
    float ms = 0.0f;
    GF_HIP(hipEventElapsedTime(&ms, r.loop.a, r.loop.b));
    j->t_iter = ms - j->t_rf;
    GF_HIP(hipEventElapsedTime(&ms, r.all.a, r.all.b));
    j->t_total = j->t_grid + ms;
    if (j->prof) j->ks.collect();
    j->reconstructed = true;
    r.begun = false;
}

bool steady_iteration(const genfire_job* j, int it)
{
    const ReconState& r = j->rs;
    const bool stage = r.cur < r.sched.size() && it == r.sched[r.cur];
    const bool mon_it = r.monitor && (it % r.L == 0);
    return !snap_on(j) && it > 1 && it + 1 < r.N && !stage && !mon_it;
}

// Steady-state iterations replay one captured HIP graph (the transform passes
// and the errK reduction, whose output slot is the device-side iteration
// index) unless per-kernel profiling is on or GENFIRE_GRAPHS=0.
void do_reconstruct(genfire_job* j)
{
    j->run_st = j->st;
    recon_begin(j);
    ReconState& r = j->rs;
    r.drop_graph();
    r.graph_launches = 0;
    const char* ge = std::getenv("GENFIRE_GRAPHS");
    const bool snap = snap_on(j);
    const bool graphs = !j->prof && !snap && !(ge && std::string(ge) == "0");
    hipStream_t st = j->run_st;
    if (snap) {
        j->cand[0].alloc(sizeof(float) * full_size(j));
        j->cand[1].alloc(sizeof(float) * full_size(j));
        j->best_box.alloc(sizeof(float) * box_size(j));
        j->best_err.alloc(sizeof(double));
        j->best_flag.alloc(sizeof(int));
    }
    std::pair<hipEvent_t, hipEvent_t> e_it;
    for (int it = 1; it <= r.N; ++it) {
        const bool stage = r.cur < r.sched.size() && it == r.sched[r.cur];
        const bool steady = graphs && steady_iteration(j, it);
        kstat_begin(j, GENFIRE_KSTAT_ITERATION, &e_it);
        if (steady) {
            if (!r.gexec) {
                GF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                recon_body(j, it);
                if (!r.half)
                    iter_errk_reduce_next(j->err_part.as<double>(), r.nparts, r.den, j->errK.as<float>(),
                                          j->errctr.as<int>(), st);
                GF_HIP(hipStreamEndCapture(st, &r.graph));
                GF_HIP(hipGraphInstantiate(&r.gexec, r.graph, nullptr, nullptr, 0));
            }
            GF_HIP(hipGraphLaunch(r.gexec, st));
            ++r.graph_launches;
        } else {
            if (snap && stage) {
                k_best_reset<<<1, 1, 0, st>>>(j->best_err.as<double>());
                GF_LAUNCH_CHECK();
            }
            recon_pre(j, it);
            recon_body(j, it);
            if (!r.half)
                iter_errk_reduce_next(j->err_part.as<double>(), r.nparts, r.den, j->errK.as<float>(),
                                      j->errctr.as<int>(), st);
            if (snap) snap_update(j, it);
        }
        kstat_end(j, GENFIRE_KSTAT_ITERATION, e_it);
        if (j->rs.monitor && (it % j->rs.L == 0) && it / j->rs.L - 1 < j->rs.nR) {
            const float* model = (j->rs.dr || j->rs.half) ? j->pfbuf.as<float>() : j->x.as<float>();
            run_rfactor(j, model, it / j->rs.L - 1);
        }
    }
    recon_end(j);
    r.drop_graph();
}

void release_grid_extras(genfire_job* j)
{
    j->ind.release();
    j->dist.release();
    j->un.release();
    j->vn.release();
    j->F.release();
    j->tiles.release();
    j->hits.release();
    j->blk_counts.release();
    j->blk_offsets.release();
    j->recifft.release();
    j->M.release();
    j->mask.release();
    j->csr_cnt.release();
    j->den_tmp.release();
    release_grid_scratch(j);
}

void release_grid_scratch(genfire_job* j)
{
    j->acc.release();
    j->sw.release();
    j->exact.release();
    j->Pt.release();
    j->nu_inv.release();
    j->Phi.release();
    j->Plo.release();
    j->Pscale.release();
    j->tile_meta.release();
}

// ---------------------------------------------------------------------------
// Job set-up (CheckPrepareData, GENFIRE_Reconstructor.m:115-171) and inputs
// ---------------------------------------------------------------------------
void grid_edges(const genfire_params& p, int dim1, int dim2, int* n1, int* n2)
{
    if (p.griddingMethod == 1) {
        // fillInFourierGrid.m:38-42: kMeasured is d x d with d from Dim2
        // (particleWindowSize(2), GENFIRE_Reconstructor.m:121)
        const double e = dim2 % 2 == 0 ? dim2 * p.oversamplingRatio : 1.0 + (dim2 - 1) * p.oversamplingRatio;
        GF_CHECK(e == std::floor(e) && e >= 1.0 && e < 65536.0, GENFIRE_ERR_ARG,
                 "GENFIRE: FFT gridding needs a whole-number grid edge (Dim2*oversamplingRatio for even Dim2, "
                 "1+(Dim2-1)*oversamplingRatio for odd), got " + std::to_string(e));
        *n1 = *n2 = (int)e;
        return;
    }
    *n1 = (int)mround(dim1 * p.oversamplingRatio);   // GENFIRE_Reconstructor.m:168
    *n2 = (int)mround(dim2 * p.oversamplingRatio);   // :169
}

void validate_job(const genfire_params* p, int dim1, int dim2, int P)
{
    GF_CHECK(dim1 >= 2 && dim2 >= 2 && P >= 1, GENFIRE_ERR_ARG,
             "GENFIRE: projections must be at least 2 x 2 and P >= 1");
    GF_CHECK(p->numIterations >= 1, GENFIRE_ERR_ARG, "GENFIRE: numIterations must be >= 1");
    GF_CHECK(p->oversamplingRatio > 0, GENFIRE_ERR_ARG, "GENFIRE: oversamplingRatio must be positive");
    GF_CHECK(p->algorithm == GENFIRE_ALG_GENFIRE || p->algorithm == GENFIRE_ALG_DR, GENFIRE_ERR_ARG,
             "GENFIRE: unknown algorithm");
    GF_CHECK(p->precision == GENFIRE_PREC_FP32 || p->precision == GENFIRE_PREC_SPLIT_BF16 ||
                 p->precision == GENFIRE_PREC_FP16,
             GENFIRE_ERR_ARG, "GENFIRE: unknown gridding precision");
    GF_CHECK(p->griddingMethod == 1 || p->griddingMethod == 2, GENFIRE_ERR_ARG,
             "GENFIRE: Unrecognized gridding method.");   // GENFIRE_Reconstructor.m:151-153
    GF_CHECK(p->projectorInterpolation == GENFIRE_INTERP_LINEAR || p->projectorInterpolation == GENFIRE_INTERP_CUBIC,
             GENFIRE_ERR_ARG, "GENFIRE: unknown projector interpolation");
    GF_CHECK(p->projectorInterpolation != GENFIRE_INTERP_CUBIC || !p->monitor_R || (dim1 >= 4 && dim2 >= 4),
             GENFIRE_ERR_ARG, "GENFIRE: interp3 'cubic' needs at least 4 nodes along every axis");
    int n1 = 0, n2 = 0;
    grid_edges(*p, dim1, dim2, &n1, &n2);
    if (p->griddingMethod == 1)
        GF_CHECK(n1 >= dim1 && n1 >= dim2, GENFIRE_ERR_ARG,   // My_paddzero.m:27-28
                 "GENFIRE: paddedsize should be equal to or larger than original volume in all dimensions!");
    GF_CHECK(n1 >= dim1 && n2 >= dim2, GENFIRE_ERR_ARG, "GENFIRE: oversampled size smaller than projection");
    GF_CHECK(fft_supported(n1) && fft_supported(n2), GENFIRE_ERR_UNSUPPORTED,
             "GENFIRE: oversampled grid " + std::to_string(n1) + "x" + std::to_string(n2) +
                 " has an FFT length the library does not support");
    (void)delay_indicators(p->constraintEnforcementMode);   // raises the reference's message for a bad mode
}

void job_setup(genfire_job* j, genfire_ctx* ctx, const genfire_params* p, int dim1, int dim2, int P)
{
    validate_job(p, dim1, dim2, P);
    j->ctx = ctx;
    j->prm = *p;
    j->dim1 = dim1;
    j->dim2 = dim2;
    j->P = P;
    grid_edges(*p, dim1, dim2, &j->n1, &j->n2);
    j->hg = half_grid(j->n1, j->n2, p->DFT_CentroSymmetricity ? 1 : 0);
    j->delays = delay_indicators(p->constraintEnforcementMode);
    GF_HIP(hipStreamCreateWithFlags(&j->st, hipStreamNonBlocking));
    j->run_st = j->st;
}

void job_set_inputs(genfire_job* j, const float* projections, const double* angles, const uint8_t* support,
                    const float* initial_object, const uint16_t* proj_f16, const float* proj_scale)
{
    GF_CHECK((projections || (proj_f16 && proj_scale)) && angles, GENFIRE_ERR_ARG,
             "GENFIRE: projections and angles are required");
    const size_t np = (size_t)j->dim1 * j->dim2 * j->P;
    j->proj.alloc(sizeof(float) * np);
    if (projections) {
        GF_HIP(hipMemcpyAsync(j->proj.p, projections, sizeof(float) * np, hipMemcpyHostToDevice, j->st));
    } else {
        DevBuf h, sc;
        h.alloc(sizeof(uint16_t) * np);
        sc.alloc(sizeof(float) * j->P);
        GF_HIP(hipMemcpyAsync(h.p, proj_f16, sizeof(uint16_t) * np, hipMemcpyHostToDevice, j->st));
        GF_HIP(hipMemcpyAsync(sc.p, proj_scale, sizeof(float) * j->P, hipMemcpyHostToDevice, j->st));
        expand_f16(h.as<uint16_t>(), sc.as<float>(), (int64_t)j->dim1 * j->dim2, j->P, j->proj.as<float>(), j->st);
        GF_HIP(hipStreamSynchronize(j->st));
    }
    j->angles.assign(angles, angles + 3 * (size_t)j->P);
    j->have_support = support != nullptr;
    if (support) {
        j->support_box.alloc(box_size(j));
        GF_HIP(hipMemcpyAsync(j->support_box.p, support, box_size(j), hipMemcpyHostToDevice, j->st));
    }
    j->have_init = initial_object != nullptr;
    if (initial_object && j->prm.griddingMethod == 1) {
        // reconstruct.m:79 pads the initial object to n1_oversampled x
        // n2_oversampled x n1_oversampled, which must be the grid's size
        const int o1 = (int)mround(j->dim1 * j->prm.oversamplingRatio);
        const int o2 = (int)mround(j->dim2 * j->prm.oversamplingRatio);
        GF_CHECK(o1 == j->n1 && o2 == j->n2, GENFIRE_ERR_ARG,
                 "GENFIRE: initial object padded to " + std::to_string(o1) + "x" + std::to_string(o2) + "x" +
                     std::to_string(o1) + " does not match the FFT-gridding grid of edge " + std::to_string(j->n1));
    }
    if (initial_object) {
        j->init_box.alloc(sizeof(float) * box_size(j));
        GF_HIP(hipMemcpyAsync(j->init_box.p, initial_object, sizeof(float) * box_size(j), hipMemcpyHostToDevice,
                              j->st));
    }
    GF_HIP(hipStreamSynchronize(j->st));
    j->Rmat.release();
    j->have_inputs = true;
    j->gridded = false;
    j->reconstructed = false;
}

// Upper estimate of the device bytes a job holds at its peak: the larger of the
// gridding phase (point lists, FP64 half-grid accumulators, measuredK and the
// recIFFT transform) and the iteration phase (iterate, half or full spectrum,
// support, best, monitor model, measured-voxel lists), plus the inputs.
int64_t job_footprint(const genfire_params& p, int dim1, int dim2, int P)
{
    int e1 = 0, e2 = 0;
    grid_edges(p, dim1, dim2, &e1, &e2);
    const int64_t n1 = e1, n2 = e2;
    const int64_t n3 = n1, nf = n1 * n2 * n3, box = (int64_t)dim1 * dim2 * dim1;
    const int64_t half_grid_vox = (n1 / 2 + 1) * (n2 + 1) * (n3 + 1);
    const double cut = std::max(0.05, p.interpolationCutoffDistance);
    const int64_t pts = (int64_t)((1.3 * cut + 0.05) * (double)(n1 * n2) * P) + 4096;
    const int64_t kp = (dim1 + 63) / 64 * 64, np = (dim2 + 15) / 16 * 16;
    const int64_t inputs = (int64_t)4 * dim1 * dim2 * P + box * 5;
    const int64_t grid = p.griddingMethod == 1
                             ? n1 * n1 * P * 8 + nf * (16 + 8 + 4 + 8 + 8) + box * 4   // spectra, d^3 sums, M, K
                             : pts * 28 +
                                   std::max<int64_t>(kp * np * 4, (int64_t)nufft_fine_len(dim1) * nufft_fine_len(dim2) * 8) * P +
                                   half_grid_vox * 36 + nf * (8 + 1 + 8) + box * 4;
    const bool half = half_supported((int)n1, (int)n2, (int)n3);
    const int64_t hn = (n1 / 2 + 1) * n2 * n3;
    const int64_t spectrum = 8 * (half ? std::max<int64_t>(2 * hn, nf) : nf);
    const int64_t meas = std::min<int64_t>(nf, 2 * pts);
    const int64_t iter = nf * (4 + 1 + 4 + 4 + 8) + spectrum + meas * 28 + n1 * n2 * 12 + box * 8 +
                         8 * zdup_floats(dim1, dim2, dim1) +    // the projector's two z-duplicated forms
                         (project_slabs(dim1, dim2, dim1) > 1 ? (int64_t)4 * P * dim1 * dim2 : 0) +   // slab sums
                         (int64_t)P * 4 * 8 * (1 + p.numIterations / std::max(1, p.monitorR_loopLength));
    return inputs + std::max(grid, iter) + (64ll << 20);
}

}  // namespace gf
