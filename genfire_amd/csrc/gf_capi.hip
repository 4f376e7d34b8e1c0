// gf_capi.hip -- the extern "C" boundary of libgenfire (include/genfire.h).
// The job phases live in gf_job.hip, the request server in gf_serve.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

#include "gf_job.h"

using namespace gf;

// ===========================================================================
// C ABI
// ===========================================================================
// Device scratch of genfire_rfactor_noalloc (the projector's R matrices, tile
// partial sums and model forms): one array per GPU, held until the process
// ends.  Host code never passes `scratch` itself to a kernel or a copy: it
// uses the device address hipGetSymbolAddress returns for the current GPU,
// taken once per device when a context is created.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kNoAllocMaxProjs = 1024;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_scratch_bind_mu;
void* g_scratch_ptr[kMaxDevices] = {};
// Serialises the no-allocation calls of the process: they share each GPU's
// scratch and the host staging arrays below.
std::mutex g_noalloc_mu;
float g_noalloc_R[(size_t)kNoAllocMaxProjs * kRStride];
double g_noalloc_sums[(size_t)kNoAllocMaxProjs * 4];

// With `device` current: its address of `scratch`, looked up once.
int bind_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return GENFIRE_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_scratch_bind_mu);
    if (g_scratch_ptr[device]) return GENFIRE_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GENFIRE_ERR_HIP;
    g_scratch_ptr[device] = p;
    return GENFIRE_OK;
}

}  // namespace

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
    p->calculate_Rfree = 0;
    p->numBinsRfree = 10;
    p->percentValuesForRfree = 0.05;
    p->RandomSeed = 0;
    p->griddingMethod = 1;
    p->allowMultipleGridMatches = 1;
    p->projectorInterpolation = GENFIRE_INTERP_LINEAR;
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
    const int rc = bind_scratch(device);
    if (rc != GENFIRE_OK) return rc;
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
        use_device(ctx);
        auto* j = new genfire_job();
        std::unique_ptr<genfire_job> guard(j);
        job_setup(j, ctx, p, dim1, dim2, num_projs);
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
        job_set_inputs(j, projections, angles, support, initial_object);
    });
}

int genfire_job_set_inputs_f16(genfire_job* j, const uint16_t* projections_f16, const float* proj_scale,
                               const double* angles, const uint8_t* support, const float* initial_object)
{
    if (!j) return GENFIRE_ERR_ARG;
    if (!projections_f16 || !proj_scale || !angles) {
        set_err(j->ctx, "GENFIRE: projections, their scales and angles are required");
        return GENFIRE_ERR_ARG;
    }
    return guarded(j->ctx, [&] {
        use_device(j->ctx);
        job_set_inputs(j, nullptr, angles, support, initial_object, projections_f16, proj_scale);
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
    if (j->prm.griddingMethod == 1) {
        set_err(j->ctx, "GENFIRE: FFT gridding forms no point list");
        return GENFIRE_ERR_STATE;
    }
    *n = j->npts;
    return GENFIRE_OK;
}

int genfire_job_get_points(genfire_job* j, int64_t* ind, int32_t* proj, float* dist, float* un, float* vn,
                           float* values)
{
    if (!j) return GENFIRE_ERR_ARG;
    return guarded(j->ctx, [&] {
        GF_CHECK(j->gridded, GENFIRE_ERR_STATE, "GENFIRE: not gridded");
        GF_CHECK(j->prm.griddingMethod != 1, GENFIRE_ERR_STATE, "GENFIRE: FFT gridding forms no point list");
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
        const int64_t n = j->prm.griddingMethod == 1 ? full_size(j) : j->hg.total();
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

int genfire_job_set_snapshots(genfire_job* j, int32_t save_loopLength, genfire_snapshot_fn fn, void* user)
{
    if (!j) return GENFIRE_ERR_ARG;
    const bool on = save_loopLength > 0 && fn != nullptr;
    j->snap_L = on ? save_loopLength : 0;
    j->snap_fn = on ? fn : nullptr;
    j->snap_user = on ? user : nullptr;
    return GENFIRE_OK;
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

int genfire_job_get_Rfree(genfire_job* j, float* out)
{
    if (!j || !out) return GENFIRE_ERR_ARG;
    if (!j->reconstructed) return GENFIRE_ERR_STATE;
    if (j->rf_bins <= 0) {
        set_err(j->ctx, "GENFIRE: Rfree was not calculated (calculate_Rfree = 0)");
        return GENFIRE_ERR_STATE;
    }
    std::memcpy(out, j->Rfree_h.data(), sizeof(float) * j->Rfree_h.size());
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

int genfire_job_graph_replays(const genfire_job* j, int64_t* replays)
{
    if (!j || !replays) return GENFIRE_ERR_ARG;
    *replays = j->rs.graph_launches;
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

namespace {

// The model's forms the projections read (project_forms), on the device.
struct ModelForms {
    DevBuf vy, zy, zx, vc;
    void build(const ProjForms& f, const float* d_vol, int d1, int d2, int d3, hipStream_t st)
    {
        if (f.cubic) {
            vc.alloc(sizeof(float) * cubic_form_floats(d1, d2, d3));
            volume_cubic_form(d_vol, d1, d2, d3, vc.as<float>(), st);
        }
        if (f.four_ray) {
            vy.alloc(sizeof(float) * (size_t)d1 * d2 * d3);
            volume_yfast(d_vol, vy.as<float>(), d1, d2, d3, st);
        }
        if (f.zy) {
            zy.alloc(sizeof(float) * zdup_floats(d1, d2, d3));
            volume_zdup(d_vol, d1, d2, d3, zy.as<float>(), true, st);
        }
        if (f.zx) {
            zx.alloc(sizeof(float) * zdup_floats(d1, d2, d3));
            volume_zdup(d_vol, d1, d2, d3, zx.as<float>(), false, st);
        }
    }
    void set(const ProjForms& f, ProjArgs* pa) const
    {
        pa->four_ray = f.four_ray ? 1 : 0;
        pa->vol_y = vy.as<float>();
        pa->zy = zy.as<float>();
        pa->zx = zx.as<float>();
        pa->vcubic = f.cubic ? vc.as<float>() : nullptr;
    }
};

// A stream that is destroyed on every way out of the call.
struct OwnStream {
    hipStream_t s = nullptr;
    OwnStream() { GF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~OwnStream()
    {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
    OwnStream(const OwnStream&) = delete;
    OwnStream& operator=(const OwnStream&) = delete;
};

void project_one(genfire_ctx* ctx, const float* vol, int d1, int d2, int d3, const double ang[3],
                 const double* vectors, float* proj, bool cubic = false)
{
    use_device(ctx);
    GF_CHECK(d1 >= 2 && d2 >= 2 && d3 >= 2, GENFIRE_ERR_ARG, "GENFIRE: volume dimensions must be >= 2");
    GF_CHECK(!cubic || (d1 >= 4 && d2 >= 4 && d3 >= 4), GENFIRE_ERR_ARG,
             "GENFIRE: interp3 'cubic' needs at least 4 nodes along every axis");
    GF_CHECK((int64_t)d1 * d2 * d3 < (1ll << 31), GENFIRE_ERR_UNSUPPORTED, "GENFIRE: volume above 2^31 voxels");
    DevBuf v, r, o;
    ModelForms mf;
    OwnStream os;   // declared last: synchronised before the buffers are freed
    hipStream_t st = os.s;
    const size_t nv = (size_t)d1 * d2 * d3;
    v.alloc(sizeof(float) * nv);
    o.alloc(sizeof(float) * (size_t)d1 * d2);
    bool any_y = false;
    std::vector<float> Rf = proj_matrices(ang, 1, &any_y, vectors);
    const ProjForms f = project_forms(Rf.data(), 1, d2, cubic);
    r.alloc(sizeof(float) * Rf.size());
    GF_HIP(hipMemcpyAsync(v.p, vol, sizeof(float) * nv, hipMemcpyHostToDevice, st));
    GF_HIP(hipMemcpyAsync(r.p, Rf.data(), sizeof(float) * Rf.size(), hipMemcpyHostToDevice, st));
    mf.build(f, v.as<float>(), d1, d2, d3, st);
    ProjArgs pa{};
    mf.set(f, &pa);
    pa.d1 = d1; pa.d2 = d2; pa.d3 = d3;
    pa.R = r.as<float>();
    pa.P = 1;
    pa.out = o.as<float>();
    pa.acc = pa.out;   // the slab sweep keeps its running ray sums in the output
    pa.m1 = d1; pa.m2 = d2;
    project(pa, st);
    GF_HIP(hipMemcpyAsync(proj, o.p, o.bytes, hipMemcpyDeviceToHost, st));
    GF_HIP(hipStreamSynchronize(st));
}

void rfactor_all(genfire_ctx* ctx, const float* model, int d1, int d2, int d3, const float* projections, int p1,
                 int p2, int P, const double* angles, const double* vectors, double* Rtotal, double* Rarr,
                 float* simu, bool cubic = false)
{
    use_device(ctx);
    GF_CHECK(P >= 1, GENFIRE_ERR_ARG, "GENFIRE: no projections");
    GF_CHECK(d1 >= 2 && d2 >= 2 && d3 >= 2, GENFIRE_ERR_ARG, "GENFIRE: model dimensions must be >= 2");
    GF_CHECK(!cubic || (d1 >= 4 && d2 >= 4 && d3 >= 4), GENFIRE_ERR_ARG,
             "GENFIRE: interp3 'cubic' needs at least 4 nodes along every axis");
    GF_CHECK(p1 >= 1 && p2 >= 1, GENFIRE_ERR_ARG, "GENFIRE: empty projections");
    GF_CHECK(p1 <= d1 && p2 <= d2, GENFIRE_ERR_ARG, "GENFIRE: projection larger than the model");
    GF_CHECK((int64_t)d1 * d2 * d3 < (1ll << 31), GENFIRE_ERR_UNSUPPORTED, "GENFIRE: model above 2^31 voxels");
    DevBuf v, r, o, m, s, s4;
    ModelForms mf;
    OwnStream os;   // declared last: synchronised before the buffers are freed
    hipStream_t st = os.s;
    const size_t nv = (size_t)d1 * d2 * d3, np = (size_t)p1 * p2 * P;
    v.alloc(sizeof(float) * nv);
    m.alloc(sizeof(float) * np);
    o.alloc(sizeof(float) * np);
    bool any_y = false;
    std::vector<float> Rf = proj_matrices(angles, P, &any_y, vectors);
    const ProjForms f = project_forms(Rf.data(), P, d2, cubic);
    r.alloc(sizeof(float) * Rf.size());
    GF_HIP(hipMemcpyAsync(v.p, model, sizeof(float) * nv, hipMemcpyHostToDevice, st));
    GF_HIP(hipMemcpyAsync(m.p, projections, sizeof(float) * np, hipMemcpyHostToDevice, st));
    GF_HIP(hipMemcpyAsync(r.p, Rf.data(), sizeof(float) * Rf.size(), hipMemcpyHostToDevice, st));
    mf.build(f, v.as<float>(), d1, d2, d3, st);
    const int nblk = project_blocks(p1, p2);
    s.alloc(sizeof(double) * (size_t)P * 4 * nblk);
    s4.alloc(sizeof(double) * (size_t)P * 4);
    ProjArgs pa{};
    mf.set(f, &pa);
    pa.d1 = d1; pa.d2 = d2; pa.d3 = d3;
    pa.R = r.as<float>();
    pa.P = P;
    pa.out = o.as<float>();
    pa.acc = pa.out;   // the slab sweep keeps its running ray sums in the output
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
    double mean = 0.0;
    for (int p = 0; p < P; ++p) {
        const double rr = h[4 * p] / h[4 * p + 2];
        if (Rarr) Rarr[p] = rr;
        mean += rr;
    }
    if (Rtotal) *Rtotal = mean / P;
}

}  // namespace

int genfire_project(genfire_ctx* ctx, const float* vol, int32_t d1, int32_t d2, int32_t d3, double phi,
                    double theta, double psi, float* proj)
{
    if (!ctx || !vol || !proj) return GENFIRE_ERR_ARG;
    const double ang[3] = {phi, theta, psi};
    return guarded(ctx, [&] { project_one(ctx, vol, d1, d2, d3, ang, nullptr, proj); });
}

int genfire_project_cubic(genfire_ctx* ctx, const float* vol, int32_t d1, int32_t d2, int32_t d3, double phi,
                          double theta, double psi, float* proj)
{
    if (!ctx || !vol || !proj) return GENFIRE_ERR_ARG;
    const double ang[3] = {phi, theta, psi};
    return guarded(ctx, [&] { project_one(ctx, vol, d1, d2, d3, ang, nullptr, proj, true); });
}

int genfire_project_general(genfire_ctx* ctx, const float* vol, int32_t d1, int32_t d2, int32_t d3, double phi,
                            double theta, double psi, const double vector1[3], const double vector2[3],
                            const double vector3[3], float* proj)
{
    if (!ctx || !vol || !proj || !vector1 || !vector2 || !vector3) return GENFIRE_ERR_ARG;
    const double ang[3] = {phi, theta, psi};
    double vec[9];
    for (int i = 0; i < 3; ++i) { vec[i] = vector1[i]; vec[3 + i] = vector2[i]; vec[6 + i] = vector3[i]; }
    return guarded(ctx, [&] { project_one(ctx, vol, d1, d2, d3, ang, vec, proj); });
}

int genfire_rfactor(genfire_ctx* ctx, const float* model, int32_t d1, int32_t d2, int32_t d3,
                    const float* projections, int32_t p1, int32_t p2, int32_t P, const double* angles,
                    double* Rtotal, double* Rarr, float* simu)
{
    if (!ctx || !model || !projections || !angles || P < 1) return GENFIRE_ERR_ARG;
    return guarded(ctx, [&] {
        rfactor_all(ctx, model, d1, d2, d3, projections, p1, p2, P, angles, nullptr, Rtotal, Rarr, simu);
    });
}

int genfire_rfactor_cubic(genfire_ctx* ctx, const float* model, int32_t d1, int32_t d2, int32_t d3,
                          const float* projections, int32_t p1, int32_t p2, int32_t P, const double* angles,
                          double* Rtotal, double* Rarr, float* simu)
{
    if (!ctx || !model || !projections || !angles || P < 1) return GENFIRE_ERR_ARG;
    return guarded(ctx, [&] {
        rfactor_all(ctx, model, d1, d2, d3, projections, p1, p2, P, angles, nullptr, Rtotal, Rarr, simu, true);
    });
}

int genfire_rfactor_general(genfire_ctx* ctx, const float* model, int32_t d1, int32_t d2, int32_t d3,
                            const float* projections, int32_t p1, int32_t p2, int32_t P, const double* angles,
                            const double vector1[3], const double vector2[3], const double vector3[3],
                            double* Rtotal, double* Rarr, float* simu)
{
    if (!ctx || !model || !projections || !angles || P < 1 || !vector1 || !vector2 || !vector3)
        return GENFIRE_ERR_ARG;
    double vec[9];
    for (int i = 0; i < 3; ++i) { vec[i] = vector1[i]; vec[3 + i] = vector2[i]; vec[6 + i] = vector3[i]; }
    return guarded(ctx, [&] {
        rfactor_all(ctx, model, d1, d2, d3, projections, p1, p2, P, angles, vec, Rtotal, Rarr, simu);
    });
}

int genfire_project_general_cubic(genfire_ctx* ctx, const float* vol, int32_t d1, int32_t d2, int32_t d3,
                                  double phi, double theta, double psi, const double vector1[3],
                                  const double vector2[3], const double vector3[3], float* proj)
{
    if (!ctx || !vol || !proj || !vector1 || !vector2 || !vector3) return GENFIRE_ERR_ARG;
    const double ang[3] = {phi, theta, psi};
    double vec[9];
    for (int i = 0; i < 3; ++i) { vec[i] = vector1[i]; vec[3 + i] = vector2[i]; vec[6 + i] = vector3[i]; }
    return guarded(ctx, [&] { project_one(ctx, vol, d1, d2, d3, ang, vec, proj, true); });
}

int genfire_rfactor_general_cubic(genfire_ctx* ctx, const float* model, int32_t d1, int32_t d2, int32_t d3,
                                  const float* projections, int32_t p1, int32_t p2, int32_t P, const double* angles,
                                  const double vector1[3], const double vector2[3], const double vector3[3],
                                  double* Rtotal, double* Rarr, float* simu)
{
    if (!ctx || !model || !projections || !angles || P < 1 || !vector1 || !vector2 || !vector3)
        return GENFIRE_ERR_ARG;
    double vec[9];
    for (int i = 0; i < 3; ++i) { vec[i] = vector1[i]; vec[3 + i] = vector2[i]; vec[6 + i] = vector3[i]; }
    return guarded(ctx, [&] {
        rfactor_all(ctx, model, d1, d2, d3, projections, p1, p2, P, angles, vec, Rtotal, Rarr, simu, true);
    });
}

int genfire_rfactor_noalloc(genfire_ctx* ctx, const float* d_model, int32_t d1, int32_t d2, int32_t d3,
                            const float* d_projections, int32_t p1, int32_t p2, int32_t P, const double* angles,
                            double* Rtotal, double* Rarr, float* d_simuProjs)
{
    if (!ctx || !d_model || !d_projections || !angles || P < 1) return GENFIRE_ERR_ARG;
    return guarded(ctx, [&] {
        use_device(ctx);
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        GF_CHECK(P <= kNoAllocMaxProjs, GENFIRE_ERR_UNSUPPORTED,
                 "GENFIRE: genfire_rfactor_noalloc takes at most " + std::to_string(kNoAllocMaxProjs) + " projections");
        GF_CHECK(d1 >= 2 && d2 >= 2 && d3 >= 2, GENFIRE_ERR_ARG, "GENFIRE: model dimensions must be >= 2");
        GF_CHECK(p1 >= 1 && p2 >= 1, GENFIRE_ERR_ARG, "GENFIRE: empty projections");
        GF_CHECK(p1 <= d1 && p2 <= d2, GENFIRE_ERR_ARG, "GENFIRE: projection larger than the model");
        GF_CHECK((int64_t)d1 * d2 * d3 < (1ll << 31), GENFIRE_ERR_UNSUPPORTED, "GENFIRE: model above 2^31 voxels");
        char* base = static_cast<char*>(g_scratch_ptr[ctx->device]);
        GF_CHECK(base, GENFIRE_ERR_STATE, "GENFIRE: no scratch address for this device");
        proj_matrices_into(angles, P, nullptr, g_noalloc_R);
        const ProjForms f = project_forms(g_noalloc_R, P, d2);
        // scratch layout, 256-byte aligned pieces: R matrices, tile partial sums,
        // per-projection sums, then the model forms the projections read
        const int nblk = project_blocks(p1, p2);
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t b_R = up(sizeof(float) * (size_t)kRStride * P);
        const size_t b_s = up(sizeof(double) * (size_t)P * 4 * nblk);
        const size_t b_s4 = up(sizeof(double) * (size_t)P * 4);
        const size_t b_vy = f.four_ray ? up(sizeof(float) * (size_t)d1 * d2 * d3) : 0;
        const size_t b_zy = f.zy ? up(sizeof(float) * (size_t)zdup_floats(d1, d2, d3)) : 0;
        const size_t b_zx = f.zx ? up(sizeof(float) * (size_t)zdup_floats(d1, d2, d3)) : 0;
        const size_t need = b_R + b_s + b_s4 + b_vy + b_zy + b_zx;
        GF_CHECK(need <= kNoAllocScratchBytes, GENFIRE_ERR_UNSUPPORTED,
                 "GENFIRE: genfire_rfactor_noalloc needs " + std::to_string(need) + " bytes of scratch for this model "
                 "and these projections; its limit is " + std::to_string(kNoAllocScratchBytes));
        float* dR = reinterpret_cast<float*>(base);
        double* ds = reinterpret_cast<double*>(base + b_R);
        double* ds4 = reinterpret_cast<double*>(base + b_R + b_s);
        char* forms = base + b_R + b_s + b_s4;
        float* dvy = f.four_ray ? reinterpret_cast<float*>(forms) : nullptr;
        float* dzy = f.zy ? reinterpret_cast<float*>(forms + b_vy) : nullptr;
        float* dzx = f.zx ? reinterpret_cast<float*>(forms + b_vy + b_zy) : nullptr;
        hipStream_t st = nullptr;   // the null stream: nothing to create
        GF_HIP(hipMemcpyAsync(dR, g_noalloc_R, sizeof(float) * (size_t)kRStride * P, hipMemcpyHostToDevice, st));
        if (dvy) volume_yfast(d_model, dvy, d1, d2, d3, st);
        if (dzy) volume_zdup(d_model, d1, d2, d3, dzy, true, st);
        if (dzx) volume_zdup(d_model, d1, d2, d3, dzx, false, st);
        ProjArgs pa{};
        pa.four_ray = f.four_ray ? 1 : 0;
        pa.vol_y = dvy;
        pa.zy = dzy;
        pa.zx = dzx;
        pa.d1 = d1; pa.d2 = d2; pa.d3 = d3;
        pa.R = dR;
        pa.P = P;
        pa.out = d_simuProjs;   // acc stays null: one pass
        pa.meas = d_projections;
        pa.m1 = p1; pa.m2 = p2;
        pa.sums = ds;
        pa.post_hoc = 1;
        project(pa, st);
        project_reduce(ds, P, nblk, ds4, st);
        GF_HIP(hipMemcpyAsync(g_noalloc_sums, ds4, sizeof(double) * (size_t)P * 4, hipMemcpyDeviceToHost, st));
        GF_HIP(hipStreamSynchronize(st));
        double mean = 0.0;
        for (int p = 0; p < P; ++p) {
            const double rr = g_noalloc_sums[4 * p] / g_noalloc_sums[4 * p + 2];
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
