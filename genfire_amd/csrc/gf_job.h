// gf_job.h -- one reconstruction job (the state of one @GENFIRE_Reconstructor
// object, GENFIRE_Reconstructor.m:5-71) and the host phases that run it on a
// HIP stream.  Shared by the C ABI (gf_capi.hip) and the request server
// (gf_serve.hip).
#pragma once

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "gf_grid.h"

struct genfire_ctx {
    int device = 0;
    std::mutex mu;
    std::string err;
};

namespace gf {

struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    void init()
    {
        if (!a) {
            GF_HIP(hipEventCreate(&a));
            GF_HIP(hipEventCreate(&b));
        }
    }
    EventTimer() = default;
    EventTimer(const EventTimer&) = delete;
    EventTimer& operator=(const EventTimer&) = delete;
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
    std::pair<hipEvent_t, hipEvent_t> take();
    void collect();
    void reset();
    ~KStats();
};

double mround(double x);                                  // MATLAB round
std::vector<int> cutoff_schedule(int N, int m);           // round(linspace(1, N, m))
std::vector<double> delay_indicators(int mode);           // GENFIRE_Reconstructor.m:156-165

// Iteration state between the phases of reconstruct.m (set up by recon_begin).
struct ReconState {
    int N = 0, L = 1, nR = 0;
    bool half = false, dr = false, monitor = false;
    int64_t hn = 0, nparts = 0;
    double den = 0.0;
    std::vector<int> sched;
    size_t cur = 0;
    IterArgs ia{};
    HalfIterArgs ha{};
    EventTimer all, loop;
    std::vector<std::unique_ptr<EventTimer>> rf;   // one pair per R-factor evaluation
    bool begun = false;
    // one captured steady-state iteration (no schedule stage, no monitor, no
    // first or last pass), replayed for every such iteration
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_launches = 0;
    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
    ReconState() = default;
    ReconState(const ReconState&) = delete;
    ReconState& operator=(const ReconState&) = delete;
    ~ReconState() { drop_graph(); }
};

}  // namespace gf

struct genfire_job {
    genfire_ctx* ctx = nullptr;
    genfire_params prm{};
    int dim1 = 0, dim2 = 0, P = 0, n1 = 0, n2 = 0;
    gf::HalfGrid hg{};
    std::vector<double> delays;
    hipStream_t st = nullptr;       // the job's own stream
    hipStream_t run_st = nullptr;   // stream its iterations run on (st, or a server's group stream)

    // inputs
    gf::DevBuf proj, support_box, init_box;
    std::vector<double> angles;  // P x 3 column-major
    bool have_inputs = false, have_support = false, have_init = false;

    // gridding
    gf::DevBuf planes, blk_counts, blk_offsets, total, ind, dist, un, vn, F, tiles, Pt, Phi, Plo, Pscale, nu_inv, scan_tmp;
    gf::DevBuf acc, sw, hits, exact;
    gf::DevBuf M, mask, recifft;
    std::vector<int64_t> proj_off;
    int64_t npts = 0;
    bool gridded = false, have_mask = false;

    // iteration
    gf::DevBuf x, K, S, best, pfbuf, err_part, err_chunk, errK, errctr, den_tmp, den, recon;
    gf::DevBuf csr_cnt, csr_off, csr_coff, csr_meta, csr_val, csr_mag, csr_d, csr_mate;
    int64_t nmeas = 0;
    gf::DevBuf vol, vol_y, zy, zx, vcubic, Rmat, rsums, rout, racc;
    // Rfree: withheld voxels per half pencil, their terms, shell order and denominators
    gf::DevBuf rf_off, rf_meta, rf_val, rf_terms, rf_order, rf_shell_off, rf_den, Rfree;
    int rf_bins = 0;
    std::vector<float> Rfree_h;
    std::vector<int64_t> tile_meta_h;    // projection offsets and first tiles (outlive their asynchronous copy)
    gf::DevBuf tile_meta;
    std::vector<float> nu_inv_h;         // host source of nu_inv, the NUFFT deconvolution factors
    gf::ProjForms rforms;         // the model forms the monitor's projections read
    std::vector<float> errK_h, errR_h;
    // save_temp snapshots (reconstruct.m:155-159): every snap_L iterations the
    // best P_R(x) since the last schedule stage goes to snap_fn
    int snap_L = 0;
    genfire_snapshot_fn snap_fn = nullptr;
    void* snap_user = nullptr;
    gf::DevBuf cand[2], best_box, best_err, best_flag;
    std::vector<float> snap_h;
    std::vector<double> Rarr_h, Rarr2_h;
    bool reconstructed = false;
    gf::ReconState rs;

    // timing
    double t_grid = 0, t_iter = 0, t_rf = 0, t_total = 0;
    bool prof = false;
    gf::KStats ks;

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

namespace gf {

void set_err(genfire_ctx* ctx, const std::string& m);
void use_device(genfire_ctx* ctx);

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

inline int64_t full_size(const genfire_job* j) { return (int64_t)j->n1 * j->n2 * j->n1; }
inline int64_t box_size(const genfire_job* j) { return (int64_t)j->dim1 * j->dim2 * j->dim1; }

// Fourier grid edges: (n1_oversampled, n2_oversampled) for DFT gridding, (d, d)
// for FFT gridding; raises when the FFT-gridding edge is not a whole number.
void grid_edges(const genfire_params& p, int dim1, int dim2, int* n1, int* n2);
// The checks of CheckPrepareData (GENFIRE_Reconstructor.m:115-171) and of this
// build's size limits; raises with the reference's message where it has one.
void validate_job(const genfire_params* p, int dim1, int dim2, int P);
// Validates the parameters and fills sizes, schedule and stream of a fresh job.
void job_setup(genfire_job* j, genfire_ctx* ctx, const genfire_params* p, int dim1, int dim2, int P);
// Uploads the inputs (readFiles, :82-112); support and initial object may be
// null; projections either FP32, or binary16 with one scale per projection.
void job_set_inputs(genfire_job* j, const float* projections, const double* angles, const uint8_t* support,
                    const float* initial_object, const uint16_t* proj_f16 = nullptr,
                    const float* proj_scale = nullptr);
// Device bytes a job of these sizes holds at its peak (an upper estimate).
int64_t job_footprint(const genfire_params& p, int dim1, int dim2, int P);

// Projector matrices, kRStride floats per projection: the ZYZ matrix of
// calculate3Dprojection_RealSpaceinterp.m:13-15, or with vectors (v1, v2, v3,
// 9 doubles) the quaternion product of ..._general.m:3-7.
std::vector<float> proj_matrices(const double* angles, int P, bool* any_y, const double* vectors = nullptr);
// The same matrices written into Rm (P x kRStride floats), allocating nothing.
void proj_matrices_into(const double* angles, int P, const double* vectors, float* Rm);

// runGridding: fillInFourierGrid_DFT + CentroSymmetricity (griddingMethod 2) or
// fillInFourierGrid + hermitianSymmetrize (griddingMethod 1), and recIFFT.
void do_grid(genfire_job* j);

// reconstruct.m in phases.  recon_begin sets up (support, initial object,
// measured-voxel lists, errK denominator); an iteration is recon_pre (schedule
// stage, first x pass), the transform passes (recon_body for one job,
// half_body_batch for several of one shape), and recon_post (errK, R-factor
// monitor); recon_end crops the result and reads the records back.  All launch
// on j->run_st.
void recon_begin(genfire_job* j);
void recon_pre(genfire_job* j, int it);
void recon_body(genfire_job* j, int it);
void recon_post(genfire_job* j, int it);
void recon_end(genfire_job* j);
void do_reconstruct(genfire_job* j);

// True when iteration `it` of the job launches the same kernels with the same
// arguments as its neighbours (no enforcement stage, no R-factor evaluation, not
// the first or the last two iterations, no snapshots): such iterations may
// replay one captured graph.
bool steady_iteration(const genfire_job* j, int it);

// The passes of iteration its[i] of every job (half-spectrum path, one shape)
// as batched launches on st.
void half_body_batch(genfire_job* const* jobs, const int* its, int count, hipStream_t st);

// The Rfree reduction of the job's current iteration (no-op without Rfree).
void rfree_step(genfire_job* j, hipStream_t st);

// Frees what the iterations no longer need (point lists, hits, measuredK).
void release_grid_extras(genfire_job* j);
// The gridding's FP64 accumulators and value operands, which a job keeps for
// its next gridding (part of release_grid_extras).
void release_grid_scratch(genfire_job* j);

}  // namespace gf
