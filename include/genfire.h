/*
 * genfire.h -- C ABI of libgenfire, the MI355X (gfx950) implementation of the
 * GENFIRE v2 reconstruction hot path: DFT gridding, the GENFIRE / Douglas-Rachford
 * iteration and the real-space R-factor.
 *
 * Conventions
 *  - Plain C types only.  Arrays are caller-allocated HOST buffers in MATLAB
 *    (column-major, dim 1 fastest) layout; complex arrays are interleaved
 *    (re, im) float pairs, the MATLAB R2018a+ interleaved-complex layout.
 *  - Every call returns GENFIRE_OK (0) or a negative status; the message of the
 *    last failure on a context is returned by genfire_last_error().  The MATLAB
 *    reference raises error('GENFIRE: ...') in the same situations.
 *  - A context owns one GPU; a job owns the device-resident state of one
 *    reconstruction (projections, k-grid, support, iterate) and its own HIP
 *    stream.  Use one job per host thread; calls return when the work is done,
 *    except genfire_job_submit, which only enqueues.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *  - @GENFIRE_Reconstructor/GENFIRE_Reconstructor.m:5-71   properties -> genfire_params
 *  - GENFIRE_Reconstructor.m:115-171 CheckPrepareData      -> genfire_job_create
 *  - GENFIRE_Reconstructor.m:172-186 runGridding(method 2) -> genfire_job_grid
 *  - fillInFourierGrid_DFT.m:85-310                          -> genfire_job_grid
 *  - CentroSymmetricity.m:2-72                               -> genfire_job_grid
 *  - reconstruct.m:28-194                                    -> genfire_job_reconstruct
 *  - src/calculate3Dprojection_RealSpaceinterp.m:1-57        -> genfire_project
 *  - src/Tian_calc_Rfactor_realspace.m:1-24                  -> genfire_rfactor
 */
#ifndef GENFIRE_H
#define GENFIRE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GENFIRE_OK 0
#define GENFIRE_ERR_ARG (-1)         /* invalid argument or size                 */
#define GENFIRE_ERR_HIP (-2)         /* HIP runtime error                         */
#define GENFIRE_ERR_NOMEM (-3)       /* device allocation failed                  */
#define GENFIRE_ERR_STATE (-4)       /* call out of order (e.g. reconstruct before grid) */
#define GENFIRE_ERR_UNSUPPORTED (-5) /* size or option the library does not handle */

#define GENFIRE_ALG_GENFIRE 0        /* reconstruct.m:117-186                    */
#define GENFIRE_ALG_DR 1             /* Douglas-Rachford on the same projections */

/* projectorInterpolation: the projector's interpolant */
#define GENFIRE_INTERP_LINEAR 0   /* trilinear (the default)       */
#define GENFIRE_INTERP_CUBIC 1    /* interp3 'cubic', as calculate3Dprojection_RealSpaceinterp.m:45,51 */

#define GENFIRE_PREC_FP32 0          /* gridding sums in FP32                    */
#define GENFIRE_PREC_SPLIT_BF16 1    /* gridding GEMM in split BF16 (hi+lo)      */
#define GENFIRE_PREC_FP16 2          /* gridding GEMM in FP16, per-projection scale */

typedef struct genfire_ctx genfire_ctx;
typedef struct genfire_job genfire_job;

/* Mirrors the GENFIRE_Reconstructor properties (GENFIRE_Reconstructor.m:30-70),
 * same names and defaults; algorithm, precision and compute_recIFFT are
 * additions of this library, and the Rfree fields define the properties
 * reconstruct.m:46-69 reads but the class never declares (numBinsRfree,
 * percentValuesForRfree, RandomSeed). */
typedef struct genfire_params {
    int32_t numIterations;              /* 50    (:42) */
    double oversamplingRatio;           /* 3     (:43) */
    double interpolationCutoffDistance; /* 0.7   (:44) */
    int32_t constraintEnforcementMode;  /* 1     (:58)  1, 2 or 3 */
    int32_t constraintPositivity;       /* 1     (:59) */
    int32_t constraintSupport;          /* 1     (:60) */
    int32_t DFT_CentroSymmetricity;     /* 1     (:63) */
    double vector1[3];                  /* [0 0 1] (:50) phi axis   */
    double vector2[3];                  /* [0 1 0] (:51) theta axis */
    double vector3[3];                  /* [1 0 0] (:52) psi axis   */
    int32_t monitor_R;                  /* 0     (:31) */
    int32_t monitorR_loopLength;        /* 10    (:32) */
    int32_t algorithm;                  /* GENFIRE_ALG_GENFIRE */
    int32_t precision;                  /* GENFIRE_PREC_FP32 */
    int32_t compute_recIFFT;            /* 1: also form recIFFT (fillInFourierGrid_DFT.m:304) */
    int32_t calculate_Rfree;            /* 0     (:47) withhold points and monitor Rfree */
    int32_t numBinsRfree;               /* 10    resolution shells (reconstruct.m:47) */
    double percentValuesForRfree;       /* 0.05  fraction withheld per shell (:58) */
    int64_t RandomSeed;                 /* 0     seed of the withheld-point permutation (:52-55) */
    int32_t griddingMethod;             /* 1     (:45)  1 FFT gridding, 2 DFT gridding */
    int32_t allowMultipleGridMatches;   /* 1     (:46)  FFT gridding: match every voxel within the cutoff */
    int32_t projectorInterpolation;     /* GENFIRE_INTERP_LINEAR: the R-factor monitor's interpolant */
} genfire_params;

/* Library version, major*10000 + minor*100 + patch. */
int genfire_version(void);

/* Fills p with the defaults listed above (GENFIRE_Reconstructor.m:30-70). */
int genfire_params_default(genfire_params* p);

/* Context: one GPU.  No counterpart in the reference (MATLAB owns its device). */
int genfire_ctx_create(int32_t device, genfire_ctx** out);
int genfire_ctx_destroy(genfire_ctx* ctx);
const char* genfire_last_error(const genfire_ctx* ctx);

/* CheckPrepareData (GENFIRE_Reconstructor.m:115-171): validates the parameters,
 * sets n1_oversampled = round(Dim1*ratio), n2_oversampled = round(Dim2*ratio)
 * (:168-169) and the enforcement schedule (:156-165), and creates a job.  The
 * Fourier grid is n1 x n2 x n1 for DFT gridding; FFT gridding builds a cube of
 * edge d = Dim2*ratio (Dim2 even) or 1 + (Dim2-1)*ratio (Dim2 odd), which must
 * be a whole number no smaller than Dim1 (fillInFourierGrid.m:38-42, 60). */
int genfire_job_create(genfire_ctx* ctx, const genfire_params* p, int32_t dim1, int32_t dim2,
                       int32_t num_projs, genfire_job** out);
int genfire_job_destroy(genfire_job* job);

/* The grid edges (n1_oversampled, n2_oversampled for DFT gridding; d, d for
 * FFT gridding) and the half-grid extent (l1, l2, l3) of
 * fillInFourierGrid_DFT.m:134-136. */
int genfire_job_sizes(const genfire_job* job, int32_t* n1, int32_t* n2, int64_t half_len[3]);

/* Uploads InputProjections (Dim1 x Dim2 x P float), InputAngles (P x 3 double,
 * column-major: phi[P], theta[P], psi[P], degrees), Support (Dim1 x Dim2 x Dim1
 * bytes, nonzero = inside; NULL = all ones, :146-148) and initialObject
 * (Dim1 x Dim2 x Dim1 float or NULL = zeros, reconstruct.m:76-80). */
int genfire_job_set_inputs(genfire_job* job, const float* projections, const double* angles,
                           const uint8_t* support, const float* initial_object);

/* As genfire_job_set_inputs with the projections stored as IEEE binary16 with
 * one scale per projection: value(a, b, p) = half(a, b, p) * proj_scale[p].  No
 * counterpart in the reference (which reads double projections,
 * GENFIRE_Reconstructor.m:82-112); for callers that keep large tilt series in
 * half precision.  The projections are expanded to FP32 on the device. */
int genfire_job_set_inputs_f16(genfire_job* job, const uint16_t* projections_f16, const float* proj_scale,
                               const double* angles, const uint8_t* support, const float* initial_object);

/* runGridding (GENFIRE_Reconstructor.m:172-186).  griddingMethod = 2:
 * fillInFourierGrid_DFT + CentroSymmetricity; griddingMethod = 1:
 * fillInFourierGrid (FFT of each padded projection, rotated nearest-voxel
 * inverse-distance average, hermitianSymmetrize); then the single cast and
 * measuredK_mask. */
int genfire_job_grid(genfire_job* job);

/* Gridded point list in projection order (the master lists of
 * fillInFourierGrid_DFT.m:257-261): 0-based half-grid linear index, projection
 * (0-based), distance D, u/n1, v/n2 and the DFT value (interleaved complex).
 * Any output pointer may be NULL.  DFT gridding only (GENFIRE_ERR_STATE after
 * FFT gridding, which forms no point list). */
int genfire_job_grid_count(const genfire_job* job, int64_t* num_points);
int genfire_job_get_points(genfire_job* job, int64_t* ind, int32_t* proj, float* dist, float* un,
                           float* vn, float* values);

/* measuredK (n1 x n2 x n1 interleaved complex, centred as in the reference) and
 * the number of gridded points that landed on each half-grid voxel (DFT
 * gridding), or on each voxel of the centred d x d x d grid before
 * hermitianSymmetrize (FFT gridding: the entries of masterInd per voxel,
 * fillInFourierGrid.m:139). */
int genfire_job_get_measuredK(genfire_job* job, float* measuredK);
int genfire_job_get_hits(genfire_job* job, int32_t* hits);
/* recIFFT (Dim1 x Dim2 x Dim1), fillInFourierGrid_DFT.m:304. */
int genfire_job_get_recIFFT(genfire_job* job, float* rec);

/* Replaces measuredK (centred, interleaved complex) and measuredK_mask (n1*n2*n1
 * bytes or NULL = measuredK ~= 0, GENFIRE_Reconstructor.m:183) on the device. */
int genfire_job_set_measuredK(genfire_job* job, const float* measuredK, const uint8_t* mask);

/* reconstruct (reconstruct.m:28-194) or the Douglas-Rachford variant. */
int genfire_job_reconstruct(genfire_job* job);
/* reconstruction (Dim1 x Dim2 x Dim1 float), reconstruct.m:192. */
int genfire_job_get_reconstruction(genfire_job* job, float* rec);
/* errK (numIterations floats), reconstruct.m:135. */
int genfire_job_get_errK(genfire_job* job, float* errK);
/* errR (floor(N/L)), Rarr_record and Rarr2_record (P x floor(N/L), column-major),
 * reconstruct.m:168-185.  Any pointer may be NULL. */
int genfire_job_get_errR(genfire_job* job, float* errR, double* Rarr, double* Rarr2);

/* Rfree_complex (reconstruct.m:137-147): numBinsRfree x numIterations
 * column-major, entry (shell, iteration) = sum|k(withheld) - M(withheld)| /
 * sum|M(withheld)| over the shell's withheld points, k the transform errK uses;
 * -1 where a shell has no measured point.  Requires calculate_Rfree. */
int genfire_job_get_Rfree(genfire_job* job, float* Rfree);

/* save_temp (reconstruct.m:155-159): with save_loopLength > 0 and fn set, every
 * save_loopLength-th iteration of genfire_job_reconstruct calls
 * fn(user, iteration, reconstruction) with the reconstruction the reference
 * would save there: My_stripzero(fftshift(best)) (Dim1 x Dim2 x Dim1 float,
 * host memory valid during the call), best being P_R(x) of the iteration with
 * the lowest errK since the last enforcement stage (:149-153).  The reference
 * writes it to [saveFilename num2str(iteration) '.mat']; the caller's fn does
 * what it wants with it.  save_loopLength <= 0 or fn NULL turns snapshots off.
 * With snapshots on, every iteration launches directly (no graph replay). */
typedef void (*genfire_snapshot_fn)(void* user, int32_t iteration, const float* reconstruction);
int genfire_job_set_snapshots(genfire_job* job, int32_t save_loopLength, genfire_snapshot_fn fn, void* user);

/* Asynchronous form for serving: grid + reconstruct enqueued on the job's
 * worker; poll returns 1 in *done when finished; wait blocks. */
int genfire_job_submit(genfire_job* job);
int genfire_job_poll(genfire_job* job, int32_t* done);
int genfire_job_wait(genfire_job* job);

/* Wall time of the last grid / reconstruct calls measured with HIP events on the
 * job's stream: ms[0] gridding, ms[1] iterations, ms[2] R-factor monitor,
 * ms[3] whole job. */
int genfire_job_timing(const genfire_job* job, double ms[4]);

/* Per-kernel HIP-event timing on the job's stream (off by default).
 * kernel ids: GENFIRE_KSTAT_*; returns launches and summed milliseconds. */
#define GENFIRE_KSTAT_GRID_VALUES 0   /* DFT value kernel, per projection launch  */
#define GENFIRE_KSTAT_FFT_PASS 1      /* one axis pass of the 3D FFT (complex)    */
#define GENFIRE_KSTAT_FFT_FUSED 2     /* fused forward/replace/inverse pass       */
#define GENFIRE_KSTAT_ITERATION 3     /* one whole iteration                      */
#define GENFIRE_KSTAT_PROJECT 4       /* projector + R sums, all projections      */
#define GENFIRE_KSTAT_COUNT 5
int genfire_job_set_profiling(genfire_job* job, int32_t on);
/* Iterations of the last genfire_job_reconstruct that replayed the captured
 * HIP graph of a steady-state iteration (no enforcement stage, no R-factor
 * monitor, not the first or last two); 0 with profiling on or GENFIRE_GRAPHS=0. */
int genfire_job_graph_replays(const genfire_job* job, int64_t* replays);
int genfire_job_kernel_stats(const genfire_job* job, int32_t kernel, int64_t* launches,
                             double* total_ms);

/* calculate3Dprojection_RealSpaceinterp(Vol, phi, theta, psi)
 * (src/calculate3Dprojection_RealSpaceinterp.m:1-57) with trilinear interpolation
 * and extrapolation value 0: Vol d1 x d2 x d3 float, proj d1 x d2 float. */
int genfire_project(genfire_ctx* ctx, const float* vol, int32_t d1, int32_t d2, int32_t d3,
                    double phi, double theta, double psi, float* proj);

/* Tian_calc_Rfactor_realspace(model, Projs, Angles) (src/Tian_calc_Rfactor_realspace.m:1-24),
 * with the projector above: model d1 x d2 x d3, projections p1 x p2 x P (p1 <= d1,
 * p2 <= d2), angles P x 3 column-major.  Outputs Rtotal, Rarr[P] and, if not
 * NULL, simuProjs (p1 x p2 x P). */
int genfire_rfactor(genfire_ctx* ctx, const float* model, int32_t d1, int32_t d2, int32_t d3,
                    const float* projections, int32_t p1, int32_t p2, int32_t num_projs,
                    const double* angles, double* Rtotal, double* Rarr, float* simuProjs);

/* The same two calls with the reference's own interpolant, interp3(...,'cubic',0)
 * (src/calculate3Dprojection_RealSpaceinterp.m:45,51): cubic convolution with
 * Keys' a = -1/2 kernel and Keys' boundary nodes f(-1) = 3f(0) - 3f(1) + f(2)
 * (MATLAB's 'cubic' on a uniform grid), extrapolation value 0 outside the node
 * box.  Every model extent must be >= 4. */
int genfire_project_cubic(genfire_ctx* ctx, const float* vol, int32_t d1, int32_t d2, int32_t d3,
                          double phi, double theta, double psi, float* proj);
int genfire_rfactor_cubic(genfire_ctx* ctx, const float* model, int32_t d1, int32_t d2, int32_t d3,
                          const float* projections, int32_t p1, int32_t p2, int32_t num_projs,
                          const double* angles, double* Rtotal, double* Rarr, float* simuProjs);

/* genfire_rfactor on device-resident buffers, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc).  d_model (d1 x d2 x d3
 * float), d_projections (p1 x p2 x P float) and d_simuProjs (p1 x p2 x P float,
 * or NULL) are device pointers on the context's GPU; angles (P x 3 double),
 * Rtotal and Rarr[P] are host memory.  The projector's matrices, partial sums
 * and model forms live in one static device array per GPU of 64 MiB, held until
 * the process ends: a call needs about 8 (d1+1)(d2+1) d3 bytes per model form
 * (one or two forms; 4 d1 d2 d3 for a tilt series about y with d2 % 4 == 0)
 * plus 8 KiB per 256 x 256 projection, and returns GENFIRE_ERR_UNSUPPORTED above
 * 64 MiB or above 1024 projections.  Calls share that array and its last
 * contents, so the library serialises them with a process-wide lock; the work
 * runs on the null stream and the call returns when it is done. */
int genfire_rfactor_noalloc(genfire_ctx* ctx, const float* d_model, int32_t d1, int32_t d2, int32_t d3,
                            const float* d_projections, int32_t p1, int32_t p2, int32_t num_projs,
                            const double* angles, double* Rtotal, double* Rarr, float* d_simuProjs);

/* calculate3Dprojection_RealSpaceinterp_general(Vol, phi, theta, psi, vector1,
 * vector2, vector3) (src/calculate3Dprojection_RealSpaceinterp_general.m):
 * R = Q(vector1, phi) Q(vector2, theta) Q(vector3, psi) (MatrixQuaternionRot),
 * trilinear as genfire_project. */
int genfire_project_general(genfire_ctx* ctx, const float* vol, int32_t d1, int32_t d2, int32_t d3,
                            double phi, double theta, double psi, const double vector1[3],
                            const double vector2[3], const double vector3[3], float* proj);

/* genfire_rfactor with the general-axis projector above (Tian's post-hoc R for
 * tilt series recorded with the gridding's own rotation axes). */
int genfire_rfactor_general(genfire_ctx* ctx, const float* model, int32_t d1, int32_t d2, int32_t d3,
                            const float* projections, int32_t p1, int32_t p2, int32_t num_projs,
                            const double* angles, const double vector1[3], const double vector2[3],
                            const double vector3[3], double* Rtotal, double* Rarr, float* simuProjs);

/* The two general-axis calls with the reference's own interp3 'cubic'
 * (calculate3Dprojection_RealSpaceinterp_general.m:37,43), as
 * genfire_project_cubic. */
int genfire_project_general_cubic(genfire_ctx* ctx, const float* vol, int32_t d1, int32_t d2, int32_t d3,
                                  double phi, double theta, double psi, const double vector1[3],
                                  const double vector2[3], const double vector3[3], float* proj);
int genfire_rfactor_general_cubic(genfire_ctx* ctx, const float* model, int32_t d1, int32_t d2, int32_t d3,
                                  const float* projections, int32_t p1, int32_t p2, int32_t num_projs,
                                  const double* angles, const double vector1[3], const double vector2[3],
                                  const double vector3[3], double* Rtotal, double* Rarr, float* simuProjs);

/* fftn (inverse = 0) or ifftn (inverse = 1, scaled by 1/N) of an n1 x n2 x n3
 * interleaved-complex array in place, with this library's 3D FFT (MATLAB's
 * fftn/ifftn at reconstruct.m:133,165 and src/my_ifft.m:11). */
int genfire_fftn(genfire_ctx* ctx, float* data, int32_t n1, int32_t n2, int32_t n3,
                 int32_t inverse);

/* ------------------------------------------------------------------------
 * Request server.  No counterpart in the reference, which reconstructs one
 * object per MATLAB session (Genfire_test.m:41-44); this is the serving form of
 * the same job for callers that submit many reconstructions: requests are
 * admitted while a slot and their HBM footprint are free, prepared (upload,
 * gridding, set-up) on worker threads, and iterated by one scheduler thread
 * that launches the transform passes of all iterating jobs of one grid shape
 * together.  Each result equals the one genfire_job_* gives for the same
 * inputs; errR and the R-factor records follow genfire_job_get_errR.
 * ------------------------------------------------------------------------ */
typedef struct genfire_server genfire_server;

typedef struct {
    int64_t hbm_budget_bytes;   /* device bytes for resident jobs; <= 0: 90% of the free memory */
    int32_t max_active;         /* jobs resident at once (slots), 1..64                      */
    int32_t max_batch;          /* jobs per batched launch, 1..8                              */
    int32_t lookahead;          /* batched iterations queued ahead of the GPU per shape, 1..16 */
    int32_t prep_threads;       /* worker threads preparing jobs, 1..8                        */
} genfire_server_config;

#define GENFIRE_REQ_QUEUED 0
#define GENFIRE_REQ_PREPARING 1
#define GENFIRE_REQ_ITERATING 2
#define GENFIRE_REQ_FINISHING 3
#define GENFIRE_REQ_DONE 4
#define GENFIRE_REQ_FAILED 5

typedef struct {
    int64_t submitted, completed, failed, queued, active;
    int64_t batched_launches;      /* scheduler iterations launched (all groups)       */
    int64_t batched_jobs;          /* job-iterations they carried                      */
    int64_t peak_footprint_bytes;  /* largest sum of admitted footprints               */
    int64_t graph_replays;         /* scheduler iterations that replayed a captured HIP graph of
                                      the group's steady-state iteration (GENFIRE_GRAPHS=0: none) */
    int64_t hbm_budget_bytes;
    int32_t max_active_seen;
    int32_t pad_;
    double p50_ms, p99_ms, mean_ms; /* submission to result, completed requests        */
    double jobs_per_hour;           /* completed / (last completion - first submission) */
} genfire_server_stats;

int genfire_server_config_default(genfire_server_config* cfg);
int genfire_server_create(genfire_ctx* ctx, const genfire_server_config* cfg, genfire_server** out);
/* waits for every submitted request, then stops the threads */
int genfire_server_destroy(genfire_server* srv);
/* copies the inputs (as genfire_job_set_inputs; support may be NULL) and queues
 * the request; fails at once on invalid parameters or a footprint above the budget */
int genfire_server_submit(genfire_server* srv, const genfire_params* params, int32_t dim1, int32_t dim2,
                          int32_t num_projs, const float* projections, const double* angles,
                          const uint8_t* support, int64_t* request_id);
int genfire_server_status(genfire_server* srv, int64_t request_id, int32_t* state);
int genfire_server_wait(genfire_server* srv, int64_t request_id);
int genfire_server_drain(genfire_server* srv);
/* reconstruction Dim1 x Dim2 x Dim1, errK[numIterations], errR[numIterations /
 * monitorR_loopLength] when monitor_R; latency_ms = {queued, preparing,
 * iterating and read-back, submission to result}; any output may be NULL */
int genfire_server_result(genfire_server* srv, int64_t request_id, float* reconstruction, float* errK,
                          float* errR, double latency_ms[4]);
/* forgets a finished request */
int genfire_server_release(genfire_server* srv, int64_t request_id);
int genfire_server_get_stats(genfire_server* srv, genfire_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* GENFIRE_H */
