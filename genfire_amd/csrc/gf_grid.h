// gf_grid.h -- internal launch interface of the gridding, FFT, iteration and
// projector kernels.  Every function enqueues on the given stream and returns.
#pragma once

#include "gf_common.h"

namespace gf {

// ---- index pass (gf_grid_index.hip) ----
int index_threads();
void grid_index_count(const PlaneConst* d_planes, int P, const HalfGrid& hg, double cutoff, int n1,
                      int n2, int nblk, uint32_t* d_blk_counts, hipStream_t st);
// exclusive scan of n counts into offsets (int64), *total the sum; d_scratch
// holds grid_index_scan_scratch(n) bytes
size_t grid_index_scan_scratch(int64_t n);
void grid_index_scan(const uint32_t* d_counts, int64_t n, int64_t* d_offsets, int64_t* d_total,
                     int64_t* d_scratch, hipStream_t st);
void grid_index_emit(const PlaneConst* d_planes, int P, const HalfGrid& hg, double cutoff, int n1,
                     int n2, int nblk, const int64_t* d_blk_offsets, uint32_t* ind, float* dist,
                     float* un, float* vn, hipStream_t st);

// ---- DFT values, scatter, finalize (gf_grid_values.hip) ----
// Tile of the value kernels: points [begin, begin+count) of projection p.
struct PointTile {
    int64_t begin;
    int32_t count;
    int32_t proj;
};
// The tiles of every projection: proj_off[P + 1] point offsets, tile_off[P + 1]
// the first tile of each projection (tile_off[P] = ntiles).
void grid_make_tiles(const int64_t* d_proj_off, const int64_t* d_tile_off, int P, int tile_pts, PointTile* d_tiles,
                     int ntiles, hipStream_t st);
// FP32 direct DFT of every point (fillInFourierGrid_DFT.m:236-248).
void grid_values_fp32(const float* d_proj, int dim1, int dim2, const float* d_un, const float* d_vn,
                      const PointTile* d_tiles, int ntiles, float2* d_F, hipStream_t st);
// MFMA form: the projections pre-transposed to row-major [a][b] (zero-padded),
// then 64-point tiles (mfma_tile_points) through v_mfma_f32_16x16x4f32.
int mfma_tile_points();
// The same values as a type-2 non-uniform FFT (gf_grid_nufft.hip): fine grid
// length M >= 2 Dim (0 when no supported FFT length fits), kernel shape beta,
// the per-mode deconvolution 1/phi^, the padded-grid load and the w x w
// interpolation over the same 64-point tiles.
int nufft_kernel_width();
int nufft_fine_len(int dim);
double nufft_beta(int dim, int M);
std::vector<float> nufft_deconv(int dim, int M);
void nufft_load(const float* d_proj, int d1, int d2, int P, int M1, int M2, const float* d_inv1,
                const float* d_inv2, float2* d_fine, hipStream_t st);
void nufft_interp(const float2* d_G, int M1, int M2, double beta1, double beta2, const float* d_un,
                  const float* d_vn, const PointTile* d_tiles, int ntiles, float2* d_F, hipStream_t st);
size_t grid_prep_floats(int dim1, int dim2, int P);
void grid_prep_proj(const float* d_proj, int dim1, int dim2, int P, float* d_Pt, hipStream_t st);
void grid_values_mfma_fp32(const float* d_Pt, int dim1, int dim2, const float* d_un, const float* d_vn,
                           const PointTile* d_tiles, int ntiles, float2* d_F, hipStream_t st);
// Reduced-precision forms (GENFIRE_PREC_SPLIT_BF16, GENFIRE_PREC_FP16) on the same
// 64-point tiles: projections split once per job into 16-bit operand arrays
// (hi, and lo for split-BF16) of grid_prep_lp_elems elements each, plus P
// per-projection scales (FP16: max|P|).
size_t grid_prep_lp_elems(int dim1, int dim2, int P);
void grid_prep_lp(int mode, const float* d_proj, int dim1, int dim2, int P, unsigned short* d_hi,
                  unsigned short* d_lo, float* d_scale, hipStream_t st);
void grid_values_mfma_lp(int mode, const unsigned short* d_hi, const unsigned short* d_lo, const float* d_scale,
                         int dim1, int dim2, const float* d_un, const float* d_vn, const PointTile* d_tiles,
                         int ntiles, float2* d_F, hipStream_t st);
// Weighted-average accumulation of one projection's points (:267-291), in order.
void grid_scatter(const uint32_t* d_ind, const float* d_dist, const float2* d_F, int64_t begin,
                  int64_t count, double2* d_acc, double* d_sw, uint32_t* d_hits, float2* d_exact,
                  hipStream_t st);
// FS -> measuredK in FFT (ifftshifted) order, with CentroSymmetricity's mirror
// rule when centro != 0, and the single cast of runGridding.
void grid_finalize(const HalfGrid& hg, int n1, int n2, int centro, const double2* d_acc,
                   const double* d_sw, const uint32_t* d_hits, const float2* d_exact, float2* d_M,
                   hipStream_t st);
constexpr uint32_t kExactBit = 0x80000000u;

// ---- FFT gridding (gf_grid_fft.hip), fillInFourierGrid.m ----
// One projection's rotation as used by the scatter: rot = the first two columns
// of the projector matrix R (rotkCoords = R'_fg * [kx; ky; 0] = kx R(:,1) +
// ky R(:,2), single); nrm = R(:,3), the slab normal; axis = its largest
// component.
struct FftPlane {
    float rot[6];
    float nrm[3];
    int axis;
};
// Padded, ifftshifted projections as complex d x d x P (FFT input).
void fftgrid_load(const float* d_proj, int dim1, int dim2, int P, int d, float2* d_F, hipStream_t st);
// Accumulates one projection (d_Fp: its d x d spectrum in FFT order) into the
// centred d^3 sums: acc += w v, sw += w, hits += 1 per matched lattice point.
void fftgrid_scatter(const FftPlane& pl, int d, int shift_max, float rho, double cutoff, const float2* d_Fp,
                     double2* d_acc, double* d_sw, uint32_t* d_hits, hipStream_t st);
// Weighted average, hermitianSymmetrize, single cast; measuredK in FFT order.
void fftgrid_finalize(int d, const double2* d_acc, const double* d_sw, float2* d_M, hipStream_t st);

// ---- layout helpers (gf_grid_values.hip) ----
// Centred <-> FFT-order conversion of an n1 x n2 x n3 complex grid
// (to_centred = 1: out = fftshift(in), 0: out = ifftshift(in)).
// out[t] = half(h[t]) * scale[t / per_proj] for P projections of per_proj values
void expand_f16(const uint16_t* d_h, const float* d_scale, int64_t per_proj, int P, float* d_out, hipStream_t st);
void shift_complex(const float2* d_in, float2* d_out, int n1, int n2, int n3, int to_centred,
                   hipStream_t st);
// Real crop of the centred d1 x d2 x d3 box out of an FFT-order n1 x n2 x n3
// grid (My_stripzero(fftshift(x), ...)); src_complex selects float2 input (real part).
void crop_centre(const void* d_src, int src_complex, int n1, int n2, int n3, float* d_dst, int d1,
                 int d2, int d3, hipStream_t st);
// Inverse of crop_centre for real data: zero-pad d1 x d2 x d3 into n1 x n2 x n3
// FFT order (ifftshift(My_paddzero(v, n))).
void pad_centre(const float* d_src, int d1, int d2, int d3, float* d_dst, int n1, int n2, int n3,
                hipStream_t st);
void pad_centre_u8(const uint8_t* d_src, int d1, int d2, int d3, uint8_t* d_dst, int n1, int n2,
                   int n3, hipStream_t st);

// ---- 3D FFT (gf_fft.hip) ----
bool fft_supported(int n);
// exp(-2 pi i m / n), m < n, FP32 from FP64, cached per device.
const float2* fft_twiddles(int n);
// Complex-to-complex pass along one axis of an n1 x n2 x n3 grid, in place.
// inverse passes scale by 1/n_axis.
void fft_axis(float2* d, int n1, int n2, int n3, int axis, int inverse, hipStream_t st);
// the same pass reading src and writing d (src and d must not overlap)
void fft_axis_from(const float2* src, float2* d, int n1, int n2, int n3, int axis, int inverse, hipStream_t st);
// In place along `axis`, visiting only the y rows (yb > 0) and z rows (zb > 0)
// of the centred box in FFT order; the other rows are left as they are.
void fft_axis_rows(float2* d, int n1, int n2, int n3, int axis, int inverse, int yb, int zb, hipStream_t st);
void fft3(float2* d, int n1, int n2, int n3, int inverse, hipStream_t st);

// ---- iteration passes (gf_fft.hip) ----
// The measured voxels (measuredK ~= 0 & measuredK_mask) are kept per z-pencil
// (x, y) in a compressed list: csr_off[x + n1*y] .. csr_off[x + n1*y + 1] index
// csr_meta (z | x << 11 | enforced << 31) and csr_val (measuredK at the voxel).
struct IterArgs {
    float* x;                 // iterate, real, FFT order
    float2* K;                // spectrum workspace
    const uint8_t* S;         // padded support, FFT order (0/1)
    const int64_t* csr_off;   // n1*n2 + 1 offsets
    const uint32_t* csr_meta; // per measured voxel
    const float2* csr_val;    // per measured voxel
    float* best;              // P_R(x) output (nullable)
    float* pf_out;            // P_F output copy (nullable)
    double* err_part;         // per-wave partial errK numerators of the fused pass
    int n1, n2, n3;
    int positivity, support, dr;
};
// Pass 1: P_R (or the DR reflection) fused into the axis-0 forward FFT.
void iter_first(const IterArgs& a, hipStream_t st);
// Axis-2 forward FFT, errK partials, replacement, axis-2 inverse FFT.
int64_t iter_fused_partials(const IterArgs& a);
void iter_fused(const IterArgs& a, hipStream_t st);
// Axis-0 inverse FFT fused with real part and the GENFIRE / DR update.
void iter_last(const IterArgs& a, hipStream_t st);
// errK[it] = sum(partials) / den.
void iter_errk_reduce(const double* d_part, int64_t n, double den, float* d_errK_it, hipStream_t st);
// errK[*ctr] = sum(part) / den, then ++*ctr (device-side iteration index)
void iter_errk_reduce_next(const double* d_part, int64_t n, double den, float* d_errK, int* d_ctr, hipStream_t st);
// measured-voxel lists: per-pencil counts, fill (after an exclusive scan of the
// counts), enforcement bits for a schedule stage ((1-Q) > delay, reconstruct.m:42,120)
// and the errK denominator sum |M(errInd)|.
void csr_count(const float2* d_M, const uint8_t* d_mask, int n1, int n2, int n3, uint32_t* d_cnt, hipStream_t st);
void csr_fill(const float2* d_M, const uint8_t* d_mask, int n1, int n2, int n3, const int64_t* d_off,
              uint32_t* d_meta, float2* d_val, hipStream_t st);
void csr_enforce(int n1, int n2, int n3, const int64_t* d_off, uint32_t* d_meta, double delay, hipStream_t st);
void csr_den(const float2* d_val, int64_t n, double* d_tmp, int ntmp, double* d_out, hipStream_t st);

// ---- half-spectrum iteration (gf_iter_half.hip) ----
// Spectrum of the real iterate for kx in [0, n1/2] in two layouts:
// A[y + n2*(kx + h*z)] and B[z + n3*(ky + n2*kx)], h = n1/2 + 1.  The measured
// voxels of each B pencil (ky + n2*kx) carry meta = z | ky << 11 | flags, the
// replacement value and |M| of the voxel and of its unstored mate.
struct HalfIterArgs {
    float* x;
    float2* A;
    float2* B;
    const uint8_t* S;
    const int64_t* csr_off;
    const uint32_t* csr_meta;
    const float2* csr_val;
    const float2* csr_mag;
    float* best;
    float* pf_out;
    double* err_part;
    // Rfree (withheld voxels, per half pencil: meta z | ky<<11, value as stored
    // or conjugated for a mate-stored voxel); null when Rfree is off
    const int64_t* rf_off;
    const uint32_t* rf_meta;
    const float2* rf_val;
    float* rf_terms;   // |k - value| per entry, written by the z pass
    int n1, n2, n3;
    int positivity, support, dr;
    // Real-space box (x, y, z extents, centred in FFT order) outside which the
    // iterate's transform input is zero; the passes skip rows that lie outside
    // it.  Equal to (n1, n2, n3) when nothing is pruned.
    int bx, by, bz;
    // DR on the box (dmode): the replacement delta D of each list entry (then
    // its Hermitian part on the kx = 0 and n1/2 planes), and for the entries of
    // those two planes the entry index of their in-plane mate
    int dmode;
    float2* csr_d;
    const int64_t* csr_mate;
    // errK of the iteration, in two fixed-order stages: on the y-inverse launch
    // (err_stage 1) workgroup c < kErrChunks (or c mod the grid) sums chunk c of
    // the z pass's err_n partials into err_chunk[c]; on the launch of the
    // iteration's last x pass (err_stage 2) workgroup 0 sets
    // errk[*errctr] = sum(err_chunk) / err_den and advances *errctr.  With dherm
    // the y-inverse launch's first workgroups also replace D on the kx = 0 and
    // n1/2 planes by its Hermitian part.  0 / null on every other launch.
    float* errk;
    int* errctr;
    double* err_chunk;
    double err_den;
    int64_t err_n;
    int err_stage;
    int dherm;
};
constexpr int kErrChunks = 64;
// Rfree[s + nbins * (*ctr)] = sum of terms of shell s (entries order[shell_off[s]
// .. shell_off[s+1]), fixed order) / den[s], or -1 for an empty shell; *ctr is
// the iteration index the errK reduction advances afterwards.
void rfree_reduce(const float* d_terms, const int32_t* d_order, const int32_t* d_shell_off, const double* d_den,
                  int nbins, float* d_Rfree, const int* d_ctr, hipStream_t st);
constexpr uint32_t kHalfTermA = 1u << 28;     // voxel measured: errK term |k| vs mag.x
constexpr uint32_t kHalfTermB = 1u << 29;     // mate measured: errK term |k| vs mag.y
constexpr uint32_t kHalfAverage = 1u << 30;   // only one of the pair measured: k = (V + k)/2
constexpr uint32_t kHalfEnforced = 1u << 31;  // replaced at the current schedule stage
constexpr uint32_t kHalfPassive = 1u << 27;   // kx = 0, n1/2 planes: unmeasured voxel whose mate is measured
bool half_supported(int n1, int n2, int n3);
int64_t half_partials(int n1, int n2, int n3);
// Up to kHalfBatchMax jobs of one grid shape in one launch (launch row
// blockIdx.y = job): the same-shape steps of a request server's jobs.
constexpr int kHalfBatchMax = 8;
struct HalfBatch {
    HalfIterArgs job[kHalfBatchMax];
    int count;
};
void half_first(const HalfBatch& b, hipStream_t st);
void half_y(const HalfBatch& b, int inverse, hipStream_t st);
void half_z(const HalfBatch& b, hipStream_t st);
void half_last(const HalfBatch& b, hipStream_t st);
void half_last_first(const HalfBatch& b, hipStream_t st);
// dmode jobs: D -> its Hermitian part on the kx = 0 and n1/2 planes (after half_z)
// dmode set-up: in-plane mate entry of every kx = 0 and n1/2 plane entry
void hcsr_mates(const int64_t* d_off, const uint32_t* d_meta, int n1, int n2, int n3, int64_t n_plane0,
                int64_t plane_m_begin, int64_t n_plane, int64_t* d_mate, hipStream_t st);
void half_first(const HalfIterArgs& a, hipStream_t st);
void half_y(const HalfIterArgs& a, int inverse, hipStream_t st);
void half_z(const HalfIterArgs& a, hipStream_t st);
void half_last(const HalfIterArgs& a, hipStream_t st);
void half_last_first(const HalfIterArgs& a, hipStream_t st);   // P5 of one iteration + P1 of the next
// The half-spectrum lists are built per (pencil, z chunk), hcsr_chunks() chunks:
// counts at (pencil * chunks + chunk), scanned into chunk offsets d_coff, the
// pencil offsets gathered from them (hcsr_pencil_off), the fill from d_coff.
int hcsr_chunks();
void hcsr_count(const float2* d_M, const uint8_t* d_mask, int n1, int n2, int n3, uint32_t* d_cnt, hipStream_t st);
void hcsr_pencil_off(const int64_t* d_coff, int64_t np, int64_t* d_off, hipStream_t st);
void hcsr_fill(const float2* d_M, const uint8_t* d_mask, int n1, int n2, int n3, const int64_t* d_coff,
               uint32_t* d_meta, float2* d_val, float2* d_mag, hipStream_t st);
void hcsr_enforce(int n1, int n2, int n3, const int64_t* d_off, uint32_t* d_meta, double delay, hipStream_t st);
void hcsr_den(const float2* d_mag, const uint32_t* d_meta, int64_t n, double* d_tmp, int ntmp, double* d_out,
              hipStream_t st);

// ---- projector (gf_project.hip) ----
// Per projection the rotation is stored as kRStride floats: the column-major
// 3x3 ZYZ matrix; at [9], 1 when the y-fastest volume copy is to be used; at
// [10], 1 when the matrix maps the y axis exactly onto itself (then [9] and
// [11] are 1); at [11], 1 when a wave's lanes run along pixel j, 0 along i,
// 2 when a wave covers an 8 x 8 pixel patch.
constexpr int kRStride = 12;
struct ProjArgs {
    const float* vol_y; // y-fastest model (index y + d2*(x + d1*z)) for the four-ray form, else nullable
    const float* zy;    // z-duplicated y-fastest model (crop_zdup / volume_zdup), nullable when unused
    const float* zx;    // z-duplicated x-fastest model, nullable when unused
    const float* vcubic; // padded form of the cubic-convolution projector (crop_cubic_form /
                         // volume_cubic_form); when set every projection reads it
    int four_ray;       // every projection is an exact tilt about y and d2 % 4 == 0: read vol_y only
    int d1, d2, d3;
    const float* R;    // P x kRStride, single
    int P;
    float* out;        // d1 x d2 x P back projections (nullable)
    const float* meas; // measured projections m1 x m2 x P (nullable)
    int m1, m2;        // measured size (crop of the back projection centre)
    double* sums;      // P x 4 x nblk partials: sum|b-m|, sum (b-m)^2, sum|m|, sum m^2
    int post_hoc;      // 1: sum ||b|-|m|| instead of |b-m| (Tian_calc_Rfactor_realspace)
    float* acc;        // m1 x m2 x P ray sums between the passes of a slab sweep (may alias out);
                       // nullable: one pass (project_slabs)
};
// The model forms a set of projector matrices (host copy, P x kRStride) reads.
struct ProjForms {
    bool four_ray = false;   // vol_y only
    bool zy = false;         // the z-duplicated y-fastest form
    bool zx = false;         // the z-duplicated x-fastest form
    bool cubic = false;      // the cubic-convolution projector's padded form (then no other form)
};
// cubic: the projections interpolate with interp3's 'cubic' (cubic convolution)
ProjForms project_forms(const float* Rm, int P, int d2, bool cubic = false);
// Cubic convolution (Keys, a = -1/2) with Keys' boundary nodes: the model padded
// by one node on every side, (d1+2) x (d2+2) x (d3+2), y fastest, node (x, y, z)
// at (y + 1) + (d2 + 2) ((x + 1) + (d1 + 2) (z + 1)).  Every extent must be >= 4.
int64_t cubic_form_floats(int d1, int d2, int d3);
void volume_cubic_form(const float* d_vol, int d1, int d2, int d3, float* d_dst, hipStream_t st);
void crop_cubic_form(const float* d_src, int n1, int n2, int n3, float* d_dst, int d1, int d2, int d3,
                     hipStream_t st);
int project_blocks(int m1, int m2);
// Launches of the trilinear projector over a d1 x d2 x d3 model: 1, or the slab
// count of a slab sweep when the z-duplicated form passes 128 MiB and
// ProjArgs::acc is set.
int project_slabs(int d1, int d2, int d3);
// sets R[9] (y-fastest forms) and R[11] (lanes along pixel j) of a projection
void project_orientation(float* R);
bool project_y_exact(const float* R);
void volume_yfast(const float* d_src, float* d_dst, int d1, int d2, int d3, hipStream_t st);
// floats of one z-duplicated form of a d1 x d2 x d3 model
int64_t zdup_floats(int d1, int d2, int d3);
// the z-duplicated form of a plain column-major volume, or of the centred
// d1 x d2 x d3 crop of an FFT-order n1 x n2 x n3 volume (fftshift then crop)
void volume_zdup(const float* d_vol, int d1, int d2, int d3, float* d_dst, bool yfast, hipStream_t st);
void crop_zdup(const float* d_src, int n1, int n2, int n3, float* d_dst, int d1, int d2, int d3, bool yfast,
               hipStream_t st);
void project(const ProjArgs& a, hipStream_t st);
void project_reduce(const double* d_sums, int P, int nblk, double* d_out4, hipStream_t st);

}  // namespace gf
