/*
 * genfire_oracle.c -- CPU restatement of the GENFIRE v2 DFT-gridding index pass,
 * the FP64 direct DFT and the real-space projector.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * genfire_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  Nothing in the product links or calls it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  -ffp-contract=off is
 * part of the definition: every FP32 expression below is evaluated one IEEE
 * operation at a time, in the order written, with no fused multiply-add, which is
 * how MATLAB evaluates the element-wise array expressions it restates.
 *
 * Definitions that MATLAB leaves to closed libraries (parity unpinned at 1-ulp
 * ties, see each function): norm() of a single 3-vector, mldivide (LU) rounding,
 * the double->single rounding of the rotation product.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------- */
/* Rotation matrices                                                          */
/* ------------------------------------------------------------------------- */

/* MatrixQuaternionRot, @GENFIRE_Reconstructor/fillInFourierGrid_DFT.m:353-363
 * (the local copy, which shadows src/MatrixQuaternionRot.m inside the gridding).
 * Active rotation by theta degrees about `vec`; column-major 3x3 output. */
static void quat_rot(const double vec[3], double theta_deg, double M[9])
{
    double th = theta_deg * M_PI / 180.0;                     /* :355 */
    double nrm = sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]); /* :356 */
    double v0 = vec[0] / nrm, v1 = vec[1] / nrm, v2 = vec[2] / nrm;
    double w = cos(th / 2.0);                                 /* :357 */
    double s = sin(th / 2.0);
    double x = -s * v0, y = -s * v1, z = -s * v2;
    /* :358-360, row by row; M[r + 3c] */
    M[0 + 3 * 0] = 1.0 - 2.0 * (y * y) - 2.0 * (z * z);
    M[0 + 3 * 1] = 2.0 * x * y + 2.0 * w * z;
    M[0 + 3 * 2] = 2.0 * x * z - 2.0 * w * y;
    M[1 + 3 * 0] = 2.0 * x * y - 2.0 * w * z;
    M[1 + 3 * 1] = 1.0 - 2.0 * (x * x) - 2.0 * (z * z);
    M[1 + 3 * 2] = 2.0 * y * z + 2.0 * w * x;
    M[2 + 3 * 0] = 2.0 * x * z + 2.0 * w * y;
    M[2 + 3 * 1] = 2.0 * y * z - 2.0 * w * x;
    M[2 + 3 * 2] = 1.0 - 2.0 * (x * x) - 2.0 * (y * y);
}

static void matmul3(const double A[9], const double B[9], double C[9])
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[i + 3 * j] = (A[i + 0] * B[0 + 3 * j] + A[i + 3] * B[1 + 3 * j]) + A[i + 6] * B[2 + 3 * j];
}

/* fillInFourierGrid_DFT.m:108-118.  R = Q(v1,phi)*Q(v2,theta)*Q(v3,psi) in FP64,
 * stored as FP32 (rotMATs is 'single', :101).  The normal is R(:,3) (:116-117,
 * the product with [0 0 1]' selects the third column exactly). */
void go_rotation(const double v1[3], const double v2[3], const double v3[3],
                 double phi, double theta, double psi, float Rf[9], double Rd[9])
{
    double A[9], B[9], C[9], AB[9], R[9];
    quat_rot(v1, phi, A);
    quat_rot(v2, theta, B);
    quat_rot(v3, psi, C);
    matmul3(A, B, AB);
    matmul3(AB, C, R);
    for (int i = 0; i < 9; ++i) {
        if (Rf) Rf[i] = (float)R[i];
        if (Rd) Rd[i] = R[i];
    }
}

/* ------------------------------------------------------------------------- */
/* Index pass (bit-exact FP32 contract)                                       */
/* ------------------------------------------------------------------------- */

typedef struct {
    float nv[3];   /* normVECs(p,:), unnormalised, single                     */
    float nh[3];   /* normvec/norm(normvec) in distancePointsPlane_YY (:332)  */
    float nn;      /* sum(n.*n) in closestpoint (:347)                         */
    float lu[9];   /* LU factors of rotMATs(:,:,p), column-major               */
    int piv[3];    /* row interchanges (LAPACK ipiv, 0-based)                  */
} plane_t;

/* norm() of a single 3-vector (:332).  Defined here as the correctly rounded
 * value float(sqrt(sum of squares in double)); MATLAB's BLAS/MKL routine is
 * closed, so its last-bit behaviour is parity unpinned. */
static float norm3f(const float v[3])
{
    double s = (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
    return (float)sqrt(s);
}

/* CP_plane = R \ CP (:182).  MATLAB's mldivide on a general square single matrix
 * is LAPACK sgesv.  Restated from the reference LAPACK algorithm (sgetrf2 on 3x3
 * reduces to: pivot = first max |a|, multipliers scaled by the FP32 reciprocal of
 * the pivot, Schur update a - l*u), no FMA.  MKL's rounding is unpinned. */
static void lu3(const float R[9], float A[9], int piv[3])
{
    memcpy(A, R, 9 * sizeof(float));
    for (int j = 0; j < 3; ++j) {
        int p = j;
        float best = fabsf(A[j + 3 * j]);
        for (int i = j + 1; i < 3; ++i)
            if (fabsf(A[i + 3 * j]) > best) { best = fabsf(A[i + 3 * j]); p = i; }
        piv[j] = p;
        if (A[p + 3 * j] != 0.0f) {
            if (p != j)
                for (int c = 0; c < 3; ++c) {
                    float t = A[j + 3 * c]; A[j + 3 * c] = A[p + 3 * c]; A[p + 3 * c] = t;
                }
            float r = 1.0f / A[j + 3 * j];
            for (int i = j + 1; i < 3; ++i) A[i + 3 * j] = A[i + 3 * j] * r;
        }
        for (int c = j + 1; c < 3; ++c) {
            float ujc = A[j + 3 * c];
            for (int i = j + 1; i < 3; ++i) A[i + 3 * c] = A[i + 3 * c] - A[i + 3 * j] * ujc;
        }
    }
}

/* sgetrs: row interchanges, unit-lower forward substitution, upper back
 * substitution (reference strsm loop order). */
static void lu3_solve(const float A[9], const int piv[3], float b[3])
{
    for (int j = 0; j < 3; ++j)
        if (piv[j] != j) { float t = b[j]; b[j] = b[piv[j]]; b[piv[j]] = t; }
    for (int k = 0; k < 3; ++k)
        if (b[k] != 0.0f)
            for (int i = k + 1; i < 3; ++i) b[i] = b[i] - b[k] * A[i + 3 * k];
    for (int k = 2; k >= 0; --k)
        if (b[k] != 0.0f) {
            b[k] = b[k] / A[k + 3 * k];
            for (int i = 0; i < k; ++i) b[i] = b[i] - b[k] * A[i + 3 * k];
        }
}

static void plane_setup(const float R[9], plane_t* pl)
{
    for (int i = 0; i < 3; ++i) pl->nv[i] = R[i + 6];
    float nrm = norm3f(pl->nv);
    for (int i = 0; i < 3; ++i) pl->nh[i] = pl->nv[i] / nrm;
    pl->nn = (pl->nv[0] * pl->nv[0] + pl->nv[1] * pl->nv[1]) + pl->nv[2] * pl->nv[2];
    lu3(R, pl->lu, pl->piv);
}

/* Exported for tests: the LU solve of one right-hand side. */
void go_lu_solve(const float R[9], const float b[3], float x[3])
{
    float A[9]; int piv[3];
    lu3(R, A, piv);
    x[0] = b[0]; x[1] = b[1]; x[2] = b[2];
    lu3_solve(A, piv, x);
}

/* Half-grid k ranges, fillInFourierGrid_DFT.m:122-136. */
void go_half_ranges(int n1, int n2, int centro, int lo[3], int len[3])
{
    if (centro) {
        lo[0] = -(n1 / 2); len[0] = n1 / 2 + 1;          /* k1 = -floor(n1/2):0           */
        lo[1] = -(n2 / 2); len[1] = 2 * (n2 / 2) + 1;    /* k2 = -floor(n2/2):floor(n2/2) */
        lo[2] = -(n1 / 2); len[2] = 2 * (n1 / 2) + 1;    /* k3 = -floor(n1/2):floor(n1/2) */
    } else {
        lo[0] = -(n1 / 2); len[0] = n1;                  /* -ceil((n-1)/2):floor((n-1)/2) */
        lo[1] = -(n2 / 2); len[1] = n2;
        lo[2] = lo[0];     len[2] = n1;
    }
}

/* One projection of fillInFourierGrid_DFT.m:175-192, scanning the WHOLE half grid
 * in MATLAB linear order (k1 fastest), exactly as the reference does:
 *   D    = |((K1*nh1 + K2*nh2) + K3*nh3)|           distancePointsPlane_YY :333
 *   Dind = find(D < cutoff)                         :178 (float vs double, exact)
 *   CP   = K + ((0 - K.n)/(n.n)) * n                closestpoint :347-348
 *   CP_plane = R \ CP                               :182
 *   keep |u| <= n1/2 & |v| <= n2/2                  :190-191
 * Outputs (0-based linear half-grid index, D, u = CP_plane(1), v = CP_plane(2)).
 * Returns the number of kept points (or -1 if cap is too small); *n_wrong counts
 * the |CP_plane(3)| > 1e-4 diagnostic of :185-187. */
int64_t go_select(int n1, int n2, int centro, const float R[9], double cutoff,
                  int64_t cap, int64_t* ind, float* D, float* u, float* v,
                  int64_t* n_wrong)
{
    plane_t pl;
    plane_setup(R, &pl);
    int lo[3], len[3];
    go_half_ranges(n1, n2, centro, lo, len);
    const float h1 = (float)n1 / 2.0f, h2 = (float)n2 / 2.0f;
    int64_t cnt = 0, wrong = 0;
    for (int i3 = 0; i3 < len[2]; ++i3) {
        const float K3 = (float)(lo[2] + i3);
        for (int i2 = 0; i2 < len[1]; ++i2) {
            const float K2 = (float)(lo[1] + i2);
            for (int i1 = 0; i1 < len[0]; ++i1) {
                const float K1 = (float)(lo[0] + i1);
                float t = (K1 * pl.nh[0] + K2 * pl.nh[1]) + K3 * pl.nh[2];
                float d = fabsf(t);
                if (!((double)d < cutoff)) continue;
                float s = (K1 * pl.nv[0] + K2 * pl.nv[1]) + K3 * pl.nv[2];
                float vv = (0.0f - s) / pl.nn;
                float cp[3];
                cp[0] = K1 + vv * pl.nv[0];
                cp[1] = K2 + vv * pl.nv[1];
                cp[2] = K3 + vv * pl.nv[2];
                lu3_solve(pl.lu, pl.piv, cp);
                if (fabsf(cp[2]) > 0.0001f) ++wrong;
                if (fabsf(cp[0]) <= h1 && fabsf(cp[1]) <= h2) {
                    if (cnt >= cap) return -1;
                    ind[cnt] = (int64_t)i1 + (int64_t)len[0] * ((int64_t)i2 + (int64_t)len[1] * i3);
                    D[cnt] = d; u[cnt] = cp[0]; v[cnt] = cp[1];
                    ++cnt;
                }
            }
        }
    }
    if (n_wrong) *n_wrong = wrong;
    return cnt;
}

/* Near-tie census for one projection: counts grid points whose FP32 D lies within
 * `ulps` ulps of the cutoff, and kept/rejected points whose |u| or |v| lies
 * within `ulps` ulps of n/2.  Reports where an unpinned 1-ulp difference in
 * norm() or mldivide could flip an index decision. */
void go_tie_census(int n1, int n2, int centro, const float R[9], double cutoff,
                   int ulps, int64_t out[4])
{
    plane_t pl;
    plane_setup(R, &pl);
    int lo[3], len[3];
    go_half_ranges(n1, n2, centro, lo, len);
    const float h1 = (float)n1 / 2.0f, h2 = (float)n2 / 2.0f;
    const float c = (float)cutoff;
    float dlo = c, dhi = c, ulo = h1, uhi = h1, vlo = h2, vhi = h2;
    for (int k = 0; k < ulps; ++k) {
        dlo = nextafterf(dlo, 0.0f); dhi = nextafterf(dhi, 1e30f);
        ulo = nextafterf(ulo, 0.0f); uhi = nextafterf(uhi, 1e30f);
        vlo = nextafterf(vlo, 0.0f); vhi = nextafterf(vhi, 1e30f);
    }
    int64_t nd = 0, nuv = 0, nkept = 0, nexact = 0;
#pragma omp parallel for schedule(static) reduction(+ : nd, nuv, nkept, nexact)
    for (int i3 = 0; i3 < len[2]; ++i3)
        for (int i2 = 0; i2 < len[1]; ++i2)
            for (int i1 = 0; i1 < len[0]; ++i1) {
                const float K1 = (float)(lo[0] + i1), K2 = (float)(lo[1] + i2), K3 = (float)(lo[2] + i3);
                float d = fabsf((K1 * pl.nh[0] + K2 * pl.nh[1]) + K3 * pl.nh[2]);
                if (d >= dlo && d <= dhi) ++nd;
                if (!((double)d < cutoff)) continue;
                float s = (K1 * pl.nv[0] + K2 * pl.nv[1]) + K3 * pl.nv[2];
                float vv = (0.0f - s) / pl.nn;
                float cp[3] = {K1 + vv * pl.nv[0], K2 + vv * pl.nv[1], K3 + vv * pl.nv[2]};
                lu3_solve(pl.lu, pl.piv, cp);
                float au = fabsf(cp[0]), av = fabsf(cp[1]);
                if ((au >= ulo && au <= uhi) || (av >= vlo && av <= vhi)) ++nuv;
                if (au <= h1 && av <= h2) ++nkept;
                if (au == h1 || av == h2) ++nexact;
            }
    out[0] = nd; out[1] = nuv; out[2] = nkept; out[3] = nexact;
}

/* ------------------------------------------------------------------------- */
/* FP64 direct DFT (the value oracle for fillInFourierGrid_DFT.m:236-248)      */
/* ------------------------------------------------------------------------- */

/* F_j = sum_{a,b} P[a,b] exp(-2 pi i (k1o[a] un_j + k2o[b] vn_j)),
 * k1o = -ceil((Dim1-1)/2):floor((Dim1-1)/2) (:147-148), un_j = u_j/n1 and
 * vn_j = v_j/n2 as the FP32 values the reference uploads (:237-238).
 * Evaluated in FP64 with exact phase reduction: k*un is exact in FP64 (a <=12-bit
 * integer times a 24-bit mantissa), its fractional part is exact, and only the
 * final sin/cos of 2*pi*frac rounds.  The sum factors exactly as
 * sum_b e2[b] * (sum_a P[a,b] e1[a]); F is interleaved (re, im). */
void go_dft(int dim1, int dim2, const float* proj, int64_t m,
            const float* un, const float* vn, double* F)
{
    const int c1 = dim1 / 2, c2 = dim2 / 2;   /* ceil((Dim-1)/2) == floor(Dim/2) */
#pragma omp parallel
    {
        double* e1 = (double*)malloc(sizeof(double) * 2 * dim1);
        double* e2 = (double*)malloc(sizeof(double) * 2 * dim2);
#pragma omp for schedule(dynamic, 16)
        for (int64_t j = 0; j < m; ++j) {
            for (int a = 0; a < dim1; ++a) {
                double t = (double)(a - c1) * (double)un[j];
                double r = t - nearbyint(t);
                e1[2 * a] = cos(2.0 * M_PI * r);
                e1[2 * a + 1] = -sin(2.0 * M_PI * r);
            }
            for (int b = 0; b < dim2; ++b) {
                double t = (double)(b - c2) * (double)vn[j];
                double r = t - nearbyint(t);
                e2[2 * b] = cos(2.0 * M_PI * r);
                e2[2 * b + 1] = -sin(2.0 * M_PI * r);
            }
            double fr = 0.0, fi = 0.0;
            for (int b = 0; b < dim2; ++b) {
                const float* col = proj + (int64_t)dim1 * b;
                double gr = 0.0, gi = 0.0;
                for (int a = 0; a < dim1; ++a) {
                    gr += (double)col[a] * e1[2 * a];
                    gi += (double)col[a] * e1[2 * a + 1];
                }
                fr += gr * e2[2 * b] - gi * e2[2 * b + 1];
                fi += gr * e2[2 * b + 1] + gi * e2[2 * b];
            }
            F[2 * j] = fr;
            F[2 * j + 1] = fi;
        }
        free(e1);
        free(e2);
    }
}

/* Same sum without the separable factorisation: one exp per (a,b) term of the
 * total phase k1o*un + k2o*vn (the literal form of :243).  Brute force for tests. */
void go_dft_brute(int dim1, int dim2, const float* proj, int64_t m,
                  const float* un, const float* vn, double* F)
{
    const int c1 = dim1 / 2, c2 = dim2 / 2;
    for (int64_t j = 0; j < m; ++j) {
        double fr = 0.0, fi = 0.0;
        for (int b = 0; b < dim2; ++b)
            for (int a = 0; a < dim1; ++a) {
                double ph = (double)(a - c1) * (double)un[j] + (double)(b - c2) * (double)vn[j];
                double r = ph - nearbyint(ph);
                double p = (double)proj[a + (int64_t)dim1 * b];
                fr += p * cos(2.0 * M_PI * r);
                fi -= p * sin(2.0 * M_PI * r);
            }
        F[2 * j] = fr;
        F[2 * j + 1] = fi;
    }
}

/* ------------------------------------------------------------------------- */
/* Real-space projector, src/calculate3Dprojection_RealSpaceinterp.m           */
/* ------------------------------------------------------------------------- */

/* MATLAB sind/cosd: exact at multiples of 90 degrees. */
static void sincosd(double x, double* s, double* c)
{
    double r = fmod(x, 360.0);
    double q = nearbyint(r / 90.0);
    double rem = r - 90.0 * q;
    double sr = sin(rem * M_PI / 180.0), cr = cos(rem * M_PI / 180.0);
    int qi = ((int)q % 4 + 4) % 4;
    switch (qi) {
        case 0: *s = sr; *c = cr; break;
        case 1: *s = cr; *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}

/* ZYZ matrix of calculate3Dprojection_RealSpaceinterp.m:13-15 (column-major). */
void go_zyz(double phi, double theta, double psi, double R[9])
{
    double sp, cp, st, ct, ss, cs;
    sincosd(phi, &sp, &cp);
    sincosd(theta, &st, &ct);
    sincosd(psi, &ss, &cs);
    R[0 + 3 * 0] = cs * ct * cp - ss * sp;
    R[0 + 3 * 1] = -ss * ct * cp - cs * sp;
    R[0 + 3 * 2] = st * cp;
    R[1 + 3 * 0] = cs * ct * sp + ss * cp;
    R[1 + 3 * 1] = -ss * ct * sp + cs * cp;
    R[1 + 3 * 2] = st * sp;
    R[2 + 3 * 0] = -cs * st;
    R[2 + 3 * 1] = ss * st;
    R[2 + 3 * 2] = ct;
}

/* interp3(...,'linear',0) semantics on one axis: nodes 0..n-1, the closed interval
 * [0, n-1] interpolates, anything outside returns the extrapolation value 0. */
static int lin_axis(double f, int n, int* i0, double* t)
{
    if (!(f >= 0.0) || f > (double)(n - 1)) return 0;
    if (n == 1) { *i0 = 0; *t = 0.0; return 1; }
    int i = (int)floor(f);
    if (i >= n - 1) i = n - 2;
    *i0 = i;
    *t = f - (double)i;
    return 1;
}

/* Proj(i,j) = sum_k Vol(R*[X_i; Y_j; Z_k]) with centred coordinates
 * X = -ceil((d1-1)/2):floor((d1-1)/2) (:17-31), trilinear interpolation,
 * extrapolation 0, sum along dim 3 (:55), all in FP64.  Vol is column-major
 * d1 x d2 x d3 FP32; out is d1 x d2 FP64. */
void go_project(const float* vol, int d1, int d2, int d3, const double R[9], double* out)
{
    const int c1 = d1 / 2, c2 = d2 / 2, c3 = d3 / 2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < d2; ++j)
        for (int i = 0; i < d1; ++i) {
            double X = (double)(i - c1), Y = (double)(j - c2);
            double acc = 0.0;
            for (int k = 0; k < d3; ++k) {
                double Z = (double)(k - c3);
                double rx = (R[0] * X + R[3] * Y) + R[6] * Z;
                double ry = (R[1] * X + R[4] * Y) + R[7] * Z;
                double rz = (R[2] * X + R[5] * Y) + R[8] * Z;
                int a0, b0, e0; double ta, tb, te;
                if (!lin_axis(rx + c1, d1, &a0, &ta)) continue;
                if (!lin_axis(ry + c2, d2, &b0, &tb)) continue;
                if (!lin_axis(rz + c3, d3, &e0, &te)) continue;
                int a1 = d1 > 1 ? a0 + 1 : a0, b1 = d2 > 1 ? b0 + 1 : b0, e1 = d3 > 1 ? e0 + 1 : e0;
#define V(a, b, e) ((double)vol[(int64_t)(a) + (int64_t)d1 * ((int64_t)(b) + (int64_t)d2 * (e))])
                double v00 = V(a0, b0, e0) * (1 - ta) + V(a1, b0, e0) * ta;
                double v10 = V(a0, b1, e0) * (1 - ta) + V(a1, b1, e0) * ta;
                double v01 = V(a0, b0, e1) * (1 - ta) + V(a1, b0, e1) * ta;
                double v11 = V(a0, b1, e1) * (1 - ta) + V(a1, b1, e1) * ta;
#undef V
                double w0 = v00 * (1 - tb) + v10 * tb;
                double w1 = v01 * (1 - tb) + v11 * tb;
                acc += w0 * (1 - te) + w1 * te;
            }
            out[i + (int64_t)d1 * j] = acc;
        }
}

/* interp3(...,'cubic',0) (src/calculate3Dprojection_RealSpaceinterp.m:45,51):
 * on a uniform grid MATLAB's 'cubic' is cubic convolution (Keys, IEEE Trans.
 * ASSP 29, 1981) with a = -1/2, weights for the nodes i-1, i, i+1, i+2 of
 * t = f - i in [0, 1]
 *   ((-t + 2) t - 1) t / 2,  ((3t - 5) t t + 2) / 2,  ((-3t + 4) t + 1) t / 2,  (t - 1) t t / 2,
 * and Keys' boundary nodes f(-1) = 3 f(0) - 3 f(1) + f(2), f(n) = 3 f(n-1) -
 * 3 f(n-2) + f(n-3) on every axis (both exact for quadratics).  A query is 0
 * unless each coordinate lies in [0, n-1]; the cell is min(floor(f), n-2).
 * MATLAB is absent here, so this kernel and boundary rule are the published
 * cubic-convolution ones, parity unpinned against MATLAB itself. */
static void keys_weights(double t, double w[4])
{
    w[0] = ((-t + 2.0) * t - 1.0) * t * 0.5;
    w[1] = ((3.0 * t - 5.0) * t * t + 2.0) * 0.5;
    w[2] = ((-3.0 * t + 4.0) * t + 1.0) * t * 0.5;
    w[3] = (t - 1.0) * t * t * 0.5;
}

/* (n+2) samples of one axis, index -1..n, from a strided row of n >= 4 */
static double keys_node(const double* row, int64_t stride, int n, int i)
{
    if (i < 0) return 3.0 * row[0] - 3.0 * row[stride] + row[2 * stride];
    if (i >= n) return 3.0 * row[(int64_t)(n - 1) * stride] - 3.0 * row[(int64_t)(n - 2) * stride] +
                       row[(int64_t)(n - 3) * stride];
    return row[(int64_t)i * stride];
}

/* The volume padded by Keys' boundary nodes, (d1+2) x (d2+2) x (d3+2),
 * x fastest, extended along x, then y, then z. */
static double* keys_pad(const float* vol, int d1, int d2, int d3)
{
    const int p1 = d1 + 2, p2 = d2 + 2, p3 = d3 + 2;
    double* a = (double*)calloc((size_t)p1 * p2 * p3, sizeof(double));
    double* row = (double*)malloc(sizeof(double) * (size_t)(d1 > d2 ? (d1 > d3 ? d1 : d3) : (d2 > d3 ? d2 : d3)));
    for (int z = 0; z < d3; ++z)
        for (int y = 0; y < d2; ++y) {
            for (int x = 0; x < d1; ++x) row[x] = vol[x + (int64_t)d1 * (y + (int64_t)d2 * z)];
            for (int x = -1; x <= d1; ++x)
                a[(x + 1) + (int64_t)p1 * ((y + 1) + (int64_t)p2 * (z + 1))] = keys_node(row, 1, d1, x);
        }
    for (int z = 0; z < d3; ++z)
        for (int x = 0; x < p1; ++x) {
            for (int y = 0; y < d2; ++y) row[y] = a[x + (int64_t)p1 * ((y + 1) + (int64_t)p2 * (z + 1))];
            for (int y = -1; y <= d2; ++y)
                a[x + (int64_t)p1 * ((y + 1) + (int64_t)p2 * (z + 1))] = keys_node(row, 1, d2, y);
        }
    for (int y = 0; y < p2; ++y)
        for (int x = 0; x < p1; ++x) {
            for (int z = 0; z < d3; ++z) row[z] = a[x + (int64_t)p1 * (y + (int64_t)p2 * (z + 1))];
            for (int z = -1; z <= d3; ++z) a[x + (int64_t)p1 * (y + (int64_t)p2 * (z + 1))] = keys_node(row, 1, d3, z);
        }
    free(row);
    return a;
}

static int cubic_axis(double f, int n, int* i0, double w[4])
{
    if (!(f >= 0.0) || f > (double)(n - 1)) return 0;
    int i = (int)floor(f);
    if (i > n - 2) i = n - 2;
    *i0 = i;
    keys_weights(f - (double)i, w);
    return 1;
}

static double cubic_at(const double* a, int d1, int d2, int d3, double fx, double fy, double fz, int* inside)
{
    int a0, b0, e0;
    double wx[4], wy[4], wz[4];
    *inside = cubic_axis(fx, d1, &a0, wx) && cubic_axis(fy, d2, &b0, wy) && cubic_axis(fz, d3, &e0, wz);
    if (!*inside) return 0.0;
    const int64_t p1 = d1 + 2, p2 = d2 + 2;
    double s = 0.0;
    for (int c = 0; c < 4; ++c)
        for (int b = 0; b < 4; ++b)
            for (int q = 0; q < 4; ++q)   /* padded index of node a0 - 1 + q is a0 + q */
                s += wz[c] * wy[b] * wx[q] * a[(a0 + q) + p1 * ((b0 + b) + p2 * (int64_t)(e0 + c))];
    return s;
}

/* the cubic interpolant at np points (x, y, z in node units), 0 outside */
void go_interp_cubic(const float* vol, int d1, int d2, int d3, const double* pts, int64_t np, double* out)
{
    double* a = keys_pad(vol, d1, d2, d3);
    for (int64_t q = 0; q < np; ++q) {
        int in;
        out[q] = cubic_at(a, d1, d2, d3, pts[3 * q], pts[3 * q + 1], pts[3 * q + 2], &in);
    }
    free(a);
}

/* calculate3Dprojection_RealSpaceinterp with interp3 'cubic' (see above) */
void go_project_cubic(const float* vol, int d1, int d2, int d3, const double R[9], double* out)
{
    const int c1 = d1 / 2, c2 = d2 / 2, c3 = d3 / 2;
    double* a = keys_pad(vol, d1, d2, d3);
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < d2; ++j)
        for (int i = 0; i < d1; ++i) {
            double X = (double)(i - c1), Y = (double)(j - c2);
            double acc = 0.0;
            for (int k = 0; k < d3; ++k) {
                double Z = (double)(k - c3);
                double rx = (R[0] * X + R[3] * Y) + R[6] * Z;
                double ry = (R[1] * X + R[4] * Y) + R[7] * Z;
                double rz = (R[2] * X + R[5] * Y) + R[8] * Z;
                int in;
                const double v = cubic_at(a, d1, d2, d3, rx + c1, ry + c2, rz + c3, &in);
                if (in) acc += v;
            }
            out[i + (int64_t)d1 * j] = acc;
        }
    free(a);
}

/* ------------------------------------------------------------------------- */
/* Type-2 NUFFT interpolation (the CPU gridding route, nufft2d2 at            */
/* fillInFourierGrid_DFT.m:249-254), for the CPU baseline                     */
/* ------------------------------------------------------------------------- */

/* F_j = sum over the w x w nearest nodes (m1, m2) of the FP64 fine grid G
 * (M1 x M2 row-major, interleaved complex, G[m1][m2]) of
 * G[m1 mod M1][m2 mod M2] phi(x_j - m1) phi(y_j - m2), x_j = un_j M1,
 * y_j = vn_j M2, with the exponential-of-semicircle kernel
 * phi(t) = exp(beta (sqrt(1 - (2t/w)^2) - 1)) on |t| < w/2 (Barnett, Magland
 * and af Klinteberg, SIAM J. Sci. Comput. 41, 2019).  The caller has divided
 * the modes by the kernel's Fourier transform and taken the fine grid's FFT. */
static double es_kernel(double t, int w, double beta)
{
    const double z = 2.0 * t / w;
    if (z * z >= 1.0) return 0.0;
    return exp(beta * (sqrt(1.0 - z * z) - 1.0));
}

void go_nufft_interp(const double* G, int M1, int M2, int w, double beta, int64_t m, const float* un,
                     const float* vn, double* F)
{
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < m; ++j) {
        const double x = (double)un[j] * M1, y = (double)vn[j] * M2;
        const int i0 = (int)ceil(x - 0.5 * w), j0 = (int)ceil(y - 0.5 * w);
        double kx[32], ky[32];
        for (int q = 0; q < w; ++q) {
            kx[q] = es_kernel(x - (i0 + q), w, beta);
            ky[q] = es_kernel(y - (j0 + q), w, beta);
        }
        double fr = 0.0, fi = 0.0;
        for (int q = 0; q < w; ++q) {
            int r1 = (i0 + q) % M1;
            if (r1 < 0) r1 += M1;
            const double* row = G + 2 * (int64_t)r1 * M2;
            double gr = 0.0, gi = 0.0;
            for (int s = 0; s < w; ++s) {
                int c = (j0 + s) % M2;
                if (c < 0) c += M2;
                gr += row[2 * c] * ky[s];
                gi += row[2 * c + 1] * ky[s];
            }
            fr += gr * kx[q];
            fi += gi * kx[q];
        }
        F[2 * j] = fr;
        F[2 * j + 1] = fi;
    }
}

/* OpenMP threads of the oracle's parallel loops (the CPU baseline's core count). */
void go_set_threads(int n);
#ifdef _OPENMP
#include <omp.h>
void go_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
#else
void go_set_threads(int n) { (void)n; }
#endif
