// gf_host_math.cpp -- host-side FP64/FP32 set-up math of the gridding and the
// projector.  Compiled for the host only; FP contraction is off so every FP32
// expression rounds once per operation in the order written.
#pragma clang fp contract(off)

#include <cmath>
#include <cstring>

#include "gf_common.h"

namespace gf {

namespace {

// MatrixQuaternionRot (fillInFourierGrid_DFT.m:353-363): active rotation by
// theta degrees about vec, FP64, column-major.
void quat_rot(const double vec[3], double theta_deg, double M[9])
{
    const double th = theta_deg * M_PI / 180.0;
    const double nrm = std::sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
    const double a0 = vec[0] / nrm, a1 = vec[1] / nrm, a2 = vec[2] / nrm;
    const double w = std::cos(th / 2.0), s = std::sin(th / 2.0);
    const double x = -s * a0, y = -s * a1, z = -s * a2;
    M[0] = 1.0 - 2.0 * (y * y) - 2.0 * (z * z);
    M[1] = 2.0 * x * y - 2.0 * w * z;
    M[2] = 2.0 * x * z + 2.0 * w * y;
    M[3] = 2.0 * x * y + 2.0 * w * z;
    M[4] = 1.0 - 2.0 * (x * x) - 2.0 * (z * z);
    M[5] = 2.0 * y * z - 2.0 * w * x;
    M[6] = 2.0 * x * z - 2.0 * w * y;
    M[7] = 2.0 * y * z + 2.0 * w * x;
    M[8] = 1.0 - 2.0 * (x * x) - 2.0 * (y * y);
}

void mul3(const double A[9], const double B[9], double C[9])
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[i + 3 * j] = (A[i] * B[3 * j] + A[i + 3] * B[1 + 3 * j]) + A[i + 6] * B[2 + 3 * j];
}

// MATLAB sind/cosd: argument reduced in degrees so multiples of 90 are exact.
void sincosd(double x, double* s, double* c)
{
    const double r = std::fmod(x, 360.0);
    const double q = std::nearbyint(r / 90.0);
    const double rem = r - 90.0 * q;
    const double sr = std::sin(rem * M_PI / 180.0), cr = std::cos(rem * M_PI / 180.0);
    switch ((((int)q) % 4 + 4) % 4) {
        case 0: *s = sr; *c = cr; break;
        case 1: *s = cr; *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}

}  // namespace

// R = Q(v1,phi)*Q(v2,theta)*Q(v3,psi) in FP64
// (src/calculate3Dprojection_RealSpaceinterp_general.m:3-7).
void rotation_f64(const double v1[3], const double v2[3], const double v3[3], double phi, double theta, double psi,
                  double R[9])
{
    double A[9], B[9], C[9], AB[9];
    quat_rot(v1, phi, A);
    quat_rot(v2, theta, B);
    quat_rot(v3, psi, C);
    mul3(A, B, AB);
    mul3(AB, C, R);
}

// fillInFourierGrid_DFT.m:108-114: R = Q(v1,phi)*Q(v2,theta)*Q(v3,psi) in FP64,
// stored as single (rotMATs, :101).
void rotation_f32(const double v1[3], const double v2[3], const double v3[3], double phi,
                  double theta, double psi, float Rf[9])
{
    double A[9], B[9], C[9], AB[9], R[9];
    quat_rot(v1, phi, A);
    quat_rot(v2, theta, B);
    quat_rot(v3, psi, C);
    mul3(A, B, AB);
    mul3(AB, C, R);
    for (int i = 0; i < 9; ++i) Rf[i] = (float)R[i];
}

// Per-projection constants of the index pass:
//  normVECs(p,:) = R(:,3) (:116-117);
//  nh = normvec / norm(normvec) (distancePointsPlane_YY, :332), with norm()
//       taken as the correctly rounded FP64 value;
//  nn = sum(n.*n) (closestpoint, :347), left to right in FP32;
//  LU factors of R for CP_plane = R \ CP (:182): partial pivoting on the first
//  largest |a|, multipliers scaled by the FP32 reciprocal of the pivot, Schur
//  update a - l*u (the LAPACK sgetrf2 order for a 3x3).
void plane_setup(const float R[9], PlaneConst* pc)
{
    std::memset(pc, 0, sizeof(*pc));
    for (int i = 0; i < 3; ++i) pc->nv[i] = R[6 + i];
    const double s2 = (double)pc->nv[0] * pc->nv[0] + (double)pc->nv[1] * pc->nv[1] +
                      (double)pc->nv[2] * pc->nv[2];
    const float nrm = (float)std::sqrt(s2);
    for (int i = 0; i < 3; ++i) pc->nh[i] = pc->nv[i] / nrm;
    pc->nn = (pc->nv[0] * pc->nv[0] + pc->nv[1] * pc->nv[1]) + pc->nv[2] * pc->nv[2];

    float A[9];
    std::memcpy(A, R, sizeof(A));
    for (int j = 0; j < 3; ++j) {
        int p = j;
        float best = std::fabs(A[j + 3 * j]);
        for (int i = j + 1; i < 3; ++i) {
            const float a = std::fabs(A[i + 3 * j]);
            if (a > best) { best = a; p = i; }
        }
        pc->piv[j] = p;
        if (A[p + 3 * j] != 0.0f) {
            if (p != j)
                for (int c = 0; c < 3; ++c) {
                    const float t = A[j + 3 * c];
                    A[j + 3 * c] = A[p + 3 * c];
                    A[p + 3 * c] = t;
                }
            const float rcp = 1.0f / A[j + 3 * j];
            for (int i = j + 1; i < 3; ++i) A[i + 3 * j] = A[i + 3 * j] * rcp;
        }
        for (int c = j + 1; c < 3; ++c) {
            const float u = A[j + 3 * c];
            for (int i = j + 1; i < 3; ++i) A[i + 3 * c] = A[i + 3 * c] - A[i + 3 * j] * u;
        }
    }
    std::memcpy(pc->lu, A, sizeof(A));

    int ax = 0;
    for (int i = 1; i < 3; ++i)
        if (std::fabs(pc->nh[i]) > std::fabs(pc->nh[ax])) ax = i;
    pc->axis = ax;
    for (int i = 0; i < 3; ++i) pc->nhd[i] = (double)pc->nh[i];
}

// ZYZ matrix of src/calculate3Dprojection_RealSpaceinterp.m:13-15 (cosd/sind).
void zyz_matrix(double phi, double theta, double psi, double R[9])
{
    double sp, cp, st, ct, ss, cs;
    sincosd(phi, &sp, &cp);
    sincosd(theta, &st, &ct);
    sincosd(psi, &ss, &cs);
    R[0] = cs * ct * cp - ss * sp;
    R[1] = cs * ct * sp + ss * cp;
    R[2] = -cs * st;
    R[3] = -ss * ct * cp - cs * sp;
    R[4] = -ss * ct * sp + cs * cp;
    R[5] = ss * st;
    R[6] = st * cp;
    R[7] = st * sp;
    R[8] = ct;
}

// fillInFourierGrid_DFT.m:122-136.
HalfGrid half_grid(int n1, int n2, int centro)
{
    HalfGrid g;
    if (centro) {
        g.lo[0] = -(n1 / 2); g.len[0] = n1 / 2 + 1;
        g.lo[1] = -(n2 / 2); g.len[1] = 2 * (n2 / 2) + 1;
        g.lo[2] = -(n1 / 2); g.len[2] = 2 * (n1 / 2) + 1;
    } else {
        g.lo[0] = -(n1 / 2); g.len[0] = n1;
        g.lo[1] = -(n2 / 2); g.len[1] = n2;
        g.lo[2] = -(n1 / 2); g.len[2] = n1;
    }
    return g;
}

}  // namespace gf
