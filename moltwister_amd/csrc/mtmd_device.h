// mtmd_device.h - device arithmetic that must match the reference CPU build.
//
// Each helper keeps the operation order and precision of the reference
// function named beside it (paths relative to the MolTwister tree;
// FCF = Cmd/MolTwisterCmdMolDyn/ForceFields/FunctorCalcForce.cpp,
// V3 = Utilities/3DVector.cpp). Contraction into FMA is disabled inside every
// helper, so the helpers stay exact even when inlined into kernels of a
// translation unit built with contraction enabled.
#pragma once
#include "mtmd_internal.h"
#include <float.h>
#include <limits.h>

#define MT_EXACT_FP _Pragma("clang fp contract(off)")

struct mt_d3 { double x, y, z; };

__device__ __forceinline__ mt_d3 mt_sub(mt_d3 a, mt_d3 b) { MT_EXACT_FP return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ mt_d3 mt_add(mt_d3 a, mt_d3 b) { MT_EXACT_FP return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ mt_d3 mt_mul(mt_d3 a, double s) { MT_EXACT_FP return {a.x * s, a.y * s, a.z * s}; }
// C3DVector::operator* (V3:298-301) and norm2/norm (V3:34-42)
__device__ __forceinline__ double mt_dot(mt_d3 a, mt_d3 b)
{
    MT_EXACT_FP
    return (a.x * b.x) + (a.y * b.y) + (a.z * b.z);
}
__device__ __forceinline__ double mt_norm(mt_d3 a)
{
    MT_EXACT_FP
    return sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}
// C3DVector::cross (V3:243-252)
__device__ __forceinline__ mt_d3 mt_cross(mt_d3 a, mt_d3 b)
{
    MT_EXACT_FP
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// int(float) / int(double) as x86-64 cvtt* instructions give them: truncation,
// INT_MIN for NaN and out-of-range values (the reference is a CPU build)
__device__ __forceinline__ int mt_f2i(float v)
{
    if (!(v >= -2147483648.0f && v < 2147483648.0f)) return INT_MIN;
    return (int)v;
}
__device__ __forceinline__ int mt_d2i(double v)
{
    if (!(v > -2147483649.0 && v < 2147483648.0)) return INT_MIN;
    return (int)v;
}

// C3DVector::moveToSameSideOfPBCAsThis (V3:133-165)
__device__ __forceinline__ mt_d3 mt_move_same_side(mt_d3 self, mt_d3 v, const double* w)
{
    MT_EXACT_FP
    double dX = self.x - v.x, dY = self.y - v.y, dZ = self.z - v.z;
    dX = (dX < 0.0) ? -dX : dX;
    dY = (dY < 0.0) ? -dY : dY;
    dZ = (dZ < 0.0) ? -dZ : dZ;
    if (dX > (w[0] / 2.0)) { if (self.x > v.x) v.x += w[0]; else v.x -= w[0]; }
    if (dY > (w[1] / 2.0)) { if (self.y > v.y) v.y += w[1]; else v.y -= w[1]; }
    if (dZ > (w[2] / 2.0)) { if (self.z > v.z) v.z += w[2]; else v.z -= w[2]; }
    return v;
}

// C3DVector::distToAcrossPBC2 (V3:68-90)
__device__ __forceinline__ double mt_dist2_pbc(mt_d3 self, mt_d3 v, const double* w)
{
    MT_EXACT_FP
    double dX = self.x - v.x, dY = self.y - v.y, dZ = self.z - v.z;
    dX = (dX < 0.0) ? -dX : dX;
    dY = (dY < 0.0) ? -dY : dY;
    dZ = (dZ < 0.0) ? -dZ : dZ;
    if (dX > (w[0] / 2.0)) dX = w[0] - dX;
    if (dY > (w[1] / 2.0)) dY = w[1] - dY;
    if (dZ > (w[2] / 2.0)) dZ = w[2] - dZ;
    return dX * dX + dY * dY + dZ * dZ;
}

// calcNonBondForceCoeffs12 / calcBondForceCoeffs12 (FCF:283-301)
__device__ __forceinline__ mt_d3 mt_coeffs12(mt_d3 r1, mt_d3 r2)
{
    MT_EXACT_FP
    mt_d3 r12 = mt_sub(r2, r1);
    double R = mt_norm(r12);
    double RInv = (R == 0.0) ? 1.0 / 1E-10 : 1.0 / R;
    return mt_mul(r12, RInv);
}

// calcAngularForceCoeffs13 (FCF:303-324)
__device__ inline mt_d3 mt_ang_coeffs13(mt_d3 r1, mt_d3 r2, mt_d3 r3)
{
    MT_EXACT_FP
    mt_d3 r12 = mt_sub(r1, r2);
    double R = mt_norm(r12);
    double RInv = (R == 0.0) ? 1.0 / 1E-10 : 1.0 / R;
    mt_d3 K = mt_mul(r12, RInv);
    mt_d3 r32 = mt_sub(r3, r2);
    double A = K.x * r32.x + K.y * r32.y + K.z * r32.z;
    double B = r32.x * r32.x + r32.y * r32.y + r32.z * r32.z;
    double sqrtB = sqrt(B);
    double sqrtB3 = (B == 0.0) ? sqrt(1E-30) : sqrtB * sqrtB * sqrtB;
    double dW_dx = (K.x * B - r32.x * A) / sqrtB3;
    double dW_dy = (K.y * B - r32.y * A) / sqrtB3;
    double dW_dz = (K.z * B - r32.z * A) / sqrtB3;
    double W = (sqrtB == 0.0) ? A / 10E-3 : A / sqrtB;
    double dDiv = sqrt(1.0 - W * W);
    double dTheta_dW = (dDiv == 0.0) ? -1.0 / 1E-10 : -1.0 / dDiv;
    return {dTheta_dW * dW_dx, dTheta_dW * dW_dy, dTheta_dW * dW_dz};
}

// calcAngularForceCoeffs2 (FCF:326-346)
__device__ inline mt_d3 mt_ang_coeffs2(mt_d3 r1, mt_d3 r2, mt_d3 r3)
{
    MT_EXACT_FP
    mt_d3 r12 = mt_sub(r1, r2);
    mt_d3 r32 = mt_sub(r3, r2);
    double A = r12.x * r12.x + r12.y * r12.y + r12.z * r12.z;
    double B = r32.x * r32.x + r32.y * r32.y + r32.z * r32.z;
    double C = r12.x * r32.x + r12.y * r32.y + r12.z * r32.z;
    double AB = A * B;
    double sqrtAB = sqrt(AB);
    double sqrtAB3 = (AB == 0.0) ? sqrt(1E-30) : sqrtAB * sqrtAB * sqrtAB;
    double dW_dx = ((2.0 * r2.x - r1.x - r3.x) * AB + (r32.x * A + r12.x * B) * C) / sqrtAB3;
    double dW_dy = ((2.0 * r2.y - r1.y - r3.y) * AB + (r32.y * A + r12.y * B) * C) / sqrtAB3;
    double dW_dz = ((2.0 * r2.z - r1.z - r3.z) * AB + (r32.z * A + r12.z * B) * C) / sqrtAB3;
    double W = (sqrtAB == 0.0) ? A / 10E-3 : C / sqrtAB;
    double dDiv = sqrt(1.0 - W * W);
    double dTheta_dW = (dDiv == 0.0) ? -1.0 / 1E-10 : -1.0 / dDiv;
    return {dTheta_dW * dW_dx, dTheta_dW * dW_dy, dTheta_dW * dW_dz};
}

// calcDihedralForceCoeffs14 before its final sign flip (FCF:348-377)
__device__ inline mt_d3 mt_dih_coeffs14(mt_d3 r1, mt_d3 r2, mt_d3 r3, mt_d3 r4)
{
    MT_EXACT_FP
    mt_d3 r21 = mt_sub(r1, r2);
    mt_d3 r23 = mt_sub(r3, r2);
    mt_d3 r32 = mt_mul(r23, -1.0);
    mt_d3 r34 = mt_sub(r4, r3);
    mt_d3 n1 = mt_cross(r23, r21);
    double n1Norm = mt_norm(n1);
    double n1NormInv = (n1Norm == 0.0) ? 1.0 / 1E-10 : 1.0 / n1Norm;
    mt_d3 K = mt_mul(n1, n1NormInv);
    double a = r32.y * r34.z - r32.z * r34.y;
    double b = r32.z * r34.x - r32.x * r34.z;
    double c = r32.x * r34.y - r32.y * r34.x;
    double A = K.x * a + K.y * b + K.z * c;
    double B = a * a + b * b + c * c;
    double sqrtBInv = (B == 0.0) ? 1.0 / 1E-20 : 1.0 / sqrt(B);
    double sqrtBInv3 = sqrtBInv * sqrtBInv * sqrtBInv;
    double W = A * sqrtBInv;
    double dW_dx = ((K.y * r32.z - K.z * r32.y) * B - (r32.z * (r32.z * r34.x - r32.x * r34.z) - r32.y * (r32.x * r34.y - r32.y * r34.x)) * A) * sqrtBInv3;
    double dW_dy = ((K.z * r32.x - K.x * r32.z) * B - (r32.x * (r32.x * r34.y - r32.y * r34.x) - r32.z * (r32.y * r34.z - r32.z * r34.y)) * A) * sqrtBInv3;
    double dW_dz = ((K.x * r32.y - K.y * r32.x) * B - (r32.y * (r32.y * r34.z - r32.z * r34.y) - r32.x * (r32.z * r34.x - r32.x * r34.z)) * A) * sqrtBInv3;
    double div = sqrt(1.0 - W * W);
    double dTheta_dW = (div == 0.0) ? 1.0 / 1E-10 : 1.0 / div;
    return {dTheta_dW * dW_dx, dTheta_dW * dW_dy, dTheta_dW * dW_dz};
}

// calcDihedralForceCoeffs23 before its final sign flip (FCF:379-424)
__device__ inline mt_d3 mt_dih_coeffs23(mt_d3 r1, mt_d3 r2, mt_d3 r3, mt_d3 r4)
{
    MT_EXACT_FP
    mt_d3 r21 = mt_sub(r1, r2);
    mt_d3 r23 = mt_sub(r3, r2);
    mt_d3 r32 = mt_mul(r23, -1.0);
    mt_d3 r34 = mt_sub(r4, r3);
    double K1 = r32.z * r34.x - r32.x * r34.z;
    double K2 = r34.z - r32.z;
    double K3 = r23.z * r21.x - r23.x * r21.z;
    double K4 = r32.x * r34.y - r32.y * r34.x;
    double K5 = r32.y - r34.y;
    double K6 = r23.x * r21.y - r23.y * r21.x;
    double L1 = r32.y * r34.z - r32.z * r34.y;
    double L2 = -K2;
    double L3 = r23.y * r21.z - r23.z * r21.y;
    double L4 = K4;
    double L5 = r34.x - r32.x;
    double L6 = K6;
    double M1 = L1, M2 = -K5, M3 = L3, M4 = K1, M5 = -L5, M6 = K3;
    double A = L3 * L3 + K3 * K3 + K6 * K6;
    double B = L1 * L1 + K1 * K1 + K4 * K4;
    double C = L3 * L1 + K3 * K1 + K6 * K4;
    double AB = A * B;
    double sqrtABInv = (AB == 0.0) ? 1.0 / 1E-20 : 1.0 / sqrt(AB);
    double sqrtABInv3 = sqrtABInv * sqrtABInv * sqrtABInv;
    double W = C * sqrtABInv;
    double dC_dx = K2 * K3 + r21.y * K4 + K5 * K6 - r21.z * K1;
    double dC_dy = L2 * L3 - r21.x * L4 + L5 * L6 + r21.z * L1;
    double dC_dz = M2 * M3 + r21.x * M4 + M5 * M6 - r21.y * M1;
    double dW_dx = (dC_dx * AB - C * ((r21.y * K6 - r21.z * K3) * B + (K1 * K2 + K4 * K5) * A)) * sqrtABInv3;
    double dW_dy = (dC_dy * AB - C * ((r21.z * L3 - r21.x * L6) * B + (L1 * L2 + L4 * L5) * A)) * sqrtABInv3;
    double dW_dz = (dC_dz * AB - C * ((r21.x * M6 - r21.y * M3) * B + (M1 * M2 + M4 * M5) * A)) * sqrtABInv3;
    double div = sqrt(1.0 - W * W);
    double dTheta_dW = (div == 0.0) ? 1.0 / 1E-10 : 1.0 / div;
    return {dTheta_dW * dW_dx, dTheta_dW * dW_dy, dTheta_dW * dW_dz};
}

// calcAngle (FCF:432-440)
__device__ __forceinline__ float mt_calc_angle(mt_d3 rk, mt_d3 ri, mt_d3 rj)
{
    MT_EXACT_FP
    mt_d3 rik = mt_sub(rk, ri), rij = mt_sub(rj, ri);
    float dot = (float)mt_dot(rik, rij);
    float RR = (float)mt_norm(rik) * (float)mt_norm(rij);
    float cs = (RR != 0.0f) ? dot / RR : 1.0f;
    return (float)acos((double)cs);
}

// calcDihedral (FCF:442-455)
__device__ __forceinline__ float mt_calc_dihedral(mt_d3 rk, mt_d3 ri, mt_d3 rj, mt_d3 rl)
{
    MT_EXACT_FP
    mt_d3 rik = mt_sub(rk, ri), rij = mt_sub(rj, ri), rjl = mt_sub(rl, rj);
    mt_d3 n1 = mt_cross(rik, rij), n2 = mt_cross(rij, rjl);
    float dot = (float)mt_dot(n1, n2);
    float nn = (float)mt_norm(n1) * (float)mt_norm(n2);
    float cs = (nn != 0.0f) ? dot / nn : 1.0f;
    if (cs < -1.0f) cs = -1.0f;
    if (cs > 1.0f) cs = 1.0f;
    return (float)acos((double)cs);
}

// interpolateLinear (FCF:457-465)
__device__ __forceinline__ float mt_interp(float r, float x1, float y1, float x2, float y2)
{
    MT_EXACT_FP
    float denom = x1 - x2;
    if (denom == 0.0f) denom = 3.0f * FLT_MIN;
    float a = (y1 - y2) / denom;
    float b = y1 - a * x1;
    return a * r + b;
}

// calcIndexOfLowerPointOfInterpolationLine (FCF:426-430) + interpolation of
// one profile (FCF:481-504, 547-567)
__device__ __forceinline__ void mt_table_eval(const float4* tab, float x, bool want_e, float& fsum, float& esum)
{
    MT_EXACT_FP
    float x0 = tab[0].x, xl = tab[MT_NPTS - 1].x;
    float delta = (xl - x0) / (float)(MT_NPTS - 1);
    int idx = mt_f2i((x - x0) / delta);
    if (idx >= 0 && idx < MT_NPTS - 1) {
        float4 p1 = tab[idx], p2 = tab[idx + 1];
        fsum += mt_interp(x, p1.x, p1.y, p2.x, p2.y);
        if (want_e) esum += mt_interp(x, p1.x, p1.z, p2.x, p2.z);
    }
}

// calcForceCoulombOn_r_k (FCF:514-535)
__device__ __forceinline__ mt_d3 mt_coulomb_pair(mt_d3 rk, mt_d3 ri, float qk, float qi, int k, int i, float& U)
{
    MT_EXACT_FP
    mt_d3 r = mt_sub(rk, ri);
    float R = (float)mt_norm(r);
    double R3 = (double)(R * R * R);
    double K = (double)(MT_COULOMB_K * qk * qi);
    if (i > k && R != 0.0f) U += (float)K / R;
    if (R3 != 0.0) return mt_mul(r, K / R3);
    return {0.0, 0.0, 0.0};
}

// scaleForcesAndPotentials (FCF:95-157) over atom k's bonded lists
__device__ inline void mt_scale_pair(const MtDev& d, const MtParams& p, mt_d3& f, float& u, int i, int molk, int k)
{
    MT_EXACT_FP
    const MtBonded& b = d.bl;
    bool hasB = false, hasA = false, hasD = false;
    if (b.has_bonds) {
        int f0 = b.b_first[k], nn = b.b_num[k];
        for (int j = 0; j < nn; j++)
            if (i == b.b_j[f0 + j]) { f = mt_mul(f, (double)p.scale12); u *= p.scale12; hasB = true; break; }
    }
    if (!hasB && b.has_angles) {
        int f0 = b.a_first[k], nn = b.a_num[k];
        for (int j = 0; j < nn; j++)
            if (!b.a_center[f0 + j] && i == b.a_j[f0 + j]) {
                f = mt_mul(f, (double)p.scale13); u *= p.scale13; hasA = true; break;
            }
    }
    if (!hasB && !hasA && b.has_dihedrals) {
        int f0 = b.d_first[k], nn = b.d_num[k];
        for (int j = 0; j < nn; j++)
            if (!b.d_center[f0 + j] && i == b.d_l[f0 + j]) {
                f = mt_mul(f, (double)p.scale14); u *= p.scale14; hasD = true; break;
            }
    }
    if (!hasB && !hasA && !hasD) {
        int moli = d.mol[i];
        if (moli >= 0 && molk >= 0 && moli == molk) { f = mt_mul(f, (double)p.scale1N); u *= p.scale1N; }
    }
}

// bonds, angles and dihedrals acting on atom k (FCF:172-232, 537-709).
// `orig` is k's original index (its bonded rows); `slot_of` maps an original
// index to the array slot holding that atom (identity when atoms are unsorted).
__device__ inline void mt_exact_bonded(const MtDev& d, const MtParams& p, int k, mt_d3& F, float& U,
                                       int orig = -1, const int* slot_of = nullptr)
{
    MT_EXACT_FP
    const MtBonded& b = d.bl;
    if (orig < 0) orig = k;
    double4 rk4 = d.pos[k];
    mt_d3 rk = {rk4.x, rk4.y, rk4.z};
#define MT_POS(o) ([&]() { double4 q = d.pos[slot_of ? slot_of[(o)] : (o)]; return mt_d3{q.x, q.y, q.z}; }())
    if (b.has_bonds) {
        int f0 = b.b_first[orig], nn = b.b_num[orig];
        for (int j = 0; j < nn; j++) {
            int ib = b.b_j[f0 + j];
            const float4* tab = d.btab + (size_t)b.b_type[f0 + j] * MT_NPTS;
            mt_d3 ri = mt_move_same_side(rk, MT_POS(ib), p.wbox);
            float r = (float)mt_norm(mt_sub(rk, ri));
            float mdU = 0.0f, e = 0.0f;
            mt_table_eval(tab, r, ib > orig, mdU, e);
            if (ib > orig) U += e;
            F = mt_add(F, mt_mul(mt_coeffs12(ri, rk), (double)mdU));
        }
    }
    if (b.has_angles) {
        int f0 = b.a_first[orig], nn = b.a_num[orig];
        for (int j = 0; j < nn; j++) {
            int g = f0 + j;
            const float4* tab = d.atab + (size_t)b.a_type[g] * MT_NPTS;
            mt_d3 ri = mt_move_same_side(rk, MT_POS(b.a_i[g]), p.wbox);
            mt_d3 rj = mt_move_same_side(rk, MT_POS(b.a_j[g]), p.wbox);
            float mdU = 0.0f, e = 0.0f;
            mt_d3 f;
            if (b.a_center[g]) {
                float theta = mt_calc_angle(ri, rk, rj);
                mt_table_eval(tab, theta, true, mdU, e);
                U += e;
                f = mt_mul(mt_ang_coeffs2(rj, rk, ri), (double)mdU);
            } else {
                float theta = mt_calc_angle(rk, ri, rj);
                mt_table_eval(tab, theta, false, mdU, e);
                f = mt_mul(mt_ang_coeffs13(rj, ri, rk), (double)mdU);
            }
            F = mt_add(F, f);
        }
    }
    if (b.has_dihedrals) {
        int f0 = b.d_first[orig], nn = b.d_num[orig];
        for (int j = 0; j < nn; j++) {
            int g = f0 + j;
            const float4* tab = d.dtab + (size_t)b.d_type[g] * MT_NPTS;
            mt_d3 ri = mt_move_same_side(rk, MT_POS(b.d_i[g]), p.wbox);
            mt_d3 rj = mt_move_same_side(rk, MT_POS(b.d_j[g]), p.wbox);
            mt_d3 rl = mt_move_same_side(rk, MT_POS(b.d_l[g]), p.wbox);
            float mdU = 0.0f, e = 0.0f;
            mt_d3 f;
            if (b.d_center[g]) {
                float theta = mt_calc_dihedral(ri, rk, rj, rl);
                mt_table_eval(tab, theta, true, mdU, e);
                U += 0.5f * e;
                f = mt_mul(mt_mul(mt_dih_coeffs23(rl, rj, rk, ri), -1.0), (double)mdU);
            } else {
                float theta = mt_calc_dihedral(rk, ri, rj, rl);
                mt_table_eval(tab, theta, false, mdU, e);
                f = mt_mul(mt_mul(mt_dih_coeffs14(rl, rj, ri, rk), -1.0), (double)mdU);
            }
            F = mt_add(F, f);
        }
    }
#undef MT_POS
}

// CMt::exp (Cmd/MolTwisterCmdMolDyn/Integrators/Math.cpp:29-38)
__device__ inline double mt_cmt_exp(double x) { return (x > 1.0) ? 2.71828 * x : exp(x); }

// One call of the Nose-Hoover chain propagator (NHChain.cpp:54-113) with
// CFctT (MDLoop.cpp:29-60): Suzuki-Yoshida weights w[3], nresp RESPA steps.
// Every momentum scaling of the chain multiplies all momenta by the same
// factor, so sum p.p/m after a scaling by c is c^2 times the sum before: the
// chain runs on the scalar KE2 = sum p.p/m and returns the accumulated factor
// S, which the next streaming kernel applies to the momenta.
__device__ inline double mt_nh_propagate(double* eta, double* peta, const double* Q, int M, int nresp, double T,
                                         double dt, double g, const double* w, double KE2)
{
    MT_EXACT_FP
    double S = 1.0;
    const double beta = 1.0 / T;
    if (M > 0 && peta[0] > 15.0) peta[0] = 10.0;
    if (M <= 0) return S;
    for (int mu = 0; mu < 3; mu++) {
        const double Dt = w[mu] * dt / (double)nresp;
        for (int i = 0; i < nresp; i++) {
#define MT_G(j) ((j) == 0 ? (KE2 - (g / beta)) : ((peta[(j) - 1] * peta[(j) - 1] / Q[(j) - 1]) - (1.0 / beta)))
            peta[M - 1] += (Dt / 4.0) * MT_G(M - 1);
            for (int j = 0; j < M - 1; j++) {
                double c = mt_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= c;
                peta[j] += (Dt / 4.0) * MT_G(j);
                c = mt_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= c;
            }
            for (int j = 0; j < M; j++) eta[j] += (Dt / 2.0) * (peta[j] / Q[j]);
            const double c0 = mt_cmt_exp(-((Dt / 2.0) * (peta[0] / Q[0])));
            S *= c0;
            KE2 *= c0 * c0;
            for (int j = M - 2; j >= 0; j--) {
                double c = mt_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= c;
                peta[j] += (Dt / 4.0) * MT_G(j);
                c = mt_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= c;
            }
            peta[M - 1] += (Dt / 4.0) * MT_G(M - 1);
#undef MT_G
        }
    }
    return S;
}

// mt_nh_propagate for a chain of exactly M links with the chain state held in
// registers: the same operations in the same order (so the same bits), every
// loop over the links unrolled with constant indices and no guard. Working on
// eta, p_eta and Q in global memory instead costs one dependent memory round
// trip per update, and loops guarded for a run-time M keep a branch around
// every update; either way the serial chain takes several times longer.
template <int M>
__device__ inline double mt_nh_propagate_fixed(double* eta_g, double* peta_g, const double* Q_g, int nresp, double T,
                                               double dt, double g, const double* w, double KE2)
{
    MT_EXACT_FP
    double eta[M], peta[M], Q[M];
#pragma unroll
    for (int j = 0; j < M; j++) {
        eta[j] = eta_g[j];
        peta[j] = peta_g[j];
        Q[j] = Q_g[j];
    }
    double S = 1.0;
    const double beta = 1.0 / T;
    if (peta[0] > 15.0) peta[0] = 10.0;
    for (int mu = 0; mu < 3; mu++) {
        const double Dt = w[mu] * dt / (double)nresp;
        for (int i = 0; i < nresp; i++) {
#define MT_G(j) ((j) == 0 ? (KE2 - (g / beta)) : ((peta[(j) - 1] * peta[(j) - 1] / Q[(j) - 1]) - (1.0 / beta)))
            peta[M - 1] += (Dt / 4.0) * MT_G(M - 1);
#pragma unroll
            for (int j = 0; j < M - 1; j++) {
                double c = mt_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= c;
                peta[j] += (Dt / 4.0) * MT_G(j);
                c = mt_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= c;
            }
#pragma unroll
            for (int j = 0; j < M; j++) eta[j] += (Dt / 2.0) * (peta[j] / Q[j]);
            const double c0 = mt_cmt_exp(-((Dt / 2.0) * (peta[0] / Q[0])));
            S *= c0;
            KE2 *= c0 * c0;
#pragma unroll
            for (int j = M - 2; j >= 0; j--) {
                double c = mt_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= c;
                peta[j] += (Dt / 4.0) * MT_G(j);
                c = mt_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= c;
            }
            peta[M - 1] += (Dt / 4.0) * MT_G(M - 1);
#undef MT_G
        }
    }
#pragma unroll
    for (int j = 0; j < M; j++) {
        eta_g[j] = eta[j];
        peta_g[j] = peta[j];
    }
    return S;
}

// The FP64/MIXED engines run the same chain to trace agreement rather than bit
// for bit: the constants of a call are formed once (1/Q, Dt/8, Dt/4, Dt/2 and
// g T for g / beta) and each scale factor is computed once and applied twice.
// CMt::exp keeps its linear branch for x > 1; its small arguments (|x| < 1/32,
// the chain's usual range, |x| = Dt |p_eta| / (8 Q)) take the Taylor series
// of exp to x^8, whose truncation is below x^9 / 9! < 1e-19, evaluated by
// Estrin's scheme (about half the dependent steps of Horner's rule: the chain
// runs on one lane, where latency is its cost).
__device__ inline double mt_exp_taylor8(double x)
{
    const double x2 = x * x, x4 = x2 * x2;
    const double p01 = 1.0 + x, p23 = fma(x, 1.0 / 6.0, 0.5);
    const double p45 = fma(x, 1.0 / 120.0, 1.0 / 24.0), p67 = fma(x, 1.0 / 5040.0, 1.0 / 720.0);
    const double q0 = fma(x2, p23, p01), q1 = fma(x2, p67, p45);
    return fma(x4, fma(x4, 1.0 / 40320.0, q1), q0);
}

// CMt::exp from its series value t = mt_exp_taylor8(x), which holds for
// |x| < 1/32; outside, the linear branch (x > 1) or exp
__device__ inline double mt_cmt_exp_from(double x, double t)
{
    if (fabs(x) < 0.03125) return t;
    return x > 1.0 ? 2.71828 * x : exp(x);
}

__device__ inline double mt_cmt_exp_fast(double x) { return mt_cmt_exp_from(x, mt_exp_taylor8(x)); }

template <int M>
__device__ inline double mt_nh_propagate_fast(double* eta_g, double* peta_g, const double* Q_g, int nresp, double T,
                                              double dt, double g, const double* w, double KE2)
{
    double eta[M], peta[M], iQ[M];
#pragma unroll
    for (int j = 0; j < M; j++) {
        eta[j] = eta_g[j];
        peta[j] = peta_g[j];
        iQ[j] = 1.0 / Q_g[j];
    }
    double S = 1.0;
    const double gT = g * T;
    if (peta[0] > 15.0) peta[0] = 10.0;
    for (int mu = 0; mu < 3; mu++) {
        const double Dt = w[mu] * dt / (double)nresp;
        const double q8 = Dt * 0.125, q4 = Dt * 0.25, q2 = Dt * 0.5;
        for (int i = 0; i < nresp; i++) {
#define MT_GF(j) ((j) == 0 ? (KE2 - gT) : (peta[(j) - 1] * peta[(j) - 1] * iQ[(j) - 1] - T))
            peta[M - 1] += q4 * MT_GF(M - 1);
            // the forward sweep's factors depend on links not yet updated in
            // it (and the last link, updated just above): all at once
            double c[M > 1 ? M - 1 : 1];
#pragma unroll
            for (int j = 0; j < M - 1; j++) {
                const double x = -q8 * (peta[j + 1] * iQ[j + 1]);
                c[j] = mt_exp_taylor8(x);
            }
#pragma unroll
            for (int j = 0; j < M - 1; j++) c[j] = mt_cmt_exp_from(-q8 * (peta[j + 1] * iQ[j + 1]), c[j]);
#pragma unroll
            for (int j = 0; j < M - 1; j++) peta[j] = (peta[j] * c[j] + q4 * MT_GF(j)) * c[j];
#pragma unroll
            for (int j = 0; j < M; j++) eta[j] += q2 * (peta[j] * iQ[j]);
            const double c0 = mt_cmt_exp_fast(-(q2 * (peta[0] * iQ[0])));
            S *= c0;
            KE2 *= c0 * c0;
#pragma unroll
            for (int j = M - 2; j >= 0; j--) {
                const double cj = mt_cmt_exp_fast(-q8 * (peta[j + 1] * iQ[j + 1]));
                peta[j] = (peta[j] * cj + q4 * MT_GF(j)) * cj;
            }
            peta[M - 1] += q4 * MT_GF(M - 1);
#undef MT_GF
        }
    }
#pragma unroll
    for (int j = 0; j < M; j++) {
        eta_g[j] = eta[j];
        peta_g[j] = peta[j];
    }
    return S;
}

// the chain propagation of a thermostat step: in registers for chains of up
// to 8 links, else on the state in memory
__device__ inline double mt_nh_step(double* eta, double* peta, const double* Q, int M, int nresp, double T,
                                    double dt, double g, const double* w, double KE2, bool fast = false)
{
    if (fast) {
        switch (M) {
        case 1: return mt_nh_propagate_fast<1>(eta, peta, Q, nresp, T, dt, g, w, KE2);
        case 2: return mt_nh_propagate_fast<2>(eta, peta, Q, nresp, T, dt, g, w, KE2);
        case 3: return mt_nh_propagate_fast<3>(eta, peta, Q, nresp, T, dt, g, w, KE2);
        case 4: return mt_nh_propagate_fast<4>(eta, peta, Q, nresp, T, dt, g, w, KE2);
        case 5: return mt_nh_propagate_fast<5>(eta, peta, Q, nresp, T, dt, g, w, KE2);
        case 6: return mt_nh_propagate_fast<6>(eta, peta, Q, nresp, T, dt, g, w, KE2);
        case 7: return mt_nh_propagate_fast<7>(eta, peta, Q, nresp, T, dt, g, w, KE2);
        case 8: return mt_nh_propagate_fast<8>(eta, peta, Q, nresp, T, dt, g, w, KE2);
        default: break;                           // longer chains: the reference order below
        }
    }
    switch (M) {
    case 1: return mt_nh_propagate_fixed<1>(eta, peta, Q, nresp, T, dt, g, w, KE2);
    case 2: return mt_nh_propagate_fixed<2>(eta, peta, Q, nresp, T, dt, g, w, KE2);
    case 3: return mt_nh_propagate_fixed<3>(eta, peta, Q, nresp, T, dt, g, w, KE2);
    case 4: return mt_nh_propagate_fixed<4>(eta, peta, Q, nresp, T, dt, g, w, KE2);
    case 5: return mt_nh_propagate_fixed<5>(eta, peta, Q, nresp, T, dt, g, w, KE2);
    case 6: return mt_nh_propagate_fixed<6>(eta, peta, Q, nresp, T, dt, g, w, KE2);
    case 7: return mt_nh_propagate_fixed<7>(eta, peta, Q, nresp, T, dt, g, w, KE2);
    case 8: return mt_nh_propagate_fixed<8>(eta, peta, Q, nresp, T, dt, g, w, KE2);
    default: return M <= 0 ? 1.0 : mt_nh_propagate(eta, peta, Q, M, nresp, T, dt, g, w, KE2);
    }
}

// The words of a 16-bit list row pass through a ring of 16 words (two
// chunks) per lane in LDS: a hit is one 16-bit store and a count, and a chunk
// that has filled is copied to the list once per group of candidates. A group
// brings at most eight words, so at most one chunk completes in it and the
// ring never holds more than 7 + 8 words. Lanes' rings are 9 dwords apart, so
// the 16-bit stores of a wavefront fall in distinct banks.
#define MT_RING_DW 9
struct MtRing16 {
    uint32_t* ring;    // this lane's ring
    uint4* out;        // chunk 0 of the atom's row (column k of the chunked layout)
    size_t stride;     // chunks between a row's chunks (n)
    int capc;          // chunks a row holds
    int cnt;           // words written
    int fl;            // chunks copied out
};

__device__ __forceinline__ void mt_ring_put(MtRing16& g, int w)
{
    reinterpret_cast<uint16_t*>(g.ring)[g.cnt & 15] = (uint16_t)w;
    g.cnt++;
}

__device__ __forceinline__ void mt_ring_flush(MtRing16& g)
{
    if ((g.cnt >> 3) > g.fl) {
        const uint32_t* c = g.ring + (g.fl & 1) * 4;
        const uint4 v = make_uint4(c[0], c[1], c[2], c[3]);
        if (g.fl < g.capc) g.out[(size_t)g.fl * g.stride] = v;
        g.fl++;
    }
}

// the chunk being filled is a window of the last words written, r0..r3 (the
// oldest lowest): a word enters at the top and the window moves down one
// word, so after a chunk's last word the window is the chunk itself
template <bool W16>
__device__ __forceinline__ void mt_neigh_push(uint32_t w, uint32_t& r0, uint32_t& r1, uint32_t& r2, uint32_t& r3)
{
    if constexpr (W16) {
        // eight 16-bit words: ((hi:lo) >> 16) moves each register down one word
        r0 = __builtin_amdgcn_alignbit(r1, r0, 16);
        r1 = __builtin_amdgcn_alignbit(r2, r1, 16);
        r2 = __builtin_amdgcn_alignbit(r3, r2, 16);
        r3 = __builtin_amdgcn_alignbit(w, r3, 16);
    } else {
        r0 = r1;
        r1 = r2;
        r2 = r3;
        r3 = w;
    }
}

// The tile passes without charges (LEAN) need no inside test: a pair at or
// beyond rc, or closer than the table's first point, indexes segment
// MT_NPTS - 1 of its row, whose force and energy coefficients are zero, so its
// terms vanish by themselves. Only the band around rc, where FP32 and the FP64
// reference predicate may disagree, is corrected afterwards: the band path
// takes back these same terms when the FP64 test puts the pair outside. Both
// places compute them with these helpers, the same operations in the same
// order, and the force-only and energy passes give the same forces.
//
// r^2 carries 1e-10 from its first product on, so it is never zero (the pad
// word of a list is the atom itself): 1/r stays finite, r = 1e-5 at the
// least, below the table's first point 0.01, and such a pair indexes the zero
// segment as the reference skips r < 1e-3. From r = 0.1 on, the 1e-10 is
// below half an FP32 ulp of r^2 and changes it only where it tips a rounding.
__device__ __forceinline__ float mt_lean_r2(float dx, float dy, float dz)
{
    return fmaf(dz, dz, fmaf(dx, dx, fmaf(dy, dy, 1.0e-10f)));
}

// 1/r of a mt_lean_r2 value and the segment of the pair
__device__ __forceinline__ float mt_lean_rinv(float r2, float x0, float invd, int& seg)
{
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const int idx = (int)(fmaf(r2, rinv, -x0) * invd);
    seg = (int)min((unsigned)idx, (unsigned)(MT_NPTS - 1));
    return rinv;
}

// the sum of v over a workgroup of NT threads (a tree in LDS, the same order
// every run); thread 0 stores it
template <int NT>
__device__ inline void mt_block_sum_store(double v, double* out)
{
    __shared__ double sh[NT];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = NT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// sum over the G consecutive lanes that share an atom (butterfly, fixed
// order, so every run gives the same bits); every lane ends with the total
template <int G>
__device__ __forceinline__ void mt_lane_sum(double& x, double& y, double& z, float& u)
{
#pragma unroll
    for (int m = 1; m < G; m <<= 1) {
        x += __shfl_xor(x, m);
        y += __shfl_xor(y, m);
        z += __shfl_xor(z, m);
        u += __shfl_xor(u, m);
    }
}

// CSimulationBox::pbcAdd (SimulationBox.cpp:135-166): one wrap step
__device__ inline void mt_pbc_add(double& pos, double L, double Lm)
{
    MT_EXACT_FP
    pos -= (double)mt_d2i((pos + L) / Lm) * Lm;
}

// Ascending sort of one cell's list range a[0..cnt) of distinct atom indices
// by one wavefront. The lists are filled by atomics, so their order is
// arbitrary until sorted. Up to 64 entries are ranked in registers; up to
// MT_CELL_SORT_LDS entries are copied into this wave's share of LDS and each
// lane ranks its entries against all of them (broadcast reads, 4 per int4);
// a larger cell falls back to an insertion sort by lane 0. Holds a block
// barrier, so every thread of the block calls it (cnt = 0 for no cell).
// on(v, rank) is called once for every entry v with its sorted position.
struct MtNoSortHook {
    __device__ void operator()(int, int) const {}
};
#define MT_CELL_SORT_LDS 2048
template <int CAP = MT_CELL_SORT_LDS, typename Hook = MtNoSortHook>
__device__ inline void mt_wave_sort_cell(int* __restrict__ a, int cnt, int* __restrict__ lds, Hook on = Hook())
{
    const int lane = threadIdx.x & 63;
    const bool in_lds = cnt > 64 && cnt <= CAP;
    const int n4 = (cnt + 3) >> 2;
    if (in_lds)
        for (int i = lane; i < 4 * n4; i += 64) lds[i] = i < cnt ? a[i] : INT_MAX;
    __syncthreads();
    if (cnt >= 1 && cnt <= 64) {
        const int v = lane < cnt ? a[lane] : INT_MAX;
        int rank = 0;
        for (int j = 0; j < cnt; j++) rank += __shfl(v, j) < v ? 1 : 0;
        if (lane < cnt) {
            a[rank] = v;
            on(v, rank);
        }
    } else if (in_lds) {
        const int4* l4 = reinterpret_cast<const int4*>(lds);
        for (int i = lane; i < cnt; i += 64) {
            const int v = lds[i];
            int rank = 0;
            for (int j = 0; j < n4; j++) {
                const int4 w = l4[j];
                rank += (w.x < v) + (w.y < v) + (w.z < v) + (w.w < v);
            }
            a[rank] = v;
            on(v, rank);
        }
    } else if (cnt > CAP && lane == 0) {
        for (int i = 1; i < cnt; i++) {
            const int v = a[i];
            int j = i - 1;
            while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; j--; }
            a[j + 1] = v;
        }
        for (int i = 0; i < cnt; i++) on(a[i], i);
    }
}

// The same for distinct keys key(v) (64-bit, unsigned order) instead of the
// entries themselves: up to 64 entries ranked in registers, up to CAP / 2 in
// this wave's share of LDS (two ints a key), a larger cell by lane 0.
template <int CAP, typename Key, typename Hook>
__device__ inline void mt_wave_sort_cell_keyed(int* __restrict__ a, int cnt, int* __restrict__ lds, Key key, Hook on)
{
    const int lane = threadIdx.x & 63;
    const bool in_lds = cnt > 64 && cnt <= CAP / 2;
    unsigned long long* lk = reinterpret_cast<unsigned long long*>(lds);
    if (in_lds)
        for (int i = lane; i < cnt; i += 64) lk[i] = key(a[i]);
    __syncthreads();
    if (cnt >= 1 && cnt <= 64) {
        const int v = lane < cnt ? a[lane] : INT_MAX;
        const unsigned long long kv = lane < cnt ? key(v) : ~0ull;
        const unsigned lo = (unsigned)kv, hi = (unsigned)(kv >> 32);
        int rank = 0;
        for (int j = 0; j < cnt; j++) {
            const unsigned long long kj = ((unsigned long long)(unsigned)__shfl((int)hi, j) << 32) |
                                          (unsigned)__shfl((int)lo, j);
            rank += kj < kv ? 1 : 0;
        }
        if (lane < cnt) {
            a[rank] = v;
            on(v, rank);
        }
    } else if (in_lds) {
        int vs[CAP / 128], rs[CAP / 128];
        int m = 0;
        for (int i = lane; i < cnt; i += 64, m++) {
            const unsigned long long kv = lk[i];
            int rank = 0;
            for (int j = 0; j < cnt; j++) rank += lk[j] < kv ? 1 : 0;
            vs[m] = a[i];
            rs[m] = rank;
        }
        __builtin_amdgcn_wave_barrier();          // (one wavefront: every read of a[] above precedes the writes)
        m = 0;
        for (int i = lane; i < cnt; i += 64, m++) {
            a[rs[m]] = vs[m];
            on(vs[m], rs[m]);
        }
    } else if (cnt > CAP / 2 && lane == 0) {
        for (int i = 1; i < cnt; i++) {
            const int v = a[i];
            const unsigned long long kv = key(v);
            int j = i - 1;
            while (j >= 0 && key(a[j]) > kv) { a[j + 1] = a[j]; j--; }
            a[j + 1] = v;
        }
        for (int i = 0; i < cnt; i++) on(a[i], i);
    }
}
