/*
 * mtmd_oracle.c - CPU restatement of MolTwister's `moldyn` hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see mtmd_oracle.h): the checker for libmtmd.so,
 * never part of the product. Compiled with -O2 -ffp-contract=off on x86-64
 * (SSE2 arithmetic, no FMA), like the reference CPU build.
 *
 * Reference paths are relative to the MolTwister source tree. Abbreviations:
 *   FCF = Cmd/MolTwisterCmdMolDyn/ForceFields/FunctorCalcForce.cpp
 *   FGC = Cmd/MolTwisterCmdMolDyn/ForceFields/FunctorGenCellList.cpp
 *   FGN = Cmd/MolTwisterCmdMolDyn/ForceFields/FunctorGenNeighList.cpp
 *   MFM = Cmd/MolTwisterCmdMolDyn/ForceFields/MDFFMatrices.cpp
 *   VV  = Cmd/MolTwisterCmdMolDyn/Integrators/VelVerlet.cpp
 *   NHC = Cmd/MolTwisterCmdMolDyn/Integrators/NHChain.cpp
 *   SB  = Cmd/MolTwisterCmdMolDyn/SimulationBox/SimulationBox.cpp
 *   MDL = Cmd/MolTwisterCmdMolDyn/MDLoop/MDLoop.cpp
 *   V3  = Utilities/3DVector.cpp
 */
#include <math.h>
#include <float.h>
#include <limits.h>
#include <stdlib.h>
#include <string.h>
#include "mtmd_oracle.h"

#define NPTS MTMD_NUM_POINTS
#define MAXFF MTMD_MAX_FF_PER_PAIR
#define CONV_T 120.272          /* Constants.h:27 */
#define CONV_t 1.0E2            /* Constants.h:30 */
#define CONV_PRESS 16388.2461   /* Constants.h:32 */
#define COULOMB_K 1389.32f      /* Constants.h:40 */
#define TO_RAD 0.017453292519943295 /* MolTwisterMDFFAngle_Harm.h:36 */

/* ------------------------------------------------------------------ */
/* 3-vectors: C3DVector arithmetic (V3:34-42,243-321)                  */
/* ------------------------------------------------------------------ */
typedef struct { double x, y, z; } v3;
static inline v3 V(double x, double y, double z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vmul(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static inline double vdot(v3 a, v3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
static inline double vnorm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline v3 vcross(v3 a, v3 b)
{
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* x86-64 cvttss2si / cvttsd2si semantics for int(float) / int(double):
 * truncation, and INT_MIN for NaN or out of range. */
static inline int f2i(float v)
{
    if (!(v >= -2147483648.0f && v < 2147483648.0f)) return INT_MIN;
    return (int)v;
}
static inline int d2i(double v)
{
    if (!(v > -2147483649.0 && v < 2147483648.0)) return INT_MIN;
    return (int)v;
}

/* C3DVector::moveToSameSideOfPBCAsThis (V3:133-165): shifts v by one box
 * width toward `self` on each axis where |self - v| > w/2. */
static inline v3 move_same_side(v3 self, v3 v, const double w[3])
{
    double dX = self.x - v.x, dY = self.y - v.y, dZ = self.z - v.z;
    dX = (dX < 0.0) ? -dX : dX;
    dY = (dY < 0.0) ? -dY : dY;
    dZ = (dZ < 0.0) ? -dZ : dZ;
    if (dX > (w[0] / 2.0)) { if (self.x > v.x) v.x += w[0]; else v.x -= w[0]; }
    if (dY > (w[1] / 2.0)) { if (self.y > v.y) v.y += w[1]; else v.y -= w[1]; }
    if (dZ > (w[2] / 2.0)) { if (self.z > v.z) v.z += w[2]; else v.z -= w[2]; }
    return v;
}

/* C3DVector::distToAcrossPBC2 (V3:68-90) */
static inline double dist2_pbc(v3 self, v3 v, const double w[3])
{
    double dX = self.x - v.x, dY = self.y - v.y, dZ = self.z - v.z;
    dX = (dX < 0.0) ? -dX : dX;
    dY = (dY < 0.0) ? -dY : dY;
    dZ = (dZ < 0.0) ? -dZ : dZ;
    if (dX > (w[0] / 2.0)) dX = w[0] - dX;
    if (dY > (w[1] / 2.0)) dY = w[1] - dY;
    if (dZ > (w[2] / 2.0)) dZ = w[2] - dZ;
    return dX * dX + dY * dY + dZ * dZ;
}

/* ------------------------------------------------------------------ */
/* Analytic force fields used to fill the tables                        */
/* ------------------------------------------------------------------ */

/* r12*RInv with the R == 0 guard: calcNonBondForceCoeffs12
 * (MolTwisterMDFFNonBonded.cpp:174-181), calcBondForceCoeffs12
 * (MolTwisterMDFFBond.cpp:181-188), FCF:283-301 */
static inline v3 coeffs12(v3 r1, v3 r2)
{
    v3 r12 = vsub(r2, r1);
    double R = vnorm(r12);
    double RInv = (R == 0.0) ? 1.0 / 1E-10 : 1.0 / R;
    return vmul(r12, RInv);
}

/* CMDFFNonBonded_*::calcForces -> f2 and ::calcPotential
 * LJ  MolTwisterMDFFNonBonded_LJ.cpp:131-162
 * Buck MolTwisterMDFFNonBonded_Buck.cpp:116-143
 * LJ1208 MolTwisterMDFFNonBonded_LJ1208.cpp:131-170 */
static v3 nb_force(int kind, const double* p, v3 r1, v3 r2)
{
    v3 r12 = vsub(r2, r1);
    double R = vnorm(r12);
    double RInv = (R == 0.0) ? 1.0 / 1E-10 : 1.0 / R;
    double RInv2 = RInv * RInv;
    double c = 0.0;
    if (kind == MTMD_NB_LJ) {
        double eps = p[0], sig = p[1], alpha = p[2];
        double RInv6 = RInv2 * RInv2 * RInv2;
        double RInv12 = RInv6 * RInv6;
        double s2 = sig * sig, s6 = s2 * s2 * s2, s12 = s6 * s6;
        c = 4.0 * eps * RInv2 * (12.0 * s12 * RInv12 - 6.0 * alpha * s6 * RInv6);
    } else if (kind == MTMD_NB_BUCK) {
        double A = p[0], rho = p[1], C = p[2];
        double RInv4 = RInv2 * RInv2;
        double RInv6 = RInv4 * RInv2;
        double rhoInv = (rho == 0.0) ? 1E-10 : 1.0 / rho;
        c = RInv2 * (A * R * rhoInv * exp(-R * rhoInv) - 6.0 * C * RInv6);
    } else if (kind == MTMD_NB_LJ1208) {
        double eps = p[0], sig = p[1], alpha = p[2];
        double RInv4 = RInv2 * RInv2;
        double RInv6 = RInv4 * RInv2;
        double RInv8 = RInv4 * RInv4;
        double RInv12 = RInv6 * RInv6;
        double s2 = sig * sig, s4 = s2 * s2, s6 = s4 * s2, s8 = s4 * s4, s12 = s6 * s6;
        c = 4.0 * eps * RInv2 * (12.0 * s12 * RInv12 - 8.0 * alpha * s8 * RInv8);
    } else if (kind == MTMD_NB_LJBUCK) {
        /* MolTwisterMDFFNonBonded_LJBuck.cpp:214-251: f1 = f2 = 0 at the end */
        return V(0.0, 0.0, 0.0);
    }
    return vmul(r12, c);
}

static double nb_pot(int kind, const double* p, v3 r1, v3 r2)
{
    v3 r12 = vsub(r2, r1);
    double R = vnorm(r12);
    double RInv = (R == 0.0) ? 1.0 / 1E-10 : 1.0 / R;
    double RInv2 = RInv * RInv;
    if (kind == MTMD_NB_LJ) {
        double eps = p[0], sig = p[1], alpha = p[2];
        double RInv6 = RInv2 * RInv2 * RInv2;
        double RInv12 = RInv6 * RInv6;
        double s2 = sig * sig, s6 = s2 * s2 * s2, s12 = s6 * s6;
        return 4.0 * eps * (s12 * RInv12 - alpha * s6 * RInv6);
    } else if (kind == MTMD_NB_BUCK) {
        double A = p[0], rho = p[1], C = p[2];
        double RInv4 = RInv2 * RInv2;
        double RInv6 = RInv4 * RInv2;
        double rhoInv = (rho == 0.0) ? 1E-10 : 1.0 / rho;
        return A * exp(-R * rhoInv) - C * RInv6;
    } else if (kind == MTMD_NB_LJ1208) {
        double eps = p[0], sig = p[1], alpha = p[2];
        double RInv4 = RInv2 * RInv2;
        double RInv6 = RInv4 * RInv2;
        double RInv8 = RInv4 * RInv4;
        double RInv12 = RInv6 * RInv6;
        double s2 = sig * sig, s4 = s2 * s2, s6 = s4 * s2, s8 = s4 * s4, s12 = s6 * s6;
        return 4.0 * eps * (s12 * RInv12 - alpha * s8 * RInv8);
    } else if (kind == MTMD_NB_LJBUCK) {
        /* MolTwisterMDFFNonBonded_LJBuck.cpp:181-211; p[0] selects LJ (0) or Buckingham (1) */
        if (p[0] == 0.0) {
            double eps = p[1], sig = p[2];
            double RInv6 = RInv2 * RInv2 * RInv2, RInv12 = RInv6 * RInv6;
            double s2 = sig * sig, s6 = s2 * s2 * s2, s12 = s6 * s6;
            return 4.0 * eps * (s12 * RInv12 - s6 * RInv6);
        }
        if (p[0] == 1.0) {
            double A = p[1], rho = p[2], C = p[3];
            double RInv4 = RInv2 * RInv2, RInv6 = RInv4 * RInv2;
            double rhoInv = (rho == 0.0) ? 1E-10 : 1.0 / rho;
            return A * exp(-R * rhoInv) - C * RInv6;
        }
    }
    return 0.0;
}

/* CMDFFBond_Harm (MolTwisterMDFFBond_Harm.cpp:90-110),
 * CMDFFBond_Morse (MolTwisterMDFFBond_Morse.cpp:95-117) */
static v3 bond_force(int kind, const double* p, v3 r1, v3 r2)
{
    v3 r12 = vsub(r2, r1);
    double R = vnorm(r12);
    double RInv = (R == 0.0) ? 1.0 / 1E-10 : 1.0 / R;
    double c = 0.0;
    if (kind == MTMD_BOND_HARM) {
        double k = p[0], r0 = p[1];
        double deltaR = R - r0;
        c = -2.0 * k * deltaR * RInv;
    } else if (kind == MTMD_BOND_MORSE) {
        double D = p[0], alpha = p[1], r0 = p[2];
        double deltaR = R - r0;
        double Exp = exp(-alpha * deltaR);
        c = -2.0 * D * alpha * RInv * Exp * (1.0 - Exp);
    }
    return vmul(r12, c);
}

static double bond_pot(int kind, const double* p, v3 r1, v3 r2)
{
    v3 r12 = vsub(r2, r1);
    double R = vnorm(r12);
    if (kind == MTMD_BOND_HARM) {
        double deltaR = R - p[1];
        return p[0] * deltaR * deltaR;
    } else if (kind == MTMD_BOND_MORSE) {
        double deltaR = R - p[2];
        double P = (1.0 - exp(-p[1] * deltaR));
        return p[0] * P * P;
    }
    return 0.0;
}

/* calcAngularForceCoeffs13: FCF:303-324, MolTwisterMDFFAngle.cpp:62-83 */
static v3 ang_coeffs13(v3 r1, v3 r2, v3 r3)
{
    v3 r12 = vsub(r1, r2);
    double R = vnorm(r12);
    double RInv = (R == 0.0) ? 1.0 / 1E-10 : 1.0 / R;
    v3 K = vmul(r12, RInv);
    v3 r32 = vsub(r3, r2);
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
    return V(dTheta_dW * dW_dx, dTheta_dW * dW_dy, dTheta_dW * dW_dz);
}

/* calcAngularForceCoeffs2: FCF:326-346 */
static v3 ang_coeffs2(v3 r1, v3 r2, v3 r3)
{
    v3 r12 = vsub(r1, r2);
    v3 r32 = vsub(r3, r2);
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
    return V(dTheta_dW * dW_dx, dTheta_dW * dW_dy, dTheta_dW * dW_dz);
}

/* calcDihedralForceCoeffs14 without the final sign flip:
 * MolTwisterMDFFDih.cpp:64-87; FCF:348-377 is the same followed by *= -1 */
static v3 dih_coeffs14(v3 r1, v3 r2, v3 r3, v3 r4)
{
    v3 r21 = vsub(r1, r2);
    v3 r23 = vsub(r3, r2);
    v3 r32 = vmul(r23, -1.0);
    v3 r34 = vsub(r4, r3);
    v3 n1 = vcross(r23, r21);
    double n1Norm = vnorm(n1);
    double n1NormInv = (n1Norm == 0.0) ? 1.0 / 1E-10 : 1.0 / n1Norm;
    v3 K = vmul(n1, n1NormInv);
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
    return V(dTheta_dW * dW_dx, dTheta_dW * dW_dy, dTheta_dW * dW_dz);
}

/* calcDihedralForceCoeffs23 without the final sign flip:
 * MolTwisterMDFFDih.cpp:89-134; FCF:379-424 flips the sign */
static v3 dih_coeffs23(v3 r1, v3 r2, v3 r3, v3 r4)
{
    v3 r21 = vsub(r1, r2);
    v3 r23 = vsub(r3, r2);
    v3 r32 = vmul(r23, -1.0);
    v3 r34 = vsub(r4, r3);
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
    double M1 = L1;
    double M2 = -K5;
    double M3 = L3;
    double M4 = K1;
    double M5 = -L5;
    double M6 = K3;
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
    return V(dTheta_dW * dW_dx, dTheta_dW * dW_dy, dTheta_dW * dW_dz);
}

/* CMDFFAngle_Harm::calcForces -> f3 and ::calcPotential
 * (MolTwisterMDFFAngle_Harm.cpp:94-130) */
static double angle_theta(v3 r1, v3 r2, v3 r3)
{
    v3 r21 = vsub(r1, r2);
    v3 r23 = vsub(r3, r2);
    double R21 = vnorm(r21), R23 = vnorm(r23);
    double R21Inv = (R21 == 0.0) ? 1.0 / 1E-10 : 1.0 / R21;
    double R23Inv = (R23 == 0.0) ? 1.0 / 1E-10 : 1.0 / R23;
    double cosTheta = vdot(r21, r23) * (R21Inv * R23Inv);
    return acos(cosTheta);
}

/* CMDFFDih_Harm::calcForces / ::calcPotential (MolTwisterMDFFDih_Harm.cpp:96-132) */
static double dihedral_phi(v3 r1, v3 r2, v3 r3, v3 r4)
{
    v3 r21 = vsub(r1, r2);
    v3 r23 = vsub(r3, r2);
    v3 r32 = vmul(r23, -1.0);
    v3 r34 = vsub(r4, r3);
    v3 n1 = vcross(r23, r21);
    v3 n2 = vcross(r32, r34);
    double n1n2 = vnorm(n1) * vnorm(n2);
    double cosPhi = (n1n2 == 0.0) ? vdot(n1, n2) / 1E-20 : vdot(n1, n2) / n1n2;
    return acos(cosPhi);
}

/* CMDFFAngle_Class2 (MolTwisterMDFFAngle_Class2.cpp:145-296) at (r1, r2, r3):
 * the force on r3 and the energy, the terms summed in the class's order.
 * q: theta0, K2, K3, K4 (Ea), M, r1, r2 (Ebb), N1, N2, r1, r2 (Eba) */
static double class2_dtheta(v3 r1, v3 r2, v3 r3, double theta0)
{
    /* Theta - Theta0 of calcPotentialHarm/calcForcesHarm (:145-190) */
    double R21 = vnorm(vsub(r1, r2)), R23 = vnorm(vsub(r3, r2));
    double R21Inv = (R21 == 0.0) ? 1.0 / 1E-10 : 1.0 / R21;
    double R23Inv = (R23 == 0.0) ? 1.0 / 1E-10 : 1.0 / R23;
    double cosTheta = vdot(vsub(r1, r2), vsub(r3, r2)) * (R21Inv * R23Inv);
    return acos(cosTheta) - theta0 * TO_RAD;
}

static v3 class2_f3(v3 r1, v3 r2, v3 r3, const double* q)
{
    v3 f3 = V(0.0, 0.0, 0.0), r12, r23, r21, u23, c13 = ang_coeffs13(r1, r2, r3);
    double R12, R23, RInv12, RInv23, coeff, cosTheta, dT;
    int degree, i;
    for (degree = 2; degree <= 4; degree++) {                         /* calcForcesHarm */
        double d = class2_dtheta(r1, r2, r3, q[0]);
        double dU = -(double)degree * q[degree - 1];
        for (i = 0; i < degree - 1; i++) dU *= d;
        f3 = vadd(f3, vmul(c13, dU));
    }
    r12 = vsub(r2, r1); r23 = vsub(r3, r2);
    R12 = vnorm(r12); R23 = vnorm(r23);
    RInv12 = (R12 == 0.0) ? 1.0 / 1E-10 : 1.0 / R12;
    RInv23 = (R23 == 0.0) ? 1.0 / 1E-10 : 1.0 / R23;
    u23 = vmul(r23, RInv23);
    f3 = vadd(f3, vmul(u23, (-1.0) * q[4] * (R12 - q[5])));           /* calcForcesCartesian */
    r21 = vmul(r12, -1.0);                                             /* calcForcesMix */
    coeff = q[7] * (R12 - q[9]) + q[8] * (R23 - q[10]);
    cosTheta = vdot(r21, r23) * (RInv12 * RInv23);
    dT = acos(cosTheta) - q[0] * TO_RAD;
    f3 = vadd(f3, vsub(vmul(u23, (-1.0) * q[8] * dT), vmul(c13, coeff)));
    return f3;
}

static double class2_u(v3 r1, v3 r2, v3 r3, const double* q)
{
    double U = 0.0, R12, R23, R21;
    int degree, i;
    for (degree = 2; degree <= 4; degree++) {                         /* calcPotentialHarm */
        double d = class2_dtheta(r1, r2, r3, q[0]), dN = 1.0;
        for (i = 0; i < degree; i++) dN *= d;
        U += q[degree - 1] * dN;
    }
    R12 = vnorm(vsub(r2, r1)); R23 = vnorm(vsub(r3, r2)); R21 = vnorm(vsub(r1, r2));
    U += q[4] * (R12 - q[5]) * (R23 - q[6]);                           /* calcPotentialCartesian */
    U += (q[7] * (R21 - q[9]) + q[8] * (R23 - q[10])) * class2_dtheta(r1, r2, r3, q[0]);   /* calcPotentialMix */
    return U;
}

/* ------------------------------------------------------------------ */
/* Profiles: calc1DForceProfile / calc1DPotentialProfile               */
/* ------------------------------------------------------------------ */
int orc_profile(int which, int kind, const double* params, float x0, float x1,
                int n, float* xo, float* fo, float* eo)
{
    double p[MTMD_BONDED_PARAMS];
    int i;
    if (n <= 0) return -1;
    for (i = 0; i < MTMD_BONDED_PARAMS; i++) p[i] = (which == 0 && i >= 4) ? 0.0 : params[i];
    float delta = (x1 - x0) / (float)(n - 1);
    for (i = 0; i < n; i++) {
        float x = x0 + (float)i * delta;
        double fval = 0.0, e = 0.0;
        if (which == 0 || which == 1) {
            /* MolTwisterMDFFNonBonded.cpp:126-172, MolTwisterMDFFBond.cpp:133-179 */
            v3 r1 = V(0.0, 0.0, 0.0), r2 = V((double)x, 0.0, 0.0);
            v3 f2 = (which == 0) ? nb_force(kind, p, r1, r2) : bond_force(kind, p, r1, r2);
            v3 c = coeffs12(r1, r2);
            fval = (c.x != 0.0) ? f2.x / c.x : 0.0;
            e = (which == 0) ? nb_pot(kind, p, r1, r2) : bond_pot(kind, p, r1, r2);
        } else if (which == 2) {
            /* MolTwisterMDFFAngle.cpp:210-257 with CMDFFAngle_Harm */
            v3 r1 = V(-1.0, 0.0, 0.0), r2 = V(0.0, 0.0, 0.0);
            v3 r3 = V(-cos((double)x), sin((double)x), 0.0);
            double Theta = angle_theta(r1, r2, r3);
            double deltaTheta = Theta - p[1] * TO_RAD;
            if (kind == MTMD_ANGLE_HARM) {
                double dU_dTheta = -2.0 * p[0] * deltaTheta;
                v3 f3 = vmul(ang_coeffs13(r1, r2, r3), dU_dTheta);
                v3 c = ang_coeffs13(r1, r2, r3);
                fval = (c.x != 0.0) ? f3.x / c.x : 0.0;
                e = p[0] * deltaTheta * deltaTheta;
            } else if (kind == MTMD_ANGLE_CLASS2) {
                v3 f3 = class2_f3(r1, r2, r3, p);
                v3 c = ang_coeffs13(r1, r2, r3);
                fval = (c.x != 0.0) ? f3.x / c.x : 0.0;
                e = class2_u(r1, r2, r3, p);
            }
        } else if (which == 3) {
            /* MolTwisterMDFFDih.cpp:272-320 with CMDFFDih_Harm */
            v3 r1 = V(0.0, 1.0, 0.0), r2 = V(0.0, 0.0, 0.0), r3 = V(0.0, 0.0, 1.0);
            v3 r4 = V(sin((double)x), cos((double)x), 1.0);
            if (kind == MTMD_DIH_HARM) {
                double K = p[0];
                int D = (int)p[1], N = (int)p[2];
                double phi = dihedral_phi(r1, r2, r3, r4);
                double dU_dPhi = -K * (double)(D * N) * sin((double)N * phi);
                v3 f4 = vmul(dih_coeffs14(r1, r2, r3, r4), dU_dPhi);
                v3 c = dih_coeffs14(r1, r2, r3, r4);
                fval = (c.x != 0.0) ? f4.x / c.x : 0.0;
                e = K * (1.0 + (double)D * cos((double)N * phi));
            } else if (kind == MTMD_DIH_FOURIER4T) {
                /* MolTwisterMDFFDih_Fourier4t.cpp:98-142 */
                double phi = dihedral_phi(r1, r2, r3, r4);
                double d1 = -0.5 * p[0] * sin(phi), d2 = p[1] * sin(2.0 * phi);
                double d3 = -1.5 * p[2] * sin(3.0 * phi), d4 = 2.0 * p[3] * sin(4.0 * phi);
                v3 f4 = vmul(dih_coeffs14(r1, r2, r3, r4), d1 + d2 + d3 + d4);
                v3 c = dih_coeffs14(r1, r2, r3, r4);
                double U1 = p[0] * (1.0 + cos(phi)), U2 = p[1] * (1.0 - cos(2.0 * phi));
                double U3 = p[2] * (1.0 + cos(3.0 * phi)), U4 = p[3] * (1.0 - cos(4.0 * phi));
                fval = (c.x != 0.0) ? f4.x / c.x : 0.0;
                e = 0.5 * (U1 + U2 + U3 + U4);
            }
        } else {
            return -1;
        }
        if (xo) xo[i] = x;
        if (fo) fo[i] = (float)fval;
        if (eo) eo[i] = (float)e;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Engine state                                                         */
/* ------------------------------------------------------------------ */
typedef struct { float x, f, e; } pt;   /* CMDFFMatrices::CPoint */

struct orc_md {
    int n, ntypes;
    int threads;       /* host threads of the per-atom loops (1: sequential) */
    v3 *r, *p;
    double* m;
    float* q;
    int* type;
    int* mol;
    unsigned char* mobile;
    double L[3];
    int ensemble;
    /* tables */
    pt* nb;            /* [((row*T + col)*MAXFF + ff)*NPTS + pt] */
    int* nbcount;      /* [row*T + col] */
    pt *bondtab, *angtab, *dihtab;
    /* per-atom bonded lists (MFM:384-520) */
    int *bfirst, *bnum, *bj, *btype;
    int *afirst, *anum, *ai, *aj, *atype; unsigned char* acenter;
    int *dfirst, *dnum, *di, *dj, *dl, *dtype; unsigned char* dcenter;
    /* cells / neighbours */
    float rc, shell;
    int nc[3];
    int *cix, *cfirst, *ccount, *clist;
    int maxn;
    int *ncount, *nlist;
    /* functor parameters */
    float scale12, scale13, scale14, scale1N;
    int coulomb;        /* MTMD_COUL_ALLPAIRS or MTMD_COUL_CUTOFF */
    int all_q_zero;
    /* integrator */
    double dt, Fcut;
    int have_forces;
    v3* F;
    float* U;
    /* NH thermostat (NHChain.h:55-66) */
    int M, nresp;
    double T, tau;
    double *eta, *peta, *Q;
    double w[3];
    /* NPT: barostat (VelVerlet.h: P_, p_eps_, eps_, W_, V0_) and its chain NH_P_ */
    double Lmax[3];
    double Pext, W, V0, eps, peps;
    int Mp, nrespP;
    double tauP;
    double *etaP, *petaP, *QP;
};

int orc_natoms(const orc_md* o) { return o->n; }

void orc_set_threads(orc_md* o, int threads) { o->threads = threads > 1 ? threads : 1; }

/* CMt::exp / CMt::sinhXoverX (Cmd/MolTwisterCmdMolDyn/Integrators/Math.cpp:27-54) */
double orc_cmt_exp(double x)
{
    if (x > 1.0) return 2.71828 * x;
    return exp(x);
}

double orc_cmt_sinhxoverx(double x)
{
    static const double a[6] = { 1.0, 1.0 / 6.0, 1.0 / 120.0, 1.0 / 5040.0, 1.0 / 362880.0, 1.0 / 39916800.0 };
    double sum = 0.0, xn = 1.0;
    int n;
    for (n = 0; n < 6; n++) { sum += a[n] * xn; xn *= (x * x); }
    return sum;
}

static uint64_t splitmix64(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
/* uniform double in [0, 1) with 53 random bits */
static double urand(uint64_t* s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* CSimulationBox::calcTemp (SB:168-180): sum p.p/m / (dim N), reduced units */
static double calc_temp_reduced(const orc_md* o)
{
    double T = 0.0;
    int k;
    for (k = 0; k < o->n; k++) T += vdot(o->p[k], o->p[k]) / o->m[k];
    return T / (3 * (double)o->n);
}

/* mol indices: connected components over the bond graph (MolecularSystemTools.cpp:238-280);
 * only equality of labels matters to the force functor (FCF:147-156) */
static int uf_find(int* par, int a) { while (par[a] != a) { par[a] = par[par[a]]; a = par[a]; } return a; }

/* a table from parameters (orc_profile) or, for MTMD_FF_PROFILE, the
 * caller's own points (what MDFFMatrices.cpp:270-382 copies in) */
static void orc_table(int which, int kind, const double* p, const float* profile, float x0, float x1, float* xs,
                      float* fs, float* es)
{
    int j;
    if (kind == MTMD_FF_PROFILE && profile) {
        for (j = 0; j < NPTS; j++) { xs[j] = profile[j]; fs[j] = profile[NPTS + j]; es[j] = profile[2 * NPTS + j]; }
        return;
    }
    orc_profile(which, kind, p, x0, x1, NPTS, xs, fs, es);
}

orc_md* orc_create(const mtmd_system* s, const mtmd_config* cfg)
{
    int n = s->n_atoms, T = s->n_types, k, j, i, c, r;
    orc_md* o = (orc_md*)calloc(1, sizeof(orc_md));
    o->n = n; o->ntypes = T;
    o->threads = 1;
    o->r = (v3*)calloc(n, sizeof(v3));
    o->p = (v3*)calloc(n, sizeof(v3));
    o->m = (double*)calloc(n, sizeof(double));
    o->q = (float*)calloc(n, sizeof(float));
    o->type = (int*)calloc(n, sizeof(int));
    o->mol = (int*)calloc(n, sizeof(int));
    o->mobile = (unsigned char*)calloc(n, 1);
    o->F = (v3*)calloc(n, sizeof(v3));
    o->U = (float*)calloc(n, sizeof(float));
    o->L[0] = s->box[0]; o->L[1] = s->box[1]; o->L[2] = s->box[2];
    o->all_q_zero = 1;
    for (k = 0; k < n; k++) {
        o->r[k] = V(s->pos[3 * k], s->pos[3 * k + 1], s->pos[3 * k + 2]);
        o->m[k] = s->mass[k];
        o->q[k] = (float)s->charge[k];
        if (o->q[k] != 0.0f) o->all_q_zero = 0;
        o->type[k] = s->type[k];
        o->mobile[k] = s->mobile ? s->mobile[k] : 1;
    }

    o->rc = (float)cfg->cutoff_radius;
    o->shell = (float)cfg->neigh_shell;
    o->scale12 = (float)cfg->scale12; o->scale13 = (float)cfg->scale13;
    o->scale14 = (float)cfg->scale14; o->scale1N = (float)cfg->scale1N;
    o->coulomb = (cfg->coulomb == MTMD_COUL_CUTOFF) ? MTMD_COUL_CUTOFF : MTMD_COUL_ALLPAIRS;
    o->ensemble = cfg->ensemble;
    o->dt = cfg->timestep / CONV_t;
    o->Fcut = cfg->cutoff_force;

    /* non-bonded table matrix (MFM:249-296) */
    o->nb = (pt*)calloc((size_t)T * T * MAXFF * NPTS, sizeof(pt));
    o->nbcount = (int*)calloc((size_t)T * T, sizeof(int));
    for (c = 0; c < T; c++)
        for (r = 0; r < T; r++) {
            int cnt = 0;
            for (i = 0; i < s->n_nonbonded; i++) {
                const mtmd_nonbonded_ff* e = &s->nonbonded[i];
                int match = (e->type_a == c && e->type_b == r) || (e->type_b == c && e->type_a == r);
                if (!match || cnt >= MAXFF) continue;
                float xs[NPTS], fs[NPTS], es[NPTS];
                orc_table(0, e->kind, e->p, e->profile, 0.01f, o->rc, xs, fs, es);
                for (j = 0; j < NPTS; j++) {
                    pt* d = &o->nb[(((size_t)r * T + c) * MAXFF + cnt) * NPTS + j];
                    d->x = xs[j]; d->f = fs[j]; d->e = es[j];
                }
                cnt++;
            }
            o->nbcount[r * T + c] = cnt;
        }
    /* bonded tables (MFM:298-382) */
    o->bondtab = (pt*)calloc((size_t)(s->n_bond_ff + 1) * NPTS, sizeof(pt));
    o->angtab = (pt*)calloc((size_t)(s->n_angle_ff + 1) * NPTS, sizeof(pt));
    o->dihtab = (pt*)calloc((size_t)(s->n_dihedral_ff + 1) * NPTS, sizeof(pt));
    for (i = 0; i < s->n_bond_ff + s->n_angle_ff + s->n_dihedral_ff; i++) {
        float xs[NPTS], fs[NPTS], es[NPTS];
        pt* dst;
        if (i < s->n_bond_ff) {
            orc_table(1, s->bond_ff[i].kind, s->bond_ff[i].p, s->bond_ff[i].profile, 0.01f, o->rc, xs, fs, es);
            dst = o->bondtab + (size_t)i * NPTS;
        } else if (i < s->n_bond_ff + s->n_angle_ff) {
            int t = i - s->n_bond_ff;
            orc_table(2, s->angle_ff[t].kind, s->angle_ff[t].p, s->angle_ff[t].profile, 0.0f, (float)(2.0 * M_PI),
                      xs, fs, es);
            dst = o->angtab + (size_t)t * NPTS;
        } else {
            int t = i - s->n_bond_ff - s->n_angle_ff;
            orc_table(3, s->dihedral_ff[t].kind, s->dihedral_ff[t].p, s->dihedral_ff[t].profile, 0.0f,
                      (float)(2.0 * M_PI), xs, fs, es);
            dst = o->dihtab + (size_t)t * NPTS;
        }
        for (j = 0; j < NPTS; j++) { dst[j].x = xs[j]; dst[j].f = fs[j]; dst[j].e = es[j]; }
    }

    /* molecule indices from all bonds */
    {
        int* par = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
        for (k = 0; k < n; k++) par[k] = k;
        for (j = 0; j < s->n_bonds; j++) {
            int a = s->bonds[3 * j], b = s->bonds[3 * j + 1];
            if (a < 0 || b < 0 || a >= n || b >= n) continue;
            a = uf_find(par, a); b = uf_find(par, b);
            if (a != b) par[a < b ? b : a] = a < b ? a : b;
        }
        for (k = 0; k < n; k++) o->mol[k] = uf_find(par, k);
        free(par);
    }

    /* per-atom bond lists (MFM:384-416) */
    o->bfirst = (int*)calloc(n, sizeof(int)); o->bnum = (int*)calloc(n, sizeof(int));
    o->bj = (int*)calloc(2 * s->n_bonds + 1, sizeof(int)); o->btype = (int*)calloc(2 * s->n_bonds + 1, sizeof(int));
    {
        int g = 0;
        for (k = 0; k < n; k++) {
            int cnt = 0;
            o->bfirst[k] = g;
            for (j = 0; j < s->n_bonds; j++) {
                int a1 = s->bonds[3 * j], a2 = s->bonds[3 * j + 1], t = s->bonds[3 * j + 2];
                int to = -1;
                if (a1 == k) to = a2;
                if (a2 == k) to = a1;
                if (to != -1 && t != -1) { o->bj[g] = to; o->btype[g] = t; g++; cnt++; }
            }
            o->bnum[k] = cnt;
        }
    }
    /* per-atom angle lists (MFM:418-462) */
    o->afirst = (int*)calloc(n, sizeof(int)); o->anum = (int*)calloc(n, sizeof(int));
    o->ai = (int*)calloc(3 * s->n_angles + 1, sizeof(int)); o->aj = (int*)calloc(3 * s->n_angles + 1, sizeof(int));
    o->atype = (int*)calloc(3 * s->n_angles + 1, sizeof(int)); o->acenter = (unsigned char*)calloc(3 * s->n_angles + 1, 1);
    {
        int g = 0;
        for (k = 0; k < n; k++) {
            int cnt = 0;
            o->afirst[k] = g;
            for (j = 0; j < s->n_angles; j++) {
                const int* a = s->angles + 4 * j;
                int I = -1, J = -1, center = 0;
                if (a[0] == k) { I = a[1]; J = a[2]; }
                if (a[1] == k) { center = 1; I = a[0]; J = a[2]; }
                if (a[2] == k) { I = a[1]; J = a[0]; }
                if (I != -1 && J != -1 && a[3] != -1) {
                    o->ai[g] = I; o->aj[g] = J; o->atype[g] = a[3]; o->acenter[g] = (unsigned char)center;
                    g++; cnt++;
                }
            }
            o->anum[k] = cnt;
        }
    }
    /* per-atom dihedral lists (MFM:464-520) */
    o->dfirst = (int*)calloc(n, sizeof(int)); o->dnum = (int*)calloc(n, sizeof(int));
    o->di = (int*)calloc(4 * s->n_dihedrals + 1, sizeof(int)); o->dj = (int*)calloc(4 * s->n_dihedrals + 1, sizeof(int));
    o->dl = (int*)calloc(4 * s->n_dihedrals + 1, sizeof(int)); o->dtype = (int*)calloc(4 * s->n_dihedrals + 1, sizeof(int));
    o->dcenter = (unsigned char*)calloc(4 * s->n_dihedrals + 1, 1);
    {
        int g = 0;
        for (k = 0; k < n; k++) {
            int cnt = 0;
            o->dfirst[k] = g;
            for (j = 0; j < s->n_dihedrals; j++) {
                const int* d = s->dihedrals + 5 * j;
                int I = -1, J = -1, Lx = -1, center = 0;
                if (d[0] == k) { I = d[1]; J = d[2]; Lx = d[3]; }
                if (d[1] == k) { center = 1; I = d[0]; J = d[2]; Lx = d[3]; }
                if (d[2] == k) { center = 1; I = d[3]; J = d[1]; Lx = d[0]; }
                if (d[3] == k) { I = d[2]; J = d[1]; Lx = d[0]; }
                if (I != -1 && J != -1 && Lx != -1 && d[4] != -1) {
                    o->di[g] = I; o->dj[g] = J; o->dl[g] = Lx; o->dtype[g] = d[4];
                    o->dcenter[g] = (unsigned char)center;
                    g++; cnt++;
                }
            }
            o->dnum[k] = cnt;
        }
    }

    /* cell grid from the initial box (MFM:43-58) */
    {
        float R = o->rc + o->shell;
        for (i = 0; i < 3; i++) {
            float wf = (float)o->L[i];
            o->nc[i] = (int)floor((double)(wf / R));
        }
    }
    o->cix = (int*)calloc(3 * (size_t)n + 3, sizeof(int));
    {
        size_t ncell = (size_t)(o->nc[0] > 0 ? o->nc[0] : 0) * (o->nc[1] > 0 ? o->nc[1] : 0) * (o->nc[2] > 0 ? o->nc[2] : 0);
        o->cfirst = (int*)calloc(ncell + 1, sizeof(int));
        o->ccount = (int*)calloc(ncell + 1, sizeof(int));
    }
    o->clist = (int*)calloc((size_t)n + 1, sizeof(int));
    o->maxn = 0; o->ncount = (int*)calloc((size_t)n + 1, sizeof(int)); o->nlist = NULL;

    /* NH thermostat (SB:53-79, NHC:32-52,115-124) */
    o->M = cfg->temperature_nh_len;
    o->nresp = cfg->temperature_respa;
    o->T = cfg->temperature / CONV_T;
    o->tau = cfg->temperature_relax / CONV_t;
    {
        const double w0 = 1.0 / (1.0 - pow(2.0, 1.0 / 3.0));
        o->w[0] = w0; o->w[1] = 1.0 - 2.0 * w0; o->w[2] = w0;
    }
    o->eta = (double*)calloc(o->M + 1, sizeof(double));
    o->peta = (double*)calloc(o->M + 1, sizeof(double));
    o->Q = (double*)calloc(o->M + 1, sizeof(double));
    if (o->M > 0) {
        o->Q[0] = (double)(3 * n) * o->T * o->tau * o->tau;
        for (j = 1; j < o->M; j++) o->Q[j] = o->T * o->tau * o->tau;
    }

    /* NPT barostat (SB:53-79, VV:31-43, NHC:115-124 with prepareArrays(1, 1)):
     * W = T tau_P^2, P in reduced units, V0 = Lx Ly Lz, eps = 0; p_eps and
     * the pressure chain start at rest (orc_set_barostat sets any state) */
    o->Lmax[0] = s->box[0]; o->Lmax[1] = s->box[1]; o->Lmax[2] = s->box[2];
    o->Mp = (o->ensemble == MTMD_NPT) ? cfg->pressure_nh_len : 0;
    o->nrespP = cfg->pressure_respa;
    o->tauP = cfg->pressure_relax / CONV_t;
    o->W = o->T * o->tauP * o->tauP;
    o->Pext = cfg->pressure / CONV_PRESS;
    o->V0 = o->Lmax[0] * o->Lmax[1] * o->Lmax[2];
    o->eps = 0.0;
    o->peps = 0.0;
    o->etaP = (double*)calloc(o->Mp + 1, sizeof(double));
    o->petaP = (double*)calloc(o->Mp + 1, sizeof(double));
    o->QP = (double*)calloc(o->Mp + 1, sizeof(double));
    if (o->Mp > 0) {
        o->QP[0] = (double)(1 * 1) * o->T * o->tauP * o->tauP;
        for (j = 1; j < o->Mp; j++) o->QP[j] = o->T * o->tauP * o->tauP;
    }

    /* momenta: given, or uniform in [-1,1] per component, immobile zeroed,
     * rescaled to T (SB:98-133) */
    if (s->mom) {
        for (k = 0; k < n; k++) o->p[k] = V(s->mom[3 * k], s->mom[3 * k + 1], s->mom[3 * k + 2]);
    } else {
        uint64_t st = s->seed;
        double Tnow, ps;
        for (k = 0; k < n; k++) {
            double px = 2.0 * urand(&st) - 1.0, py = 2.0 * urand(&st) - 1.0, pz = 2.0 * urand(&st) - 1.0;
            if (px == 0.0) px = 1.0;
            if (py == 0.0) py = 1.0;
            if (pz == 0.0) pz = 1.0;
            o->p[k] = V(px, py, pz);
        }
        for (k = 0; k < n; k++) if (!o->mobile[k]) o->p[k] = V(0.0, 0.0, 0.0);
        Tnow = calc_temp_reduced(o);
        ps = sqrt(o->T / Tnow);
        for (k = 0; k < n; k++) { o->p[k].x *= ps; o->p[k].y *= ps; o->p[k].z *= ps; }
    }
    return o;
}

void orc_destroy(orc_md* o)
{
    if (!o) return;
    free(o->r); free(o->p); free(o->m); free(o->q); free(o->type); free(o->mol); free(o->mobile);
    free(o->F); free(o->U); free(o->nb); free(o->nbcount); free(o->bondtab); free(o->angtab); free(o->dihtab);
    free(o->bfirst); free(o->bnum); free(o->bj); free(o->btype);
    free(o->afirst); free(o->anum); free(o->ai); free(o->aj); free(o->atype); free(o->acenter);
    free(o->dfirst); free(o->dnum); free(o->di); free(o->dj); free(o->dl); free(o->dtype); free(o->dcenter);
    free(o->cix); free(o->cfirst); free(o->ccount); free(o->clist); free(o->ncount); free(o->nlist);
    free(o->eta); free(o->peta); free(o->Q);
    free(o->etaP); free(o->petaP); free(o->QP);
    free(o);
}

void orc_set_state(orc_md* o, const double* pos, const double* mom)
{
    int k;
    for (k = 0; k < o->n; k++) {
        if (pos) o->r[k] = V(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]);
        if (mom) o->p[k] = V(mom[3 * k], mom[3 * k + 1], mom[3 * k + 2]);
    }
    if (pos) o->have_forces = 0;
}

void orc_get_state(const orc_md* o, double* pos, double* mom)
{
    int k;
    for (k = 0; k < o->n; k++) {
        if (pos) { pos[3 * k] = o->r[k].x; pos[3 * k + 1] = o->r[k].y; pos[3 * k + 2] = o->r[k].z; }
        if (mom) { mom[3 * k] = o->p[k].x; mom[3 * k + 1] = o->p[k].y; mom[3 * k + 2] = o->p[k].z; }
    }
}

void orc_set_thermostat(orc_md* o, const double* eta, const double* p_eta)
{
    int j;
    for (j = 0; j < o->M; j++) { if (eta) o->eta[j] = eta[j]; if (p_eta) o->peta[j] = p_eta[j]; }
}

void orc_get_thermostat(const orc_md* o, double* eta, double* p_eta)
{
    int j;
    for (j = 0; j < o->M; j++) { if (eta) eta[j] = o->eta[j]; if (p_eta) p_eta[j] = o->peta[j]; }
}

void orc_set_barostat(orc_md* o, double eps, double p_eps, const double* eta, const double* p_eta)
{
    int j;
    o->eps = eps;
    o->peps = p_eps;
    for (j = 0; j < o->Mp; j++) { if (eta) o->etaP[j] = eta[j]; if (p_eta) o->petaP[j] = p_eta[j]; }
}

void orc_get_barostat(const orc_md* o, double* eps, double* p_eps, double* eta, double* p_eta, double* box)
{
    int j;
    if (eps) *eps = o->eps;
    if (p_eps) *p_eps = o->peps;
    for (j = 0; j < o->Mp; j++) { if (eta) eta[j] = o->etaP[j]; if (p_eta) p_eta[j] = o->petaP[j]; }
    if (box) { box[0] = o->L[0]; box[1] = o->L[1]; box[2] = o->L[2]; }
}

void orc_get_mol(const orc_md* o, int32_t* mol) { memcpy(mol, o->mol, sizeof(int) * o->n); }

/* ------------------------------------------------------------------ */
/* Cell list: CCellList::updatePBC (MFM:67-76), CFunctorGenCellList    */
/* (FGC:49-73), kernelAssembleList (FGC:75-142)                         */
/* ------------------------------------------------------------------ */
int orc_build_cells(orc_md* o, int32_t* ncells_xyz, int32_t* ixyz,
                    int32_t* cell_first, int32_t* cell_count, int32_t* cell_list)
{
    int k, i, n = o->n;
    int Nx = o->nc[0], Ny = o->nc[1], Nz = o->nc[2];
    int total = Nx * Ny * Nz;
    float wf[3], lf[3];
    int err = 0, next = 0;
    if (Nx < 1 || Ny < 1 || Nz < 1) return -1;   /* the reference indexes out of bounds here */
    for (i = 0; i < 3; i++) { wf[i] = (float)o->L[i]; lf[i] = -wf[i] / 2.0f; }
    for (i = 0; i < total; i++) { o->cfirst[i] = -1; o->ccount[i] = 0; }   /* resetCellList, MFM:78-85 */
    for (k = 0; k < n; k++) {
        v3 r = o->r[k];
        int ix = d2i(floor((r.x - (double)lf[0]) * (double)Nx / (double)wf[0]));
        int iy = d2i(floor((r.y - (double)lf[1]) * (double)Ny / (double)wf[1]));
        int iz = d2i(floor((r.z - (double)lf[2]) * (double)Nz / (double)wf[2]));
        ix = (ix < 0) ? 0 : ((ix >= Nx) ? Nx - 1 : ix);
        iy = (iy < 0) ? 0 : ((iy >= Ny) ? Ny - 1 : iy);
        iz = (iz < 0) ? 0 : ((iz >= Nz) ? Nz - 1 : iz);
        o->cix[3 * k] = ix; o->cix[3 * k + 1] = iy; o->cix[3 * k + 2] = iz;
    }
    for (k = 0; k < n; k++) {
        int c = o->cix[3 * k] + Nx * (o->cix[3 * k + 1] + o->cix[3 * k + 2] * Ny);
        o->ccount[c]++;
    }
    for (i = 0; i < total; i++)
        if (o->ccount[i] != 0) { o->cfirst[i] = next; next += o->ccount[i]; o->ccount[i] = 0; }
    for (k = 0; k < n; k++) {
        int c = o->cix[3 * k] + Nx * (o->cix[3 * k + 1] + o->cix[3 * k + 2] * Ny);
        if (c < total) {
            if (o->cfirst[c] != -1) {
                int e = o->cfirst[c] + o->ccount[c];
                if (e < n) { o->clist[e] = k; o->ccount[c]++; } else err = 2;
            }
        } else err = 1;
    }
    if (ncells_xyz) { ncells_xyz[0] = Nx; ncells_xyz[1] = Ny; ncells_xyz[2] = Nz; }
    if (ixyz) memcpy(ixyz, o->cix, sizeof(int) * 3 * (size_t)n);
    if (cell_first) memcpy(cell_first, o->cfirst, sizeof(int) * (size_t)total);
    if (cell_count) memcpy(cell_count, o->ccount, sizeof(int) * (size_t)total);
    if (cell_list) memcpy(cell_list, o->clist, sizeof(int) * (size_t)n);
    return err;
}

/* ------------------------------------------------------------------ */
/* Neighbour list: CFunctorGenNeighList::operator() (FGN:72-151).       */
/* Rows are kept complete (the reference's grow-and-retry loop,         */
/* MFM:161-179, changes only the capacity, not the content).            */
/* ------------------------------------------------------------------ */
static int neigh_row(const orc_md* o, int k, int* out, int cap)
{
    int Nx = o->nc[0], Ny = o->nc[1], Nz = o->nc[2];
    float wf[3], lf[3];
    double w[3];
    float rc2 = o->rc;
    int i, cnt = 0, ix, iy, iz;
    rc2 *= rc2;
    for (i = 0; i < 3; i++) {
        wf[i] = (float)o->L[i]; lf[i] = -wf[i] / 2.0f;
        w[i] = (double)(lf[i] + wf[i]) - (double)lf[i];
    }
    for (ix = o->cix[3 * k] - 1; ix <= o->cix[3 * k] + 1; ix++) {
        int Ix = (ix >= 0) ? ((ix < Nx) ? ix : 0) : Nx - 1;
        for (iy = o->cix[3 * k + 1] - 1; iy <= o->cix[3 * k + 1] + 1; iy++) {
            int Iy = (iy >= 0) ? ((iy < Ny) ? iy : 0) : Ny - 1;
            for (iz = o->cix[3 * k + 2] - 1; iz <= o->cix[3 * k + 2] + 1; iz++) {
                int Iz = (iz >= 0) ? ((iz < Nz) ? iz : 0) : Nz - 1;
                int c = Ix + Nx * (Iy + Iz * Ny);
                int num = o->ccount[c], first = o->cfirst[c], e;
                for (e = 0; e < num; e++) {
                    int a = o->clist[first + e];
                    double d2 = dist2_pbc(o->r[k], o->r[a], w);
                    if (d2 < (double)rc2) {
                        if (a == k) continue;
                        if (cnt < cap) out[cnt] = a;
                        cnt++;
                    }
                }
            }
        }
    }
    return cnt;
}

int orc_build_neighbors(orc_md* o, int32_t max_per_atom, int32_t* count, int32_t* list)
{
    int k, mx = 0;
    int* tmp;
    /* rows are independent: with orc_set_threads(t > 1) they are built by t
     * threads, each row by the same operations in the same order */
#pragma omp parallel for schedule(dynamic, 256) reduction(max : mx) if (o->threads > 1) num_threads(o->threads)
    for (k = 0; k < o->n; k++) {
        int c = neigh_row(o, k, NULL, 0);
        o->ncount[k] = c;
        if (c > mx) mx = c;
    }
    free(o->nlist);
    o->maxn = mx > 0 ? mx : 1;
    o->nlist = (int*)calloc((size_t)o->n * o->maxn, sizeof(int));
#pragma omp parallel for schedule(dynamic, 256) if (o->threads > 1) num_threads(o->threads)
    for (k = 0; k < o->n; k++) neigh_row(o, k, o->nlist + (size_t)k * o->maxn, o->maxn);
    if (count) memcpy(count, o->ncount, sizeof(int) * (size_t)o->n);
    if (list) {
        tmp = (int*)list;
        for (k = 0; k < o->n; k++) {
            int c = o->ncount[k] < max_per_atom ? o->ncount[k] : max_per_atom;
            memcpy(tmp + (size_t)k * max_per_atom, o->nlist + (size_t)k * o->maxn, sizeof(int) * (size_t)c);
        }
    }
    return mx;
}

/* ------------------------------------------------------------------ */
/* Force functor (FCF:95-281, 426-709)                                  */
/* ------------------------------------------------------------------ */

/* calcIndexOfLowerPointOfInterpolationLine (FCF:426-430) */
static inline int lower_index(float x, const pt* first, const pt* last)
{
    float delta = (last->x - first->x) / (float)(NPTS - 1);
    return f2i((x - first->x) / delta);
}

/* interpolateLinear (FCF:457-465) */
static inline float interp(float r, float x1, float y1, float x2, float y2)
{
    float denom = x1 - x2;
    float a, b;
    if (denom == 0.0f) denom = 3.0f * FLT_MIN;
    a = (y1 - y2) / denom;
    b = y1 - a * x1;
    return a * r + b;
}

/* calcAngle (FCF:432-440) */
static float calc_angle(v3 rk, v3 ri, v3 rj)
{
    v3 rik = vsub(rk, ri), rij = vsub(rj, ri);
    float dot = (float)vdot(rik, rij);
    float RR = (float)vnorm(rik) * (float)vnorm(rij);
    float cs = (RR != 0.0f) ? dot / RR : 1.0f;
    return (float)acos((double)cs);
}

/* calcDihedral (FCF:442-455) */
static float calc_dihedral(v3 rk, v3 ri, v3 rj, v3 rl)
{
    v3 rik = vsub(rk, ri), rij = vsub(rj, ri), rjl = vsub(rl, rj);
    v3 n1 = vcross(rik, rij), n2 = vcross(rij, rjl);
    float dot = (float)vdot(n1, n2);
    float nn = (float)vnorm(n1) * (float)vnorm(n2);
    float cs = (nn != 0.0f) ? dot / nn : 1.0f;
    if (cs < -1.0f) cs = -1.0f;
    if (cs > 1.0f) cs = 1.0f;
    return (float)acos((double)cs);
}

/* table force on r_k along r_k - r_i, optionally energy: the shared body of
 * calcForceNonBondedOn_r_k per FF entry (FCF:481-504) and calcForceBondOn_r_k
 * (FCF:547-567) */
static inline void table_eval(const pt* tab, float x, int want_e, float* fsum, float* esum)
{
    int idx = lower_index(x, &tab[0], &tab[NPTS - 1]);
    if (idx >= 0 && idx < NPTS - 1) {
        const pt* p1 = &tab[idx];
        const pt* p2 = &tab[idx + 1];
        *fsum += interp(x, p1->x, p1->f, p2->x, p2->f);
        if (want_e) *esum += interp(x, p1->x, p1->e, p2->x, p2->e);
    }
}

/* scaleForcesAndPotentials (FCF:95-157) */
static void scale_pair(const orc_md* o, v3* f, float* u, int i, int molk, int k)
{
    int j, hasB = 0, hasA = 0, hasD = 0;
    for (j = 0; j < o->bnum[k]; j++)
        if (i == o->bj[o->bfirst[k] + j]) {
            *f = vmul(*f, (double)o->scale12); *u *= o->scale12; hasB = 1; break;
        }
    if (!hasB)
        for (j = 0; j < o->anum[k]; j++) {
            int g = o->afirst[k] + j;
            if (!o->acenter[g] && i == o->aj[g]) {
                *f = vmul(*f, (double)o->scale13); *u *= o->scale13; hasA = 1; break;
            }
        }
    if (!hasB && !hasA)
        for (j = 0; j < o->dnum[k]; j++) {
            int g = o->dfirst[k] + j;
            if (!o->dcenter[g] && i == o->dl[g]) {
                *f = vmul(*f, (double)o->scale14); *u *= o->scale14; hasD = 1; break;
            }
        }
    if (!hasB && !hasA && !hasD) {
        int moli = o->mol[i];
        if (moli >= 0 && molk >= 0 && moli == molk) { *f = vmul(*f, (double)o->scale1N); *u *= o->scale1N; }
    }
}

/* calcForceCoulombOn_r_k (FCF:514-535) */
static inline v3 coulomb_pair(const orc_md* o, v3 rk, v3 ri, int k, int i, float* U)
{
    v3 r = vsub(rk, ri);
    float qi = o->q[i], qk = o->q[k];
    float R = (float)vnorm(r);
    double R3 = (double)(R * R * R);
    double K = (double)(COULOMB_K * qk * qi);
    if (i > k && R != 0.0f) *U += (float)K / R;
    if (R3 != 0.0) return vmul(r, K / R3);
    return V(0.0, 0.0, 0.0);
}

static void force_atom(const orc_md* o, int k, const double wbox[3], v3* Fout, float* Uout)
{
    v3 F = V(0.0, 0.0, 0.0), f;
    float U = 0.0f;
    int j, i, molk;
    v3 rk;
    *Fout = F; *Uout = 0.0f;
    if (!o->mobile[k]) return;
    rk = o->r[k];
    molk = o->mol[k];

    /* bonds (FCF:172-184, 537-575) */
    for (j = 0; j < o->bnum[k]; j++) {
        int g = o->bfirst[k] + j, ib = o->bj[g];
        const pt* tab = o->bondtab + (size_t)o->btype[g] * NPTS;
        v3 ri = move_same_side(rk, o->r[ib], wbox);
        float r = (float)vnorm(vsub(rk, ri));
        float mdU = 0.0f, e = 0.0f;
        table_eval(tab, r, ib > k, &mdU, &e);
        if (ib > k) U += e;
        f = vmul(coeffs12(ri, rk), (double)mdU);
        F = vadd(F, f);
    }
    /* angles (FCF:186-207, 577-642) */
    for (j = 0; j < o->anum[k]; j++) {
        int g = o->afirst[k] + j;
        const pt* tab = o->angtab + (size_t)o->atype[g] * NPTS;
        v3 ri = move_same_side(rk, o->r[o->ai[g]], wbox);
        v3 rj = move_same_side(rk, o->r[o->aj[g]], wbox);
        float mdU = 0.0f, e = 0.0f;
        if (o->acenter[g]) {
            /* calcForceAngularOn_r_i(r_i, r_k, r_j): angle at k */
            float theta = calc_angle(ri, rk, rj);
            table_eval(tab, theta, 1, &mdU, &e);
            U += e;
            f = vmul(ang_coeffs2(rj, rk, ri), (double)mdU);
        } else {
            float theta = calc_angle(rk, ri, rj);
            table_eval(tab, theta, 0, &mdU, &e);
            f = vmul(ang_coeffs13(rj, ri, rk), (double)mdU);
        }
        F = vadd(F, f);
    }
    /* dihedrals (FCF:209-232, 644-709) */
    for (j = 0; j < o->dnum[k]; j++) {
        int g = o->dfirst[k] + j;
        const pt* tab = o->dihtab + (size_t)o->dtype[g] * NPTS;
        v3 ri = move_same_side(rk, o->r[o->di[g]], wbox);
        v3 rj = move_same_side(rk, o->r[o->dj[g]], wbox);
        v3 rl = move_same_side(rk, o->r[o->dl[g]], wbox);
        float mdU = 0.0f, e = 0.0f;
        if (o->dcenter[g]) {
            /* calcForceDihedralOn_r_i(r_i, r_k, r_j, r_l) */
            float theta = calc_dihedral(ri, rk, rj, rl);
            table_eval(tab, theta, 1, &mdU, &e);
            U += 0.5f * e;
            f = vmul(vmul(dih_coeffs23(rl, rj, rk, ri), -1.0), (double)mdU);
        } else {
            float theta = calc_dihedral(rk, ri, rj, rl);
            table_eval(tab, theta, 0, &mdU, &e);
            f = vmul(vmul(dih_coeffs14(rl, rj, ri, rk), -1.0), (double)mdU);
        }
        F = vadd(F, f);
    }
    /* neighbours (FCF:234-256, 467-512) */
    for (j = 0; j < o->ncount[k]; j++) {
        int ia = o->nlist[(size_t)k * o->maxn + j];
        float u = 0.0f, r, mdU = 0.0f;
        v3 ri = move_same_side(rk, o->r[ia], wbox);
        r = (float)vnorm(vsub(rk, ri));
        if (r < 1E-3f) {
            f = V(0.0, 0.0, 0.0);
        } else {
            int T = o->ntypes, row = o->type[k], col = o->type[ia], ff;
            int cnt = o->nbcount[row * T + col];
            for (ff = 0; ff < cnt; ff++) {
                const pt* tab = o->nb + (((size_t)row * T + col) * MAXFF + ff) * NPTS;
                table_eval(tab, r, ia > k, &mdU, &u);
            }
            f = vmul(coeffs12(ri, rk), (double)mdU);
        }
        scale_pair(o, &f, &u, ia, molk, k);
        F = vadd(F, f);
        U += u;
    }
    /* Coulomb (FCF:258-278) */
    if (!o->all_q_zero) {
        if (o->coulomb == MTMD_COUL_ALLPAIRS) {
            for (i = 0; i < o->n; i++) {
                float u = 0.0f;
                v3 ri = move_same_side(rk, o->r[i], wbox);
                f = coulomb_pair(o, rk, ri, k, i, &u);
                scale_pair(o, &f, &u, i, molk, k);
                F = vadd(F, f);
                U += u;
            }
        } else {
            for (j = 0; j < o->ncount[k]; j++) {
                int ia = o->nlist[(size_t)k * o->maxn + j];
                float u = 0.0f;
                v3 ri = move_same_side(rk, o->r[ia], wbox);
                f = coulomb_pair(o, rk, ri, k, ia, &u);
                scale_pair(o, &f, &u, ia, molk, k);
                F = vadd(F, f);
                U += u;
            }
        }
    }
    *Fout = F;
    *Uout = U;
}

static void compute_forces_internal(orc_md* o)
{
    int k;
    double wbox[3];
    orc_build_cells(o, NULL, NULL, NULL, NULL, NULL);
    orc_build_neighbors(o, 0, NULL, NULL);
    for (k = 0; k < 3; k++) {
        /* CFunctorCalcForce pbc_ (FCF:41): [-Lf/2, Lf/2] in double */
        double Lf = (double)(float)o->L[k];
        double hi = Lf / 2.0, lo = -Lf / 2.0;
        wbox[k] = hi - lo;
    }
#pragma omp parallel for schedule(dynamic, 256) if (o->threads > 1) num_threads(o->threads)
    for (k = 0; k < o->n; k++) force_atom(o, k, wbox, &o->F[k], &o->U[k]);
    o->have_forces = 1;
}

void orc_compute_forces(orc_md* o, double* F, float* U)
{
    compute_forces_internal(o);
    orc_get_forces(o, F, U);
}

void orc_get_forces(const orc_md* o, double* F, float* U)
{
    int k;
    for (k = 0; k < o->n; k++) {
        if (F) { F[3 * k] = o->F[k].x; F[3 * k + 1] = o->F[k].y; F[3 * k + 2] = o->F[k].z; }
        if (U) U[k] = o->U[k];
    }
}

/* ------------------------------------------------------------------ */
/* Nose-Hoover chain (NHC:54-113) with CFctT (MDL:29-60)                */
/* ------------------------------------------------------------------ */
/* G of CFctT (MDL:29-52): j = 0 from the particles' kinetic energy */
static double fctT_G0(orc_md* o, double beta)
{
    const double g = (double)(3 * o->n);
    double s = 0.0;
    int k;
    for (k = 0; k < o->n; k++) s += vdot(o->p[k], o->p[k]) / o->m[k];
    return s - (g / beta);
}

static void fctT_scale(orc_md* o, double coeff)
{
    int k;
    for (k = 0; k < o->n; k++) o->p[k] = vmul(o->p[k], coeff);
}

/* G and scaleMomentum of CFctP (MDL:63-83): the barostat momentum, g = 1 */
static double fctP_G0(orc_md* o, double beta)
{
    return o->peps * o->peps / o->W - (1.0 / beta);
}

static void fctP_scale(orc_md* o, double coeff) { o->peps *= coeff; }

/* CNHChain::propagator (NHC:54-113) for a chain of length M acting through
 * G0 / scale; G(j > 0) = p_eta[j-1]^2 / Q[j-1] - 1/beta for both functors */
static void chain_propagate(orc_md* o, int M, int nresp, double T, double* eta, double* peta, const double* Q,
                            double (*G0)(orc_md*, double), void (*scale)(orc_md*, double))
{
    double beta = 1.0 / T, coeff;
    int mu, i, j;
#define CHAIN_G(jj) ((jj) == 0 ? G0(o, beta) : (peta[(jj) - 1] * peta[(jj) - 1] / Q[(jj) - 1]) - (1.0 / beta))
    if (M > 0 && peta[0] > 15.0) peta[0] = 10.0;
    if (M <= 0) return;
    for (mu = 0; mu < 3; mu++) {
        double Dt = o->w[mu] * o->dt / (double)nresp;
        for (i = 0; i < nresp; i++) {
            peta[M - 1] += (Dt / 4.0) * CHAIN_G(M - 1);
            for (j = 0; j < M - 1; j++) {
                coeff = orc_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= coeff;
                peta[j] += (Dt / 4.0) * CHAIN_G(j);
                coeff = orc_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= coeff;
            }
            for (j = 0; j < M; j++) eta[j] += (Dt / 2.0) * (peta[j] / Q[j]);
            coeff = orc_cmt_exp(-((Dt / 2.0) * (peta[0] / Q[0])));
            scale(o, coeff);
            for (j = M - 2; j >= 0; j--) {
                coeff = orc_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= coeff;
                peta[j] += (Dt / 4.0) * CHAIN_G(j);
                coeff = orc_cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] *= coeff;
            }
            peta[M - 1] += (Dt / 4.0) * CHAIN_G(M - 1);
        }
    }
#undef CHAIN_G
}

void orc_nh_propagate(orc_md* o)
{
    chain_propagate(o, o->M, o->nresp, o->T, o->eta, o->peta, o->Q, fctT_G0, fctT_scale);
}

/* ------------------------------------------------------------------ */
/* Integrator (VV:55-122), wrap (SB:135-166), MD loop (MDL:124-138)     */
/* ------------------------------------------------------------------ */
static void cut_forces(orc_md* o)
{
    int k;
    double Fc = o->Fcut;
    for (k = 0; k < o->n; k++) {
        v3* F = &o->F[k];
        if (fabs(F->x) > Fc) F->x = ((F->x >= 0.0) ? 1.0 : -1.0) * Fc;
        if (fabs(F->y) > Fc) F->y = ((F->y >= 0.0) ? 1.0 : -1.0) * Fc;
        if (fabs(F->z) > Fc) F->z = ((F->z >= 0.0) ? 1.0 : -1.0) * Fc;
    }
}

static inline void pbc_add(double* pos, double L, double Lm) { *pos -= (double)d2i((*pos + L) / Lm) * Lm; }

static void pbc_wrap(orc_md* o)
{
    int k;
    double LX = o->L[0], LY = o->L[1], LZ = o->L[2];
    double X2 = LX / 2.0, Y2 = LY / 2.0, Z2 = LZ / 2.0;
    for (k = 0; k < o->n; k++) {
        v3* r = &o->r[k];
        if (r->x >= X2) pbc_add(&r->x, X2, LX);
        if (r->x < -X2) pbc_add(&r->x, -X2, LX);
        if (r->y >= Y2) pbc_add(&r->y, Y2, LY);
        if (r->y < -Y2) pbc_add(&r->y, -Y2, LY);
        if (r->z >= Z2) pbc_add(&r->z, Z2, LZ);
        if (r->z < -Z2) pbc_add(&r->z, -Z2, LZ);
    }
}

static void vv_step(orc_md* o)
{
    double dt_2 = o->dt / 2.0;
    int k;
    for (k = 0; k < o->n; k++) {
        double mk = o->m[k];
        o->p[k] = vadd(o->p[k], vmul(o->F[k], dt_2));
        o->r[k] = vadd(o->r[k], vmul(o->p[k], o->dt / mk));
    }
    compute_forces_internal(o);
    if (o->Fcut > 0) cut_forces(o);
    for (k = 0; k < o->n; k++) o->p[k] = vadd(o->p[k], vmul(o->F[k], dt_2));
}

/* NPT (VV:57-71,124-150,160-174,192-211): the barostat variables and the
 * box follow CVelVerlet operation for operation */
static double npt_volume(const orc_md* o) { return o->V0 * exp(3.0 * o->eps); }   /* getV, VV:192-197 */

static void npt_box(orc_md* o)                                                      /* getSimBox, VV:199-211 */
{
    double Vmax = o->Lmax[0] * o->Lmax[1] * o->Lmax[2];
    double etaCube = npt_volume(o) / Vmax;
    etaCube = pow(etaCube, 1.0 / 3.0);
    o->L[0] = etaCube * o->Lmax[0];
    o->L[1] = etaCube * o->Lmax[1];
    o->L[2] = etaCube * o->Lmax[2];
}

static double npt_G_eps(const orc_md* o)                                            /* G_eps, VV:160-174 */
{
    double V = npt_volume(o), sum_p = 0.0, sum_f = 0.0;
    int k;
    for (k = 0; k < o->n; k++) {
        double p2 = vdot(o->p[k], o->p[k]);
        sum_p += (p2 / o->m[k]);
        sum_f += vdot(o->F[k], o->r[k]);
    }
    return (1.0 + 1.0 / (double)o->n) * sum_p + sum_f - 3.0 * o->Pext * V;
}

static void npt_prop_p(orc_md* o)                                                   /* prop_p, VV:124-136 */
{
    double u_eps = o->peps / o->W;
    double alpha = (1.0 + 1.0 / (double)o->n);
    double parm = alpha * u_eps * o->dt / 4.0;
    double coeff = orc_cmt_exp(-parm);
    double coeff2 = coeff * coeff;
    double coeff3 = orc_cmt_sinhxoverx(parm);
    double cf = (o->dt / 2.0) * coeff * coeff3;
    int k;
    for (k = 0; k < o->n; k++) o->p[k] = vadd(vmul(o->p[k], coeff2), vmul(o->F[k], cf));
}

static void npt_prop_r(orc_md* o)                                                   /* prop_r, VV:138-150 */
{
    double u_eps = o->peps / o->W;
    double parm = u_eps * o->dt / 2.0;
    double coeff = orc_cmt_exp(parm);
    double coeff2 = coeff * coeff;
    double coeff3 = orc_cmt_sinhxoverx(parm) * coeff;
    double cr = o->dt * coeff3;
    int k;
    for (k = 0; k < o->n; k++) {
        v3 u = vmul(o->p[k], 1.0 / o->m[k]);
        o->r[k] = vadd(vmul(o->r[k], coeff2), vmul(u, cr));
    }
}

static void npt_step(orc_md* o)                                                     /* VV:57-71 */
{
    o->peps += ((o->dt / 2.0) * npt_G_eps(o));
    npt_prop_p(o);
    npt_prop_r(o);
    o->eps += (o->dt * o->peps / o->W);
    npt_box(o);
    compute_forces_internal(o);
    npt_prop_p(o);
    o->peps += ((o->dt / 2.0) * npt_G_eps(o));
}

static void orc_nhp_propagate(orc_md* o)
{
    chain_propagate(o, o->Mp, o->nrespP, o->T, o->etaP, o->petaP, o->QP, fctP_G0, fctP_scale);
}

void orc_step(orc_md* o, int nsteps)
{
    int t;
    const int npt = (o->ensemble == MTMD_NPT), nh_t = (o->ensemble == MTMD_NVT || npt);
    if (!o->have_forces) compute_forces_internal(o);   /* calcInitialForces, MDL:189-192 */
    for (t = 0; t < nsteps; t++) {
        if (npt) orc_nhp_propagate(o);
        if (nh_t) orc_nh_propagate(o);
        if (npt) npt_step(o); else vv_step(o);
        if (nh_t) orc_nh_propagate(o);
        if (npt) orc_nhp_propagate(o);
        pbc_wrap(o);                                   /* with calcSimBox: o->L is the current box */
    }
}

/* calcTemp, calcPress, calcV, calcUtot, calcCurrentKineticEnergy (SB:168-231) */
void orc_thermo(orc_md* o, double* out)
{
    double sp = 0.0, spf = 0.0, Kt = 0.0, Ut = 0.0, V, Tred, P;
    int k;
    if (!o->have_forces) compute_forces_internal(o);
    for (k = 0; k < o->n; k++) sp += vdot(o->p[k], o->p[k]) / o->m[k];
    Tred = sp / (3 * (double)o->n);
    V = (o->ensemble == MTMD_NPT) ? npt_volume(o) : o->L[0] * o->L[1] * o->L[2];
    for (k = 0; k < o->n; k++) {
        double p2 = vdot(o->p[k], o->p[k]) / o->m[k];
        double f = vdot(o->F[k], o->r[k]);
        spf += (p2 + f);
    }
    P = (1.0 / (3.0 * V)) * spf;
    for (k = 0; k < o->n; k++) Kt += vdot(o->p[k], o->p[k]) / o->m[k];
    Kt = 0.5 * Kt;
    for (k = 0; k < o->n; k++) Ut += (double)o->U[k];
    out[0] = Tred * CONV_T;
    out[1] = P * CONV_PRESS;
    out[2] = V;
    out[3] = Ut;
    out[4] = Kt;
    out[5] = Ut + Kt;
    out[6] = sp;
    {
        double sf = 0.0;
        for (k = 0; k < o->n; k++) sf += vdot(o->F[k], o->r[k]);
        out[7] = sf;
    }
}

/* CMDLoop::runOptimization (MDL:143-173), gradientDescentStep (SB:197-208) */
int orc_optimize(orc_md* o, int max_steps, double lr, double accuracy, int out_stride)
{
    double Un = 0.0, Unp1, Udiff;
    int i, k;
    compute_forces_internal(o);
    for (k = 0; k < o->n; k++) Un += (double)o->U[k];
    for (i = 0; i < max_steps; i++) {
        compute_forces_internal(o);
        for (k = 0; k < o->n; k++) o->r[k] = vadd(o->r[k], vmul(o->F[k], lr));
        Unp1 = 0.0;
        for (k = 0; k < o->n; k++) Unp1 += (double)o->U[k];
        Udiff = fabs(Unp1 - Un);
        Un = Unp1;
        if ((Udiff < accuracy) && (i > out_stride)) { i++; break; }
    }
    o->have_forces = 0;
    return i;
}

/* ------------------------------------------------------------------ */
/* Topology search (MolTwisterStateTools.cpp:544-724,767-920,993-1190): */
/* for each rule, the loops of its detection criterion over the atoms   */
/* of the rule's first type in index order (getAtomsWithID,             */
/* MolTwisterState.cpp:598-610), then the duplicate removal of          */
/* MolecularSystemTools.cpp:25-235 (first occurrence kept, an entry     */
/* equals its reverse; angles and dihedrals naming an atom twice are    */
/* dropped). Brute force: small systems only.                           */
/* ------------------------------------------------------------------ */
static double topo_d2(const double* pos, int a, int b, const double* w)
{
    double dx = pos[3 * a] - pos[3 * b], dy = pos[3 * a + 1] - pos[3 * b + 1], dz = pos[3 * a + 2] - pos[3 * b + 2];
    if (w) {   /* CAtom::getDistanceTo2UsingPBC, MolTwisterAtom.cpp:374-394 */
        dx = (dx < 0.0) ? -dx : dx;
        dy = (dy < 0.0) ? -dy : dy;
        dz = (dz < 0.0) ? -dz : dz;
        if (dx > (w[0] / 2.0)) dx = w[0] - dx;
        if (dy > (w[1] / 2.0)) dy = w[1] - dy;
        if (dz > (w[2] / 2.0)) dz = w[2] - dz;
    }
    return (dx * dx + dy * dy + dz * dz);
}

static int topo_same(const int* a, const int* b, int na)
{
    int q, fwd = 1, rev = 1;
    for (q = 0; q < na; q++) { fwd &= (a[q] == b[q]); rev &= (a[q] == b[na - 1 - q]); }
    return fwd || rev;
}

/* CAtom::findAtomsInMolecule (MolTwisterAtom.cpp:578-605): the atom, then
 * recursively each bond destination in the atom's bond order, each atom
 * once (the visited map) */
static void topo_mol_visit(int a, const int32_t* bs, const int32_t* bd, char* seen, int* order, int* cnt)
{
    int i;
    if (seen[a]) return;
    seen[a] = 1;
    order[(*cnt)++] = a;
    for (i = bs[a]; i < bs[a + 1]; i++) topo_mol_visit(bd[i], bs, bd, seen, order, cnt);
}

typedef struct { int* buf; int cnt, cap, width; } topo_out;

static void topo_emit(topo_out* t, int a, int b, int c, int d, int r)
{
    int v[5] = {a, b, c, d, r};
    if (t->cnt == t->cap) { t->cap *= 2; t->buf = (int*)realloc(t->buf, sizeof(int) * (size_t)t->cap * t->width); }
    memcpy(t->buf + (size_t)t->cnt * t->width, v, sizeof(int) * t->width);
    t->cnt++;
}

int orc_find_topology(int which, int n, const double* pos, const int32_t* type, const double* box,
                      const int32_t* bond_start, const int32_t* bond_dest, int n_rules,
                      const mtmd_topo_rule* rules, int max_out, int32_t* out)
{
    const int na = which + 1, width = which + 2;
    topo_out t;
    int r, j, k, l, m, e, f;
    int* lists[4];
    int nlist[4];
    char* seen = (char*)calloc((size_t)n + 1, 1);
    int* mol = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    t.cap = 1024; t.cnt = 0; t.width = width;
    t.buf = (int*)malloc(sizeof(int) * (size_t)t.cap * width);
    for (r = 0; r < n_rules; r++) {
        const double R02 = rules[r].R0 * rules[r].R0;
        const int* ty = rules[r].types;
        const int crit = rules[r].crit;
        int q;
        for (q = 0; q < na; q++) {          /* CMolTwisterState::getAtomsWithID */
            lists[q] = (int*)malloc(sizeof(int) * (size_t)(n + 1));
            nlist[q] = 0;
            for (j = 0; j < n; j++) if (type[j] == ty[q]) lists[q][nlist[q]++] = j;
        }
        if (crit == MTMD_CRIT_ALL_R) {      /* :577-603, :805-839, :1033-1077 */
            for (j = 0; j < nlist[0]; j++) {
                int a = lists[0][j];
                for (k = 0; k < nlist[1]; k++) {
                    int b = lists[1][k];
                    if (a == b) continue;
                    if (!(topo_d2(pos, a, b, box) < R02)) continue;
                    if (which == 1) { topo_emit(&t, a, b, r, 0, 0); continue; }
                    for (l = 0; l < nlist[2]; l++) {
                        int c = lists[2][l];
                        if (c == b || c == a) continue;
                        if (!(topo_d2(pos, b, c, box) < R02)) continue;
                        if (which == 2) { topo_emit(&t, a, b, c, r, 0); continue; }
                        for (m = 0; m < nlist[3]; m++) {
                            int d = lists[3][m];
                            if (d == c || d == b || d == a) continue;
                            if (topo_d2(pos, c, d, box) < R02) topo_emit(&t, a, b, c, d, r);
                        }
                    }
                }
            }
        } else if (crit == MTMD_CRIT_MOL_R) {   /* :605-636, :843-888, :1084-1144 */
            for (j = 0; j < nlist[0]; j++) {
                int a = lists[0][j], nm = 0, i;
                topo_mol_visit(a, bond_start, bond_dest, seen, mol, &nm);
                for (i = 0; i < nm; i++) seen[mol[i]] = 0;
                for (k = 0; k < nm; k++) {
                    int b = mol[k];
                    if (a == b) continue;
                    if (!(type[b] == ty[1] && topo_d2(pos, a, b, box) < R02)) continue;
                    if (which == 1) { topo_emit(&t, a, b, r, 0, 0); continue; }
                    for (l = 0; l < nm; l++) {
                        int c = mol[l];
                        if (c == a || c == b) continue;
                        if (!(type[c] == ty[2] && topo_d2(pos, b, c, box) < R02)) continue;
                        if (which == 2) { topo_emit(&t, a, b, c, r, 0); continue; }
                        for (m = 0; m < nm; m++) {
                            int d = mol[m];
                            if (d == a || d == b || d == c) continue;
                            if (type[d] == ty[3] && topo_d2(pos, c, d, box) < R02) topo_emit(&t, a, b, c, d, r);
                        }
                    }
                }
            }
        } else if (crit == MTMD_CRIT_VISIBLE) {  /* :637-666, :889-920, :1145-1198 */
            for (j = 0; j < nlist[0]; j++) {
                int a = lists[0][j];
                for (k = bond_start[a]; k < bond_start[a + 1]; k++) {
                    int b = bond_dest[k];
                    if (type[b] != ty[1]) continue;
                    if (which == 1) { topo_emit(&t, a, b, r, 0, 0); continue; }
                    for (l = bond_start[b]; l < bond_start[b + 1]; l++) {
                        int c = bond_dest[l];
                        if (type[c] != ty[2]) continue;
                        if (which == 2) { topo_emit(&t, a, b, c, r, 0); continue; }
                        for (m = bond_start[c]; m < bond_start[c + 1]; m++) {
                            int d = bond_dest[m];
                            if (type[d] == ty[3]) topo_emit(&t, a, b, c, d, r);
                        }
                    }
                }
            }
        } else if (crit == MTMD_CRIT_14 && which == 1) {   /* :667-716 */
            for (j = 0; j < nlist[0]; j++) {
                int a = lists[0][j];
                for (k = bond_start[a]; k < bond_start[a + 1]; k++) {
                    int b = bond_dest[k];
                    for (l = bond_start[b]; l < bond_start[b + 1]; l++) {
                        int c = bond_dest[l];
                        for (m = bond_start[c]; m < bond_start[c + 1]; m++) {
                            int d = bond_dest[m];
                            if (type[d] == ty[1] && d != c && d != b && d != a) topo_emit(&t, a, d, r, 0, 0);
                        }
                    }
                }
            }
        }
        /* any other criterion: the reference prints an error and returns what it has (none here) */
        for (q = 0; q < na; q++) free(lists[q]);
    }
    /* removeDuplicate*Indices: keep the first of equal or reversed entries,
     * then drop angles / dihedrals naming an atom twice */
    {
        int kept = 0;
        for (e = 0; e < t.cnt; e++) {
            int dup = 0;
            for (f = 0; f < kept && !dup; f++) dup = topo_same(t.buf + (size_t)f * width, t.buf + (size_t)e * width, na);
            if (dup) continue;
            if (kept != e) memmove(t.buf + (size_t)kept * width, t.buf + (size_t)e * width, sizeof(int) * width);
            kept++;
        }
        t.cnt = kept;
        if (which >= 2) {
            kept = 0;
            for (e = 0; e < t.cnt; e++) {
                const int* x = t.buf + (size_t)e * width;
                int p2, p3, rep = 0;
                for (p2 = 0; p2 < na; p2++)
                    for (p3 = p2 + 1; p3 < na; p3++) rep |= (x[p2] == x[p3]);
                if (rep) continue;
                if (kept != e) memmove(t.buf + (size_t)kept * width, x, sizeof(int) * width);
                kept++;
            }
            t.cnt = kept;
        }
    }
    if (out) memcpy(out, t.buf, sizeof(int) * (size_t)(t.cnt < max_out ? t.cnt : max_out) * width);
    free(t.buf);
    free(seen);
    free(mol);
    return t.cnt;
}
