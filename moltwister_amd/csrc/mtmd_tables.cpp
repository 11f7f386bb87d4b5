// mtmd_tables.cpp - 1000-point force/energy profiles of the MDFF classes.
//
// The engine interpolates these float tables exactly as the reference does, so
// they must be bit-identical to what the reference generates on the host:
// each profile follows CMDFF*::calc1DForceProfile / calc1DPotentialProfile
// (MolTwisterMDFFNonBonded.cpp:126-172, MolTwisterMDFFBond.cpp:133-179,
// MolTwisterMDFFAngle.cpp:210-257, MolTwisterMDFFDih.cpp:272-320) with the
// analytic force fields of the classes named below. Built with
// -ffp-contract=off (x86-64 SSE2 arithmetic, as the reference host build).
#include <cmath>
#include <array>
#include "../../include/mtmd.h"

namespace {

struct V3 {
    double x, y, z;
    V3 operator-(const V3& b) const { return {x - b.x, y - b.y, z - b.z}; }
    V3 operator*(double s) const { return {x * s, y * s, z * s}; }
    double dot(const V3& b) const { return (x * b.x) + (y * b.y) + (z * b.z); }
    double norm() const { return std::sqrt(x * x + y * y + z * z); }
    V3 cross(const V3& b) const { return {y * b.z - z * b.y, z * b.x - x * b.z, x * b.y - y * b.x}; }
};

inline double inv_or_big(double R) { return (R == 0.0) ? 1.0 / 1E-10 : 1.0 / R; }

// (-dU/dr)/r and U of the non-bonded classes at separation vector r12 = r2 - r1
// LJ: MolTwisterMDFFNonBonded_LJ.cpp:131-162; Buck: _Buck.cpp:116-143; LJ1208: _LJ1208.cpp:131-170
struct NonBonded {
    int kind; std::array<double, 4> p;
    double coef(const V3& r12) const
    {
        const double R = r12.norm(), RInv = inv_or_big(R), RInv2 = RInv * RInv;
        switch (kind) {
        case MTMD_NB_LJ: {
            const double RInv6 = RInv2 * RInv2 * RInv2, RInv12 = RInv6 * RInv6;
            const double s2 = p[1] * p[1], s6 = s2 * s2 * s2, s12 = s6 * s6;
            return 4.0 * p[0] * RInv2 * (12.0 * s12 * RInv12 - 6.0 * p[2] * s6 * RInv6);
        }
        case MTMD_NB_BUCK: {
            const double RInv4 = RInv2 * RInv2, RInv6 = RInv4 * RInv2;
            const double rhoInv = (p[1] == 0.0) ? 1E-10 : 1.0 / p[1];
            return RInv2 * (p[0] * R * rhoInv * std::exp(-R * rhoInv) - 6.0 * p[2] * RInv6);
        }
        case MTMD_NB_LJ1208: {
            const double RInv4 = RInv2 * RInv2, RInv6 = RInv4 * RInv2, RInv8 = RInv4 * RInv4, RInv12 = RInv6 * RInv6;
            const double s2 = p[1] * p[1], s4 = s2 * s2, s6 = s4 * s2, s8 = s4 * s4, s12 = s6 * s6;
            return 4.0 * p[0] * RInv2 * (12.0 * s12 * RInv12 - 8.0 * p[2] * s8 * RInv8);
        }
        case MTMD_NB_LJBUCK:
            // CMDFFNonBonded_LJBuck::calcForces ends by zeroing both forces (_LJBuck.cpp:250)
            return 0.0;
        }
        return 0.0;
    }
    double energy(const V3& r12) const
    {
        const double R = r12.norm(), RInv = inv_or_big(R), RInv2 = RInv * RInv;
        switch (kind) {
        case MTMD_NB_LJ: {
            const double RInv6 = RInv2 * RInv2 * RInv2, RInv12 = RInv6 * RInv6;
            const double s2 = p[1] * p[1], s6 = s2 * s2 * s2, s12 = s6 * s6;
            return 4.0 * p[0] * (s12 * RInv12 - p[2] * s6 * RInv6);
        }
        case MTMD_NB_BUCK: {
            const double RInv4 = RInv2 * RInv2, RInv6 = RInv4 * RInv2;
            const double rhoInv = (p[1] == 0.0) ? 1E-10 : 1.0 / p[1];
            return p[0] * std::exp(-R * rhoInv) - p[2] * RInv6;
        }
        case MTMD_NB_LJ1208: {
            const double RInv4 = RInv2 * RInv2, RInv6 = RInv4 * RInv2, RInv8 = RInv4 * RInv4, RInv12 = RInv6 * RInv6;
            const double s2 = p[1] * p[1], s4 = s2 * s2, s6 = s4 * s2, s8 = s4 * s4, s12 = s6 * s6;
            return 4.0 * p[0] * (s12 * RInv12 - p[2] * s8 * RInv8);
        }
        case MTMD_NB_LJBUCK: {
            // CMDFFNonBonded_LJBuck::calcPotential (_LJBuck.cpp:181-211): LJ (alpha 1) or Buckingham
            if (p[0] == 0.0) {
                const double RInv6 = RInv2 * RInv2 * RInv2, RInv12 = RInv6 * RInv6;
                const double s2 = p[2] * p[2], s6 = s2 * s2 * s2, s12 = s6 * s6;
                return 4.0 * p[1] * (s12 * RInv12 - s6 * RInv6);
            }
            if (p[0] == 1.0) {
                const double RInv4 = RInv2 * RInv2, RInv6 = RInv4 * RInv2;
                const double rhoInv = (p[2] == 0.0) ? 1E-10 : 1.0 / p[2];
                return p[1] * std::exp(-R * rhoInv) - p[3] * RInv6;
            }
            return 0.0;
        }
        }
        return 0.0;
    }
};

// Harm: MolTwisterMDFFBond_Harm.cpp:90-110; Morse: MolTwisterMDFFBond_Morse.cpp:95-117
struct Bond {
    int kind; std::array<double, 4> p;
    double coef(const V3& r12) const
    {
        const double R = r12.norm(), RInv = inv_or_big(R);
        if (kind == MTMD_BOND_HARM) return -2.0 * p[0] * (R - p[1]) * RInv;
        if (kind == MTMD_BOND_MORSE) {
            const double E = std::exp(-p[1] * (R - p[2]));
            return -2.0 * p[0] * p[1] * RInv * E * (1.0 - E);
        }
        return 0.0;
    }
    double energy(const V3& r12) const
    {
        const double R = r12.norm();
        if (kind == MTMD_BOND_HARM) { const double d = R - p[1]; return p[0] * d * d; }
        if (kind == MTMD_BOND_MORSE) { const double q = (1.0 - std::exp(-p[1] * (R - p[2]))); return p[0] * q * q; }
        return 0.0;
    }
};

// CMDFFAngle::calcAngularForceCoeffs13 (MolTwisterMDFFAngle.cpp:62-83)
V3 angle_dtheta_dr3(const V3& r1, const V3& r2, const V3& r3)
{
    const V3 K = (r1 - r2) * inv_or_big((r1 - r2).norm());
    const V3 r32 = r3 - r2;
    const double A = K.x * r32.x + K.y * r32.y + K.z * r32.z;
    const double B = r32.x * r32.x + r32.y * r32.y + r32.z * r32.z;
    const double sB = std::sqrt(B);
    const double sB3 = (B == 0.0) ? std::sqrt(1E-30) : sB * sB * sB;
    const double W = (sB == 0.0) ? A / 10E-3 : A / sB;
    const double dv = std::sqrt(1.0 - W * W);
    const double g = (dv == 0.0) ? -1.0 / 1E-10 : -1.0 / dv;
    return {g * ((K.x * B - r32.x * A) / sB3), g * ((K.y * B - r32.y * A) / sB3), g * ((K.z * B - r32.z * A) / sB3)};
}

// CMDFFDih::calcDihedralForceCoeffs14 (MolTwisterMDFFDih.cpp:64-87)
V3 dihedral_dphi_dr4(const V3& r1, const V3& r2, const V3& r3, const V3& r4)
{
    const V3 r21 = r1 - r2, r23 = r3 - r2, r32 = r23 * (-1.0), r34 = r4 - r3;
    const V3 n1 = r23.cross(r21);
    const V3 K = n1 * inv_or_big(n1.norm());
    const double a = r32.y * r34.z - r32.z * r34.y;
    const double b = r32.z * r34.x - r32.x * r34.z;
    const double c = r32.x * r34.y - r32.y * r34.x;
    const double A = K.x * a + K.y * b + K.z * c;
    const double B = a * a + b * b + c * c;
    const double sBi = (B == 0.0) ? 1.0 / 1E-20 : 1.0 / std::sqrt(B);
    const double sBi3 = sBi * sBi * sBi;
    const double W = A * sBi;
    const double dx = ((K.y * r32.z - K.z * r32.y) * B - (r32.z * (r32.z * r34.x - r32.x * r34.z) - r32.y * (r32.x * r34.y - r32.y * r34.x)) * A) * sBi3;
    const double dy = ((K.z * r32.x - K.x * r32.z) * B - (r32.x * (r32.x * r34.y - r32.y * r34.x) - r32.z * (r32.y * r34.z - r32.z * r34.y)) * A) * sBi3;
    const double dz = ((K.x * r32.y - K.y * r32.x) * B - (r32.y * (r32.y * r34.z - r32.z * r34.y) - r32.x * (r32.z * r34.x - r32.x * r34.z)) * A) * sBi3;
    const double dv = std::sqrt(1.0 - W * W);
    const double g = (dv == 0.0) ? 1.0 / 1E-10 : 1.0 / dv;
    return {g * dx, g * dy, g * dz};
}

constexpr double kToRad = 0.017453292519943295;   // MolTwisterMDFFAngle_Harm.h:36, _Class2.h:76

// theta - theta0 at (r1, r2, r3) as the angle classes compute it
double angle_delta(const V3& r1, const V3& r2, const V3& r3, double theta0_deg)
{
    const V3 r21 = r1 - r2, r23 = r3 - r2;
    const double cosT = r21.dot(r23) * (inv_or_big(r21.norm()) * inv_or_big(r23.norm()));
    return std::acos(cosT) - theta0_deg * kToRad;
}

// CMDFFAngle_Class2 (MolTwisterMDFFAngle_Class2.cpp:145-296): the force on r3 and
// the energy of Ea (degrees 2, 3, 4), Ebb and Eba, summed in the class's order.
// q: theta0, K2, K3, K4, M, r1, r2 (Ebb), N1, N2, r1, r2 (Eba)
V3 class2_force3(const V3& r1, const V3& r2, const V3& r3, const double* q)
{
    V3 f3{0.0, 0.0, 0.0};
    auto add = [&](const V3& a) { f3 = {f3.x + a.x, f3.y + a.y, f3.z + a.z}; };
    const V3 c13 = angle_dtheta_dr3(r1, r2, r3);
    for (int degree = 2; degree <= 4; degree++) {                  // calcForcesHarm
        const double dT = angle_delta(r1, r2, r3, q[0]);
        double dU = -double(degree) * q[degree - 1];
        for (int i = 0; i < degree - 1; i++) dU *= dT;
        add(c13 * dU);
    }
    const V3 r12 = r2 - r1, r23 = r3 - r2;
    const double R12 = r12.norm(), R23 = r23.norm();
    const double RInv12 = (R12 == 0.0) ? 1.0 / 1E-10 : 1.0 / R12, RInv23 = (R23 == 0.0) ? 1.0 / 1E-10 : 1.0 / R23;
    const V3 u23 = r23 * RInv23;
    add(u23 * ((-1.0) * q[4] * (R12 - q[5])));                      // calcForcesCartesian
    {                                                               // calcForcesMix
        const V3 r21 = r12 * (-1.0);
        const double coeff = q[7] * (R12 - q[9]) + q[8] * (R23 - q[10]);
        const double cosT = r21.dot(r23) * (RInv12 * RInv23);
        const double dT = std::acos(cosT) - q[0] * kToRad;
        const V3 a = u23 * ((-1.0) * q[8] * dT), b = c13 * coeff;
        add({a.x - b.x, a.y - b.y, a.z - b.z});
    }
    return f3;
}

double class2_energy(const V3& r1, const V3& r2, const V3& r3, const double* q)
{
    double U = 0.0;
    for (int degree = 2; degree <= 4; degree++) {                  // calcPotentialHarm
        const double dT = angle_delta(r1, r2, r3, q[0]);
        double dN = 1.0;
        for (int i = 0; i < degree; i++) dN *= dT;
        U += q[degree - 1] * dN;
    }
    const double R12 = (r2 - r1).norm(), R23 = (r3 - r2).norm();
    U += q[4] * (R12 - q[5]) * (R23 - q[6]);                        // calcPotentialCartesian
    const double R21 = (r1 - r2).norm();
    U += (q[7] * (R21 - q[9]) + q[8] * (R23 - q[10])) * angle_delta(r1, r2, r3, q[0]);   // calcPotentialMix
    return U;
}

} // namespace

extern "C" int mtmd_profile(int which, int kind, const double* params, float x0, float x1, int n, float* xo,
                            float* fo, float* eo)
{
    if (n <= 0 || !params) return MTMD_EINVAL;
    if (which == 0 && kind != MTMD_NB_LJ && kind != MTMD_NB_BUCK && kind != MTMD_NB_LJ1208 && kind != MTMD_NB_LJBUCK)
        return MTMD_EINVAL;
    if (which == 1 && kind != MTMD_BOND_HARM && kind != MTMD_BOND_MORSE) return MTMD_EINVAL;
    if (which == 2 && kind != MTMD_ANGLE_HARM && kind != MTMD_ANGLE_CLASS2) return MTMD_EINVAL;
    if (which == 3 && kind != MTMD_DIH_HARM && kind != MTMD_DIH_FOURIER4T) return MTMD_EINVAL;
    if (which < 0 || which > 3) return MTMD_EINVAL;
    // non-bonded kinds take 4 parameters, bonded kinds MTMD_BONDED_PARAMS
    double q[MTMD_BONDED_PARAMS] = {0.0};
    for (int i = 0; i < (which == 0 ? 4 : MTMD_BONDED_PARAMS); i++) q[i] = params[i];
    std::array<double, 4> p{q[0], q[1], q[2], q[3]};
    const float step = (x1 - x0) / float(n - 1);
    for (int i = 0; i < n; i++) {
        const float x = x0 + float(i) * step;
        double fval = 0.0, e = 0.0;
        if (which <= 1) {
            // force on r2 = (x,0,0) from r1 = 0, divided by dr/dx2 = unit vector x
            const V3 r1{0.0, 0.0, 0.0}, r2{double(x), 0.0, 0.0};
            const V3 r12 = r2 - r1;
            const double c = (which == 0) ? NonBonded{kind, p}.coef(r12) : Bond{kind, p}.coef(r12);
            const V3 F = r12 * c;
            const V3 u = r12 * inv_or_big(r12.norm());
            fval = (u.x != 0.0) ? F.x / u.x : 0.0;
            e = (which == 0) ? NonBonded{kind, p}.energy(r12) : Bond{kind, p}.energy(r12);
        } else if (which == 2 && kind == MTMD_ANGLE_CLASS2) {
            const V3 r1{-1.0, 0.0, 0.0}, r2{0.0, 0.0, 0.0}, r3{-std::cos(double(x)), std::sin(double(x)), 0.0};
            const V3 f3 = class2_force3(r1, r2, r3, q);
            const V3 c = angle_dtheta_dr3(r1, r2, r3);
            fval = (c.x != 0.0) ? f3.x / c.x : 0.0;
            e = class2_energy(r1, r2, r3, q);
        } else if (which == 2) {
            // CMDFFAngle_Harm (MolTwisterMDFFAngle_Harm.cpp:94-130) at r1=(-1,0,0), r2=0, r3=(-cos, sin, 0)
            const V3 r1{-1.0, 0.0, 0.0}, r2{0.0, 0.0, 0.0}, r3{-std::cos(double(x)), std::sin(double(x)), 0.0};
            const V3 r21 = r1 - r2, r23 = r3 - r2;
            const double cosT = r21.dot(r23) * (inv_or_big(r21.norm()) * inv_or_big(r23.norm()));
            const double dTheta = std::acos(cosT) - p[1] * kToRad;
            const V3 c = angle_dtheta_dr3(r1, r2, r3);
            const V3 f3 = c * (-2.0 * p[0] * dTheta);
            fval = (c.x != 0.0) ? f3.x / c.x : 0.0;
            e = p[0] * dTheta * dTheta;
        } else {
            // CMDFFDih_Harm (MolTwisterMDFFDih_Harm.cpp:96-132) at r1=(0,1,0), r2=0, r3=(0,0,1), r4=(sin, cos, 1)
            const int D = int(p[1]), N = int(p[2]);
            const V3 r1{0.0, 1.0, 0.0}, r2{0.0, 0.0, 0.0}, r3{0.0, 0.0, 1.0};
            const V3 r4{std::sin(double(x)), std::cos(double(x)), 1.0};
            const V3 r23 = r3 - r2, r32 = r23 * (-1.0);
            const V3 n1 = r23.cross(r1 - r2), n2 = r32.cross(r4 - r3);
            const double nn = n1.norm() * n2.norm();
            const double cosPhi = (nn == 0.0) ? n1.dot(n2) / 1E-20 : n1.dot(n2) / nn;
            const double phi = std::acos(cosPhi);
            const V3 c = dihedral_dphi_dr4(r1, r2, r3, r4);
            if (kind == MTMD_DIH_FOURIER4T) {
                // CMDFFDih_Fourier4t (MolTwisterMDFFDih_Fourier4t.cpp:98-142)
                const double dU = -0.5 * p[0] * std::sin(phi) + p[1] * std::sin(2.0 * phi) -
                                  1.5 * p[2] * std::sin(3.0 * phi) + 2.0 * p[3] * std::sin(4.0 * phi);
                const V3 f4 = c * dU;
                fval = (c.x != 0.0) ? f4.x / c.x : 0.0;
                const double U1 = p[0] * (1.0 + std::cos(phi)), U2 = p[1] * (1.0 - std::cos(2.0 * phi));
                const double U3 = p[2] * (1.0 + std::cos(3.0 * phi)), U4 = p[3] * (1.0 - std::cos(4.0 * phi));
                e = 0.5 * (U1 + U2 + U3 + U4);
            } else {
                const V3 f4 = c * (-p[0] * double(D * N) * std::sin(double(N) * phi));
                fval = (c.x != 0.0) ? f4.x / c.x : 0.0;
                e = p[0] * (1.0 + double(D) * std::cos(double(N) * phi));
            }
        }
        if (xo) xo[i] = x;
        if (fo) fo[i] = float(fval);
        if (eo) eo[i] = float(e);
    }
    return MTMD_OK;
}
