// mtmd_draws.cpp - the reference's initial random state, reproduced.
//
// CSimulationBox::initSystem (SimulationBox.cpp:53-75) draws the initial state
// with the C library's rand(), in this order:
//   1. thermostat chain positions, eta_j = rand() / RAND_MAX (NHChain.cpp:126-132);
//   2. the barostat chain's positions, the same;
//   3. momenta, each component a rand() - 1 with a = 2 / RAND_MAX, a zero
//      replaced by 1, immobile atoms zeroed, then all scaled to the temperature
//      (SimulationBox.cpp:98-133);
//   4. thermostat chain momenta, p_eta_j = (a rand() - 1) Q_j / tau, a zero
//      replaced by Q_j / tau (NHChain.cpp:134-143);
//   5. the barostat chain's momenta, the same;
//   6. the barostat momentum, p_eps = (a rand() - 1) W / tau_P, W = T tau_P^2
//      (VelVerlet.cpp:176-183, SimulationBox.cpp:73).
// The masses (Q) follow CNHChain::prepareArrays (NHChain.cpp:115-124): the
// thermostat's Q_0 = 3 N T tau^2, the barostat chain's (prepared for one
// particle in one dimension, SimulationBox.cpp:81-86) Q_0 = T tau_P^2, and
// Q_j = T tau^2 for j > 0. The temperature scaling of step 3 runs while every
// particle still has the default mass of 1 (CParticle3D::CParticle3D; the
// masses are copied in afterwards by copyPosFromState, SimulationBox.cpp:74,
// 88-96), so the momenta are scaled to T with unit masses.
//
// The generator is glibc's rand(): the additive feedback generator of
// random_r (TYPE_3, degree 31, separation 3), seeded as srandom_r seeds it. A
// first `moldyn run` in a fresh MolTwister process draws with seed 1, the C
// library's initial state (CMDLoop::runSimulation reseeds from the clock only
// after the box is built, MDLoop.cpp:118).
#include "mtmd_internal.h"
#include "../../include/mtmd.h"

#include <cmath>
#include <cstdint>
#include <vector>

namespace {

// glibc rand(): r[i] = r[i - 31] + r[i - 3] (mod 2^32), output r[i] >> 1.
// srand(s): r[0] = s (1 for s = 0), r[i] = 16807 r[i - 1] mod (2^31 - 1) for
// 0 < i < 31 by Schrage's factorisation (127773 = m div 16807, 2836 = m mod
// 16807) on a signed 32-bit word, then 310 outputs discarded.
class GlibcRand {
public:
    explicit GlibcRand(uint32_t seed)
    {
        if (seed == 0) seed = 1;
        r_[0] = seed;
        int32_t w = (int32_t)seed;            // a 32-bit signed word: seeds above 2^31 start negative
        for (int i = 1; i < 31; i++) {
            const int64_t hi = w / 127773, lo = w % 127773;
            w = (int32_t)(16807 * lo - 2836 * hi);
            if (w < 0) w += 2147483647;
            r_[i] = (uint32_t)w;
        }
        f_ = 3;
        b_ = 0;
        for (int i = 0; i < 310; i++) next();
    }
    int32_t next()
    {
        r_[f_] += r_[b_];
        const int32_t out = (int32_t)(r_[f_] >> 1);
        f_ = f_ == 30 ? 0 : f_ + 1;
        b_ = b_ == 30 ? 0 : b_ + 1;
        return out;
    }

private:
    uint32_t r_[31];
    int f_, b_;
};

const double kRandMax = 2147483647.0;   // glibc RAND_MAX

} // namespace

extern "C" int mtmd_reference_draws(const mtmd_system* s, const mtmd_config* c, uint32_t seed, double* mom,
                                    double* eta_T, double* p_eta_T, double* eta_P, double* p_eta_P, double* p_eps)
{
    if (!s || !c || s->n_atoms < 1 || c->temperature_nh_len < 0 || c->pressure_nh_len < 0) return MTMD_EINVAL;
    const int n = s->n_atoms, mt = c->temperature_nh_len, mp = c->pressure_nh_len;
    const double T = c->temperature / MT_CONV_T;
    const double tau_t = c->temperature_relax / MT_CONV_t, tau_p = c->pressure_relax / MT_CONV_t;
    const double a = 2.0 / kRandMax;
    std::vector<double> qt(mt), qp(mp);
    for (int j = 0; j < mt; j++) qt[j] = (j == 0 ? double(3 * n) : 1.0) * T * tau_t * tau_t;
    for (int j = 0; j < mp; j++) qp[j] = (j == 0 ? double(1 * 1) : 1.0) * T * tau_p * tau_p;

    GlibcRand g(seed);
    std::vector<double> et(mt), ep(mp), pt(mt), pp(mp), p(3 * (size_t)n);
    for (int j = 0; j < mt; j++) et[j] = double(g.next()) / kRandMax;
    for (int j = 0; j < mp; j++) ep[j] = double(g.next()) / kRandMax;
    for (int k = 0; k < n; k++)
        for (int d = 0; d < 3; d++) {
            double v = a * double(g.next()) - 1.0;
            if (v == 0.0) v = 1.0;
            p[3 * (size_t)k + d] = v;
        }
    if (s->mobile)
        for (int k = 0; k < n; k++)
            if (!s->mobile[k]) p[3 * (size_t)k] = p[3 * (size_t)k + 1] = p[3 * (size_t)k + 2] = 0.0;
    double tsum = 0.0;
    for (int k = 0; k < n; k++) {
        const double* q = &p[3 * (size_t)k];
        tsum += ((q[0] * q[0]) + (q[1] * q[1]) + (q[2] * q[2])) / 1.0;   // unit masses (see above)
    }
    const double tnow = tsum / (3 * double(n));
    const double scale = std::sqrt(T / tnow);
    for (double& v : p) v *= scale;
    for (int j = 0; j < mt; j++) {
        pt[j] = (a * double(g.next()) - 1.0) * (qt[j] / tau_t);
        if (pt[j] == 0.0) pt[j] = qt[j] / tau_t;
    }
    for (int j = 0; j < mp; j++) {
        pp[j] = (a * double(g.next()) - 1.0) * (qp[j] / tau_p);
        if (pp[j] == 0.0) pp[j] = qp[j] / tau_p;
    }
    const double W = T * tau_p * tau_p;
    double pe = (a * double(g.next()) - 1.0) * (W / tau_p);
    if (pe == 0.0) pe = W / tau_p;

    if (mom) for (size_t i = 0; i < p.size(); i++) mom[i] = p[i];
    for (int j = 0; j < mt; j++) {
        if (eta_T) eta_T[j] = et[j];
        if (p_eta_T) p_eta_T[j] = pt[j];
    }
    for (int j = 0; j < mp; j++) {
        if (eta_P) eta_P[j] = ep[j];
        if (p_eta_P) p_eta_P[j] = pp[j];
    }
    if (p_eps) *p_eps = pe;
    return MTMD_OK;
}
