// mtmd_fm.hip - force matching: loss and its gradient with respect to the
// force-field parameters over a set of frames, by a hand-written adjoint.
//
// Model (the analytic forms of the reference's MDFF classes; the engine's
// 1000-point tables are not differentiable at the step sizes a gradient
// check needs):
//   LJ         U = 4 eps [(sigma/r)^12 - (sigma/r)^6]     (MolTwisterMDFFNonBonded_LJ.cpp:46-60, alpha = 1)
//              between atoms of different molecules, truncated at the cutoff,
//              minimum image (same-molecule pairs carry the default scale 0)
//   bond       U = k (r - r0)^2                           (MolTwisterMDFFBond_Harm.cpp:43-51)
//   angle      U = k (theta - theta0)^2, theta0 in degrees (MolTwisterMDFFAngle_Harm.cpp:47-55)
// Charges are not part of the model: a parameter-free Coulomb term is the
// same in the model and in the target and cancels in the residual.
//
// Loss over F frames of N atoms:  L = sum_f sum_i |F_i - F_i^ref|^2 / (3 N F).
// With R_i = 2 (F_i - F_i^ref) / (3 N F), dL/dtheta = sum_i R_i . dF_i/dtheta:
// a forward pass gives F, then an adjoint pass contracts R with the
// parameter derivatives of every force term (pair terms over the full pair
// set, each ordered pair once, so both ends are covered). Pairs are found
// again in every pass by a tiled minimum-image search over the atoms that
// carry LJ parameters; no neighbour list is stored.
//
// Parameter vector: for each non-bonded entry (eps, sigma), then for each
// bond type (k, r0), then for each angle type (k, theta0 [deg]).
// Template R = float: pair geometry and force terms in FP32, per-atom and
// gradient accumulation in FP64. R = double: FP64 throughout (the FP64
// reference gradient, and the model for finite differences).
#include "mtmd_internal.h"
#include "../../include/mtmd.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#define FM_TPB 256
#define FM_MAXE 4                                   // LJ entries
#define FM_MAXB 4                                   // bond types
#define FM_MAXA 4                                   // angle types
#define FM_MAXP (2 * (FM_MAXE + FM_MAXB + FM_MAXA))

static inline int fm_blk(int n, int b) { return (n + b - 1) / b; }

struct FmDev {
    int n, nf, nlj, T;
    double L[3];
    double rc2;
    const double* pos;      // [nf][n][3]
    const double* fref;     // [nf][n][3]
    double* F;              // [nf][n][3] model forces of the current parameters
    const int* lj_atoms;    // atoms with an LJ-carrying type, ascending
    const int* type; const int* mol;
    const int* pair_param;  // [T*T] first parameter index (eps) of the pair's entry, -1 if none
    // per-atom CSR: bonds (partner, type) and angles (other1, centre, other2, type, role)
    const int* b_first; const int* b_num; const int* b_j; const int* b_t;
    const int* a_first; const int* a_num; const int* a_idx;   // into the global angle list
    const int* ang;          // [na][4]: a, centre, b, type
    int bond_p0, angle_p0;   // first parameter index of bond / angle types
    const double* prm;       // device copy of the parameter vector
    double* part;            // per-block partial sums [blocks][np + 1]
    int np;
};

template <typename R>
__device__ __forceinline__ R fm_img(R d, double L)
{
    const R Lr = (R)L;
    return d - Lr * rint(d / Lr);
}

// forward: LJ part of F for every LJ atom of every frame, tiled search
template <typename R>
__global__ void __launch_bounds__(FM_TPB) k_fm_lj_forward(FmDev d)
{
    __shared__ R sx[FM_TPB], sy[FM_TPB], sz[FM_TPB];
    __shared__ int st[FM_TPB], sm[FM_TPB];
    const int f = blockIdx.y;
    const int li = blockIdx.x * FM_TPB + threadIdx.x;
    const bool act = li < d.nlj;
    const int i = act ? d.lj_atoms[li] : 0;
    const double* P = d.pos + (size_t)f * d.n * 3;
    const R xi = (R)P[3 * i], yi = (R)P[3 * i + 1], zi = (R)P[3 * i + 2];
    const int ti = d.type[i], mi = d.mol[i];
    double Fx = 0.0, Fy = 0.0, Fz = 0.0;
    for (int base = 0; base < d.nlj; base += FM_TPB) {
        __syncthreads();
        const int lj = base + threadIdx.x;
        if (lj < d.nlj) {
            const int j = d.lj_atoms[lj];
            sx[threadIdx.x] = (R)P[3 * j]; sy[threadIdx.x] = (R)P[3 * j + 1]; sz[threadIdx.x] = (R)P[3 * j + 2];
            st[threadIdx.x] = d.type[j]; sm[threadIdx.x] = d.mol[j];
        }
        __syncthreads();
        if (!act) continue;
        const int lim = min(FM_TPB, d.nlj - base);
        for (int e = 0; e < lim; e++) {
            if (sm[e] == mi) continue;
            const int pp = d.pair_param[ti * d.T + st[e]];
            if (pp < 0) continue;
            const R dx = fm_img<R>(xi - sx[e], d.L[0]), dy = fm_img<R>(yi - sy[e], d.L[1]),
                    dz = fm_img<R>(zi - sz[e], d.L[2]);
            const R r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < (R)d.rc2) || r2 == (R)0) continue;
            const R eps = (R)d.prm[pp], sig = (R)d.prm[pp + 1];
            const R s2 = sig * sig / r2, s6 = s2 * s2 * s2, s12 = s6 * s6;
            const R fr = (R)24 * eps * ((R)2 * s12 - s6) / r2;       // f(r)/r
            Fx += (double)(fr * dx); Fy += (double)(fr * dy); Fz += (double)(fr * dz);
        }
    }
    if (act) {
        double* F = d.F + ((size_t)f * d.n + i) * 3;
        F[0] = Fx; F[1] = Fy; F[2] = Fz;
    }
}

// harmonic angle at centre c for the pair of arms u = ra - rc, v = rb - rc:
// theta and d theta / d ra, d theta / d rb (d/d rc = minus their sum)
template <typename R>
__device__ __forceinline__ R fm_angle(const R u[3], const R v[3], R ga[3], R gb[3])
{
    const R uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const R vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const R uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
    const R nu = sqrt(uu), nv = sqrt(vv);
    R c = uv / (nu * nv);
    c = c > (R)1 ? (R)1 : (c < (R)-1 ? (R)-1 : c);
    const R th = acos(c);
    R s = sqrt((R)1 - c * c);
    s = s < (R)1e-12 ? (R)1e-12 : s;
    for (int q = 0; q < 3; q++) {
        ga[q] = -(v[q] / (nu * nv) - c * u[q] / uu) / s;
        gb[q] = -(u[q] / (nu * nv) - c * v[q] / vv) / s;
    }
    return th;
}

template <typename R>
__device__ __forceinline__ void fm_load(const double* P, int a, R x[3])
{
    x[0] = (R)P[3 * a]; x[1] = (R)P[3 * a + 1]; x[2] = (R)P[3 * a + 2];
}

// forward: bonded forces added to F (every atom of every frame)
template <typename R>
__global__ void __launch_bounds__(FM_TPB) k_fm_bonded_forward(FmDev d)
{
    const int f = blockIdx.y;
    const int i = blockIdx.x * FM_TPB + threadIdx.x;
    if (i >= d.n) return;
    const double* P = d.pos + (size_t)f * d.n * 3;
    R xi[3];
    fm_load<R>(P, i, xi);
    double F[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < d.b_num[i]; q++) {
        const int j = d.b_j[d.b_first[i] + q], bt = d.b_t[d.b_first[i] + q];
        R xj[3], dd[3];
        fm_load<R>(P, j, xj);
        for (int a = 0; a < 3; a++) dd[a] = fm_img<R>(xi[a] - xj[a], d.L[a]);
        const R r = sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
        const R k = (R)d.prm[d.bond_p0 + 2 * bt], r0 = (R)d.prm[d.bond_p0 + 2 * bt + 1];
        const R fr = (R)-2 * k * (r - r0) / r;
        for (int a = 0; a < 3; a++) F[a] += (double)(fr * dd[a]);
    }
    for (int q = 0; q < d.a_num[i]; q++) {
        const int g = d.a_idx[d.a_first[i] + q];
        const int A = d.ang[4 * g], C = d.ang[4 * g + 1], B = d.ang[4 * g + 2], at = d.ang[4 * g + 3];
        R xa[3], xb[3], xc[3], u[3], v[3], ga[3], gb[3];
        fm_load<R>(P, A, xa); fm_load<R>(P, B, xb); fm_load<R>(P, C, xc);
        for (int a = 0; a < 3; a++) { u[a] = fm_img<R>(xa[a] - xc[a], d.L[a]); v[a] = fm_img<R>(xb[a] - xc[a], d.L[a]); }
        const R th = fm_angle<R>(u, v, ga, gb);
        const R k = (R)d.prm[d.angle_p0 + 2 * at];
        const R th0 = (R)(d.prm[d.angle_p0 + 2 * at + 1] * (M_PI / 180.0));
        const R m = (R)-2 * k * (th - th0);
        for (int a = 0; a < 3; a++) {
            const R gi = (i == A) ? ga[a] : ((i == B) ? gb[a] : -(ga[a] + gb[a]));
            F[a] += (double)(m * gi);
        }
    }
    double* Fo = d.F + ((size_t)f * d.n + i) * 3;
    Fo[0] += F[0]; Fo[1] += F[1]; Fo[2] += F[2];
}

// block sums of each thread's nv partial sums (nv <= FM_MAXP + 1) into out[nv]
__device__ __forceinline__ void fm_block_store(const double (&vals)[FM_MAXP + 1], int nv, double* out)
{
    __shared__ double sh[FM_MAXP + 1][FM_TPB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < FM_MAXP + 1; q++) {
        double v = vals[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) sh[q][w] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < nv) {
        double s = 0.0;
        for (int ww = 0; ww < FM_TPB / 64; ww++) s += sh[threadIdx.x][ww];
        out[threadIdx.x] = s;
    }
}

__device__ __forceinline__ void fm_residual(const FmDev& d, int f, int i, double scale, double R_[3], double& l)
{
    const double* F = d.F + ((size_t)f * d.n + i) * 3;
    const double* G = d.fref + ((size_t)f * d.n + i) * 3;
    l = 0.0;
    for (int a = 0; a < 3; a++) {
        const double e = F[a] - G[a];
        l += e * e;
        R_[a] = 2.0 * scale * e;
    }
    l *= scale;
}

// adjoint of the LJ terms: sum_i R_i . d F_i / d (eps, sigma) per entry
template <typename R>
__global__ void __launch_bounds__(FM_TPB) k_fm_lj_adjoint(FmDev d, double scale)
{
    __shared__ R sx[FM_TPB], sy[FM_TPB], sz[FM_TPB];
    __shared__ int st[FM_TPB], sm[FM_TPB];
    const int f = blockIdx.y;
    const int li = blockIdx.x * FM_TPB + threadIdx.x;
    const bool act = li < d.nlj;
    const int i = act ? d.lj_atoms[li] : 0;
    const double* P = d.pos + (size_t)f * d.n * 3;
    const R xi = (R)P[3 * i], yi = (R)P[3 * i + 1], zi = (R)P[3 * i + 2];
    const int ti = d.type[i], mi = d.mol[i];
    double Ri[3], dummy;
    fm_residual(d, f, i, scale, Ri, dummy);
    const R rx = (R)Ri[0], ry = (R)Ri[1], rz = (R)Ri[2];
    double g[FM_MAXP + 1];
#pragma unroll
    for (int q = 0; q < FM_MAXP + 1; q++) g[q] = 0.0;
    for (int base = 0; base < d.nlj; base += FM_TPB) {
        __syncthreads();
        const int lj = base + threadIdx.x;
        if (lj < d.nlj) {
            const int j = d.lj_atoms[lj];
            sx[threadIdx.x] = (R)P[3 * j]; sy[threadIdx.x] = (R)P[3 * j + 1]; sz[threadIdx.x] = (R)P[3 * j + 2];
            st[threadIdx.x] = d.type[j]; sm[threadIdx.x] = d.mol[j];
        }
        __syncthreads();
        if (!act) continue;
        const int lim = min(FM_TPB, d.nlj - base);
        for (int e = 0; e < lim; e++) {
            if (sm[e] == mi) continue;
            const int pp = d.pair_param[ti * d.T + st[e]];
            if (pp < 0) continue;
            const R dx = fm_img<R>(xi - sx[e], d.L[0]), dy = fm_img<R>(yi - sy[e], d.L[1]),
                    dz = fm_img<R>(zi - sz[e], d.L[2]);
            const R r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < (R)d.rc2) || r2 == (R)0) continue;
            const R eps = (R)d.prm[pp], sig = (R)d.prm[pp + 1];
            const R s2 = sig * sig / r2, s6 = s2 * s2 * s2, s12 = s6 * s6;
            const R proj = rx * dx + ry * dy + rz * dz;                // R_i . (r_i - r_j)
            // f/r = 24 eps (2 s12 - s6) / r^2
            const R dfe = (R)24 * ((R)2 * s12 - s6) / r2;                // d(f/r)/d eps
            const R dfs = (R)144 * eps * ((R)4 * s12 - s6) / (r2 * sig); // d(f/r)/d sigma
            // the same pair is counted again from j with -R_j: both ends covered;
            // the accumulators stay in registers (constant indices after unrolling)
#pragma unroll
            for (int q = 0; q < FM_MAXE; q++)
                if (2 * q == pp) {
                    g[2 * q] += (double)(dfe * proj);
                    g[2 * q + 1] += (double)(dfs * proj);
                }
        }
    }
    double* out = d.part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (d.np + 1);
    if (!act) {
#pragma unroll
        for (int q = 0; q < FM_MAXP + 1; q++) g[q] = 0.0;
    }
    fm_block_store(g, d.np + 1, out);
}

// adjoint of the bonded terms, plus the loss itself (one thread per atom)
template <typename R>
__global__ void __launch_bounds__(FM_TPB) k_fm_bonded_adjoint(FmDev d, double scale)
{
    const int f = blockIdx.y;
    const int i = blockIdx.x * FM_TPB + threadIdx.x;
    const bool act = i < d.n;
    double g[FM_MAXP + 1];
#pragma unroll
    for (int q = 0; q < FM_MAXP + 1; q++) g[q] = 0.0;
    double lsum = 0.0;
    if (act) {
        const double* P = d.pos + (size_t)f * d.n * 3;
        double Ri[3], l;
        fm_residual(d, f, i, scale, Ri, l);
        lsum = l;
        R xi[3];
        fm_load<R>(P, i, xi);
        for (int q = 0; q < d.b_num[i]; q++) {
            const int j = d.b_j[d.b_first[i] + q], bt = d.b_t[d.b_first[i] + q];
            R xj[3], dd[3];
            fm_load<R>(P, j, xj);
            for (int a = 0; a < 3; a++) dd[a] = fm_img<R>(xi[a] - xj[a], d.L[a]);
            const R r = sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
            const R k = (R)d.prm[d.bond_p0 + 2 * bt], r0 = (R)d.prm[d.bond_p0 + 2 * bt + 1];
            R proj = 0;
            for (int a = 0; a < 3; a++) proj += (R)Ri[a] * dd[a];
            // F_i = -2 k (r - r0) dd / r
            const double gk = (double)((R)-2 * (r - r0) / r * proj), gr = (double)((R)2 * k / r * proj);
#pragma unroll
            for (int q = 0; q < FM_MAXB; q++)
                if (q == bt) { g[2 * FM_MAXE + 2 * q] += gk; g[2 * FM_MAXE + 2 * q + 1] += gr; }
        }
        for (int q = 0; q < d.a_num[i]; q++) {
            const int gg = d.a_idx[d.a_first[i] + q];
            const int A = d.ang[4 * gg], C = d.ang[4 * gg + 1], B = d.ang[4 * gg + 2], at = d.ang[4 * gg + 3];
            R xa[3], xb[3], xc[3], u[3], v[3], ga[3], gb[3];
            fm_load<R>(P, A, xa); fm_load<R>(P, B, xb); fm_load<R>(P, C, xc);
            for (int a = 0; a < 3; a++) {
                u[a] = fm_img<R>(xa[a] - xc[a], d.L[a]);
                v[a] = fm_img<R>(xb[a] - xc[a], d.L[a]);
            }
            const R th = fm_angle<R>(u, v, ga, gb);
            const R k = (R)d.prm[d.angle_p0 + 2 * at];
            const R th0 = (R)(d.prm[d.angle_p0 + 2 * at + 1] * (M_PI / 180.0));
            R proj = 0;
            for (int a = 0; a < 3; a++) {
                const R gi = (i == A) ? ga[a] : ((i == B) ? gb[a] : -(ga[a] + gb[a]));
                proj += (R)Ri[a] * gi;
            }
            // F_i = -2 k (theta - theta0) d theta / d r_i
            const double gk = (double)((R)-2 * (th - th0) * proj);
            const double gt = (double)((R)2 * k * proj) * (M_PI / 180.0);
#pragma unroll
            for (int q = 0; q < FM_MAXA; q++)
                if (q == at) { g[2 * (FM_MAXE + FM_MAXB) + 2 * q] += gk; g[2 * (FM_MAXE + FM_MAXB) + 2 * q + 1] += gt; }
        }
    }
    // compact the fixed-slot accumulators into the parameter layout
    double v[FM_MAXP + 1];
    const int ne = d.bond_p0 / 2, nbt = (d.angle_p0 - d.bond_p0) / 2, nat = (d.np - d.angle_p0) / 2;
#pragma unroll
    for (int q = 0; q < FM_MAXP + 1; q++) v[q] = 0.0;
#pragma unroll
    for (int q = 0; q < FM_MAXB; q++)
        if (q < nbt) {
#pragma unroll
            for (int t = 0; t < 2 * FM_MAXE + 2 * FM_MAXB; t++)
                if (t == 2 * ne + 2 * q) { v[t] = g[2 * FM_MAXE + 2 * q]; v[t + 1] = g[2 * FM_MAXE + 2 * q + 1]; }
        }
#pragma unroll
    for (int q = 0; q < FM_MAXA; q++)
        if (q < nat) {
#pragma unroll
            for (int t = 0; t < FM_MAXP; t++)
                if (t == 2 * (ne + nbt) + 2 * q) {
                    v[t] = g[2 * (FM_MAXE + FM_MAXB) + 2 * q];
                    v[t + 1] = g[2 * (FM_MAXE + FM_MAXB) + 2 * q + 1];
                }
        }
#pragma unroll
    for (int t = 0; t < FM_MAXP + 1; t++)
        if (t == d.np) v[t] = lsum;
    double* out = d.part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (d.np + 1);
    fm_block_store(v, d.np + 1, out);
}

__global__ void k_fm_zero(double* a, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = 0.0;
}

// ---------------------------------------------------------------------------
namespace {
thread_local std::string g_fm_err;

struct FmMem {
    std::vector<void*> ptrs;
    template <class T> T* alloc(size_t n)
    {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
    ~FmMem() { for (void* p : ptrs) (void)hipFree(p); }
};
} // namespace

struct mtmd_fm {
    FmMem mem;
    FmDev d{};
    hipStream_t stream = nullptr;
    int np = 0, nb_lj = 0, nb_atoms = 0;
    double* h_part = nullptr;
    std::string err;
    int fail(int code, const char* fmt, ...)
    {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    ~mtmd_fm()
    {
        if (h_part) (void)hipHostFree(h_part);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

#define FM_HIP(h, call)                                                                                       \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess) return (h)->fail(MTMD_EHIP, "%s: %s", #call, hipGetErrorString(e_));            \
    } while (0)

extern "C" const char* mtmd_fm_last_error(const mtmd_fm* h) { return h ? h->err.c_str() : g_fm_err.c_str(); }

extern "C" int mtmd_fm_num_params(const mtmd_fm* h) { return h ? h->np : -1; }

extern "C" int mtmd_fm_create(const mtmd_system* s, double cutoff, int n_frames, const double* frames,
                              const double* target, mtmd_fm** out)
{
    auto bad = [&](int code, const char* m) { g_fm_err = m; return code; };
    if (!s || !out || !frames || n_frames < 1 || !(cutoff > 0.0) || s->n_atoms < 1) return bad(MTMD_EINVAL, "invalid arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return bad(MTMD_ENODEV, "no HIP device");
    const int n = s->n_atoms, T = s->n_types;
    for (int a = 0; a < 3; a++)
        if (!(s->box[a] > 2.0 * cutoff)) return bad(MTMD_EBOX, "box below twice the cutoff");
    // parameters: per non-bonded entry (eps, sigma), per bond type (k, r0), per angle type (k, theta0)
    std::vector<int> pair_param((size_t)T * T, -1);
    for (int e = 0; e < s->n_nonbonded; e++) {
        const auto& nb = s->nonbonded[e];
        if (nb.kind != MTMD_NB_LJ) return bad(MTMD_EINVAL, "force matching takes LJ non-bonded entries");
        if (nb.type_a < 0 || nb.type_b < 0 || nb.type_a >= T || nb.type_b >= T) return bad(MTMD_EINVAL, "type out of range");
        if (pair_param[nb.type_a * T + nb.type_b] >= 0) return bad(MTMD_EINVAL, "one LJ entry per type pair");
        pair_param[nb.type_a * T + nb.type_b] = pair_param[nb.type_b * T + nb.type_a] = 2 * e;
    }
    for (int q = 0; q < s->n_bond_ff; q++)
        if (s->bond_ff[q].kind != MTMD_BOND_HARM) return bad(MTMD_EINVAL, "force matching takes harmonic bonds");
    for (int q = 0; q < s->n_angle_ff; q++)
        if (s->angle_ff[q].kind != MTMD_ANGLE_HARM) return bad(MTMD_EINVAL, "force matching takes harmonic angles");
    if (s->n_dihedrals) return bad(MTMD_EINVAL, "force matching takes no dihedrals");
    if (s->n_nonbonded > FM_MAXE || s->n_bond_ff > FM_MAXB || s->n_angle_ff > FM_MAXA)
        return bad(MTMD_EINVAL, "force matching takes at most 4 LJ entries, 4 bond and 4 angle types");
    const int np = 2 * (s->n_nonbonded + s->n_bond_ff + s->n_angle_ff);

    // molecules from bonds; per-atom bond and angle lists
    std::vector<int> par(n);
    for (int k = 0; k < n; k++) par[k] = k;
    auto find = [&](int a) { while (par[a] != a) { par[a] = par[par[a]]; a = par[a]; } return a; };
    for (int j = 0; j < s->n_bonds; j++) {
        int a = s->bonds[3 * j], b = s->bonds[3 * j + 1];
        if (a < 0 || b < 0 || a >= n || b >= n || s->bonds[3 * j + 2] < 0 || s->bonds[3 * j + 2] >= s->n_bond_ff)
            return bad(MTMD_EINVAL, "bond index out of range");
        a = find(a); b = find(b);
        if (a != b) par[std::max(a, b)] = std::min(a, b);
    }
    std::vector<int> mol(n), type(s->type, s->type + n);
    for (int k = 0; k < n; k++) mol[k] = find(k);
    std::vector<std::vector<int>> bl(n), bt(n), al(n);
    for (int j = 0; j < s->n_bonds; j++) {
        const int a = s->bonds[3 * j], b = s->bonds[3 * j + 1], t = s->bonds[3 * j + 2];
        bl[a].push_back(b); bt[a].push_back(t);
        bl[b].push_back(a); bt[b].push_back(t);
    }
    std::vector<int> ang(4 * (size_t)s->n_angles);
    for (int j = 0; j < s->n_angles; j++) {
        for (int q = 0; q < 4; q++) ang[4 * j + q] = s->angles[4 * j + q];
        for (int q = 0; q < 3; q++) {
            if (ang[4 * j + q] < 0 || ang[4 * j + q] >= n) return bad(MTMD_EINVAL, "angle index out of range");
            al[ang[4 * j + q]].push_back(j);
        }
        if (ang[4 * j + 3] < 0 || ang[4 * j + 3] >= s->n_angle_ff) return bad(MTMD_EINVAL, "angle type out of range");
    }
    std::vector<int> b_first(n), b_num(n), b_j, b_t, a_first(n), a_num(n), a_idx;
    for (int k = 0; k < n; k++) {
        b_first[k] = (int)b_j.size(); b_num[k] = (int)bl[k].size();
        b_j.insert(b_j.end(), bl[k].begin(), bl[k].end()); b_t.insert(b_t.end(), bt[k].begin(), bt[k].end());
        a_first[k] = (int)a_idx.size(); a_num[k] = (int)al[k].size();
        a_idx.insert(a_idx.end(), al[k].begin(), al[k].end());
    }
    std::vector<int> lj_atoms;
    for (int k = 0; k < n; k++) {
        if (type[k] < 0 || type[k] >= T) return bad(MTMD_EINVAL, "atom type out of range");
        bool any = false;
        for (int t = 0; t < T; t++) any = any || pair_param[type[k] * T + t] >= 0;
        if (any) lj_atoms.push_back(k);
    }

    auto h = new mtmd_fm();
    std::unique_ptr<mtmd_fm> guard(h);
    h->np = np;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return bad(MTMD_EHIP, "stream");
    FmDev& d = h->d;
    d.n = n; d.nf = n_frames; d.nlj = (int)lj_atoms.size(); d.T = T; d.np = np;
    for (int a = 0; a < 3; a++) d.L[a] = s->box[a];
    d.rc2 = cutoff * cutoff;
    d.bond_p0 = 2 * s->n_nonbonded;
    d.angle_p0 = d.bond_p0 + 2 * s->n_bond_ff;
    const size_t fsz = (size_t)n_frames * n * 3;
    double* pos = h->mem.alloc<double>(fsz);
    double* fref = h->mem.alloc<double>(fsz);
    d.F = h->mem.alloc<double>(fsz);
    int* lja = h->mem.alloc<int>(lj_atoms.size());
    int* typ = h->mem.alloc<int>(n); int* mo = h->mem.alloc<int>(n);
    int* pp = h->mem.alloc<int>(pair_param.size());
    int* bf = h->mem.alloc<int>(n); int* bn = h->mem.alloc<int>(n);
    int* bj = h->mem.alloc<int>(b_j.size()); int* btt = h->mem.alloc<int>(b_t.size());
    int* af = h->mem.alloc<int>(n); int* an = h->mem.alloc<int>(n); int* ai = h->mem.alloc<int>(a_idx.size());
    int* ag = h->mem.alloc<int>(ang.size());
    double* prm = h->mem.alloc<double>(np);
    h->nb_lj = fm_blk(std::max(d.nlj, 1), FM_TPB);
    h->nb_atoms = fm_blk(n, FM_TPB);
    const size_t nparts = (size_t)n_frames * (h->nb_lj + h->nb_atoms) * (np + 1);
    d.part = h->mem.alloc<double>(nparts);
    if (!pos || !fref || !d.F || !lja || !typ || !mo || !pp || !bf || !bn || !bj || !btt || !af || !an || !ai || !ag ||
        !prm || !d.part || hipHostMalloc((void**)&h->h_part, nparts * sizeof(double), 0) != hipSuccess)
        return bad(MTMD_ENOMEM, "device allocation failed");
    auto up = [&](void* dst, const void* src, size_t bytes) {
        return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    bool ok = up(pos, frames, fsz * sizeof(double)) && up(lja, lj_atoms.data(), lj_atoms.size() * sizeof(int)) &&
              up(typ, type.data(), n * sizeof(int)) && up(mo, mol.data(), n * sizeof(int)) &&
              up(pp, pair_param.data(), pair_param.size() * sizeof(int)) &&
              up(bf, b_first.data(), n * sizeof(int)) && up(bn, b_num.data(), n * sizeof(int)) &&
              up(bj, b_j.data(), b_j.size() * sizeof(int)) && up(btt, b_t.data(), b_t.size() * sizeof(int)) &&
              up(af, a_first.data(), n * sizeof(int)) && up(an, a_num.data(), n * sizeof(int)) &&
              up(ai, a_idx.data(), a_idx.size() * sizeof(int)) && up(ag, ang.data(), ang.size() * sizeof(int));
    if (target) ok = ok && up(fref, target, fsz * sizeof(double));
    else ok = ok && hipMemset(fref, 0, fsz * sizeof(double)) == hipSuccess;
    if (!ok) return bad(MTMD_EHIP, "upload failed");
    d.pos = pos; d.fref = fref; d.lj_atoms = lja; d.type = typ; d.mol = mo; d.pair_param = pp;
    d.b_first = bf; d.b_num = bn; d.b_j = bj; d.b_t = btt; d.a_first = af; d.a_num = an; d.a_idx = ai; d.ang = ag;
    d.prm = prm;
    *out = guard.release();
    return MTMD_OK;
}

extern "C" void mtmd_fm_destroy(mtmd_fm* h) { delete h; }

extern "C" int mtmd_fm_set_target(mtmd_fm* h, const double* target)
{
    if (!h || !target) return MTMD_EINVAL;
    FM_HIP(h, hipMemcpy((void*)h->d.fref, target, (size_t)h->d.nf * h->d.n * 3 * sizeof(double),
                        hipMemcpyHostToDevice));
    return MTMD_OK;
}

static int fm_forward(mtmd_fm* h, const double* params, int fp64)
{
    FmDev& d = h->d;
    FM_HIP(h, hipMemcpyAsync((void*)d.prm, params, d.np * sizeof(double), hipMemcpyHostToDevice, h->stream));
    const size_t fsz = (size_t)d.nf * d.n * 3;
    hipLaunchKernelGGL(k_fm_zero, dim3((unsigned)std::min<size_t>((fsz + 255) / 256, 65536)), dim3(256), 0,
                       h->stream, d.F, fsz);
    const dim3 glj(h->nb_lj, d.nf), gat(h->nb_atoms, d.nf);
    if (d.nlj > 0) {
        if (fp64) hipLaunchKernelGGL(k_fm_lj_forward<double>, glj, dim3(FM_TPB), 0, h->stream, d);
        else hipLaunchKernelGGL(k_fm_lj_forward<float>, glj, dim3(FM_TPB), 0, h->stream, d);
    }
    if (fp64) hipLaunchKernelGGL(k_fm_bonded_forward<double>, gat, dim3(FM_TPB), 0, h->stream, d);
    else hipLaunchKernelGGL(k_fm_bonded_forward<float>, gat, dim3(FM_TPB), 0, h->stream, d);
    FM_HIP(h, hipGetLastError());
    return MTMD_OK;
}

// model forces of one frame for the given parameters (layout [3 * n])
extern "C" int mtmd_fm_forces(mtmd_fm* h, const double* params, int32_t frame, int32_t fp64, double* F)
{
    if (!h || !params || !F || frame < 0 || frame >= h->d.nf) return MTMD_EINVAL;
    int rc = fm_forward(h, params, fp64);
    if (rc) return rc;
    FM_HIP(h, hipMemcpyAsync(F, h->d.F + (size_t)frame * h->d.n * 3, (size_t)h->d.n * 3 * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
    FM_HIP(h, hipStreamSynchronize(h->stream));
    return MTMD_OK;
}

// model forces of every frame for the given parameters, made the target
// (the synthetic reference forces of a force-matching problem)
extern "C" int mtmd_fm_target_from_params(mtmd_fm* h, const double* params)
{
    if (!h || !params) return MTMD_EINVAL;
    int rc = fm_forward(h, params, 1);
    if (rc) return rc;
    FM_HIP(h, hipMemcpyAsync((void*)h->d.fref, h->d.F, (size_t)h->d.nf * h->d.n * 3 * sizeof(double),
                             hipMemcpyDeviceToDevice, h->stream));
    FM_HIP(h, hipStreamSynchronize(h->stream));
    return MTMD_OK;
}

// loss and gradient for the given parameters; fp64 selects the FP64 model
extern "C" int mtmd_fm_loss_grad(mtmd_fm* h, const double* params, int32_t fp64, double* loss, double* grad)
{
    if (!h || !params) return MTMD_EINVAL;
    int rc = fm_forward(h, params, fp64);
    if (rc) return rc;
    FmDev& d = h->d;
    const double scale = 1.0 / (3.0 * (double)d.n * (double)d.nf);
    const size_t nlj = (size_t)d.nf * h->nb_lj * (d.np + 1), nat = (size_t)d.nf * h->nb_atoms * (d.np + 1);
    FmDev da = d;
    if (d.nlj > 0) {
        const dim3 glj(h->nb_lj, d.nf);
        if (fp64) hipLaunchKernelGGL(k_fm_lj_adjoint<double>, glj, dim3(FM_TPB), 0, h->stream, da, scale);
        else hipLaunchKernelGGL(k_fm_lj_adjoint<float>, glj, dim3(FM_TPB), 0, h->stream, da, scale);
    } else {
        hipLaunchKernelGGL(k_fm_zero, dim3(1), dim3(256), 0, h->stream, d.part, nlj);
    }
    FmDev db = d;
    db.part = d.part + nlj;
    const dim3 gat(h->nb_atoms, d.nf);
    if (fp64) hipLaunchKernelGGL(k_fm_bonded_adjoint<double>, gat, dim3(FM_TPB), 0, h->stream, db, scale);
    else hipLaunchKernelGGL(k_fm_bonded_adjoint<float>, gat, dim3(FM_TPB), 0, h->stream, db, scale);
    FM_HIP(h, hipGetLastError());
    FM_HIP(h, hipMemcpyAsync(h->h_part, d.part, (nlj + nat) * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    FM_HIP(h, hipStreamSynchronize(h->stream));
    // fixed-order sum of the block partials
    std::vector<double> g(d.np + 1, 0.0);
    for (size_t b = 0; b < (nlj + nat) / (d.np + 1); b++)
        for (int q = 0; q <= d.np; q++) g[q] += h->h_part[b * (d.np + 1) + q];
    if (loss) *loss = g[d.np];
    if (grad) std::copy(g.begin(), g.begin() + d.np, grad);
    return MTMD_OK;
}
