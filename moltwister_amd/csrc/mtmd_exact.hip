// mtmd_exact.hip - reference-exact kernels for gfx950 and the shared
// integrator/thermostat/reduction kernels.
//
// Built with -ffp-contract=off: every expression below keeps the precision and
// operation order of the reference CPU build (x86-64 SSE, no FMA), so cell
// lists, neighbour lists and per-atom forces come out bit-identical.
// Reference paths are relative to the MolTwister tree:
//   FCF = Cmd/MolTwisterCmdMolDyn/ForceFields/FunctorCalcForce.cpp
//   FGC = Cmd/MolTwisterCmdMolDyn/ForceFields/FunctorGenCellList.cpp
//   FGN = Cmd/MolTwisterCmdMolDyn/ForceFields/FunctorGenNeighList.cpp
//   V3  = Utilities/3DVector.cpp
#include "mtmd_internal.h"
#include "mtmd_device.h"
#include <float.h>

#define TPB 256

// ---------------------------------------------------------------------------
// small utilities
// ---------------------------------------------------------------------------
__global__ void k_fill_i32(int* p, int v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static inline int nblk(int n, int b) { return (n + b - 1) / b; }

int mt_reduce_blocks(int n) { int b = nblk(n, TPB); return b < 1 ? 1 : (b > 1024 ? 1024 : b); }

// ---------------------------------------------------------------------------
// cell list: CFunctorGenCellList::operator() (FGC:49-73) + kernelAssembleList
// (FGC:75-142) as count / scan / scatter / per-cell stable order
// ---------------------------------------------------------------------------
__global__ void k_exact_cell_keys(const double4* __restrict__ pos, int n, int Nx, int Ny, int Nz,
                                  float wx, float wy, float wz, float lx, float ly, float lz,
                                  int* __restrict__ key, int* __restrict__ cxyz, int* __restrict__ count,
                                  unsigned int* flags)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 r = pos[k];
    // non-finite coordinates would all clamp into cell 0; flag them so the
    // host reports MTMD_ENAN instead of building quadratic lists
    if (!(isfinite(r.x) && isfinite(r.y) && isfinite(r.z))) atomicOr(&flags[3], 1u);
    int ix = mt_d2i(floor((r.x - (double)lx) * (double)Nx / (double)wx));
    int iy = mt_d2i(floor((r.y - (double)ly) * (double)Ny / (double)wy));
    int iz = mt_d2i(floor((r.z - (double)lz) * (double)Nz / (double)wz));
    ix = (ix < 0) ? 0 : ((ix >= Nx) ? Nx - 1 : ix);
    iy = (iy < 0) ? 0 : ((iy >= Ny) ? Ny - 1 : iy);
    iz = (iz < 0) ? 0 : ((iz >= Nz) ? Nz - 1 : iz);
    int c = ix + Nx * (iy + iz * Ny);
    key[k] = c;
    if (cxyz) { cxyz[3 * k] = ix; cxyz[3 * k + 1] = iy; cxyz[3 * k + 2] = iz; }
    atomicAdd(&count[c], 1);
}

// single-workgroup exclusive scan over the cells: first = -1 for empty cells
// (the reference assigns starts only to non-empty cells, FGC:93-103)
__global__ void __launch_bounds__(1024) k_scan_cells(const int* __restrict__ count, int* __restrict__ first,
                                                     int* __restrict__ fill, int ncell)
{
    __shared__ int sums[1024];
    int t = threadIdx.x;
    int chunk = (ncell + 1023) / 1024;
    int b = t * chunk, e = min(b + chunk, ncell);
    int s = 0;
    for (int c = b; c < e; c++) s += count[c];
    sums[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = (t >= off) ? sums[t - off] : 0;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    int run = sums[t] - s;
    for (int c = b; c < e; c++) {
        int cnt = count[c];
        first[c] = cnt ? run : -1;
        fill[c] = 0;
        run += cnt;
    }
}

__global__ void k_cell_scatter(const int* __restrict__ key, int n, const int* __restrict__ first,
                               int* __restrict__ fill, int* __restrict__ list)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int c = key[k];
    int slot = atomicAdd(&fill[c], 1);
    list[first[c] + slot] = k;
}

// ascending atom index inside each cell (the reference fills cells in atom order)
__global__ void __launch_bounds__(TPB) k_cell_sort(const int* __restrict__ first, const int* __restrict__ count,
                                                   int* __restrict__ list, int ncell, const unsigned int* flags)
{
    __shared__ __align__(16) int lds[(TPB / 64) * MT_CELL_SORT_LDS];
    if (flags[3]) return;                                        // uniform over the grid
    const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int cnt = c < ncell ? count[c] : 0;
    mt_wave_sort_cell(list + (c < ncell ? first[c] : 0), cnt, lds + (threadIdx.x >> 6) * MT_CELL_SORT_LDS);
}

void mt_exact_cells(const MtDev& d, const MtParams& p, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(d.ncells, TPB)), dim3(TPB), 0, s, d.cell_count, 0, d.ncells);
    hipLaunchKernelGGL(k_exact_cell_keys, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.pos, d.n, d.nc[0], d.nc[1],
                       d.nc[2], p.Lf[0], p.Lf[1], p.Lf[2], p.lowf[0], p.lowf[1], p.lowf[2], d.cell_key,
                       d.cell_xyz, d.cell_count, d.flags);
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, s, d.cell_count, d.cell_first, d.cell_fill, d.ncells);
    hipLaunchKernelGGL(k_cell_scatter, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.cell_key, d.n, d.cell_first,
                       d.cell_fill, d.cell_list);
    hipLaunchKernelGGL(k_cell_sort, dim3(nblk(d.ncells, TPB / 64)), dim3(TPB), 0, s, d.cell_first, d.cell_count,
                       d.cell_list, d.ncells, d.flags);
}

// ---------------------------------------------------------------------------
// neighbour list: CFunctorGenNeighList::operator() (FGN:72-151), same cell
// traversal (ix outer, iz inner, wrapped with the reference's rule), same
// predicate d2 < double(float rc^2), self skipped. Rows are column-major.
// ---------------------------------------------------------------------------
__global__ void k_exact_neigh(const double4* __restrict__ pos, int n, const int* __restrict__ cxyz,
                              const int* __restrict__ cfirst, const int* __restrict__ ccount,
                              const int* __restrict__ clist, int Nx, int Ny, int Nz,
                              double wx, double wy, double wz, float rc2, int ncap,
                              int* __restrict__ ncount, uint32_t* __restrict__ nl, int* __restrict__ nreq,
                              const unsigned int* flags)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (flags[3]) { ncount[k] = 0; return; }
    const double w[3] = {wx, wy, wz};
    double4 rk4 = pos[k];
    mt_d3 rk = {rk4.x, rk4.y, rk4.z};
    int cx = cxyz[3 * k], cy = cxyz[3 * k + 1], cz = cxyz[3 * k + 2];
    int cnt = 0;
    for (int ix = cx - 1; ix <= cx + 1; ix++) {
        int Ix = (ix >= 0) ? ((ix < Nx) ? ix : 0) : Nx - 1;
        for (int iy = cy - 1; iy <= cy + 1; iy++) {
            int Iy = (iy >= 0) ? ((iy < Ny) ? iy : 0) : Ny - 1;
            for (int iz = cz - 1; iz <= cz + 1; iz++) {
                int Iz = (iz >= 0) ? ((iz < Nz) ? iz : 0) : Nz - 1;
                int c = Ix + Nx * (Iy + Iz * Ny);
                int num = ccount[c], f0 = cfirst[c];
                for (int e = 0; e < num; e++) {
                    int a = clist[f0 + e];
                    double4 ra4 = pos[a];
                    mt_d3 ra = {ra4.x, ra4.y, ra4.z};
                    double d2 = mt_dist2_pbc(rk, ra, w);
                    if (d2 < (double)rc2) {
                        if (a == k) continue;
                        if (cnt < ncap) nl[(size_t)cnt * n + k] = (uint32_t)a;
                        cnt++;
                    }
                }
            }
        }
    }
    ncount[k] = cnt;
    if (cnt > ncap) atomicMax(nreq, cnt);
}

void mt_exact_neighbors(const MtDev& d, const MtParams& p, hipStream_t s)
{
    hipLaunchKernelGGL(k_exact_neigh, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.pos, d.n, d.cell_xyz, d.cell_first,
                       d.cell_count, d.cell_list, d.nc[0], d.nc[1], d.nc[2], p.wnl[0], p.wnl[1], p.wnl[2], p.rc2f,
                       d.ncap, d.ncount, d.nlist, d.nreq, d.flags);
}

// ---------------------------------------------------------------------------
// force functor (FCF:159-281): bonded, neighbours, then (separate kernel)
// all-pairs Coulomb continuing the same per-atom accumulation in index order
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(TPB) k_exact_force(MtDev d, MtParams p)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.n) return;
    mt_d3 F = {0.0, 0.0, 0.0};
    float U = 0.0f;
    if (!d.mobile[k]) { d.frc[k] = make_double4(0.0, 0.0, 0.0, 0.0); d.epot[k] = 0.0f; return; }
    mt_exact_bonded(d, p, k, F, U);
    int molk = d.mol[k];
    double4 rk4 = d.pos[k];
    mt_d3 rk = {rk4.x, rk4.y, rk4.z};
    int cnt = d.ncount[k];
    int T = d.T, row = d.type[k];
    for (int j = 0; j < cnt; j++) {
        int ia = (int)d.nlist[(size_t)j * d.n + k];
        double4 ri4 = d.pos[ia];
        mt_d3 ri = mt_move_same_side(rk, {ri4.x, ri4.y, ri4.z}, p.wbox);
        mt_d3 f;
        float u = 0.0f;
        float r = (float)mt_norm(mt_sub(rk, ri));
        if (r < 1E-3f) {
            f = {0.0, 0.0, 0.0};
        } else {
            int col = d.type[ia];
            int nff = d.nbcnt[row * T + col];
            float mdU = 0.0f;
            for (int ff = 0; ff < nff; ff++) {
                const float4* tab = d.nbtab + ((size_t)(row * T + col) * MT_MAXFF + ff) * MT_NPTS;
                mt_table_eval(tab, r, ia > k, mdU, u);
            }
            f = mt_mul(mt_coeffs12(ri, rk), (double)mdU);
        }
        mt_scale_pair(d, p, f, u, ia, molk, k);
        F = mt_add(F, f);
        U += u;
    }
    if (p.coulomb_on && !p.coulomb_allpairs) {
        for (int j = 0; j < cnt; j++) {
            int ia = (int)d.nlist[(size_t)j * d.n + k];
            double4 ri4 = d.pos[ia];
            mt_d3 ri = mt_move_same_side(rk, {ri4.x, ri4.y, ri4.z}, p.wbox);
            float u = 0.0f;
            mt_d3 f = mt_coulomb_pair(rk, ri, (float)rk4.w, (float)ri4.w, k, ia, u);
            mt_scale_pair(d, p, f, u, ia, molk, k);
            F = mt_add(F, f);
            U += u;
        }
    }
    d.frc[k] = make_double4(F.x, F.y, F.z, 0.0);
    d.epot[k] = U;
}

// all-pairs minimum-image Coulomb (FCF:258-278, 514-535): each workgroup owns
// TPB atoms k and walks all atoms i in index order through LDS tiles
__global__ void __launch_bounds__(TPB) k_exact_coulomb_allpairs(MtDev d, MtParams p)
{
    __shared__ double sx[TPB], sy[TPB], sz[TPB];
    __shared__ float sq[TPB];
    __shared__ int smol[TPB];
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = (k < d.n) && d.mobile[k];
    mt_d3 F = {0.0, 0.0, 0.0}, rk = {0.0, 0.0, 0.0};
    float U = 0.0f, qk = 0.0f;
    int molk = -1, xk = 0;
    if (active) {
        double4 f4 = d.frc[k];
        F = {f4.x, f4.y, f4.z};
        U = d.epot[k];
        double4 r4 = d.pos[k];
        rk = {r4.x, r4.y, r4.z};
        qk = (float)r4.w;
        molk = d.mol[k];
        xk = d.xmol[k];
    }
    for (int base = 0; base < d.n; base += TPB) {
        int i = base + threadIdx.x;
        __syncthreads();
        if (i < d.n) {
            double4 r4 = d.pos[i];
            sx[threadIdx.x] = r4.x; sy[threadIdx.x] = r4.y; sz[threadIdx.x] = r4.z;
            sq[threadIdx.x] = (float)r4.w;
            smol[threadIdx.x] = d.mol[i];
        }
        __syncthreads();
        if (!active || qk == 0.0f) continue;   // every term is +-0: the sum is unchanged
        int lim = min(TPB, d.n - base);
        for (int t = 0; t < lim; t++) {
            float qi = sq[t];
            int ia = base + t;
            if (qi == 0.0f) continue;
            mt_d3 ri = mt_move_same_side(rk, {sx[t], sy[t], sz[t]}, p.wbox);
            float u = 0.0f;
            mt_d3 f = mt_coulomb_pair(rk, ri, qk, qi, k, ia, u);
            if (xk || (smol[t] == molk && molk >= 0)) mt_scale_pair(d, p, f, u, ia, molk, k);
            F = mt_add(F, f);
            U += u;
        }
    }
    if (active) {
        d.frc[k] = make_double4(F.x, F.y, F.z, 0.0);
        d.epot[k] = U;
    }
}

void mt_exact_forces(const MtDev& d, const MtParams& p, hipStream_t s)
{
    hipLaunchKernelGGL(k_exact_force, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d, p);
    if (p.coulomb_on && p.coulomb_allpairs)
        hipLaunchKernelGGL(k_exact_coulomb_allpairs, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d, p);
}

// ---------------------------------------------------------------------------
// integrator: CVelVerlet::propagator NVE/NVT branch (VelVerlet.cpp:74-93),
// cutForces (:96-122), pbcWrap (SimulationBox.cpp:135-166)
// ---------------------------------------------------------------------------

__device__ inline void block_sum_store(double v, double* out) { mt_block_sum_store<TPB>(v, out); }

// p *= S (thermostat), p += F dt/2, r += p dt/m; optional max |r - r_ref|^2
__global__ void __launch_bounds__(TPB) k_kick_drift(double4* __restrict__ pos, double4* __restrict__ mom,
                                                    const double4* __restrict__ frc, int n, const double* nh,
                                                    int scale_slot, double dt, const double4* __restrict__ posref,
                                                    unsigned int* flags)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    float d2 = 0.0f;
    if (k < n) {
        double S = (scale_slot >= 0) ? nh[scale_slot] : 1.0;
        double4 p = mom[k], r = pos[k], F = frc[k];
        double dt_2 = dt / 2.0;
        if (scale_slot >= 0) { p.x *= S; p.y *= S; p.z *= S; }
        p.x = p.x + F.x * dt_2; p.y = p.y + F.y * dt_2; p.z = p.z + F.z * dt_2;
        double c = dt / p.w;
        r.x = r.x + p.x * c; r.y = r.y + p.y * c; r.z = r.z + p.z * c;
        mom[k] = p;
        pos[k] = r;
        if (posref) {
            double4 q = posref[k];
            double dx = r.x - q.x, dy = r.y - q.y, dz = r.z - q.z;
            d2 = (float)(dx * dx + dy * dy + dz * dz);
        }
    }
    if (posref) {
        // wave-level max, one atomic per wavefront
        for (int off = 32; off > 0; off >>= 1) d2 = fmaxf(d2, __shfl_xor(d2, off));
        if ((threadIdx.x & 63) == 0 && d2 > 0.0f) atomicMax(&flags[0], __float_as_uint(d2));
    }
}

void mt_kick_drift(const MtDev& d, const MtParams& p, int scale_slot, int track_disp, hipStream_t s)
{
    hipLaunchKernelGGL(k_kick_drift, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.pos, d.mom, d.frc, d.n, d.nh,
                       scale_slot, p.dt, track_disp ? d.posref : (const double4*)nullptr, d.flags);
}

// cap F per component (cutForces), p += F dt/2, optional block partials of p.p/m
__global__ void __launch_bounds__(TPB) k_cap_kick(double4* __restrict__ mom, double4* __restrict__ frc, int n,
                                                  double Fcut, double dt, double* red, int want_ke)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0;
    if (k < n) {
        double4 F = frc[k], pm = mom[k];
        if (Fcut > 0) {
            if (fabs(F.x) > Fcut) F.x = ((F.x >= 0.0) ? 1.0 : -1.0) * Fcut;
            if (fabs(F.y) > Fcut) F.y = ((F.y >= 0.0) ? 1.0 : -1.0) * Fcut;
            if (fabs(F.z) > Fcut) F.z = ((F.z >= 0.0) ? 1.0 : -1.0) * Fcut;
            frc[k] = F;
        }
        double dt_2 = dt / 2.0;
        pm.x = pm.x + F.x * dt_2; pm.y = pm.y + F.y * dt_2; pm.z = pm.z + F.z * dt_2;
        mom[k] = pm;
        ke = ((pm.x * pm.x) + (pm.y * pm.y) + (pm.z * pm.z)) / pm.w;
    }
    if (want_ke) block_sum_store(ke, &red[blockIdx.x]);
}

void mt_cap_kick(const MtDev& d, const MtParams& p, int want_ke, hipStream_t s)
{
    hipLaunchKernelGGL(k_cap_kick, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.mom, d.frc, d.n, p.Fcut, p.dt, d.red,
                       want_ke);
}

// Nose-Hoover chain propagator (NHChain.cpp:54-113) with CFctT (MDLoop.cpp:29-60),
// mt_nh_propagate in mtmd_device.h; one chain, nh layout: eta[M], p_eta[M],
// Q[M], scale. Sums the block partials of p.p/m first.
// fast (FP64/MIXED, trace agreement): the chain state staged in LDS while the
// partials load, the partials summed by a butterfly of shuffles; with
// from_slot the kinetic energy is nh[3M + 1] instead of the partials. The
// kinetic energy after this scale, KE S^2, goes to nh[3M + 1] for the next
// propagation.
__global__ void k_nh_chain(double* nh, const double* red, int nblocks, int M, int nresp, double T, double dt,
                           double g, double w0, double w1, double w2, int fast, int from_slot)
{
    __shared__ double sh[64];
    __shared__ double snh[3 * 8];
    const double w[3] = {w0, w1, w2};
    double s = 0.0;
    if (fast) {
        const bool staged = M <= 8;
        if (staged)
            for (int i = threadIdx.x; i < 3 * M; i += 64) snh[i] = nh[i];
        if (!from_slot) {
            for (int b = threadIdx.x; b < nblocks; b += 64) s += red[b];
            for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        }
        __syncthreads();
        if (threadIdx.x != 0) return;
        if (from_slot) s = nh[3 * M + 1];
        double* c = staged ? snh : nh;
        const double S = mt_nh_step(c, c + M, c + 2 * M, M, nresp, T, dt, g, w, s, true);
        if (staged)
            for (int i = 0; i < 2 * M; i++) nh[i] = snh[i];
        nh[3 * M] = S;
        nh[3 * M + 1] = s * S * S;
        return;
    }
    for (int b = threadIdx.x; b < nblocks; b += 64) s += red[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double KE2 = 0.0;
    for (int t = 0; t < 64; t++) KE2 += sh[t];
    nh[3 * M] = mt_nh_step(nh, nh + M, nh + 2 * M, M, nresp, T, dt, g, w, KE2, fast != 0);
}

// FP64/MIXED, between two steps of one call: the chain at the end of step t
// on the summed p.p/m after the kick (its scale Sb into nh[3M]) and, at once,
// the chain at the start of step t + 1 on that sum times Sb^2, the kinetic
// energy after Sb is applied (its scale Sf into nh[3M + 1]); the next front
// half applies Sb and Sf. The second propagation's kinetic energy is formed
// analytically rather than summed again, as k_nh_chain's from_slot takes it
// when a call ends and the next begins, so the steps of one call and the same
// steps in separate calls give the same bits.
__global__ void k_nh_chain2(double* nh, const double* red, int nblocks, int M, int nresp, double T, double dt,
                            double g, double w0, double w1, double w2)
{
    __shared__ double snh[3 * 8];
    const double w[3] = {w0, w1, w2};
    const bool staged = M <= 8;
    if (staged)
        for (int i = threadIdx.x; i < 3 * M; i += 64) snh[i] = nh[i];
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 64) s += red[b];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    __syncthreads();
    if (threadIdx.x != 0) return;
    double* c = staged ? snh : nh;
    const double Sb = mt_nh_step(c, c + M, c + 2 * M, M, nresp, T, dt, g, w, s, true);
    const double ke = s * Sb * Sb;
    const double Sf = mt_nh_step(c, c + M, c + 2 * M, M, nresp, T, dt, g, w, ke, true);
    if (staged)
        for (int i = 0; i < 2 * M; i++) nh[i] = snh[i];
    nh[3 * M] = Sb;
    nh[3 * M + 1] = Sf;
}

void mt_nh_chain2(const MtDev& d, int M, int nresp, double T, double dt, double g, const double* w, int nblocks,
                  hipStream_t s)
{
    hipLaunchKernelGGL(k_nh_chain2, dim3(1), dim3(64), 0, s, d.nh, d.red, nblocks, M, nresp, T, dt, g, w[0], w[1],
                       w[2]);
}

void mt_nh_chain(const MtDev& d, int M, int nresp, double T, double dt, double g, const double* w, int nblocks,
                 hipStream_t s, int fast, int from_slot)
{
    hipLaunchKernelGGL(k_nh_chain, dim3(1), dim3(64), 0, s, d.nh, d.red, nblocks, M, nresp, T, dt, g, w[0], w[1],
                       w[2], fast, from_slot);
}

// p *= S; pbcWrap (SimulationBox.cpp:135-166); optional partials of p.p/m

__global__ void __launch_bounds__(TPB) k_scale_wrap(double4* __restrict__ pos, double4* __restrict__ mom, int n,
                                                    const double* nh, int scale_slot, double LX, double LY,
                                                    double LZ, double* red, int want_ke)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0;
    if (k < n) {
        double4 r = pos[k];
        double X2 = LX / 2.0, Y2 = LY / 2.0, Z2 = LZ / 2.0;
        if (r.x >= X2) mt_pbc_add(r.x, X2, LX);
        if (r.x < -X2) mt_pbc_add(r.x, -X2, LX);
        if (r.y >= Y2) mt_pbc_add(r.y, Y2, LY);
        if (r.y < -Y2) mt_pbc_add(r.y, -Y2, LY);
        if (r.z >= Z2) mt_pbc_add(r.z, Z2, LZ);
        if (r.z < -Z2) mt_pbc_add(r.z, -Z2, LZ);
        pos[k] = r;
        double4 pm = mom[k];
        if (scale_slot >= 0) {
            double S = nh[scale_slot];
            pm.x *= S; pm.y *= S; pm.z *= S;
            mom[k] = pm;
        }
        ke = ((pm.x * pm.x) + (pm.y * pm.y) + (pm.z * pm.z)) / pm.w;
    }
    if (want_ke) block_sum_store(ke, &red[blockIdx.x]);
}

void mt_scale_wrap(const MtDev& d, const MtParams& p, int scale_slot, int want_ke, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_wrap, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.pos, d.mom, d.n, d.nh, scale_slot,
                       p.L[0], p.L[1], p.L[2], d.red, want_ke);
}

__global__ void __launch_bounds__(TPB) k_reduce_ke(const double4* __restrict__ mom, int n, double* red)
{
    double ke = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        double4 pm = mom[k];
        ke += ((pm.x * pm.x) + (pm.y * pm.y) + (pm.z * pm.z)) / pm.w;
    }
    block_sum_store(ke, &red[blockIdx.x]);
}

void mt_reduce_ke(const MtDev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_ke, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.mom, d.n, d.red);
}

// partials of sum p.p/m, sum F.r, sum U (SimulationBox.cpp:168-231) in red[3*b..3*b+2]
__global__ void __launch_bounds__(TPB) k_reduce_thermo(const double4* __restrict__ pos,
                                                       const double4* __restrict__ mom,
                                                       const double4* __restrict__ frc,
                                                       const float* __restrict__ epot, int n, double* red)
{
    double a = 0.0, b = 0.0, c = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        double4 pm = mom[k], r = pos[k], F = frc[k];
        a += ((pm.x * pm.x) + (pm.y * pm.y) + (pm.z * pm.z)) / pm.w;
        b += (F.x * r.x) + (F.y * r.y) + (F.z * r.z);
        c += (double)epot[k];
    }
    block_sum_store(a, &red[3 * blockIdx.x]);
    __syncthreads();
    block_sum_store(b, &red[3 * blockIdx.x + 1]);
    __syncthreads();
    block_sum_store(c, &red[3 * blockIdx.x + 2]);
}

void mt_reduce_thermo_to(const MtDev& d, double* red, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_thermo, dim3(mt_reduce_blocks(d.n)), dim3(TPB), 0, s, d.pos, d.mom, d.frc, d.epot,
                       d.n, red);
}

void mt_reduce_thermo(const MtDev& d, hipStream_t s) { mt_reduce_thermo_to(d, d.red, s); }

// CSimulationBox::gradientDescentStep (SimulationBox.cpp:197-208): r += F*lr
__global__ void k_gradient_step(double4* __restrict__ pos, const double4* __restrict__ frc, int n, double lr)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 r = pos[k], F = frc[k];
    r.x = r.x + F.x * lr; r.y = r.y + F.y * lr; r.z = r.z + F.z * lr;
    pos[k] = r;
}

void mt_gradient_step(const MtDev& d, double lr, hipStream_t s)
{
    hipLaunchKernelGGL(k_gradient_step, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.pos, d.frc, d.n, lr);
}

// CMDLoop::appendToMomentumDistribution (MDLoop.cpp:304-322), one step: each
// momentum component of every atom goes to bin int(((p + maxP)/(2 maxP))*nbins)
// of its axis (hist[axis*nbins + bin], counted when the bin is in range), and
// the first thermostat momentum, when peta is given, to hist[3*nbins + bin]
// with its own maximum. Counts gather in LDS per workgroup, then one atomic
// per non-empty bin goes to global memory.
#define MT_DISTR_MAXBINS 256
__global__ void __launch_bounds__(TPB) k_pdistr(const double4* __restrict__ mom, int n, double maxP, int nbins,
                                                const double* __restrict__ peta, double maxPeta,
                                                int* __restrict__ hist)
{
    __shared__ int sh[3 * MT_DISTR_MAXBINS];
    for (int i = threadIdx.x; i < 3 * nbins; i += blockDim.x) sh[i] = 0;
    __syncthreads();
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const double4 pm = mom[k];
        const double c[3] = {pm.x, pm.y, pm.z};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int b = mt_d2i(((c[a] + maxP) / (2.0 * maxP)) * (double)nbins);
            if (b >= 0 && b < nbins) atomicAdd(&sh[a * nbins + b], 1);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * nbins; i += blockDim.x)
        if (sh[i]) atomicAdd(&hist[i], sh[i]);
    if (peta && blockIdx.x == 0 && threadIdx.x == 0) {
        const int b = mt_d2i(((peta[0] + maxPeta) / (2.0 * maxPeta)) * (double)nbins);
        if (b >= 0 && b < nbins) atomicAdd(&hist[3 * nbins + b], 1);
    }
}

int mt_pdistr(const MtDev& d, double maxP, int nbins, const double* peta, double maxPeta, int* hist, hipStream_t s)
{
    if (nbins < 1 || nbins > MT_DISTR_MAXBINS) return -1;
    int nb = nblk(d.n, TPB);
    nb = nb < 1 ? 1 : (nb > 1024 ? 1024 : nb);
    hipLaunchKernelGGL(k_pdistr, dim3(nb), dim3(TPB), 0, s, d.mom, d.n, maxP, nbins, peta, maxPeta, hist);
    return 0;
}

// ---------------------------------------------------------------------------
// NPT: the particle half of CVelVerlet::propagator's MTK branch
// (VelVerlet.cpp:57-71,124-174); the barostat scalars (p_eps, eps, box) are
// advanced on the host between these kernels
// ---------------------------------------------------------------------------
// the NH_T scale applied to the momenta, then block partials of sum p.p/m
// (red[b]) and sum Fpi.r (red[gridDim.x + b]) for G_eps (VelVerlet.cpp:160-174)
__global__ void __launch_bounds__(TPB) k_npt_scale_sums(double4* __restrict__ mom, const double4* __restrict__ pos,
                                                        const double4* __restrict__ frc, int n, const double* nh,
                                                        int scale_slot, double* red)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (k < n) {
        double4 pm = mom[k];
        if (scale_slot >= 0) {
            const double S = nh[scale_slot];
            pm.x *= S; pm.y *= S; pm.z *= S;
            mom[k] = pm;
        }
        const double4 r = pos[k], F = frc[k];
        a = ((pm.x * pm.x) + (pm.y * pm.y) + (pm.z * pm.z)) / pm.w;
        b = (F.x * r.x) + (F.y * r.y) + (F.z * r.z);
    }
    block_sum_store(a, &red[blockIdx.x]);
    __syncthreads();
    block_sum_store(b, &red[gridDim.x + blockIdx.x]);
}

// prop_p then prop_r (VelVerlet.cpp:124-150): p = p c2 + F cf, r = r r2 + (p/m) rr
__global__ void __launch_bounds__(TPB) k_npt_prop_pr(double4* __restrict__ pos, double4* __restrict__ mom,
                                                     const double4* __restrict__ frc, int n, double cp2, double cpf,
                                                     double cr2, double crr)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 pm = mom[k], r = pos[k];
    const double4 F = frc[k];
    pm.x = pm.x * cp2 + F.x * cpf;
    pm.y = pm.y * cp2 + F.y * cpf;
    pm.z = pm.z * cp2 + F.z * cpf;
    const double im = 1.0 / pm.w;
    const double ux = pm.x * im, uy = pm.y * im, uz = pm.z * im;
    r.x = r.x * cr2 + ux * crr;
    r.y = r.y * cr2 + uy * crr;
    r.z = r.z * cr2 + uz * crr;
    mom[k] = pm;
    pos[k] = r;
}

// prop_p after the force call, then the partials of G_eps as k_npt_scale_sums
__global__ void __launch_bounds__(TPB) k_npt_prop_p_sums(double4* __restrict__ mom, const double4* __restrict__ pos,
                                                         const double4* __restrict__ frc, int n, double cp2,
                                                         double cpf, double* red)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (k < n) {
        double4 pm = mom[k];
        const double4 r = pos[k], F = frc[k];
        pm.x = pm.x * cp2 + F.x * cpf;
        pm.y = pm.y * cp2 + F.y * cpf;
        pm.z = pm.z * cp2 + F.z * cpf;
        mom[k] = pm;
        a = ((pm.x * pm.x) + (pm.y * pm.y) + (pm.z * pm.z)) / pm.w;
        b = (F.x * r.x) + (F.y * r.y) + (F.z * r.z);
    }
    block_sum_store(a, &red[blockIdx.x]);
    __syncthreads();
    block_sum_store(b, &red[gridDim.x + blockIdx.x]);
}

int mt_npt_blocks(int n) { return nblk(n, TPB); }

void mt_npt_scale_sums(const MtDev& d, int scale_slot, hipStream_t s)
{
    hipLaunchKernelGGL(k_npt_scale_sums, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.mom, d.pos, d.frc, d.n, d.nh,
                       scale_slot, d.red);
}

void mt_npt_prop_pr(const MtDev& d, double cp2, double cpf, double cr2, double crr, hipStream_t s)
{
    hipLaunchKernelGGL(k_npt_prop_pr, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.pos, d.mom, d.frc, d.n, cp2, cpf,
                       cr2, crr);
}

void mt_npt_prop_p_sums(const MtDev& d, double cp2, double cpf, hipStream_t s)
{
    hipLaunchKernelGGL(k_npt_prop_p_sums, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.mom, d.pos, d.frc, d.n, cp2, cpf,
                       d.red);
}
