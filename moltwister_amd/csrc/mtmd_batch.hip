// mtmd_batch.hip - many independent systems (replicas) advanced together.
//
// A batch is n independent `moldyn` runs (each a CSimulationBox + CMDLoop of
// the reference) that share the integration settings (time step, cutoff,
// skin, ensemble, chain length) and differ in size, box, temperature and
// state. Atoms of all replicas are concatenated; every kernel of a step
// covers every replica in one launch (per-replica data is looked up through
// the atom's replica id), and the whole step is captured once in a hipGraph
// and replayed. For comparison the same kernels can be launched replica by
// replica (mtmd_batch_step_separate), the way n separate engines would run.
//
// Precision is that of the MIXED engine: FP32 pair geometry and table
// interpolation, FP64 force accumulation, integration, Nose-Hoover chains and
// reductions. Pair vectors are minimum images in FP32 of FP32 positions
// (replica boxes are small, |x| < 20 sigma for the target workload), so small
// boxes with fewer than three cells per axis are handled without the
// reference's duplicate-neighbour behaviour: each distinct cell is visited
// once. Neighbour lists are skin lists rebuilt per replica on displacement.
//
// Supported per replica: tabulated non-bonded force fields (any kind, or
// profiles), no charges, no bonded terms, all atoms mobile.
#include "mtmd_internal.h"
#include "mtmd_device.h"
#include "../../include/mtmd.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#define MB_TPB 256
#define MB_MAXM 8

static inline int mb_blk(int n, int b) { return (n + b - 1) / b; }
static inline float mb_bits_as_float(int v) { float f; std::memcpy(&f, &v, sizeof f); return f; }

// per-replica constants (device array)
struct MbRep {
    double L[3];            // box widths (centred on the origin)
    float Lf[3], iLf[3];    // FP32 widths and their inverses (minimum image)
    int nc[3];              // cell grid
    int cell0, ncell;       // first global cell, number of cells
    int sk;                 // persistent kernel: cells in y and z per list radius (1 or 2), the
                            // stencil's half-width on those axes
    int atom0, n;           // atom range
    double T;               // target temperature (reduced units)
    double g;               // 3 n degrees of freedom
};

struct MbDev {
    int n, nrep, ncells, T, M;
    double4* pos; double4* mom; double4* frc; double4* posref;
    float4* posf; float* epot; int* rid;
    int* cell_key; int* cell_count; int* cell_first; int* cell_fill; int* cell_list; int* cell_rid;
    int ncap; int* ncount; uint32_t* nlist; int* nreq;
    float2* nbcoef2; float4* nbcoef;
    MbRep* rep;
    double* nh;             // per replica: eta[M], p_eta[M], Q[M], S, KE2  (stride 3M + 2)
    double* red;            // per replica: sum p.p/m, sum F.r, sum U
    unsigned int* rflag;    // per replica: rebuild the lists this step
    unsigned int* flags;    // [0] list overflow, [1] non-finite position
};

struct MbPar {
    double dt, Fcut;
    float rc2, rlist2, x0, inv_delta, half_skin2;
    int nresp;
    double w[3];
    int nvt;
};

// ---------------------------------------------------------------------------
// kernels; ranges [a0, a1) atoms, [c0, c1) cells, [r0, r1) replicas
// ---------------------------------------------------------------------------

// per-replica sums (one workgroup per replica): p.p/m, F.r, U
__global__ void __launch_bounds__(MB_TPB) k_mb_reduce(MbDev d, int r0, int want_all)
{
    __shared__ double sh[3][MB_TPB];
    const int r = r0 + blockIdx.x;
    const MbRep rp = d.rep[r];
    double a = 0.0, b = 0.0, c = 0.0;
    for (int k = rp.atom0 + threadIdx.x; k < rp.atom0 + rp.n; k += MB_TPB) {
        const double4 pm = d.mom[k];
        a += ((pm.x * pm.x) + (pm.y * pm.y) + (pm.z * pm.z)) / pm.w;
        if (want_all) {
            const double4 F = d.frc[k], x = d.pos[k];
            b += (F.x * x.x) + (F.y * x.y) + (F.z * x.z);
            c += (double)d.epot[k];
        }
    }
    sh[0][threadIdx.x] = a; sh[1][threadIdx.x] = b; sh[2][threadIdx.x] = c;
    __syncthreads();
    for (int off = MB_TPB / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int q = 0; q < 3; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        d.red[3 * r] = sh[0][0];
        if (want_all) { d.red[3 * r + 1] = sh[1][0]; d.red[3 * r + 2] = sh[2][0]; }
    }
}

// one Nose-Hoover chain propagation per replica (one thread per replica)
__global__ void k_mb_nh(MbDev d, MbPar p, int r0, int r1)
{
    const int r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= r1) return;
    const int M = d.M, stride = 3 * M + 2;
    double* nh = d.nh + (size_t)r * stride;
    const MbRep rp = d.rep[r];
    nh[3 * M] = mt_nh_step(nh, nh + M, nh + 2 * M, M, p.nresp, rp.T, p.dt, rp.g, p.w, d.red[3 * r]);
}

// thermostat scale, half kick, drift, FP32 copy, per-replica displacement flag
__global__ void __launch_bounds__(MB_TPB) k_mb_kick_drift(MbDev d, MbPar p, int a0, int a1)
{
    const int k = a0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a1) return;
    const int r = d.rid[k];
    double4 pm = d.mom[k], x = d.pos[k];
    const double4 F = d.frc[k];
    if (p.nvt) {
        const double S = d.nh[(size_t)r * (3 * d.M + 2) + 3 * d.M];
        pm.x *= S; pm.y *= S; pm.z *= S;
    }
    const double dt_2 = p.dt / 2.0;
    pm.x += F.x * dt_2; pm.y += F.y * dt_2; pm.z += F.z * dt_2;
    const double c = p.dt / pm.w;
    x.x += pm.x * c; x.y += pm.y * c; x.z += pm.z * c;
    d.mom[k] = pm;
    d.pos[k] = x;
    const float4 old = d.posf[k];
    d.posf[k] = make_float4((float)x.x, (float)x.y, (float)x.z, old.w);
    const double4 q = d.posref[k];
    const MbRep& rp = d.rep[r];
    double dx = x.x - q.x, dy = x.y - q.y, dz = x.z - q.z;
    dx -= rp.L[0] * rint(dx / rp.L[0]);
    dy -= rp.L[1] * rint(dy / rp.L[1]);
    dz -= rp.L[2] * rint(dz / rp.L[2]);
    if ((float)(dx * dx + dy * dy + dz * dz) > p.half_skin2 && d.rflag[r] == 0u) d.rflag[r] = 1u;
}

// ---- conditional list rebuild (replicas with rflag set) ----

__global__ void __launch_bounds__(MB_TPB) k_mb_cells_clear(MbDev d, int c0, int c1)
{
    const int c = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= c1 || !d.rflag[d.cell_rid[c]]) return;
    d.cell_count[c] = 0;
    d.cell_fill[c] = 0;
}

__device__ __forceinline__ void mb_key_runs(int key, int& head_lane, int& len)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1);
    const unsigned long long heads = __ballot(lane == 0 || prev != key);
    head_lane = 63 - __clzll(heads & ((2ull << lane) - 1ull));
    const unsigned long long above = heads & ~((2ull << head_lane) - 1ull);
    len = (above ? __ffsll(above) - 1 : 64) - head_lane;
}

__global__ void __launch_bounds__(MB_TPB) k_mb_cell_keys(MbDev d, int a0, int a1)
{
    const int k = a0 + blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    if (k < a1) {
        const int r = d.rid[k];
        if (d.rflag[r]) {
            const MbRep& rp = d.rep[r];
            const double4 x = d.pos[k];
            if (!(isfinite(x.x) && isfinite(x.y) && isfinite(x.z))) atomicOr(&d.flags[1], 1u);
            int i3[3];
            const double xs[3] = {x.x, x.y, x.z};
            for (int a = 0; a < 3; a++) {
                // position folded into the box first: between wraps an atom can sit just outside
                double u = xs[a] - rp.L[a] * floor(xs[a] / rp.L[a] + 0.5);
                int ic = (int)floor((u + rp.L[a] / 2.0) * (double)rp.nc[a] / rp.L[a]);
                i3[a] = ic < 0 ? 0 : (ic >= rp.nc[a] ? rp.nc[a] - 1 : ic);
            }
            key = rp.cell0 + i3[0] + rp.nc[0] * (i3[1] + rp.nc[1] * i3[2]);
            d.cell_key[k] = key;
        }
    }
    int hl, len;
    mb_key_runs(key, hl, len);
    if (key >= 0 && (int)(threadIdx.x & 63) == hl) atomicAdd(&d.cell_count[key], len);
}

// exclusive scan of one replica's cell counts per workgroup; the cells of
// replica r hold exactly its atoms, so its first slot is atom0
__global__ void __launch_bounds__(MB_TPB) k_mb_scan(MbDev d, int r0)
{
    __shared__ int ws[MB_TPB / 64];
    __shared__ int carry;
    const int r = r0 + blockIdx.x;
    if (!d.rflag[r]) return;
    const MbRep rp = d.rep[r];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) carry = rp.atom0;
    __syncthreads();
    for (int b = 0; b < rp.ncell; b += MB_TPB) {
        const int c = b + t;
        const int v = c < rp.ncell ? d.cell_count[rp.cell0 + c] : 0;
        int incl = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o);
            if (lane >= o) incl += u;
        }
        if (lane == 63) ws[w] = incl;
        __syncthreads();
        int wpre = 0;
        for (int j = 0; j < w; j++) wpre += ws[j];
        if (c < rp.ncell) d.cell_first[rp.cell0 + c] = carry + wpre + incl - v;
        __syncthreads();
        if (t == MB_TPB - 1) carry += wpre + incl;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(MB_TPB) k_mb_scatter(MbDev d, int a0, int a1)
{
    const int k = a0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int key = (k < a1 && d.rflag[d.rid[k]]) ? d.cell_key[k] : -1;
    int hl, len;
    mb_key_runs(key, hl, len);
    int base = 0;
    if (key >= 0 && lane == hl) base = atomicAdd(&d.cell_fill[key], len);
    base = __shfl(base, hl);
    if (key >= 0) d.cell_list[d.cell_first[key] + base + (lane - hl)] = k;
}

// ascending atom order inside each cell (reproducible lists): one wavefront per cell
__global__ void __launch_bounds__(MB_TPB) k_mb_cell_sort(MbDev d, int c0, int c1)
{
    __shared__ __align__(16) int lds[(MB_TPB / 64) * MT_CELL_SORT_LDS];
    if (d.flags[1]) return;                                      // uniform over the grid
    const int c = c0 + (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const bool live = c < c1 && d.rflag[d.cell_rid[c]];
    mt_wave_sort_cell(d.cell_list + (live ? d.cell_first[c] : 0), live ? d.cell_count[c] : 0,
                      lds + (threadIdx.x >> 6) * MT_CELL_SORT_LDS);
}

__device__ __forceinline__ float mb_min_image(float d, float L, float iL) { return fmaf(-L, rintf(d * iL), d); }

// skin list of one atom: every distinct cell within one cell of its own (per
// axis: three offsets when the replica has three or more cells, else every
// cell once), FP32 minimum-image distances; chunked column-major output
// (word j of atom k at nlist[4 * ((j / 4) * n + k) + j % 4], unused words of
// the last chunk hold k itself, which gives r = 0 and no force)
__global__ void __launch_bounds__(MB_TPB) k_mb_neigh(MbDev d, MbPar p, int a0, int a1)
{
    const int k = a0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a1) return;
    const int r = d.rid[k];
    if (!d.rflag[r]) return;
    if (d.flags[1]) { d.ncount[k] = 0; return; }
    const MbRep& rp = d.rep[r];
    const float4 pk = d.posf[k];
    const int key = d.cell_key[k] - rp.cell0;
    const int hx = key % rp.nc[0], hy = (key / rp.nc[0]) % rp.nc[1], hz = key / (rp.nc[0] * rp.nc[1]);
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        if (rp.nc[a] >= 3) { lo[a] = -1; hi[a] = 1; }
        else { lo[a] = 0; hi[a] = rp.nc[a] - 1; }
    }
    const float Lx = rp.Lf[0], Ly = rp.Lf[1], Lz = rp.Lf[2];
    const float iLx = rp.iLf[0], iLy = rp.iLf[1], iLz = rp.iLf[2];
    const int n = d.n, ncap = d.ncap;
    uint4* __restrict__ nl4 = reinterpret_cast<uint4*>(d.nlist);
    int cnt = 0;
    uint32_t b0 = 0, b1 = 0, b2 = 0;
    for (int oz = lo[2]; oz <= hi[2]; oz++) {
        int cz = hz + oz; cz = cz < 0 ? cz + rp.nc[2] : (cz >= rp.nc[2] ? cz - rp.nc[2] : cz);
        for (int oy = lo[1]; oy <= hi[1]; oy++) {
            int cy = hy + oy; cy = cy < 0 ? cy + rp.nc[1] : (cy >= rp.nc[1] ? cy - rp.nc[1] : cy);
            for (int ox = lo[0]; ox <= hi[0]; ox++) {
                int cx = hx + ox; cx = cx < 0 ? cx + rp.nc[0] : (cx >= rp.nc[0] ? cx - rp.nc[0] : cx);
                const int cc = rp.cell0 + cx + rp.nc[0] * (cy + rp.nc[1] * cz);
                const int f = d.cell_first[cc], m = d.cell_count[cc];
                for (int e = 0; e < m; e++) {
                    const int a = d.cell_list[f + e];
                    const float4 pa = d.posf[a];
                    const float dx = mb_min_image(pk.x - pa.x, Lx, iLx);
                    const float dy = mb_min_image(pk.y - pa.y, Ly, iLy);
                    const float dz = mb_min_image(pk.z - pa.z, Lz, iLz);
                    const bool hit = a != k && dx * dx + dy * dy + dz * dz < p.rlist2;
                    if (hit) {
                        const uint32_t w = (uint32_t)a;
                        const int slot = cnt & 3;
                        b0 = slot == 0 ? w : b0;
                        b1 = slot == 1 ? w : b1;
                        b2 = slot == 2 ? w : b2;
                        if (slot == 3 && cnt < ncap) nl4[(size_t)(cnt >> 2) * n + k] = make_uint4(b0, b1, b2, w);
                        cnt++;
                    }
                }
            }
        }
    }
    const int slot = cnt & 3;
    if (slot && cnt < ncap) {
        const uint32_t pad = (uint32_t)k;
        nl4[(size_t)(cnt >> 2) * n + k] = make_uint4(b0, slot > 1 ? b1 : pad, slot > 2 ? b2 : pad, pad);
    }
    d.ncount[k] = cnt;
    d.posref[k] = d.pos[k];
    if (cnt > ncap) { atomicOr(&d.flags[0], 1u); atomicMax(d.nreq, cnt); }
}

__global__ void k_mb_clear_rflags(MbDev d, int r0, int r1)
{
    const int r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (r < r1) d.rflag[r] = 0u;
}

// pair forces over the chunked list: FP32 minimum image and table
// interpolation, FP64 accumulation; four gathers in flight per chunk
template <bool WANT_E>
__global__ void __launch_bounds__(512) k_mb_force(MbDev d, MbPar p, int a0, int a1)
{
    extern __shared__ float2 stab[];
    const int T = d.T;
    if (!WANT_E) {
        for (int i = threadIdx.x; i < T * T * MT_NPTS; i += blockDim.x) stab[i] = d.nbcoef2[i];
        __syncthreads();
    }
    const int k = a0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a1) return;
    const MbRep& rp = d.rep[d.rid[k]];
    const float Lx = rp.Lf[0], Ly = rp.Lf[1], Lz = rp.Lf[2];
    const float iLx = rp.iLf[0], iLy = rp.iLf[1], iLz = rp.iLf[2];
    const float4 pk = d.posf[k];
    const int rowoff = __float_as_int(pk.w) * T * MT_NPTS;
    double Fx = 0.0, Fy = 0.0, Fz = 0.0;
    float U = 0.0f;
    const int nch = (min(d.ncount[k], d.ncap) + 3) >> 2;
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4* __restrict__ nl4 = reinterpret_cast<const u32x4*>(d.nlist) + k;
    const size_t stride = (size_t)d.n;
    for (int q = 0; q < nch; q++) {
        const u32x4 w = __builtin_nontemporal_load(nl4 + (size_t)q * stride);
        float4 a4[4] = {d.posf[w.x], d.posf[w.y], d.posf[w.z], d.posf[w.w]};
        asm volatile("" : "+v"(a4[0].w), "+v"(a4[1].w), "+v"(a4[2].w), "+v"(a4[3].w));
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float4 pa = a4[j];
            const float dx = mb_min_image(pk.x - pa.x, Lx, iLx);
            const float dy = mb_min_image(pk.y - pa.y, Ly, iLy);
            const float dz = mb_min_image(pk.z - pa.z, Lz, iLz);
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < p.rc2 && r2 >= 1.0e-6f) {
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * rinv;
                const int idx = (int)((r - p.x0) * p.inv_delta);
                if (idx >= 0 && idx < MT_NPTS - 1) {
                    const int ci = rowoff + __float_as_int(pa.w) * MT_NPTS + idx;
                    float fr;
                    if (WANT_E) {
                        const float4 c = d.nbcoef[ci];
                        fr = c.x + c.y * rinv;
                        U += 0.5f * (c.z * r + c.w);
                    } else {
                        const float2 c = stab[ci];
                        fr = c.x + c.y * rinv;
                    }
                    Fx += (double)(dx * fr);
                    Fy += (double)(dy * fr);
                    Fz += (double)(dz * fr);
                }
            }
        }
    }
    d.frc[k] = make_double4(Fx, Fy, Fz, 0.0);
    if (WANT_E) d.epot[k] = U;
}

// force cap (cutForces) and half kick
__global__ void __launch_bounds__(MB_TPB) k_mb_cap_kick(MbDev d, MbPar p, int a0, int a1)
{
    const int k = a0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a1) return;
    double4 F = d.frc[k], pm = d.mom[k];
    if (p.Fcut > 0) {
        if (fabs(F.x) > p.Fcut) F.x = ((F.x >= 0.0) ? 1.0 : -1.0) * p.Fcut;
        if (fabs(F.y) > p.Fcut) F.y = ((F.y >= 0.0) ? 1.0 : -1.0) * p.Fcut;
        if (fabs(F.z) > p.Fcut) F.z = ((F.z >= 0.0) ? 1.0 : -1.0) * p.Fcut;
        d.frc[k] = F;
    }
    const double dt_2 = p.dt / 2.0;
    pm.x += F.x * dt_2; pm.y += F.y * dt_2; pm.z += F.z * dt_2;
    d.mom[k] = pm;
}

// thermostat scale and PBC wrap into the replica's box
__global__ void __launch_bounds__(MB_TPB) k_mb_scale_wrap(MbDev d, MbPar p, int a0, int a1)
{
    const int k = a0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a1) return;
    const int r = d.rid[k];
    const MbRep& rp = d.rep[r];
    double4 x = d.pos[k];
    double* xs[3] = {&x.x, &x.y, &x.z};
    for (int a = 0; a < 3; a++) {
        const double h = rp.L[a] / 2.0;
        if (*xs[a] >= h || *xs[a] < -h) *xs[a] -= rp.L[a] * floor(*xs[a] / rp.L[a] + 0.5);
    }
    d.pos[k] = x;
    if (p.nvt) {
        const double S = d.nh[(size_t)r * (3 * d.M + 2) + 3 * d.M];
        double4 pm = d.mom[k];
        pm.x *= S; pm.y *= S; pm.z *= S;
        d.mom[k] = pm;
    }
}

__global__ void k_mb_fill_u32(unsigned int* a, unsigned int v, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------
// The persistent replica kernel: one workgroup per replica advances it through
// whole steps, the replica's FP32 positions (in cell order) and the pair table
// resident in LDS. A step of replica r (MDLoop.cpp:124-138): the chain and
// scale at its start, half kick and drift (FP64, global state in atom order
// through the replica's slot -> atom map), the displacement test over the
// replica, when due the cell sort in LDS and the 16-bit skin lists (words are
// LDS places), the pair forces from LDS with the back half kick, wrap and
// force cap in the lane that reduced the force, then the chain and scale at
// the step's end. Nothing is shared between replicas, so the workgroups never
// wait for each other; one launch runs every step of a call.
// ---------------------------------------------------------------------------
#define MB_RUN_MAXN 4096      // atoms a replica may hold (64 KB of staged float4 positions)
#define MB_RUN_MAXCELLS 1024  // cells of a replica's grid
#define MB_RUN_TPB 1024       // largest workgroup
#define MB_RUN_CHUNK 256      // steps a launch of a call runs at most (the unit a list growth runs again)

struct MbRun {
    const int* order;         // replica of workgroup b: large replicas first, then small ones, so the two
                              // workgroups a CU holds together carry a large and a small replica
    int* perm;                // [n] slot -> atom of each replica segment (slots in cell order)
    int* perm2;               // [n] the new order of a rebuild
    uint32_t* nl16;           // 16-bit rows: chunk q (8 words) of slot a0 + s at uint4 q * n + a0 + s
    int* ncount16;            // [n] words in each slot's row
    int ncap16;               // words a row holds (a multiple of 8)
    int nlimit;               // longest row accepted without a regrow (ncap16, or less with no slack)
    int maxn, maxcells;       // staging capacities of the launch (largest replica)
};

__device__ __forceinline__ double mb_block_sum(double v, double* sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nw; i++) t += sh[i];      // fixed order: the same bits in every launch
    return t;
}

__device__ __forceinline__ float mb_block_max(float v, float* sh)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    float t = 0.0f;
    for (int i = 0; i < nw; i++) t = fmaxf(t, sh[i]);
    return t;
}

// a position folded into the replica's box (k_mb_cell_keys' formula); the
// staged positions and the cells come from the same bits
__device__ __forceinline__ double3 mb_fold(const MbRep& rp, const double4& x)
{
    return make_double3(x.x - rp.L[0] * floor(x.x / rp.L[0] + 0.5), x.y - rp.L[1] * floor(x.y / rp.L[1] + 0.5),
                        x.z - rp.L[2] * floor(x.z / rp.L[2] + 0.5));
}

__device__ __forceinline__ int mb_cell_of(const MbRep& rp, const double3& f)
{
    const double us[3] = {f.x, f.y, f.z};
    int i3[3];
    for (int a = 0; a < 3; a++) {
        const int ic = (int)floor((us[a] + rp.L[a] / 2.0) * (double)rp.nc[a] / rp.L[a]);
        i3[a] = ic < 0 ? 0 : (ic >= rp.nc[a] ? rp.nc[a] - 1 : ic);
    }
    return i3[0] + rp.nc[0] * (i3[1] + rp.nc[1] * i3[2]);
}

// the candidates of places [e0, e1) for the slot at place s (position px, py,
// pz with the run's periodic shift applied), MB_GROUP at a time: the distances
// first, as a mask of hits, then the hits in place order into the slot's ring
// of 16-bit list words, a completed chunk copied out after each group. MI:
// minimum image per pair, for replicas with fewer than three cells on an axis.
// The last group's loads may pass e1 by up to MB_GROUP - 1 places (the staging
// is padded by 16); those are masked off.
#ifndef MB_GROUP
#define MB_GROUP 8
#endif
template <bool MI>
__device__ __forceinline__ void mb_scan_run(const float4* __restrict__ spos, int e0, int e1, int s, float px,
                                            float py, float pz, const MbRep& rp, float rl2, MtRing16& g)
{
    for (int e = e0; e < e1; e += MB_GROUP) {
        uint32_t m = 0;
#pragma unroll
        for (int i = 0; i < MB_GROUP; i++) {
            const float4 pa = spos[e + i];
            float dx = px - pa.x, dy = py - pa.y, dz = pz - pa.z;
            if (MI) {
                dx = mb_min_image(dx, rp.Lf[0], rp.iLf[0]);
                dy = mb_min_image(dy, rp.Lf[1], rp.iLf[1]);
                dz = mb_min_image(dz, rp.Lf[2], rp.iLf[2]);
            }
            m |= (fmaf(dz, dz, fmaf(dy, dy, dx * dx)) < rl2 ? 1u : 0u) << i;
        }
        const int ne = e1 - e;
        if (ne < MB_GROUP) m &= (1u << ne) - 1u;
        const unsigned so = (unsigned)(s - e);
        if (so < (unsigned)MB_GROUP) m &= ~(1u << so);
        while (m) {
            mt_ring_put(g, e + __builtin_ctz(m));
            m &= m - 1u;
        }
        mt_ring_flush(g);
    }
}

// ascending sort of one cell's entries a[0, cnt) (places of the previous
// order) by their staged x, ties by atom index (perm[place]), by one
// wavefront: two entries a lane ranked in registers (cnt <= 128), lane 0 by
// insertion beyond that
__device__ inline void mb_sort_cell_x(int* __restrict__ a, int cnt, const float4* __restrict__ spos,
                                      const int* __restrict__ perm)
{
    const int lane = threadIdx.x & 63;
    auto key = [&](int q) -> unsigned long long {
        const unsigned b = __float_as_uint(spos[q].x);
        const unsigned o = (b & 0x80000000u) ? ~b : (b | 0x80000000u);   // float order as unsigned order
        return ((unsigned long long)o << 32) | (unsigned)perm[q];
    };
    if (cnt <= 128) {
        const int v0 = lane < cnt ? a[lane] : 0, v1 = lane + 64 < cnt ? a[lane + 64] : 0;
        const unsigned long long k0 = lane < cnt ? key(v0) : ~0ull, k1 = lane + 64 < cnt ? key(v1) : ~0ull;
        int q0 = 0, q1 = 0;
        const int n0 = min(cnt, 64);
        for (int j = 0; j < n0; j++) {
            const unsigned long long kj = ((unsigned long long)(unsigned)__shfl((int)(k0 >> 32), j) << 32) |
                                          (unsigned)__shfl((int)k0, j);
            q0 += kj < k0 ? 1 : 0;
            q1 += kj < k1 ? 1 : 0;
        }
        for (int j = 0; j < cnt - 64; j++) {
            const unsigned long long kj = ((unsigned long long)(unsigned)__shfl((int)(k1 >> 32), j) << 32) |
                                          (unsigned)__shfl((int)k1, j);
            q0 += kj < k0 ? 1 : 0;
            q1 += kj < k1 ? 1 : 0;
        }
        __builtin_amdgcn_wave_barrier();              // (one wavefront: every read of a[] above precedes the writes)
        if (lane < cnt) a[q0] = v0;
        if (lane + 64 < cnt) a[q1] = v1;
    } else if (lane == 0) {
        for (int i = 1; i < cnt; i++) {
            const int v = a[i];
            const unsigned long long kv = key(v);
            int j = i - 1;
            while (j >= 0 && key(a[j]) > kv) { a[j + 1] = a[j]; j--; }
            a[j + 1] = v;
        }
    }
}

// first place of [e0, e1) whose staged x is at least t (e1 if none); the
// places are in ascending x
__device__ __forceinline__ int mb_lower_x(const float4* __restrict__ spos, int e0, int e1, float t)
{
    int len = e1 - e0;
    while (len > 0) {
        const int h = len >> 1;
        if (spos[e0 + h].x < t) { e0 += h + 1; len -= h + 1; } else len = h;
    }
    return e0;
}

// a replica's chain (fast arithmetic; chains of 1..8, which the batch takes)
__device__ __forceinline__ double mb_nh_fast(double* nh, int M, const MbPar& p, const MbRep& rp, double KE2)
{
    double *eta = nh, *peta = nh + M;
    const double* Q = nh + 2 * M;
    switch (M) {
    case 1: return mt_nh_propagate_fast<1>(eta, peta, Q, p.nresp, rp.T, p.dt, rp.g, p.w, KE2);
    case 2: return mt_nh_propagate_fast<2>(eta, peta, Q, p.nresp, rp.T, p.dt, rp.g, p.w, KE2);
    case 3: return mt_nh_propagate_fast<3>(eta, peta, Q, p.nresp, rp.T, p.dt, rp.g, p.w, KE2);
    case 4: return mt_nh_propagate_fast<4>(eta, peta, Q, p.nresp, rp.T, p.dt, rp.g, p.w, KE2);
    case 5: return mt_nh_propagate_fast<5>(eta, peta, Q, p.nresp, rp.T, p.dt, rp.g, p.w, KE2);
    case 6: return mt_nh_propagate_fast<6>(eta, peta, Q, p.nresp, rp.T, p.dt, rp.g, p.w, KE2);
    case 7: return mt_nh_propagate_fast<7>(eta, peta, Q, p.nresp, rp.T, p.dt, rp.g, p.w, KE2);
    case 8: return mt_nh_propagate_fast<8>(eta, peta, Q, p.nresp, rp.T, p.dt, rp.g, p.w, KE2);
    default: return 1.0;
    }
}

// mode 0: prime (list build, forces with energies, no motion); 1: nsteps MD
// steps, energies in the last when want_e_last; 2: forces only, 3: list build
// only (timing)
template <bool NVT, int G>
__global__ void __launch_bounds__(MB_RUN_TPB) k_mb_run(MbDev d, MbPar p, MbRun u, int nsteps, int mode,
                                                       int want_e_last)
{
    extern __shared__ float4 smem4[];
    __shared__ double shd[MB_RUN_TPB / 64];
    __shared__ float shf[MB_RUN_TPB / 64];
    __shared__ double s_scale[2];                     // [0] end scale of the last step, [1] start scale
    const int T = d.T;
    float4* spos = smem4;                             // [maxn + 16] staged positions; .w the type's row offset in bytes
    float2* stab = reinterpret_cast<float2*>(spos + u.maxn + 16);   // [T T MT_NPTS] (a_f, b_f)
    int* cst = reinterpret_cast<int*>(stab + (size_t)T * T * MT_NPTS);   // [maxcells + 1] cell starts
    int* cfill = cst + u.maxcells + 1;                // [maxcells]
    uint32_t* sring = reinterpret_cast<uint32_t*>(cfill + u.maxcells);   // [block * MT_RING_DW] list-word rings
    const int tid = threadIdx.x, nt = blockDim.x, wave = tid >> 6, nwav = nt >> 6;
    const int r = u.order[blockIdx.x];
    const MbRep rp = d.rep[r];
    const int n = rp.n, a0 = rp.atom0, ntot = d.n;
    const int M = d.M, stride = 3 * M + 2;
    double* nh = d.nh + (size_t)r * stride;
    for (int i = tid; i < T * T * MT_NPTS; i += nt) stab[i] = d.nbcoef2[i];
    const float Lx = rp.Lf[0], Ly = rp.Lf[1], Lz = rp.Lf[2];
    const float iLx = rp.iLf[0], iLy = rp.iLf[1], iLz = rp.iLf[2];
    const double dt_2 = p.dt / 2.0;
    const float x0 = p.x0, invd = p.inv_delta;
    uint4* __restrict__ nl4 = reinterpret_cast<uint4*>(u.nl16);

    auto stage = [&]() {
        for (int s = tid; s < n; s += nt) {
            const int k = u.perm[a0 + s];
            const double3 x = mb_fold(rp, d.pos[k]);
            const int ty = __float_as_int(d.posf[k].w);
            spos[s] = make_float4((float)x.x, (float)x.y, (float)x.z, __int_as_float(ty * MT_NPTS * (int)sizeof(float2)));
        }
    };

    // cell sort of the replica (counting sort in LDS, then ascending atom
    // order inside each cell) and its 16-bit skin lists
    auto rebuild = [&]() {
        const int ncell = rp.ncell;
        for (int c = tid; c < ncell; c += nt) { cst[c] = 0; cfill[c] = 0; }
        __syncthreads();
        for (int s = tid; s < n; s += nt) {
            const double3 f = mb_fold(rp, d.pos[u.perm[a0 + s]]);
            const int key = mb_cell_of(rp, f);
            spos[s].x = (float)f.x;                   // the sort's key (restaged below)
            spos[s].w = __int_as_float(key);
            atomicAdd(&cst[key], 1);
        }
        __syncthreads();
        if (wave == 0) {
            // exclusive scan of the counts by one wavefront: a run of cells a lane
            const int lane = tid & 63, per = (ncell + 63) >> 6, c0 = lane * per;
            int sum = 0;
            for (int c = c0; c < min(c0 + per, ncell); c++) sum += cst[c];
            int incl = sum;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            int run = incl - sum;
            for (int c = c0; c < min(c0 + per, ncell); c++) {
                const int v = cst[c];
                cst[c] = run;
                run += v;
            }
            if (lane == 63) cst[ncell] = incl;
        }
        __syncthreads();
        for (int s = tid; s < n; s += nt) {
            const int key = __float_as_int(spos[s].w);
            u.perm2[a0 + cst[key] + atomicAdd(&cfill[key], 1)] = s;
        }
        __syncthreads();
        // each cell in ascending x (ties by atom index), a wavefront per cell
        for (int c = wave; c < ncell; c += nwav)
            mb_sort_cell_x(u.perm2 + a0 + cst[c], cst[c + 1] - cst[c], spos, u.perm + a0);
        __syncthreads();
        for (int s = tid; s < n; s += nt) u.perm2[a0 + s] = u.perm[a0 + u.perm2[a0 + s]];   // places -> atoms
        __syncthreads();
        for (int s = tid; s < n; s += nt) {
            const int k = u.perm2[a0 + s];
            u.perm[a0 + s] = k;
            d.posref[k] = d.pos[k];
        }
        __syncthreads();
        stage();
        __syncthreads();
        // lists: every distinct cell within one cell of the slot's own (per
        // axis three offsets with three or more cells, else every cell once),
        // in place order within each row of cells; with three or more cells
        // on every axis only the rows and x windows that can hold a pair
        // within the list radius are tested (the same lists)
        const int cap = u.ncap16;
        for (int s = tid; s < n; s += nt) {
            int lo = 0, hi = ncell;                   // the cell holding place s
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (cst[mid] <= s) lo = mid; else hi = mid;
            }
            const int hx = lo % rp.nc[0], hy = (lo / rp.nc[0]) % rp.nc[1], hz = lo / (rp.nc[0] * rp.nc[1]);
            int olo[3], ohi[3];
            for (int a = 0; a < 3; a++) {
                const int h = a == 0 ? 1 : rp.sk;    // (two per list radius only with 5 or more cells)
                if (rp.nc[a] >= 3) { olo[a] = -h; ohi[a] = h; }
                else { olo[a] = 0; ohi[a] = rp.nc[a] - 1; }
            }
            const float4 pk = spos[s];
            MtRing16 g{sring + (size_t)tid * MT_RING_DW, nl4 + a0 + s, (size_t)ntot, cap >> 3, 0, 0};
            const int nc0 = rp.nc[0], nc1 = rp.nc[1], nc2 = rp.nc[2];
            const bool mi = nc0 < 3 || nc1 < 3 || nc2 < 3;
            const float rl2 = p.rlist2, wy = Ly / (float)nc1, wz = Lz / (float)nc2;
            for (int oz = olo[2]; oz <= ohi[2]; oz++) {
                // neighbour cells across the box's faces: the slot's position
                // moved by a box length instead of a minimum image per pair
                int cz = nc2 >= 3 ? hz + oz : oz;
                float pz = pk.z;
                if (cz < 0) { cz += nc2; pz += Lz; } else if (cz >= nc2) { cz -= nc2; pz -= Lz; }
                // distance from the slot to the cells' z range, less a margin
                // for the rounding of the edges (so never more than the truth)
                const float zlo = fmaf((float)cz, wz, -0.5f * Lz);
                const float dz = fmaxf(fmaxf(zlo - pz, pz - (zlo + wz)) - 1e-3f, 0.f);
                for (int oy = olo[1]; oy <= ohi[1]; oy++) {
                    int cy = nc1 >= 3 ? hy + oy : oy;
                    float py = pk.y;
                    if (cy < 0) { cy += nc1; py += Ly; } else if (cy >= nc1) { cy -= nc1; py -= Ly; }
                    const float ylo = fmaf((float)cy, wy, -0.5f * Ly);
                    const float dy = fmaxf(fmaxf(ylo - py, py - (ylo + wy)) - 1e-3f, 0.f);
                    const float r2yz = fmaf(dy, dy, dz * dz);
                    // no atom of the row within the list radius
                    if (!mi && r2yz >= rl2) continue;
                    // half-width of the x window the row's atoms can fall in
                    const float W = sqrtf(fmaxf(rl2 - r2yz, 0.f)) + 1e-3f;
                    // the row's cells within one of hx are contiguous places:
                    // one run, or two where the row wraps
                    const int row = nc0 * (cy + nc1 * cz);
                    int ea0, ea1, eb0 = 0, eb1 = 0;
                    float pxb = pk.x;
                    if (nc0 < 3) {
                        ea0 = cst[row]; ea1 = cst[row + nc0];
                    } else if (hx == 0) {
                        ea0 = cst[row]; ea1 = cst[row + 2];
                        eb0 = cst[row + nc0 - 1]; eb1 = cst[row + nc0]; pxb = pk.x + Lx;
                    } else if (hx == nc0 - 1) {
                        ea0 = cst[row + hx - 1]; ea1 = cst[row + nc0];
                        eb0 = cst[row]; eb1 = cst[row + 1]; pxb = pk.x - Lx;
                    } else {
                        ea0 = cst[row + hx - 1]; ea1 = cst[row + hx + 2];
                    }
                    for (int part = 0; part < 2; part++) {
                        int e0 = part ? eb0 : ea0, e1 = part ? eb1 : ea1;
                        const float px = part ? pxb : pk.x;
                        if (mi) {
                            mb_scan_run<true>(spos, e0, e1, s, px, py, pz, rp, rl2, g);
                        } else {
                            // the run is in ascending x: only its x window
                            e0 = mb_lower_x(spos, e0, e1, px - W);
                            e1 = mb_lower_x(spos, e0, e1, px + W);
                            mb_scan_run<false>(spos, e0, e1, s, px, py, pz, rp, rl2, g);
                        }
                    }
                }
            }
            const int cnt = g.cnt;
            if (cnt & 7) {
                for (int i = cnt & 7; i < 8; i++) mt_ring_put(g, s);   // pad: the atom itself (r = 0, no force)
                mt_ring_flush(g);
            }
            u.ncount16[a0 + s] = cnt;
            if (cnt > u.nlimit) { atomicOr(&d.flags[0], 1u); atomicMax(d.nreq, cnt); }
        }
        __syncthreads();
    };

    // pair forces of every slot (G lanes a slot): FP32 minimum image on the
    // staged positions, the table's segment (its last one, zero, beyond rc or
    // below the first point), eight pairs of a chunk summed in FP32 and the
    // chunk sums in FP64. Then, in the lane that holds the force: the cap, the
    // half kick and the wrap (integrate), or the force alone (mode 0 and 2).
    // Returns this thread's part of sum p.p/m after the kick.
    auto forces = [&](bool want_e, bool integrate) -> double {
        const int gl = tid % G;
        double ke = 0.0;
        for (int v = tid / G; v < n; v += nt / G) {   // the G lanes of a slot share v
            const float4 pk = spos[v];
            const int tk = __float_as_int(pk.w);      // home row: its type's offset times T
            const char* row = reinterpret_cast<const char*>(stab) + (size_t)tk * T;
            const int nch = (min(u.ncount16[a0 + v], u.ncap16) + 7) >> 3;
            double Fx = 0.0, Fy = 0.0, Fz = 0.0;
            float U = 0.0f;
            for (int q = gl; q < nch; q += G) {
                const uint4 w = nl4[(size_t)q * ntot + a0 + v];
                const int li[8] = {(int)(w.x & 0xffffu), (int)(w.x >> 16), (int)(w.y & 0xffffu), (int)(w.y >> 16),
                                   (int)(w.z & 0xffffu), (int)(w.z >> 16), (int)(w.w & 0xffffu), (int)(w.w >> 16)};
                float cx = 0.f, cy = 0.f, cz = 0.f;
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const float4 a = spos[li[i]];
                    const float dx = mb_min_image(pk.x - a.x, Lx, iLx);
                    const float dy = mb_min_image(pk.y - a.y, Ly, iLy);
                    const float dz = mb_min_image(pk.z - a.z, Lz, iLz);
                    const float r2 = mt_lean_r2(dx, dy, dz);
                    int seg;
                    const float rinv = mt_lean_rinv(r2, x0, invd, seg);
                    const int off = __float_as_int(a.w) + seg * (int)sizeof(float2);
                    const float2 c = *reinterpret_cast<const float2*>(row + off);
                    const float fr = fmaf(c.y, rinv, c.x);
                    if (want_e) {
                        const float4 c4 = d.nbcoef[((size_t)tk * T + __float_as_int(a.w)) / sizeof(float2) + seg];
                        U += 0.5f * fmaf(c4.z, r2 * rinv, c4.w);
                    }
                    cx = fmaf(dx, fr, cx);
                    cy = fmaf(dy, fr, cy);
                    cz = fmaf(dz, fr, cz);
                }
                Fx += (double)cx;
                Fy += (double)cy;
                Fz += (double)cz;
            }
            mt_lane_sum<G>(Fx, Fy, Fz, U);
            if (gl == 0) {
                const int k = u.perm[a0 + v];
                double4 F = make_double4(Fx, Fy, Fz, 0.0);
                if (p.Fcut > 0) {                     // cutForces (VelVerlet.cpp:81,96-122)
                    if (fabs(F.x) > p.Fcut) F.x = ((F.x >= 0.0) ? 1.0 : -1.0) * p.Fcut;
                    if (fabs(F.y) > p.Fcut) F.y = ((F.y >= 0.0) ? 1.0 : -1.0) * p.Fcut;
                    if (fabs(F.z) > p.Fcut) F.z = ((F.z >= 0.0) ? 1.0 : -1.0) * p.Fcut;
                }
                d.frc[k] = F;
                if (want_e) d.epot[k] = U;
                if (integrate) {
                    double4 pm = d.mom[k], x = d.pos[k];
                    pm.x += F.x * dt_2; pm.y += F.y * dt_2; pm.z += F.z * dt_2;
                    double* xs[3] = {&x.x, &x.y, &x.z};
                    for (int a = 0; a < 3; a++) {     // pbcWrap into the replica's box
                        const double h = rp.L[a] / 2.0;
                        if (*xs[a] >= h || *xs[a] < -h) *xs[a] -= rp.L[a] * floor(*xs[a] / rp.L[a] + 0.5);
                    }
                    d.mom[k] = pm;
                    d.pos[k] = x;
                    ke += ((pm.x * pm.x) + (pm.y * pm.y) + (pm.z * pm.z)) / pm.w;
                }
            }
        }
        return ke;
    };

    stage();
    __syncthreads();
    if (mode != 1) {
        if (mode == 0 || mode == 3) rebuild();
        if (mode != 3) forces(mode == 0, false);
        return;
    }
    for (int t = 0; t < nsteps; t++) {
        const bool last = (t == nsteps - 1);
        // the step's start: chain on the kinetic energy after the previous
        // step's end scale (kept analytically, KE2 S^2), one thread
        if (NVT) {
            if (tid == 0) {
                s_scale[0] = nh[3 * M];               // end scale of the previous step, not yet applied
                const double Sf = mb_nh_fast(nh, M, p, rp, nh[3 * M + 1]);
                nh[3 * M + 1] *= Sf * Sf;
                nh[3 * M] = 1.0;
                s_scale[1] = Sf;
            }
            __syncthreads();
        }
        float d2 = 0.0f;
        for (int s = tid; s < n; s += nt) {
            const int k = u.perm[a0 + s];
            double4 pm = d.mom[k], x = d.pos[k];
            const double4 F = d.frc[k];
            if (NVT) {
                const double Sb = s_scale[0], Sf = s_scale[1];
                pm.x *= Sb; pm.y *= Sb; pm.z *= Sb;
                pm.x *= Sf; pm.y *= Sf; pm.z *= Sf;
            }
            pm.x += F.x * dt_2; pm.y += F.y * dt_2; pm.z += F.z * dt_2;
            const double c = p.dt / pm.w;
            x.x += pm.x * c; x.y += pm.y * c; x.z += pm.z * c;
            d.mom[k] = pm;
            d.pos[k] = x;
            spos[s].x = (float)x.x; spos[s].y = (float)x.y; spos[s].z = (float)x.z;
            const double4 q = d.posref[k];
            double dx = x.x - q.x, dy = x.y - q.y, dz = x.z - q.z;
            dx -= rp.L[0] * rint(dx / rp.L[0]);
            dy -= rp.L[1] * rint(dy / rp.L[1]);
            dz -= rp.L[2] * rint(dz / rp.L[2]);
            const float dd = (float)(dx * dx + dy * dy + dz * dz);
            if (isnan(dd)) atomicOr(&d.flags[1], 1u);  // a non-finite position (NaN or inf)
            d2 = fmaxf(d2, dd);
        }
        if (mb_block_max(d2, shf) > p.half_skin2) rebuild();
        __syncthreads();
        const double ke = forces(last && want_e_last, true);
        if (NVT) {
            // the step's end: chain on sum p.p/m after the kick; its scale is
            // applied at the next step's start, or below after the last step
            const double KE2 = mb_block_sum(ke, shd);
            if (tid == 0) {
                const double Sb = mb_nh_fast(nh, M, p, rp, KE2);
                nh[3 * M] = Sb;
                nh[3 * M + 1] = KE2 * Sb * Sb;
            }
        }
        __syncthreads();
    }
    if (NVT) {
        const double Sb = nh[3 * M];
        for (int s = tid; s < n; s += nt) {
            const int k = u.perm[a0 + s];
            double4 pm = d.mom[k];
            pm.x *= Sb; pm.y *= Sb; pm.z *= Sb;
            d.mom[k] = pm;
        }
        __syncthreads();
        if (tid == 0) nh[3 * M] = 1.0;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

thread_local std::string g_batch_err;

struct DevMem {
    std::vector<void*> ptrs;
    template <class T> T* alloc(size_t n)
    {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
    void release(void* p)
    {
        for (size_t i = 0; i < ptrs.size(); i++)
            if (ptrs[i] == p) { (void)hipFree(p); ptrs.erase(ptrs.begin() + (long)i); return; }
    }
    ~DevMem() { for (void* p : ptrs) (void)hipFree(p); }
};

} // namespace

std::vector<float4> mtmd_host_table(int which, int kind, const double* prm, const float* profile, float x0, float x1);
void mtmd_host_segment(const float4* t, int i, float& af, float& bf, float& ae, float& be);

struct mtmd_batch {
    DevMem mem;
    MbDev d{};
    MbPar p{};
    hipStream_t stream = nullptr;
    std::vector<MbRep> rep;
    int nrep = 0, n = 0;
    int force_block = 256, list_block = 128;
    // the persistent replica kernel (k_mb_run), used when every replica fits
    // its staging: workgroup size and lanes per atom (autotuned)
    bool run = false;
    MbRun u{};
    int run_block = 512, run_g = 2, run_split = 1;
    float rlist = 0.f;
    int* order_batch = nullptr;   // workgroup -> replica of a whole-batch launch
    int* order_ident = nullptr;   // replica i at i (one-replica launches)
    // the state a call starts from, to run it again after the lists grow
    double4 *snap_pos = nullptr, *snap_mom = nullptr, *snap_frc = nullptr;
    double* snap_nh = nullptr;
    bool list_slack = true;       // MTMD_BATCH_LIST_SLACK=0: rows sized to the longest, no margin
    bool have_forces = false, ke_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int64_t steps = 0;
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
    ~mtmd_batch()
    {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

#define MB_HIP(b, call)                                                                                       \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess) return (b)->fail(MTMD_EHIP, "%s: %s", #call, hipGetErrorString(e_));            \
    } while (0)

// launches of one step piece over replicas [r0, r1) (all replicas: batched)
static void mb_rebuild(mtmd_batch* b, int r0, int r1, hipStream_t s)
{
    const MbDev& d = b->d;
    const int a0 = b->rep[r0].atom0, a1 = b->rep[r1 - 1].atom0 + b->rep[r1 - 1].n;
    const int c0 = b->rep[r0].cell0, c1 = b->rep[r1 - 1].cell0 + b->rep[r1 - 1].ncell;
    const int na = a1 - a0, nc = c1 - c0;
    hipLaunchKernelGGL(k_mb_cells_clear, dim3(mb_blk(nc, MB_TPB)), dim3(MB_TPB), 0, s, d, c0, c1);
    hipLaunchKernelGGL(k_mb_cell_keys, dim3(mb_blk(na, MB_TPB)), dim3(MB_TPB), 0, s, d, a0, a1);
    hipLaunchKernelGGL(k_mb_scan, dim3(r1 - r0), dim3(MB_TPB), 0, s, d, r0);
    hipLaunchKernelGGL(k_mb_scatter, dim3(mb_blk(na, MB_TPB)), dim3(MB_TPB), 0, s, d, a0, a1);
    hipLaunchKernelGGL(k_mb_cell_sort, dim3(mb_blk(nc, MB_TPB / 64)), dim3(MB_TPB), 0, s, d, c0, c1);
    hipLaunchKernelGGL(k_mb_neigh, dim3(mb_blk(na, b->list_block)), dim3(b->list_block), 0, s, d, b->p, a0, a1);
    hipLaunchKernelGGL(k_mb_clear_rflags, dim3(mb_blk(r1 - r0, 64)), dim3(64), 0, s, d, r0, r1);
}

static void mb_forces(mtmd_batch* b, int r0, int r1, bool want_e, hipStream_t s)
{
    const MbDev& d = b->d;
    const int a0 = b->rep[r0].atom0, a1 = b->rep[r1 - 1].atom0 + b->rep[r1 - 1].n;
    const int bs = b->force_block;
    if (want_e)
        hipLaunchKernelGGL(k_mb_force<true>, dim3(mb_blk(a1 - a0, bs)), dim3(bs), 0, s, d, b->p, a0, a1);
    else
        hipLaunchKernelGGL(k_mb_force<false>, dim3(mb_blk(a1 - a0, bs)), dim3(bs),
                           (size_t)d.T * d.T * MT_NPTS * sizeof(float2), s, d, b->p, a0, a1);
}

// one MD step of replicas [r0, r1) (MDLoop.cpp:124-138 for each replica)
static void mb_step(mtmd_batch* b, int r0, int r1, bool want_e, hipStream_t s)
{
    const MbDev& d = b->d;
    const int a0 = b->rep[r0].atom0, a1 = b->rep[r1 - 1].atom0 + b->rep[r1 - 1].n, na = a1 - a0;
    const int nr = r1 - r0;
    if (b->p.nvt) hipLaunchKernelGGL(k_mb_nh, dim3(mb_blk(nr, 64)), dim3(64), 0, s, d, b->p, r0, r1);
    hipLaunchKernelGGL(k_mb_kick_drift, dim3(mb_blk(na, MB_TPB)), dim3(MB_TPB), 0, s, d, b->p, a0, a1);
    mb_rebuild(b, r0, r1, s);
    mb_forces(b, r0, r1, want_e, s);
    hipLaunchKernelGGL(k_mb_cap_kick, dim3(mb_blk(na, MB_TPB)), dim3(MB_TPB), 0, s, d, b->p, a0, a1);
    if (b->p.nvt) {
        hipLaunchKernelGGL(k_mb_reduce, dim3(nr), dim3(MB_TPB), 0, s, d, r0, 0);
        hipLaunchKernelGGL(k_mb_nh, dim3(mb_blk(nr, 64)), dim3(64), 0, s, d, b->p, r0, r1);
    }
    hipLaunchKernelGGL(k_mb_scale_wrap, dim3(mb_blk(na, MB_TPB)), dim3(MB_TPB), 0, s, d, b->p, a0, a1);
    if (b->p.nvt) hipLaunchKernelGGL(k_mb_reduce, dim3(nr), dim3(MB_TPB), 0, s, d, r0, 0);
}

// dynamic LDS of a k_mb_run launch with `block` threads
static size_t mb_run_lds(const mtmd_batch* b, int block)
{
    size_t bytes = (size_t)(b->u.maxn + 16) * sizeof(float4) + (size_t)b->d.T * b->d.T * MT_NPTS * sizeof(float2) +
                   (size_t)(2 * b->u.maxcells + 1) * sizeof(int) + (size_t)block * MT_RING_DW * sizeof(uint32_t);
    return (bytes + 15) & ~(size_t)15;
}

template <bool NVT, int G>
static hipError_t mb_run_one(mtmd_batch* b, const MbRun& u, int grid, int nsteps, int mode, int want_e,
                             hipStream_t s)
{
    const size_t lds = mb_run_lds(b, b->run_block);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mb_run<NVT, G>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_mb_run<NVT, G>), dim3(grid), dim3(b->run_block), lds, s, b->d, b->p, u, nsteps, mode,
                       want_e);
    return hipGetLastError();
}

// k_mb_run over the replicas order[0, grid)
static hipError_t mb_run_launch(mtmd_batch* b, const int* order, int grid, int nsteps, int mode, int want_e,
                                hipStream_t s)
{
    MbRun u = b->u;
    u.order = order;
    if (b->p.nvt) {
        if (b->run_g == 1) return mb_run_one<true, 1>(b, u, grid, nsteps, mode, want_e, s);
        if (b->run_g == 2) return mb_run_one<true, 2>(b, u, grid, nsteps, mode, want_e, s);
        return mb_run_one<true, 4>(b, u, grid, nsteps, mode, want_e, s);
    }
    if (b->run_g == 1) return mb_run_one<false, 1>(b, u, grid, nsteps, mode, want_e, s);
    if (b->run_g == 2) return mb_run_one<false, 2>(b, u, grid, nsteps, mode, want_e, s);
    return mb_run_one<false, 4>(b, u, grid, nsteps, mode, want_e, s);
}

// the persistent kernel's cell grids: x cells at least a list radius wide,
// y and z cells `split` per list radius where each of those axes then holds
// five or more cells and x three or more (else one per radius); uploads the
// replicas and sets the staging's cell capacity. False when a grid would
// exceed MB_RUN_MAXCELLS cells.
static bool mb_set_split(mtmd_batch* b, int split, float rlist)
{
    int maxcells = 0;
    std::vector<MbRep> rep = b->rep;
    for (MbRep& rp : rep) {
        for (int a = 0; a < 3; a++) rp.nc[a] = std::max(1, (int)std::floor(rp.L[a] / (double)rlist));
        rp.sk = 1;
        if (split == 2 && rp.nc[0] >= 3) {
            const int ny = (int)std::floor(2.0 * rp.L[1] / (double)rlist), nz = (int)std::floor(2.0 * rp.L[2] / (double)rlist);
            if (ny >= 5 && nz >= 5) { rp.nc[1] = ny; rp.nc[2] = nz; rp.sk = 2; }
        }
        rp.ncell = rp.nc[0] * rp.nc[1] * rp.nc[2];
        if (rp.ncell > MB_RUN_MAXCELLS) return false;
        maxcells = std::max(maxcells, rp.ncell);
    }
    if (hipMemcpy(b->d.rep, rep.data(), rep.size() * sizeof(MbRep), hipMemcpyHostToDevice) != hipSuccess) return false;
    b->u.maxcells = maxcells;
    b->run_split = split;
    return true;
}

// the chain's kinetic energy slot from the reduced sums, no scale pending
__global__ void k_mb_ke_slot(MbDev d)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d.nrep || d.M == 0) return;
    double* nh = d.nh + (size_t)r * (3 * d.M + 2);
    nh[3 * d.M] = 1.0;
    nh[3 * d.M + 1] = d.red[3 * r];
}

static int mb_check(mtmd_batch* b)
{
    unsigned int fl[2];
    MB_HIP(b, hipMemcpyAsync(fl, b->d.flags, sizeof fl, hipMemcpyDeviceToHost, b->stream));
    MB_HIP(b, hipStreamSynchronize(b->stream));
    if (fl[1]) return b->fail(MTMD_ENAN, "non-finite atom position");
    if (fl[0]) {
        int need = 0;
        MB_HIP(b, hipMemcpy(&need, b->d.nreq, sizeof(int), hipMemcpyDeviceToHost));
        return b->fail(MTMD_ENEIGH, "neighbour list overflow (capacity %d, needed %d)",
                       b->run ? b->u.ncap16 : b->d.ncap, need);
    }
    return MTMD_OK;
}

// list overflow of the last launches: the longest row requested (0: none);
// non-finite positions are an error
static int mb_overflow(mtmd_batch* b, int& need)
{
    unsigned int fl[2];
    need = 0;
    MB_HIP(b, hipMemcpyAsync(fl, b->d.flags, sizeof fl, hipMemcpyDeviceToHost, b->stream));
    MB_HIP(b, hipStreamSynchronize(b->stream));
    if (fl[1]) return b->fail(MTMD_ENAN, "non-finite atom position");
    if (fl[0]) {
        MB_HIP(b, hipMemcpy(&need, b->d.nreq, sizeof(int), hipMemcpyDeviceToHost));
        need = std::max(need, b->u.nlimit + 1);
    }
    return MTMD_OK;
}

// the persistent kernel's rows grown to hold `need` words (with a fifth more
// and 16 words of margin unless MTMD_BATCH_LIST_SLACK=0, which accepts rows
// up to `need` only, so that any longer row grows them again), the overflow
// flag and request cleared (CMDFFMatrices grows its list and builds it again,
// MDFFMatrices.cpp:163-178)
static int mb_grow16(mtmd_batch* b, int need)
{
    int cap = b->list_slack ? (int)std::ceil(need * 1.2) + 16 : need;
    cap = (cap + 7) & ~7;
    if (cap > 65536) return b->fail(MTMD_ENEIGH, "neighbour list rows above 65536 words requested (%d)", need);
    if (cap > b->u.ncap16) {
        b->mem.release(b->u.nl16);
        b->u.nl16 = b->mem.alloc<uint32_t>((size_t)(cap / 2) * b->n);
        if (!b->u.nl16) return b->fail(MTMD_ENOMEM, "list growth to %d words a row failed", cap);
        b->u.ncap16 = cap;
    }
    b->u.nlimit = b->list_slack ? b->u.ncap16 : need;
    MB_HIP(b, hipMemsetAsync(b->d.flags, 0, sizeof(unsigned int), b->stream));
    MB_HIP(b, hipMemsetAsync(b->d.nreq, 0, sizeof(int), b->stream));
    return MTMD_OK;
}

static int mb_snapshot(mtmd_batch* b, bool restore)
{
    const size_t a = (size_t)b->n * sizeof(double4), h = (size_t)b->nrep * (3 * b->d.M + 2) * sizeof(double);
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    if (restore) {
        MB_HIP(b, hipMemcpyAsync(b->d.pos, b->snap_pos, a, k, b->stream));
        MB_HIP(b, hipMemcpyAsync(b->d.mom, b->snap_mom, a, k, b->stream));
        MB_HIP(b, hipMemcpyAsync(b->d.frc, b->snap_frc, a, k, b->stream));
        MB_HIP(b, hipMemcpyAsync(b->d.nh, b->snap_nh, h, k, b->stream));
    } else {
        MB_HIP(b, hipMemcpyAsync(b->snap_pos, b->d.pos, a, k, b->stream));
        MB_HIP(b, hipMemcpyAsync(b->snap_mom, b->d.mom, a, k, b->stream));
        MB_HIP(b, hipMemcpyAsync(b->snap_frc, b->d.frc, a, k, b->stream));
        MB_HIP(b, hipMemcpyAsync(b->snap_nh, b->d.nh, h, k, b->stream));
    }
    return MTMD_OK;
}

// launches of a call on the persistent kernel (run(): whole batch or replica
// by replica), with list growth: after an overflow the rows grow, the state
// the call started from is restored, lists are built from it, and the call
// runs again
template <class Run>
static int mb_run_call(mtmd_batch* b, Run run)
{
    int rc = mb_snapshot(b, false);
    if (rc) return rc;
    for (int attempt = 0; attempt < 16; attempt++) {
        MB_HIP(b, run());
        int need = 0;
        if ((rc = mb_overflow(b, need))) return rc;
        if (!need) return MTMD_OK;
        // grow until a build from the restored state fits, then run again
        for (;;) {
            if ((rc = mb_grow16(b, need)) || (rc = mb_snapshot(b, true))) return rc;
            MB_HIP(b, mb_run_launch(b, b->order_batch, b->nrep, 0, 3, 0, b->stream));
            if ((rc = mb_overflow(b, need))) return rc;
            if (!need) break;
        }
        if (std::getenv("MTMD_BATCH_LOG"))
            std::fprintf(stderr, "batch: lists grown to %d words a row, call run again\n", b->u.ncap16);
    }
    return b->fail(MTMD_ENEIGH, "neighbour lists kept overflowing (capacity %d)", b->u.ncap16);
}

// lists built and forces and energies computed on the current state (mode 0),
// the rows grown until the lists fit
static int mb_build_fit(mtmd_batch* b)
{
    for (int attempt = 0; attempt <= 16; attempt++) {
        MB_HIP(b, mb_run_launch(b, b->order_batch, b->nrep, 0, 0, 1, b->stream));
        int need = 0, rc = mb_overflow(b, need);
        if (rc) return rc;
        if (!need) return MTMD_OK;
        if ((rc = mb_grow16(b, need))) return rc;
    }
    return b->fail(MTMD_ENEIGH, "neighbour lists kept overflowing (capacity %d)", b->u.ncap16);
}

// initial lists, forces and kinetic energies
static int mb_prime(mtmd_batch* b)
{
    MbDev& d = b->d;
    if (b->run) {
        int rc = mb_build_fit(b);
        if (rc) return rc;
        hipLaunchKernelGGL(k_mb_reduce, dim3(b->nrep), dim3(MB_TPB), 0, b->stream, d, 0, 0);
        hipLaunchKernelGGL(k_mb_ke_slot, dim3(mb_blk(b->nrep, 64)), dim3(64), 0, b->stream, d);
        MB_HIP(b, hipGetLastError());
        if ((rc = mb_check(b))) return rc;
        b->have_forces = true;
        return MTMD_OK;
    }
    hipLaunchKernelGGL(k_mb_fill_u32, dim3(mb_blk(b->nrep, 64)), dim3(64), 0, b->stream, d.rflag, 1u, b->nrep);
    mb_rebuild(b, 0, b->nrep, b->stream);
    mb_forces(b, 0, b->nrep, true, b->stream);
    hipLaunchKernelGGL(k_mb_reduce, dim3(b->nrep), dim3(MB_TPB), 0, b->stream, d, 0, 0);
    MB_HIP(b, hipGetLastError());
    int rc = mb_check(b);
    if (rc) return rc;
    b->have_forces = true;
    return MTMD_OK;
}

extern "C" const char* mtmd_batch_last_error(const mtmd_batch* b)
{
    return b ? b->err.c_str() : g_batch_err.c_str();
}

extern "C" int mtmd_batch_create(int n_systems, const mtmd_system* systems, const mtmd_config* cfgs,
                                 mtmd_batch** out)
{
    auto bad = [&](int code, const char* msg) { g_batch_err = msg; return code; };
    if (!out || !systems || !cfgs || n_systems < 1) return bad(MTMD_EINVAL, "invalid arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return bad(MTMD_ENODEV, "no HIP device");
    const mtmd_config& c0 = cfgs[0];
    for (int i = 0; i < n_systems; i++) {
        const mtmd_config& c = cfgs[i];
        if (c.ensemble != c0.ensemble || c.timestep != c0.timestep || c.cutoff_radius != c0.cutoff_radius ||
            c.skin != c0.skin || c.temperature_nh_len != c0.temperature_nh_len ||
            c.temperature_respa != c0.temperature_respa || c.cutoff_force != c0.cutoff_force)
            return bad(MTMD_EINVAL, "replicas must share ensemble, timestep, cutoff, skin, chain length, RESPA "
                                    "steps and force cap");
        if (c.temperature_relax != c0.temperature_relax)
            return bad(MTMD_EINVAL, "replicas must share the thermostat relaxation time");
    }
    if (c0.ensemble == MTMD_NPT) return bad(MTMD_EINVAL, "ensemble npt is not supported");
    if (c0.temperature_nh_len < 0 || c0.temperature_nh_len > MB_MAXM) return bad(MTMD_EINVAL, "chain length 0..8");
    if (!(c0.skin > 0.0) || !(c0.cutoff_radius > 0.0) || !(c0.timestep > 0.0))
        return bad(MTMD_EINVAL, "cutoff, skin and time step must be positive");

    auto b = new mtmd_batch();
    std::unique_ptr<mtmd_batch> guard(b);
    b->nrep = n_systems;
    const float rc = (float)c0.cutoff_radius;
    const float rlist = (float)(c0.cutoff_radius + c0.skin);
    b->rlist = rlist;
    b->p.dt = c0.timestep / MT_CONV_t;
    b->p.Fcut = c0.cutoff_force;
    b->p.rc2 = rc * rc;
    b->p.rlist2 = rlist * rlist;
    b->p.half_skin2 = (float)((c0.skin / 2.0) * (c0.skin / 2.0));
    b->p.x0 = 0.01f;
    b->p.inv_delta = 1.0f / ((rc - 0.01f) / float(MT_NPTS - 1));
    b->p.nresp = c0.temperature_respa;
    {
        const double w0 = 1.0 / (1.0 - std::pow(2.0, 1.0 / 3.0));   // NHChain.cpp:45-52
        b->p.w[0] = w0; b->p.w[1] = 1.0 - 2.0 * w0; b->p.w[2] = w0;
    }
    b->p.nvt = (c0.ensemble == MTMD_NVT && c0.temperature_nh_len > 0) ? 1 : 0;
    const int M = b->p.nvt ? c0.temperature_nh_len : 0;

    // the union of the type sets: every replica must define the same type pairs
    int T = 0;
    for (int i = 0; i < n_systems; i++) T = std::max(T, systems[i].n_types);
    if (T < 1) return bad(MTMD_EINVAL, "no atom types");
    std::vector<float4> nbtab((size_t)T * T * MT_MAXFF * MT_NPTS, make_float4(0, 0, 0, 0));
    std::vector<int> nbcnt((size_t)T * T, 0);
    {
        const mtmd_system& s = systems[0];
        for (int i = 0; i < s.n_nonbonded; i++) {
            const auto& e = s.nonbonded[i];
            if (e.type_a < 0 || e.type_b < 0 || e.type_a >= T || e.type_b >= T)
                return bad(MTMD_EINVAL, "non-bonded entry: type out of range");
            if (e.kind == MTMD_FF_PROFILE && (!e.profile || e.profile[0] != 0.01f || e.profile[MT_NPTS - 1] != rc))
                return bad(MTMD_EINVAL, "non-bonded profile must span 0.01 .. cutoff");
            auto t = mtmd_host_table(0, e.kind, e.p, e.profile, 0.01f, rc);
            if (t.empty()) return bad(MTMD_EINVAL, "non-bonded entry: unknown kind");
            const int pairs[2][2] = {{e.type_a, e.type_b}, {e.type_b, e.type_a}};
            for (int q = 0; q < (e.type_a == e.type_b ? 1 : 2); q++) {
                const int pr = pairs[q][0] * T + pairs[q][1];
                if (nbcnt[pr] >= MT_MAXFF) return bad(MTMD_EINVAL, "more than 5 force fields for a type pair");
                std::copy(t.begin(), t.end(), nbtab.begin() + ((size_t)pr * MT_MAXFF + nbcnt[pr]) * MT_NPTS);
                nbcnt[pr]++;
            }
        }
    }
    std::vector<float4> coef((size_t)T * T * MT_NPTS, make_float4(0, 0, 0, 0));
    for (int pr = 0; pr < T * T; pr++)
        for (int ff = 0; ff < nbcnt[pr]; ff++)
            for (int i = 0; i < MT_NPTS - 1; i++) {
                float af, bf, ae, be;
                mtmd_host_segment(&nbtab[((size_t)pr * MT_MAXFF + ff) * MT_NPTS], i, af, bf, ae, be);
                float4& o = coef[(size_t)pr * MT_NPTS + i];
                o.x += af; o.y += bf; o.z += ae; o.w += be;
            }
    std::vector<float2> coef2(coef.size());
    for (size_t i = 0; i < coef.size(); i++) coef2[i] = make_float2(coef[i].x, coef[i].y);
    if (coef2.size() * sizeof(float2) > 65536) return bad(MTMD_EINVAL, "too many atom types for the batch engine");

    // replicas: boxes, cell grids, atoms
    std::vector<double4> pos, mom;
    std::vector<float4> posf;
    std::vector<int> rid, cell_rid;
    int natoms = 0, ncells = 0;
    double nexp_max = 0.0;
    for (int r = 0; r < n_systems; r++) {
        const mtmd_system& s = systems[r];
        if (s.n_atoms < 1 || !s.pos || !s.mass || !s.type) return bad(MTMD_EINVAL, "replica without atoms");
        if (s.n_bonds || s.n_angles || s.n_dihedrals) return bad(MTMD_EINVAL, "the batch engine takes no bonded terms");
        if (s.n_nonbonded != systems[0].n_nonbonded) return bad(MTMD_EINVAL, "replicas must share the force fields");
        for (int k = 0; k < s.n_atoms; k++) {
            if (s.charge && s.charge[k] != 0.0) return bad(MTMD_EINVAL, "the batch engine takes no charges");
            if (s.mobile && !s.mobile[k]) return bad(MTMD_EINVAL, "the batch engine takes mobile atoms only");
            if (s.type[k] < 0 || s.type[k] >= T) return bad(MTMD_EINVAL, "atom type out of range");
            if (!(s.mass[k] > 0.0)) return bad(MTMD_EINVAL, "atom mass must be positive");
        }
        MbRep rp{};
        for (int a = 0; a < 3; a++) {
            if (!(s.box[a] > 2.0 * rlist)) return bad(MTMD_EBOX, "replica box below twice the list radius");
            rp.L[a] = s.box[a];
            rp.Lf[a] = (float)s.box[a];
            rp.iLf[a] = 1.0f / rp.Lf[a];
            rp.nc[a] = std::max(1, (int)std::floor(s.box[a] / (double)rlist));
        }
        rp.cell0 = ncells;
        rp.ncell = rp.nc[0] * rp.nc[1] * rp.nc[2];
        rp.sk = 1;
        rp.atom0 = natoms;
        rp.n = s.n_atoms;
        rp.T = cfgs[r].temperature / MT_CONV_T;
        rp.g = 3.0 * s.n_atoms;
        b->rep.push_back(rp);
        const double V = s.box[0] * s.box[1] * s.box[2];
        nexp_max = std::max(nexp_max, (double)s.n_atoms / V * 4.0 / 3.0 * M_PI * rlist * rlist * rlist);
        // momenta: given, or uniform draws scaled to T (SimulationBox.cpp:109-133 in spirit, seeded)
        std::vector<double> p3(3 * (size_t)s.n_atoms);
        if (s.mom) {
            std::copy(s.mom, s.mom + 3 * (size_t)s.n_atoms, p3.begin());
        } else {
            uint64_t st = s.seed * 6364136223846793005ull + 1442695040888963407ull;
            auto u = [&]() {
                st = st * 6364136223846793005ull + 1442695040888963407ull;
                return (double)(st >> 11) * (2.0 / 9007199254740992.0) - 1.0;
            };
            double e2 = 0.0;
            for (int k = 0; k < s.n_atoms; k++)
                for (int a = 0; a < 3; a++) { p3[3 * k + a] = u(); e2 += p3[3 * k + a] * p3[3 * k + a] / s.mass[k]; }
            const double sc = e2 > 0.0 ? std::sqrt(3.0 * s.n_atoms * rp.T / e2) : 0.0;
            for (double& v : p3) v *= sc;
        }
        for (int k = 0; k < s.n_atoms; k++) {
            double x[3] = {s.pos[3 * k], s.pos[3 * k + 1], s.pos[3 * k + 2]};
            for (int a = 0; a < 3; a++) x[a] -= s.box[a] * std::floor(x[a] / s.box[a] + 0.5);
            pos.push_back(make_double4(x[0], x[1], x[2], 0.0));
            mom.push_back(make_double4(p3[3 * k], p3[3 * k + 1], p3[3 * k + 2], s.mass[k]));
            posf.push_back(make_float4((float)x[0], (float)x[1], (float)x[2], mb_bits_as_float(s.type[k])));
            rid.push_back(r);
        }
        for (int c = 0; c < rp.ncell; c++) cell_rid.push_back(r);
        natoms += s.n_atoms;
        ncells += rp.ncell;
    }
    if (natoms >= (1 << 30)) return bad(MTMD_EINVAL, "too many atoms");
    b->n = natoms;

    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bad(MTMD_EHIP, "stream creation failed");
    MbDev& d = b->d;
    d.n = natoms; d.nrep = n_systems; d.ncells = ncells; d.T = T; d.M = M;
    d.pos = b->mem.alloc<double4>(natoms); d.mom = b->mem.alloc<double4>(natoms);
    d.frc = b->mem.alloc<double4>(natoms); d.posref = b->mem.alloc<double4>(natoms);
    d.posf = b->mem.alloc<float4>(natoms); d.epot = b->mem.alloc<float>(natoms); d.rid = b->mem.alloc<int>(natoms);
    d.cell_key = b->mem.alloc<int>(natoms); d.cell_list = b->mem.alloc<int>(natoms);
    d.cell_count = b->mem.alloc<int>(ncells); d.cell_first = b->mem.alloc<int>(ncells);
    d.cell_fill = b->mem.alloc<int>(ncells); d.cell_rid = b->mem.alloc<int>(ncells);
    d.ncount = b->mem.alloc<int>(natoms); d.nreq = b->mem.alloc<int>(1);
    d.ncap = ((int)std::ceil(nexp_max * 1.35) + 16 + 3) & ~3;
    d.nlist = b->mem.alloc<uint32_t>((size_t)d.ncap * natoms);
    d.nbcoef2 = b->mem.alloc<float2>(coef2.size()); d.nbcoef = b->mem.alloc<float4>(coef.size());
    d.rep = b->mem.alloc<MbRep>(n_systems);
    const int stride = 3 * M + 2;
    d.nh = b->mem.alloc<double>((size_t)n_systems * stride);
    d.red = b->mem.alloc<double>(3 * (size_t)n_systems);
    d.rflag = b->mem.alloc<unsigned int>(n_systems);
    d.flags = b->mem.alloc<unsigned int>(4);
    if (!d.pos || !d.mom || !d.frc || !d.posref || !d.posf || !d.epot || !d.rid || !d.cell_key || !d.cell_list ||
        !d.cell_count || !d.cell_first || !d.cell_fill || !d.cell_rid || !d.ncount || !d.nreq || !d.nlist ||
        !d.nbcoef2 || !d.nbcoef || !d.rep || !d.nh || !d.red || !d.rflag || !d.flags)
        return bad(MTMD_ENOMEM, "device allocation failed");
    std::vector<double> nh((size_t)n_systems * stride, 0.0);
    for (int r = 0; r < n_systems; r++) {
        // chain at rest, masses of CNHChain (NHChain.cpp:115-143): Q0 = g T tau^2, Qj = T tau^2
        const double tau = c0.temperature_relax / MT_CONV_t;
        double* q = nh.data() + (size_t)r * stride + 2 * M;
        for (int j = 0; j < M; j++) q[j] = (j == 0 ? b->rep[r].g : 1.0) * b->rep[r].T * tau * tau;
        nh[(size_t)r * stride + 3 * M] = 1.0;
    }
    auto up = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    bool ok = up(d.pos, pos.data(), natoms * sizeof(double4)) && up(d.posref, pos.data(), natoms * sizeof(double4)) &&
              up(d.mom, mom.data(), natoms * sizeof(double4)) && up(d.posf, posf.data(), natoms * sizeof(float4)) &&
              up(d.rid, rid.data(), natoms * sizeof(int)) && up(d.cell_rid, cell_rid.data(), ncells * sizeof(int)) &&
              up(d.nbcoef2, coef2.data(), coef2.size() * sizeof(float2)) &&
              up(d.nbcoef, coef.data(), coef.size() * sizeof(float4)) &&
              up(d.rep, b->rep.data(), n_systems * sizeof(MbRep)) && up(d.nh, nh.data(), nh.size() * sizeof(double)) &&
              hipMemset(d.flags, 0, 4 * sizeof(unsigned int)) == hipSuccess &&
              hipMemset(d.nreq, 0, sizeof(int)) == hipSuccess &&
              hipMemset(d.frc, 0, natoms * sizeof(double4)) == hipSuccess &&
              hipMemset(d.epot, 0, natoms * sizeof(float)) == hipSuccess;
    if (!ok) return bad(MTMD_EHIP, "upload failed");
    // the persistent replica kernel when every replica fits its staging
    {
        int maxn = 0, maxcells = 0;
        for (const MbRep& rp : b->rep) { maxn = std::max(maxn, rp.n); maxcells = std::max(maxcells, rp.ncell); }
        b->u.maxn = maxn;
        b->u.maxcells = maxcells;
        b->run = maxn <= MB_RUN_MAXN && maxcells <= MB_RUN_MAXCELLS && T <= 2 &&
                 mb_run_lds(b, MB_RUN_TPB) <= 160 * 1024 && std::getenv("MTMD_BATCH_KERNELS") == nullptr;
    }
    if (b->run) {
        MbRun& u = b->u;
        const char* ls = std::getenv("MTMD_BATCH_LIST_SLACK");
        b->list_slack = !(ls && std::strcmp(ls, "0") == 0);
        u.ncap16 = b->list_slack ? (((int)std::ceil(nexp_max * 1.35) + 16 + 7) & ~7) : 16;
        u.nlimit = u.ncap16;
        b->snap_pos = b->mem.alloc<double4>(natoms);
        b->snap_mom = b->mem.alloc<double4>(natoms);
        b->snap_frc = b->mem.alloc<double4>(natoms);
        b->snap_nh = b->mem.alloc<double>((size_t)n_systems * stride);
        if (!b->snap_pos || !b->snap_mom || !b->snap_frc || !b->snap_nh)
            return bad(MTMD_ENOMEM, "device allocation failed");
        u.perm = b->mem.alloc<int>(natoms);
        u.perm2 = b->mem.alloc<int>(natoms);
        u.ncount16 = b->mem.alloc<int>(natoms);
        u.nl16 = b->mem.alloc<uint32_t>((size_t)(u.ncap16 / 2) * natoms);
        b->order_batch = b->mem.alloc<int>(n_systems);
        b->order_ident = b->mem.alloc<int>(n_systems);
        if (!u.perm || !u.perm2 || !u.ncount16 || !u.nl16 || !b->order_batch || !b->order_ident)
            return bad(MTMD_ENOMEM, "device allocation failed");
        std::vector<int> ident(natoms);
        for (int i = 0; i < natoms; i++) ident[i] = i;
        // whole-batch order: the larger half of the replicas by size, largest
        // first, then the smaller half, smallest first; the two workgroups a CU
        // runs together then carry about the average size between them
        std::vector<int> bysize(n_systems), ord(n_systems), id(n_systems);
        for (int i = 0; i < n_systems; i++) bysize[i] = id[i] = i;
        std::stable_sort(bysize.begin(), bysize.end(), [&](int x, int y) { return b->rep[x].n > b->rep[y].n; });
        const int half = (n_systems + 1) / 2;
        for (int i = 0; i < n_systems; i++) ord[i] = i < half ? bysize[i] : bysize[n_systems - 1 - (i - half)];
        ok = up(u.perm, ident.data(), natoms * sizeof(int)) && up(b->order_batch, ord.data(), n_systems * sizeof(int)) &&
             up(b->order_ident, id.data(), n_systems * sizeof(int));
        if (!ok) return bad(MTMD_EHIP, "upload failed");
        // a cell shape chosen before any autotune (MTMD_BATCH_CELL_SPLIT=2)
        const char* sp = std::getenv("MTMD_BATCH_CELL_SPLIT");
        if (sp && std::atoi(sp) == 2 && !mb_set_split(b, 2, rlist)) return bad(MTMD_EHIP, "cell grid upload failed");
    }
    int rc0 = mb_prime(b);
    if (rc0) { g_batch_err = b->err; return rc0; }
    *out = guard.release();
    return MTMD_OK;
}

extern "C" void mtmd_batch_destroy(mtmd_batch* b) { delete b; }

// nsteps of every replica: the persistent kernel, MB_RUN_CHUNK steps a
// launch; or (graph path) replays of one captured step graph, the last step
// eagerly with energies
extern "C" int mtmd_batch_step(mtmd_batch* b, int32_t nsteps)
{
    if (!b || nsteps < 0) return MTMD_EINVAL;
    if (nsteps == 0) return MTMD_OK;
    if (b->run) {
        // every step of every replica in one launch a chunk of up to
        // MB_RUN_CHUNK steps (a list growth runs one chunk again), energies
        // in the last step of the call
        for (int t = 0; t < nsteps; t += MB_RUN_CHUNK) {
            const int k = std::min(MB_RUN_CHUNK, nsteps - t), last = t + k == nsteps;
            const int rc = mb_run_call(b, [&]() { return mb_run_launch(b, b->order_batch, b->nrep, k, 1, last, b->stream); });
            if (rc) return rc;
            b->steps += k;
        }
        return MTMD_OK;
    }
    if (!b->graph_exec) {
        MB_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        mb_step(b, 0, b->nrep, false, b->stream);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (e != hipSuccess) return b->fail(MTMD_EHIP, "step capture failed: %s", hipGetErrorString(e));
        b->graph = g;
        MB_HIP(b, hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
    }
    for (int t = 0; t < nsteps - 1; t++) MB_HIP(b, hipGraphLaunch(b->graph_exec, b->stream));
    mb_step(b, 0, b->nrep, true, b->stream);
    MB_HIP(b, hipGetLastError());
    b->steps += nsteps;
    return mb_check(b);
}

// the same steps with every kernel launched once per replica, in turn
extern "C" int mtmd_batch_step_separate(mtmd_batch* b, int32_t nsteps)
{
    if (!b || nsteps < 0) return MTMD_EINVAL;
    if (b->run) {
        // one launch a replica a step, as n separate engines would run
        const int rc = mb_run_call(b, [&]() -> hipError_t {
            for (int t = 0; t < nsteps; t++)
                for (int r = 0; r < b->nrep; r++) {
                    const hipError_t e = mb_run_launch(b, b->order_ident + r, 1, 1, 1, t == nsteps - 1, b->stream);
                    if (e != hipSuccess) return e;
                }
            return hipSuccess;
        });
        if (!rc) b->steps += nsteps;
        return rc;
    }
    for (int t = 0; t < nsteps; t++)
        for (int r = 0; r < b->nrep; r++) mb_step(b, r, r + 1, t == nsteps - 1, b->stream);
    MB_HIP(b, hipGetLastError());
    b->steps += nsteps;
    return mb_check(b);
}

extern "C" int mtmd_batch_synchronize(mtmd_batch* b)
{
    if (!b) return MTMD_EINVAL;
    MB_HIP(b, hipStreamSynchronize(b->stream));
    return MTMD_OK;
}

extern "C" int mtmd_batch_info(const mtmd_batch* b, int32_t* n_systems, int32_t* n_atoms, int32_t* force_block,
                               int32_t* list_block)
{
    if (!b) return MTMD_EINVAL;
    if (n_systems) *n_systems = b->nrep;
    if (n_atoms) *n_atoms = b->n;
    if (force_block) *force_block = b->run ? b->run_block : b->force_block;
    if (list_block) *list_block = b->run ? b->run_g : b->list_block;
    return MTMD_OK;
}

extern "C" int mtmd_batch_tuning(const mtmd_batch* b, int32_t* persistent, int32_t* block, int32_t* lanes,
                                 int32_t* cell_split)
{
    if (!b) return MTMD_EINVAL;
    if (persistent) *persistent = b->run ? 1 : 0;
    if (block) *block = b->run ? b->run_block : b->force_block;
    if (lanes) *lanes = b->run ? b->run_g : 1;
    if (cell_split) *cell_split = b->run ? b->run_split : 1;
    return MTMD_OK;
}

// thermo of every replica (out[n_systems]) on the current state and the
// forces and energies of the last step
extern "C" int mtmd_batch_get_thermo(mtmd_batch* b, mtmd_thermo* out)
{
    if (!b || !out) return MTMD_EINVAL;
    hipLaunchKernelGGL(k_mb_reduce, dim3(b->nrep), dim3(MB_TPB), 0, b->stream, b->d, 0, 1);
    MB_HIP(b, hipGetLastError());
    std::vector<double> red(3 * (size_t)b->nrep);
    MB_HIP(b, hipMemcpyAsync(red.data(), b->d.red, red.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    MB_HIP(b, hipStreamSynchronize(b->stream));
    for (int r = 0; r < b->nrep; r++) {
        const MbRep& rp = b->rep[r];
        const double V = rp.L[0] * rp.L[1] * rp.L[2];
        const double sp = red[3 * r], sf = red[3 * r + 1], su = red[3 * r + 2];
        mtmd_thermo& t = out[r];
        std::memset(&t, 0, sizeof t);
        t.sum_p2m = sp; t.sum_fr = sf;
        t.T = sp / (3.0 * rp.n) * MT_CONV_T;
        t.P = (1.0 / (3.0 * V)) * (sp + sf) * MT_CONV_PRESS;
        t.V = V; t.U = su; t.K = 0.5 * sp; t.Etot = su + 0.5 * sp;
    }
    return MTMD_OK;
}

// state of replica i (layouts [3 * n_i])
extern "C" int mtmd_batch_get_state(mtmd_batch* b, int32_t i, double* pos, double* mom)
{
    if (!b || i < 0 || i >= b->nrep) return MTMD_EINVAL;
    const MbRep& rp = b->rep[i];
    std::vector<double4> x(rp.n), p(rp.n);
    MB_HIP(b, hipMemcpyAsync(x.data(), b->d.pos + rp.atom0, rp.n * sizeof(double4), hipMemcpyDeviceToHost, b->stream));
    MB_HIP(b, hipMemcpyAsync(p.data(), b->d.mom + rp.atom0, rp.n * sizeof(double4), hipMemcpyDeviceToHost, b->stream));
    MB_HIP(b, hipStreamSynchronize(b->stream));
    for (int k = 0; k < rp.n; k++) {
        if (pos) { pos[3 * k] = x[k].x; pos[3 * k + 1] = x[k].y; pos[3 * k + 2] = x[k].z; }
        if (mom) { mom[3 * k] = p[k].x; mom[3 * k + 1] = p[k].y; mom[3 * k + 2] = p[k].z; }
    }
    return MTMD_OK;
}

// forces and per-atom energies of replica i from the last evaluation (capped
// as the integrator used them); layouts [3 * n_i], [n_i]
extern "C" int mtmd_batch_get_forces(mtmd_batch* b, int32_t i, double* F, float* U)
{
    if (!b || i < 0 || i >= b->nrep) return MTMD_EINVAL;
    const MbRep& rp = b->rep[i];
    std::vector<double4> f(rp.n);
    MB_HIP(b, hipMemcpyAsync(f.data(), b->d.frc + rp.atom0, rp.n * sizeof(double4), hipMemcpyDeviceToHost, b->stream));
    if (U) MB_HIP(b, hipMemcpyAsync(U, b->d.epot + rp.atom0, rp.n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    MB_HIP(b, hipStreamSynchronize(b->stream));
    if (F)
        for (int k = 0; k < rp.n; k++) { F[3 * k] = f[k].x; F[3 * k + 1] = f[k].y; F[3 * k + 2] = f[k].z; }
    return MTMD_OK;
}

// Block-size autotuner: times the pair-force and list-build kernels of the
// whole batch for each candidate workgroup size (HIP events, `reps` launches
// each, lists forced to rebuild) and keeps the fastest. Drops the captured
// step graph, which is re-captured with the new shapes on the next step.
extern "C" int mtmd_batch_autotune(mtmd_batch* b, int reps, double* force_ms, double* list_ms)
{
    if (!b || reps < 1) return MTMD_EINVAL;
    hipEvent_t e0, e1;
    MB_HIP(b, hipEventCreate(&e0));
    MB_HIP(b, hipEventCreate(&e1));
    if (b->run) {
        // the persistent kernel: every candidate runs the same steps from the
        // same state (kept on the device, restored and its lists built again
        // before each), timed with HIP events; the state is restored at the
        // end. The chosen shape's force pass (mode 2) and list build (mode 3)
        // are timed for the report.
        int rc = mb_snapshot(b, false);
        if (rc) return rc;
        const int ns = std::max(8, 8 * reps);
        auto steps_ms = [&](double& ms_out) -> int {
            for (int attempt = 0; attempt < 16; attempt++) {
                int r2 = mb_snapshot(b, true);
                if (r2) return r2;
                MB_HIP(b, mb_run_launch(b, b->order_batch, b->nrep, 0, 3, 0, b->stream));
                MB_HIP(b, hipEventRecord(e0, b->stream));
                MB_HIP(b, mb_run_launch(b, b->order_batch, b->nrep, ns, 1, 0, b->stream));
                MB_HIP(b, hipEventRecord(e1, b->stream));
                MB_HIP(b, hipEventSynchronize(e1));
                int need = 0;
                if ((r2 = mb_overflow(b, need))) return r2;
                if (need) {                       // lists grown: measure again
                    if ((r2 = mb_grow16(b, need))) return r2;
                    continue;
                }
                float ms = 0.f;
                MB_HIP(b, hipEventElapsedTime(&ms, e0, e1));
                ms_out = ms / ns;
                return MTMD_OK;
            }
            return b->fail(MTMD_ENEIGH, "neighbour lists kept overflowing (capacity %d)", b->u.ncap16);
        };
        auto pass_ms = [&](int mode, double& ms_out) -> int {
            MB_HIP(b, mb_run_launch(b, b->order_batch, b->nrep, 0, mode, 0, b->stream));
            MB_HIP(b, hipEventRecord(e0, b->stream));
            for (int i = 0; i < reps; i++) MB_HIP(b, mb_run_launch(b, b->order_batch, b->nrep, 0, mode, 0, b->stream));
            MB_HIP(b, hipEventRecord(e1, b->stream));
            MB_HIP(b, hipEventSynchronize(e1));
            float ms = 0.f;
            MB_HIP(b, hipEventElapsedTime(&ms, e0, e1));
            ms_out = ms / reps;
            return MTMD_OK;
        };
        const int bcand[3] = {256, 512, 1024}, gcand[3] = {1, 2, 4};
        double best = 1e30;
        int bb = b->run_block, bg = b->run_g, bsp = b->run_split;
        for (int split = 1; split <= 2; split++) {
            if (!mb_set_split(b, split, b->rlist)) continue;
            for (int bs : bcand)
                for (int g : gcand) {
                    if (mb_run_lds(b, bs) > 160 * 1024) continue;
                    b->run_block = bs;
                    b->run_g = g;
                    double t = 0.0;
                    if ((rc = steps_ms(t))) return rc;
                    if (std::getenv("MTMD_BATCH_TUNE_LOG"))
                        std::fprintf(stderr, "batch tune: cells/radius %d block %d lanes %d step %.4f ms\n", split, bs,
                                     g, t);
                    if (t < best) { best = t; bb = bs; bg = g; bsp = split; }
                }
        }
        if (!mb_set_split(b, bsp, b->rlist)) return b->fail(MTMD_EHIP, "cell grid upload failed");
        b->run_block = bb;
        b->run_g = bg;
        double tf = 0.0, tb = 0.0;
        if ((rc = mb_snapshot(b, true))) return rc;
        if ((rc = pass_ms(3, tb)) || (rc = pass_ms(2, tf))) return rc;
        if (force_ms) *force_ms = tf;
        if (list_ms) *list_ms = tb;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        // the starting state with its lists, forces and energies
        if ((rc = mb_snapshot(b, true))) return rc;
        return mb_build_fit(b);
    }
    auto time_it = [&](auto&& launch, double& ms_out) -> int {
        launch();   // warm
        MB_HIP(b, hipEventRecord(e0, b->stream));
        for (int i = 0; i < reps; i++) launch();
        MB_HIP(b, hipEventRecord(e1, b->stream));
        MB_HIP(b, hipEventSynchronize(e1));
        float ms = 0.f;
        MB_HIP(b, hipEventElapsedTime(&ms, e0, e1));
        ms_out = ms / reps;
        return MTMD_OK;
    };
    const int fcand[4] = {64, 128, 256, 512}, lcand[3] = {64, 128, 256};
    double best = 1e30, t = 0.0;
    int bestb = b->force_block;
    for (int bs : fcand) {
        b->force_block = bs;
        int rc = time_it([&]() { mb_forces(b, 0, b->nrep, false, b->stream); }, t);
        if (rc) return rc;
        if (t < best) { best = t; bestb = bs; }
    }
    b->force_block = bestb;
    if (force_ms) *force_ms = best;
    best = 1e30;
    bestb = b->list_block;
    for (int bs : lcand) {
        b->list_block = bs;
        int rc = time_it([&]() {
            hipLaunchKernelGGL(k_mb_fill_u32, dim3(mb_blk(b->nrep, 64)), dim3(64), 0, b->stream, b->d.rflag, 1u,
                               b->nrep);
            mb_rebuild(b, 0, b->nrep, b->stream);
        }, t);
        if (rc) return rc;
        if (t < best) { best = t; bestb = bs; }
    }
    b->list_block = bestb;
    if (list_ms) *list_ms = best;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (b->graph_exec) { (void)hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    // the timed force passes ran without energies: refresh them for the thermo
    mb_forces(b, 0, b->nrep, true, b->stream);
    MB_HIP(b, hipGetLastError());
    return mb_check(b);
}
