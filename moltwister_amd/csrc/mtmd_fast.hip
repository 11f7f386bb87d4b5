// mtmd_fast.hip - FP64 and MIXED engines for gfx950: spatially sorted atoms,
// neighbour lists at cutoff + skin rebuilt on displacement, pair forces
// gathered per atom over column-major lists (no atomics, deterministic).
//
// Per step the conditional rebuild kernels read flags[0] (the largest squared
// displacement since the last rebuild, written by the integrator's front
// half) and return at once when it is below (skin/2)^2, so a step never waits
// on the host. The pair-force kernel, which runs after the last of them,
// resets flags[0] for the next step's front half (MT_RESET_DISP).
//
// MIXED keeps an FP32 copy of every position relative to the centre of the
// atom's home cell (the cell it sat in at the last rebuild). Each neighbour
// word carries the home-cell offset code of the pair, so the pair vector is
//   d = xl_k - xl_i - offset * cell_size
// in FP32 from two small numbers: no minimum-image branch and no loss of
// precision to the box size. Pair forces and table interpolation run in FP32,
// the four pairs of a list chunk are summed in FP32, and the per-atom sum of
// the chunk sums, integration and reductions are FP64.
#include "mtmd_internal.h"
#include "mtmd_device.h"
#include <algorithm>
#include <climits>

#define TPB 256

static inline int nblk(int n, int b) { return (n + b - 1) / b; }

// Workgroups of a launch are dispatched round-robin over the 8 XCDs of the
// chip (workgroup b on XCD b % 8), and each XCD has its own L2. Atoms are
// spatially sorted, so giving every XCD one contiguous range of logical blocks
// keeps the positions its workgroups gather (a slab of the box plus a halo)
// inside that XCD's L2. The map is a bijection on [0, nb).
__device__ __forceinline__ int mt_xcd_block(int b, int nb)
{
    const int x = b & 7, q = nb >> 3, r = nb & 7;
    return x * q + min(x, r) + (b >> 3);
}

// cell-offset shift (ox, oy, oz) * cell size of neighbour-word code 0..26
__device__ __forceinline__ void mt_fill_shifts(float4* sh, const MtParams& p)
{
    for (int c = threadIdx.x; c < 27; c += blockDim.x)
        sh[c] = make_float4((float)(c % 3 - 1) * p.cs[0], (float)((c / 3) % 3 - 1) * p.cs[1],
                            (float)(c / 9 - 1) * p.cs[2], 0.f);
}

// The rebuild flag has two slots. The conditional list kernels read
// flags[dslot]; a kernel that moves atoms (a front half) first clears that
// slot, whose readers have all run, and sets flags[dslot ^ 1], which no
// kernel of its launch reads, when an atom has moved more than half the skin;
// the host then flips dslot. So no launch both clears and sets one slot.
#define MT_WRITER_CLEAR(d, p)                              \
    if (blockIdx.x == 0 && threadIdx.x == 0) (d).flags[(p).dslot] = 0u

__device__ __forceinline__ bool mt_skip_rebuild(int cond, const unsigned int* flags, int dslot, float thr)
{
    return cond && !(__uint_as_float(flags[dslot]) > thr);
}

__global__ void k_fast_fill(int* p, int v, int n, int cond, const unsigned int* flags, int dslot, float thr)
{
    if (mt_skip_rebuild(cond, flags, dslot, thr)) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Lanes of a wavefront that hold equal keys in a row form a run; atoms are
// spatially sorted, so runs are long and one atomic per run replaces one per
// atom. Gives the run's head lane and length.
__device__ __forceinline__ void mt_key_runs(int key, int& head_lane, int& len)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1);
    const unsigned long long heads = __ballot(lane == 0 || prev != key);
    head_lane = 63 - __clzll(heads & ((2ull << lane) - 1ull));
    const unsigned long long above = heads & ~((2ull << head_lane) - 1ull);
    len = (above ? __ffsll(above) - 1 : 64) - head_lane;
}

// The rebuild kernels are launched every step and return at once unless a
// rebuild is due; such a launch costs GPU time in proportion to its
// workgroups, so the per-atom and per-cell ones take at most MT_REBUILD_GRID
// workgroups and stride over the rest
#define MT_REBUILD_GRID 1024
static inline int nblk_rebuild(long long n, int b) { return (int)std::min<long long>(nblk(n, b), MT_REBUILD_GRID); }

__global__ void __launch_bounds__(TPB) k_fast_cell_keys(const double4* __restrict__ pos, int n, int Nx, int Ny,
                                                        int Nz, double LX, double LY, double LZ,
                                                        int* __restrict__ key, int* __restrict__ count, int cond,
                                                        unsigned int* flags, int dslot, float thr)
{
    if (mt_skip_rebuild(cond, flags, dslot, thr)) return;
    for (int k0 = blockIdx.x * blockDim.x; k0 < n; k0 += gridDim.x * blockDim.x) {   // uniform over the block
        const int k = k0 + (int)threadIdx.x;
        int c = -1;
        if (k < n) {
            const double4 r = pos[k];
            if (!(isfinite(r.x) && isfinite(r.y) && isfinite(r.z))) atomicOr(&flags[3], 1u);
            int ix = mt_d2i(floor((r.x + LX / 2.0) * (double)Nx / LX));
            int iy = mt_d2i(floor((r.y + LY / 2.0) * (double)Ny / LY));
            int iz = mt_d2i(floor((r.z + LZ / 2.0) * (double)Nz / LZ));
            ix = ix < 0 ? 0 : (ix >= Nx ? Nx - 1 : ix);
            iy = iy < 0 ? 0 : (iy >= Ny ? Ny - 1 : iy);
            iz = iz < 0 ? 0 : (iz >= Nz ? Nz - 1 : iz);
            c = ix + Nx * (iy + Ny * iz);
            key[k] = c;
        }
        int hl, len;
        mt_key_runs(c, hl, len);
        if (c >= 0 && (int)(threadIdx.x & 63) == hl) atomicAdd(&count[c], len);
    }
}

// exclusive scan of the cell counts in blocks of SCAN_B cells: block sums,
// then each block adds the sums of the blocks before it to its own scan
#define SCAN_B 1024
__global__ void __launch_bounds__(SCAN_B) k_fast_scan_sums(const int* __restrict__ count, int* __restrict__ bsum,
                                                           int ncell, int cond, const unsigned int* flags, int dslot, float thr)
{
    if (mt_skip_rebuild(cond, flags, dslot, thr)) return;
    __shared__ int ws[SCAN_B / 64];
    const int c = blockIdx.x * SCAN_B + threadIdx.x;
    int v = c < ncell ? count[c] : 0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int sum = 0;
        for (int w = 0; w < SCAN_B / 64; w++) sum += ws[w];
        bsum[blockIdx.x] = sum;
    }
}

__global__ void __launch_bounds__(SCAN_B) k_fast_scan_apply(const int* __restrict__ count,
                                                            const int* __restrict__ bsum, int* __restrict__ first,
                                                            int* __restrict__ fill, int ncell, int cond,
                                                            const unsigned int* flags, int dslot, float thr)
{
    if (mt_skip_rebuild(cond, flags, dslot, thr)) return;
    __shared__ int ws[SCAN_B / 64];
    __shared__ int s_base;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    // sum of the block sums before this block
    int pre = 0;
    for (int j = t; j < (int)blockIdx.x; j += SCAN_B) pre += bsum[j];
    for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o);
    if (lane == 0) ws[w] = pre;
    __syncthreads();
    if (t == 0) {
        int sum = 0;
        for (int j = 0; j < SCAN_B / 64; j++) sum += ws[j];
        s_base = sum;
    }
    __syncthreads();
    // inclusive scan inside the wavefront, then across the block's wavefronts
    const int c = blockIdx.x * SCAN_B + t;
    const int v = c < ncell ? count[c] : 0;
    int incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
    }
    __syncthreads();
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    int wpre = 0;
    for (int j = 0; j < w; j++) wpre += ws[j];
    if (c < ncell) {
        first[c] = s_base + wpre + incl - v;
        fill[c] = 0;
    }
}

// places the atoms of each cell at cell_first. With `down` set, fill holds the
// cell counts and each run of a cell takes the places below the count it
// finds, so the counts end at zero (the state k_fast_cell_keys counts from);
// otherwise fill starts at zero and counts up.
__global__ void __launch_bounds__(TPB) k_fast_scatter(const int* __restrict__ key, int n,
                                                      const int* __restrict__ first, int* __restrict__ fill,
                                                      int* __restrict__ list, int down, int cond,
                                                      const unsigned int* flags, int dslot, float thr)
{
    if (mt_skip_rebuild(cond, flags, dslot, thr)) return;
    const int lane = threadIdx.x & 63;
    for (int k0 = blockIdx.x * blockDim.x; k0 < n; k0 += gridDim.x * blockDim.x) {   // uniform over the block
        const int k = k0 + (int)threadIdx.x;
        const int c = k < n ? key[k] : -1;
        int hl, len;
        mt_key_runs(c, hl, len);
        int base = 0;
        if (c >= 0 && lane == hl) base = down ? atomicSub(&fill[c], len) - len : atomicAdd(&fill[c], len);
        base = __shfl(base, hl);
        if (c >= 0) list[first[c] + base + (lane - hl)] = k;
    }
}

// the per-atom set-up of a rebuild for atom s, the rank-th atom of cell c in
// cell order (u = first[c] + rank): home cell, reference position for the
// displacement test, its place in cell order, and the MIXED cell-local float
// position (.w: the bits of type * MT_NPTS, the type's offset in a row of
// segment coefficients) and charge at that place
// the cell-local position of atom s in cell c (relative to the cell's centre,
// wrapped), rounded to FP32 as the MIXED engine stages it
__device__ __forceinline__ float3 mt_fast_local(const MtDev& d, const MtParams& p, int s, int c)
{
    const double4 r = d.pos[s];
    const int ix = c % d.nc[0], iy = (c / d.nc[0]) % d.nc[1], iz = c / (d.nc[0] * d.nc[1]);
    const double cx = -p.L[0] / 2.0 + ((double)ix + 0.5) * p.cw[0];
    const double cy = -p.L[1] / 2.0 + ((double)iy + 0.5) * p.cw[1];
    const double cz = -p.L[2] / 2.0 + ((double)iz + 0.5) * p.cw[2];
    double dx = r.x - cx, dy = r.y - cy, dz = r.z - cz;
    dx -= p.L[0] * rint(dx * p.invL[0]);
    dy -= p.L[1] * rint(dy * p.invL[1]);
    dz -= p.L[2] * rint(dz * p.invL[2]);
    return make_float3((float)dx, (float)dy, (float)dz);
}

__device__ __forceinline__ void mt_fast_home_atom(const MtDev& d, const MtParams& p, int s, int c, int u)
{
    d.home[s] = c;
    d.posref[s] = d.pos[s];
    d.sidx[s] = u;
    d.qfs[u] = d.qf[s];
    const float3 l = mt_fast_local(d, p, s, c);
    d.posf[u] = make_float4(l.x, l.y, l.z, __int_as_float(d.type[s] * MT_NPTS));
}

// ascending slot order inside each cell, so lists are reproducible run to run;
// with `home` set the per-atom set-up of every atom at its sorted place
// follows (mt_fast_home_atom). In the MIXED tile mode (d.tile16) the atoms of
// a cell are ordered by their cell-local FP32 x instead (slot as the
// tie-break), so the three cells of an x row of staged cells form a run whose
// staged x never decreases inside a cell; the 16-bit list build uses that to
// skip the ends of the row that no pair of its screen can reach.
//
// Each wavefront takes a contiguous range of cells, whose atoms are
// contiguous in cell_list, in chunks of whole cells of at most MT_SORT_CHUNK
// atoms. A chunk's counts, list entries and keys are each fetched with every
// load of the kind in flight at once (one wait on memory per kind, not one a
// cell), the keys go to this wavefront's share of LDS, and an atom's place in
// its cell is the number of smaller keys there. The places are distinct
// because the keys are (slots are distinct). A cell of more than
// MT_SORT_CHUNK atoms is sorted by lane 0 (insertion). No workgroup barrier:
// the wavefronts of a workgroup work independently.
#define MT_SORT_CHUNK 512
#define MT_SORT_PER (MT_SORT_CHUNK / 64)

// the sort key of slot v in cell c, and its cell-local position when xk
__device__ __forceinline__ unsigned long long mt_sort_key(const MtDev& d, const MtParams& p, bool xk, int v, int c,
                                                          float3& l)
{
    if (!xk) return (unsigned long long)(unsigned)v;
    l = mt_fast_local(d, p, v, c);
    uint32_t u = __float_as_uint(l.x);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);          // float order as unsigned order
    return ((unsigned long long)u << 32) | (uint32_t)v;
}

__device__ __forceinline__ void mt_sort_set_up(const MtDev& d, const MtParams& p, bool xk, int v, int c, int u,
                                               const float3& l)
{
    if (!xk) {
        mt_fast_home_atom(d, p, v, c, u);
        return;
    }
    d.home[v] = c;
    d.posref[v] = d.pos[v];
    d.sidx[v] = u;
    d.qfs[u] = d.qf[v];
    d.posf[u] = make_float4(l.x, l.y, l.z, __int_as_float(d.type[v] * MT_NPTS));
}

__global__ void __launch_bounds__(TPB) k_fast_cell_sort(MtDev d, MtParams p, int cond, int home)
{
    __shared__ unsigned long long s_key[TPB / 64][MT_SORT_CHUNK];
    __shared__ unsigned char s_cell[TPB / 64][MT_SORT_CHUNK];   // chunk place -> cell of the chunk (0..63)
    __shared__ int s_off[TPB / 64][65];                           // cell of the chunk -> its first chunk place
    if (mt_skip_rebuild(cond, d.flags, p.dslot, p.half_skin2) || d.flags[3]) return;   // uniform over the grid
    if (home && blockIdx.x == 0 && threadIdx.x == 0) d.flags[4] += 1;        // rebuild count
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long long nwave = (long long)gridDim.x * (blockDim.x >> 6);
    const long long gw = (long long)blockIdx.x * (blockDim.x >> 6) + wv;
    const int cb = (int)(d.ncells * gw / nwave), ce = (int)(d.ncells * (gw + 1) / nwave);
    const bool xk = home && d.tile16;
    unsigned long long* sk = s_key[wv];
    unsigned char* sc = s_cell[wv];
    int* so = s_off[wv];
    for (int cur = cb; cur < ce;) {                   // uniform over the wavefront
        const int cj = cur + lane;
        const int cnt = cj < ce ? d.cell_count[cj] : 0;
        const int fst = cj < ce ? d.cell_first[cj] : 0;
        int incl = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o);
            if (lane >= o) incl += u;
        }
        // the cells of this chunk: a prefix of the lanes (incl never decreases)
        const int ncj = __popcll(__ballot(cj < ce && incl <= MT_SORT_CHUNK));
        const int f0 = __shfl(fst, 0);
        if (ncj == 0) {
            // the first cell alone holds more than a chunk
            const int n0 = __shfl(cnt, 0);
            if (lane == 0) {
                int* a = d.cell_list + f0;
                float3 l;
                for (int i = 1; i < n0; i++) {
                    const int v = a[i];
                    const unsigned long long kv = mt_sort_key(d, p, xk, v, cur, l);
                    int j = i - 1;
                    while (j >= 0 && mt_sort_key(d, p, xk, a[j], cur, l) > kv) { a[j + 1] = a[j]; j--; }
                    a[j + 1] = v;
                }
                if (home)
                    for (int i = 0; i < n0; i++) {
                        const int v = a[i];
                        mt_sort_key(d, p, xk, v, cur, l);
                        mt_sort_set_up(d, p, xk, v, cur, f0 + i, l);
                    }
            }
            cur += 1;
            continue;
        }
        const int A = __shfl(incl, ncj - 1);           // atoms of the chunk
        const int excl = incl - cnt;
        // place -> cell: zeros, each non-empty cell's index at its first
        // place, then a running maximum over the places
        for (int m = 0; m < MT_SORT_PER; m++)
            if (lane + 64 * m < A) sc[lane + 64 * m] = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < ncj) {
            so[lane] = excl;
            if (cnt > 0) sc[excl] = (unsigned char)lane;
        }
        if (lane == 0) so[ncj] = A;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        int run[MT_SORT_PER], lmax = 0;
#pragma unroll
        for (int m = 0; m < MT_SORT_PER; m++) {
            const int i = MT_SORT_PER * lane + m;
            lmax = max(lmax, i < A ? (int)sc[i] : 0);
            run[m] = lmax;
        }
        int pmax = lmax;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(pmax, o);
            if (lane >= o) pmax = max(pmax, u);
        }
        pmax = __shfl_up(pmax, 1);
        if (lane == 0) pmax = 0;
#pragma unroll
        for (int m = 0; m < MT_SORT_PER; m++) {
            const int i = MT_SORT_PER * lane + m;
            if (i < A) sc[i] = (unsigned char)max(pmax, run[m]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the chunk's slots and keys (place lane + 64 m), every load in flight
        int vs[MT_SORT_PER], js[MT_SORT_PER];
        unsigned long long ks[MT_SORT_PER];
        float3 ls[MT_SORT_PER];
#pragma unroll
        for (int m = 0; m < MT_SORT_PER; m++) {
            const int i = lane + 64 * m;
            vs[m] = i < A ? d.cell_list[f0 + i] : 0;
            js[m] = i < A ? (int)sc[i] : 0;
        }
#pragma unroll
        for (int m = 0; m < MT_SORT_PER; m++) {
            const int i = lane + 64 * m;
            ls[m] = make_float3(0.f, 0.f, 0.f);
            if (i < A) {
                ks[m] = mt_sort_key(d, p, xk, vs[m], cur + js[m], ls[m]);
                sk[i] = ks[m];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // place in the cell: the number of smaller keys of the cell; every
        // read of cell_list above has completed (its values are used), so
        // the writes below cannot overtake it
#pragma unroll
        for (int m = 0; m < MT_SORT_PER; m++) {
            const int i = lane + 64 * m;
            if (i < A) {
                const int o0 = so[js[m]], o1 = so[js[m] + 1];
                int rank = 0, t = o0;
                for (; t + 4 <= o1; t += 4) {         // four reads in flight
                    const unsigned long long k0 = sk[t], k1 = sk[t + 1], k2 = sk[t + 2], k3 = sk[t + 3];
                    rank += (k0 < ks[m] ? 1 : 0) + (k1 < ks[m] ? 1 : 0) + (k2 < ks[m] ? 1 : 0) + (k3 < ks[m] ? 1 : 0);
                }
                for (; t < o1; t++) rank += sk[t] < ks[m] ? 1 : 0;
                const int u = f0 + o0 + rank;
                d.cell_list[u] = vs[m];
                if (home) mt_sort_set_up(d, p, xk, vs[m], cur + js[m], u, ls[m]);
            }
        }
        // this chunk's LDS is rewritten by the next one
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        cur += ncj;
    }
}

// The exclusive scan of the counts of a rebuild in one workgroup (grids of at
// most MT_SCAN_ONE_BLOCK cells): k_fast_cell_keys has counted into cell_next,
// which is all zero between rebuilds (k_fast_scatter counts each cell back
// down to zero as it places the cell's atoms). The scan writes cell_first and
// copies the counts to cell_count. Every tile of 4 SCAN_B cells (four
// consecutive cells a thread, one 16-byte load) is loaded at once, so the
// workgroup waits on memory once; each tile's thread sums are scanned across
// the wavefront, and the wavefront totals of all tiles by wavefront 0, one
// barrier before and one after. (A one-pass variant in which the last
// workgroup of the key kernel scans, found by a ticket, needs a device-scope
// fence in every workgroup; with one L2 per XCD that cost 0.38 ms a build on
// the 1M-atom mixture.)
#define MT_SCAN_ONE_BLOCK 65536
#define MT_SCAN_ONE_TILES (MT_SCAN_ONE_BLOCK / (4 * SCAN_B))
__global__ void __launch_bounds__(SCAN_B) k_fast_scan_one(MtDev d, int cond, int dslot, float thr)
{
    if (mt_skip_rebuild(cond, d.flags, dslot, thr)) return;
    constexpr int NW = SCAN_B / 64;
    __shared__ int s_w[MT_SCAN_ONE_TILES * NW];       // wavefront totals, then their exclusive prefixes
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nc = d.ncells;
    const int ntile = (nc + 4 * SCAN_B - 1) / (4 * SCAN_B);
    int4 v[MT_SCAN_ONE_TILES];
    int ex[MT_SCAN_ONE_TILES];
#pragma unroll
    for (int t = 0; t < MT_SCAN_ONE_TILES; t++) {
        const int c = 4 * (t * SCAN_B + (int)threadIdx.x);
        v[t] = make_int4(0, 0, 0, 0);
        if (c + 3 < nc) {
            v[t] = *reinterpret_cast<const int4*>(d.cell_next + c);   // c is a multiple of 4: 16-byte aligned
        } else if (c < nc) {
            v[t].x = d.cell_next[c];
            if (c + 1 < nc) v[t].y = d.cell_next[c + 1];
            if (c + 2 < nc) v[t].z = d.cell_next[c + 2];
        }
    }
#pragma unroll
    for (int t = 0; t < MT_SCAN_ONE_TILES; t++) {
        // (tiles past the last hold zeros: scanned, not stored)
        const int sum = v[t].x + v[t].y + v[t].z + v[t].w;
        int incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o);
            if (lane >= o) incl += u;
        }
        ex[t] = incl - sum;
        if (lane == 63 && t < ntile) s_w[t * NW + w] = incl;
    }
    __syncthreads();
    if (w == 0) {
        // exclusive scan of the ntile * NW totals in tile, then wavefront order:
        // lane j takes entries 4j .. 4j + 3
        const int m = ntile * NW;
        int e[4], tot = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int j = 4 * lane + i;
            e[i] = j < m ? s_w[j] : 0;
            tot += e[i];
        }
        int incl = tot;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o);
            if (lane >= o) incl += u;
        }
        int run = incl - tot;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int j = 4 * lane + i;
            if (j < m) s_w[j] = run;
            run += e[i];
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MT_SCAN_ONE_TILES; t++) {
        const int c = 4 * (t * SCAN_B + (int)threadIdx.x);
        if (c >= nc) continue;
        const int b = s_w[t * NW + w] + ex[t];
        const int4 f = make_int4(b, b + v[t].x, b + v[t].x + v[t].y, b + v[t].x + v[t].y + v[t].z);
        if (c + 3 < nc) {
            *reinterpret_cast<int4*>(d.cell_first + c) = f;
            *reinterpret_cast<int4*>(d.cell_count + c) = v[t];
        } else if (c < nc) {
            d.cell_first[c] = f.x;
            d.cell_count[c] = v[t].x;
            if (c + 1 < nc) { d.cell_first[c + 1] = f.y; d.cell_count[c + 1] = v[t].y; }
            if (c + 2 < nc) { d.cell_first[c + 2] = f.z; d.cell_count[c + 2] = v[t].z; }
        }
    }
}

void mt_fast_cells(const MtDev& d, const MtParams& p, int cond, int with_home, hipStream_t s)
{
    const float thr = p.half_skin2;
    if (d.ncells <= MT_SCAN_ONE_BLOCK) {
        hipLaunchKernelGGL(k_fast_cell_keys, dim3(nblk_rebuild(d.n, TPB)), dim3(TPB), 0, s, d.pos, d.n, d.nc[0],
                           d.nc[1], d.nc[2], p.L[0], p.L[1], p.L[2], d.cell_key, d.cell_next, cond, d.flags, p.dslot,
                           thr);
        hipLaunchKernelGGL(k_fast_scan_one, dim3(1), dim3(SCAN_B), 0, s, d, cond, p.dslot, thr);
        hipLaunchKernelGGL(k_fast_scatter, dim3(nblk_rebuild(d.n, TPB)), dim3(TPB), 0, s, d.cell_key, d.n,
                           d.cell_first, d.cell_next, d.cell_list, 1, cond, d.flags, p.dslot, thr);
        hipLaunchKernelGGL(k_fast_cell_sort, dim3(nblk_rebuild(d.ncells, TPB / 64)), dim3(TPB), 0, s, d, p, cond,
                           with_home);
        return;
    }
    const int nsb = nblk(d.ncells, SCAN_B);
    hipLaunchKernelGGL(k_fast_fill, dim3(nblk(d.ncells, TPB)), dim3(TPB), 0, s, d.cell_count, 0, d.ncells, cond,
                       d.flags, p.dslot, thr);
    hipLaunchKernelGGL(k_fast_cell_keys, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.pos, d.n, d.nc[0], d.nc[1],
                       d.nc[2], p.L[0], p.L[1], p.L[2], d.cell_key, d.cell_count, cond, d.flags, p.dslot, thr);
    hipLaunchKernelGGL(k_fast_scan_sums, dim3(nsb), dim3(SCAN_B), 0, s, d.cell_count, d.cell_bsum, d.ncells, cond,
                       d.flags, p.dslot, thr);
    hipLaunchKernelGGL(k_fast_scan_apply, dim3(nsb), dim3(SCAN_B), 0, s, d.cell_count, d.cell_bsum, d.cell_first,
                       d.cell_fill, d.ncells, cond, d.flags, p.dslot, thr);
    hipLaunchKernelGGL(k_fast_scatter, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.cell_key, d.n, d.cell_first,
                       d.cell_fill, d.cell_list, 0, cond, d.flags, p.dslot, thr);
    hipLaunchKernelGGL(k_fast_cell_sort, dim3(nblk(d.ncells, TPB / 64)), dim3(TPB), 0, s, d, p, cond, with_home);
}

// gather every per-atom array into cell order: new slot s <- old slot list[s]
__global__ void k_fast_sort_apply(MtDev d)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.n) return;
    int o = d.cell_list[s];
    d.pos2[s] = d.pos[o];
    d.mom2[s] = d.mom[o];
    d.frc2[s] = d.frc[o];
    d.epot2[s] = d.epot[o];
    d.type2[s] = d.type[o];
    d.mol2[s] = d.mol[o];
    d.mobile2[s] = d.mobile[o];
    d.qf2[s] = d.qf[o];
    int orig = d.perm[o];
    d.perm2[s] = orig;
    d.inv[orig] = s;
}

void mt_fast_sort_apply(const MtDev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_fast_sort_apply, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d);
}

// scatter back to original order: slot perm[s] <- s
__global__ void k_fast_unsort(MtDev d)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.n) return;
    int o = d.perm[s];
    d.pos2[o] = d.pos[s];
    d.mom2[o] = d.mom[s];
    d.frc2[o] = d.frc[s];
    d.epot2[o] = d.epot[s];
    d.type2[o] = d.type[s];
    d.mol2[o] = d.mol[s];
    d.mobile2[o] = d.mobile[s];
    d.qf2[o] = d.qf[s];
}

void mt_fast_unsort(const MtDev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_fast_unsort, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d);
}

__global__ void k_fast_identity(int* perm, int* inv, int n)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) { perm[s] = s; inv[s] = s; }
}

// positions and momenta in original atom order (sp, sm), and back into the
// current slot order: the state a multi-step call starts from, restored when
// a list build inside the call overflowed
__global__ void __launch_bounds__(TPB) k_fast_snapshot(MtDev d, double4* __restrict__ sp, double4* __restrict__ sm)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.n) return;
    const int o = d.perm[s];
    sp[o] = d.pos[s];
    sm[o] = d.mom[s];
}

__global__ void __launch_bounds__(TPB) k_fast_restore(MtDev d, const double4* __restrict__ sp,
                                                      const double4* __restrict__ sm)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.n) return;
    const int o = d.perm[s];
    d.pos[s] = sp[o];
    d.mom[s] = sm[o];
}

void mt_fast_snapshot(const MtDev& d, double4* sp, double4* sm, hipStream_t s)
{
    hipLaunchKernelGGL(k_fast_snapshot, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d, sp, sm);
}

void mt_fast_restore(const MtDev& d, const double4* sp, const double4* sm, hipStream_t s)
{
    hipLaunchKernelGGL(k_fast_restore, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d, sp, sm);
}

void mt_fast_identity_perm(const MtDev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_fast_identity, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d.perm, d.inv, d.n);
}

// home cell, reference position and local coordinates at a rebuild
// is atom `ia` (slot) a 1-2/1-3/1-4 partner of atom k (slot)? k's rows are in
// original indexing (FunctorCalcForce.cpp:95-145 lists)
__device__ inline bool mt_is_partner(const MtDev& d, int k, int ia)
{
    const MtBonded& b = d.bl;
    int ok = d.perm[k], oi = d.perm[ia];
    if (b.has_bonds)
        for (int j = 0; j < b.b_num[ok]; j++) if (b.b_j[b.b_first[ok] + j] == oi) return true;
    if (b.has_angles)
        for (int j = 0; j < b.a_num[ok]; j++) {
            int g = b.a_first[ok] + j;
            if (!b.a_center[g] && b.a_j[g] == oi) return true;
        }
    if (b.has_dihedrals)
        for (int j = 0; j < b.d_num[ok]; j++) {
            int g = b.d_first[ok] + j;
            if (!b.d_center[g] && b.d_l[g] == oi) return true;
        }
    return false;
}

// Neighbour list at rlist. One 256-thread workgroup builds the lists of a
// 2x2x2 block of home cells, one cell per half-wave (a cell holds ~26 atoms at
// rho* = 1.2 with cells of cutoff + skin). The 4x4x4 cells around the block
// (the union of the eight 27-cell neighbourhoods) are staged in LDS as
// (x, y, z, word index) in cell order; every lane (one atom of a home cell)
// then scans the staged cells of its own neighbourhood with broadcast reads,
// four candidates at a time, and writes its hits as 16-byte chunks of the
// chunked column-major layout. In the MIXED tile mode a block whose staged
// cells fit FT_CAP gets 16-bit words (the neighbour's staged place, eight a
// chunk, mt_neigh_block_atoms16); every other block gets 32-bit words (word
// index | offset code, four a chunk, mt_neigh_block_atoms), its cells staged
// in batches of whole cells when they hold more than NB_CAP atoms.
//
// MODE 0: no molecules. MODE 1: same-molecule pairs are left out (molecule
// ids staged in LDS beside the positions). MODE 2: as 1, and bonded partners
// in other molecules are left out too (a global-memory search of the atom's
// bonded lists). Left-out pairs belong to the special-pair pass. Keeping
// global loads out of the scan loop matters: on this architecture the
// counter that orders loads also counts stores, so a load in the loop would
// make every iteration wait for the previous chunk store.
#define NB_CAP 2048
#define NB_PAD_CODE 13u      // offset code of (0, 0, 0): a pad word is the atom itself
static_assert(NB_CAP == MT_NB_STAGE_ATOMS, "staging capacity of the 32-bit list build");
static_assert(MT_FT_CAP <= NB_CAP && MT_FT_CAP % 8 == 0, "16-bit blocks fit the 32-bit staging");
#define FT_CAP MT_FT_CAP

// slot of a list-word index (the index itself unless the words are in cell order)
__device__ __forceinline__ int mt_id_slot(const MtDev& d, int id) { return d.wsorted ? d.cell_list[id] : id; }

// a candidate inside rlist that is not the atom itself. k: the home atom's
// slot (its list column); e: the candidate's place in the staged cells.
// 32-bit words: the candidate's word index | offset code, four to a chunk.
// W16: e itself, eight to a chunk.
template <int MODE, bool W16>
__device__ __forceinline__ void mt_neigh_hit(const MtDev& d, uint4* nl4, float4 c, uint32_t code, int k, int molk,
                                             const int* smol, int e, int n, int cap, int& cnt, uint32_t& r0,
                                             uint32_t& r1, uint32_t& r2, uint32_t& r3)
{
    const int a = (int)__float_as_uint(c.w);
    if constexpr (MODE >= 1) {
        if (smol[e] == molk) return;
    }
    if constexpr (MODE == 2) {
        if (mt_is_partner(d, k, mt_id_slot(d, a))) return;
    }
    constexpr int lg = W16 ? 3 : 2;                   // log2 of the words a chunk
    mt_neigh_push<W16>(W16 ? (uint32_t)e : ((uint32_t)a | code), r0, r1, r2, r3);
    if ((cnt & ((1 << lg) - 1)) == (1 << lg) - 1 && cnt < cap)
        nl4[(size_t)(cnt >> lg) * n + k] = make_uint4(r0, r1, r2, r3);
    cnt++;
}

// the 32-bit lists of one workgroup's home atoms (the scan below the staging
// set-up of k_fast_neigh_block), staging the 64 cells in batches of whole
// cells that fit SCAP atoms
template <int MODE, int SCAP>
__device__ __forceinline__ void mt_neigh_block_atoms(const MtDev& d, const MtParams& p, float4* spos, int* smol,
                                                     const int* s_first, const int* s_cnt, const int* s_off, int hxo,
                                                     int hyo, int hzo, int l, int f0, int cnt_c, int cnt_max)
{
    const int t = threadIdx.x;
    const int n = d.n;
    const int cap = d.ncap;                           // words a row holds
    const float rl2 = p.rlist2;
    uint4* __restrict__ nl4 = reinterpret_cast<uint4*>(d.nlist);
    int staged_b = -1, staged_e = -1;
    for (int ib = 0; ib < cnt_max; ib += 32) {
        const bool act = ib + l < cnt_c;
        const int uk = f0 + ib + l;                 // place in cell order
        const int k = act ? d.cell_list[uk] : 0;
        const int self = d.wsorted ? uk : k;
        const float4 pk = act ? d.posf[uk] : make_float4(0.f, 0.f, 0.f, 0.f);
        const int molk = (MODE >= 1 && act) ? d.mol[k] : -1;
        // MODE 2: only atoms with a bonded partner outside their molecule search
        const bool xk = MODE == 2 && act && d.xmol[d.perm[k]] != 0;
        int cnt = 0;
        uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
        int cb = 0;
        while (cb < 64) {
            // batch of whole staged cells [cb, ce) that fits SCAP (uniform)
            int ce = cb + 1;
            while (ce < 64 && s_off[ce + 1] - s_off[cb] <= SCAP) ce++;
            const int base = s_off[cb];
            if (s_off[ce] - base > SCAP) {
                // a single cell larger than the staging buffer: this block's
                // lists are not built, and its atoms' counts are zeroed, so no
                // pass reads rows of an earlier build before the host sees
                // the flag (the condition is uniform over the workgroup)
                if (t == 0) d.flags[5] = 1u;
                for (int i = l; i < cnt_c; i += 32) d.ncount[d.cell_list[f0 + i]] = 0;
                return;
            }
            if (cb != staged_b || ce != staged_e) {
                __syncthreads();
                for (int i = base + t; i < s_off[ce]; i += 256) {
                    // staged cell of item i: binary search over s_off[cb..ce]
                    int lo = cb, hi = ce;
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (s_off[mid] <= i) lo = mid; else hi = mid;
                    }
                    const int ui = s_first[lo] + (i - s_off[lo]);
                    const float4 pa = d.posf[ui];
                    const int a = (MODE >= 1 || !d.wsorted) ? d.cell_list[ui] : ui;
                    spos[i - base] = make_float4(pa.x, pa.y, pa.z, __uint_as_float((uint32_t)(d.wsorted ? ui : a)));
                    if constexpr (MODE >= 1) smol[i - base] = d.mol[a];
                }
                __syncthreads();
                staged_b = cb;
                staged_e = ce;
            }
            if (act) {
                for (int oz = -1; oz <= 1; oz++)
                    for (int oy = -1; oy <= 1; oy++)
                        for (int ox = -1; ox <= 1; ox++) {
                            const int q = (ox + 1 + hxo) + 4 * ((oy + 1 + hyo) + 4 * (oz + 1 + hzo));
                            if (q < cb || q >= ce) continue;
                            const uint32_t code = (uint32_t)((ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)) << MT_NW_IDX_BITS;
                            const float qx = pk.x - (float)ox * p.cs[0];
                            const float qy = pk.y - (float)oy * p.cs[1];
                            const float qz = pk.z - (float)oz * p.cs[2];
                            const int e0 = s_off[q] - base, e1 = e0 + s_cnt[q];
                            // four candidates a pass: their LDS reads are issued
                            // together, and the hit test (inside rlist, not the
                            // atom itself) is one predicate per candidate
                            int e = e0;
                            for (; e + 3 < e1; e += 4) {
                                float4 c[4];
#pragma unroll
                                for (int i = 0; i < 4; i++) c[i] = spos[e + i];
                                bool hv[4];
#pragma unroll
                                for (int i = 0; i < 4; i++) {
                                    const float dx = qx - c[i].x, dy = qy - c[i].y, dz = qz - c[i].z;
                                    hv[i] = dx * dx + dy * dy + dz * dz < rl2 && __float_as_int(c[i].w) != self;
                                }
#pragma unroll
                                for (int i = 0; i < 4; i++)
                                    if (hv[i] && (MODE != 2 || !xk ||
                                                  !mt_is_partner(d, k, mt_id_slot(d, __float_as_int(c[i].w)))))
                                        mt_neigh_hit<MODE == 2 ? 1 : MODE, false>(d, nl4, c[i], code, k, molk, smol,
                                                                                  e + i, n, cap, cnt, r0, r1, r2,
                                                                                  r3);
                            }
                            for (; e < e1; e++) {
                                const float4 c0 = spos[e];
                                const float dx = qx - c0.x, dy = qy - c0.y, dz = qz - c0.z;
                                if (dx * dx + dy * dy + dz * dz < rl2 && __float_as_int(c0.w) != self &&
                                    (MODE != 2 || !xk || !mt_is_partner(d, k, mt_id_slot(d, __float_as_int(c0.w)))))
                                    mt_neigh_hit<MODE == 2 ? 1 : MODE, false>(d, nl4, c0, code, k, molk, smol, e, n,
                                                                              cap, cnt, r0, r1, r2, r3);
                            }
                        }
            }
            cb = ce;
        }
        if (act) {
            // last partial chunk, padded with the atom itself (its word with
            // the offset code of (0, 0, 0))
            const int slot = cnt & 3;
            if (slot && cnt < cap) {
                const uint32_t pad = (uint32_t)self | (NB_PAD_CODE << MT_NW_IDX_BITS);
                for (int i = slot; i < 4; i++) mt_neigh_push<false>(pad, r0, r1, r2, r3);
                nl4[(size_t)(cnt >> 2) * n + k] = make_uint4(r0, r1, r2, r3);
            }
            d.ncount[k] = cnt;
            if (cnt > cap) { d.flags[2] = 1; atomicMax(d.nreq, cnt); }
        }
    }
}

// FP16 pair of (x, y), for the 16-bit list build's distance screen
typedef _Float16 mt_h2 __attribute__((ext_vector_type(2)));

// half an FP16 ulp at magnitude v > 0 (2^(exponent(v) - 11))
__device__ __forceinline__ float mt_half_ulp16(float v)
{
    return __uint_as_float((uint32_t)((int)(__float_as_uint(v) >> 23) - 11) << 23);
}

// the 16-bit lists of a block whose 64 staged cells fit FT_CAP: the cells are
// staged once, in the frame of the centre of the block's first cell (as
// k_fast_force_tile stages them: cell offsets times cell widths, exact in
// FP32), so the three cells ox = -1..1 of a (oy, oz) row of a home cell's
// neighbourhood are one contiguous run of staged atoms with no per-cell
// shift: nine runs a home atom. The distance screen takes x and y in FP16
// (one packed subtraction and one dot product with an FP32 z term, five
// instructions a candidate) against rlist^2 widened by a margin that covers
// the FP16 rounding, so a list holds every pair inside rlist and a few just
// beyond it, which the force kernels see as outside rc.
template <int MODE, bool SELF>
__device__ __forceinline__ void mt_neigh_run16(const MtDev& d, uint4* nl4, const uint2* sxz, const int* sword,
                                               const int* smol, int e, int e1, mt_h2 pxy, float pz, float thr,
                                               int lself, int k, int molk, bool xk, int n, int cap, int& cnt,
                                               uint32_t& r0, uint32_t& r1, uint32_t& r2, uint32_t& r3)
{
    const float4 dummy = make_float4(0.f, 0.f, 0.f, 0.f);
    for (; e + 3 < e1; e += 4) {
        uint2 c[4];
#pragma unroll
        for (int i = 0; i < 4; i++) c[i] = sxz[e + i];
        bool hv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const mt_h2 dxy = pxy - __builtin_bit_cast(mt_h2, c[i].x);
            const float dz = pz - __uint_as_float(c[i].y);
            hv[i] = __builtin_amdgcn_fdot2(dxy, dxy, dz * dz, false) < thr && (!SELF || e + i != lself);
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (hv[i] && (MODE != 2 || !xk || !mt_is_partner(d, k, mt_id_slot(d, sword[e + i]))))
                mt_neigh_hit<MODE == 2 ? 1 : MODE, true>(d, nl4, dummy, 0u, k, molk, smol, e + i, n, cap, cnt, r0,
                                                          r1, r2, r3);
    }
    for (; e < e1; e++) {
        const uint2 c0 = sxz[e];
        const mt_h2 dxy = pxy - __builtin_bit_cast(mt_h2, c0.x);
        const float dz = pz - __uint_as_float(c0.y);
        if (__builtin_amdgcn_fdot2(dxy, dxy, dz * dz, false) < thr && (!SELF || e != lself) &&
            (MODE != 2 || !xk || !mt_is_partner(d, k, mt_id_slot(d, sword[e]))))
            mt_neigh_hit<MODE == 2 ? 1 : MODE, true>(d, nl4, dummy, 0u, k, molk, smol, e, n, cap, cnt, r0, r1, r2,
                                                      r3);
    }
}

// the hits of staged places [e, e1) of a run, in order, into the ring (the
// screen of mt_neigh_run16)
template <int MODE, bool SELF>
__device__ __forceinline__ void mt_neigh_run16r(const MtDev& d, const uint2* sxz, const int* sword, const int* smol,
                                                int e, int e1, mt_h2 pxy, float pz, float thr, int lself, int k,
                                                int molk, bool xk, MtRing16& g)
{
    for (; e + 3 < e1; e += 4) {
        uint2 c[4];
#pragma unroll
        for (int i = 0; i < 4; i++) c[i] = sxz[e + i];
        bool hv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const mt_h2 dxy = pxy - __builtin_bit_cast(mt_h2, c[i].x);
            const float dz = pz - __uint_as_float(c[i].y);
            hv[i] = __builtin_amdgcn_fdot2(dxy, dxy, dz * dz, false) < thr && (!SELF || e + i != lself);
            if constexpr (MODE >= 1) hv[i] = hv[i] && smol[e + i] != molk;
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (hv[i] && (MODE != 2 || !xk || !mt_is_partner(d, k, mt_id_slot(d, sword[e + i])))) mt_ring_put(g, e + i);
        mt_ring_flush(g);
    }
    for (; e < e1; e++) {
        const uint2 c0 = sxz[e];
        const mt_h2 dxy = pxy - __builtin_bit_cast(mt_h2, c0.x);
        const float dz = pz - __uint_as_float(c0.y);
        if (__builtin_amdgcn_fdot2(dxy, dxy, dz * dz, false) < thr && (!SELF || e != lself) &&
            (MODE < 1 || smol[e] != molk) &&
            (MODE != 2 || !xk || !mt_is_partner(d, k, mt_id_slot(d, sword[e])))) {
            mt_ring_put(g, e);
            mt_ring_flush(g);
        }
    }
}

// first place e in [lo, hi) of a staged cell whose FP16 x difference to the
// home atom, dx = px - x[e] rounded as the screen rounds it, satisfies
// dx < B (LEFT: the run's ox = -1 cell) or dx <= -B (right: its ox = +1 cell);
// hi when there is none. The cell's staged x never decreases (tile-mode cells
// are ordered by x), so dx never increases and each predicate is false and
// then true along the cell.
template <bool LEFT>
__device__ __forceinline__ int mt_run_bound16(const uint2* sxz, int lo, int hi, mt_h2 pxy, _Float16 B)
{
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const _Float16 dx = (pxy - __builtin_bit_cast(mt_h2, sxz[mid].x)).x;
        const bool t = LEFT ? (dx < B) : (dx <= -B);
        lo = t ? lo : mid + 1;
        hi = t ? mid : hi;
    }
    return lo;
}

template <int MODE>
__device__ __forceinline__ void mt_neigh_block_atoms16(const MtDev& d, const MtParams& p, float4* spos, int* smol,
                                                       int* sword, const int* s_first, const int* s_cnt,
                                                       const int* s_off, const int* s_hfirst, const int* s_hcnt,
                                                       uint32_t* sring)
{
    const int t = threadIdx.x;
    const int n = d.n;
    const int cap = 2 * d.ncap;                       // 16-bit words a row holds
    uint4* __restrict__ nl4 = reinterpret_cast<uint4*>(d.nlist);
    // the staging buffer (FT_CAP / 2 float4s) as FT_CAP (FP16 x | FP16 y,
    // FP32 z) pairs; MODE 2 keeps the word indices in sword
    uint2* sxz = reinterpret_cast<uint2*>(spos);
    // the screen's threshold: staged |x|, |y| <= 2.5 cs (cells at offsets -1..2
    // around the first cell's centre, atoms within half a cell of their
    // cell's centre at a build) round by at most h(2.6 cs) each in FP16, the
    // difference of a pair inside rlist by h(rlist), so |error of dx, dy| <=
    // delta and |error of r^2| <= 2 sqrt(2) rlist delta + 2 delta^2
    const float csm = fmaxf(p.cs[0], fmaxf(p.cs[1], p.cs[2]));
    const float delta = 2.0f * mt_half_ulp16(2.6f * csm) + mt_half_ulp16(p.rlist);
    const float thr = p.rlist2 + 3.0f * p.rlist * delta + 2.0f * delta * delta;
    // B: the least FP16 value with B^2 >= thr. A candidate whose FP16 x
    // difference has |dx| >= B fails the screen whatever dy and dz are (dx^2 is
    // exact in FP32, and the screen's sum, rounded, is at least dx^2 >= thr),
    // so the ends of a run beyond +-B can be skipped without changing a word
    _Float16 B = (_Float16)sqrtf(thr);
    while ((float)B * (float)B < thr) B = __builtin_bit_cast(_Float16, (uint16_t)(__builtin_bit_cast(uint16_t, B) + 1));
    {
        // four threads a staged cell, each taking every fourth atom
        const int q = t >> 2;
        const float sx = (float)((q & 3) - 1) * p.cs[0], sy = (float)(((q >> 2) & 3) - 1) * p.cs[1];
        const float sz = (float)((q >> 4) - 1) * p.cs[2];
        for (int m = t & 3; m < s_cnt[q]; m += 4) {
            const int ui = s_first[q] + m, i = s_off[q] + m;
            const float4 pa = d.posf[ui];
            const mt_h2 xy = {(_Float16)(pa.x + sx), (_Float16)(pa.y + sy)};
            sxz[i] = make_uint2(__builtin_bit_cast(uint32_t, xy), __float_as_uint(pa.z + sz));
            if constexpr (MODE >= 1) {
                const int a = d.cell_list[ui];
                smol[i] = d.mol[a];
                if constexpr (MODE == 2) sword[i] = d.wsorted ? ui : a;
            }
        }
    }
    __syncthreads();
    // the block's home atoms, cell after cell, over all the workgroup's lanes
    // (a lane per atom, as k_fast_force_tile takes them): a home cell with
    // more atoms than a half-wave holds costs no extra pass of the wavefront
    int hoff[9];
    hoff[0] = 0;
#pragma unroll
    for (int h = 0; h < 8; h++) hoff[h + 1] = hoff[h] + s_hcnt[h];
    for (int v = t; v < hoff[8]; v += blockDim.x) {
        int h = 0;
        while (h < 7 && v >= hoff[h + 1]) h++;
        const int m = v - hoff[h];                    // place in the home cell
        const int hxo = h & 1, hyo = (h >> 1) & 1, hzo = h >> 2;
        // the home cell's place among the staged cells
        const int qh = (hxo + 1) + 4 * ((hyo + 1) + 4 * (hzo + 1));
        const int uk = s_hfirst[h] + m;               // place in cell order
        const int k = d.cell_list[uk];
        const int lself = s_off[qh] + m;              // own staged place (the pad word)
        const uint2 hv = sxz[lself];
        const mt_h2 pxy = __builtin_bit_cast(mt_h2, hv.x);
        const float pz = __uint_as_float(hv.y);
        const int molk = MODE >= 1 ? d.mol[k] : -1;
        const bool xk = MODE == 2 && d.xmol[d.perm[k]] != 0;
        MtRing16 g{sring + (size_t)t * MT_RING_DW, nl4 + k, (size_t)n, cap >> 3, 0, 0};
        for (int oz = 0; oz < 3; oz++)
            for (int oy = 0; oy < 3; oy++) {
                const int q = hxo + 4 * ((oy + hyo) + 4 * (oz + hzo));     // the row's ox = -1 cell
                // the run's candidates within +-B in x: the head of the left
                // cell and the tail of the right cell are out of reach
                const int e0 = mt_run_bound16<true>(sxz, s_off[q], s_off[q + 1], pxy, B);
                const int e1 = mt_run_bound16<false>(sxz, s_off[q + 2], s_off[q + 3], pxy, B);
                // only the middle run holds the atom itself
                if (oy == 1 && oz == 1)
                    mt_neigh_run16r<MODE, true>(d, sxz, sword, smol, e0, e1, pxy, pz, thr, lself, k, molk, xk, g);
                else
                    mt_neigh_run16r<MODE, false>(d, sxz, sword, smol, e0, e1, pxy, pz, thr, lself, k, molk, xk, g);
            }
        // last partial chunk, padded with the atom's own staged place
        const int cnt = g.cnt;
        const int slot = cnt & 7;
        if (slot && cnt < cap) {
            for (int i = slot; i < 8; i++) mt_ring_put(g, lself);
            mt_ring_flush(g);
        }
        d.ncount[k] = cnt;
        // the capacity request is in 32-bit words
        if (cnt > cap) { d.flags[2] = 1; atomicMax(d.nreq, (cnt + 1) >> 1); }
    }
}

// TILE: the launch of the MIXED tile mode (d.tile16), whose blocks mostly take
// 16-bit words: FT_CAP FP16/FP32 coordinate pairs of staging (16 KB), twice
// the workgroups a CU of the full 32 KB buffer; the blocks that keep 32-bit
// words then stage their cells in batches of FT_CAP / 2 atoms
template <int MODE, bool TILE>
__global__ void __launch_bounds__(256) k_fast_neigh_block(MtDev d, MtParams p, int cond)
{
    if (mt_skip_rebuild(cond, d.flags, p.dslot, p.half_skin2)) return;
    constexpr int SCAP = TILE ? MT_NB_STAGE_ATOMS_TILE : NB_CAP;
    __shared__ float4 spos[SCAP];
    __shared__ int smol_buf[MODE >= 1 ? NB_CAP : 1];
    __shared__ int sword_buf[(TILE && MODE == 2) ? NB_CAP : 1];
    __shared__ int s_first[64], s_cnt[64], s_off[65];
    __shared__ int s_hfirst[8], s_hcnt[8];
    __shared__ uint32_t sring[TILE ? 256 * MT_RING_DW : 1];   // 16-bit rows: a ring of list words a lane
    const int t = threadIdx.x;
    const int Nx = d.nc[0], Ny = d.nc[1], Nz = d.nc[2];
    const int bx = (Nx + 1) >> 1, by = (Ny + 1) >> 1;
    const int g = mt_xcd_block(blockIdx.x, gridDim.x);
    const int x0 = 2 * (g % bx), y0 = 2 * ((g / bx) % by), z0 = 2 * (g / (bx * by));
    // this half-wave's home cell: (x0 + hxo, y0 + hyo, z0 + hzo)
    const int hc = t >> 5, l = t & 31;
    const int hxo = hc & 1, hyo = (hc >> 1) & 1, hzo = hc >> 2;
    const int hx = x0 + hxo, hy = y0 + hyo, hz = z0 + hzo;
    const bool home_ok = hx < Nx && hy < Ny && hz < Nz;
    if (t < 64) {
        // staged cell q = (ox+1) + 4*((oy+1) + 4*(oz+1)), o* in -1..2 around (x0, y0, z0)
        const int ox = (t & 3) - 1, oy = ((t >> 2) & 3) - 1, oz = (t >> 4) - 1;
        int cx = x0 + ox, cy = y0 + oy, cz = z0 + oz;
        cx = cx < 0 ? cx + Nx : (cx >= Nx ? cx - Nx : cx);
        cy = cy < 0 ? cy + Ny : (cy >= Ny ? cy - Ny : cy);
        cz = cz < 0 ? cz + Nz : (cz >= Nz ? cz - Nz : cz);
        const int cc = cx + Nx * (cy + Ny * cz);
        const int cn = d.cell_count[cc];
        s_first[t] = d.cell_first[cc];
        s_cnt[t] = cn;
        // inclusive scan of the 64 counts across the wavefront
        int v = cn;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(v, o);
            if (t >= o) v += u;
        }
        s_off[t + 1] = v;
        if (t == 0) s_off[0] = 0;
    }
    if ((t & 31) == 0) {
        const int hcell = home_ok ? hx + Nx * (hy + Ny * hz) : 0;
        s_hfirst[hc] = home_ok ? d.cell_first[hcell] : 0;
        s_hcnt[hc] = home_ok ? d.cell_count[hcell] : 0;
    }
    __syncthreads();
    const int f0 = s_hfirst[hc], cnt_c = s_hcnt[hc];
    int cnt_max = 0;
    for (int i = 0; i < 8; i++) cnt_max = max(cnt_max, s_hcnt[i]);
    if (cnt_max == 0) return;
    if (d.flags[3]) {
        for (int i = l; i < cnt_c; i += 32) d.ncount[d.cell_list[f0 + i]] = 0;
        return;
    }
    // the word format of this block's lists (uniform; k_fast_force_tile makes
    // the same choice from the same counts)
    if (TILE && s_off[64] <= FT_CAP)
        mt_neigh_block_atoms16<MODE>(d, p, spos, smol_buf, sword_buf, s_first, s_cnt, s_off, s_hfirst, s_hcnt, sring);
    else
        mt_neigh_block_atoms<MODE, SCAP>(d, p, spos, smol_buf, s_first, s_cnt, s_off, hxo, hyo, hzo, l, f0, cnt_c,
                                         cnt_max);
}

#define NBW_CAP 1024

// Neighbour list for dense cells (tens to hundreds of atoms per cell, e.g.
// water with a 10 A cutoff), with enough workgroups to fill the GPU: a
// workgroup takes a batch of NBD_ATOMS home atoms of one cell (grid: cells x
// batches, each workgroup looping over the batches y, y + gridDim.y, ...),
// NBD_G lanes per home atom. The 27 neighbour cells are staged in LDS one at
// a time (in chunks of NBW_CAP atoms); the NBD_G lanes of an atom test NBD_G
// consecutive candidates per step, and a wavefront ballot gives every hit its
// rank among the atom's hits of that step, so the words land in candidate
// order: the list is the one k_fast_neigh_block's sequential scan would write.
#define NBD_G 16
#define NBD_ATOMS (256 / NBD_G)
template <int MODE>
__global__ void __launch_bounds__(256) k_fast_neigh_dense(MtDev d, MtParams p, int cond)
{
    if (mt_skip_rebuild(cond, d.flags, p.dslot, p.half_skin2)) return;
    __shared__ float4 spos[NBW_CAP];
    __shared__ int smol_buf[MODE >= 1 ? NBW_CAP : 1];
    const int t = threadIdx.x, lane = t & 63;
    const int a = t / NBD_G, g = t % NBD_G;
    const int Nx = d.nc[0], Ny = d.nc[1], Nz = d.nc[2];
    const int c = blockIdx.x;
    const int hx = c % Nx, hy = (c / Nx) % Ny, hz = c / (Nx * Ny);
    const int f0 = d.cell_first[c], cnt_c = d.cell_count[c];
    if (d.flags[3]) {
        if (blockIdx.y == 0)
            for (int i = t; i < cnt_c; i += 256) d.ncount[d.cell_list[f0 + i]] = 0;
        return;
    }
    const int n = d.n, ncap = d.ncap;
    const float rl2 = p.rlist2;
    uint32_t* __restrict__ nl = d.nlist;
    // the lanes of this atom's group within the wavefront, and those below g
    const int gbase = lane & ~(NBD_G - 1);
    const uint64_t below = (1ull << g) - 1ull;
    for (int hb = (int)blockIdx.y * NBD_ATOMS; hb < cnt_c; hb += (int)gridDim.y * NBD_ATOMS) {
        const bool act = hb + a < cnt_c;
        const int uk = f0 + hb + a;
        const int k = act ? d.cell_list[uk] : 0;
        const int self = d.wsorted ? uk : k;
        const float4 pk = act ? d.posf[uk] : make_float4(0.f, 0.f, 0.f, 0.f);
        const int molk = (MODE >= 1 && act) ? d.mol[k] : -1;
        const bool xk = MODE == 2 && act && d.xmol[d.perm[k]] != 0;
        int cnt = 0;
        for (int oz = -1; oz <= 1; oz++) {
            int cz = hz + oz; cz = cz < 0 ? cz + Nz : (cz >= Nz ? cz - Nz : cz);
            for (int oy = -1; oy <= 1; oy++) {
                int cy = hy + oy; cy = cy < 0 ? cy + Ny : (cy >= Ny ? cy - Ny : cy);
                for (int ox = -1; ox <= 1; ox++) {
                    int cx = hx + ox; cx = cx < 0 ? cx + Nx : (cx >= Nx ? cx - Nx : cx);
                    const int cc = cx + Nx * (cy + Ny * cz);
                    const uint32_t code = (uint32_t)((ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)) << MT_NW_IDX_BITS;
                    const float qx = pk.x - (float)ox * p.cs[0];
                    const float qy = pk.y - (float)oy * p.cs[1];
                    const float qz = pk.z - (float)oz * p.cs[2];
                    const int g0 = d.cell_first[cc], gn = d.cell_count[cc];
                    for (int jb = 0; jb < gn; jb += NBW_CAP) {
                        const int m = min(NBW_CAP, gn - jb);
                        __syncthreads();
                        for (int e = t; e < m; e += 256) {
                            const int ui = g0 + jb + e;
                            const float4 pa = d.posf[ui];
                            const int at = (MODE >= 1 || !d.wsorted) ? d.cell_list[ui] : ui;
                            spos[e] = make_float4(pa.x, pa.y, pa.z, __uint_as_float((uint32_t)(d.wsorted ? ui : at)));
                            if constexpr (MODE >= 1) smol_buf[e] = d.mol[at];
                        }
                        __syncthreads();
                        for (int e0 = 0; e0 < m; e0 += NBD_G) {
                            const int e = e0 + g;
                            bool hit = false;
                            uint32_t id = 0u;
                            if (act && e < m) {
                                const float4 q4 = spos[e];
                                const float dx = qx - q4.x, dy = qy - q4.y, dz = qz - q4.z;
                                id = __float_as_uint(q4.w);
                                hit = dx * dx + dy * dy + dz * dz < rl2 && (int)id != self;
                                if constexpr (MODE >= 1) hit = hit && smol_buf[e] != molk;
                                if constexpr (MODE == 2)
                                    hit = hit && !(xk && mt_is_partner(d, k, mt_id_slot(d, (int)id)));
                            }
                            const uint64_t bal = __ballot(hit);
                            const uint64_t grp = (bal >> gbase) & ((NBD_G == 64) ? ~0ull : ((1ull << NBD_G) - 1ull));
                            if (hit) {
                                const int j = cnt + __popcll(grp & below);
                                if (j < ncap) nl[(((size_t)(j >> 2) * n + k) << 2) + (j & 3)] = id | code;
                            }
                            cnt += __popcll(grp);
                        }
                    }
                }
            }
        }
        if (act) {
            // the last partial chunk, padded with the atom itself (offset code 0,0,0)
            const int j = cnt + g;
            if (g < ((4 - (cnt & 3)) & 3) && j < ncap)
                nl[(((size_t)(j >> 2) * n + k) << 2) + (j & 3)] = (uint32_t)self | (NB_PAD_CODE << MT_NW_IDX_BITS);
            if (g == 0) {
                d.ncount[k] = cnt;
                if (cnt > ncap) { d.flags[2] = 1; atomicMax(d.nreq, cnt); }
            }
        }
    }
}

void mt_fast_neighbors(const MtDev& d, const MtParams& p, int cond, hipStream_t s)
{
    if (d.force_dense || d.n > 40 * d.ncells) {
        // dense cells: workgroups of NBD_ATOMS home atoms, about 1.5x the mean
        // cell's batches per cell (a fuller cell loops)
        const int mean = (d.n + d.ncells - 1) / d.ncells;
        const dim3 grid(d.ncells, std::max(1, (3 * mean / 2 + NBD_ATOMS - 1) / NBD_ATOMS));
        if (d.has_xmol)
            hipLaunchKernelGGL(k_fast_neigh_dense<2>, grid, dim3(256), 0, s, d, p, cond);
        else if (d.has_molecules)
            hipLaunchKernelGGL(k_fast_neigh_dense<1>, grid, dim3(256), 0, s, d, p, cond);
        else
            hipLaunchKernelGGL(k_fast_neigh_dense<0>, grid, dim3(256), 0, s, d, p, cond);
        return;
    }
    const int nb = ((d.nc[0] + 1) / 2) * ((d.nc[1] + 1) / 2) * ((d.nc[2] + 1) / 2);
    if (d.tile16) {
        if (d.has_xmol)
            hipLaunchKernelGGL((k_fast_neigh_block<2, true>), dim3(nb), dim3(256), 0, s, d, p, cond);
        else if (d.has_molecules)
            hipLaunchKernelGGL((k_fast_neigh_block<1, true>), dim3(nb), dim3(256), 0, s, d, p, cond);
        else
            hipLaunchKernelGGL((k_fast_neigh_block<0, true>), dim3(nb), dim3(256), 0, s, d, p, cond);
    } else {
        if (d.has_xmol)
            hipLaunchKernelGGL((k_fast_neigh_block<2, false>), dim3(nb), dim3(256), 0, s, d, p, cond);
        else if (d.has_molecules)
            hipLaunchKernelGGL((k_fast_neigh_block<1, false>), dim3(nb), dim3(256), 0, s, d, p, cond);
        else
            hipLaunchKernelGGL((k_fast_neigh_block<0, false>), dim3(nb), dim3(256), 0, s, d, p, cond);
    }
}

// ---------------------------------------------------------------------------
// pair forces
// ---------------------------------------------------------------------------

// The FP64/MIXED neighbour lists are chunked column-major: word j of atom k
// sits in 16-byte chunk j/4 at nlist[4 * ((j/4) * n + k) + j%4], so a
// wavefront (consecutive atoms) loads one contiguous 1 KB line of chunks per
// four neighbours. Unused words of an atom's last chunk hold the atom itself
// with offset code (0, 0, 0), which gives r = 0 and no force.
typedef unsigned int mt_u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t mt_fast_word(const uint32_t* nl, int n, int k, int j)
{
    return nl[(((size_t)(j >> 2) * n + k) << 2) + (j & 3)];
}

// the float4 positions of a chunk's four neighbours
__device__ __forceinline__ void mt_gather4(const float4* posf, mt_u32x4 w, float4& a0, float4& a1, float4& a2,
                                           float4& a3)
{
    a0 = posf[w.x & MT_NW_IDX_MASK];
    a1 = posf[w.y & MT_NW_IDX_MASK];
    a2 = posf[w.z & MT_NW_IDX_MASK];
    a3 = posf[w.w & MT_NW_IDX_MASK];
}

// force (and half energy) of one MIXED pair inside rc from its FP32
// separation: segment coefficients (a_f, b_f[, a_e, b_e]) of the pair's summed
// tables, plus the cutoff Coulomb term
template <bool WANT_E, bool COUL, typename CoefT>
__device__ __forceinline__ void mt_mixed_force(float dx, float dy, float dz, float r2, const float4& pa, float qa,
                                               const CoefT* row, float x0, float invd, float qk, float& Fx,
                                               float& Fy, float& Fz, float& U)
{
    if (!(r2 >= 1.0e-6f)) return;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    float fr = 0.0f;
    const int idx = (int)((r - x0) * invd);
    if ((unsigned)idx < (unsigned)(MT_NPTS - 1)) {
        // pa.w holds the neighbour's type times MT_NPTS: its offset in the row
        const CoefT c = row[__float_as_int(pa.w) + idx];
        fr = c.x + c.y * rinv;
        if constexpr (WANT_E) U += 0.5f * (c.z * r + c.w);
    }
    if (COUL) {
        const float K = MT_COULOMB_K * qk * qa;
        fr += K * rinv * rinv * rinv;
        if (WANT_E) U += 0.5f * K * rinv;
    }
    Fx = fmaf(dx, fr, Fx);
    Fy = fmaf(dy, fr, Fy);
    Fz = fmaf(dz, fr, Fz);
}

// one MIXED pair: FP32 geometry on local coordinates, FP32 force. The pair
// potentials are truncated at rc without a shift, so a pair within FP32
// rounding of rc would be counted or not depending on precision. A pair in a
// narrow band around rc is therefore left out here and the call returns true;
// mt_mixed_band later takes its inside test in FP64 with the FP64 engine's
// (the reference's) predicate, after the chunk's other pairs, so the double
// positions are loaded only for those rare pairs.
template <bool WANT_E, bool COUL, typename CoefT>
__device__ __forceinline__ bool mt_mixed_pair_o(const float4& pa, const float4& o, float qa, const float4& pk,
                                                const CoefT* row, float rc2, float x0, float invd, float qk,
                                                float& Fx, float& Fy, float& Fz, float& U)
{
    const float dx = pk.x - pa.x - o.x;
    const float dy = pk.y - pa.y - o.y;
    const float dz = pk.z - pa.z - o.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    // the band is r2 within 2e-5 rc2 of rc2 (loop-invariant bounds, hoisted):
    // one compare on the common path
    const float lo = rc2 * (1.0f - 2.0e-5f), hi = rc2 * (1.0f + 2.0e-5f);
    if (r2 < lo) {
        mt_mixed_force<WANT_E, COUL>(dx, dy, dz, r2, pa, qa, row, x0, invd, qk, Fx, Fy, Fz, U);
        return false;
    }
    return r2 < hi;
}

// the same with the shift of the word's offset code and the charge from
// global memory
template <bool WANT_E, bool COUL, typename CoefT>
__device__ __forceinline__ bool mt_mixed_pair(uint32_t w, const float4& pa, const float4& pk, const float4* sh,
                                              const CoefT* row, float rc2, float x0, float invd, float qk,
                                              const float* qf, float& Fx, float& Fy, float& Fz, float& U)
{
    const float qa = COUL ? qf[w & MT_NW_IDX_MASK] : 0.0f;
    return mt_mixed_pair_o<WANT_E, COUL>(pa, sh[w >> MT_NW_IDX_BITS], qa, pk, row, rc2, x0, invd, qk, Fx, Fy, Fz,
                                         U);
}

// a band pair of mt_mixed_pair: FP64 inside test from the double positions,
// then the same FP32 force
template <bool WANT_E, bool COUL, typename CoefT>
__device__ __forceinline__ void mt_mixed_band(uint32_t w, const float4& pk, const float4* sh, const CoefT* row,
                                           float x0, float invd, float qk, float& Fx, float& Fy, float& Fz,
                                           float& U, const MtDev& d, const MtParams& p, int k)
{
    const int u = (int)(w & MT_NW_IDX_MASK);          // place in cell order
    const double4 rk4 = d.pos[k], ra4 = d.pos[d.cell_list[u]];
    if (!(mt_dist2_pbc({rk4.x, rk4.y, rk4.z}, {ra4.x, ra4.y, ra4.z}, p.wnl) < (double)p.rc2f)) return;
    const float4 pa = d.posf[u];
    const float4 o = sh[w >> MT_NW_IDX_BITS];
    const float dx = pk.x - pa.x - o.x;
    const float dy = pk.y - pa.y - o.y;
    const float dz = pk.z - pa.z - o.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    mt_mixed_force<WANT_E, COUL>(dx, dy, dz, r2, pa, COUL ? d.qfs[u] : 0.0f, row, x0, invd, qk, Fx, Fy, Fz, U);
}

// the pair loop of one atom over chunks q0, q0 + qs, ... of its chunked list
// (qs lanes share an atom when the system is too small to fill the GPU with
// one lane per atom): four gathers in flight together, and the next chunk's
// words loaded while the current chunk is computed
template <bool WANT_E, bool COUL, typename CoefT>
__device__ __forceinline__ void mt_mixed_atom(const MtDev& d, const MtParams& p, int k, const float4& pk,
                                              const float4* sh, const CoefT* row, float qk, double& Fx,
                                              double& Fy, double& Fz, float& U, int q0, int qs)
{
    const int nch = (min(d.ncount[k], d.ncap) + 3) >> 2;
    if (q0 >= nch) return;
    const mt_u32x4* __restrict__ nl4 = reinterpret_cast<const mt_u32x4*>(d.nlist) + k;
    const size_t stride = (size_t)d.n;
    const float rc2 = p.rc2f, x0 = p.x0, invd = p.inv_delta;
    mt_u32x4 wn = __builtin_nontemporal_load(nl4 + (size_t)q0 * stride);
    for (int q = q0; q < nch; q += qs) {
        const mt_u32x4 w = wn;
        float4 a0, a1, a2, a3;
        mt_gather4(d.posf, w, a0, a1, a2, a3);
        // unconditional (the last iteration reloads a chunk of this atom's
        // list) so the load needs no branch and stays in flight across the
        // pair math
        wn = __builtin_nontemporal_load(nl4 + (size_t)min(q + qs, nch - 1) * stride);
        // the empty asm makes the four whole 16-byte gathers complete here
        // (while the next chunk's load stays in flight), so the type bits in
        // .w are not fetched by a second, dependent load inside a pair branch
        asm volatile("" : "+v"(a0.w), "+v"(a1.w), "+v"(a2.w), "+v"(a3.w));
        // the four pair forces of a chunk are summed in FP32, the chunk sums in FP64
        float cx = 0.f, cy = 0.f, cz = 0.f;
        const bool b0 = mt_mixed_pair<WANT_E, COUL>(w.x, a0, pk, sh, row, rc2, x0, invd, qk, d.qfs, cx, cy, cz, U);
        const bool b1 = mt_mixed_pair<WANT_E, COUL>(w.y, a1, pk, sh, row, rc2, x0, invd, qk, d.qfs, cx, cy, cz, U);
        const bool b2 = mt_mixed_pair<WANT_E, COUL>(w.z, a2, pk, sh, row, rc2, x0, invd, qk, d.qfs, cx, cy, cz, U);
        const bool b3 = mt_mixed_pair<WANT_E, COUL>(w.w, a3, pk, sh, row, rc2, x0, invd, qk, d.qfs, cx, cy, cz, U);
        if (b0 | b1 | b2 | b3) {
            if (b0) mt_mixed_band<WANT_E, COUL>(w.x, pk, sh, row, x0, invd, qk, cx, cy, cz, U, d, p, k);
            if (b1) mt_mixed_band<WANT_E, COUL>(w.y, pk, sh, row, x0, invd, qk, cx, cy, cz, U, d, p, k);
            if (b2) mt_mixed_band<WANT_E, COUL>(w.z, pk, sh, row, x0, invd, qk, cx, cy, cz, U, d, p, k);
            if (b3) mt_mixed_band<WANT_E, COUL>(w.w, pk, sh, row, x0, invd, qk, cx, cy, cz, U, d, p, k);
        }
        Fx += (double)cx;
        Fy += (double)cy;
        Fz += (double)cz;
    }
}


// MIXED with energies: coefficients (a_f, b_f, a_e, b_e) read from global
// memory; G lanes per atom
template <bool WANT_E, bool COUL, int G>
__global__ void __launch_bounds__(512) k_fast_force_mixed(MtDev d, MtParams p)
{
    __shared__ float4 sh[27];
    mt_fill_shifts(sh, p);
    __syncthreads();
    const int t = mt_xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int k = t / G, g = t % G;
    if (k >= d.n) return;                       // all G lanes of an atom leave together
    const float4 pk = d.posf[d.sidx[k]];
    const float4* __restrict__ row = d.nbcoef + (size_t)__float_as_int(pk.w) * d.T;
    const float qk = COUL ? d.qf[k] : 0.0f;
    double Fx = 0.0, Fy = 0.0, Fz = 0.0;
    float U = 0.0f;
    mt_mixed_atom<WANT_E, COUL>(d, p, k, pk, sh, row, qk, Fx, Fy, Fz, U, g, G);
    mt_lane_sum<G>(Fx, Fy, Fz, U);
    if (g != 0) return;
    if (!d.mobile[k]) { Fx = Fy = Fz = 0.0; U = 0.0f; }
    d.frc[k] = make_double4(Fx, Fy, Fz, 0.0);
    if (WANT_E) d.epot[k] = U;
}

// MIXED with the force-only segment coefficients (a_f, b_f) of every type pair
// staged in LDS once per workgroup. The grid holds as many workgroups as can
// be resident at once; workgroup w owns atoms [w n / g, (w+1) n / g), and the
// ranges are XCD-contiguous. G lanes per atom. Used when T^2 * 1000 * 8 bytes
// fits the LDS budget and energies are not needed.
template <bool COUL, int G>
__global__ void __launch_bounds__(512) k_fast_force_mixed_lds(MtDev d, MtParams p)
{
    extern __shared__ float2 stab[];
    __shared__ float4 sh[27];
    const int T = d.T;
    const int ntab = T * T * MT_NPTS;
    for (int i = threadIdx.x; i < ntab; i += blockDim.x) stab[i] = d.nbcoef2[i];
    mt_fill_shifts(sh, p);
    __syncthreads();
    const int lw = mt_xcd_block(blockIdx.x, gridDim.x);
    const int kb = (int)((long long)lw * d.n / gridDim.x), ke = (int)((long long)(lw + 1) * d.n / gridDim.x);
    const int g = threadIdx.x % G;
    for (int k = kb + (int)threadIdx.x / G; k < ke; k += blockDim.x / G) {
        const float4 pk = d.posf[d.sidx[k]];
        const float2* __restrict__ row = stab + __float_as_int(pk.w) * T;
        const float qk = COUL ? d.qf[k] : 0.0f;
        double Fx = 0.0, Fy = 0.0, Fz = 0.0;
        float U = 0.0f;
        mt_mixed_atom<false, COUL>(d, p, k, pk, sh, row, qk, Fx, Fy, Fz, U, g, G);
        mt_lane_sum<G>(Fx, Fy, Fz, U);
        if (g == 0) {
            if (!d.mobile[k]) { Fx = Fy = Fz = 0.0; }
            d.frc[k] = make_double4(Fx, Fy, Fz, 0.0);
        }
    }
}

// MIXED pair forces with neighbour-cell tiles in LDS. One workgroup per
// 2x2x2 block of home cells (the list build's blocking). The positions (and
// charges) of the 4x4x4 cells around the block, the union of the eight
// 27-cell neighbourhoods, are staged once in LDS in cell order, moved from
// cell-centre coordinates into the frame of the centre of the block's first
// cell (x0, y0, z0): the staged cells at offsets -1..2 are shifted by
// -1..2 cell widths, products that are exact in FP32, so the separation of
// a pair is one difference of staged positions with no per-pair shift. The
// list words of such a block (k_fast_neigh_block, W16) are the neighbour's
// LDS place itself, 16 bits, eight to a 16-byte chunk: half the list bytes
// of the coded 32-bit words, and one dependent LDS read per pair. A
// workgroup whose staged cells hold more than FT_CAP atoms has 32-bit coded
// words and gathers from global memory instead (per-offset shifts). G lanes
// per home atom.

// four pairs of a lane, without a branch on the common path: a pair outside
// rc (or the atom itself) still reads a segment and its force is selected
// away, and a pair outside the table reads segment MT_NPTS - 1 of its row,
// which is zero, so the four coefficient reads are issued together and
// waited for once. The arithmetic per pair is mt_mixed_pair_o's. Pairs in
// the band around rc come back as bits 0..3.
template <bool WANT_E, bool COUL, bool SHIFT, typename CoefT>
__device__ __forceinline__ uint32_t mt_tile_pairs4(const float4 (&a)[4], const float4 (&o)[4], const float (&qa)[4],
                                                   const float4& pk, const CoefT* row, float lo, float hi, float x0,
                                                   float invd, float qk, float& cx, float& cy, float& cz, float& U)
{
    constexpr bool LEAN = !COUL;
    float dx[4], dy[4], dz[4], rinv[4], r[4];
    bool in[4];
    int ix[4];
    uint32_t band = 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        dx[i] = SHIFT ? pk.x - a[i].x - o[i].x : pk.x - a[i].x;
        dy[i] = SHIFT ? pk.y - a[i].y - o[i].y : pk.y - a[i].y;
        dz[i] = SHIFT ? pk.z - a[i].z - o[i].z : pk.z - a[i].z;
        int seg;
        if constexpr (LEAN) {
            const float r2 = mt_lean_r2(dx[i], dy[i], dz[i]);
            band |= (r2 >= lo && r2 < hi) ? 1u << i : 0u;
            rinv[i] = mt_lean_rinv(r2, x0, invd, seg);
            in[i] = true;
            r[i] = WANT_E ? r2 * rinv[i] : 0.0f;
        } else {
            const float r2 = dx[i] * dx[i] + dy[i] * dy[i] + dz[i] * dz[i];
            const bool lt = r2 < lo;
            in[i] = lt && r2 >= 1.0e-6f;
            band |= (!lt && r2 < hi) ? 1u << i : 0u;
            rinv[i] = __builtin_amdgcn_rsqf(fmaxf(r2, 1.0e-6f));
            r[i] = r2 * rinv[i];
            const int idx = (int)((r[i] - x0) * invd);
            seg = (int)min((unsigned)idx, (unsigned)(MT_NPTS - 1));
        }
        // .w: the neighbour's type times MT_NPTS, its offset in the row; in
        // elements for global neighbours (SHIFT), in bytes for staged ones
        ix[i] = __float_as_int(a[i].w) + (SHIFT ? seg : seg * (int)sizeof(CoefT));
    }
    CoefT c[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
        c[i] = SHIFT ? row[ix[i]] : *reinterpret_cast<const CoefT*>(reinterpret_cast<const char*>(row) + ix[i]);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if constexpr (LEAN) {
            const float fr = fmaf(c[i].y, rinv[i], c[i].x);
            if constexpr (WANT_E) U += 0.5f * fmaf(c[i].z, r[i], c[i].w);
            cx = fmaf(dx[i], fr, cx);
            cy = fmaf(dy[i], fr, cy);
            cz = fmaf(dz[i], fr, cz);
            continue;
        }
        float fr = c[i].x + c[i].y * rinv[i];
        float K = 0.0f;
        if constexpr (COUL) {
            K = MT_COULOMB_K * qk * qa[i];
            fr += K * rinv[i] * rinv[i] * rinv[i];
        }
        fr = in[i] ? fr : 0.0f;
        if constexpr (WANT_E) {
            U += in[i] ? 0.5f * (c[i].z * r[i] + c[i].w) : 0.0f;
            if constexpr (COUL) U += in[i] ? 0.5f * K * rinv[i] : 0.0f;
        }
        cx = fmaf(dx[i], fr, cx);
        cy = fmaf(dy[i], fr, cy);
        cz = fmaf(dz[i], fr, cz);
    }
    return band;
}

// a band pair of a 16-bit block list: the staged cell of LDS place l gives
// the neighbour's cell-order place, the inside test is taken in FP64 with the
// reference predicate (as mt_mixed_band), the force from the staged positions
template <bool WANT_E, bool COUL, typename CoefT>
__device__ __forceinline__ void mt_tile_band16(int l, const float4& a, float qa, const float4& pk, const CoefT* row,
                                               float x0, float invd, float qk, float& cx, float& cy, float& cz,
                                               float& U, const MtDev& d, const MtParams& p, int k,
                                               const int* s_first, const int* s_off)
{
    int lo = 0, hi = 64;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_off[mid] <= l) lo = mid; else hi = mid;
    }
    const int u = s_first[lo] + (l - s_off[lo]);
    const double4 rk4 = d.pos[k], ra4 = d.pos[d.cell_list[u]];
    const bool inside = mt_dist2_pbc({rk4.x, rk4.y, rk4.z}, {ra4.x, ra4.y, ra4.z}, p.wnl) < (double)p.rc2f;
    if constexpr (!COUL) {
        // the pair loop added this pair's table terms: they stand when FP64
        // puts the pair inside, and are taken back otherwise
        if (inside) return;
        const float dx = pk.x - a.x, dy = pk.y - a.y, dz = pk.z - a.z;
        int seg;
        const float r2 = mt_lean_r2(dx, dy, dz);
        const float rinv = mt_lean_rinv(r2, x0, invd, seg);
        const CoefT c = *reinterpret_cast<const CoefT*>(reinterpret_cast<const char*>(row) + __float_as_int(a.w) +
                                                        seg * (int)sizeof(CoefT));
        const float fr = fmaf(c.y, rinv, c.x);
        if constexpr (WANT_E) U -= 0.5f * fmaf(c.z, r2 * rinv, c.w);
        cx = fmaf(-dx, fr, cx);
        cy = fmaf(-dy, fr, cy);
        cz = fmaf(-dz, fr, cz);
        return;
    }
    if (!inside) return;
    const float dx = pk.x - a.x, dy = pk.y - a.y, dz = pk.z - a.z;
    const float r2 = dx * dx + dy * dy + dz * dz;
    // the staged type offset is in bytes, mt_mixed_force's in elements
    const float4 ae = make_float4(a.x, a.y, a.z, __int_as_float(__float_as_int(a.w) / (int)sizeof(CoefT)));
    mt_mixed_force<WANT_E, COUL>(dx, dy, dz, r2, ae, qa, row, x0, invd, qk, cx, cy, cz, U);
}

// the pair loop of a home atom of a 16-bit block over chunks q0, q0 + qs, ...
// (the next chunk's words in flight across the current chunk's pairs); the
// pair forces of a chunk's eight pairs are summed in FP32, those sums in FP64
template <bool WANT_E, bool COUL, typename CoefT>
__device__ __forceinline__ void mt_tile_atom16(const MtDev& d, const MtParams& p, int k, const float4& pk,
                                               const float4* spos, const float* sq, const int* s_first,
                                               const int* s_off, const CoefT* row, float qk, double& Fx, double& Fy,
                                               double& Fz, float& U, int q0, int qs)
{
    const int nch = (min(d.ncount[k], 2 * d.ncap) + 7) >> 3;
    if (q0 >= nch) return;
    const mt_u32x4* __restrict__ nl4 = reinterpret_cast<const mt_u32x4*>(d.nlist) + k;
    const size_t stride = (size_t)d.n;
    const float x0 = p.x0, invd = p.inv_delta;
    const float lo = p.rc2f * (1.0f - 2.0e-5f), hi = p.rc2f * (1.0f + 2.0e-5f);
    mt_u32x4 wn = __builtin_nontemporal_load(nl4 + (size_t)q0 * stride);
    for (int q = q0; q < nch; q += qs) {
        const mt_u32x4 w = wn;
        const int li[8] = {(int)(w.x & 0xffffu), (int)(w.x >> 16), (int)(w.y & 0xffffu), (int)(w.y >> 16),
                           (int)(w.z & 0xffffu), (int)(w.z >> 16), (int)(w.w & 0xffffu), (int)(w.w >> 16)};
        float4 a[8];
        float qa[8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            a[i] = spos[li[i]];
            qa[i] = COUL ? sq[li[i]] : 0.0f;
        }
        // unconditional (the last iteration reloads a chunk of this atom's list)
        wn = __builtin_nontemporal_load(nl4 + (size_t)min(q + qs, nch - 1) * stride);
        float cx = 0.f, cy = 0.f, cz = 0.f;
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
            const float4 ah[4] = {a[4 * hh], a[4 * hh + 1], a[4 * hh + 2], a[4 * hh + 3]};
            const float qh[4] = {qa[4 * hh], qa[4 * hh + 1], qa[4 * hh + 2], qa[4 * hh + 3]};
            const uint32_t band =
                mt_tile_pairs4<WANT_E, COUL, false>(ah, ah, qh, pk, row, lo, hi, x0, invd, qk, cx, cy, cz, U);
            if (band) {
#pragma unroll
                for (int i = 0; i < 4; i++)
                    if (band & (1u << i))
                        mt_tile_band16<WANT_E, COUL>(li[4 * hh + i], ah[i], qh[i], pk, row, x0, invd, qk, cx, cy, cz,
                                                     U, d, p, k, s_first, s_off);
            }
        }
        Fx += (double)cx;
        Fy += (double)cy;
        Fz += (double)cz;
    }
}

// a band pair of a 32-bit block list in a pass without charges: the pair
// loop's terms (mt_tile_pairs4 with SHIFT) are taken back when the FP64
// reference predicate puts the pair outside rc
template <bool WANT_E, typename CoefT>
__device__ __forceinline__ void mt_tile_band32_lean(uint32_t w, const float4& pk, const float4* shd, const CoefT* row,
                                                    float x0, float invd, float& cx, float& cy, float& cz, float& U,
                                                    const MtDev& d, const MtParams& p, int k)
{
    const int u = (int)(w & MT_NW_IDX_MASK);          // place in cell order
    const double4 rk4 = d.pos[k], ra4 = d.pos[d.cell_list[u]];
    if (mt_dist2_pbc({rk4.x, rk4.y, rk4.z}, {ra4.x, ra4.y, ra4.z}, p.wnl) < (double)p.rc2f) return;
    const float4 a = d.posf[u];
    const float4 o = shd[__builtin_amdgcn_ubfe(w, MT_NW_IDX_BITS, 5)];
    const float dx = pk.x - a.x - o.x, dy = pk.y - a.y - o.y, dz = pk.z - a.z - o.z;
    int seg;
    const float r2 = mt_lean_r2(dx, dy, dz);
    const float rinv = mt_lean_rinv(r2, x0, invd, seg);
    const CoefT c = row[__float_as_int(a.w) + seg];
    const float fr = fmaf(c.y, rinv, c.x);
    if constexpr (WANT_E) U -= 0.5f * fmaf(c.z, r2 * rinv, c.w);
    cx = fmaf(-dx, fr, cx);
    cy = fmaf(-dy, fr, cy);
    cz = fmaf(-dz, fr, cz);
}

// the same for a block over FT_CAP atoms: 32-bit coded words, positions and
// charges gathered from global memory, the offset's shift from shd
template <bool WANT_E, bool COUL, typename CoefT>
__device__ __forceinline__ void mt_tile_atom32(const MtDev& d, const MtParams& p, int k, const float4& pk,
                                               const float4* shd, const CoefT* row, float qk, double& Fx, double& Fy,
                                               double& Fz, float& U, int q0, int qs)
{
    const int nch = (min(d.ncount[k], d.ncap) + 3) >> 2;
    if (q0 >= nch) return;
    const mt_u32x4* __restrict__ nl4 = reinterpret_cast<const mt_u32x4*>(d.nlist) + k;
    const size_t stride = (size_t)d.n;
    const float x0 = p.x0, invd = p.inv_delta;
    const float lo = p.rc2f * (1.0f - 2.0e-5f), hi = p.rc2f * (1.0f + 2.0e-5f);
    mt_u32x4 wn = __builtin_nontemporal_load(nl4 + (size_t)q0 * stride);
    for (int q = q0; q < nch; q += qs) {
        const mt_u32x4 w = wn;
        const uint32_t wv[4] = {w.x, w.y, w.z, w.w};
        float4 a[4], o[4];
        float qa[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int ia = (int)(wv[i] & MT_NW_IDX_MASK);
            o[i] = shd[__builtin_amdgcn_ubfe(wv[i], MT_NW_IDX_BITS, 5)];
            a[i] = d.posf[ia];
            qa[i] = COUL ? d.qfs[ia] : 0.0f;
        }
        wn = __builtin_nontemporal_load(nl4 + (size_t)min(q + qs, nch - 1) * stride);
        float cx = 0.f, cy = 0.f, cz = 0.f;
        const uint32_t band = mt_tile_pairs4<WANT_E, COUL, true>(a, o, qa, pk, row, lo, hi, x0, invd, qk, cx, cy, cz, U);
        if (band) {
#pragma unroll
            for (int i = 0; i < 4; i++)
                if (band & (1u << i)) {
                    if constexpr (!COUL)
                        mt_tile_band32_lean<WANT_E>(wv[i], pk, shd, row, x0, invd, cx, cy, cz, U, d, p, k);
                    else
                        mt_mixed_band<WANT_E, COUL>(wv[i], pk, shd, row, x0, invd, qk, cx, cy, cz, U, d, p, k);
                }
        }
        Fx += (double)cx;
        Fy += (double)cy;
        Fz += (double)cz;
    }
}

// the staged cells of the block whose first cell is (x0, y0, z0): staged cell
// q = (ox+1) + 4 (oy+1) + 16 (oz+1), o* in -1..2, has its first cell-order
// place in s_first[q], its count in s_cnt[q] and its first staged place in
// s_off[q] (s_off[64] = total); home cell h = (h&1, (h>>1)&1, h>>2) has
// first place s_hfirst[h] and first home index s_hoff[h] (s_hoff[8] = home
// atoms). Threads 0..63 write; the caller synchronises.
__device__ __forceinline__ void mt_tile_cells(const MtDev& d, int x0, int y0, int z0, int* s_first, int* s_cnt,
                                              int* s_off, int* s_hfirst, int* s_hoff)
{
    const int t = threadIdx.x;
    if (t >= 64) return;
    const int Nx = d.nc[0], Ny = d.nc[1], Nz = d.nc[2];
    const int ox = (t & 3) - 1, oy = ((t >> 2) & 3) - 1, oz = (t >> 4) - 1;
    int cx = x0 + ox, cy = y0 + oy, cz = z0 + oz;
    cx = cx < 0 ? cx + Nx : (cx >= Nx ? cx - Nx : cx);
    cy = cy < 0 ? cy + Ny : (cy >= Ny ? cy - Ny : cy);
    cz = cz < 0 ? cz + Nz : (cz >= Nz ? cz - Nz : cz);
    const int cc = cx + Nx * (cy + Ny * cz);
    const int cn = d.cell_count[cc];
    s_first[t] = d.cell_first[cc];
    s_cnt[t] = cn;
    int v = cn;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o);
        if (t >= o) v += u;
    }
    s_off[t + 1] = v;
    if (t == 0) s_off[0] = 0;
    // home cells (the first eight lanes): counts scanned across them
    int hn = 0, hf = 0;
    if (t < 8) {
        const int hx = x0 + (t & 1), hy = y0 + ((t >> 1) & 1), hz = z0 + (t >> 2);
        if (hx < Nx && hy < Ny && hz < Nz) {
            const int hcell = hx + Nx * (hy + Ny * hz);
            hn = d.cell_count[hcell];
            hf = d.cell_first[hcell];
        }
    }
    int hv = hn;
    for (int o = 1; o < 8; o <<= 1) {
        const int u = __shfl_up(hv, o);
        if (t >= o) hv += u;
    }
    if (t < 8) {
        s_hfirst[t] = hf;
        s_hoff[t + 1] = hv;
        if (t == 0) s_hoff[0] = 0;
    }
}

// the block of workgroup blockIdx.x (XCD-contiguous order) and its first cell
__device__ __forceinline__ void mt_tile_block(const MtDev& d, int& x0, int& y0, int& z0)
{
    const int bx = (d.nc[0] + 1) >> 1, by = (d.nc[1] + 1) >> 1;
    const int g = mt_xcd_block(blockIdx.x, gridDim.x);
    x0 = 2 * (g % bx);
    y0 = 2 * ((g / bx) % by);
    z0 = 2 * (g / (bx * by));
}

// home atom v of a block: its home cell
__device__ __forceinline__ int mt_tile_home(const int* s_hoff, int v)
{
    int h = 0;
    while (h < 7 && v >= s_hoff[h + 1]) h++;
    return h;
}

// WANT_E: per-atom energies too, with the (a_f, b_f, a_e, b_e) coefficients
// read from global memory; otherwise (a_f, b_f) of every type pair staged in LDS
// the NVE integration of one atom and the rebuild flag (defined with the
// integrator kernels below)
__device__ __forceinline__ void mt_fast_back_half(const MtParams& p, double4& F, double4& pm, double4& r,
                                                  bool& capped);
__device__ __forceinline__ void mt_fast_kick_drift_atom(const MtParams& p, double4& pm, double4& r, const double4& F);
__device__ __forceinline__ float mt_fast_locate_at(const MtDev& d, const MtParams& p, const double4& q,
                                                   const double4& r, int mixed, int c0, int u, int tw,
                                                   float4* posf_out, bool with_w);
__device__ __forceinline__ void mt_fast_flag_disp(const MtDev& d, const MtParams& p, float d2);

// FUSE (NVE steps of a call but the last, no forces added after this pass):
// each home atom's force is followed at once, in the lane that reduced it, by
// the back half of this step and the front half of the next
// (k_fast_nve_fused's arithmetic, the same bits), the new positions written
// to pos2 and posf2, since the staged positions of this pass must not change
// under other workgroups; the host then swaps the buffers
template <bool WANT_E, bool COUL, int G, bool FUSE = false>
__global__ void __launch_bounds__(256 * G) __attribute__((amdgpu_waves_per_eu(4)))
k_fast_force_tile(MtDev d, MtParams p)
{
    if constexpr (FUSE) MT_WRITER_CLEAR(d, p);
    __shared__ float4 spos[FT_CAP];
    __shared__ float sq_buf[COUL ? FT_CAP : 1];
    __shared__ float4 shd[27];                        // 32-bit blocks: shift of each offset code
    __shared__ int s_first[64], s_cnt[64], s_off[65];
    __shared__ int s_hfirst[8], s_hoff[9];
    const int t = threadIdx.x;
    const int T = d.T;
    // the segment coefficients in LDS (read from global memory instead, L1/L2
    // hits, the pass measured 1.7x slower on the 1M-atom mixture)
    // T * T rows of MT_NPTS segment coefficients, or 2T - 1 symmetric rows
    // (d.tabsym: row ti + tj, so a pair's row offset is the sum of the two
    // staged type offsets)
    extern __shared__ float2 stab[];
    const int trow = d.tabsym ? 1 : T;                // row offset factor of the home atom's type
    if constexpr (!WANT_E) {
        const int ntab = (d.tabsym ? 2 * T - 1 : T * T) * MT_NPTS;
        const float2* src = d.tabsym ? d.nbcoef2s : d.nbcoef2;
        for (int i = t; i < ntab; i += blockDim.x) stab[i] = src[i];
    }
    int x0, y0, z0;
    mt_tile_block(d, x0, y0, z0);
    mt_tile_cells(d, x0, y0, z0, s_first, s_cnt, s_off, s_hfirst, s_hoff);
    __syncthreads();
    const int total = s_off[64];
    const bool w16 = total <= FT_CAP;                 // uniform: the list build's choice for this block
    if (w16) {
        // blockDim / 64 threads a staged cell, each taking every (blockDim / 64)-th atom
        const int per = (int)blockDim.x >> 6, q = t / per;
        const float sx = (float)((q & 3) - 1) * p.cs[0], sy = (float)(((q >> 2) & 3) - 1) * p.cs[1];
        const float sz = (float)((q >> 4) - 1) * p.cs[2];
        for (int m = t % per; m < s_cnt[q]; m += per) {
            const int u = s_first[q] + m, i = s_off[q] + m;
            const float4 a = d.posf[u];
            // .w: the type's offset in a coefficient row, in bytes
            spos[i] = make_float4(a.x + sx, a.y + sy, a.z + sz,
                                  __int_as_float(__float_as_int(a.w) * (WANT_E ? (int)sizeof(float4) : (int)sizeof(float2))));
            if constexpr (COUL) sq_buf[i] = d.qfs[u];
        }
    } else if (t < 27) {
        shd[t] = make_float4((float)(t % 3 - 1) * p.cs[0], (float)((t / 3) % 3 - 1) * p.cs[1],
                             (float)(t / 9 - 1) * p.cs[2], 0.0f);
    }
    __syncthreads();
    const int nh = s_hoff[8];
    const int gl = t % G;
    float d2max = 0.0f;                               // FUSE: largest squared displacement of this lane's atoms
    for (int v = t / G; v < nh; v += (int)blockDim.x / G) {
        const int h = mt_tile_home(s_hoff, v);
        const int u = s_hfirst[h] + (v - s_hoff[h]);
        const int k = d.cell_list[u];
        const float qk = COUL ? d.qf[k] : 0.0f;
        double Fx = 0.0, Fy = 0.0, Fz = 0.0;
        float U = 0.0f;
        int tw16;                                     // .w of the home atom's entry
        if (w16) {
            // the home atom's own staged entry (home cell h is staged cell qh)
            const int qh = (1 + (h & 1)) + 4 * ((1 + ((h >> 1) & 1)) + 4 * (1 + (h >> 2)));
            const float4 pk = spos[s_off[qh] + (v - s_hoff[h])];
            tw16 = __float_as_int(pk.w);
            // the home atom's row: its type offset (bytes) times T, or itself
            // in symmetric rows (the energy pass reads the full rows)
            const size_t rowb = (size_t)__float_as_int(pk.w) * (WANT_E ? T : trow);
            if constexpr (WANT_E)
                mt_tile_atom16<true, COUL>(d, p, k, pk, spos, sq_buf, s_first, s_off,
                                           reinterpret_cast<const float4*>(reinterpret_cast<const char*>(d.nbcoef) + rowb),
                                           qk, Fx, Fy, Fz, U, gl, G);
            else
                mt_tile_atom16<false, COUL>(d, p, k, pk, spos, sq_buf, s_first, s_off,
                                            reinterpret_cast<const float2*>(reinterpret_cast<const char*>(stab) + rowb),
                                            qk, Fx, Fy, Fz, U, gl, G);
        } else {
            const float4 pk = d.posf[u];
            tw16 = __float_as_int(pk.w);
            if constexpr (WANT_E)
                mt_tile_atom32<true, COUL>(d, p, k, pk, shd, d.nbcoef + (size_t)__float_as_int(pk.w) * T, qk, Fx, Fy,
                                           Fz, U, gl, G);
            else
                mt_tile_atom32<false, COUL>(d, p, k, pk, shd, stab + __float_as_int(pk.w) * trow, qk, Fx, Fy, Fz, U,
                                            gl, G);
        }
        mt_lane_sum<G>(Fx, Fy, Fz, U);
        if (gl == 0) {
            if (!d.mobile[k]) { Fx = Fy = Fz = 0.0; U = 0.0f; }
            if constexpr (FUSE) {
                // the lane that holds the force integrates the atom at once;
                // the force itself is not stored (the next pass computes the
                // next one, and a call's last step is not fused). The atom's
                // home cell, place and type come from the block, not from
                // home, sidx and type (the same values), so the state loads
                // below are all the tail waits for.
                const double4 q = d.posref[k];
                double4 F = make_double4(Fx, Fy, Fz, 0.0), pm = d.mom[k], r = d.pos[k];
                const int hc = (x0 + (h & 1)) + d.nc[0] * ((y0 + ((h >> 1) & 1)) + d.nc[1] * (z0 + (h >> 2)));
                const int tw = w16 ? (int)(tw16 / (WANT_E ? (int)sizeof(float4) : (int)sizeof(float2))) : tw16;
                bool capped;
                mt_fast_back_half(p, F, pm, r, capped);
                mt_fast_kick_drift_atom(p, pm, r, F);
                d2max = fmaxf(d2max, mt_fast_locate_at(d, p, q, r, 1, hc, u, tw, d.posf2, true));
                d.mom[k] = pm;
                d.pos2[k] = r;
            } else {
                d.frc[k] = make_double4(Fx, Fy, Fz, 0.0);
                if (WANT_E) d.epot[k] = U;
            }
        }
    }
    if constexpr (FUSE) mt_fast_flag_disp(d, p, d2max);
}

// FP64 arithmetic on the tiled design, for measurement (mtmd_bench_kernel 4):
// the force-only pass of k_fast_force_tile without charges, with everything
// its pair math touches in FP64. Positions are staged in the block frame from
// the FP64 positions (the cell-local coordinate of mt_fast_local, kept in
// FP64, plus the staged cell's offset in FP64 cell widths), r comes from an
// FP64 reciprocal square root, the segment coefficients are the same table's
// (FP32 in LDS, widened at use), and forces are summed in FP64. Two lanes an
// atom take alternate chunks of the same 16-bit block lists, as the FP32 pass
// does, so beside k_fast_force_tile it prices the precision alone. Blocks
// whose staged cells exceed FT_CAP atoms (32-bit lists) are left out and
// counted in flags[6].
#define FT64_TPB 512
#define FT64_G 2
__global__ void __launch_bounds__(FT64_TPB) k_fast_force_tile64(MtDev d, MtParams p)
{
    __shared__ double sx[FT_CAP], sy[FT_CAP], sz[FT_CAP];   // block-frame positions
    __shared__ unsigned char sty[FT_CAP];                   // types
    __shared__ int s_first[64], s_cnt[64], s_off[65];
    __shared__ int s_hfirst[8], s_hoff[9];
    extern __shared__ float2 stab[];                        // the coefficient rows (symmetric or T x T)
    const int t = threadIdx.x;
    const int T = d.T;
    {
        const int ntab = (d.tabsym ? 2 * T - 1 : T * T) * MT_NPTS;
        const float2* src = d.tabsym ? d.nbcoef2s : d.nbcoef2;
        for (int i = t; i < ntab; i += blockDim.x) stab[i] = src[i];
    }
    int x0, y0, z0;
    mt_tile_block(d, x0, y0, z0);
    mt_tile_cells(d, x0, y0, z0, s_first, s_cnt, s_off, s_hfirst, s_hoff);
    __syncthreads();
    if (s_off[64] > FT_CAP) {                         // uniform: a block of 32-bit lists
        if (t == 0) atomicAdd(&d.flags[6], 1u);
        return;
    }
    const int Nx = d.nc[0], Ny = d.nc[1], Nz = d.nc[2];
    const int per = (int)blockDim.x >> 6, q = t / per;
    {
        const int ox = (q & 3) - 1, oy = ((q >> 2) & 3) - 1, oz = (q >> 4) - 1;
        int cx = x0 + ox, cy = y0 + oy, cz = z0 + oz;
        cx = cx < 0 ? cx + Nx : (cx >= Nx ? cx - Nx : cx);
        cy = cy < 0 ? cy + Ny : (cy >= Ny ? cy - Ny : cy);
        cz = cz < 0 ? cz + Nz : (cz >= Nz ? cz - Nz : cz);
        const double ccx = -p.L[0] / 2.0 + ((double)cx + 0.5) * p.cw[0];
        const double ccy = -p.L[1] / 2.0 + ((double)cy + 0.5) * p.cw[1];
        const double ccz = -p.L[2] / 2.0 + ((double)cz + 0.5) * p.cw[2];
        const double ox_ = (double)ox * p.cw[0], oy_ = (double)oy * p.cw[1], oz_ = (double)oz * p.cw[2];
        for (int m = t % per; m < s_cnt[q]; m += per) {
            const int u = s_first[q] + m, k = d.cell_list[u], i = s_off[q] + m;
            const double4 r = d.pos[k];
            double dx = r.x - ccx, dy = r.y - ccy, dz = r.z - ccz;
            dx -= p.L[0] * rint(dx * p.invL[0]);
            dy -= p.L[1] * rint(dy * p.invL[1]);
            dz -= p.L[2] * rint(dz * p.invL[2]);
            sx[i] = dx + ox_;
            sy[i] = dy + oy_;
            sz[i] = dz + oz_;
            sty[i] = (unsigned char)d.type[k];
        }
    }
    __syncthreads();
    const int nh = s_hoff[8];
    const double x0d = (double)p.x0, invd = (double)p.inv_delta;
    const mt_u32x4* __restrict__ nl = reinterpret_cast<const mt_u32x4*>(d.nlist);
    const int gl = t % FT64_G;
    for (int v = t / FT64_G; v < nh; v += (int)blockDim.x / FT64_G) {
        const int h = mt_tile_home(s_hoff, v);
        const int u = s_hfirst[h] + (v - s_hoff[h]);
        const int k = d.cell_list[u];
        const int qh = (1 + (h & 1)) + 4 * ((1 + ((h >> 1) & 1)) + 4 * (1 + (h >> 2)));
        const int pl = s_off[qh] + (v - s_hoff[h]);
        const double px = sx[pl], py = sy[pl], pz = sz[pl];
        const int tk = sty[pl], rowbase = (d.tabsym ? tk : tk * T) * MT_NPTS;
        const int nch = (min(d.ncount[k], 2 * d.ncap) + 7) >> 3;
        double Fx = 0.0, Fy = 0.0, Fz = 0.0;
        float U = 0.0f;
        for (int c = gl; c < nch; c += FT64_G) {
            const mt_u32x4 w = __builtin_nontemporal_load(nl + (size_t)c * d.n + k);
            const int li[8] = {(int)(w.x & 0xffffu), (int)(w.x >> 16), (int)(w.y & 0xffffu), (int)(w.y >> 16),
                               (int)(w.z & 0xffffu), (int)(w.z >> 16), (int)(w.w & 0xffffu), (int)(w.w >> 16)};
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int l = li[i];
                const double dx = px - sx[l], dy = py - sy[l], dz = pz - sz[l];
                const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                // the atom itself (row padding): rinv 0, below the table, no force
                const double rinv = r2 > 0.0 ? rsqrt(r2) : 0.0;
                const int idx = (int)((r2 * rinv - x0d) * invd);
                const int seg = (int)min((unsigned)idx, (unsigned)(MT_NPTS - 1));   // beyond rc or below x0: zero row end
                const float2 cf = stab[rowbase + (int)sty[l] * MT_NPTS + seg];
                const double fr = fma((double)cf.y, rinv, (double)cf.x);
                Fx = fma(dx, fr, Fx);
                Fy = fma(dy, fr, Fy);
                Fz = fma(dz, fr, Fz);
            }
        }
        mt_lane_sum<FT64_G>(Fx, Fy, Fz, U);
        if (gl != 0) continue;
        if (!d.mobile[k]) Fx = Fy = Fz = 0.0;
        d.frc[k] = make_double4(Fx, Fy, Fz, 0.0);
    }
}

int mt_fast_force_tile64(const MtDev& d, const MtParams& p, hipStream_t s)
{
    if (!d.tile16 || d.T < 1 || d.T > 255) return -1;
    const int nb = ((d.nc[0] + 1) / 2) * ((d.nc[1] + 1) / 2) * ((d.nc[2] + 1) / 2);
    const size_t lds = (size_t)(d.tabsym ? 2 * d.T - 1 : d.T * d.T) * MT_NPTS * sizeof(float2);
    if (lds + FT_CAP * (3 * sizeof(double) + 1) + 2048 > 160 * 1024) return -1;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fast_force_tile64),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return -1;
    hipLaunchKernelGGL(k_fast_force_tile64, dim3(nb), dim3(FT64_TPB), lds, s, d, p);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// FP64: the reference's per-pair arithmetic (FunctorCalcForce.cpp:467-535)
// on the skin list, restricted to pairs inside rc by the reference predicate.
// G lanes per atom: lane g takes the list's 4-word chunks g, g + G, ...
template <bool WANT_E, bool COUL, int G>
__global__ void __launch_bounds__(256) k_fast_force_fp64(MtDev d, MtParams p)
{
    const int t = mt_xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int k = t / G, g = t % G;
    if (k >= d.n) return;                       // all G lanes of an atom leave together
    const double4 rk4 = d.pos[k];
    const mt_d3 rk = {rk4.x, rk4.y, rk4.z};
    const int row = d.type[k], T = d.T;
    const float qk = (float)rk4.w;
    mt_d3 F = {0.0, 0.0, 0.0};
    float U = 0.0f;
    const int cnt = min(d.ncount[k], d.ncap);
    for (int j = 4 * g; j < cnt; j += ((j & 3) == 3) ? 4 * G - 3 : 1) {
        const int a = (int)(mt_fast_word(d.nlist, d.n, k, j) & MT_NW_IDX_MASK);
        const double4 ra4 = d.pos[a];
        const mt_d3 ra = {ra4.x, ra4.y, ra4.z};
        if (!(mt_dist2_pbc(rk, ra, p.wnl) < (double)p.rc2f)) continue;
        const mt_d3 ri = mt_move_same_side(rk, ra, p.wbox);
        float u = 0.0f;
        const float r = (float)mt_norm(mt_sub(rk, ri));
        if (!(r < 1E-3f)) {
            const int col = d.type[a];
            const int nff = d.nbcnt[row * T + col];
            float mdU = 0.0f;
            for (int ff = 0; ff < nff; ff++) {
                const float4* tab = d.nbtab + ((size_t)(row * T + col) * MT_MAXFF + ff) * MT_NPTS;
                mt_table_eval(tab, r, WANT_E, mdU, u);
            }
            F = mt_add(F, mt_mul(mt_coeffs12(ri, rk), (double)mdU));
        }
        if (COUL) {
            float uc = 0.0f;
            // i > k selects the energy half; pass (a, -1) so the energy is always formed
            mt_d3 fc = mt_coulomb_pair(rk, ri, qk, (float)ra4.w, -1, 0, uc);
            F = mt_add(F, fc);
            u += uc;
        }
        if (WANT_E) U += 0.5f * u;
    }
    mt_lane_sum<G>(F.x, F.y, F.z, U);
    if (g != 0) return;
    if (!d.mobile[k]) { F = {0.0, 0.0, 0.0}; U = 0.0f; }
    d.frc[k] = make_double4(F.x, F.y, F.z, 0.0);
    if (WANT_E) d.epot[k] = U;
}

// geometry of the MIXED all-pairs term: box widths as the reference's
// minimum image sees them (float L), their halves, and the band around the
// image boundary where the image is decided in FP64 instead
struct MtApGeom {
    float w[3], iw[3], hw[3];
    float edge;
};

__device__ __forceinline__ MtApGeom mt_ap_geom(const MtParams& p)
{
    MtApGeom g;
    float wmax = 0.f;
    for (int a = 0; a < 3; a++) {
        g.w[a] = p.Lf[a];
        g.iw[a] = 1.0f / p.Lf[a];
        g.hw[a] = 0.5f * p.Lf[a];
        wmax = fmaxf(wmax, p.Lf[a]);
    }
    // far above the FP32 rounding of a coordinate difference (~1e-7 * w)
    g.edge = 1.0e-5f * wmax;
    return g;
}

// a position within 0.75 w - edge of the box centre on every axis: two such
// positions are less than 3w/2 - edge apart, with the rounding of their FP32
// difference to spare
__device__ __forceinline__ bool mt_ap_inner(float x, float y, float z, const MtApGeom& g)
{
    return fabsf(x) <= 1.5f * g.hw[0] - g.edge && fabsf(y) <= 1.5f * g.hw[1] - g.edge &&
           fabsf(z) <= 1.5f * g.hw[2] - g.edge;
}

// One pair term of the MIXED all-pairs Coulomb sum, added to (fx, fy, fz, u):
// F_k += K q_k q_i d / |d|^3, u += K q_k q_i / |d|, d = r_k - image(r_i), in
// FP32 from FP32 positions, the image taken by rounding d / w. Within `edge` of
// the image boundary (|d| = w/2, or 3w/2 for positions far outside the box)
// the image is decided in FP64 from the double positions with the reference's
// rule (C3DVector::moveToSameSideOfPBCAsThis, 3DVector.cpp:133-165), so FP32
// rounding never picks the other image of a pair than the FP64 engine does.
// Contraction is off and the FMAs are explicit: the special-pair pass
// recomputes this term for subtraction and must get the same bits.
// FAR = false leaves out the test for |d| near 3w/2, for pairs whose positions
// are both within 0.75 w - edge of the box centre (mt_ap_inner), where it
// cannot hold: the same bits as with it.
template <bool FAR = true>
__device__ __forceinline__ void mt_ap_term(float xk, float yk, float zk, const mt_d3& rk, float qks, float4 q, int i,
                                           const double4* pos, const MtParams& p, const MtApGeom& g, float& fx,
                                           float& fy, float& fz, float& u)
{
#pragma clang fp contract(off)
    float dx = xk - q.x, dy = yk - q.y, dz = zk - q.z;
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    bool nearb = fabsf(ax - g.hw[0]) < g.edge || fabsf(ay - g.hw[1]) < g.edge || fabsf(az - g.hw[2]) < g.edge;
    if constexpr (FAR)
        nearb = nearb || ax > 3.0f * g.hw[0] - g.edge || ay > 3.0f * g.hw[1] - g.edge || az > 3.0f * g.hw[2] - g.edge;
    if (nearb) {
        const double4 r4 = pos[i];
        const mt_d3 ri = mt_move_same_side(rk, {r4.x, r4.y, r4.z}, p.wbox);
        dx = (float)(rk.x - ri.x);
        dy = (float)(rk.y - ri.y);
        dz = (float)(rk.z - ri.z);
    } else {
        dx = fmaf(-g.w[0], rintf(dx * g.iw[0]), dx);
        dy = fmaf(-g.w[1], rintf(dy * g.iw[1]), dy);
        dz = fmaf(-g.w[2], rintf(dz * g.iw[2]), dz);
    }
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float rinv = r2 > 0.0f ? __builtin_amdgcn_rsqf(r2) : 0.0f;   // the atom itself: 0
    const float kr = (qks * q.w) * rinv;
    const float sc = (kr * rinv) * rinv;
    fx = fmaf(dx, sc, fx);
    fy = fmaf(dy, sc, fy);
    fz = fmaf(dz, sc, fz);
    u += kr;
}

// bonded terms and special (same-molecule / scaled) pairs of atom k with the
// exact FP64 helpers, added on top of its pair force into Fk (with its
// energy, want_e); false, and Fk untouched, when the atom has none
__device__ __forceinline__ bool mt_fast_bonded_atom(const MtDev& d, const MtParams& p, int k, int want_e, int coul,
                                                    int mixed, double4& Fk)
{
    if (!d.mobile[k]) return false;
    const int o = d.perm[k];
    const MtBonded& b = d.bl;
    const bool has_b = (b.has_bonds && b.b_num[o]) || (b.has_angles && b.a_num[o]) ||
                       (b.has_dihedrals && b.d_num[o]);
    const bool has_s = d.sp.n_entries > 0 && d.sp.num[o] > 0;
    if (!has_b && !has_s) return false;
    mt_d3 F = {0.0, 0.0, 0.0};
    float U = 0.0f;
    if (has_b) mt_exact_bonded(d, p, k, F, U, o, d.inv);
    if (has_s) {
        const double4 rk4 = d.pos[k];
        const mt_d3 rk = {rk4.x, rk4.y, rk4.z};
        const int row = d.type[k], T = d.T;
        for (int j = 0; j < d.sp.num[o]; j++) {
            const int po = d.sp.partner[d.sp.first[o] + j];
            const float sc = d.sp.scale[d.sp.first[o] + j];
            const int a = d.inv[po];
            const double4 ra4 = d.pos[a];
            const mt_d3 ra = {ra4.x, ra4.y, ra4.z};
            const bool inside = mt_dist2_pbc(rk, ra, p.wnl) < (double)p.rc2f;
            const bool coul_here = coul && (p.coulomb_allpairs || inside);
            if (!inside && !coul_here) continue;
            const mt_d3 ri = mt_move_same_side(rk, ra, p.wbox);
            float u = 0.0f;
            mt_d3 f = {0.0, 0.0, 0.0};
            const float r = (float)mt_norm(mt_sub(rk, ri));
            if (inside && sc != 0.0f && !(r < 1E-3f)) {
                const int col = d.type[a];
                float mdU = 0.0f;
                for (int ff = 0; ff < d.nbcnt[row * T + col]; ff++)
                    mt_table_eval(d.nbtab + ((size_t)(row * T + col) * MT_MAXFF + ff) * MT_NPTS, r, true, mdU, u);
                f = mt_mul(mt_coeffs12(ri, rk), (double)mdU);
            }
            F = mt_add(F, mt_mul(f, (double)sc));
            U += 0.5f * (u * sc);
            if (coul_here) {
                float uc = 0.0f;
                const mt_d3 fc = mt_coulomb_pair(rk, ri, (float)rk4.w, (float)ra4.w, -1, 0, uc);
                if (p.coulomb_allpairs && mixed) {
                    // the all-pairs sum holds this pair's FP32 term with weight 1:
                    // take out exactly that term, put in the scaled FP64 one
                    float tx = 0.f, ty = 0.f, tz = 0.f, tu = 0.f;
                    mt_ap_term((float)rk4.x, (float)rk4.y, (float)rk4.z, rk, MT_COULOMB_K * d.qf[k],
                               make_float4((float)ra4.x, (float)ra4.y, (float)ra4.z, d.qf[a]), a, d.pos, p,
                               mt_ap_geom(p), tx, ty, tz, tu);
                    F = mt_add(F, mt_mul(fc, (double)sc));
                    F = {F.x - (double)tx, F.y - (double)ty, F.z - (double)tz};
                    U += 0.5f * (uc * sc) - 0.5f * tu;
                } else {
                    // FP64 all-pairs holds the pair's FP64 term with weight 1
                    const float cs = p.coulomb_allpairs ? sc - 1.0f : sc;
                    F = mt_add(F, mt_mul(fc, (double)cs));
                    U += 0.5f * (uc * cs);
                }
            }
        }
    }
    const double4 f0 = d.frc[k];
    Fk = make_double4(f0.x + F.x, f0.y + F.y, f0.z + F.z, 0.0);
    if (want_e) d.epot[k] += U;
    return true;
}

// one thread an atom. KICK (NVT steps): then the back half's force cap and
// half kick of every atom and the block partials of p.p/m, as k_cap_kick
// computes them (the same operations on the same values, the same blocks), in
// place of that kernel
template <bool KICK>
__global__ void __launch_bounds__(TPB) k_fast_bonded_special(MtDev d, MtParams p, int want_e, int coul,
                                                             int mixed)
{
    MT_EXACT_FP
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0;
    if (k < d.n) {
        double4 F;
        const bool had = mt_fast_bonded_atom(d, p, k, want_e, coul, mixed, F);
        if constexpr (!KICK) {
            if (had) d.frc[k] = F;
        } else {
            if (!had) F = d.frc[k];
            const double Fcut = p.Fcut;
            if (Fcut > 0) {
                if (fabs(F.x) > Fcut) F.x = ((F.x >= 0.0) ? 1.0 : -1.0) * Fcut;
                if (fabs(F.y) > Fcut) F.y = ((F.y >= 0.0) ? 1.0 : -1.0) * Fcut;
                if (fabs(F.z) > Fcut) F.z = ((F.z >= 0.0) ? 1.0 : -1.0) * Fcut;
            }
            if (had || Fcut > 0) d.frc[k] = F;
            double4 pm = d.mom[k];
            const double dt_2 = p.dt / 2.0;
            pm.x = pm.x + F.x * dt_2; pm.y = pm.y + F.y * dt_2; pm.z = pm.z + F.z * dt_2;
            d.mom[k] = pm;
            ke = ((pm.x * pm.x) + (pm.y * pm.y) + (pm.z * pm.z)) / pm.w;
        }
    }
    if constexpr (KICK) mt_block_sum_store<TPB>(ke, &d.red[blockIdx.x]);
}

// ---------------------------------------------------------------------------
// all-pairs minimum-image Coulomb for the sorted engines (the reference's
// Coulomb, FunctorCalcForce.cpp:258-278, 514-535). Each workgroup owns 256
// atoms k and one of `nsplit` slices of the atoms i, staged through LDS;
// slice sums land in part[split][k] and k_ap_sum adds them in split order, so
// the result does not depend on scheduling. Every pair is summed here with
// weight 1, same-molecule pairs and bonded partners included: the special-pair
// pass adds (scale - 1) times their Coulomb term, which keeps exclusion tests
// out of the O(N^2) loop.
// ---------------------------------------------------------------------------
#define AP_TPB 256

__global__ void __launch_bounds__(AP_TPB) k_ap_coulomb_mixed(MtDev d, MtParams p, int want_e)
{
    __shared__ float4 tq[AP_TPB];
    const int t = threadIdx.x;
    const int k = blockIdx.x * AP_TPB + t;
    const bool act = k < d.n;
    const int nsplit = gridDim.y, split = blockIdx.y;
    const int per = (((d.n + nsplit - 1) / nsplit) + AP_TPB - 1) / AP_TPB * AP_TPB;
    const int i0 = split * per, i1 = min(d.n, i0 + per);
    const MtApGeom g = mt_ap_geom(p);
    float xk = 0.f, yk = 0.f, zk = 0.f, qks = 0.f;
    mt_d3 rk = {0.0, 0.0, 0.0};
    if (act) {
        const double4 r4 = d.pos[k];
        rk = {r4.x, r4.y, r4.z};
        xk = (float)r4.x; yk = (float)r4.y; zk = (float)r4.z;
        qks = MT_COULOMB_K * d.qf[k];
    }
    double Fx = 0.0, Fy = 0.0, Fz = 0.0, U = 0.0;
    // the wavefront's own atoms all inside 0.75 w (uniform over the wavefront)
    const bool own_in = __all(!act || mt_ap_inner(xk, yk, zk, g));
    for (int base = i0; base < i1; base += AP_TPB) {
        __syncthreads();
        const int i = base + t;
        bool in = true;
        if (i < i1) {
            const double4 q = d.pos[i];
            tq[t] = make_float4((float)q.x, (float)q.y, (float)q.z, d.qf[i]);
            in = mt_ap_inner(tq[t].x, tq[t].y, tq[t].z, g);
        } else {
            tq[t] = make_float4(xk, yk, zk, 0.f);       // zero charge: no contribution
        }
        // the tile's atoms all inside 0.75 w (uniform over the workgroup)
        const bool tile_in = __syncthreads_and(in);
        if (qks == 0.0f) continue;
        float fx = 0.f, fy = 0.f, fz = 0.f, u = 0.f;
        const int lim = min(AP_TPB, i1 - base);
        if (tile_in && own_in) {
            // no pair can be near 3w/2: the shorter test, the same bits
#pragma unroll 4
            for (int e = 0; e < AP_TPB; e++) {
                const int ie = e < lim ? base + e : k;   // padded entries: a zero charge at the atom itself
                mt_ap_term<false>(xk, yk, zk, rk, qks, tq[e], ie, d.pos, p, g, fx, fy, fz, u);
            }
        } else {
#pragma unroll 4
            for (int e = 0; e < AP_TPB; e++) {
                const int ie = e < lim ? base + e : k;
                mt_ap_term<true>(xk, yk, zk, rk, qks, tq[e], ie, d.pos, p, g, fx, fy, fz, u);
            }
        }
        Fx += (double)fx; Fy += (double)fy; Fz += (double)fz;
        U += 0.5 * (double)u;
    }
    if (act) d.ap_part[(size_t)split * d.n + k] = make_double4(Fx, Fy, Fz, want_e ? U : 0.0);
}

// FP64 engine: the reference's pair arithmetic (mt_coulomb_pair) on double positions
__global__ void __launch_bounds__(AP_TPB) k_ap_coulomb_fp64(MtDev d, MtParams p, int want_e)
{
    __shared__ double sx[AP_TPB], sy[AP_TPB], sz[AP_TPB];
    __shared__ float sq[AP_TPB];
    const int t = threadIdx.x;
    const int k = blockIdx.x * AP_TPB + t;
    const bool act = k < d.n;
    const int nsplit = gridDim.y, split = blockIdx.y;
    const int per = (((d.n + nsplit - 1) / nsplit) + AP_TPB - 1) / AP_TPB * AP_TPB;
    const int i0 = split * per, i1 = min(d.n, i0 + per);
    mt_d3 rk = {0.0, 0.0, 0.0};
    float qk = 0.f;
    if (act) {
        const double4 r4 = d.pos[k];
        rk = {r4.x, r4.y, r4.z};
        qk = (float)r4.w;
    }
    mt_d3 F = {0.0, 0.0, 0.0};
    float U = 0.0f;
    for (int base = i0; base < i1; base += AP_TPB) {
        __syncthreads();
        const int i = base + t;
        if (i < i1) {
            const double4 q = d.pos[i];
            sx[t] = q.x; sy[t] = q.y; sz[t] = q.z;
            sq[t] = (float)q.w;
        }
        __syncthreads();
        if (!act || qk == 0.0f) continue;
        const int lim = min(AP_TPB, i1 - base);
        for (int e = 0; e < lim; e++) {
            const float qi = sq[e];
            if (qi == 0.0f || base + e == k) continue;
            const mt_d3 ri = mt_move_same_side(rk, {sx[e], sy[e], sz[e]}, p.wbox);
            float u = 0.0f;
            F = mt_add(F, mt_coulomb_pair(rk, ri, qk, qi, -1, 0, u));
            U += 0.5f * u;
        }
    }
    if (act) d.ap_part[(size_t)split * d.n + k] = make_double4(F.x, F.y, F.z, (double)U);
}

__global__ void k_ap_sum(MtDev d, int nsplit, int want_e)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.n || !d.mobile[k]) return;
    double fx = 0.0, fy = 0.0, fz = 0.0, u = 0.0;
    for (int sp = 0; sp < nsplit; sp++) {
        const double4 v = d.ap_part[(size_t)sp * d.n + k];
        fx += v.x; fy += v.y; fz += v.z; u += v.w;
    }
    const double4 f0 = d.frc[k];
    d.frc[k] = make_double4(f0.x + fx, f0.y + fy, f0.z + fz, 0.0);
    if (want_e) d.epot[k] += (float)u;
}

// Lanes per atom for the pair-force kernels: one while the atoms alone fill
// the GPU (32 waves per CU), else 4 or 16 so that the lanes share out the
// atoms' list chunks, keeping at least ~4 chunks per lane. The choice depends
// only on n, the list capacity and the CU count, so it is fixed for a run of
// steps (the per-atom sums, and so the bits, depend on it).
static int mt_lanes_per_atom(int n, int ncap, int ncu, int mixed)
{
    const long long fill = (long long)ncu * 2048;
    const int chunks = (ncap + 3) / 4;
    if ((long long)n >= fill || chunks < 16) return 1;
    if ((long long)n * 4 >= fill || chunks < 64) return 4;
    // MIXED: (the 30k-atom water box at a 10 A cutoff, 159 chunks an atom:
    // its force pass measured 0.061 ms at 8 lanes, 0.064 at 16, 0.073 at 32;
    // the FP64 pass was slower at 8)
    return (mixed && chunks < 256) ? 8 : 16;
}

static int mt_cu_count()
{
    static int ncu = 0;
    if (ncu == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            ncu = v;
        else
            return 256;
    }
    return ncu;
}

#ifndef MT_FORCE_ITERS
#define MT_FORCE_ITERS 1
#endif
template <bool COUL, int G>
static void mt_launch_mixed_lds(const MtDev& d, const MtParams& p, size_t lds_tab, hipStream_t s)
{
    const int bs = 512;
    // each workgroup takes MT_FORCE_ITERS passes of bs / G atoms; the LDS
    // table is loaded once per workgroup
    const long long need = ((long long)d.n * G + bs - 1) / bs;
    const long long per = (need + MT_FORCE_ITERS - 1) / MT_FORCE_ITERS;
    // a multiple of 8 workgroups so every XCD gets the same share
    const int g = (int)std::max<long long>(8, (per + 7) / 8 * 8);
    hipLaunchKernelGGL((k_fast_force_mixed_lds<COUL, G>), dim3(g), dim3(bs), lds_tab, s, d, p);
}

template <bool WANT_E, bool COUL, int G>
static void mt_launch_mixed(const MtDev& d, const MtParams& p, int bs, hipStream_t s)
{
    hipLaunchKernelGGL((k_fast_force_mixed<WANT_E, COUL, G>), dim3(nblk((int)((long long)d.n * G), bs)), dim3(bs), 0,
                       s, d, p);
}

template <bool WANT_E, bool COUL, int G>
static void mt_launch_fp64(const MtDev& d, const MtParams& p, int bs, hipStream_t s)
{
    hipLaunchKernelGGL((k_fast_force_fp64<WANT_E, COUL, G>), dim3(nblk((int)((long long)d.n * G), bs)), dim3(bs), 0, s,
                       d, p);
}

#ifndef MT_TILE_G
#define MT_TILE_G 2
#endif
// the LDS-tiled MIXED kernel (k_fast_force_tile and the block lists it reads):
// sparse cells (the block list build's regime), the coefficient table within
// 32 KB of LDS, and enough 2x2x2 cell blocks to fill the GPU
int mt_fast_tile_mode(const MtDev& d, const double* L, int mixed)
{
    const size_t lds_tab = (size_t)d.T * d.T * MT_NPTS * sizeof(float2);
    if (!mixed || !d.wsorted || d.force_dense || lds_tab > 32768 || d.n > 40 * d.ncells) return 0;
    // the 16-bit list build screens staged coordinates (up to 2.5 cell widths)
    // in FP16: keep them far inside its range
    for (int i = 0; i < 3; i++)
        if (!(2.6 * L[i] / d.nc[i] < 16384.0)) return 0;
    const long long nb = (long long)((d.nc[0] + 1) / 2) * ((d.nc[1] + 1) / 2) * ((d.nc[2] + 1) / 2);
    return nb >= 2LL * mt_cu_count() ? 1 : 0;
}

template <int G>
static void mt_launch_pairs(const MtDev& d, const MtParams& p, int mixed, int want_e, bool coul, int block,
                            size_t lds_tab, hipStream_t s)
{
    if (mixed && d.tile16) {
        // every MIXED pass of these lists, energies included
        const int nb = ((d.nc[0] + 1) / 2) * ((d.nc[1] + 1) / 2) * ((d.nc[2] + 1) / 2);
        const dim3 grid(nb), tpb(256 * MT_TILE_G);
        if (d.tabsym) lds_tab = (size_t)(2 * d.T - 1) * MT_NPTS * sizeof(float2);
        if (want_e) {
            if (coul) hipLaunchKernelGGL((k_fast_force_tile<true, true, MT_TILE_G>), grid, tpb, 0, s, d, p);
            else hipLaunchKernelGGL((k_fast_force_tile<true, false, MT_TILE_G>), grid, tpb, 0, s, d, p);
        } else {
            if (coul) hipLaunchKernelGGL((k_fast_force_tile<false, true, MT_TILE_G>), grid, tpb, lds_tab, s, d, p);
            else hipLaunchKernelGGL((k_fast_force_tile<false, false, MT_TILE_G>), grid, tpb, lds_tab, s, d, p);
        }
        return;
    }
    if (mixed && !want_e && lds_tab <= 65536) {
        if (coul) mt_launch_mixed_lds<true, G>(d, p, lds_tab, s);
        else mt_launch_mixed_lds<false, G>(d, p, lds_tab, s);
    } else if (mixed) {
        const int bs = (block == 64 || block == 128 || block == 256 || block == 512) ? block : 256;
        if (want_e) {
            if (coul) mt_launch_mixed<true, true, G>(d, p, bs, s);
            else mt_launch_mixed<true, false, G>(d, p, bs, s);
        } else {
            if (coul) mt_launch_mixed<false, true, G>(d, p, bs, s);
            else mt_launch_mixed<false, false, G>(d, p, bs, s);
        }
    } else {
        const int bs = (block == 64 || block == 128 || block == 256) ? block : 256;
        if (want_e) {
            if (coul) mt_launch_fp64<true, true, G>(d, p, bs, s);
            else mt_launch_fp64<true, false, G>(d, p, bs, s);
        } else {
            if (coul) mt_launch_fp64<false, true, G>(d, p, bs, s);
            else mt_launch_fp64<false, false, G>(d, p, bs, s);
        }
    }
}

int mt_fast_lanes(const MtDev& d, int mixed) { return mt_lanes_per_atom(d.n, d.ncap, mt_cu_count(), mixed); }

int mt_fast_can_fuse(const MtDev& d, const MtParams& p, int mixed)
{
    const bool ap = p.coulomb_on && p.coulomb_allpairs;
    const bool extra = d.bl.has_bonds || d.bl.has_angles || d.bl.has_dihedrals || d.sp.n_entries > 0;
    return (mixed && d.tile16 && d.posf2 && !ap && !extra) ? 1 : 0;
}

void mt_fast_forces_fused(const MtDev& d, MtParams& p, hipStream_t s)
{
    const int nb = ((d.nc[0] + 1) / 2) * ((d.nc[1] + 1) / 2) * ((d.nc[2] + 1) / 2);
    const size_t lds_tab = (size_t)(d.tabsym ? 2 * d.T - 1 : d.T * d.T) * MT_NPTS * sizeof(float2);
    const dim3 grid(nb), tpb(256 * MT_TILE_G);
    if (p.coulomb_on)
        hipLaunchKernelGGL((k_fast_force_tile<false, true, MT_TILE_G, true>), grid, tpb, lds_tab, s, d, p);
    else
        hipLaunchKernelGGL((k_fast_force_tile<false, false, MT_TILE_G, true>), grid, tpb, lds_tab, s, d, p);
    p.dslot ^= 1;
}

int mt_fast_forces(const MtDev& d, const MtParams& p, int mixed, int want_e, int block, hipStream_t s, int kick)
{
    const bool coul_all = p.coulomb_on != 0;
    const bool ap = coul_all && p.coulomb_allpairs;
    const bool coul = coul_all && !ap;      // Coulomb inside the list kernels
    const size_t lds_tab = (size_t)d.T * d.T * MT_NPTS * sizeof(float2);
    const int G = mt_fast_lanes(d, mixed);
    if (G == 16) mt_launch_pairs<16>(d, p, mixed, want_e, coul, block, lds_tab, s);
    else if (G == 8) mt_launch_pairs<8>(d, p, mixed, want_e, coul, block, lds_tab, s);
    else if (G == 4) mt_launch_pairs<4>(d, p, mixed, want_e, coul, block, lds_tab, s);
    else mt_launch_pairs<1>(d, p, mixed, want_e, coul, block, lds_tab, s);
    if (ap) {
        dim3 g(nblk(d.n, AP_TPB), d.ap_nsplit);
        if (mixed) hipLaunchKernelGGL(k_ap_coulomb_mixed, g, dim3(AP_TPB), 0, s, d, p, want_e);
        else hipLaunchKernelGGL(k_ap_coulomb_fp64, g, dim3(AP_TPB), 0, s, d, p, want_e);
        hipLaunchKernelGGL(k_ap_sum, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d, d.ap_nsplit, want_e);
    }
    if (d.bl.has_bonds || d.bl.has_angles || d.bl.has_dihedrals || d.sp.n_entries > 0) {
        if (kick)
            hipLaunchKernelGGL(k_fast_bonded_special<true>, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d, p, want_e,
                               coul_all ? 1 : 0, mixed);
        else
            hipLaunchKernelGGL(k_fast_bonded_special<false>, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d, p, want_e,
                               coul_all ? 1 : 0, mixed);
        return kick ? 1 : 0;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// integrator front half for the sorted engines: thermostat scale, kick, drift,
// minimum-image displacement since the last rebuild, MIXED local coordinates
// ---------------------------------------------------------------------------
// the per-atom arithmetic of the front half, shared with the fused NVE kernel
// (contraction off in both, so fused and separate steps agree bit for bit)
// displacement since the last rebuild (squared, minimum image) and, MIXED,
// the position relative to the centre of the atom's home cell
// posf_out: where the MIXED local coordinates go (at the atom's place in cell
// order); with_w: the whole float4, the type offset in .w included (a buffer
// the rebuild did not write), else x, y, z only
// q: the atom's reference position of the last rebuild; MIXED: c0 its home
// cell, u its place in cell order and tw (with_w) the bits of type * MT_NPTS,
// as the rebuild stored them in home, sidx and type
__device__ __forceinline__ float mt_fast_locate_at(const MtDev& d, const MtParams& p, const double4& q,
                                                   const double4& r, int mixed, int c0, int u, int tw,
                                                   float4* posf_out, bool with_w)
{
    MT_EXACT_FP
    double dx = r.x - q.x, dy = r.y - q.y, dz = r.z - q.z;
    // nearest image by a multiplication with 1/L: the count differs from
    // rint(dx / L) only where dx / L lies within rounding of a half-integer,
    // and a displacement is far below L / 2 or (across the box) near L
    dx -= p.L[0] * rint(dx * p.invL[0]);
    dy -= p.L[1] * rint(dy * p.invL[1]);
    dz -= p.L[2] * rint(dz * p.invL[2]);
    if (mixed) {
        const int ix = c0 % d.nc[0], iy = (c0 / d.nc[0]) % d.nc[1], iz = c0 / (d.nc[0] * d.nc[1]);
        // the cell centre with cw = L / nc computed once (the same bits)
        double lx = r.x - (-p.L[0] / 2.0 + ((double)ix + 0.5) * p.cw[0]);
        double ly = r.y - (-p.L[1] / 2.0 + ((double)iy + 0.5) * p.cw[1]);
        double lz = r.z - (-p.L[2] / 2.0 + ((double)iz + 0.5) * p.cw[2]);
        // a local coordinate is within half a cell of zero, or near +-L once
        // the atom's wrap has moved it across the box
        lx -= p.L[0] * rint(lx * p.invL[0]);
        ly -= p.L[1] * rint(ly * p.invL[1]);
        lz -= p.L[2] * rint(lz * p.invL[2]);
        if (with_w) {
            posf_out[u] = make_float4((float)lx, (float)ly, (float)lz, __int_as_float(tw));
        } else {
            // x, y, z only: .w (the type offset) stays, and the line is not read
            float* pf = reinterpret_cast<float*>(posf_out + u);
            *reinterpret_cast<float2*>(pf) = make_float2((float)lx, (float)ly);
            pf[2] = (float)lz;
        }
    }
    return (float)(dx * dx + dy * dy + dz * dz);
}

// the same for atom k, its rebuild data read back
__device__ __forceinline__ float mt_fast_locate_atom_to(const MtDev& d, const MtParams& p, int k, const double4& r,
                                                        int mixed, float4* posf_out, bool with_w)
{
    return mt_fast_locate_at(d, p, d.posref[k], r, mixed, mixed ? d.home[k] : 0, mixed ? d.sidx[k] : 0,
                             (mixed && with_w) ? d.type[k] * MT_NPTS : 0, posf_out, with_w);
}

__device__ __forceinline__ float mt_fast_locate_atom(const MtDev& d, const MtParams& p, int k, const double4& r,
                                                     int mixed)
{
    return mt_fast_locate_atom_to(d, p, k, r, mixed, d.posf, false);
}

// half kick and drift
__device__ __forceinline__ void mt_fast_kick_drift_atom(const MtParams& p, double4& pm, double4& r, const double4& F)
{
    MT_EXACT_FP
    const double dt_2 = p.dt / 2.0;
    pm.x += F.x * dt_2; pm.y += F.y * dt_2; pm.z += F.z * dt_2;
    const double c = p.dt / pm.w;
    r.x += pm.x * c; r.y += pm.y * c; r.z += pm.z * c;
}

__device__ __forceinline__ float mt_fast_front_to(const MtDev& d, const MtParams& p, int k, double4& pm, double4& r,
                                                  const double4& F, int mixed, float4* posf_out, bool with_w)
{
    mt_fast_kick_drift_atom(p, pm, r, F);
    return mt_fast_locate_atom_to(d, p, k, r, mixed, posf_out, with_w);
}

__device__ __forceinline__ float mt_fast_front(const MtDev& d, const MtParams& p, int k, double4& pm, double4& r,
                                               const double4& F, int mixed)
{
    return mt_fast_front_to(d, p, k, pm, r, F, mixed, d.posf, false);
}

// The back half of step t (force cap, half kick, PBC wrap; CVelVerlet +
// pbcWrap, VelVerlet.cpp:96-122, SimulationBox.cpp:135-166) for the force F
// of step t, which comes back capped
__device__ __forceinline__ void mt_fast_back_half(const MtParams& p, double4& F, double4& pm, double4& r,
                                                  bool& capped)
{
    MT_EXACT_FP
    capped = false;
    if (p.Fcut > 0) {
        if (fabs(F.x) > p.Fcut) { F.x = ((F.x >= 0.0) ? 1.0 : -1.0) * p.Fcut; capped = true; }
        if (fabs(F.y) > p.Fcut) { F.y = ((F.y >= 0.0) ? 1.0 : -1.0) * p.Fcut; capped = true; }
        if (fabs(F.z) > p.Fcut) { F.z = ((F.z >= 0.0) ? 1.0 : -1.0) * p.Fcut; capped = true; }
    }
    const double dt_2 = p.dt / 2.0;
    pm.x = pm.x + F.x * dt_2; pm.y = pm.y + F.y * dt_2; pm.z = pm.z + F.z * dt_2;
    const double X2 = p.L[0] / 2.0, Y2 = p.L[1] / 2.0, Z2 = p.L[2] / 2.0;
    if (r.x >= X2) mt_pbc_add(r.x, X2, p.L[0]);
    if (r.x < -X2) mt_pbc_add(r.x, -X2, p.L[0]);
    if (r.y >= Y2) mt_pbc_add(r.y, Y2, p.L[1]);
    if (r.y < -Y2) mt_pbc_add(r.y, -Y2, p.L[1]);
    if (r.z >= Z2) mt_pbc_add(r.z, Z2, p.L[2]);
    if (r.z < -Z2) mt_pbc_add(r.z, -Z2, p.L[2]);
}

// the back half of step t and the front half of step t + 1 (half kick,
// drift, displacement, local coordinates) of atom k. Returns the squared
// displacement since the last rebuild.
__device__ __forceinline__ float mt_fast_nve_atom(const MtDev& d, const MtParams& p, int k, double4& F, double4& pm,
                                                  double4& r, bool& capped, int mixed, float4* posf_out, bool with_w)
{
    mt_fast_back_half(p, F, pm, r, capped);
    return mt_fast_front_to(d, p, k, pm, r, F, mixed, posf_out, with_w);
}

// only "has any atom moved beyond skin/2" matters: one atomic per wavefront
// that crossed the threshold, and none once the flag is already raised
__device__ __forceinline__ void mt_fast_flag_disp(const MtDev& d, const MtParams& p, float d2)
{
    for (int off = 32; off > 0; off >>= 1) d2 = fmaxf(d2, __shfl_xor(d2, off));
    unsigned int* f = d.flags + (p.dslot ^ 1);
    if ((threadIdx.x & 63) == 0 && d2 > p.half_skin2 && !(__uint_as_float(*f) > p.half_skin2))
        atomicMax(f, __float_as_uint(d2));
}

// NVE steps inside one call: the back half of step t (force cap, half kick,
// PBC wrap; CVelVerlet + pbcWrap, VelVerlet.cpp:96-122, SimulationBox.cpp:135-166)
// and the front half of step t + 1 (half kick, drift) in one pass. The
// capped force is written back only where the cap changed it.
__global__ void __launch_bounds__(TPB) k_fast_nve_fused(MtDev d, MtParams p, int mixed)
{
    MT_WRITER_CLEAR(d, p);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    float d2 = 0.0f;
    if (k < d.n) {
        double4 F = d.frc[k], pm = d.mom[k], r = d.pos[k];
        bool capped;
        d2 = mt_fast_nve_atom(d, p, k, F, pm, r, capped, mixed, d.posf, false);
        if (capped) d.frc[k] = F;
        d.mom[k] = pm;
        d.pos[k] = r;
    }
    mt_fast_flag_disp(d, p, d2);
}

__global__ void __launch_bounds__(TPB) k_fast_locate(MtDev d, MtParams p, int mixed)
{
    MT_WRITER_CLEAR(d, p);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    float d2 = 0.0f;
    if (k < d.n) d2 = mt_fast_locate_atom(d, p, k, d.pos[k], mixed);
    mt_fast_flag_disp(d, p, d2);
}

void mt_fast_locate(const MtDev& d, MtParams& p, int mixed, hipStream_t s)
{
    hipLaunchKernelGGL(k_fast_locate, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d, p, mixed);
    p.dslot ^= 1;
}

void mt_fast_nve_fused(const MtDev& d, MtParams& p, int mixed, hipStream_t s)
{
    hipLaunchKernelGGL(k_fast_nve_fused, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d, p, mixed);
    p.dslot ^= 1;
}

__global__ void __launch_bounds__(TPB) k_fast_kick_drift(MtDev d, MtParams p, int scale_slot, int mixed)
{
    MT_EXACT_FP
    MT_WRITER_CLEAR(d, p);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    float d2 = 0.0f;
    if (k < d.n) {
        double4 pm = d.mom[k], r = d.pos[k];
        const double4 F = d.frc[k];
        if (scale_slot >= 0) {
            const double S = d.nh[scale_slot];
            pm.x *= S; pm.y *= S; pm.z *= S;
        }
        d2 = mt_fast_front(d, p, k, pm, r, F, mixed);
        d.mom[k] = pm;
        d.pos[k] = r;
    }
    mt_fast_flag_disp(d, p, d2);
}

void mt_fast_kick_drift(const MtDev& d, MtParams& p, int scale_slot, int mixed, hipStream_t s)
{
    hipLaunchKernelGGL(k_fast_kick_drift, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d, p, scale_slot, mixed);
    p.dslot ^= 1;
}

// NVT between two steps of one call: the end of step t (momenta scaled by
// nh[sb], positions wrapped as k_scale_wrap wraps them, SimulationBox.cpp:135-166)
// and the front half of step t + 1 (momenta scaled by nh[sf], kick, drift),
// the same operations in the same order as k_scale_wrap then k_fast_kick_drift
__global__ void __launch_bounds__(TPB) k_fast_scale_kick_drift(MtDev d, MtParams p, int sb, int sf, int mixed)
{
    MT_EXACT_FP
    MT_WRITER_CLEAR(d, p);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    float d2 = 0.0f;
    if (k < d.n) {
        double4 pm = d.mom[k], r = d.pos[k];
        const double4 F = d.frc[k];
        const double X2 = p.L[0] / 2.0, Y2 = p.L[1] / 2.0, Z2 = p.L[2] / 2.0;
        if (r.x >= X2) mt_pbc_add(r.x, X2, p.L[0]);
        if (r.x < -X2) mt_pbc_add(r.x, -X2, p.L[0]);
        if (r.y >= Y2) mt_pbc_add(r.y, Y2, p.L[1]);
        if (r.y < -Y2) mt_pbc_add(r.y, -Y2, p.L[1]);
        if (r.z >= Z2) mt_pbc_add(r.z, Z2, p.L[2]);
        if (r.z < -Z2) mt_pbc_add(r.z, -Z2, p.L[2]);
        const double Sb = d.nh[sb], Sf = d.nh[sf];
        pm.x *= Sb; pm.y *= Sb; pm.z *= Sb;
        pm.x *= Sf; pm.y *= Sf; pm.z *= Sf;
        d2 = mt_fast_front(d, p, k, pm, r, F, mixed);
        d.mom[k] = pm;
        d.pos[k] = r;
    }
    mt_fast_flag_disp(d, p, d2);
}

void mt_fast_scale_kick_drift(const MtDev& d, MtParams& p, int sb, int sf, int mixed, hipStream_t s)
{
    hipLaunchKernelGGL(k_fast_scale_kick_drift, dim3(nblk(d.n, TPB)), dim3(TPB), 0, s, d, p, sb, sf, mixed);
    p.dslot ^= 1;
}

// ---------------------------------------------------------------------------
// U_cut = sum over the pairs inside rc of U(rc) of the pair's force fields,
// scaled like the pair. The reference's pair potentials are truncated without
// a shift, so U jumps by U(rc) whenever a pair crosses the cutoff while the
// force does not: Etot - U_cut has no such jumps and is the energy velocity
// Verlet conserves for these forces, so its drift measures the integration
// (time step, precision) rather than the truncation. Every engine; the EXACT
// lists hold plain atom indices at rc, the FP64/MIXED lists coded words at
// rc + skin plus the special pairs.
// ---------------------------------------------------------------------------
// the cutoff Coulomb energy of a pair at the cutoff, K qk qa / rc in FP32 as
// the pair term forms it (FunctorCalcForce.cpp:258-278); 0 unless Coulomb is
// on and truncated at rc
__device__ __forceinline__ float mt_ucut_coul(const MtParams& p, float qk, float qa)
{
    return (p.coulomb_on && !p.coulomb_allpairs) ? MT_COULOMB_K * qk * qa / p.rc : 0.0f;
}

// the special pairs (1-2/1-3/1-4 and same-molecule pairs) of slot k inside rc
__device__ __forceinline__ double mt_ucut_special(const MtDev& d, const MtParams& p, int k, const mt_d3& rk, int row,
                                                  float qk)
{
    double u = 0.0;
    if (d.sp.n_entries > 0) {
        const int o = d.perm[k];
        for (int j = 0; j < d.sp.num[o]; j++) {
            const int a = d.inv[d.sp.partner[d.sp.first[o] + j]];
            const double4 ra4 = d.pos[a];
            if (mt_dist2_pbc(rk, {ra4.x, ra4.y, ra4.z}, p.wnl) < (double)p.rc2f)
                u += 0.5 * (double)((d.ecut[row * d.T + d.type[a]] + mt_ucut_coul(p, qk, (float)ra4.w)) *
                                    d.sp.scale[d.sp.first[o] + j]);
        }
    }
    return u;
}

__device__ __forceinline__ void mt_block_sum_to(double u, double* out)
{
    __shared__ double sh[TPB];
    sh[threadIdx.x] = u;
    __syncthreads();
    for (int off = TPB / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

__global__ void __launch_bounds__(TPB) k_ucut(MtDev d, MtParams p, int exact, double* __restrict__ red)
{
    double u = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < d.n; k += gridDim.x * blockDim.x) {
        if (!d.mobile[k]) continue;
        const double4 rk4 = d.pos[k];
        const mt_d3 rk = {rk4.x, rk4.y, rk4.z};
        const float qk = (float)rk4.w;
        const int T = d.T, row = d.type[k], molk = d.mol[k];
        const int cnt = min(d.ncount[k], d.ncap);
        for (int j = 0; j < cnt; j++) {
            const int a = exact ? (int)d.nlist[(size_t)j * d.n + k]
                                : mt_id_slot(d, (int)(mt_fast_word(d.nlist, d.n, k, j) & MT_NW_IDX_MASK));
            const double4 ra4 = d.pos[a];
            float e = d.ecut[row * T + d.type[a]] + mt_ucut_coul(p, qk, (float)ra4.w);
            if (exact) {
                mt_d3 f = {0.0, 0.0, 0.0};
                mt_scale_pair(d, p, f, e, a, molk, k);
            } else if (!(mt_dist2_pbc(rk, {ra4.x, ra4.y, ra4.z}, p.wnl) < (double)p.rc2f)) {
                continue;
            }
            u += 0.5 * (double)e;
        }
        if (!exact) u += mt_ucut_special(d, p, k, rk, row, qk);
    }
    mt_block_sum_to(u, red + blockIdx.x);
}

// type pairs whose U(rc) k_ucut_tile keeps in LDS; the tile mode's 32 KB
// coefficient table allows at most 4 (two types)
#define MT_UCUT_TT 64

// the same over the MIXED block lists (d.tile16): workgroups take the 2x2x2
// blocks g = blockIdx.x, blockIdx.x + gridDim.x, ...; a block on 16-bit words
// has the FP64 positions and types of its staged cells copied to LDS, so a
// word is two LDS reads
__global__ void __launch_bounds__(TPB) k_ucut_tile(MtDev d, MtParams p, double* __restrict__ red)
{
    __shared__ double s_x[FT_CAP], s_y[FT_CAP], s_z[FT_CAP];
    __shared__ int s_ty[FT_CAP];
    __shared__ float s_q[FT_CAP];
    __shared__ int s_first[64], s_cnt[64], s_off[65];
    __shared__ int s_hfirst[8], s_hoff[9];
    __shared__ float s_ecut[MT_UCUT_TT];
    const int t = threadIdx.x;
    const int bx = (d.nc[0] + 1) >> 1, by = (d.nc[1] + 1) >> 1, bz = (d.nc[2] + 1) >> 1;
    const int nb = bx * by * bz;
    for (int i = t; i < d.T * d.T; i += blockDim.x) s_ecut[i] = d.ecut[i];
    double u = 0.0;
    for (int g = blockIdx.x; g < nb; g += gridDim.x) {
        __syncthreads();                              // the previous block's tables are read
        mt_tile_cells(d, 2 * (g % bx), 2 * ((g / bx) % by), 2 * (g / (bx * by)), s_first, s_cnt, s_off, s_hfirst,
                      s_hoff);
        __syncthreads();
        const bool w16 = s_off[64] <= FT_CAP;
        if (w16) {
            // TPB / 64 threads a staged cell
            const int per = TPB >> 6, q = t / per;
            for (int m = t % per; m < s_cnt[q]; m += per) {
                const int a = d.cell_list[s_first[q] + m], i = s_off[q] + m;
                const double4 r = d.pos[a];
                s_x[i] = r.x;
                s_y[i] = r.y;
                s_z[i] = r.z;
                s_ty[i] = d.type[a];
                s_q[i] = (float)r.w;
            }
            __syncthreads();
        }
        const int nh = s_hoff[8];
        for (int v = t; v < nh; v += blockDim.x) {
            const int h = mt_tile_home(s_hoff, v);
            const int k = d.cell_list[s_hfirst[h] + (v - s_hoff[h])];
            if (!d.mobile[k]) continue;
            const double4 rk4 = d.pos[k];
            const mt_d3 rk = {rk4.x, rk4.y, rk4.z};
            const float qk = (float)rk4.w;
            const int row = d.type[k];
            const int cnt = min(d.ncount[k], w16 ? 2 * d.ncap : d.ncap);
            if (w16) {
                // a chunk of eight words a load; U(rc) of the pair's types from LDS
                const uint4* __restrict__ nl4 = reinterpret_cast<const uint4*>(d.nlist) + k;
                for (int j0 = 0; j0 < cnt; j0 += 8) {
                    const uint4 w = nl4[(size_t)(j0 >> 3) * d.n];
                    const uint32_t wv[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                    for (int i = 0; i < 8; i++) {
                        if (j0 + i >= cnt) break;
                        const int l = (int)((wv[i >> 1] >> (16 * (i & 1))) & 0xffffu);
                        if (mt_dist2_pbc(rk, {s_x[l], s_y[l], s_z[l]}, p.wnl) < (double)p.rc2f)
                            u += 0.5 * (double)(s_ecut[row * d.T + s_ty[l]] + mt_ucut_coul(p, qk, s_q[l]));
                    }
                }
            } else {
                for (int j = 0; j < cnt; j++) {
                    const int a = mt_id_slot(d, (int)(mt_fast_word(d.nlist, d.n, k, j) & MT_NW_IDX_MASK));
                    const double4 ra4 = d.pos[a];
                    if (mt_dist2_pbc(rk, {ra4.x, ra4.y, ra4.z}, p.wnl) < (double)p.rc2f)
                        u += 0.5 * (double)(d.ecut[row * d.T + d.type[a]] + mt_ucut_coul(p, qk, (float)ra4.w));
                }
            }
            u += mt_ucut_special(d, p, k, rk, row, qk);
        }
    }
    mt_block_sum_to(u, red + blockIdx.x);
}

int mt_ucut_blocks(const MtDev& d, int exact)
{
    if (!exact && d.tile16) {
        const int nb = ((d.nc[0] + 1) / 2) * ((d.nc[1] + 1) / 2) * ((d.nc[2] + 1) / 2);
        return std::max(1, std::min(nb, mt_reduce_blocks(d.n)));
    }
    return mt_reduce_blocks(d.n);
}

int mt_ucut_to(const MtDev& d, const MtParams& p, int exact, double* red, hipStream_t s)
{
    const int g = mt_ucut_blocks(d, exact);
    if (!exact && d.tile16)
        hipLaunchKernelGGL(k_ucut_tile, dim3(g), dim3(TPB), 0, s, d, p, red);
    else
        hipLaunchKernelGGL(k_ucut, dim3(g), dim3(TPB), 0, s, d, p, exact, red);
    return g;
}

int mt_ucut(const MtDev& d, const MtParams& p, int exact, hipStream_t s) { return mt_ucut_to(d, p, exact, d.red, s); }

// The lists of the last build in the caller's atom indices: row perm[k] of
// `out` (max_per_atom entries) receives atom k's neighbours in list order,
// cnt[perm[k]] its word count. Workgroups take the 2x2x2 blocks as the list
// build does, so a block's word format (16-bit staged places or 32-bit coded
// words) is decided from the same cell counts; a 16-bit word is mapped back
// through the block's staged cells to its cell-order place, then its slot.
__global__ void __launch_bounds__(TPB) k_fast_decode_lists(MtDev d, int* __restrict__ out, int* __restrict__ cnt_out,
                                                           int maxp)
{
    __shared__ int s_first[64], s_cnt[64], s_off[65];
    __shared__ int s_hfirst[8], s_hoff[9];
    const int t = threadIdx.x;
    const int bx = (d.nc[0] + 1) >> 1, by = (d.nc[1] + 1) >> 1, bz = (d.nc[2] + 1) >> 1;
    const int nb = bx * by * bz;
    for (int g = blockIdx.x; g < nb; g += gridDim.x) {
        __syncthreads();                              // the previous block's tables are read
        mt_tile_cells(d, 2 * (g % bx), 2 * ((g / bx) % by), 2 * (g / (bx * by)), s_first, s_cnt, s_off, s_hfirst,
                      s_hoff);
        __syncthreads();
        const bool w16 = d.tile16 && s_off[64] <= FT_CAP;
        const int nh = s_hoff[8];
        for (int v = t; v < nh; v += blockDim.x) {
            const int h = mt_tile_home(s_hoff, v);
            const int k = d.cell_list[s_hfirst[h] + (v - s_hoff[h])];
            const int ok = d.perm[k];
            const int c = d.ncount[k];
            cnt_out[ok] = c;
            const int m = min(min(c, w16 ? 2 * d.ncap : d.ncap), maxp);
            int* row = out + (size_t)ok * maxp;
            for (int j = 0; j < m; j++) {
                int slot;
                if (w16) {
                    const uint32_t w = d.nlist[(((size_t)(j >> 3) * d.n + k) << 2) + ((j & 7) >> 1)];
                    const int l = (int)((w >> (16 * (j & 1))) & 0xffffu);
                    int lo = 0, hi = 64;
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (s_off[mid] <= l) lo = mid; else hi = mid;
                    }
                    slot = d.cell_list[s_first[lo] + (l - s_off[lo])];
                } else {
                    slot = mt_id_slot(d, (int)(mt_fast_word(d.nlist, d.n, k, j) & MT_NW_IDX_MASK));
                }
                row[j] = d.perm[slot];
            }
        }
    }
}

void mt_fast_decode_lists(const MtDev& d, int* out, int* cnt, int maxp, hipStream_t s)
{
    const int nb = ((d.nc[0] + 1) / 2) * ((d.nc[1] + 1) / 2) * ((d.nc[2] + 1) / 2);
    hipLaunchKernelGGL(k_fast_decode_lists, dim3(std::max(1, std::min(nb, 4096))), dim3(TPB), 0, s, d, out, cnt,
                       maxp);
}

__global__ void k_fast_clear(unsigned int* flags) { flags[0] = 0u; flags[1] = 0u; }

void mt_fast_clear_flags(const MtDev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_fast_clear, dim3(1), dim3(1), 0, s, d.flags);
}

// plain streaming pass over the neighbour-list buffer, read and written back
// in place in 16-byte words (ncap * n * 4 bytes each way): the reference point
// for the HBM byte counters and the achievable streaming bandwidth. The buffer
// is far larger than the caches, so every byte crosses HBM.
__global__ void __launch_bounds__(TPB) k_copy_u4(uint4* __restrict__ buf, size_t nwords)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x) {
        uint4 v = buf[i];
        // opaque to the compiler, so the write-back of the same value is kept
        asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
        buf[i] = v;
    }
}

void mt_copy_lists(const MtDev& d, hipStream_t s)
{
    const size_t nwords = (size_t)d.ncap * d.n / 4;
    const int g = (int)std::min<size_t>((nwords + TPB - 1) / TPB, 65536);
    hipLaunchKernelGGL(k_copy_u4, dim3(g), dim3(TPB), 0, s, reinterpret_cast<uint4*>(d.nlist), nwords);
}
