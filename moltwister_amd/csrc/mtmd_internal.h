// mtmd_internal.h - device data layout and kernel launchers of libmtmd.so.
//
// Layout in HBM (structure of arrays, one entry per atom unless noted):
//   pos    double4  x, y, z, q            (q: the atom's float charge, widened)
//   mom    double4  px, py, pz, m
//   frc    double4  Fx, Fy, Fz, 0         (F == Fpi in every engine)
//   epot   float    per-atom potential energy
//   type, mol int32; mobile uint8
//   posf   float4   xl, yl, zl, bits of type*NPTS, in CELL ORDER (entry u is
//                   the atom cell_list[u] of the last rebuild): position
//                   relative to the centre of the atom's home cell, written
//                   at each rebuild and by the integrator's front half
//   home   int32    home cell of the atom at the last neighbour rebuild
//   sidx   int32    the atom's place u in cell order; qfs float: charges in
//                   cell order
// Tables (float4 per point, 16-byte aligned):
//   nbtab  [((row*T + col)*MAXFF + ff)*NPTS + pt]  x, -dU/dx, U, 0
//   nbcnt  [row*T + col]
//   btab/atab/dtab [type*NPTS + pt]
//   nbcoef [(row*T + col)*NPTS + pt]   a_f, b_f, a_e, b_e of the segment
//          starting at pt, summed over the pair's force fields (MIXED)
// Neighbour lists are column-major: nl[j*ncap_stride + k], so the lanes of a
// wavefront (consecutive atoms) read consecutive words at every j. FP64/MIXED
// lists are chunked (4 words per 16-byte chunk) and their words carry a 5-bit
// cell-offset code above a 27-bit index: the neighbour's slot (FP64) or its
// place in cell order (MIXED).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MT_NPTS 1000
#define MT_MAXFF 5
#define MT_COULOMB_K 1389.32f
#define MT_CONV_T 120.272
#define MT_CONV_t 1.0E2
#define MT_CONV_PRESS 16388.2461

// neighbour word of the FP64/MIXED lists: atom index in the low 27 bits, the
// home-cell offset code (dx+1) + 3*(dy+1) + 9*(dz+1) in the top 5 bits
// atoms the block list build stages in LDS at once: a cell holding more than
// this (the tile-mode figure when d.tile16 is set) cannot be listed (flags[5])
// MT_FT_CAP: the most atoms the 4x4x4 staged cells of a 2x2x2 block may hold
// for the block's lists to take 16-bit words (MIXED tile mode)
#ifndef MT_FT_CAP
#define MT_FT_CAP 2048
#endif
#define MT_NB_STAGE_ATOMS 2048
#define MT_NB_STAGE_ATOMS_TILE (MT_FT_CAP / 2)
#define MT_NW_IDX_BITS 27
#define MT_NW_IDX_MASK ((1u << MT_NW_IDX_BITS) - 1u)

struct MtBonded {
    // per-atom CSR in atom (original) index order, reference order inside a row
    int* b_first; int* b_num; int* b_j; int* b_type;
    int* a_first; int* a_num; int* a_i; int* a_j; int* a_type; uint8_t* a_center;
    int* d_first; int* d_num; int* d_i; int* d_j; int* d_l; int* d_type; uint8_t* d_center;
    int has_bonds, has_angles, has_dihedrals;
};

struct MtSpecial {
    // pairs evaluated outside the neighbour list by the FP64/MIXED engines:
    // per atom (original index) CSR of (partner original index, scale)
    int* first; int* num; int* partner; float* scale;
    int n_entries;
};

struct MtDev {
    int n, T;
    double4* pos; double4* mom; double4* frc; float* epot;
    int* type; int* mol; uint8_t* mobile; uint8_t* xmol;  // xmol: bonded partner outside own molecule
    int has_molecules;     // some molecule has more than one atom
    int has_xmol;          // some atom has a bonded partner outside its molecule
    float4* nbtab; int* nbcnt; float4* btab; float4* atab; float4* dtab;
    float4* nbcoef;
    float2* nbcoef2;       // force-only a_f, b_f per segment (MIXED, staged in LDS)
    float2* nbcoef2s;      // the same in symmetric rows (tabsym): row ti + tj, 2T - 1 rows
    int tabsym;            // T <= 2 and the pair coefficients are symmetric in the two types
    float* ecut;           // [row*T + col] pair energy at the cutoff, summed over force fields
    MtBonded bl;
    MtSpecial sp;
    // cells
    int nc[3]; int ncells;
    int* cell_key; int* cell_count; int* cell_first; int* cell_fill; int* cell_list;
    int* cell_xyz;         // EXACT: ix, iy, iz per atom
    int* cell_bsum;        // FP64/MIXED: per-block sums of the cell-count scan
    int* cell_next;        // FP64/MIXED: counts of the rebuild in progress, all zero between rebuilds
    // neighbours
    int ncap; int* ncount; uint32_t* nlist; int* nreq;
    // MIXED block lists (mt_fast_tile_mode): the lists of every 2x2x2 block of
    // home cells whose 4x4x4 staged cells hold at most 2048 atoms are 16-bit
    // words, the neighbour's place in those staged cells (8 words a chunk,
    // 2 ncap words a row); the other blocks keep the 32-bit coded words
    int tile16;
    // set once a block list build met a cell larger than its staging (one cell
    // over 2048 atoms in an otherwise sparse system): from then on the lists
    // come from the dense-cell build, which takes cells of any size, and the
    // tile mode is off
    int force_dense;
    // sorting (FP64/MIXED): perm[s] = original index of sorted slot s, inv[o] = slot
    int* perm; int* inv;
    int* perm2;
    double4* pos2; double4* mom2; double4* frc2; float* epot2; int* type2; int* mol2; uint8_t* mobile2;
    float* qf; float* qf2;  // float charges in slot order
    double4* posref;       // positions at the last rebuild
    float4* posf; int* home;
    float4* posf2;         // MIXED tile mode: the local coordinates a fused force pass writes
    // posf is kept in cell order: posf[u] belongs to atom cell_list[u] (u is
    // the atom's place in the cell lists of the last rebuild, sidx[slot] = u),
    // so the atoms of a cell are one contiguous run of posf; qfs = charges in
    // the same order. With wsorted set (MIXED) the neighbour-list words carry u
    // instead of the slot.
    int* sidx; float* qfs; int wsorted;
    double4* ap_part;      // all-pairs Coulomb partial sums [ap_nsplit][n] (FP64/MIXED)
    int ap_nsplit;
    // reductions / scalars
    double* red;           // block partials
    int red_blocks;
    double* nh;            // NH chain state: eta[M], p_eta[M], Q[M], scale, KE2
    unsigned int* flags;   // [0], [1] max displacement^2 (float bits) above the rebuild threshold,
                           // one slot read, the other written (MtParams::dslot), [2] list overflow,
                           // [3] non-finite position, [4] rebuild count, [5] cell over the staging
                           // capacity
};

struct MtParams {
    float Lf[3];           // float(L): box widths as the reference's functors see them
    double L[3];           // double box (integrator, wrap)
    double wbox[3];        // force min-image widths: double(Lf)/2 - (-double(Lf)/2)
    double wnl[3];         // neighbour-search widths: double(low + Lf) - double(low)
    float lowf[3];         // -Lf/2
    float rc, rc2f;        // float cutoff and its float square
    float rlist, rlist2;   // FP64/MIXED list radius
    float scale12, scale13, scale14, scale1N;
    int coulomb_allpairs, coulomb_on;
    double dt, Fcut;
    float x0, inv_delta;   // non-bonded table start and 1/spacing (MIXED)
    float cs[3];           // cell size per axis (MIXED local coordinates)
    double cw[3];          // the same in double, L / nc (cell centres)
    double invL[3];        // 1 / L (nearest-image counts of displacements and local coordinates)
    float half_skin2;      // rebuild threshold on the squared displacement
    int dslot;             // flags[dslot]: the rebuild flag the conditional list kernels read; the
                           // kernels that move atoms clear it and set flags[dslot ^ 1] (the host
                           // flips dslot after launching one of them)
};

// ---- launchers (mtmd_exact.hip) ----
void mt_exact_cells(const MtDev& d, const MtParams& p, hipStream_t s);
void mt_exact_neighbors(const MtDev& d, const MtParams& p, hipStream_t s);
void mt_exact_forces(const MtDev& d, const MtParams& p, hipStream_t s);
void mt_kick_drift(const MtDev& d, const MtParams& p, int scale_slot, int track_disp, hipStream_t s);
void mt_cap_kick(const MtDev& d, const MtParams& p, int want_ke, hipStream_t s);
// fast: the FP64/MIXED form of the chain (mt_nh_propagate_fast), else the
// reference's operation order
void mt_nh_chain(const MtDev& d, int M, int nresp, double T, double dt, double g, const double* w,
                 int nblocks, hipStream_t s, int fast = 0, int from_slot = 0);
// FP64/MIXED NVT between two steps of a call: both chain propagations (end of
// step t into nh[3M], start of t + 1 into nh[3M + 1]) in one launch
void mt_nh_chain2(const MtDev& d, int M, int nresp, double T, double dt, double g, const double* w, int nblocks,
                  hipStream_t s);
void mt_scale_wrap(const MtDev& d, const MtParams& p, int scale_slot, int want_ke, hipStream_t s);
void mt_reduce_ke(const MtDev& d, hipStream_t s);
void mt_reduce_thermo(const MtDev& d, hipStream_t s);
// the same partials (3 a block: sum p.p/m, sum F.r, sum U) into red[0, 3 mt_reduce_blocks(n))
void mt_reduce_thermo_to(const MtDev& d, double* red, hipStream_t s);
void mt_gradient_step(const MtDev& d, double lr, hipStream_t s);
// NPT particle kernels (all engines): partials of sum p.p/m in red[0, nb) and
// sum F.r in red[nb, 2 nb), nb = mt_npt_blocks(n)
int mt_npt_blocks(int n);
void mt_npt_scale_sums(const MtDev& d, int scale_slot, hipStream_t s);
void mt_npt_prop_pr(const MtDev& d, double cp2, double cpf, double cr2, double crr, hipStream_t s);
void mt_npt_prop_p_sums(const MtDev& d, double cp2, double cpf, hipStream_t s);
// one step of the momentum histograms of mtmd_run (hist [4*nbins] ints on the
// device, nbins <= 256; peta NULL skips the thermostat bin); -1 on bad nbins
int mt_pdistr(const MtDev& d, double maxP, int nbins, const double* peta, double maxPeta, int* hist, hipStream_t s);

// ---- launchers (mtmd_fast.hip) ----
// `conditional` kernels return at once unless this step's maximum
// displacement since the last rebuild exceeds half the skin (flags[0] vs p.half_skin2)
// with_home: the cell sort also does the per-atom set-up of the rebuild (home
// cell, reference position, cell-order place, MIXED float position)
void mt_fast_cells(const MtDev& d, const MtParams& p, int conditional, int with_home, hipStream_t s);
void mt_fast_sort_apply(const MtDev& d, hipStream_t s);   // writes the *2 buffers, perm2 and inv
void mt_fast_unsort(const MtDev& d, hipStream_t s);       // writes the *2 buffers in original order
void mt_fast_identity_perm(const MtDev& d, hipStream_t s);
void mt_fast_snapshot(const MtDev& d, double4* sp, double4* sm, hipStream_t s);        // pos, mom by original index
void mt_fast_restore(const MtDev& d, const double4* sp, const double4* sm, hipStream_t s);
void mt_fast_neighbors(const MtDev& d, const MtParams& p, int conditional, hipStream_t s);
// 1 when the MIXED pair forces run on 2x2x2 cell blocks staged in LDS (sparse
// cells, the coefficient table within 32 KB, enough blocks to fill the GPU,
// cell widths of box L within the FP16 range of the list build's screen);
// the host stores it in d.tile16 whenever the cell grid is set, before a build
int mt_fast_tile_mode(const MtDev& d, const double* L, int mixed);
// the tiled force-only pass in FP64 arithmetic (measurement); -1 when the
// tile mode is off or its LDS does not fit
int mt_fast_force_tile64(const MtDev& d, const MtParams& p, hipStream_t s);
// pair, all-pairs Coulomb and bonded forces. kick (NVT): the pass that adds
// the bonded terms also applies the force cap and half kick and leaves the
// block partials of p.p/m (mt_cap_kick's work); returns 1 when it did, 0
// when the system has no bonded or special terms (then call mt_cap_kick)
int mt_fast_forces(const MtDev& d, const MtParams& p, int mixed, int want_energy, int block, hipStream_t s,
                   int kick = 0);
// 1 when the NVE step may run as one fused pass (mt_fast_forces_fused): the
// MIXED tile mode with every force of the step inside that pass
int mt_fast_can_fuse(const MtDev& d, const MtParams& p, int mixed);
// pair forces of the step, then the back half of the step and the front half
// of the next: new positions in pos2 and local coordinates in posf2, which
// the host swaps with pos and posf; flips p.dslot
void mt_fast_forces_fused(const MtDev& d, MtParams& p, hipStream_t s);
// lanes per atom the pair-force kernels use for this system (1, 4, 8 or 16;
// 8 for MIXED only)
int mt_fast_lanes(const MtDev& d, int mixed);
// the kernels that move atoms (front halves) set the rebuild flag: their
// launchers flip p.dslot
void mt_fast_kick_drift(const MtDev& d, MtParams& p, int scale_slot, int mixed, hipStream_t s);
// the end of step t (scale nh[sb], wrap) and the front half of t + 1 (scale nh[sf], kick, drift)
void mt_fast_scale_kick_drift(const MtDev& d, MtParams& p, int sb, int sf, int mixed, hipStream_t s);
void mt_fast_clear_flags(const MtDev& d, hipStream_t s);
// NPT: displacement since the last rebuild and MIXED local coordinates of the
// current positions (the locating half of the front-half kernel)
void mt_fast_locate(const MtDev& d, MtParams& p, int mixed, hipStream_t s);
// NVE: back half of one step and front half of the next in one pass
void mt_fast_nve_fused(const MtDev& d, MtParams& p, int mixed, hipStream_t s);
// in-place read + write of the whole neighbour-list buffer (HBM reference)
void mt_copy_lists(const MtDev& d, hipStream_t s);
// the lists of the last build in original atom indices: out[o * maxp + j]
// (maxp 0: counts only), cnt[o] = word count of atom o
void mt_fast_decode_lists(const MtDev& d, int* out, int* cnt, int maxp, hipStream_t s);
// per-block partial sums of U_cut into d.red[0 .. returned count)
int mt_ucut(const MtDev& d, const MtParams& p, int exact, hipStream_t s);
// the same into red[0 .. returned count); the count is mt_ucut_blocks (<= 1024)
int mt_ucut_to(const MtDev& d, const MtParams& p, int exact, double* red, hipStream_t s);
int mt_ucut_blocks(const MtDev& d, int exact);

int mt_reduce_blocks(int n);
