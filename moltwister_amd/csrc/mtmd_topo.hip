// mtmd_topo.hip - topology search of the MD force fields on the GPU.
//
// Replaces CMolTwisterStateTools::getAllMDBondsInSystem / getAllMDAnglesInSystem
// / getAllMDDihedralsInSystem (Cmd/Tools/MolTwisterStateTools.cpp:544-724,
// 767-920, 993-1190) for the four detection criteria of the MDFF bond, angle
// and dihedral types (MDFF/Bonds/MolTwisterMDFFBond.h:43).
//
// all_r_less_than: the reference tests every atom of the first type against
// every atom of the second (and, for angles and dihedrals, every atom of the
// next type against each hit): O(N_1 N_2) distance tests per force field. Here
// the atoms within the largest R0 of each atom are found once through a cell
// grid, each row sorted by atom index, and the chains j - k (- l (- m)) are
// enumerated from those rows, so the output comes out in the reference's order:
// force field, then j, k, l, m ascending.
//
// mol_r_less_than: the second and later atoms come from the first atom's
// molecule in the depth-first order of CAtom::findAtomsInMolecule
// (MolTwisterAtom.cpp:578-605). One thread per first atom walks the bond graph
// iteratively (an explicit stack replaces the recursion; the visiting order is
// the same) into a per-thread slot that records each reached atom's rank in that
// order. The candidates of a link are the atoms within R0 of the previous atom,
// so each link takes them from the distance row in ascending rank, which is the
// reference's walk over the whole molecule list with the others skipped; a row
// longer than ROW_SCAN falls back to that walk over the slot's order list.
//
// only_visible_bonds and only_14_bonds walk the bond graph (CAtom::bonds_ in the
// caller's order) from each first atom: no distances.
//
// Duplicates are then removed keeping the first occurrence, as
// CMolecularSystemTools::removeDuplicateBondIndices / AngleIndices / DihIndices
// do (Cmd/Tools/MolecularSystemTools.cpp:25-235): a bond equals its reverse, an
// angle (a, b, c) equals (c, b, a), a dihedral (a, b, c, d) equals (d, c, b, a),
// and angles and dihedrals that name an atom twice are dropped.
//
// Distances are the reference's in double: |r_j - r_k|^2, or with box widths
// given, each |component| folded once to w - |d| when it exceeds w/2
// (CAtom::getDistanceTo2UsingPBC, MolTwisterAtom.cpp:374-394). The force-field
// matrices search with the periodic distance (MDFFMatrices.cpp:123).
// Built with -ffp-contract=off.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/mtmd.h"

namespace {

struct TopoDev {
    const double* pos;
    const int* type;
    int n;
    int pbc;
    double w[3];
    // cell grid
    int nc[3];
    double lo[3], cs[3];
    const int* cell_start;     // [ncells + 1]
    const int* cell_atoms;     // atoms sorted by cell
    // neighbour rows within R0max, ascending atom index
    const int* row_start;      // [n + 1]
    int* rows;
    double r0max2;
    // bond graph (CAtom::bonds_ of each atom, in its order)
    const int* bs;             // [n + 1]
    const int* bd;
    // molecules of the undirected graph: id and index inside the molecule
    const int* comp;
    const int* loc;
};

__device__ __forceinline__ double topo_d2(const TopoDev& t, int a, int b)
{
    double dx = t.pos[3 * a] - t.pos[3 * b];
    double dy = t.pos[3 * a + 1] - t.pos[3 * b + 1];
    double dz = t.pos[3 * a + 2] - t.pos[3 * b + 2];
    if (t.pbc) {
        dx = (dx < 0.0) ? -dx : dx;
        dy = (dy < 0.0) ? -dy : dy;
        dz = (dz < 0.0) ? -dz : dz;
        if (dx > (t.w[0] / 2.0)) dx = t.w[0] - dx;
        if (dy > (t.w[1] / 2.0)) dy = t.w[1] - dy;
        if (dz > (t.w[2] / 2.0)) dz = t.w[2] - dz;
    }
    return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ int topo_cell_axis(const TopoDev& t, int ax, double x)
{
    long long c = (long long)floor((x - t.lo[ax]) / t.cs[ax]);
    if (t.pbc) {
        c %= t.nc[ax];
        if (c < 0) c += t.nc[ax];
    } else {
        c = c < 0 ? 0 : (c >= t.nc[ax] ? t.nc[ax] - 1 : c);
    }
    return (int)c;
}

// the distinct neighbour cells along one axis: -1..1 around c, wrapped (pbc)
// or clipped, each cell once when the axis has fewer than three cells
__device__ __forceinline__ int topo_axis_cells(const TopoDev& t, int ax, int c, int out[3])
{
    int m = 0;
    for (int o = -1; o <= 1; o++) {
        int v = c + o;
        if (t.pbc) v = (v + t.nc[ax]) % t.nc[ax];
        else if (v < 0 || v >= t.nc[ax]) continue;
        bool seen = false;
        for (int i = 0; i < m; i++) seen |= (out[i] == v);
        if (!seen) out[m++] = v;
    }
    return m;
}

__global__ void k_topo_cells(TopoDev t, int* cell)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= t.n) return;
    const int cx = topo_cell_axis(t, 0, t.pos[3 * a]);
    const int cy = topo_cell_axis(t, 1, t.pos[3 * a + 1]);
    const int cz = topo_cell_axis(t, 2, t.pos[3 * a + 2]);
    cell[a] = cx + t.nc[0] * (cy + t.nc[1] * cz);
}

// rows of the atoms within R0max; count pass (rows == nullptr) or fill pass
__global__ void k_topo_rows(TopoDev t, int* count)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= t.n) return;
    const int cx = topo_cell_axis(t, 0, t.pos[3 * a]);
    const int cy = topo_cell_axis(t, 1, t.pos[3 * a + 1]);
    const int cz = topo_cell_axis(t, 2, t.pos[3 * a + 2]);
    int xs[3], ys[3], zs[3];
    const int nx = topo_axis_cells(t, 0, cx, xs), ny = topo_axis_cells(t, 1, cy, ys), nz = topo_axis_cells(t, 2, cz, zs);
    int found = 0;
    int* out = (count == nullptr) ? t.rows + t.row_start[a] : nullptr;
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                const int c = xs[ix] + t.nc[0] * (ys[iy] + t.nc[1] * zs[iz]);
                for (int q = t.cell_start[c]; q < t.cell_start[c + 1]; q++) {
                    const int b = t.cell_atoms[q];
                    if (b == a || topo_d2(t, a, b) >= t.r0max2) continue;
                    if (out) out[found] = b;
                    found++;
                }
            }
    if (count) {
        count[a] = found;
        return;
    }
    // ascending atom index (the reference walks the atoms of a type in index order)
    for (int i = 1; i < found; i++) {
        const int v = out[i];
        int j = i - 1;
        while (j >= 0 && out[j] > v) { out[j + 1] = out[j]; j--; }
        out[j + 1] = v;
    }
}

struct Rule {
    int types[4];
    double r02;
};

// the chains j - k (- l (- m)) from atom j of the rule's first type, in the
// reference's loop order; count pass (out == nullptr) or write pass
template <int DEPTH>
__global__ void k_topo_paths(TopoDev t, Rule r, int rule_index, const int* offset, int* count, int* out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= t.n) return;
    int found = 0;
    if (t.type[j] == r.types[0]) {
        int* o = out ? out + (size_t)(DEPTH + 3) * offset[j] : nullptr;   // entries of DEPTH + 2 atoms and the rule
        for (int qk = t.row_start[j]; qk < t.row_start[j + 1]; qk++) {
            const int k = t.rows[qk];
            if (t.type[k] != r.types[1] || !(topo_d2(t, j, k) < r.r02)) continue;
            if (DEPTH == 0) {
                if (o) { o[0] = j; o[1] = k; o[2] = rule_index; o += 3; }
                found++;
                continue;
            }
            for (int ql = t.row_start[k]; ql < t.row_start[k + 1]; ql++) {
                const int l = t.rows[ql];
                if (l == j || t.type[l] != r.types[2] || !(topo_d2(t, k, l) < r.r02)) continue;
                if (DEPTH == 1) {
                    if (o) { o[0] = j; o[1] = k; o[2] = l; o[3] = rule_index; o += 4; }
                    found++;
                    continue;
                }
                for (int qm = t.row_start[l]; qm < t.row_start[l + 1]; qm++) {
                    const int m = t.rows[qm];
                    if (m == k || m == j || t.type[m] != r.types[3] || !(topo_d2(t, l, m) < r.r02)) continue;
                    if (o) { o[0] = j; o[1] = k; o[2] = l; o[3] = m; o[4] = rule_index; o += 5; }
                    found++;
                }
            }
        }
    }
    if (count) count[j] = found;
}

// ---- only_visible_bonds / only_14_bonds: walks along the bond graph ----
// DEPTH + 2 atoms a chain (ONLY14: bonds a - d, with d three bonds from a);
// count pass (out == nullptr) or write pass at offset[a]
template <int DEPTH, bool ONLY14>
__global__ void k_topo_graph(TopoDev t, Rule r, int rule_index, const int* offset, int* count, int* out)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= t.n) return;
    int found = 0;
    if (t.type[a] == r.types[0]) {
        int* o = out ? out + (size_t)(DEPTH + 3) * offset[a] : nullptr;
        for (int qb = t.bs[a]; qb < t.bs[a + 1]; qb++) {
            const int b = t.bd[qb];
            if (!ONLY14) {
                if (t.type[b] != r.types[1]) continue;
                if (DEPTH == 0) {
                    if (o) { o[0] = a; o[1] = b; o[2] = rule_index; o += 3; }
                    found++;
                    continue;
                }
            }
            for (int qc = t.bs[b]; qc < t.bs[b + 1]; qc++) {
                const int c = t.bd[qc];
                if (!ONLY14) {
                    if (t.type[c] != r.types[2]) continue;
                    if (DEPTH == 1) {
                        if (o) { o[0] = a; o[1] = b; o[2] = c; o[3] = rule_index; o += 4; }
                        found++;
                        continue;
                    }
                }
                for (int qd = t.bs[c]; qd < t.bs[c + 1]; qd++) {
                    const int d = t.bd[qd];
                    if (ONLY14) {
                        if (t.type[d] != r.types[1] || d == c || d == b || d == a) continue;
                        if (o) { o[0] = a; o[1] = d; o[2] = rule_index; o += 3; }
                    } else {
                        if (t.type[d] != r.types[3]) continue;
                        if (o) { o[0] = a; o[1] = b; o[2] = c; o[3] = d; o[4] = rule_index; o += 5; }
                    }
                    found++;
                }
            }
        }
    }
    if (count) count[a] = found;
}

// ---- mol_r_less_than: links inside the first atom's molecule ----
// a distance row longer than this is not scanned for the next rank; the slot's
// order list is walked instead (same atoms in the same order)
constexpr int ROW_SCAN = 64;

struct MolSlots {
    int* rank;     // [S * mmax]: rank of each molecule atom in the current walk, -1 outside it
    int* order;    // [S * mmax]: atoms in walk order
    int* stk;      // [S * mmax * 2]: (atom, next bond) of the walk's open atoms
    int mmax;      // atoms in the largest molecule
    int S;         // slots (threads)
};

// CAtom::findAtomsInMolecule (MolTwisterAtom.cpp:578-605) from a: a, then each
// bond destination not yet reached, depth first, in the atoms' bond order.
// Returns the number of atoms reached; their ranks are set in rank[loc].
__device__ int mol_walk(const TopoDev& t, int a, int* rank, int* order, int* stk)
{
    int nm = 0, sp = 0;
    rank[t.loc[a]] = nm;
    order[nm++] = a;
    stk[0] = a;
    stk[1] = t.bs[a];
    while (sp >= 0) {
        const int x = stk[2 * sp], it = stk[2 * sp + 1];
        if (it < t.bs[x + 1]) {
            stk[2 * sp + 1] = it + 1;
            const int y = t.bd[it];
            if (rank[t.loc[y]] < 0) {
                rank[t.loc[y]] = nm;
                order[nm++] = y;
                sp++;
                stk[2 * sp] = y;
                stk[2 * sp + 1] = t.bs[y];
            }
        } else {
            sp--;
        }
    }
    return nm;
}

// the atom of smallest walk rank above `after` that is of type ty, is none of
// e0, e1, e2 and lies closer than sqrt(r02) to x; -1 when there is none
__device__ int mol_next(const TopoDev& t, const int* rank, const int* order, int nm, int comp_a, int x, int after,
                        int ty, int e0, int e1, int e2, double r02)
{
    const int q0 = t.row_start[x], q1 = t.row_start[x + 1];
    if (q1 - q0 <= ROW_SCAN) {
        int best = -1, best_rank = 0x7fffffff;
        for (int q = q0; q < q1; q++) {
            const int y = t.rows[q];
            if (t.comp[y] != comp_a) continue;
            const int ry = rank[t.loc[y]];
            if (ry <= after || ry >= best_rank) continue;   // (after >= -1: atoms outside the walk have -1)
            if (y == e0 || y == e1 || y == e2 || t.type[y] != ty || !(topo_d2(t, x, y) < r02)) continue;
            best = y;
            best_rank = ry;
        }
        return best;
    }
    for (int k = after + 1; k < nm; k++) {
        const int y = order[k];
        if (y == e0 || y == e1 || y == e2 || t.type[y] != ty) continue;
        if (topo_d2(t, x, y) < r02) return y;
    }
    return -1;
}

// one slot a thread; first atoms a = slot, slot + S, ...; count or write pass
template <int DEPTH>
__global__ void k_topo_mol(TopoDev t, Rule r, int rule_index, MolSlots ms, const int* offset, int* count, int* out)
{
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= ms.S) return;
    int* rank = ms.rank + (size_t)slot * ms.mmax;
    int* order = ms.order + (size_t)slot * ms.mmax;
    int* stk = ms.stk + (size_t)slot * ms.mmax * 2;
    for (int a = slot; a < t.n; a += ms.S) {
        if (t.type[a] != r.types[0]) {
            if (count) count[a] = 0;
            continue;
        }
        const int nm = mol_walk(t, a, rank, order, stk);
        const int ca = t.comp[a];
        int found = 0;
        int* o = out ? out + (size_t)(DEPTH + 3) * offset[a] : nullptr;
        for (int b = mol_next(t, rank, order, nm, ca, a, -1, r.types[1], a, a, a, r.r02); b >= 0;
             b = mol_next(t, rank, order, nm, ca, a, rank[t.loc[b]], r.types[1], a, a, a, r.r02)) {
            if (DEPTH == 0) {
                if (o) { o[0] = a; o[1] = b; o[2] = rule_index; o += 3; }
                found++;
                continue;
            }
            for (int c = mol_next(t, rank, order, nm, ca, b, -1, r.types[2], a, b, b, r.r02); c >= 0;
                 c = mol_next(t, rank, order, nm, ca, b, rank[t.loc[c]], r.types[2], a, b, b, r.r02)) {
                if (DEPTH == 1) {
                    if (o) { o[0] = a; o[1] = b; o[2] = c; o[3] = rule_index; o += 4; }
                    found++;
                    continue;
                }
                for (int d = mol_next(t, rank, order, nm, ca, c, -1, r.types[3], a, b, c, r.r02); d >= 0;
                     d = mol_next(t, rank, order, nm, ca, c, rank[t.loc[d]], r.types[3], a, b, c, r.r02)) {
                    if (o) { o[0] = a; o[1] = b; o[2] = c; o[3] = d; o[4] = rule_index; o += 5; }
                    found++;
                }
            }
        }
        if (count) count[a] = found;
        for (int k = 0; k < nm; k++) rank[t.loc[order[k]]] = -1;   // the slot is clean for the next atom
    }
}

struct DevMem {
    std::vector<void*> ptrs;
    template <class T> T* alloc(size_t n)
    {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
    ~DevMem() { for (void* p : ptrs) (void)hipFree(p); }
};

thread_local std::string g_topo_error;

int fail(int code, const std::string& msg)
{
    g_topo_error = msg;
    return code;
}

// exclusive scan of counts into offsets[0..n], on the host
long long scan(const std::vector<int>& c, std::vector<int>& off)
{
    off.assign(c.size() + 1, 0);
    long long s = 0;
    for (size_t i = 0; i < c.size(); i++) {
        off[i] = (int)s;
        s += c[i];
    }
    off[c.size()] = (int)std::min<long long>(s, 0x7fffffff);
    return s;
}

} // namespace

extern "C" const char* mtmd_topology_last_error(void) { return g_topo_error.c_str(); }

extern "C" int mtmd_find_topology(int32_t which, int32_t n, const double* pos, const int32_t* type, const double* box,
                                  const int32_t* bond_start, const int32_t* bond_dest, int32_t n_rules,
                                  const mtmd_topo_rule* rules, int32_t max_out, int32_t* out, int32_t* n_out)
{
    g_topo_error.clear();
    if (which < 1 || which > 3) return fail(MTMD_EINVAL, "which must be 1 (bonds), 2 (angles) or 3 (dihedrals)");
    if (n < 0 || (n > 0 && (!pos || !type)) || n_rules < 0 || (n_rules > 0 && !rules) || !n_out)
        return fail(MTMD_EINVAL, "atoms, types, rules and n_out are required");
    *n_out = 0;
    double r0max = 0.0;
    bool need_rows = false, need_graph = false, need_mol = false;
    for (int i = 0; i < n_rules; i++) {
        const int crit = rules[i].crit;
        if (crit < MTMD_CRIT_ALL_R || crit > MTMD_CRIT_14 || (crit == MTMD_CRIT_14 && which != 1))
            // the reference prints "unsupported determination criteria" and stops the search
            return fail(MTMD_EINVAL, "rule " + std::to_string(i) + ": unsupported detection criterion " +
                                         std::to_string(crit) + (crit == MTMD_CRIT_14 ? " (1-4 is for bonds)" : ""));
        if (crit == MTMD_CRIT_ALL_R || crit == MTMD_CRIT_MOL_R) {
            if (!(rules[i].R0 > 0.0) || !std::isfinite(rules[i].R0)) return fail(MTMD_EINVAL, "R0 must be positive");
            r0max = std::max(r0max, rules[i].R0);
            need_rows = true;
        }
        if (crit != MTMD_CRIT_ALL_R) need_graph = true;
        if (crit == MTMD_CRIT_MOL_R) need_mol = true;
    }
    if (n == 0 || n_rules == 0) return MTMD_OK;
    for (int k = 0; k < 3 * n; k++)
        if (!std::isfinite(pos[k])) return fail(MTMD_EINVAL, "non-finite position");
    int nbond = 0;
    if (need_graph) {
        if (!bond_start || !bond_dest) return fail(MTMD_EINVAL, "the molecule, visible-bond and 1-4 criteria need the bond graph");
        if (bond_start[0] != 0) return fail(MTMD_EINVAL, "bond_start[0] must be 0");
        for (int a = 0; a < n; a++)
            if (bond_start[a + 1] < bond_start[a]) return fail(MTMD_EINVAL, "bond_start must not decrease");
        nbond = bond_start[n];
        for (int q = 0; q < nbond; q++)
            if (bond_dest[q] < 0 || bond_dest[q] >= n) return fail(MTMD_EINVAL, "bond destination out of range");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MTMD_ENODEV, "no HIP device");

    TopoDev t{};
    t.n = n;
    t.pbc = box ? 1 : 0;
    if (!need_rows) r0max = 1.0;                  // (grid unused)
    t.r0max2 = r0max * r0max * (1.0 + 1e-12);     // rows are a superset; each rule tests its own R0
    // grid: cells of at least R0max over the periodic box, or over the atoms' extent
    for (int ax = 0; ax < 3; ax++) {
        double lo = 0.0, hi = 0.0;
        if (box) {
            if (!(box[ax] > 0.0)) return fail(MTMD_EINVAL, "box widths must be positive");
            t.w[ax] = box[ax];
            lo = -box[ax] / 2.0;
            hi = box[ax] / 2.0;
        } else {
            lo = hi = pos[ax];
            for (int k = 1; k < n; k++) { lo = std::min(lo, pos[3 * k + ax]); hi = std::max(hi, pos[3 * k + ax]); }
            hi += 1e-9 * std::max(1.0, std::fabs(hi));
        }
        const double width = std::max(hi - lo, r0max);
        int nc = (int)std::floor(width / r0max);
        nc = std::max(1, std::min(nc, 1024));
        t.nc[ax] = nc;
        t.lo[ax] = lo;
        t.cs[ax] = width / nc;
        if (t.cs[ax] < r0max) t.cs[ax] = r0max;   // (only reached through rounding)
    }
    // at most a few cells per atom: coarser cells stay correct (they only hold more atoms)
    const long long max_cells = std::max<long long>(64, 4LL * n);
    while ((long long)t.nc[0] * t.nc[1] * t.nc[2] > max_cells) {
        int ax = 0;
        for (int q = 1; q < 3; q++) if (t.nc[q] > t.nc[ax]) ax = q;
        const double width = t.cs[ax] * t.nc[ax];
        t.nc[ax] = std::max(1, (t.nc[ax] * 3) / 4);
        t.cs[ax] = width / t.nc[ax];
    }
    const long long ncells = (long long)t.nc[0] * t.nc[1] * t.nc[2];

    DevMem mem;
    hipStream_t s = nullptr;
    double* d_pos = mem.alloc<double>(3 * (size_t)n);
    int* d_type = mem.alloc<int>(n);
    int* d_cell = mem.alloc<int>(n);
    int* d_count = mem.alloc<int>(n);
    int* d_cstart = mem.alloc<int>(ncells + 1);
    int* d_catoms = mem.alloc<int>(n);
    int* d_rstart = mem.alloc<int>((size_t)n + 1);
    if (!d_pos || !d_type || !d_cell || !d_count || !d_cstart || !d_catoms || !d_rstart)
        return fail(MTMD_ENOMEM, "device allocation failed");
#define TCHK(expr)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return fail(MTMD_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
    TCHK(hipMemcpy(d_pos, pos, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    TCHK(hipMemcpy(d_type, type, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    t.pos = d_pos;
    t.type = d_type;
    const int B = 256, G = (n + B - 1) / B;
    std::vector<int> cnt(n), rstart;

    if (need_graph) {
        int* d_bs = mem.alloc<int>((size_t)n + 1);
        int* d_bd = mem.alloc<int>((size_t)nbond);
        if (!d_bs || !d_bd) return fail(MTMD_ENOMEM, "device allocation failed");
        TCHK(hipMemcpy(d_bs, bond_start, ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice));
        if (nbond) TCHK(hipMemcpy(d_bd, bond_dest, (size_t)nbond * sizeof(int), hipMemcpyHostToDevice));
        t.bs = d_bs;
        t.bd = d_bd;
    }

    if (need_rows) {
        // cells, then a counting sort of the atoms by cell on the host (set-up work, O(N))
        hipLaunchKernelGGL(k_topo_cells, dim3(G), dim3(B), 0, s, t, d_cell);
        std::vector<int> cell(n);
        TCHK(hipMemcpy(cell.data(), d_cell, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
        std::vector<int> cstart(ncells + 1, 0), catoms(n);
        for (int a = 0; a < n; a++) {
            if (cell[a] < 0 || cell[a] >= ncells) return fail(MTMD_EHIP, "cell index out of range");
            cstart[cell[a] + 1]++;
        }
        for (long long c = 0; c < ncells; c++) cstart[c + 1] += cstart[c];
        {
            std::vector<int> fill(cstart.begin(), cstart.end() - 1);
            for (int a = 0; a < n; a++) catoms[fill[cell[a]]++] = a;
        }
        TCHK(hipMemcpy(d_cstart, cstart.data(), cstart.size() * sizeof(int), hipMemcpyHostToDevice));
        TCHK(hipMemcpy(d_catoms, catoms.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
        t.cell_start = d_cstart;
        t.cell_atoms = d_catoms;

        // rows of the atoms within R0max
        hipLaunchKernelGGL(k_topo_rows, dim3(G), dim3(B), 0, s, t, d_count);
        TCHK(hipMemcpy(cnt.data(), d_count, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
        const long long nrow = scan(cnt, rstart);
        if (nrow > 0x7fffffffLL) return fail(MTMD_ENOMEM, "too many pairs within R0");
        int* d_rows = mem.alloc<int>(nrow);
        if (!d_rows) return fail(MTMD_ENOMEM, "device allocation failed");
        TCHK(hipMemcpy(d_rstart, rstart.data(), rstart.size() * sizeof(int), hipMemcpyHostToDevice));
        t.row_start = d_rstart;
        t.rows = d_rows;
        hipLaunchKernelGGL(k_topo_rows, dim3(G), dim3(B), 0, s, t, (int*)nullptr);
        TCHK(hipGetLastError());
    }

    MolSlots ms{};
    if (need_mol) {
        // molecules of the undirected bond graph (union-find); each walk stays inside one
        std::vector<int> parent(n);
        for (int a = 0; a < n; a++) parent[a] = a;
        auto root = [&](int a) {
            while (parent[a] != a) { parent[a] = parent[parent[a]]; a = parent[a]; }
            return a;
        };
        for (int a = 0; a < n; a++)
            for (int q = bond_start[a]; q < bond_start[a + 1]; q++) {
                const int ra = root(a), rb = root(bond_dest[q]);
                if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
            }
        std::vector<int> comp(n), loc(n), size(n, 0);
        int mmax = 0;
        for (int a = 0; a < n; a++) {
            comp[a] = root(a);
            loc[a] = size[comp[a]]++;
            mmax = std::max(mmax, size[comp[a]]);
        }
        int* d_comp = mem.alloc<int>(n);
        int* d_loc = mem.alloc<int>(n);
        if (!d_comp || !d_loc) return fail(MTMD_ENOMEM, "device allocation failed");
        TCHK(hipMemcpy(d_comp, comp.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
        TCHK(hipMemcpy(d_loc, loc.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
        t.comp = d_comp;
        t.loc = d_loc;
        // slots: 16 bytes per molecule atom each, within 256 MiB
        const long long per_slot = 4LL * mmax;
        long long S = (256LL << 20) / (4 * per_slot);
        S = std::max<long long>(64, std::min<long long>(S, 65536));
        S = std::min<long long>(S, n);
        ms.mmax = mmax;
        ms.S = (int)S;
        ms.rank = mem.alloc<int>((size_t)S * mmax);
        ms.order = mem.alloc<int>((size_t)S * mmax);
        ms.stk = mem.alloc<int>((size_t)S * mmax * 2);
        if (!ms.rank || !ms.order || !ms.stk) return fail(MTMD_ENOMEM, "device allocation failed");
        TCHK(hipMemset(ms.rank, 0xff, (size_t)S * mmax * sizeof(int)));   // -1: outside every walk
    }

    // chains per force field, in force-field order
    const int width = which + 2;
    std::vector<int> all;
    std::vector<int> off;
    int* d_off2 = mem.alloc<int>((size_t)n + 1);
    if (!d_off2) return fail(MTMD_ENOMEM, "device allocation failed");
    for (int ri = 0; ri < n_rules; ri++) {
        Rule r{};
        for (int q = 0; q < 4; q++) r.types[q] = rules[ri].types[q];
        r.r02 = rules[ri].R0 * rules[ri].R0;
        const int crit = rules[ri].crit;
        const int GM = (ms.S + B - 1) / B;
        auto launch = [&](const int* o, int* c, int* w) {
            if (crit == MTMD_CRIT_ALL_R) {
                if (which == 1) hipLaunchKernelGGL((k_topo_paths<0>), dim3(G), dim3(B), 0, s, t, r, ri, o, c, w);
                else if (which == 2) hipLaunchKernelGGL((k_topo_paths<1>), dim3(G), dim3(B), 0, s, t, r, ri, o, c, w);
                else hipLaunchKernelGGL((k_topo_paths<2>), dim3(G), dim3(B), 0, s, t, r, ri, o, c, w);
            } else if (crit == MTMD_CRIT_MOL_R) {
                if (which == 1) hipLaunchKernelGGL((k_topo_mol<0>), dim3(GM), dim3(B), 0, s, t, r, ri, ms, o, c, w);
                else if (which == 2) hipLaunchKernelGGL((k_topo_mol<1>), dim3(GM), dim3(B), 0, s, t, r, ri, ms, o, c, w);
                else hipLaunchKernelGGL((k_topo_mol<2>), dim3(GM), dim3(B), 0, s, t, r, ri, ms, o, c, w);
            } else if (crit == MTMD_CRIT_VISIBLE) {
                if (which == 1) hipLaunchKernelGGL((k_topo_graph<0, false>), dim3(G), dim3(B), 0, s, t, r, ri, o, c, w);
                else if (which == 2) hipLaunchKernelGGL((k_topo_graph<1, false>), dim3(G), dim3(B), 0, s, t, r, ri, o, c, w);
                else hipLaunchKernelGGL((k_topo_graph<2, false>), dim3(G), dim3(B), 0, s, t, r, ri, o, c, w);
            } else {
                hipLaunchKernelGGL((k_topo_graph<0, true>), dim3(G), dim3(B), 0, s, t, r, ri, o, c, w);
            }
        };
        launch(nullptr, d_count, nullptr);
        TCHK(hipGetLastError());
        TCHK(hipMemcpy(cnt.data(), d_count, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
        const long long np = scan(cnt, off);
        if (np == 0) continue;
        if ((long long)all.size() + np * width > 0x7fffffffLL) return fail(MTMD_ENOMEM, "too many entries");
        int* d_out = mem.alloc<int>((size_t)np * width);
        if (!d_out) return fail(MTMD_ENOMEM, "device allocation failed");
        TCHK(hipMemcpy(d_off2, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
        launch(d_off2, nullptr, d_out);
        TCHK(hipGetLastError());
        const size_t base = all.size();
        all.resize(base + (size_t)np * width);
        TCHK(hipMemcpy(all.data() + base, d_out, (size_t)np * width * sizeof(int), hipMemcpyDeviceToHost));
    }
    TCHK(hipDeviceSynchronize());
#undef TCHK

    // keep the first occurrence of each entry and of its reverse; drop angles and
    // dihedrals that name an atom twice (only the graph walks make them). The
    // reference drops those after the duplicate pass; an entry naming an atom
    // twice, and its reverse, can only equal another such entry, so dropping
    // them first gives the same list.
    struct Key {
        int v[4];
        bool operator==(const Key& o) const { return std::memcmp(v, o.v, sizeof v) == 0; }
    };
    struct Hash {
        size_t operator()(const Key& k) const
        {
            size_t h = 1469598103934665603ull;
            for (int x : k.v) h = (h ^ (size_t)(unsigned)x) * 1099511628211ull;
            return h;
        }
    };
    std::unordered_set<Key, Hash> seen;
    const size_t total = all.size() / width;
    seen.reserve(total * 2 + 1);
    long long kept = 0;
    for (size_t e = 0; e < total; e++) {
        const int* x = &all[e * width];
        Key fwd{{-1, -1, -1, -1}}, rev{{-1, -1, -1, -1}};
        bool repeat = false;
        for (int q = 0; q < which + 1; q++) {
            fwd.v[q] = x[q];
            rev.v[q] = x[which - q];
            for (int q2 = 0; q2 < q && which >= 2; q2++) repeat |= (x[q2] == x[q]);
        }
        if (repeat || seen.count(fwd) || seen.count(rev)) continue;
        seen.insert(fwd);
        if (out && kept < max_out) std::memcpy(out + kept * width, x, width * sizeof(int));
        kept++;
    }
    *n_out = (int32_t)kept;
    if (out && kept > max_out)
        return fail(MTMD_EINVAL, "output holds " + std::to_string(max_out) + " entries, " + std::to_string(kept) +
                                     " are needed");
    return MTMD_OK;
}
