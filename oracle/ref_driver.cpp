// ref_driver.cpp - drives MolTwister's own CPU sources as a checker.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile into oracle/_ref/libmtref.so
// together with unmodified reference translation units compiled straight from
// the MolTwister tree (paths below are relative to it). It exposes a small C
// API so tests can compare the oracle (oracle/mtmd_oracle.c) and the HIP
// engine with the reference's own arithmetic:
//   - profiles:  CMDFFNonBonded_LJ/_Buck/_LJ1208/_LJBuck::calc1D*Profile,
//                CMDFFBond_Harm/_Morse, CMDFFAngle_Harm/_Class2,
//                CMDFFDih_Harm/_Fourier4t
//   - cells:     CFunctorGenCellList::operator() + assembleCellList
//   - neighbours:CFunctorGenNeighList::operator() with the grow-and-retry loop
//   - forces:    CFunctorCalcForce::operator()
//   - NH chain:  CNHChain::propagator, CMt::exp/sinhXoverX, and its initial
//                draws (setRandNHPos, setRandNHMom)
//   - DCD files: CDCDFile::createDCDFileIfNotExists + appendToDCDFile, called
//                as CMDLoop::appendToDCDFile calls them (MDLoop.cpp:248-258)
//
// What the driver restates (the reference code for it needs CMolTwisterState,
// whose C3DView translation unit needs GLUT, absent in this image):
//   - the device-array set-up of CMDFFMatrices::prepareFFMatrices
//     (Cmd/MolTwisterCmdMolDyn/ForceFields/MDFFMatrices.cpp:181-536): atom
//     list, table layout, per-atom bonded lists, molecule indices;
//   - CFctT::G / scaleMomentum (Cmd/MolTwisterCmdMolDyn/MDLoop/MDLoop.cpp:29-60).
// The functors' private members are set directly (they are filled from
// CMDFFMatrices by setForceFieldMatrices in the reference), which is why the
// reference headers are included with access control lifted.
#include <vector>
#include <string>
#include <memory>
#include <functional>
#include <map>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>

#define private public
#define protected public
#include "Cmd/MolTwisterCmdMolDyn/ForceFields/FunctorCalcForce.h"
#include "Cmd/MolTwisterCmdMolDyn/ForceFields/FunctorGenCellList.h"
#include "Cmd/MolTwisterCmdMolDyn/ForceFields/FunctorGenNeighList.h"
#include "Cmd/MolTwisterCmdMolDyn/Integrators/NHChain.h"
#include "Cmd/MolTwisterCmdMolDyn/Integrators/Math.h"
#include "MDFF/Non-Bonded/MolTwisterMDFFNonBonded_LJ.h"
#include "MDFF/Non-Bonded/MolTwisterMDFFNonBonded_Buck.h"
#include "MDFF/Non-Bonded/MolTwisterMDFFNonBonded_LJ1208.h"
#include "MDFF/Bonds/MolTwisterMDFFBond_Harm.h"
#include "MDFF/Bonds/MolTwisterMDFFBond_Morse.h"
#include "MDFF/Angles/MolTwisterMDFFAngle_Harm.h"
#include "MDFF/Dihedrals/MolTwisterMDFFDih_Harm.h"
#include "MDFF/Non-Bonded/MolTwisterMDFFNonBonded_LJBuck.h"
#include "MDFF/Angles/MolTwisterMDFFAngle_Class2.h"
#include "MDFF/Dihedrals/MolTwisterMDFFDih_Fourier4t.h"
#include "Utilities/DCDFile.h"
#undef private
#undef protected

#include "../include/mtmd.h"

// defined in FunctorGenCellList.cpp (CPU build: a plain function)
void kernelAssembleList(CMDFFMatrices::CCellListIndex* devAtomCellIndices, int numAtomCellIndices, int* devCellList,
                        CMDFFMatrices::CListPointer* devCellListEntryPointers, int Nx, int Ny, int Nz, int* err);

#define EXPORT extern "C" __attribute__((visibility("default")))

static const int NP = NUM_POINTS_IN_PROFILES;
static const int MAXFF = MAX_FF_PER_ATOMIC_SET;

static std::vector<std::pair<float, float>> profile_of(int which, int kind, const double* p, float x0, float x1, int n,
                                                       bool energy)
{
    if (which == 0) {
        if (kind == MTMD_NB_LJ) {
            CMDFFNonBonded_LJ f; f.epsilon_ = p[0]; f.sigma_ = p[1]; f.alpha_ = p[2];
            return energy ? f.calc1DPotentialProfile(x0, x1, n) : f.calc1DForceProfile(x0, x1, n);
        }
        if (kind == MTMD_NB_BUCK) {
            CMDFFNonBonded_Buck f; f.A_ = p[0]; f.rho_ = p[1]; f.C_ = p[2];
            return energy ? f.calc1DPotentialProfile(x0, x1, n) : f.calc1DForceProfile(x0, x1, n);
        }
        if (kind == MTMD_NB_LJ1208) {
            CMDFFNonBonded_LJ1208 f; f.epsilon_ = p[0]; f.sigma_ = p[1]; f.alpha_ = p[2];
            return energy ? f.calc1DPotentialProfile(x0, x1, n) : f.calc1DForceProfile(x0, x1, n);
        }
        if (kind == MTMD_NB_LJBUCK) {
            CMDFFNonBonded_LJBuck f;
            f.ljBuckType_ = (int)p[0];
            if (f.ljBuckType_ == 0) { f.epsilon_ = p[1]; f.sigma_ = p[2]; }
            else { f.A_ = p[1]; f.rho_ = p[2]; f.C_ = p[3]; }
            return energy ? f.calc1DPotentialProfile(x0, x1, n) : f.calc1DForceProfile(x0, x1, n);
        }
    } else if (which == 1) {
        if (kind == MTMD_BOND_HARM) {
            CMDFFBond_Harm f; f.k_ = p[0]; f.r0_ = p[1];
            return energy ? f.calc1DPotentialProfile(x0, x1, n) : f.calc1DForceProfile(x0, x1, n);
        }
        if (kind == MTMD_BOND_MORSE) {
            CMDFFBond_Morse f; f.D_ = p[0]; f.alpha_ = p[1]; f.r0_ = p[2];
            return energy ? f.calc1DPotentialProfile(x0, x1, n) : f.calc1DForceProfile(x0, x1, n);
        }
    } else if (which == 2) {
        if (kind == MTMD_ANGLE_HARM) {
            CMDFFAngle_Harm f; f.k_ = p[0]; f.theta0_ = p[1];
            return energy ? f.calc1DPotentialProfile(x0, x1, n) : f.calc1DForceProfile(x0, x1, n);
        }
        if (kind == MTMD_ANGLE_CLASS2) {
            CMDFFAngle_Class2 f;
            f.Ea_.theta0_ = p[0]; f.Ea_.K2_ = p[1]; f.Ea_.K3_ = p[2]; f.Ea_.K4_ = p[3];
            f.Ebb_.M_ = p[4]; f.Ebb_.r1_ = p[5]; f.Ebb_.r2_ = p[6];
            f.Eba_.N1_ = p[7]; f.Eba_.N2_ = p[8]; f.Eba_.r1_ = p[9]; f.Eba_.r2_ = p[10];
            return energy ? f.calc1DPotentialProfile(x0, x1, n) : f.calc1DForceProfile(x0, x1, n);
        }
    } else if (which == 3) {
        if (kind == MTMD_DIH_HARM) {
            CMDFFDih_Harm f; f.K_ = p[0]; f.D_ = (int)p[1]; f.N_ = (int)p[2];
            return energy ? f.calc1DPotentialProfile(x0, x1, n) : f.calc1DForceProfile(x0, x1, n);
        }
        if (kind == MTMD_DIH_FOURIER4T) {
            CMDFFDih_Fourier4t f;
            for (int i = 0; i < 4; i++) f.V_[i] = p[i];
            return energy ? f.calc1DPotentialProfile(x0, x1, n) : f.calc1DForceProfile(x0, x1, n);
        }
    }
    return {};
}

EXPORT int ref_profile(int which, int kind, const double* p, float x0, float x1, int n, float* x, float* f, float* e)
{
    auto F = profile_of(which, kind, p, x0, x1, n, false);
    auto E = profile_of(which, kind, p, x0, x1, n, true);
    if ((int)F.size() != n || (int)E.size() != n) return -1;
    for (int i = 0; i < n; i++) {
        if (x) x[i] = F[i].first;
        if (f) f[i] = F[i].second;
        if (e) e[i] = E[i].second;
    }
    return 0;
}

EXPORT double ref_cmt_exp(double x) { return CMt::exp(x); }
EXPORT double ref_cmt_sinhxoverx(double x) { return CMt::sinhXoverX(x); }

struct ref_md {
    int n = 0, T = 0;
    float rc = 0, shell = 0;
    double L[3] = {0, 0, 0};
    float scale[4] = {0, 0, 0.5f, 0};
    std::vector<CMDFFMatrices::CAtom> atoms;
    std::vector<CMDFFMatrices::CPoint> nb, bondtab, angtab, dihtab;
    std::vector<size_t> nbcount;
    std::vector<CMDFFMatrices::CBond> blist;
    std::vector<CMDFFMatrices::CAngle> alist;
    std::vector<CMDFFMatrices::CDihedral> dlist;
    std::vector<CMDFFMatrices::CListPointer> bptr, aptr, dptr;
    // cells and neighbours
    int nc[3] = {0, 0, 0};
    std::vector<int> cellList;
    std::vector<CMDFFMatrices::CListPointer> cellPtr;
    std::vector<CMDFFMatrices::CCellListIndex> cellIdx;
    int maxNeigh = 10;
    std::vector<int> neighList, neighCount;
    int retries = 0;
    // NH
    std::vector<double> p, m;  // momenta [3n], masses
    CNHChain nh;
    double dt = 0;
};

// CFctT (MDLoop.cpp:29-60) over the driver's momentum array
class DrvFctT : public CFct
{
public:
    explicit DrvFctT(ref_md* s) : s_(s) {}
    double G(int j, std::vector<double>& p_eta, std::vector<double>& Q, double beta) override
    {
        const double g = double(3 * s_->n);
        const double g_tilde = 1.0;
        double ret = 0.0;
        if (j == 0) {
            double Sum_p2_m = 0.0;
            for (int k = 0; k < s_->n; k++) {
                C3DVector p(s_->p[3 * k], s_->p[3 * k + 1], s_->p[3 * k + 2]);
                double p2 = p * p;
                Sum_p2_m += (p2 / s_->m[k]);
            }
            ret = Sum_p2_m - (g / beta);
        } else {
            ret = (p_eta[j - 1] * p_eta[j - 1] / Q[j - 1]) - (g_tilde / beta);
        }
        return ret;
    }
    void scaleMomentum(double coeff) override
    {
        for (int k = 0; k < s_->n; k++) {
            C3DVector p(s_->p[3 * k], s_->p[3 * k + 1], s_->p[3 * k + 2]);
            p *= coeff;
            s_->p[3 * k] = p.x_; s_->p[3 * k + 1] = p.y_; s_->p[3 * k + 2] = p.z_;
        }
    }
private:
    ref_md* s_;
};

EXPORT ref_md* ref_create(const mtmd_system* s, const mtmd_config* cfg)
{
    ref_md* o = new ref_md();
    int n = s->n_atoms, T = s->n_types;
    o->n = n; o->T = T;
    o->rc = (float)cfg->cutoff_radius;
    o->shell = (float)cfg->neigh_shell;
    for (int i = 0; i < 3; i++) o->L[i] = s->box[i];
    o->scale[0] = (float)cfg->scale12; o->scale[1] = (float)cfg->scale13;
    o->scale[2] = (float)cfg->scale14; o->scale[3] = (float)cfg->scale1N;

    // atom list (MDFFMatrices.cpp:200-213)
    o->atoms.resize(n);
    for (int i = 0; i < n; i++) {
        auto& a = o->atoms[i];
        a.isMobile_ = s->mobile ? (s->mobile[i] != 0) : true;
        a.r_.x_ = s->pos[3 * i]; a.r_.y_ = s->pos[3 * i + 1]; a.r_.z_ = s->pos[3 * i + 2];
        a.m_ = (float)s->mass[i];
        a.q_ = (float)s->charge[i];
        a.index_ = i;
        a.typeIndex_ = s->type[i];
    }
    // molecule indices (MDFFMatrices.cpp:215-231, MolecularSystemTools.cpp:238-280)
    {
        std::vector<std::vector<int>> dest(n);
        for (int j = 0; j < s->n_bonds; j++) {
            int a = s->bonds[3 * j], b = s->bonds[3 * j + 1];
            if (a < n) dest[a].push_back(b);
            if (b < n) dest[b].push_back(a);
        }
        std::vector<int> mol(n, -1);
        int next = 0;
        for (int i = 0; i < n; i++) {
            if (mol[i] != -1) continue;
            std::vector<int> stack{i};
            std::map<int, int> seen;
            while (!stack.empty()) {
                int a = stack.back(); stack.pop_back();
                if (seen.count(a)) continue;
                seen[a] = 1;
                mol[a] = next;
                for (int b : dest[a]) if (b < n) stack.push_back(b);
            }
            next++;
        }
        for (int i = 0; i < n; i++) o->atoms[i].molIndex_ = mol[i];
    }
    // non-bonded matrix (MDFFMatrices.cpp:249-296)
    o->nb.assign((size_t)T * T * MAXFF * NP, CMDFFMatrices::CPoint());
    o->nbcount.assign((size_t)T * T, 0);
    for (int c = 0; c < T; c++)
        for (int r = 0; r < T; r++) {
            std::vector<int> idx;
            for (int i = 0; i < s->n_nonbonded; i++) {
                const auto& e = s->nonbonded[i];
                if (e.type_a == c && e.type_b == r) { idx.push_back(i); continue; }
                if (e.type_b == c && e.type_a == r) { idx.push_back(i); continue; }
            }
            if ((int)idx.size() > MAXFF) idx.resize(MAXFF);
            o->nbcount[CMDFFMatrices::toIndexNonBond(r, c, T)] = idx.size();
            for (size_t i = 0; i < idx.size(); i++) {
                const auto& e = s->nonbonded[idx[i]];
                auto F = profile_of(0, e.kind, e.p, 0.01f, o->rc, NP, false);
                auto E = profile_of(0, e.kind, e.p, 0.01f, o->rc, NP, true);
                for (int j = 0; j < NP; j++)
                    o->nb[CMDFFMatrices::toIndexNonBond(r, c, i, j, T, T, MAXFF)] =
                        CMDFFMatrices::CPoint(F[j].first, F[j].second, E[j].second);
            }
        }
    // bonded tables (MDFFMatrices.cpp:298-382)
    auto fill = [&](std::vector<CMDFFMatrices::CPoint>& dst, int which, const mtmd_bonded_ff* ff, int cnt, float x0, float x1) {
        dst.assign((size_t)(cnt + 1) * NP, CMDFFMatrices::CPoint());
        for (int i = 0; i < cnt; i++) {
            auto F = profile_of(which, ff[i].kind, ff[i].p, x0, x1, NP, false);
            auto E = profile_of(which, ff[i].kind, ff[i].p, x0, x1, NP, true);
            for (int j = 0; j < NP; j++)
                dst[CMDFFMatrices::toIndexBonded(i, j, NP)] = CMDFFMatrices::CPoint(F[j].first, F[j].second, E[j].second);
        }
    };
    fill(o->bondtab, 1, s->bond_ff, s->n_bond_ff, 0.01f, o->rc);
    fill(o->angtab, 2, s->angle_ff, s->n_angle_ff, 0.0, float(2.0 * M_PI));
    fill(o->dihtab, 3, s->dihedral_ff, s->n_dihedral_ff, 0.0, float(2.0 * M_PI));

    // per-atom bonded lists (MDFFMatrices.cpp:384-520)
    o->bptr.resize(n); o->aptr.resize(n); o->dptr.resize(n);
    int g = 0;
    for (int k = 0; k < n; k++) {
        o->bptr[k].indexFirstEntry_ = g;
        int cnt = 0;
        for (int j = 0; j < s->n_bonds; j++) {
            int a1 = s->bonds[3 * j], a2 = s->bonds[3 * j + 1], t = s->bonds[3 * j + 2];
            int to = -1;
            if (a1 == k) to = a2;
            if (a2 == k) to = a1;
            if (to != -1 && t != -1) {
                CMDFFMatrices::CBond b; b.atomIndex1_ = k; b.atomIndex2_ = to; b.bondType_ = t;
                o->blist.push_back(b); g++; cnt++;
            }
        }
        o->bptr[k].numEntries_ = cnt;
    }
    g = 0;
    for (int k = 0; k < n; k++) {
        o->aptr[k].indexFirstEntry_ = g;
        int cnt = 0;
        for (int j = 0; j < s->n_angles; j++) {
            const int* a = s->angles + 4 * j;
            int I = -1, J = -1; bool center = false;
            if (a[0] == k) { I = a[1]; J = a[2]; }
            if (a[1] == k) { center = true; I = a[0]; J = a[2]; }
            if (a[2] == k) { I = a[1]; J = a[0]; }
            if (I != -1 && J != -1 && a[3] != -1) {
                CMDFFMatrices::CAngle an; an.atomIndex1_ = k; an.atomIndex2_ = I; an.atomIndex3_ = J;
                an.angleType_ = a[3]; an.assocAtomIsAtCenterOfAngle_ = center;
                o->alist.push_back(an); g++; cnt++;
            }
        }
        o->aptr[k].numEntries_ = cnt;
    }
    g = 0;
    for (int k = 0; k < n; k++) {
        o->dptr[k].indexFirstEntry_ = g;
        int cnt = 0;
        for (int j = 0; j < s->n_dihedrals; j++) {
            const int* d = s->dihedrals + 5 * j;
            int I = -1, J = -1, Lx = -1; bool center = false;
            if (d[0] == k) { I = d[1]; J = d[2]; Lx = d[3]; }
            if (d[1] == k) { center = true; I = d[0]; J = d[2]; Lx = d[3]; }
            if (d[2] == k) { center = true; I = d[3]; J = d[1]; Lx = d[0]; }
            if (d[3] == k) { I = d[2]; J = d[1]; Lx = d[0]; }
            if (I != -1 && J != -1 && Lx != -1 && d[4] != -1) {
                CMDFFMatrices::CDihedral dh; dh.atomIndex1_ = k; dh.atomIndex2_ = I; dh.atomIndex3_ = J;
                dh.atomIndex4_ = Lx; dh.dihedralType_ = d[4]; dh.assocAtomIsAtCenterOfDihedral_ = center;
                o->dlist.push_back(dh); g++; cnt++;
            }
        }
        o->dptr[k].numEntries_ = cnt;
    }
    // keep the vectors non-empty so &v[0] is valid
    if (o->blist.empty()) o->blist.resize(1);
    if (o->alist.empty()) o->alist.resize(1);
    if (o->dlist.empty()) o->dlist.resize(1);

    // cell list init (MDFFMatrices.cpp:43-63)
    {
        float R = o->rc + o->shell;
        for (int i = 0; i < 3; i++) o->nc[i] = (int)floor(double(float(o->L[i]) / R));
        int tot = o->nc[0] * o->nc[1] * o->nc[2];
        o->cellList.assign(n, -1);
        o->cellPtr.assign(tot > 0 ? tot : 1, CMDFFMatrices::CListPointer());
        o->cellIdx.assign(n, CMDFFMatrices::CCellListIndex());
    }
    o->neighList.assign((size_t)n * o->maxNeigh, 0);
    o->neighCount.assign(n, 0);

    // momenta, masses and NH chain (SimulationBox.cpp:53-79, NHChain.cpp:115-124)
    o->p.assign(3 * (size_t)n, 0.0);
    if (s->mom) for (int i = 0; i < 3 * n; i++) o->p[i] = s->mom[i];
    o->m.assign(s->mass, s->mass + n);
    o->nh.M_ = cfg->temperature_nh_len;
    o->nh.n_ = cfg->temperature_respa;
    o->nh.T_ = cfg->temperature / 120.272;
    o->nh.tau_ = cfg->temperature_relax / 1.0E2;
    o->nh.prepareArrays(n, 3);
    o->dt = cfg->timestep / 1.0E2;
    return o;
}

EXPORT void ref_destroy(ref_md* o) { delete o; }

EXPORT void ref_set_state(ref_md* o, const double* pos, const double* mom)
{
    for (int i = 0; i < o->n; i++) {
        if (pos) { o->atoms[i].r_.x_ = pos[3 * i]; o->atoms[i].r_.y_ = pos[3 * i + 1]; o->atoms[i].r_.z_ = pos[3 * i + 2]; }
    }
    if (mom) for (int i = 0; i < 3 * o->n; i++) o->p[i] = mom[i];
}

EXPORT void ref_get_mol(const ref_md* o, int* mol)
{
    for (int i = 0; i < o->n; i++) mol[i] = o->atoms[i].molIndex_;
}

// CMDFFMatrices::genNeighList, cell half (MDFFMatrices.cpp:151-159)
EXPORT int ref_build_cells(ref_md* o, int* ncells_xyz, int* ixyz, int* cell_first, int* cell_count, int* cell_list)
{
    int tot = o->nc[0] * o->nc[1] * o->nc[2];
    if (o->nc[0] < 1 || o->nc[1] < 1 || o->nc[2] < 1) return -1;
    for (auto& cp : o->cellPtr) cp = CMDFFMatrices::CListPointer(-1, 0);   // resetCellList
    CFunctorGenCellList gen;
    gen.devCellList_ = o->cellList.data();
    gen.devCellListEntryPointers_ = o->cellPtr.data();
    gen.cellCountX_ = o->nc[0]; gen.cellCountY_ = o->nc[1]; gen.cellCountZ_ = o->nc[2];
    // updatePBC (MDFFMatrices.cpp:67-76)
    gen.pbcWidthX_ = (float)o->L[0]; gen.pbcWidthY_ = (float)o->L[1]; gen.pbcWidthZ_ = (float)o->L[2];
    gen.pbcLowX_ = -gen.pbcWidthX_ / 2.0f; gen.pbcLowY_ = -gen.pbcWidthY_ / 2.0f; gen.pbcLowZ_ = -gen.pbcWidthZ_ / 2.0f;
    std::transform(o->atoms.begin(), o->atoms.end(), o->cellIdx.begin(), gen);
    int err = 0;
    kernelAssembleList(o->cellIdx.data(), (int)o->cellIdx.size(), o->cellList.data(), o->cellPtr.data(),
                       o->nc[0], o->nc[1], o->nc[2], &err);
    if (ncells_xyz) { ncells_xyz[0] = o->nc[0]; ncells_xyz[1] = o->nc[1]; ncells_xyz[2] = o->nc[2]; }
    for (int k = 0; k < o->n; k++) {
        if (ixyz) { ixyz[3 * k] = o->cellIdx[k].ix_; ixyz[3 * k + 1] = o->cellIdx[k].iy_; ixyz[3 * k + 2] = o->cellIdx[k].iz_; }
        if (cell_list) cell_list[k] = o->cellList[k];
    }
    for (int c = 0; c < tot; c++) {
        if (cell_first) cell_first[c] = o->cellPtr[c].indexFirstEntry_;
        if (cell_count) cell_count[c] = o->cellPtr[c].numEntries_;
    }
    return err;
}

static void setup_neigh_functor(ref_md* o, CFunctorGenNeighList& gen)
{
    gen.devCellList_ = o->cellList.data();
    gen.devCellListEntryPointers_ = o->cellPtr.data();
    gen.devAtomCellIndicesRaw_ = o->cellIdx.data();
    gen.devNeighList_ = o->neighList.data();
    gen.devNeighListCount_ = o->neighCount.data();
    gen.devAtomList_ = o->atoms.data();
    gen.numAtoms_ = o->n;
    gen.cellCountX_ = o->nc[0]; gen.cellCountY_ = o->nc[1]; gen.cellCountZ_ = o->nc[2];
    gen.maxNeighbors_ = o->maxNeigh;
    gen.rCutoff2_ = o->rc;
    gen.rCutoff2_ *= gen.rCutoff2_;
    float w[3], lo[3];
    for (int i = 0; i < 3; i++) { w[i] = (float)o->L[i]; lo[i] = -w[i] / 2.0f; }
    gen.pbc_.rLow_.x_ = (double)lo[0]; gen.pbc_.rLow_.y_ = (double)lo[1]; gen.pbc_.rLow_.z_ = (double)lo[2];
    gen.pbc_.rHigh_.x_ = double(lo[0] + w[0]); gen.pbc_.rHigh_.y_ = double(lo[1] + w[1]); gen.pbc_.rHigh_.z_ = double(lo[2] + w[2]);
}

// CMDFFMatrices::genNeighList, neighbour half with the retry loop (MDFFMatrices.cpp:161-179)
EXPORT int ref_build_neighbors(ref_md* o, int max_per_atom, int* count, int* list, int* capacity)
{
    int storage = -1;
    do {
        std::fill(o->neighCount.begin(), o->neighCount.end(), 0);   // resetNeighList
        CFunctorGenNeighList gen;
        setup_neigh_functor(o, gen);
        std::transform(o->atoms.begin(), o->atoms.end(), o->neighCount.begin(), gen);
        storage = gen.checkRequiredMaxNumNeighbours(o->neighCount);
        if (storage != -1) {
            o->maxNeigh = storage + 10;   // resizeNeighList (MDFFMatrices.cpp:110-115)
            o->neighList.assign((size_t)o->n * o->maxNeigh, 0);
            o->retries++;
        }
    } while (storage != -1);
    int mx = 0;
    for (int k = 0; k < o->n; k++) {
        mx = std::max(mx, o->neighCount[k]);
        if (count) count[k] = o->neighCount[k];
        if (list) {
            int c = std::min(o->neighCount[k], max_per_atom);
            for (int j = 0; j < c; j++)
                list[(size_t)k * max_per_atom + j] = o->neighList[CFunctorGenNeighList::neighIndexToFlatIndex(k, j, o->maxNeigh)];
        }
    }
    if (capacity) *capacity = o->maxNeigh;
    return mx;
}

// CVelVerlet::calcParticleForces (VelVerlet.cpp:213-222) minus the host copies
EXPORT int ref_compute_forces(ref_md* o, double* F, double* Fpi, float* U)
{
    if (ref_build_cells(o, nullptr, nullptr, nullptr, nullptr, nullptr) < 0) return -1;
    ref_build_neighbors(o, 0, nullptr, nullptr, nullptr);
    CFunctorCalcForce calc(3, (float)o->L[0], (float)o->L[1], (float)o->L[2], 1000.0f, o->rc,
                           o->scale[0], o->scale[1], o->scale[2], o->scale[3]);
    calc.devAtomList_ = o->atoms.data();
    calc.devNonBondFFMatrix_ = o->nb.data();
    calc.devNonBondFFMatrixFFCount_ = o->nbcount.data();
    calc.devBondFFList_ = o->bondtab.data();
    calc.devAngleFFList_ = o->angtab.data();
    calc.devDihedralFFList_ = o->dihtab.data();
    calc.devNeighList_ = o->neighList.data();
    calc.devNeighListCount_ = o->neighCount.data();
    calc.devBondsForAtomListPointers_ = o->bptr.data();
    calc.devBondsForAtomLists_ = o->blist.data();
    calc.devAnglesForAtomListPointers_ = o->aptr.data();
    calc.devAnglesForAtomLists_ = o->alist.data();
    calc.devDihedralsForAtomListPointers_ = o->dptr.data();
    calc.devDihedralsForAtomLists_ = o->dlist.data();
    calc.Natoms_ = o->n;
    calc.Natomtypes_ = o->T;
    calc.maxNeighbours_ = o->maxNeigh;
    std::vector<CMDFFMatrices::CForces> out(o->n);
    std::transform(o->atoms.begin(), o->atoms.end(), out.begin(), calc);
    for (int k = 0; k < o->n; k++) {
        if (F) { F[3 * k] = out[k].F_.x_; F[3 * k + 1] = out[k].F_.y_; F[3 * k + 2] = out[k].F_.z_; }
        if (Fpi) { Fpi[3 * k] = out[k].Fpi_.x_; Fpi[3 * k + 1] = out[k].Fpi_.y_; Fpi[3 * k + 2] = out[k].Fpi_.z_; }
        if (U) U[k] = out[k].U_;
    }
    return 0;
}

EXPORT void ref_set_thermostat(ref_md* o, const double* eta, const double* p_eta)
{
    for (int j = 0; j < o->nh.M_; j++) {
        if (eta) o->nh.eta_[j] = eta[j];
        if (p_eta) o->nh.p_eta_[j] = p_eta[j];
    }
}

EXPORT void ref_get_thermostat(const ref_md* o, double* eta, double* p_eta)
{
    for (int j = 0; j < o->nh.M_; j++) {
        if (eta) eta[j] = o->nh.eta_[j];
        if (p_eta) p_eta[j] = o->nh.p_eta_[j];
    }
}

// CNHChain's own initial draws with the C library's rand() in its current
// state: prepareArrays(n, dim) (NHChain.cpp:115-124), then setRandNHPos
// (which = 0, NHChain.cpp:126-132) or setRandNHMom (which = 1, :134-143).
// A caller interleaves these with its own rand() calls to replay
// CSimulationBox::initSystem's sequence (SimulationBox.cpp:66-71).
EXPORT void ref_nh_draw(int which, int M, int n, int dim, double T, double tau, double* eta, double* p_eta)
{
    CNHChain c;
    c.M_ = M;
    c.T_ = T;
    c.tau_ = tau;
    c.prepareArrays(n, dim);
    if (which == 0) c.setRandNHPos(); else c.setRandNHMom();
    for (int j = 0; j < M; j++) {
        if (eta) eta[j] = c.eta_[j];
        if (p_eta) p_eta[j] = c.p_eta_[j];
    }
}

// one CNHChain::propagator call with CFctT (NHChain.cpp:54-113)
EXPORT void ref_nh_propagate(ref_md* o, double* mom_out)
{
    DrvFctT f(o);
    o->nh.propagator(o->n, 3, o->dt, f);
    if (mom_out) for (int i = 0; i < 3 * o->n; i++) mom_out[i] = o->p[i];
}

// CMDLoop::appendToDCDFile (MDLoop.cpp:248-258): header when the file is new,
// then one record of the positions xyz[3n] with the box narrowed to int
EXPORT void ref_dcd_append(const char* path, int n_steps, int stride, double timestep, int n, double box_x,
                           double box_y, double box_z, const double* xyz)
{
    std::function<std::tuple<double, double, double>(const int&)> get = [xyz](const int& i) {
        return std::tuple<double, double, double>(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    };
    CDCDFile::createDCDFileIfNotExists(path, n_steps, stride, timestep, n);
    CDCDFile::appendToDCDFile(path, n, {box_x, box_y, box_z}, get);
}
