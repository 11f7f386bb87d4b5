/*
 * mtmd_oracle.h - CPU restatement of MolTwister's `moldyn` hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the oracle the HIP engine is checked
 * against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker (or the timed CPU baseline). The
 * product library (libmtmd.so) never links, loads or falls back to it.
 *
 * Every function follows the reference file:line named beside it, operation
 * for operation (same precision per operation, same order), so that cell
 * lists, neighbour lists and per-atom forces are bit-identical to the
 * reference CPU build. It is compiled with -ffp-contract=off.
 *
 * Pinning (see tests/test_oracle_ref.py, tests/test_oracle_math.py):
 *  - profiles, cells, neighbour lists, per-atom forces, NH chain propagator:
 *    bit-compared against the reference's own sources compiled by
 *    oracle/Makefile (oracle/_ref) and against committed fixtures
 *    (tests/golden) written from that build;
 *  - velocity-Verlet step, PBC wrap, thermo reductions, gradient-descent step:
 *    pinned by closed forms and invariants (time reversibility, exact
 *    harmonic orbits, sums against numpy) because the reference's
 *    CVelVerlet/CSimulationBox need CMolTwisterState + C3DView, whose
 *    translation unit needs GLUT, absent here;
 *  - NVT and NPT steps (thermostat and barostat chains, MTK barostat): the
 *    trajectory converges with order 2 in the time step to the MTK equations
 *    of motion integrated independently (tests/test_oracle_npt.py);
 *  - per-atom bonded lists and molecule indices: restated twice (here and in
 *    the reference driver, MDFFMatrices.cpp:384-520), pinned through the
 *    force comparison with the reference functor.
 */
#ifndef MTMD_ORACLE_H
#define MTMD_ORACLE_H

#include <stdint.h>
#include "../include/mtmd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_md orc_md;

/* profiles (which: 0 non-bonded, 1 bond, 2 angle, 3 dihedral) */
int orc_profile(int which, int kind, const double* params, float x0, float x1,
                int n, float* x, float* f, float* e);

orc_md* orc_create(const mtmd_system* sys, const mtmd_config* cfg);
void    orc_destroy(orc_md* o);
int     orc_natoms(const orc_md* o);
/* host threads for the per-atom loops of the neighbour-list build and the
 * force functor (default 1). Every row and every atom's force is computed by
 * the same operations in the same order whatever the thread count, so the
 * results are identical; only the CPU baseline's all-core figure uses t > 1. */
void    orc_set_threads(orc_md* o, int threads);

void orc_set_state(orc_md* o, const double* pos, const double* mom);
void orc_get_state(const orc_md* o, double* pos, double* mom);
void orc_set_thermostat(orc_md* o, const double* eta, const double* p_eta);
void orc_get_thermostat(const orc_md* o, double* eta, double* p_eta);
/* NPT barostat: eps, p_eps and the pressure chain (arrays of pressure_nh_len);
 * box receives the current simulation box (getSimBox) */
void orc_set_barostat(orc_md* o, double eps, double p_eps, const double* eta, const double* p_eta);
void orc_get_barostat(const orc_md* o, double* eps, double* p_eps, double* eta, double* p_eta, double* box);
/* raw per-atom data, for tests */
void orc_get_mol(const orc_md* o, int32_t* mol);

/* cells on current positions; returns reference error code (0, 1, 2) */
int orc_build_cells(orc_md* o, int32_t* ncells_xyz, int32_t* ixyz,
                    int32_t* cell_first, int32_t* cell_count, int32_t* cell_list);
/* neighbour lists (after orc_build_cells); returns max count */
int orc_build_neighbors(orc_md* o, int32_t max_per_atom, int32_t* count, int32_t* list);

/* full force call: cells + neighbours + per-atom functor (uncapped) */
void orc_compute_forces(orc_md* o, double* F, float* U);
/* one NH-T propagator call (NHChain.cpp:54-113 with CFctT, MDLoop.cpp:29-60) */
void orc_nh_propagate(orc_md* o);
/* nsteps of the MD loop (MDLoop.cpp:124-138) */
void orc_step(orc_md* o, int nsteps);
/* thermo: T[K], P[atm], V, U, K, Etot, sum p2/m, sum F.r */
void orc_thermo(orc_md* o, double* out8);
void orc_get_forces(const orc_md* o, double* F, float* U);
/* gradient descent (MDLoop.cpp:143-173); returns steps done */
int  orc_optimize(orc_md* o, int max_steps, double lr, double accuracy, int out_stride);

/* topology search by the rules' detection criteria (brute force);
 * bond_start / bond_dest: the bond graph (NULL when every rule is
 * MTMD_CRIT_ALL_R); returns the number of entries, writes at most max_out */
int orc_find_topology(int which, int n, const double* pos, const int32_t* type, const double* box,
                      const int32_t* bond_start, const int32_t* bond_dest, int n_rules,
                      const mtmd_topo_rule* rules, int max_out, int32_t* out);

/* reference scalar helpers */
double orc_cmt_exp(double x);
double orc_cmt_sinhxoverx(double x);

#ifdef __cplusplus
}
#endif
#endif
