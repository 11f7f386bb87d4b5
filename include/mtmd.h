/*
 * mtmd.h - C ABI of the MI355X molecular-dynamics engine (libmtmd.so).
 *
 * This is the drop-in boundary for MolTwister's `moldyn` hot path. Every entry
 * point names the reference interface it replaces (paths are relative to the
 * MolTwister source tree). Plain C types only: pointers, sizes, PODs.
 *
 * Units are the reference's internal ones (Cmd/MolTwisterCmdMolDyn/Integrators/
 * Constants.h:23-40): length in Angstrom, mass in g/mol, energy in kJ/mol,
 * charge in e, time in 100 fs, momentum in (g/mol)*Angstrom/(100 fs).
 * Configuration fields keep the reference's user units (fs, K, atm).
 *
 * Ownership: the caller owns every host buffer; mtmd_create deep-copies the
 * system, and getters write only into caller-provided buffers. A handle is
 * single-threaded (one HIP stream); different handles may be driven from
 * different host threads.
 *
 * Error convention: every int-returning call returns MTMD_OK (0) or a negative
 * MTMD_E* code; mtmd_last_error(h) gives the message (mtmd_last_error(NULL)
 * gives the message of the last failed mtmd_create on this thread).
 */
#ifndef MTMD_H
#define MTMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MTMD_ABI_VERSION 4

/* Number of points in every force/energy profile and the maximum number of
 * non-bonded force fields per atom-type pair
 * (Cmd/MolTwisterCmdMolDyn/ForceFields/MDFFMatrices.h:26-27). */
#define MTMD_NUM_POINTS 1000
#define MTMD_MAX_FF_PER_PAIR 5

/* status codes */
enum {
    MTMD_OK = 0,
    MTMD_EINVAL = -1,        /* invalid argument or unsupported option       */
    MTMD_EHIP = -2,          /* HIP runtime error                            */
    MTMD_ENEIGH = -3,        /* neighbour list overflow after regrowth       */
    MTMD_ENAN = -4,          /* non-finite value detected                    */
    MTMD_EBOX = -5,          /* box too small for the cell/neighbour scheme  */
    MTMD_ENOMEM = -6,        /* device or host allocation failed             */
    MTMD_ENODEV = -7         /* no HIP device                                */
};

/* ensembles: same values as SMolDynConfigStruct::Ensemble
 * (Cmd/MolTwisterCmdMolDyn/Config/MolDynConfigStruct.h:26) */
enum { MTMD_NVT = 0, MTMD_NPT = 1, MTMD_NVE = 2 };

/* engines (no counterpart in the reference; the reference has one algorithm):
 *  EXACT - the reference algorithm: cell and full neighbour lists rebuilt at rc
 *          before every force call, tabulated FP32 profiles with FP64 geometry
 *          and accumulation, the reference summation order per atom.
 *  FP64  - same arithmetic as EXACT per pair, but neighbour lists at rc + skin
 *          rebuilt on displacement, atoms spatially sorted at each rebuild.
 *  MIXED - as FP64 but pair geometry and table interpolation in FP32, the
 *          four pairs of a list chunk summed in FP32 and the chunk sums per
 *          atom in FP64; integration and reductions in FP64. */
enum { MTMD_ENGINE_EXACT = 0, MTMD_ENGINE_FP64 = 1, MTMD_ENGINE_MIXED = 2 };

/* Coulomb treatment.
 *  ALLPAIRS - minimum-image sum over all atoms in index order, no cutoff
 *             (Cmd/MolTwisterCmdMolDyn/ForceFields/FunctorCalcForce.cpp:258-278).
 *  CUTOFF   - the same pair term restricted to pairs inside the non-bonded cutoff.
 *  AUTO     - ALLPAIRS for the EXACT engine, CUTOFF otherwise. Systems whose
 *             charges are all zero skip the Coulomb pass (every term is +-0). */
enum { MTMD_COUL_AUTO = 0, MTMD_COUL_ALLPAIRS = 1, MTMD_COUL_CUTOFF = 2 };

/* force-field kinds */
enum { MTMD_NB_LJ = 1, MTMD_NB_BUCK = 2, MTMD_NB_LJ1208 = 3, MTMD_NB_LJBUCK = 4 };
/* Any force field given as its tabulated profile instead of parameters: the
 * MTMD_NUM_POINTS points (x, -dU/dx, U) that the reference's
 * calc1DForceProfile / calc1DPotentialProfile give for it, as they are fed to
 * the force matrices (MDFFMatrices.cpp:270-290 non-bonded, 296-382 bonded).
 * `profile` then points at 3 * MTMD_NUM_POINTS floats: x[], f[], e[]. Every
 * MDFF type of the reference can be passed this way, the tables are then
 * bit-identical by construction. The x range is the one the reference
 * samples: non-bonded and bonds 0.01 .. cutoff, angles and dihedrals
 * 0 .. 2 pi (the FP64/MIXED engines check the non-bonded range). */
#define MTMD_FF_PROFILE 100
enum { MTMD_BOND_HARM = 1, MTMD_BOND_MORSE = 2 };
enum { MTMD_ANGLE_HARM = 1, MTMD_ANGLE_CLASS2 = 2 };
enum { MTMD_DIH_HARM = 1, MTMD_DIH_FOURIER4T = 2 };

/* Run configuration: the fields of SMolDynConfigStruct
 * (Cmd/MolTwisterCmdMolDyn/Config/MolDynConfigStruct.h:24-65) that the engine
 * reads, with the defaults of CMolDynConfig::resetToDefaults
 * (Cmd/MolTwisterCmdMolDyn/Config/MolDynConfig.cpp:289-328), plus engine
 * options at the end. */
typedef struct mtmd_config {
    int32_t ensemble;               /* MTMD_NVT | MTMD_NPT | MTMD_NVE          */
    double  temperature;            /* K                                       */
    double  temperature_relax;      /* fs, Nose-Hoover tau                     */
    int32_t temperature_nh_len;     /* chain length M                          */
    int32_t temperature_respa;      /* RESPA steps n                           */
    double  pressure;               /* atm                                     */
    double  pressure_relax;         /* fs                                      */
    int32_t pressure_nh_len;
    int32_t pressure_respa;
    double  cutoff_radius;          /* Angstrom                                */
    double  neigh_shell;            /* Angstrom, sizes the cells               */
    double  cutoff_force;           /* per-component force cap, <= 0 disables  */
    int32_t n_steps;
    double  timestep;               /* fs                                      */
    int32_t out_stride;
    double  scale12, scale13, scale14, scale1N;
    double  opt_learning_rate;
    int32_t opt_max_steps;
    double  opt_accuracy;
    int32_t verbose;
    char    info_file[256];         /* thermo table written by mtmd_run        */
    /* output files of mtmd_run / mtmd_run_optimize (CMDLoop, MDLoop.cpp:84-102,
     * 370-431); paths as the `moldyn cfg set` keys xyzfile, dcdfile, xtcfile,
     * pdistrfile, vdistrfile (the last two are prefixes: <prefix>_p_x.dat ...,
     * <prefix>_v.dat) */
    char    xyz_file[256];
    char    dcd_file[256];
    char    xtc_file[256];
    char    pdistr_file[256];
    char    vdistr_file[256];
    int32_t include_xyz;            /* default 0                                */
    int32_t include_dcd;            /* default 1                                */
    int32_t include_xtc;            /* default 0; no XTC library: writes DCD    */
    float   xtc_precision;          /* default 1000                             */
    int32_t include_pdistr;         /* momentum histograms after step 1000      */
    int32_t include_vdistr;         /* volume histogram after step 1000         */
    double  max_pdistr;             /* AA*g/(fs*mol), default 0.3               */
    double  max_vdistr;             /* AA^3, default 1e5                        */
    /* engine options (no counterpart in the reference) */
    int32_t engine;                 /* MTMD_ENGINE_*                           */
    int32_t coulomb;                /* MTMD_COUL_*                             */
    double  skin;                   /* Angstrom, list radius = cutoff + skin   */
    int32_t block_size;             /* force-kernel workgroup size, 0 = auto   */
    int32_t reserved[7];
} mtmd_config;

/* One non-bonded force field between two atom types. Parameters:
 *  LJ:     p[0]=epsilon, p[1]=sigma, p[2]=alpha (default 1)
 *          (MDFF/Non-Bonded/MolTwisterMDFFNonBonded_LJ.cpp:131-162)
 *  BUCK:   p[0]=A, p[1]=rho, p[2]=C (MolTwisterMDFFNonBonded_Buck.cpp:116-143)
 *  LJ1208: p[0]=epsilon, p[1]=sigma, p[2]=alpha (default 1)
 *          (MolTwisterMDFFNonBonded_LJ1208.cpp:131-170)
 *  LJBUCK: p[0]=0 (LJ: p[1]=epsilon, p[2]=sigma) or 1 (Buckingham: p[1]=A,
 *          p[2]=rho, p[3]=C) (MolTwisterMDFFNonBonded_LJBuck.cpp:181-251).
 *          The reference's calcForces zeroes both forces at its end
 *          (:250), so the force profile of this kind is 0 and only the
 *          energy is tabulated; the engines reproduce that. */
typedef struct mtmd_nonbonded_ff {
    int32_t type_a, type_b, kind, pad;
    double  p[4];
    const float* profile;      /* kind == MTMD_FF_PROFILE only, else NULL   */
} mtmd_nonbonded_ff;

/* One bonded force-field type. Parameters:
 *  BOND_HARM:    p[0]=k, p[1]=r0          U = k (r - r0)^2
 *  BOND_MORSE:   p[0]=D, p[1]=alpha, p[2]=r0
 *  ANGLE_HARM:   p[0]=k, p[1]=theta0[deg] U = k (theta - theta0)^2
 *  ANGLE_CLASS2: p[0]=theta0[deg], p[1..3]=K2, K3, K4 (Ea); p[4..6]=M, r1, r2
 *                (Ebb); p[7..10]=N1, N2, r1, r2 (Eba)
 *                (MDFF/Angles/MolTwisterMDFFAngle_Class2.cpp:145-296)
 *  DIH_HARM:     p[0]=K, p[1]=D, p[2]=N   U = K (1 + D cos(N phi))
 *  DIH_FOURIER4T: p[0..3]=V1..V4          U = (V1 (1 + cos phi) + V2 (1 - cos 2phi)
 *                + V3 (1 + cos 3phi) + V4 (1 - cos 4phi)) / 2
 *                (MDFF/Dihedrals/MolTwisterMDFFDih_Fourier4t.cpp:98-142) */
#define MTMD_BONDED_PARAMS 12
typedef struct mtmd_bonded_ff {
    int32_t kind, pad;
    double  p[MTMD_BONDED_PARAMS];
    const float* profile;      /* kind == MTMD_FF_PROFILE only, else NULL   */
} mtmd_bonded_ff;

/* The system: the inputs CMDFFMatrices::prepareFFMatrices and CSimulationBox
 * read from CMolTwisterState (MDFFMatrices.cpp:181-536,
 * SimulationBox.cpp:27-133). Atom types are indices in first-appearance order
 * of the atom names (MolTwisterState.cpp:459-484); the box is the periodic cell
 * centred on the origin with widths box[0..2] (C3DView::calcPBC with a user
 * PBC). Topology lists are the global MD bond/angle/dihedral lists in the order
 * CMolTwisterStateTools::getAllMD*InSystem returns them; molecule membership is
 * derived from the bonds (MolecularSystemTools.cpp:238-280). */
typedef struct mtmd_system {
    int32_t n_atoms;
    int32_t n_types;
    const double*  pos;        /* [3*n_atoms] x,y,z                               */
    const double*  mom;        /* [3*n_atoms] or NULL: random momenta scaled to T */
    const double*  mass;       /* [n_atoms]                                       */
    const double*  charge;     /* [n_atoms]                                       */
    const int32_t* type;       /* [n_atoms]                                       */
    const uint8_t* mobile;     /* [n_atoms] or NULL (all mobile)                  */
    double box[3];
    int32_t n_nonbonded;   const mtmd_nonbonded_ff* nonbonded;
    int32_t n_bond_ff;     const mtmd_bonded_ff*    bond_ff;
    int32_t n_angle_ff;    const mtmd_bonded_ff*    angle_ff;
    int32_t n_dihedral_ff; const mtmd_bonded_ff*    dihedral_ff;
    int32_t n_bonds;       const int32_t* bonds;      /* [3*n]: a, b, ff          */
    int32_t n_angles;      const int32_t* angles;     /* [4*n]: a, centre, c, ff  */
    int32_t n_dihedrals;   const int32_t* dihedrals;  /* [5*n]: a, b, c, d, ff    */
    uint64_t seed;             /* for random momenta when mom == NULL            */
    /* [n_types] atom-type names (the atom IDs of CMolTwisterState, which
     * define the types) for the XYZ writer, or NULL: types are then named by
     * their index */
    const char* const* type_names;
} mtmd_system;

/* Thermodynamic output (SimulationBox.cpp:168-231, MDLoop.cpp:391-404).
 * T and P use the reference definitions: T from dim*N degrees of freedom,
 * P from the per-atom sum of Fpi.r with wrapped coordinates. */
typedef struct mtmd_thermo {
    double T;        /* K                   */
    double P;        /* atm                 */
    double V;        /* Angstrom^3          */
    double U;        /* kJ/mol              */
    double K;        /* kJ/mol              */
    double Etot;     /* kJ/mol              */
    double sum_p2m;  /* sum p.p/m (reduced) */
    double sum_fr;   /* sum F.r   (reduced) */
    double U_cut;    /* kJ/mol: sum over pairs inside the cutoff of the pair
                        energy at the cutoff (scaled like the pair): the
                        tabulated terms and, with MTMD_COUL_CUTOFF, the
                        Coulomb term K qi qj / rc. The pair potentials are
                        truncated without a shift, so U jumps by U(rc) when a
                        pair crosses rc; Etot - U_cut does not jump and is
                        what the integrator conserves. No counterpart in the
                        reference.                                           */
} mtmd_thermo;

typedef struct mtmd_handle mtmd_handle;

/* ---- library ---- */
int         mtmd_abi_version(void);
const char* mtmd_last_error(const mtmd_handle* h);
const char* mtmd_status_string(int status);

/* CMolDynConfig::resetToDefaults (MolDynConfig.cpp:289-328) */
void mtmd_config_defaults(mtmd_config* cfg);
/* CMolDynConfig::set (MolDynConfig.cpp:95-287): keys and values as the
 * `moldyn cfg set <key> <value>` command; returns MTMD_OK or MTMD_EINVAL with
 * the reference's "Error: ..." text in err (if err != NULL). */
int mtmd_config_set(mtmd_config* cfg, const char* key, const char* value,
                    char* err, size_t err_len);

/* Profiles: CMDFF*::calc1DForceProfile / calc1DPotentialProfile
 * (MolTwisterMDFFNonBonded.cpp:126-172, MolTwisterMDFFBond.cpp:133-179,
 * MolTwisterMDFFAngle.cpp:210-257, MolTwisterMDFFDih.cpp:272-320).
 * which: 0 non-bonded, 1 bond, 2 angle, 3 dihedral; params holds 4 values
 * (non-bonded) or MTMD_BONDED_PARAMS values (bonded). Writes n points of
 * x, -dU/dx and U as float. This is also what `moldyn ff *profile` prints. */
int mtmd_profile(int which, int kind, const double* params, float x0, float x1,
                 int n, float* x, float* f, float* e);

/* ---- engine ----
 * Replaces CSimulationBox::CSimulationBox + CVelVerlet::CVelVerlet +
 * CMDFFMatrices::CMDFFMatrices (SimulationBox.cpp:27-79, VelVerlet.cpp:31-48,
 * MDFFMatrices.cpp:117-133). */
int  mtmd_create(const mtmd_system* sys, const mtmd_config* cfg, mtmd_handle** out);
void mtmd_destroy(mtmd_handle* h);

/* particles_[k].r_/p_ (SimulationBox.h:77); layouts [3*n] */
int mtmd_set_state(mtmd_handle* h, const double* pos, const double* mom);
int mtmd_get_state(mtmd_handle* h, double* pos, double* mom);
/* NH_T_.eta_/p_eta_ (NHChain.h:59-61); arrays of temperature_nh_len */
int mtmd_set_thermostat(mtmd_handle* h, const double* eta, const double* p_eta);
int mtmd_get_thermostat(mtmd_handle* h, double* eta, double* p_eta);

/* The reference's initial random state, drawn as CSimulationBox::initSystem
 * draws it (SimulationBox.cpp:53-75) with glibc's rand() after srand(seed):
 * thermostat chain positions (NHChain.cpp:126-132), barostat chain positions,
 * momenta (SimulationBox.cpp:98-133: a rand() - 1 per component, a zero
 * replaced by 1, immobile atoms zeroed, scaled to the temperature with unit
 * masses, as the reference's particles still have the default mass of 1 at
 * that point), thermostat chain momenta (NHChain.cpp:134-143), barostat chain
 * momenta, barostat momentum (VelVerlet.cpp:176-183). A first `moldyn run` in a
 * fresh MolTwister process on glibc draws with seed 1. Host only (no device).
 * Outputs, any may be NULL: mom[3 n], eta_T and p_eta_T [temperature_nh_len],
 * eta_P and p_eta_P [pressure_nh_len], p_eps. Pass them to mtmd_create (as
 * sys->mom), mtmd_set_thermostat and mtmd_set_barostat to start from the
 * reference's state. */
int mtmd_reference_draws(const mtmd_system* s, const mtmd_config* c, uint32_t seed, double* mom,
                         double* eta_T, double* p_eta_T, double* eta_P, double* p_eta_P, double* p_eps);

/* NPT barostat of CVelVerlet (VelVerlet.h: eps_, p_eps_) and its chain NH_P_
 * (NHChain.h:59-61, arrays of pressure_nh_len). The reference draws p_eps and
 * the chain at random (VelVerlet.cpp:176-183, NHChain.cpp:126-143); a handle
 * starts them at rest, and these calls set or read any state. Setting eps
 * rescales the box to Lmax (V0 exp(3 eps) / V0)^(1/3) (getSimBox,
 * VelVerlet.cpp:199-211); box[3] receives the current simulation box.
 * mtmd_set_barostat fails outside the npt ensemble. */
int mtmd_set_barostat(mtmd_handle* h, double eps, double p_eps, const double* eta, const double* p_eta);
int mtmd_get_barostat(mtmd_handle* h, double* eps, double* p_eps, double* eta, double* p_eta, double* box);

/* CMDFFMatrices::genNeighList, cell half (MDFFMatrices.cpp:151-159,
 * FunctorGenCellList.cpp:49-169) on the current positions. Any output may be
 * NULL. ixyz [3*n]: cell of each atom; cell_first/cell_count [ncells]
 * (first = -1 for empty cells); cell_list [n]. ncells_xyz[3] receives the grid. */
int mtmd_build_cells(mtmd_handle* h, int32_t* ncells_xyz, int32_t* ixyz,
                     int32_t* cell_first, int32_t* cell_count, int32_t* cell_list);

/* CMDFFMatrices::genNeighList, neighbour half (MDFFMatrices.cpp:161-179,
 * FunctorGenNeighList.cpp:72-191): full lists at the cutoff radius in the
 * reference order. count [n]; list [n*max_per_atom] row-major (may be NULL);
 * *required receives the largest count. Returns MTMD_ENEIGH when a row does
 * not fit max_per_atom (count and required are still filled). */
int mtmd_build_neighbors(mtmd_handle* h, int32_t max_per_atom, int32_t* count,
                         int32_t* list, int32_t* required);

/* CVelVerlet::calcParticleForces (VelVerlet.cpp:152-158,213-222): forces on
 * the current positions, uncapped. F and Fpi [3*n], U [n]; any may be NULL. */
int mtmd_compute_forces(mtmd_handle* h, double* F, double* Fpi, float* U);

/* CMDLoop::runSimulation inner loop (MDLoop.cpp:124-138): nsteps of
 * NH_P, NH_T, velocity Verlet (with force cap in NVT/NVE; the MTK barostat
 * in NPT, VelVerlet.cpp:57-71), NH_T, NH_P, PBC wrap in the current box.
 * The first call computes the initial forces (MDLoop.cpp:121). */
int mtmd_step(mtmd_handle* h, int32_t nsteps);

/* calcTemp/calcPress/calcV/calcUtot/calcCurrentKineticEnergy
 * (SimulationBox.cpp:168-231) on the current state and last forces. */
int mtmd_get_thermo(mtmd_handle* h, mtmd_thermo* out);

/* mtmd_get_thermo for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Same reductions and the same sums, bit for
 * bit, with the block partials in a fixed device array of the library (32 KiB
 * on each GPU, held until the process ends) instead of the handle's buffer;
 * calls no allocating routine, device or host. Needs the forces of the
 * current positions (after mtmd_compute_forces or mtmd_step; MTMD_EINVAL
 * otherwise, since computing them may grow the neighbour lists) and rejects
 * reductions of more workgroups than the array holds. Calls that overlap on
 * one GPU share the array and race. */
int mtmd_get_thermo_noalloc(mtmd_handle* h, mtmd_thermo* out);

/* Per-atom forces and energies of the last force evaluation (capped, as the
 * integrator uses them). */
int mtmd_get_forces(mtmd_handle* h, double* F, float* U);

/* CMDLoop::runOptimization (MDLoop.cpp:143-173) with
 * CSimulationBox::gradientDescentStep (SimulationBox.cpp:197-208): at most
 * max_steps gradient-descent steps r += lr*F; stops when |dU| < accuracy and
 * the step index exceeds out_stride. *steps_done receives the count. */
int mtmd_optimize(mtmd_handle* h, int32_t max_steps, double learning_rate,
                  double accuracy, int32_t* steps_done);

/* Diagnostics: neighbour-list rebuilds since create, current list capacity. */
int mtmd_get_stats(mtmd_handle* h, int64_t* rebuilds, int64_t* steps,
                   int32_t* max_neighbors);

/* Diagnostics, no counterpart in the reference (whose lists are the public
 * members cellList_ / neighList_ of CMDFFMatrices, MDFFMatrices.h:228-242):
 * the FP64/MIXED engine's own neighbour lists as of its last build, at the
 * list radius cutoff + skin, in the engine's word order, as the caller's atom
 * indices (16-bit block-list words are decoded through their block's staged
 * cells). count [n]: words of each row; list [n*max_per_atom] row-major or
 * NULL; *required receives the largest count. Builds the lists first when
 * there are none. MTMD_EINVAL on the EXACT engine (mtmd_build_neighbors gives
 * its lists); MTMD_ENEIGH when a row does not fit max_per_atom. */
int mtmd_get_skin_lists(mtmd_handle* h, int32_t max_per_atom, int32_t* count, int32_t* list, int32_t* required);

/* Blocks until all queued device work of the handle is done. */
int mtmd_synchronize(mtmd_handle* h);

/* Diagnostics, no counterpart in the reference (the reference's only timer is
 * the per-step wall clock of MDLoop.cpp:434-456): average duration in ms of
 * one launch of a kernel group, timed with HIP events on the handle's stream
 * over `reps` back-to-back launches. which: 0 pair-force pass, 1 full
 * neighbour-list build, 2 one MD step, 3 in-place read + write of the whole
 * neighbour-list buffer, 4 * capacity bytes per atom each way (FP64/MIXED
 * engines; calibrates the HBM byte counters), 4 the MIXED engine's tiled
 * force-only pass with FP64 arithmetic (block-frame FP64 positions, FP64
 * square root and sums, the same lists and table; prices precision alone
 * beside which = 0; its forces replace the handle's until the next force
 * call; MIXED tile mode only, no charges). */
int mtmd_bench_kernel(mtmd_handle* h, int which, int reps, double* avg_ms);

/* CMDLoop::appendToXYZFile (MDLoop/MDLoop.cpp:229-246): appends one frame of
 * the current positions, headed "Frame <t>", to the text file at path. */
int mtmd_append_xyz(mtmd_handle* h, const char* path, int32_t t);

/* CMDLoop::appendToDCDFile (MDLoop/MDLoop.cpp:248-258) with
 * CDCDFile::createDCDFileIfNotExists / appendToDCDFile
 * (Utilities/DCDFile.cpp:670-729): writes the CHARMM header (n_sets =
 * n_steps / stride, write frequency = stride, time step in fs) when the file
 * does not exist, then appends one record of the current positions as float
 * with the current box widths narrowed to int. */
int mtmd_append_dcd(mtmd_handle* h, const char* path, int32_t n_steps, int32_t stride);

/* CMDSimulator::run (Simulator/MDSimulator.cpp:35-47): creates the engine,
 * computes initial forces and runs cfg->n_steps, writing the reference's
 * thermo table every out_stride steps to cfg->info_file (and to `log` when
 * not NULL), the XYZ / DCD frames at the same steps when enabled, and the
 * momentum and volume histograms (150 bins, accumulated every step after
 * step 1000) at the end (MDLoop.cpp:107-141,370-431). */
int mtmd_run(const mtmd_config* cfg, const mtmd_system* sys, void* log_file);

/* CMDSimulator::optimize (Simulator/MDSimulator.cpp:49-66): gradient descent
 * (MDLoop.cpp:143-173) with a thermo line and the XYZ / DCD frames every
 * out_stride steps. */
int mtmd_run_optimize(const mtmd_config* cfg, const mtmd_system* sys, void* log_file);

/* ---- topology search ----
 * CMolTwisterStateTools::getAllMDBondsInSystem / getAllMDAnglesInSystem /
 * getAllMDDihedralsInSystem (Cmd/Tools/MolTwisterStateTools.cpp:544-724,
 * 767-920, 993-1140): the bonds, angles or dihedrals of each force-field
 * rule, by the rule's detection criterion (CMDFFBond / CMDFFAngle / CMDFFDih
 * ::EDetCrit, MDFF/Bonds/MolTwisterMDFFBond.h:43):
 *   MTMD_CRIT_ALL_R    all_r_less_than: chains of atoms of the rule's types,
 *                      each link shorter than R0, first atom, then second, ...
 *                      ascending (:577-603, :805-839, :1033-1077);
 *   MTMD_CRIT_MOL_R    mol_r_less_than: the same links, but the second and
 *                      later atoms are taken from the first atom's molecule in
 *                      the depth-first order of CAtom::findAtomsInMolecule
 *                      (MolTwisterAtom.cpp:578-605) (:605-636, :843-888,
 *                      :1084-1144);
 *   MTMD_CRIT_VISIBLE  only_visible_bonds: chains along the bond graph, each
 *                      atom's bonds in its own order, types matched, no
 *                      distance test (:637-666, :889-930, :1145-1198);
 *   MTMD_CRIT_14       only_14_bonds (bonds only): the first atom and each
 *                      atom of the second type three bonds away along the
 *                      graph, other than the three before it (:667-716).
 * Duplicates and reversed duplicates are removed keeping the first, and
 * angles / dihedrals that name one atom twice are dropped
 * (MolecularSystemTools.cpp:25-235). which: 1 bonds, 2 angles, 3 dihedrals;
 * rule r gives the types of its atoms (the first which + 1 entries), R0 and
 * crit; types are indices into the caller's atom-name table (-1 matches no
 * atom). box: periodic widths for the folded distance of
 * CAtom::getDistanceTo2UsingPBC (what the force-field matrices use), or NULL
 * for the plain distance. bond_start [n_atoms + 1] / bond_dest: the bond graph
 * (CAtom::bonds_, in the order CAtom::attachBondTo appended them), needed when
 * a rule uses MTMD_CRIT_MOL_R, _VISIBLE or _14, else NULL. out:
 * [(which + 2) * max_out] entries laid out as mtmd_system's bonds / angles /
 * dihedrals (atoms, then the rule index), or NULL to count only; *n_out
 * receives the count. Returns MTMD_EINVAL when out is too short. Runs on the
 * first HIP device; mtmd_topology_last_error() gives the message of a failed
 * call. */
#define MTMD_CRIT_ALL_R   0
#define MTMD_CRIT_MOL_R   1
#define MTMD_CRIT_VISIBLE 2
#define MTMD_CRIT_14      3

typedef struct mtmd_topo_rule {
    int32_t types[4];
    double  R0;
    int32_t crit;      /* MTMD_CRIT_* */
    int32_t reserved;  /* 0 */
} mtmd_topo_rule;

int mtmd_find_topology(int32_t which, int32_t n_atoms, const double* pos, const int32_t* type, const double* box,
                       const int32_t* bond_start, const int32_t* bond_dest,
                       int32_t n_rules, const mtmd_topo_rule* rules, int32_t max_out, int32_t* out,
                       int32_t* n_out);
const char* mtmd_topology_last_error(void);

/* ---- batches of independent systems (replicas) ----
 * No counterpart in the reference, which advances one system per `moldyn run`
 * (one CSimulationBox + CMDLoop, SimulationBox.cpp:27-133, MDLoop.cpp:107-141).
 * A batch advances n such systems in lockstep: the replicas share ensemble,
 * time step, cutoff, skin, chain length, RESPA steps, force cap and thermostat
 * relaxation time (cfgs[i] must agree on them; temperatures may differ), and
 * each replica has its own box, atoms, momenta and Nose-Hoover chain, with the
 * reference's per-step sequence (MDLoop.cpp:124-138). Arithmetic as the MIXED
 * engine. Replicas take tabulated non-bonded force fields only (all replicas
 * the same list), no charges, no bonded terms, all atoms mobile; boxes with
 * fewer than three cells per axis are allowed (each distinct cell is searched
 * once) but must exceed twice cutoff + skin. */
typedef struct mtmd_batch mtmd_batch;

int  mtmd_batch_create(int n_systems, const mtmd_system* systems, const mtmd_config* cfgs, mtmd_batch** out);
void mtmd_batch_destroy(mtmd_batch* b);
const char* mtmd_batch_last_error(const mtmd_batch* b);
/* nsteps of every replica, energies computed in the last. When every replica
 * holds at most 4096 atoms and 1024 cells and there are at most two atom
 * types, one launch of a persistent kernel runs up to 256 steps: a workgroup
 * per replica, its FP32 positions and the pair table resident in LDS, list
 * rebuilds, forces, integration and chain inside the workgroup; a list that
 * overflows is grown and the launch's steps run again from their starting
 * state, as the reference grows its list and builds it again
 * (MDFFMatrices.cpp:163-178). Otherwise (or
 * with MTMD_BATCH_KERNELS set in the environment at create time) a captured
 * hipGraph of the whole batch's step is replayed, the last step eagerly. */
int mtmd_batch_step(mtmd_batch* b, int32_t nsteps);
/* the same steps launched replica by replica, one launch a replica a step
 * (the cost of running n separate systems); same results, bit for bit */
int mtmd_batch_step_separate(mtmd_batch* b, int32_t nsteps);
int mtmd_batch_synchronize(mtmd_batch* b);
/* thermo of every replica, out[n_systems] (T, P, V, U, K, Etot, sums; U_cut 0) */
int mtmd_batch_get_thermo(mtmd_batch* b, mtmd_thermo* out);
/* state, forces (capped, as integrated) and per-atom energies of replica i */
int mtmd_batch_get_state(mtmd_batch* b, int32_t i, double* pos, double* mom);
int mtmd_batch_get_forces(mtmd_batch* b, int32_t i, double* F, float* U);
/* autotuner over the whole batch; returns the chosen shape's force-pass and
 * list-build times. Persistent kernel: the cell shape (one or two cells per
 * list radius along y and z) times the workgroup size (256, 512, 1024) times
 * the lanes per atom (1, 2, 4), each timed on 8 reps MD steps from the same
 * state, which is restored afterwards. Graph path: the pair-force kernel's
 * block (64..512) and the list build's (64..256), `reps` launches each. */
int mtmd_batch_autotune(mtmd_batch* b, int reps, double* force_ms, double* list_ms);
/* force_block / list_block: the persistent kernel's workgroup size and lanes
 * per atom, or the graph path's force and list kernel blocks */
int mtmd_batch_info(const mtmd_batch* b, int32_t* n_systems, int32_t* n_atoms, int32_t* force_block,
                    int32_t* list_block);
/* the shape in use: persistent kernel (1) or graph path (0), workgroup size,
 * lanes per atom, and cells per list radius along y and z (1, or 2 where a
 * replica's box holds five or more such cells on both axes and three cells
 * along x) */
int mtmd_batch_tuning(const mtmd_batch* b, int32_t* persistent, int32_t* block, int32_t* lanes, int32_t* cell_split);

/* ---- force matching ----
 * No counterpart in the reference (it has no gradients). Loss
 *   L = sum_frames sum_i |F_i(theta) - F_i^ref|^2 / (3 N n_frames)
 * and dL/dtheta by a hand-written adjoint, for the analytic forms of the
 * reference's MDFF classes: LJ 4 eps [(s/r)^12 - (s/r)^6] between atoms of
 * different molecules within `cutoff` (MolTwisterMDFFNonBonded_LJ.cpp:46-60),
 * harmonic bonds k (r - r0)^2 (MolTwisterMDFFBond_Harm.cpp:43-51) and angles
 * k (theta - theta0)^2 (MolTwisterMDFFAngle_Harm.cpp:47-55), minimum image.
 * The system gives topology, types and box (its force-field parameters are
 * ignored; kinds must be LJ / BOND_HARM / ANGLE_HARM, at most 4 of each).
 * theta = for each non-bonded entry (eps, sigma), for each bond type (k, r0),
 * for each angle type (k, theta0 in degrees). frames and target are
 * [n_frames][3 n] and stay in device memory; pairs are searched again in
 * every pass (no stored lists). fp64 = 0: FP32 pair terms with FP64
 * accumulation; fp64 = 1: FP64 throughout. */
typedef struct mtmd_fm mtmd_fm;

int  mtmd_fm_create(const mtmd_system* sys, double cutoff, int n_frames, const double* frames,
                    const double* target, mtmd_fm** out);
void mtmd_fm_destroy(mtmd_fm* h);
const char* mtmd_fm_last_error(const mtmd_fm* h);
int  mtmd_fm_num_params(const mtmd_fm* h);
int  mtmd_fm_set_target(mtmd_fm* h, const double* target);
/* target := model forces of every frame at `params` (synthetic references) */
int  mtmd_fm_target_from_params(mtmd_fm* h, const double* params);
/* model forces of one frame, [3 n] */
int  mtmd_fm_forces(mtmd_fm* h, const double* params, int32_t frame, int32_t fp64, double* F);
int  mtmd_fm_loss_grad(mtmd_fm* h, const double* params, int32_t fp64, double* loss, double* grad);

#ifdef __cplusplus
}
#endif
#endif /* MTMD_H */
