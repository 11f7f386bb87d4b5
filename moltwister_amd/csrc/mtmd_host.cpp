// mtmd_host.cpp - C ABI of libmtmd.so (include/mtmd.h): configuration, system
// set-up, device memory, and the step driver for the three engines.
//
// Replaces, on the host side, CMolDynConfig (Config/MolDynConfig.cpp),
// CMDFFMatrices::prepareFFMatrices (ForceFields/MDFFMatrices.cpp:181-536),
// CSimulationBox (SimulationBox/SimulationBox.cpp), CVelVerlet
// (Integrators/VelVerlet.cpp), CMDLoop::runSimulation/runOptimization
// (MDLoop/MDLoop.cpp:107-173) and CMDSimulator (Simulator/MDSimulator.cpp).
// All paths are relative to Cmd/MolTwisterCmdMolDyn/ of the MolTwister tree
// unless stated otherwise. The state stays resident in HBM; only thermo
// scalars and explicitly requested arrays cross PCIe.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/mtmd.h"
#include "mtmd_internal.h"
#include "mtmd_output.h"

namespace {

thread_local std::string g_create_error;

struct DevBuf {
    std::vector<void*> ptrs;
    template <class T> T* alloc(size_t count)
    {
        void* p = nullptr;
        if (count == 0) count = 1;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    void release()
    {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
};

uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double urand(uint64_t& s) { return double(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

} // namespace


struct mtmd_handle {
    int n = 0, T = 0;
    int device = 0;               // the HIP device current at create (stream, memory)
    mtmd_config cfg{};
    int engine = MTMD_ENGINE_EXACT;
    MtDev d{};
    MtParams p{};
    hipStream_t stream = nullptr;
    DevBuf mem;
    std::string err;

    // NPT: the barostat of CVelVerlet (VelVerlet.h: P_, p_eps_, eps_, W_, V0_)
    // and its chain NH_P_, advanced on the host with the reference's scalar
    // arithmetic; Lmax is the initial box, Lrb the box of the last list build
    bool npt = false;
    double Lmax[3] = {0, 0, 0}, Lrb[3] = {0, 0, 0};
    double Pext = 0, W = 0, V0 = 0, eps = 0, peps = 0;
    int Mp = 0, nrespP = 0;
    double tauP = 0;
    std::vector<double> etaP, petaP, QP;
    unsigned int rebuild_count = 0;

    // atom types and type names (XYZ frames)
    std::vector<int32_t> type_host;
    std::vector<std::string> type_names;

    // thermostat (NHChain.h:55-66)
    int M = 0, nresp = 0;
    double Tred = 0, tau = 0, g = 0, w[3] = {0, 0, 0};
    int nh_scale_slot = -1;

    bool have_forces = false;     // forces of the current positions exist
    bool ke_valid = false;        // d.red holds per-block sums of p.p/m of the current momenta
    bool ke_slot = false;         // FP64/MIXED NVT: nh[3M + 1] holds p.p/m of the current momenta (formed by
                                  // the last chain propagation), taken by the next call's first chain
    int64_t rebuilds = 0, steps = 0;

    // FP64/MIXED engines
    bool fast = false;
    bool need_sort = true;
    int sort_every = 100;
    int steps_since_sort = 0;
    double half_skin2 = 0;
    std::vector<int> perm_host;   // valid after download_perm()
    int* h_status = nullptr;      // pinned status words

    // captured plain steps (no re-sort, not the last of a call), one step or a
    // block of kBlockSteps a graph (a replay costs a few microseconds of GPU
    // time of its own, which a block shares). The kernels take the buffers and the rebuild-flag slot by
    // value, and a step may swap them (the re-sort swaps the slot buffers, a
    // fused NVE pass pos/pos2 and posf/posf2, a front half flips dslot; the
    // three together reach four states): a graph is found by the state it
    // starts from and leaves the state its captured step left; dropped when
    // the neighbour buffer moves
    struct StepState {
        const void* pos = nullptr; const void* pos2 = nullptr; const void* mom = nullptr;
        const void* posf = nullptr; const void* posf2 = nullptr; int dslot = 0;
        bool operator==(const StepState& o) const
        {
            return pos == o.pos && pos2 == o.pos2 && mom == o.mom && posf == o.posf && posf2 == o.posf2 &&
                   dslot == o.dslot;
        }
    };
    struct StepGraph {
        bool valid = false;
        int nstep = 1;                // plain steps captured in the graph
        StepState from, to;
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
    };
    // four states, each with a graph of one step and one of kBlockSteps steps
    static constexpr int kGraphsPerEnsemble = 8;
    static constexpr int kBlockSteps = 8;
    StepGraph graphs[2 * kGraphsPerEnsemble];
    int graph_next[2] = {0, 0};
    StepState step_state() const { return {d.pos, d.pos2, d.mom, d.posf, d.posf2, p.dslot}; }
    void set_step_state(const StepState& st)
    {
        d.pos = (double4*)st.pos; d.pos2 = (double4*)st.pos2; d.mom = (double4*)st.mom;
        d.posf = (float4*)st.posf; d.posf2 = (float4*)st.posf2; p.dslot = st.dslot;
    }
    bool use_graphs = true;
    bool use_fuse = true;
    bool list_slack = true;       // lists regrow with a margin over the longest row
    // the state a multi-step call of the FP64/MIXED engines starts from (device
    // copies in original atom order, the thermostat chain, the host-side
    // barostat and counters), restored to re-run the call when a list build
    // inside it overflowed
    struct Snapshot {
        double4* pos = nullptr;
        double4* mom = nullptr;
        double* nh = nullptr;
        double eps = 0, peps = 0, L[3] = {0, 0, 0};
        std::vector<double> etaP, petaP;
        int64_t steps = 0;
        int steps_since_sort = 0;
        bool ke_slot = false;
    } snap;
    void drop_graphs()
    {
        for (auto& sg : graphs) {
            if (sg.x) (void)hipGraphExecDestroy(sg.x);
            if (sg.g) (void)hipGraphDestroy(sg.g);
            sg = StepGraph();
        }
    }

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
};

#define HIPCHK(h, expr)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return (h)->fail(MTMD_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// library
// ---------------------------------------------------------------------------
extern "C" int mtmd_abi_version(void) { return MTMD_ABI_VERSION; }

extern "C" const char* mtmd_last_error(const mtmd_handle* h)
{
    return h ? h->err.c_str() : g_create_error.c_str();
}

extern "C" const char* mtmd_status_string(int s)
{
    switch (s) {
    case MTMD_OK: return "ok";
    case MTMD_EINVAL: return "invalid argument";
    case MTMD_EHIP: return "HIP runtime error";
    case MTMD_ENEIGH: return "neighbour list overflow";
    case MTMD_ENAN: return "non-finite value";
    case MTMD_EBOX: return "box too small";
    case MTMD_ENOMEM: return "out of memory";
    case MTMD_ENODEV: return "no HIP device";
    }
    return "unknown status";
}

// CMolDynConfig::resetToDefaults (Config/MolDynConfig.cpp:289-328)
extern "C" void mtmd_config_defaults(mtmd_config* c)
{
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->timestep = 1.0;
    c->out_stride = 100;
    c->ensemble = MTMD_NVT;
    c->temperature = 298.0;
    c->temperature_relax = 20.0 * c->timestep;
    c->temperature_nh_len = 4;
    c->temperature_respa = 4;
    c->pressure = 1.0;
    c->pressure_relax = 5000.0 * c->timestep;
    c->pressure_nh_len = 4;
    c->pressure_respa = 4;
    c->cutoff_radius = 10.0;
    c->neigh_shell = 2.0;
    c->cutoff_force = 1000.0;
    c->n_steps = 50000;
    std::snprintf(c->info_file, sizeof c->info_file, "%s", "out.txt");
    std::snprintf(c->xyz_file, sizeof c->xyz_file, "%s", "traj.xyz");
    std::snprintf(c->dcd_file, sizeof c->dcd_file, "%s", "traj.dcd");
    std::snprintf(c->xtc_file, sizeof c->xtc_file, "%s", "traj.xtc");
    std::snprintf(c->pdistr_file, sizeof c->pdistr_file, "%s", "distr");
    std::snprintf(c->vdistr_file, sizeof c->vdistr_file, "%s", "distr");
    c->include_xyz = 0;
    c->include_dcd = 1;
    c->include_xtc = 0;
    c->xtc_precision = 1000.0f;
    c->include_pdistr = 0;
    c->include_vdistr = 0;
    c->max_pdistr = 0.3;
    c->max_vdistr = 100.0E3;
    c->scale12 = 0.0;
    c->scale13 = 0.0;
    c->scale14 = 0.5;
    c->scale1N = 0.0;
    c->opt_learning_rate = 1e-6;
    c->opt_max_steps = 5000;
    c->opt_accuracy = 0.0;
    c->engine = MTMD_ENGINE_EXACT;
    c->coulomb = MTMD_COUL_AUTO;
    c->skin = 0.3;
}

// CMolDynConfig::set (Config/MolDynConfig.cpp:95-287): the reference's keys,
// atoi/atof parsing and error strings, plus the engine options.
extern "C" int mtmd_config_set(mtmd_config* c, const char* key, const char* value, char* err, size_t err_len)
{
    auto fail = [&](const char* msg) {
        if (err && err_len) std::snprintf(err, err_len, "%s", msg);
        return MTMD_EINVAL;
    };
    if (!c || !key || !value) return fail("Error: unknown parameter!");
    const std::string k = key, v = value;
    struct IntKey { const char* name; int32_t* field; };
    struct DblKey { const char* name; double* field; };
    const IntKey ik[] = {{"timesteps", &c->n_steps}, {"outstride", &c->out_stride},
                         {"temperaturenhlen", &c->temperature_nh_len}, {"temperaturerespa", &c->temperature_respa},
                         {"pressurenhlen", &c->pressure_nh_len}, {"pressurerespa", &c->pressure_respa},
                         {"optmaxsteps", &c->opt_max_steps}, {"blocksize", &c->block_size}};
    const DblKey dk[] = {{"timestep", &c->timestep}, {"temperature", &c->temperature},
                         {"temperaturerelax", &c->temperature_relax}, {"pressure", &c->pressure},
                         {"pressurerelax", &c->pressure_relax}, {"cutoffradius", &c->cutoff_radius},
                         {"neighshell", &c->neigh_shell}, {"cutoffforce", &c->cutoff_force},
                         {"scale12", &c->scale12}, {"scale13", &c->scale13}, {"scale14", &c->scale14},
                         {"scale1N", &c->scale1N}, {"optlearningrate", &c->opt_learning_rate},
                         {"optaccuracy", &c->opt_accuracy}, {"skin", &c->skin}};
    for (const auto& e : ik) if (k == e.name) { *e.field = std::atoi(value); return MTMD_OK; }
    for (const auto& e : dk) if (k == e.name) { *e.field = std::atof(value); return MTMD_OK; }
    if (k == "ensemble") {
        if (v == "nvt") c->ensemble = MTMD_NVT;
        else if (v == "npt") c->ensemble = MTMD_NPT;
        else if (v == "nve") c->ensemble = MTMD_NVE;
        else return fail("Error: unknown ensemble!");
        return MTMD_OK;
    }
    struct FileKey { const char* name; char* field; size_t size; };
    const FileKey fk[] = {{"infofile", c->info_file, sizeof c->info_file}, {"xyzfile", c->xyz_file, sizeof c->xyz_file},
                          {"dcdfile", c->dcd_file, sizeof c->dcd_file}, {"xtcfile", c->xtc_file, sizeof c->xtc_file},
                          {"pdistrfile", c->pdistr_file, sizeof c->pdistr_file},
                          {"vdistrfile", c->vdistr_file, sizeof c->vdistr_file}};
    for (const auto& e : fk)
        if (k == e.name) { std::snprintf(e.field, e.size, "%s", value); return MTMD_OK; }
    if (k == "xtcprecision") { c->xtc_precision = (float)std::atof(value); return MTMD_OK; }
    if (k == "maxpdistr") { c->max_pdistr = std::atof(value); return MTMD_OK; }
    if (k == "maxvdistr") { c->max_vdistr = std::atof(value); return MTMD_OK; }
    struct FlagKey { const char* name; int32_t* field; };
    const FlagKey gk[] = {{"includexyzfile", &c->include_xyz}, {"includedcdfile", &c->include_dcd},
                          {"includextcfile", &c->include_xtc}, {"includepdistrfile", &c->include_pdistr},
                          {"includevdistrfile", &c->include_vdistr}, {"verboseoutput", &c->verbose}};
    for (const auto& e : gk) {
        if (k != e.name) continue;
        if (v != "yes" && v != "no") {
            char msg[160];
            std::snprintf(msg, sizeof msg, "Error: '%s' parameter should be either 'yes' or 'no'!", key);
            return fail(msg);
        }
        *e.field = (v == "yes");
        return MTMD_OK;
    }
    if (k == "engine") {
        if (v == "exact") c->engine = MTMD_ENGINE_EXACT;
        else if (v == "fp64") c->engine = MTMD_ENGINE_FP64;
        else if (v == "mixed") c->engine = MTMD_ENGINE_MIXED;
        else return fail("Error: unknown engine!");
        return MTMD_OK;
    }
    if (k == "coulomb") {
        if (v == "auto") c->coulomb = MTMD_COUL_AUTO;
        else if (v == "allpairs") c->coulomb = MTMD_COUL_ALLPAIRS;
        else if (v == "cutoff") c->coulomb = MTMD_COUL_CUTOFF;
        else return fail("Error: unknown coulomb mode!");
        return MTMD_OK;
    }
    return fail("Error: unknown parameter!");
}

// ---------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------
namespace {

template <class T> int upload(mtmd_handle* h, T*& dst, const std::vector<T>& src)
{
    dst = h->mem.alloc<T>(src.size());
    if (!dst) return h->fail(MTMD_ENOMEM, "device allocation of %zu bytes failed", src.size() * sizeof(T));
    if (!src.empty()) HIPCHK(h, hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return MTMD_OK;
}

// a caller-supplied profile: x[], f[], e[] of MT_NPTS floats each
std::vector<float4> table_from_profile(const float* pr)
{
    std::vector<float4> t(MT_NPTS);
    for (int i = 0; i < MT_NPTS; i++) t[i] = make_float4(pr[i], pr[MT_NPTS + i], pr[2 * MT_NPTS + i], 0.0f);
    return t;
}

std::vector<float4> make_table(int which, int kind, const double* p, float x0, float x1)
{
    std::vector<float> x(MT_NPTS), f(MT_NPTS), e(MT_NPTS);
    std::vector<float4> t(MT_NPTS);
    mtmd_profile(which, kind, p, x0, x1, MT_NPTS, x.data(), f.data(), e.data());
    for (int i = 0; i < MT_NPTS; i++) t[i] = make_float4(x[i], f[i], e[i], 0.0f);
    return t;
}

// segment coefficients a, b of interpolateLinear (FunctorCalcForce.cpp:457-465)
// for force and energy, for the MIXED engine's single-multiply evaluation
void segment_coeffs(const float4* t, int i, float& af, float& bf, float& ae, float& be)
{
    float denom = t[i].x - t[i + 1].x;
    if (denom == 0.0f) denom = 3.0f * 1.17549435e-38f;
    af = (t[i].y - t[i + 1].y) / denom;
    bf = t[i].y - af * t[i].x;
    ae = (t[i].z - t[i + 1].z) / denom;
    be = t[i].z - ae * t[i].x;
}

} // namespace

// table builders shared with the batch engine (mtmd_batch.hip): a profile
// from parameters or from the caller's points; empty for an unknown kind
std::vector<float4> mtmd_host_table(int which, int kind, const double* prm, const float* profile, float x0, float x1)
{
    if (kind == MTMD_FF_PROFILE) return profile ? table_from_profile(profile) : std::vector<float4>();
    if (mtmd_profile(which, kind, prm, x0, x1, 2, nullptr, nullptr, nullptr) != MTMD_OK) return {};
    return make_table(which, kind, prm, x0, x1);
}

void mtmd_host_segment(const float4* t, int i, float& af, float& bf, float& ae, float& be)
{
    segment_coeffs(t, i, af, bf, ae, be);
}

static int setup(mtmd_handle* h, const mtmd_system* s)
{
    const int n = s->n_atoms, T = s->n_types;
    mtmd_config& c = h->cfg;
    h->n = n; h->T = T;
    h->engine = c.engine;
    h->fast = (c.engine == MTMD_ENGINE_FP64 || c.engine == MTMD_ENGINE_MIXED);
    {
        // MTMD_GRAPHS=0 launches every step's kernels directly (same results)
        const char* g = std::getenv("MTMD_GRAPHS");
        h->use_graphs = !(g && std::strcmp(g, "0") == 0);
        // MTMD_FUSE=0 keeps the MIXED NVE integration out of the pair pass
        // (same results: the same per-atom arithmetic in a kernel of its own)
        const char* f = std::getenv("MTMD_FUSE");
        h->use_fuse = !(f && std::strcmp(f, "0") == 0);
        // MTMD_LIST_SLACK=0 sizes the lists to the longest row of the latest
        // build with no margin, starting from 16 words (every list growth path,
        // the re-run of a multi-step call included, then runs early)
        const char* ls = std::getenv("MTMD_LIST_SLACK");
        h->list_slack = !(ls && std::strcmp(ls, "0") == 0);
    }
    if (n <= 0 || T <= 0 || !s->pos || !s->mass || !s->charge || !s->type)
        return h->fail(MTMD_EINVAL, "system needs atoms, types, positions, masses and charges");
    if (n >= (1 << MT_NW_IDX_BITS)) return h->fail(MTMD_EINVAL, "at most %d atoms", (1 << MT_NW_IDX_BITS) - 1);
    if (c.ensemble != MTMD_NVT && c.ensemble != MTMD_NVE && c.ensemble != MTMD_NPT)
        return h->fail(MTMD_EINVAL, "unknown ensemble");
    if (c.cutoff_radius <= 0) return h->fail(MTMD_EINVAL, "cutoffradius must be positive");
    if (c.timestep <= 0) return h->fail(MTMD_EINVAL, "timestep must be positive");
    if (c.ensemble != MTMD_NVE && (c.temperature_nh_len < 1 || c.temperature_respa < 1 || c.temperature <= 0 ||
                                   c.temperature_relax <= 0))
        return h->fail(MTMD_EINVAL, "NVT and NPT need temperature, temperaturerelax, temperaturenhlen and "
                                    "temperaturerespa > 0");
    if (c.ensemble == MTMD_NPT && (c.pressure_nh_len < 1 || c.pressure_respa < 1 || c.pressure_relax <= 0))
        return h->fail(MTMD_EINVAL, "NPT needs pressurerelax, pressurenhlen and pressurerespa > 0");
    for (int k = 0; k < n; k++)
        if (s->type[k] < 0 || s->type[k] >= T || !(s->mass[k] > 0) || !std::isfinite(s->pos[3 * k]) ||
            !std::isfinite(s->pos[3 * k + 1]) || !std::isfinite(s->pos[3 * k + 2]))
            return h->fail(MTMD_EINVAL, "atom %d: type out of range, non-positive mass or non-finite position", k);
    for (int i = 0; i < 3; i++)
        if (!(s->box[i] > 0)) return h->fail(MTMD_EINVAL, "box widths must be positive");

    MtParams& p = h->p;
    p.rc = (float)c.cutoff_radius;
    p.rc2f = p.rc;
    p.rc2f *= p.rc2f;
    for (int i = 0; i < 3; i++) {
        p.L[i] = s->box[i];
        p.Lf[i] = (float)s->box[i];
        p.lowf[i] = -p.Lf[i] / 2.0f;
        // CFunctorCalcForce::pbc_ (FunctorCalcForce.cpp:41)
        const double Lf = (double)p.Lf[i];
        p.wbox[i] = (Lf / 2.0) - (-Lf / 2.0);
        // CFunctorGenNeighList::pbc_ (FunctorGenNeighList.cpp:63-69)
        p.wnl[i] = (double)(p.lowf[i] + p.Lf[i]) - (double)p.lowf[i];
    }
    p.scale12 = (float)c.scale12; p.scale13 = (float)c.scale13;
    p.scale14 = (float)c.scale14; p.scale1N = (float)c.scale1N;
    p.dt = c.timestep / MT_CONV_t;
    p.Fcut = c.cutoff_force;
    bool any_q = false;
    for (int k = 0; k < n; k++) if ((float)s->charge[k] != 0.0f) any_q = true;
    int cm = c.coulomb;
    if (cm == MTMD_COUL_AUTO) cm = h->fast ? MTMD_COUL_CUTOFF : MTMD_COUL_ALLPAIRS;
    p.coulomb_allpairs = (cm == MTMD_COUL_ALLPAIRS);
    p.coulomb_on = any_q ? 1 : 0;

    // --- tables (MDFFMatrices.cpp:249-382) ---
    std::vector<float4> nbtab((size_t)T * T * MT_MAXFF * MT_NPTS, make_float4(0, 0, 0, 0));
    std::vector<int> nbcnt((size_t)T * T, 0);
    for (int col = 0; col < T; col++)
        for (int row = 0; row < T; row++) {
            int cnt = 0;
            for (int i = 0; i < s->n_nonbonded; i++) {
                const auto& e = s->nonbonded[i];
                const bool match = (e.type_a == col && e.type_b == row) || (e.type_b == col && e.type_a == row);
                if (!match) continue;
                if (cnt >= MT_MAXFF)
                    return h->fail(MTMD_EINVAL, "more than %d non-bonded force fields for types %d-%d", MT_MAXFF,
                                   row, col);
                if (e.kind != MTMD_NB_LJ && e.kind != MTMD_NB_BUCK && e.kind != MTMD_NB_LJ1208 &&
                    e.kind != MTMD_NB_LJBUCK && e.kind != MTMD_FF_PROFILE)
                    return h->fail(MTMD_EINVAL, "non-bonded entry %d: unknown kind %d", i, e.kind);
                if (e.kind == MTMD_FF_PROFILE && !e.profile)
                    return h->fail(MTMD_EINVAL, "non-bonded entry %d: profile kind without a profile", i);
                if (e.kind == MTMD_FF_PROFILE && h->fast &&
                    (e.profile[0] != 0.01f || e.profile[MT_NPTS - 1] != p.rc))
                    return h->fail(MTMD_EINVAL, "non-bonded entry %d: the profile must span 0.01 .. cutoff", i);
                auto t = e.kind == MTMD_FF_PROFILE ? table_from_profile(e.profile)
                                                   : make_table(0, e.kind, e.p, 0.01f, p.rc);
                std::copy(t.begin(), t.end(), nbtab.begin() + ((size_t)(row * T + col) * MT_MAXFF + cnt) * MT_NPTS);
                cnt++;
            }
            nbcnt[row * T + col] = cnt;
        }
    auto bonded_tables = [&](int which, const mtmd_bonded_ff* ff, int cnt, float x0, float x1,
                             std::vector<float4>& out) -> int {
        out.assign((size_t)std::max(cnt, 1) * MT_NPTS, make_float4(0, 0, 0, 0));
        for (int i = 0; i < cnt; i++) {
            if (ff[i].kind == MTMD_FF_PROFILE) {
                if (!ff[i].profile)
                    return h->fail(MTMD_EINVAL, "bonded force field %d (table %d): profile kind without a profile",
                                   i, which);
                auto t = table_from_profile(ff[i].profile);
                std::copy(t.begin(), t.end(), out.begin() + (size_t)i * MT_NPTS);
                continue;
            }
            // a two-point profile with no outputs only validates the kind
            if (mtmd_profile(which, ff[i].kind, ff[i].p, x0, x1, 2, nullptr, nullptr, nullptr) != MTMD_OK)
                return h->fail(MTMD_EINVAL, "bonded force field %d (table %d): unknown kind %d", i, which, ff[i].kind);
            auto t = make_table(which, ff[i].kind, ff[i].p, x0, x1);
            std::copy(t.begin(), t.end(), out.begin() + (size_t)i * MT_NPTS);
        }
        return MTMD_OK;
    };
    std::vector<float4> btab, atab, dtab;
    const float two_pi = float(2.0 * M_PI);
    int rc_;
    if ((rc_ = bonded_tables(1, s->bond_ff, s->n_bond_ff, 0.01f, p.rc, btab))) return rc_;
    if ((rc_ = bonded_tables(2, s->angle_ff, s->n_angle_ff, 0.0f, two_pi, atab))) return rc_;
    if ((rc_ = bonded_tables(3, s->dihedral_ff, s->n_dihedral_ff, 0.0f, two_pi, dtab))) return rc_;

    // --- molecule indices from the bond graph (MolecularSystemTools.cpp:238-280) ---
    std::vector<int> par(n);
    std::iota(par.begin(), par.end(), 0);
    auto find = [&](int a) { while (par[a] != a) { par[a] = par[par[a]]; a = par[a]; } return a; };
    for (int j = 0; j < s->n_bonds; j++) {
        int a = s->bonds[3 * j], b = s->bonds[3 * j + 1];
        if (a < 0 || b < 0 || a >= n || b >= n) return h->fail(MTMD_EINVAL, "bond %d: atom index out of range", j);
        a = find(a); b = find(b);
        if (a != b) par[std::max(a, b)] = std::min(a, b);
    }
    std::vector<int> mol(n);
    for (int k = 0; k < n; k++) mol[k] = find(k);

    // --- per-atom bonded lists in the reference order (MDFFMatrices.cpp:384-520):
    // bucket the global lists by atom, keeping global order inside each bucket ---
    for (int j = 0; j < s->n_bonds; j++)
        if (s->bonds[3 * j + 2] >= s->n_bond_ff) return h->fail(MTMD_EINVAL, "bond %d: force field out of range", j);
    for (int j = 0; j < s->n_angles; j++) {
        for (int q = 0; q < 3; q++)
            if (s->angles[4 * j + q] < 0 || s->angles[4 * j + q] >= n)
                return h->fail(MTMD_EINVAL, "angle %d: atom index out of range", j);
        if (s->angles[4 * j + 3] >= s->n_angle_ff) return h->fail(MTMD_EINVAL, "angle %d: force field out of range", j);
    }
    for (int j = 0; j < s->n_dihedrals; j++) {
        for (int q = 0; q < 4; q++)
            if (s->dihedrals[5 * j + q] < 0 || s->dihedrals[5 * j + q] >= n)
                return h->fail(MTMD_EINVAL, "dihedral %d: atom index out of range", j);
        if (s->dihedrals[5 * j + 4] >= s->n_dihedral_ff)
            return h->fail(MTMD_EINVAL, "dihedral %d: force field out of range", j);
    }
    std::vector<std::vector<int>> bonds_of(n), angles_of(n), dihs_of(n);
    for (int j = 0; j < s->n_bonds; j++) {
        const int a = s->bonds[3 * j], b = s->bonds[3 * j + 1];
        bonds_of[a].push_back(j);
        if (b != a) bonds_of[b].push_back(j);
    }
    for (int j = 0; j < s->n_angles; j++) {
        const int* a = s->angles + 4 * j;
        for (int q = 0; q < 3; q++) {
            bool dup = false;
            for (int r = 0; r < q; r++) dup |= (a[r] == a[q]);
            if (!dup) angles_of[a[q]].push_back(j);
        }
    }
    for (int j = 0; j < s->n_dihedrals; j++) {
        const int* a = s->dihedrals + 5 * j;
        for (int q = 0; q < 4; q++) {
            bool dup = false;
            for (int r = 0; r < q; r++) dup |= (a[r] == a[q]);
            if (!dup) dihs_of[a[q]].push_back(j);
        }
    }
    std::vector<int> b_first(n), b_num(n), b_j, b_type;
    std::vector<int> a_first(n), a_num(n), a_i, a_j, a_type;
    std::vector<uint8_t> a_center;
    std::vector<int> d_first(n), d_num(n), d_i, d_j, d_l, d_type;
    std::vector<uint8_t> d_center;
    for (int k = 0; k < n; k++) {
        b_first[k] = (int)b_j.size();
        for (int j : bonds_of[k]) {
            const int a1 = s->bonds[3 * j], a2 = s->bonds[3 * j + 1], t = s->bonds[3 * j + 2];
            int to = -1;
            if (a1 == k) to = a2;
            if (a2 == k) to = a1;
            if (to != -1 && t != -1) { b_j.push_back(to); b_type.push_back(t); }
        }
        b_num[k] = (int)b_j.size() - b_first[k];
        a_first[k] = (int)a_i.size();
        for (int j : angles_of[k]) {
            const int* a = s->angles + 4 * j;
            int I = -1, J = -1, centre = 0;
            if (a[0] == k) { I = a[1]; J = a[2]; }
            if (a[1] == k) { centre = 1; I = a[0]; J = a[2]; }
            if (a[2] == k) { I = a[1]; J = a[0]; }
            if (I != -1 && J != -1 && a[3] != -1) {
                a_i.push_back(I); a_j.push_back(J); a_type.push_back(a[3]); a_center.push_back((uint8_t)centre);
            }
        }
        a_num[k] = (int)a_i.size() - a_first[k];
        d_first[k] = (int)d_i.size();
        for (int j : dihs_of[k]) {
            const int* a = s->dihedrals + 5 * j;
            int I = -1, J = -1, L = -1, centre = 0;
            if (a[0] == k) { I = a[1]; J = a[2]; L = a[3]; }
            if (a[1] == k) { centre = 1; I = a[0]; J = a[2]; L = a[3]; }
            if (a[2] == k) { centre = 1; I = a[3]; J = a[1]; L = a[0]; }
            if (a[3] == k) { I = a[2]; J = a[1]; L = a[0]; }
            if (I != -1 && J != -1 && L != -1 && a[4] != -1) {
                d_i.push_back(I); d_j.push_back(J); d_l.push_back(L); d_type.push_back(a[4]);
                d_center.push_back((uint8_t)centre);
            }
        }
        d_num[k] = (int)d_i.size() - d_first[k];
    }
    // xmol: atom has a scaled partner (1-2, 1-3, 1-4) outside its own molecule
    std::vector<uint8_t> xmol(n, 0);
    h->d.has_molecules = 0;
    for (int k = 0; k < n; k++) if (mol[k] != k) { h->d.has_molecules = 1; break; }
    for (int k = 0; k < n; k++) {
        for (int j = 0; j < b_num[k]; j++) if (mol[b_j[b_first[k] + j]] != mol[k]) xmol[k] = 1;
        for (int j = 0; j < a_num[k]; j++)
            if (!a_center[a_first[k] + j] && mol[a_j[a_first[k] + j]] != mol[k]) xmol[k] = 1;
        for (int j = 0; j < d_num[k]; j++)
            if (!d_center[d_first[k] + j] && mol[d_l[d_first[k] + j]] != mol[k]) xmol[k] = 1;
    }

    h->d.has_xmol = 0;
    for (int k = 0; k < n; k++) if (xmol[k]) { h->d.has_xmol = 1; break; }

    // --- special pairs for the FP64/MIXED engines: every same-molecule pair and
    // every scaled cross-molecule partner, with the reference's scale
    // (scaleForcesAndPotentials, FunctorCalcForce.cpp:95-157); zero scales dropped
    // unless all-pairs Coulomb needs them ---
    std::vector<int> sp_first(n, 0), sp_num(n, 0), sp_partner;
    std::vector<float> sp_scale;
    if (h->fast) {
        std::vector<std::vector<int>> members;
        {
            std::vector<int> root_to_idx(n, -1);
            for (int k = 0; k < n; k++) {
                int r = mol[k];
                if (root_to_idx[r] < 0) { root_to_idx[r] = (int)members.size(); members.emplace_back(); }
                members[root_to_idx[r]].push_back(k);
            }
            for (auto& m : members)
                if (m.size() > 512)
                    return h->fail(MTMD_EINVAL, "molecule of %zu atoms: the fp64/mixed engines take molecules of at "
                                                "most 512 atoms (use engine exact)", m.size());
            for (int k = 0; k < n; k++) {
                sp_first[k] = (int)sp_partner.size();
                auto scale_of = [&](int i) -> float {
                    for (int j = 0; j < b_num[k]; j++) if (b_j[b_first[k] + j] == i) return p.scale12;
                    for (int j = 0; j < a_num[k]; j++)
                        if (!a_center[a_first[k] + j] && a_j[a_first[k] + j] == i) return p.scale13;
                    for (int j = 0; j < d_num[k]; j++)
                        if (!d_center[d_first[k] + j] && d_l[d_first[k] + j] == i) return p.scale14;
                    return (mol[i] == mol[k]) ? p.scale1N : 1.0f;
                };
                std::vector<int> cand = members[root_to_idx[mol[k]]];
                if (xmol[k]) {
                    for (int j = 0; j < b_num[k]; j++) cand.push_back(b_j[b_first[k] + j]);
                    for (int j = 0; j < a_num[k]; j++) if (!a_center[a_first[k] + j]) cand.push_back(a_j[a_first[k] + j]);
                    for (int j = 0; j < d_num[k]; j++) if (!d_center[d_first[k] + j]) cand.push_back(d_l[d_first[k] + j]);
                }
                std::sort(cand.begin(), cand.end());
                cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
                for (int i : cand) {
                    if (i == k) continue;
                    float sc = scale_of(i);
                    // all-pairs Coulomb sums every pair with weight 1, so a zero
                    // scale is kept there: the special pass subtracts the pair
                    if (sc != 0.0f || (p.coulomb_on && p.coulomb_allpairs)) {
                        sp_partner.push_back(i);
                        sp_scale.push_back(sc);
                    }
                }
                sp_num[k] = (int)sp_partner.size() - sp_first[k];
            }
        }
        // cross-molecule partners must not also appear in the neighbour list:
        // the fast neighbour build skips same-molecule pairs and, for xmol
        // atoms, their bonded partners
    }

    // --- cell grid ---
    MtDev& d = h->d;
    d.n = n; d.T = T;
    if (!h->fast) {
        // CCellList::init (MDFFMatrices.cpp:43-63) on the initial box
        const float R = (float)c.cutoff_radius + (float)c.neigh_shell;
        for (int i = 0; i < 3; i++) d.nc[i] = (int)std::floor((double)(p.Lf[i] / R));
        if (d.nc[0] < 1 || d.nc[1] < 1 || d.nc[2] < 1)
            return h->fail(MTMD_EBOX, "box smaller than cutoffradius + neighshell: no cell fits");
        if (c.verbose && (d.nc[0] < 3 || d.nc[1] < 3 || d.nc[2] < 3))
            std::fprintf(stderr, "mtmd: fewer than 3 cells on an axis; the reference visits such cells more than "
                                 "once and so does the exact engine\n");
    } else {
        if (!(c.skin >= 0)) return h->fail(MTMD_EINVAL, "skin must be >= 0");
        p.rlist = (float)(c.cutoff_radius + c.skin);
        p.rlist2 = p.rlist * p.rlist;
        for (int i = 0; i < 3; i++) {
            d.nc[i] = (int)std::floor(p.L[i] / (double)p.rlist);
            if (d.nc[i] < 3)
                return h->fail(MTMD_EBOX, "the fp64/mixed engines need at least 3 cells of cutoff+skin per axis "
                                          "(box %.4g, cell %.4g)", p.L[i], (double)p.rlist);
            if (p.L[i] < 2.0 * p.rlist) return h->fail(MTMD_EBOX, "box below twice the list radius");
            p.cs[i] = (float)(p.L[i] / d.nc[i]);
            p.cw[i] = p.L[i] / d.nc[i];
            p.invL[i] = 1.0 / p.L[i];
        }
        h->half_skin2 = (c.skin / 2.0) * (c.skin / 2.0);
        p.half_skin2 = (float)h->half_skin2;
        p.x0 = 0.01f;
        {
            // same spacing the tables use: (rc - 0.01f)/999 in float
            const float delta = (p.rc - 0.01f) / float(MT_NPTS - 1);
            p.inv_delta = 1.0f / delta;
        }
    }
    d.ncells = d.nc[0] * d.nc[1] * d.nc[2];

    // --- device memory ---
    std::vector<double4> pos(n), mom(n);
    std::vector<int> type(s->type, s->type + n);
    std::vector<uint8_t> mobile(n, 1);
    if (s->mobile) for (int k = 0; k < n; k++) mobile[k] = s->mobile[k] ? 1 : 0;
    for (int k = 0; k < n; k++) {
        pos[k] = make_double4(s->pos[3 * k], s->pos[3 * k + 1], s->pos[3 * k + 2], (double)(float)s->charge[k]);
        mom[k] = make_double4(0, 0, 0, s->mass[k]);
    }
    if (s->mom) {
        for (int k = 0; k < n; k++) { mom[k].x = s->mom[3 * k]; mom[k].y = s->mom[3 * k + 1]; mom[k].z = s->mom[3 * k + 2]; }
    } else {
        // CSimulationBox::setRandMom (SimulationBox.cpp:98-133) from a seeded generator
        uint64_t st = s->seed;
        for (int k = 0; k < n; k++) {
            double px = 2.0 * urand(st) - 1.0, py = 2.0 * urand(st) - 1.0, pz = 2.0 * urand(st) - 1.0;
            mom[k].x = (px == 0.0) ? 1.0 : px;
            mom[k].y = (py == 0.0) ? 1.0 : py;
            mom[k].z = (pz == 0.0) ? 1.0 : pz;
        }
        for (int k = 0; k < n; k++) if (!mobile[k]) { mom[k].x = mom[k].y = mom[k].z = 0.0; }
        double Tsum = 0.0;
        for (int k = 0; k < n; k++)
            Tsum += ((mom[k].x * mom[k].x) + (mom[k].y * mom[k].y) + (mom[k].z * mom[k].z)) / mom[k].w;
        const double Tnow = Tsum / (3 * (double)n);
        const double ps = std::sqrt((c.temperature / MT_CONV_T) / Tnow);
        for (int k = 0; k < n; k++) { mom[k].x *= ps; mom[k].y *= ps; mom[k].z *= ps; }
    }
    int rc2;
#define UP(dst, src) if ((rc2 = upload(h, dst, src))) return rc2
    UP(d.pos, pos); UP(d.mom, mom); UP(d.type, type); UP(d.mol, mol); UP(d.mobile, mobile); UP(d.xmol, xmol);
    {
        // U at the last table point (x = rc) summed over each pair's force fields
        std::vector<float> ecut((size_t)T * T, 0.0f);
        for (int pr = 0; pr < T * T; pr++)
            for (int ff = 0; ff < nbcnt[pr]; ff++)
                ecut[pr] += nbtab[((size_t)pr * MT_MAXFF + ff) * MT_NPTS + MT_NPTS - 1].z;
        UP(d.ecut, ecut);
    }
    UP(d.nbtab, nbtab); UP(d.nbcnt, nbcnt); UP(d.btab, btab); UP(d.atab, atab); UP(d.dtab, dtab);
    UP(d.bl.b_first, b_first); UP(d.bl.b_num, b_num); UP(d.bl.b_j, b_j); UP(d.bl.b_type, b_type);
    UP(d.bl.a_first, a_first); UP(d.bl.a_num, a_num); UP(d.bl.a_i, a_i); UP(d.bl.a_j, a_j);
    UP(d.bl.a_type, a_type); UP(d.bl.a_center, a_center);
    UP(d.bl.d_first, d_first); UP(d.bl.d_num, d_num); UP(d.bl.d_i, d_i); UP(d.bl.d_j, d_j); UP(d.bl.d_l, d_l);
    UP(d.bl.d_type, d_type); UP(d.bl.d_center, d_center);
    d.bl.has_bonds = !b_j.empty();
    d.bl.has_angles = !a_i.empty();
    d.bl.has_dihedrals = !d_i.empty();
    if (h->fast) {
        UP(d.sp.first, sp_first); UP(d.sp.num, sp_num); UP(d.sp.partner, sp_partner); UP(d.sp.scale, sp_scale);
        d.sp.n_entries = (int)sp_partner.size();
        // MIXED segment coefficients, summed over each pair's force fields
        std::vector<float4> coef((size_t)T * T * MT_NPTS, make_float4(0, 0, 0, 0));
        for (int pr = 0; pr < T * T; pr++)
            for (int ff = 0; ff < nbcnt[pr]; ff++) {
                const float4* t = &nbtab[((size_t)pr * MT_MAXFF + ff) * MT_NPTS];
                for (int i = 0; i < MT_NPTS - 1; i++) {
                    float af, bf, ae, be;
                    segment_coeffs(t, i, af, bf, ae, be);
                    float4& o = coef[(size_t)pr * MT_NPTS + i];
                    o.x += af; o.y += bf; o.z += ae; o.w += be;
                }
            }
        UP(d.nbcoef, coef);
        std::vector<float2> coef2(coef.size());
        for (size_t i = 0; i < coef.size(); i++) coef2[i] = make_float2(coef[i].x, coef[i].y);
        UP(d.nbcoef2, coef2);
        // the tile kernel's LDS table in symmetric rows when the two types'
        // coefficients agree bit for bit (row ti + tj holds the pair (min, max))
        d.tabsym = 0;
        if (T <= 2) {
            bool sym = true;
            if (T == 2)
                for (int i = 0; i < MT_NPTS && sym; i++)
                    sym = std::memcmp(&coef2[(size_t)1 * MT_NPTS + i], &coef2[(size_t)2 * MT_NPTS + i],
                                      sizeof(float2)) == 0;
            if (sym) {
                std::vector<float2> cs((size_t)(2 * T - 1) * MT_NPTS);
                for (int r = 0; r < 2 * T - 1; r++) {
                    const int a = r < T ? 0 : r - T + 1, b = r - a;
                    std::memcpy(&cs[(size_t)r * MT_NPTS], &coef2[((size_t)a * T + b) * MT_NPTS],
                                MT_NPTS * sizeof(float2));
                }
                UP(d.nbcoef2s, cs);
                d.tabsym = 1;
            }
        }
        std::vector<int> ident(n);
        std::iota(ident.begin(), ident.end(), 0);
        UP(d.perm, ident); UP(d.inv, ident);
        d.perm2 = h->mem.alloc<int>(n);
        d.frc2 = h->mem.alloc<double4>(n); d.epot2 = h->mem.alloc<float>(n);
        d.pos2 = h->mem.alloc<double4>(n); d.mom2 = h->mem.alloc<double4>(n);
        d.type2 = h->mem.alloc<int>(n); d.mol2 = h->mem.alloc<int>(n); d.mobile2 = h->mem.alloc<uint8_t>(n);
        d.posref = h->mem.alloc<double4>(n);
        d.posf = h->mem.alloc<float4>(n);
        d.posf2 = (c.engine == MTMD_ENGINE_MIXED) ? h->mem.alloc<float4>(n) : nullptr;
        d.home = h->mem.alloc<int>(n);
        d.sidx = h->mem.alloc<int>(n);
        d.qfs = h->mem.alloc<float>(n);
        d.wsorted = (c.engine == MTMD_ENGINE_MIXED) ? 1 : 0;
        d.tile16 = mt_fast_tile_mode(d, p.L, c.engine == MTMD_ENGINE_MIXED);
        d.qf = h->mem.alloc<float>(n);
        d.qf2 = h->mem.alloc<float>(n);
        if (p.coulomb_on && p.coulomb_allpairs) {
            // enough (atom block, slice) workgroups to fill the chip: ~4 per CU
            const int blocks = (n + 255) / 256;
            d.ap_nsplit = std::max(1, std::min(64, (256 * 4 + blocks - 1) / blocks));
            d.ap_part = h->mem.alloc<double4>((size_t)d.ap_nsplit * n);
            if (!d.ap_part) return h->fail(MTMD_ENOMEM, "device allocation failed");
        }
        std::vector<float> qf(n);
        for (int k = 0; k < n; k++) qf[k] = (float)s->charge[k];
        HIPCHK(h, hipMemcpy(d.qf, qf.data(), n * sizeof(float), hipMemcpyHostToDevice));
        if (!d.pos2 || !d.mom2 || !d.type2 || !d.mol2 || !d.mobile2 || !d.posref || !d.posf || !d.home || !d.qf || !d.sidx || !d.qfs ||
            !d.qf2 || !d.perm2 || !d.frc2 || !d.epot2 || (c.engine == MTMD_ENGINE_MIXED && !d.posf2))
            return h->fail(MTMD_ENOMEM, "device allocation failed");
    }
#undef UP
    d.frc = h->mem.alloc<double4>(n);
    d.epot = h->mem.alloc<float>(n);
    d.cell_key = h->mem.alloc<int>(n);
    d.cell_xyz = h->fast ? nullptr : h->mem.alloc<int>(3 * (size_t)n);
    d.cell_count = h->mem.alloc<int>(d.ncells);
    d.cell_first = h->mem.alloc<int>(d.ncells);
    d.cell_fill = h->mem.alloc<int>(d.ncells);
    d.cell_bsum = h->mem.alloc<int>((d.ncells + 1023) / 1024 + 1);
    d.cell_next = h->mem.alloc<int>(d.ncells);
    d.cell_list = h->mem.alloc<int>(n);
    d.ncount = h->mem.alloc<int>(n);
    d.nreq = h->mem.alloc<int>(1);
    d.red_blocks = std::max(2 * ((n + 255) / 256), 3 * 1024);
    d.red = h->mem.alloc<double>(d.red_blocks);
    d.flags = h->mem.alloc<unsigned int>(8);
    if (!d.frc || !d.epot || !d.cell_key || !d.cell_count || !d.cell_first || !d.cell_fill || !d.cell_list ||
        !d.cell_bsum || !d.cell_next ||
        !d.ncount || !d.nreq || !d.red || !d.flags || (!h->fast && !d.cell_xyz))
        return h->fail(MTMD_ENOMEM, "device allocation failed");
    HIPCHK(h, hipMemset(d.frc, 0, n * sizeof(double4)));
    HIPCHK(h, hipMemset(d.epot, 0, n * sizeof(float)));
    HIPCHK(h, hipMemset(d.flags, 0, 8 * sizeof(unsigned int)));
    HIPCHK(h, hipMemset(d.nreq, 0, sizeof(int)));       // the longest row a build could not hold
    HIPCHK(h, hipMemset(d.cell_next, 0, d.ncells * sizeof(int)));
    // neighbour capacity: the expected full count at the list radius plus margin
    {
        const double V = p.L[0] * p.L[1] * p.L[2];
        const double rl = h->fast ? (double)p.rlist : c.cutoff_radius;
        const double nexp = (double)n / V * (4.0 / 3.0) * M_PI * rl * rl * rl;
        d.ncap = std::max(16, (int)std::ceil(nexp * 1.35) + 16);
        if (!h->fast && (d.nc[0] < 3 || d.nc[1] < 3 || d.nc[2] < 3)) d.ncap *= 2;
        d.ncap = std::min(d.ncap, n * 27 + 16);
        if (!h->list_slack) d.ncap = 16;
        d.ncap = (d.ncap + 3) & ~3;          // whole 16-byte chunks (FP64/MIXED list layout)
        d.nlist = h->mem.alloc<uint32_t>((size_t)d.ncap * n);
        if (!d.nlist) return h->fail(MTMD_ENOMEM, "neighbour list allocation failed");
    }

    // --- thermostat (SimulationBox.cpp:53-79, NHChain.cpp:32-52,115-143) ---
    h->M = (c.ensemble == MTMD_NVT || c.ensemble == MTMD_NPT) ? c.temperature_nh_len : 0;
    h->nresp = c.temperature_respa;
    h->Tred = c.temperature / MT_CONV_T;
    h->tau = c.temperature_relax / MT_CONV_t;
    h->g = (double)(3 * n);
    {
        const double w0 = 1.0 / (1.0 - std::pow(2.0, 1.0 / 3.0));   // NHChain.cpp:45-52
        h->w[0] = w0; h->w[1] = 1.0 - 2.0 * w0; h->w[2] = w0;
    }
    const int M = std::max(h->M, 1);
    std::vector<double> nh(3 * M + 2, 0.0);
    if (h->M > 0) {
        nh[2 * M] = (double)(3 * n) * h->Tred * h->tau * h->tau;
        for (int j = 1; j < M; j++) nh[2 * M + j] = h->Tred * h->tau * h->tau;
        // The chain starts at rest (eta = p_eta = 0). The reference draws
        // p_eta[j] uniformly in +-Q_j/tau (NHChain.cpp:126-143); since Q_0
        // grows with N, p_eta[0]^2/Q_0 does too, and a negative draw drives
        // the chain unstable for large systems. mtmd_set_thermostat sets any
        // state, the reference's draw included.
    }
    nh[3 * M] = 1.0;
    h->nh_scale_slot = 3 * M;
    // --- barostat (SimulationBox.cpp:53-79, VelVerlet.cpp:31-43, NHChain.cpp:115-124
    // with prepareArrays(1, 1)): p_eps and the pressure chain start at rest
    // (the reference draws them at random); mtmd_set_barostat sets any state
    h->npt = (c.ensemble == MTMD_NPT);
    for (int i = 0; i < 3; i++) h->Lmax[i] = h->Lrb[i] = s->box[i];
    if (h->npt) {
        h->Mp = c.pressure_nh_len;
        h->nrespP = c.pressure_respa;
        h->tauP = c.pressure_relax / MT_CONV_t;
        h->W = h->Tred * h->tauP * h->tauP;
        h->Pext = c.pressure / MT_CONV_PRESS;
        h->V0 = h->Lmax[0] * h->Lmax[1] * h->Lmax[2];
        h->etaP.assign(h->Mp, 0.0);
        h->petaP.assign(h->Mp, 0.0);
        h->QP.assign(h->Mp, h->Tred * h->tauP * h->tauP);
        h->QP[0] = (double)(1 * 1) * h->Tred * h->tauP * h->tauP;
    }
    h->type_host.assign(s->type, s->type + n);
    h->type_names.assign(T, std::string());
    for (int t = 0; t < T; t++)
        h->type_names[t] = (s->type_names && s->type_names[t]) ? std::string(s->type_names[t]) : std::to_string(t);
    if ((rc2 = upload(h, d.nh, nh))) return rc2;
    h->h_status = nullptr;
    HIPCHK(h, hipHostMalloc((void**)&h->h_status, 16 * sizeof(int), hipHostMallocDefault));
    return MTMD_OK;
}

extern "C" int mtmd_create(const mtmd_system* sys, const mtmd_config* cfg, mtmd_handle** out)
{
    g_create_error.clear();
    if (!out) { g_create_error = "out is NULL"; return MTMD_EINVAL; }
    *out = nullptr;
    if (!sys || !cfg) { g_create_error = "system and config are required"; return MTMD_EINVAL; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_error = "no HIP device";
        return MTMD_ENODEV;
    }
    mtmd_handle* h = new mtmd_handle();
    h->cfg = *cfg;
    if (hipGetDevice(&h->device) != hipSuccess) h->device = 0;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        delete h;
        return MTMD_EHIP;
    }
    int rc = setup(h, sys);
    if (rc != MTMD_OK) {
        g_create_error = h->err;
        mtmd_destroy(h);
        return rc;
    }
    *out = h;
    return MTMD_OK;
}

extern "C" void mtmd_destroy(mtmd_handle* h)
{
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->drop_graphs();
    h->mem.release();
    if (h->h_status) (void)hipHostFree(h->h_status);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

static void swap_slots(MtDev& d);

// ---------------------------------------------------------------------------
// state transfer (positions/momenta in original atom order)
// ---------------------------------------------------------------------------
static int download_perm(mtmd_handle* h)
{
    h->perm_host.resize(h->n);
    if (!h->fast) { std::iota(h->perm_host.begin(), h->perm_host.end(), 0); return MTMD_OK; }
    HIPCHK(h, hipMemcpyAsync(h->perm_host.data(), h->d.perm, h->n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return MTMD_OK;
}

static int get_d4(mtmd_handle* h, const double4* src, double* out, bool xyz_only = true)
{
    std::vector<double4> tmp(h->n);
    HIPCHK(h, hipMemcpyAsync(tmp.data(), src, h->n * sizeof(double4), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    int rc = download_perm(h);
    if (rc) return rc;
    (void)xyz_only;
    for (int s = 0; s < h->n; s++) {
        const int o = h->perm_host[s];
        out[3 * o] = tmp[s].x; out[3 * o + 1] = tmp[s].y; out[3 * o + 2] = tmp[s].z;
    }
    return MTMD_OK;
}

// bring the FP64/MIXED engines back to the original atom order
static int unsort(mtmd_handle* h)
{
    if (!h->fast) return MTMD_OK;
    mt_fast_unsort(h->d, h->stream);
    swap_slots(h->d);
    mt_fast_identity_perm(h->d, h->stream);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->need_sort = true;
    h->ke_valid = false;
    return MTMD_OK;
}

extern "C" int mtmd_set_state(mtmd_handle* h, const double* pos, const double* mom)
{
    if (!h) return MTMD_EINVAL;
    int rc = unsort(h);
    if (rc) return rc;
    const int n = h->n;
    std::vector<double4> tmp(n);
    if (pos) {
        HIPCHK(h, hipMemcpy(tmp.data(), h->d.pos, n * sizeof(double4), hipMemcpyDeviceToHost));
        for (int k = 0; k < n; k++) { tmp[k].x = pos[3 * k]; tmp[k].y = pos[3 * k + 1]; tmp[k].z = pos[3 * k + 2]; }
        HIPCHK(h, hipMemcpy(h->d.pos, tmp.data(), n * sizeof(double4), hipMemcpyHostToDevice));
        h->have_forces = false;
    }
    if (mom) {
        HIPCHK(h, hipMemcpy(tmp.data(), h->d.mom, n * sizeof(double4), hipMemcpyDeviceToHost));
        for (int k = 0; k < n; k++) { tmp[k].x = mom[3 * k]; tmp[k].y = mom[3 * k + 1]; tmp[k].z = mom[3 * k + 2]; }
        HIPCHK(h, hipMemcpy(h->d.mom, tmp.data(), n * sizeof(double4), hipMemcpyHostToDevice));
        h->ke_valid = false;
        h->ke_slot = false;
    }
    return MTMD_OK;
}

extern "C" int mtmd_get_state(mtmd_handle* h, double* pos, double* mom)
{
    if (!h) return MTMD_EINVAL;
    int rc;
    if (pos && (rc = get_d4(h, h->d.pos, pos))) return rc;
    if (mom && (rc = get_d4(h, h->d.mom, mom))) return rc;
    return MTMD_OK;
}

extern "C" int mtmd_set_thermostat(mtmd_handle* h, const double* eta, const double* p_eta)
{
    if (!h) return MTMD_EINVAL;
    if (h->M == 0) return MTMD_OK;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (eta) HIPCHK(h, hipMemcpy(h->d.nh, eta, h->M * sizeof(double), hipMemcpyHostToDevice));
    if (p_eta) HIPCHK(h, hipMemcpy(h->d.nh + h->M, p_eta, h->M * sizeof(double), hipMemcpyHostToDevice));
    return MTMD_OK;
}

// ---------------------------------------------------------------------------
// NPT (MDLoop.cpp:124-138 with VelVerlet.cpp:57-71,124-174,192-211)
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
// the box every kernel sees: widths as the functors take them (float) and as
// the integrator and the wrap take them (double)
static void set_box(mtmd_handle* h, const double L[3])
{
    MtParams& p = h->p;
    for (int i = 0; i < 3; i++) {
        p.L[i] = L[i];
        p.Lf[i] = (float)L[i];
        p.lowf[i] = -p.Lf[i] / 2.0f;
        const double Lf = (double)p.Lf[i];
        p.wbox[i] = (Lf / 2.0) - (-Lf / 2.0);
        p.wnl[i] = (double)(p.lowf[i] + p.Lf[i]) - (double)p.lowf[i];
        if (h->fast) {
            p.cs[i] = (float)(L[i] / h->d.nc[i]);
            p.cw[i] = L[i] / h->d.nc[i];
            p.invL[i] = 1.0 / L[i];
        }
    }
}

// CMt::exp / CMt::sinhXoverX (Integrators/Math.cpp:27-54)
static double cmt_exp(double x) { return (x > 1.0) ? 2.71828 * x : std::exp(x); }
static double cmt_sinhxoverx(double x)
{
    static const double a[6] = {1.0, 1.0 / 6.0, 1.0 / 120.0, 1.0 / 5040.0, 1.0 / 362880.0, 1.0 / 39916800.0};
    double sum = 0.0, xn = 1.0;
    for (int n = 0; n < 6; n++) { sum += a[n] * xn; xn *= (x * x); }
    return sum;
}

static double npt_volume(const mtmd_handle* h) { return h->V0 * std::exp(3.0 * h->eps); }   // getV

static void npt_box(const mtmd_handle* h, double L[3])                                      // getSimBox
{
    const double Vmax = h->Lmax[0] * h->Lmax[1] * h->Lmax[2];
    double etaCube = npt_volume(h) / Vmax;
    etaCube = std::pow(etaCube, 1.0 / 3.0);
    for (int i = 0; i < 3; i++) L[i] = etaCube * h->Lmax[i];
}

// G_eps (VelVerlet.cpp:160-174) from the sums of p.p/m and Fpi.r
static double npt_G(const mtmd_handle* h, double sum_p, double sum_f)
{
    return (1.0 + 1.0 / (double)h->n) * sum_p + sum_f - 3.0 * h->Pext * npt_volume(h);
}

// NH_P_.propagator with CFctP (NHChain.cpp:54-113, MDLoop.cpp:63-83): a chain
// on the barostat momentum, G(0) = p_eps^2/W - T, scaleMomentum scales p_eps
static void npt_nhp(mtmd_handle* h)
{
    const int M = h->Mp;
    double* peta = h->petaP.data();
    double* eta = h->etaP.data();
    const double* Q = h->QP.data();
    const double beta = 1.0 / h->Tred;
    auto G = [&](int j) {
        if (j == 0) return h->peps * h->peps / h->W - (1.0 / beta);
        return (peta[j - 1] * peta[j - 1] / Q[j - 1]) - (1.0 / beta);
    };
    if (M > 0 && peta[0] > 15.0) peta[0] = 10.0;
    if (M <= 0) return;
    for (int mu = 0; mu < 3; mu++) {
        const double Dt = h->w[mu] * h->p.dt / (double)h->nrespP;
        for (int i = 0; i < h->nrespP; i++) {
            peta[M - 1] += (Dt / 4.0) * G(M - 1);
            for (int j = 0; j < M - 1; j++) {
                peta[j] *= cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] += (Dt / 4.0) * G(j);
                peta[j] *= cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
            }
            for (int j = 0; j < M; j++) eta[j] += (Dt / 2.0) * (peta[j] / Q[j]);
            h->peps *= cmt_exp(-((Dt / 2.0) * (peta[0] / Q[0])));
            for (int j = M - 2; j >= 0; j--) {
                peta[j] *= cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
                peta[j] += (Dt / 4.0) * G(j);
                peta[j] *= cmt_exp(-(Dt / 8.0) * (peta[j + 1] / Q[j + 1]));
            }
            peta[M - 1] += (Dt / 4.0) * G(M - 1);
        }
    }
}
#pragma clang fp contract(on)

// FP64/MIXED under NPT: coarsen the cell grid when a cell of the current box
// would be narrower than the list radius (the device arrays were sized for the
// initial, finer grid)
static int npt_regrid(mtmd_handle* h)
{
    MtDev& d = h->d;
    bool changed = false;
    for (int i = 0; i < 3; i++) {
        const int nc = (int)std::floor(h->p.L[i] / (double)h->p.rlist);
        if (nc < 3) return h->fail(MTMD_EBOX, "the box shrank below 3 cells of cutoff+skin per axis");
        if (nc < d.nc[i]) { d.nc[i] = nc; changed = true; }
    }
    if (changed) {
        d.ncells = d.nc[0] * d.nc[1] * d.nc[2];
        d.tile16 = mt_fast_tile_mode(d, h->p.L, h->engine == MTMD_ENGINE_MIXED);   // lists of the new grid
        set_box(h, h->p.L);                          // cell sizes of the new grid
        h->drop_graphs();
    }
    return MTMD_OK;
}

extern "C" int mtmd_set_barostat(mtmd_handle* h, double eps, double p_eps, const double* eta, const double* p_eta)
{
    if (!h) return MTMD_EINVAL;
    if (!h->npt) return h->fail(MTMD_EINVAL, "the barostat exists in the npt ensemble only");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->eps = eps;
    h->peps = p_eps;
    for (int j = 0; j < h->Mp; j++) {
        if (eta) h->etaP[j] = eta[j];
        if (p_eta) h->petaP[j] = p_eta[j];
    }
    // the box follows eps (getSimBox); forces of the old box are stale
    double box[3];
    npt_box(h, box);
    set_box(h, box);
    if (h->fast) {
        int rc = npt_regrid(h);
        if (rc) return rc;
    }
    h->have_forces = false;
    h->need_sort = h->fast;
    return MTMD_OK;
}

extern "C" int mtmd_get_barostat(mtmd_handle* h, double* eps, double* p_eps, double* eta, double* p_eta,
                                 double* box)
{
    if (!h) return MTMD_EINVAL;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (eps) *eps = h->eps;
    if (p_eps) *p_eps = h->peps;
    for (int j = 0; j < h->Mp; j++) {
        if (eta) eta[j] = h->etaP[j];
        if (p_eta) p_eta[j] = h->petaP[j];
    }
    if (box) for (int i = 0; i < 3; i++) box[i] = h->p.L[i];
    return MTMD_OK;
}

extern "C" int mtmd_get_thermostat(mtmd_handle* h, double* eta, double* p_eta)
{
    if (!h) return MTMD_EINVAL;
    if (h->M == 0) return MTMD_OK;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (eta) HIPCHK(h, hipMemcpy(eta, h->d.nh, h->M * sizeof(double), hipMemcpyDeviceToHost));
    if (p_eta) HIPCHK(h, hipMemcpy(p_eta, h->d.nh + h->M, h->M * sizeof(double), hipMemcpyDeviceToHost));
    return MTMD_OK;
}

// ---------------------------------------------------------------------------
// force evaluation
// ---------------------------------------------------------------------------
static int grow_neighbors(mtmd_handle* h, int need)
{
    MtDev& d = h->d;
    need = std::max(need, d.ncap);                    // (grows, never shrinks)
    const int cap = (h->list_slack ? need + need / 8 + 16 + 3 : need + 3) & ~3;
    void* old = d.nlist;
    uint32_t* nl = nullptr;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (hipMalloc((void**)&nl, (size_t)cap * h->n * sizeof(uint32_t)) != hipSuccess)
        return h->fail(MTMD_ENOMEM, "neighbour list of %d rows does not fit", cap);
    // the captured step graphs hold the old buffer as a kernel argument: they
    // go before the buffer does
    h->drop_graphs();
    for (auto& p : h->mem.ptrs) if (p == old) { (void)hipFree(p); p = nl; }
    d.nlist = nl;
    d.ncap = cap;
    return MTMD_OK;
}

// EXACT: CMDFFMatrices::genNeighList + the force functor, every call
static int exact_cells_checked(mtmd_handle* h)
{
    MtDev& d = h->d;
    HIPCHK(h, hipMemsetAsync(d.flags + 3, 0, sizeof(unsigned int), h->stream));
    mt_exact_cells(d, h->p, h->stream);
    return MTMD_OK;
}

static int exact_forces(mtmd_handle* h)
{
    MtDev& d = h->d;
    int rc0 = exact_cells_checked(h);
    if (rc0) return rc0;
    for (int attempt = 0; attempt < 4; attempt++) {
        HIPCHK(h, hipMemsetAsync(d.nreq, 0, sizeof(int), h->stream));
        mt_exact_neighbors(d, h->p, h->stream);
        HIPCHK(h, hipMemcpyAsync(h->h_status, d.nreq, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->h_status + 1, d.flags + 3, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (h->h_status[1]) return h->fail(MTMD_ENAN, "non-finite atom position");
        const int need = h->h_status[0];
        if (need <= d.ncap) break;
        int rc = grow_neighbors(h, need);
        if (rc) return rc;
        if (attempt == 3) return h->fail(MTMD_ENEIGH, "neighbour list keeps overflowing");
    }
    mt_exact_forces(d, h->p, h->stream);
    HIPCHK(h, hipGetLastError());
    h->rebuilds++;
    return MTMD_OK;
}

static void swap_slots(MtDev& d)
{
    std::swap(d.pos, d.pos2); std::swap(d.mom, d.mom2); std::swap(d.frc, d.frc2); std::swap(d.epot, d.epot2);
    std::swap(d.type, d.type2); std::swap(d.mol, d.mol2); std::swap(d.mobile, d.mobile2); std::swap(d.qf, d.qf2);
    std::swap(d.perm, d.perm2);
}

// FP64/MIXED: full rebuild (optionally with spatial re-sort) or conditional
static int fast_rebuild(mtmd_handle* h, bool force_rebuild, bool sort)
{
    MtDev& d = h->d;
    const int cond = force_rebuild ? 0 : 1;
    if (sort) {
        // spatial re-sort: order atoms by cell (ascending original slot inside a cell)
        mt_fast_cells(d, h->p, 0, 0, h->stream);
        mt_fast_sort_apply(d, h->stream);
        swap_slots(d);
        h->steps_since_sort = 0;
        h->need_sort = false;
    }
    mt_fast_cells(d, h->p, cond, 1, h->stream);
    mt_fast_neighbors(d, h->p, cond, h->stream);
    HIPCHK(h, hipGetLastError());
    return MTMD_OK;
}

static int fast_check_capacity(mtmd_handle* h, bool* overflowed)
{
    MtDev& d = h->d;
    HIPCHK(h, hipMemcpyAsync(h->h_status, d.flags, 8 * sizeof(unsigned int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *overflowed = h->h_status[2] != 0;
    h->rebuilds = (int64_t)(unsigned)h->h_status[4];
    if (h->h_status[5]) {
        if (d.force_dense)
            return h->fail(MTMD_EBOX, "a cell holds more atoms than the neighbour-list build stages at once (%d)",
                           d.tile16 ? MT_NB_STAGE_ATOMS_TILE : MT_NB_STAGE_ATOMS);
        // a cell too large for the block build: the lists come from the
        // dense-cell build from now on (32-bit words, tile mode off); the
        // caller builds them again, as after an overflow
        d.force_dense = 1;
        d.tile16 = 0;
        h->drop_graphs();
        HIPCHK(h, hipMemsetAsync(d.flags + 5, 0, sizeof(unsigned int), h->stream));
        *overflowed = true;
    }
    if (h->h_status[3]) {
        HIPCHK(h, hipMemsetAsync(d.flags + 3, 0, sizeof(unsigned int), h->stream));
        h->have_forces = false;
        return h->fail(MTMD_ENAN, "non-finite atom position");
    }
    return MTMD_OK;
}

// kicked (NVT steps): set when the forces' bonded pass also did the back
// half's cap and kick with the kinetic-energy partials (see mt_fast_forces)
static int fast_forces(mtmd_handle* h, bool want_energy, bool* kicked = nullptr)
{
    const int k = mt_fast_forces(h->d, h->p, h->engine == MTMD_ENGINE_MIXED, want_energy ? 1 : 0,
                                 h->cfg.block_size, h->stream, kicked ? 1 : 0);
    if (kicked) *kicked = k != 0;
    HIPCHK(h, hipGetLastError());
    return MTMD_OK;
}

// a full, unconditional list build (with re-sort) and forces of the current positions
static int fast_full_forces(mtmd_handle* h)
{
    MtDev& d = h->d;
    int rc = fast_rebuild(h, true, true);
    if (rc) return rc;
    for (int attempt = 0;; attempt++) {
        bool ov = false;
        if ((rc = fast_check_capacity(h, &ov))) return rc;
        if (!ov) break;
        if (attempt == 2) return h->fail(MTMD_ENEIGH, "neighbour list overflow after regrowth");
        // the build told us the largest row (nreq), or moved to the
        // dense-cell build: regrow and build again
        HIPCHK(h, hipMemcpy(h->h_status, d.nreq, sizeof(int), hipMemcpyDeviceToHost));
        if ((rc = grow_neighbors(h, h->h_status[0]))) return rc;
        HIPCHK(h, hipMemsetAsync(d.flags + 2, 0, sizeof(unsigned int), h->stream));
        if ((rc = fast_rebuild(h, true, false))) return rc;
    }
    return fast_forces(h, true);
}

static int compute_forces_now(mtmd_handle* h)
{
    int rc = h->fast ? fast_full_forces(h) : exact_forces(h);
    if (rc) return rc;
    h->have_forces = true;
    return MTMD_OK;
}

static int get_forces_host(mtmd_handle* h, double* F, double* Fpi, float* U)
{
    const int n = h->n;
    int rc;
    std::vector<double> tmp(3 * (size_t)n);
    if (F || Fpi) {
        if ((rc = get_d4(h, h->d.frc, tmp.data()))) return rc;
        if (F) std::memcpy(F, tmp.data(), tmp.size() * sizeof(double));
        if (Fpi) std::memcpy(Fpi, tmp.data(), tmp.size() * sizeof(double));
    }
    if (U) {
        std::vector<float> u(n);
        HIPCHK(h, hipMemcpyAsync(u.data(), h->d.epot, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if ((rc = download_perm(h))) return rc;
        for (int s = 0; s < n; s++) U[h->perm_host[s]] = u[s];
    }
    return MTMD_OK;
}

extern "C" int mtmd_compute_forces(mtmd_handle* h, double* F, double* Fpi, float* U)
{
    if (!h) return MTMD_EINVAL;
    int rc = compute_forces_now(h);
    if (rc) return rc;
    return get_forces_host(h, F, Fpi, U);
}

extern "C" int mtmd_get_forces(mtmd_handle* h, double* F, float* U)
{
    if (!h) return MTMD_EINVAL;
    if (!h->have_forces) {
        int rc = compute_forces_now(h);
        if (rc) return rc;
    }
    return get_forces_host(h, F, nullptr, U);
}

extern "C" int mtmd_build_cells(mtmd_handle* h, int32_t* ncxyz, int32_t* ixyz, int32_t* cfirst, int32_t* ccount,
                                int32_t* clist)
{
    if (!h) return MTMD_EINVAL;
    if (h->fast) return h->fail(MTMD_EINVAL, "reference cell lists are produced by the exact engine");
    MtDev& d = h->d;
    int rc0 = exact_cells_checked(h);
    if (rc0) return rc0;
    HIPCHK(h, hipGetLastError());
    if (ncxyz) { ncxyz[0] = d.nc[0]; ncxyz[1] = d.nc[1]; ncxyz[2] = d.nc[2]; }
    if (ixyz) HIPCHK(h, hipMemcpyAsync(ixyz, d.cell_xyz, 3 * (size_t)h->n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    if (cfirst) HIPCHK(h, hipMemcpyAsync(cfirst, d.cell_first, d.ncells * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    if (ccount) HIPCHK(h, hipMemcpyAsync(ccount, d.cell_count, d.ncells * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    if (clist) HIPCHK(h, hipMemcpyAsync(clist, d.cell_list, h->n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->h_status + 1, d.flags + 3, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->h_status[1]) return h->fail(MTMD_ENAN, "non-finite atom position");
    return MTMD_OK;
}

extern "C" int mtmd_build_neighbors(mtmd_handle* h, int32_t max_per_atom, int32_t* count, int32_t* list,
                                    int32_t* required)
{
    if (!h) return MTMD_EINVAL;
    if (h->fast) return h->fail(MTMD_EINVAL, "reference neighbour lists are produced by the exact engine");
    MtDev& d = h->d;
    int rc0 = exact_cells_checked(h);
    if (rc0) return rc0;
    for (int attempt = 0; attempt < 4; attempt++) {
        HIPCHK(h, hipMemsetAsync(d.nreq, 0, sizeof(int), h->stream));
        mt_exact_neighbors(d, h->p, h->stream);
        HIPCHK(h, hipMemcpyAsync(h->h_status, d.nreq, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->h_status + 1, d.flags + 3, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (h->h_status[1]) return h->fail(MTMD_ENAN, "non-finite atom position");
        if (h->h_status[0] <= d.ncap) break;
        int rc = grow_neighbors(h, h->h_status[0]);
        if (rc) return rc;
    }
    std::vector<int> cnt(h->n);
    HIPCHK(h, hipMemcpy(cnt.data(), d.ncount, h->n * sizeof(int), hipMemcpyDeviceToHost));
    int mx = 0;
    for (int v : cnt) mx = std::max(mx, v);
    if (required) *required = mx;
    if (count) std::memcpy(count, cnt.data(), h->n * sizeof(int));
    if (list && max_per_atom > 0) {
        std::vector<uint32_t> nl((size_t)d.ncap * h->n);
        HIPCHK(h, hipMemcpy(nl.data(), d.nlist, nl.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (int k = 0; k < h->n; k++) {
            const int c = std::min(cnt[k], max_per_atom);
            for (int j = 0; j < c; j++) list[(size_t)k * max_per_atom + j] = (int32_t)nl[(size_t)j * h->n + k];
        }
    }
    h->have_forces = false;
    return (mx > max_per_atom && list) ? h->fail(MTMD_ENEIGH, "row of %d neighbours exceeds %d", mx, max_per_atom)
                                       : MTMD_OK;
}


// NPT steps, all engines: NH_P and the barostat scalars on the host, the NH_T
// chain and the particles on the device. Two host waits per step read the
// block partials of G_eps (and, FP64/MIXED, the list-build counter).
// FP64/MIXED lists: a step rebuilds when an atom moved more than skin/2 less
// the box change since the last build (which bounds the error of the
// minimum-image displacement of atoms wrapped in between), when the box
// changed by more than skin/4 since that build, or when a cell would drop
// below the list radius (the grid is then coarsened).
static int npt_steps(mtmd_handle* h, int nsteps, bool energy_last)
{
    MtDev& d = h->d;
    MtParams& p = h->p;
    const int nb = mt_npt_blocks(h->n);
    const int mixed = h->engine == MTMD_ENGINE_MIXED;
    const double skin = h->cfg.skin;
    std::vector<double> part(2 * (size_t)nb);
    int rc;
    auto sums = [&](double& sp, double& sf) -> int {
        HIPCHK(h, hipMemcpyAsync(part.data(), d.red, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        if (h->fast)
            HIPCHK(h, hipMemcpyAsync(h->h_status, d.flags, 8 * sizeof(unsigned int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        sp = 0.0;
        sf = 0.0;
        for (int b = 0; b < nb; b++) { sp += part[b]; sf += part[nb + b]; }
        return MTMD_OK;
    };
    if (h->fast) {
        // the device's list-build counter at the start of the call
        HIPCHK(h, hipMemcpyAsync(h->h_status, d.flags, 8 * sizeof(unsigned int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        h->rebuild_count = h->h_status[4];
        mt_fast_clear_flags(d, h->stream);
    }
    for (int t = 0; t < nsteps; t++) {
        const bool last = (t == nsteps - 1);
        npt_nhp(h);
        if (!h->ke_valid) mt_reduce_ke(d, h->stream);
        mt_nh_chain(d, h->M, h->nresp, h->Tred, p.dt, h->g, h->w, nb, h->stream, h->fast ? 1 : 0);
        mt_npt_scale_sums(d, h->nh_scale_slot, h->stream);
        double sp, sf;
        if ((rc = sums(sp, sf))) return rc;
        h->peps += ((p.dt / 2.0) * npt_G(h, sp, sf));
        // prop_p (VelVerlet.cpp:124-136) and prop_r (:138-150) coefficients
        const double u_eps = h->peps / h->W;
        const double parm_p = (1.0 + 1.0 / (double)h->n) * u_eps * p.dt / 4.0;
        const double cp = cmt_exp(-parm_p);
        const double cp2 = cp * cp, cpf = (p.dt / 2.0) * cp * cmt_sinhxoverx(parm_p);
        const double parm_r = u_eps * p.dt / 2.0;
        const double cr = cmt_exp(parm_r);
        const double cr2 = cr * cr, crr = p.dt * (cmt_sinhxoverx(parm_r) * cr);
        h->eps += (p.dt * h->peps / h->W);
        double box[3];
        npt_box(h, box);
        set_box(h, box);
        mt_npt_prop_pr(d, cp2, cpf, cr2, crr, h->stream);
        if (!h->fast) {
            if ((rc = exact_forces(h))) return rc;
        } else {
            double dL = 0.0;
            bool regrid = false;
            for (int i = 0; i < 3; i++) {
                dL = std::max(dL, std::fabs(box[i] - h->Lrb[i]));
                if (box[i] / d.nc[i] < (double)p.rlist) regrid = true;
            }
            const bool sort = (++h->steps_since_sort >= h->sort_every);
            const bool force = regrid || sort || dL > skin / 4.0;
            if (regrid && (rc = npt_regrid(h))) return rc;
            const double thr = std::max(0.0, skin / 2.0 - dL);
            p.half_skin2 = (float)(thr * thr);
            mt_fast_locate(d, p, mixed, h->stream);
            if ((rc = fast_rebuild(h, force, sort))) return rc;
            if (force) for (int i = 0; i < 3; i++) h->Lrb[i] = box[i];
            if ((rc = fast_forces(h, last && energy_last))) return rc;
        }
        mt_npt_prop_p_sums(d, cp2, cpf, h->stream);
        if ((rc = sums(sp, sf))) return rc;
        h->peps += ((p.dt / 2.0) * npt_G(h, sp, sf));
        if (h->fast) {
            // a list build on the device (the displacement test) this step
            if (h->h_status[4] != h->rebuild_count) {
                h->rebuild_count = h->h_status[4];
                for (int i = 0; i < 3; i++) h->Lrb[i] = box[i];
            }
            if (h->h_status[2]) break;                // list overflow: reported below
        }
        mt_nh_chain(d, h->M, h->nresp, h->Tred, p.dt, h->g, h->w, nb, h->stream, h->fast ? 1 : 0);
        npt_nhp(h);
        mt_scale_wrap(d, p, h->nh_scale_slot, 1, h->stream);
        h->ke_valid = true;
        h->steps++;
    }
    HIPCHK(h, hipGetLastError());
    return MTMD_OK;
}

// ---------------------------------------------------------------------------
// MD loop (MDLoop.cpp:124-138): NH_T, velocity Verlet, NH_T, wrap
// ---------------------------------------------------------------------------
// md_steps_once: a list build of the call overflowed; the lists have grown
static constexpr int MT_LIST_RERUN = 1000;

static int md_steps_once(mtmd_handle* h, int nsteps, bool energy_last)
{
    MtDev& d = h->d;
    const bool nvt = (h->M > 0);
    const int nblocks = (h->n + 255) / 256;
    int rc;
    if (!h->have_forces) {                              // calcInitialForces (MDLoop.cpp:189-192)
        if ((rc = compute_forces_now(h))) return rc;
    }
    if (h->fast && h->need_sort) {
        // re-sort now so the step loop runs on spatially ordered atoms
        if ((rc = fast_full_forces(h))) return rc;
        for (int i = 0; i < 3; i++) h->Lrb[i] = h->p.L[i];
    }
    const int mixed = h->engine == MTMD_ENGINE_MIXED;
    if (h->npt) {
        h->ke_slot = false;                             // (the barostat scales the momenta too)
        if ((rc = npt_steps(h, nsteps, energy_last))) return rc;
    } else if (!h->fast) {
        for (int t = 0; t < nsteps; t++) {
            if (nvt) {
                if (!h->ke_valid) mt_reduce_ke(d, h->stream);
                mt_nh_chain(d, h->M, h->nresp, h->Tred, h->p.dt, h->g, h->w, nblocks, h->stream, h->fast ? 1 : 0);
            }
            mt_kick_drift(d, h->p, nvt ? h->nh_scale_slot : -1, 0, h->stream);
            if ((rc = exact_forces(h))) return rc;
            mt_cap_kick(d, h->p, nvt ? 1 : 0, h->stream);
            if (nvt) mt_nh_chain(d, h->M, h->nresp, h->Tred, h->p.dt, h->g, h->w, nblocks, h->stream, h->fast ? 1 : 0);
            mt_scale_wrap(d, h->p, nvt ? h->nh_scale_slot : -1, nvt ? 1 : 0, h->stream);
            h->ke_valid = nvt;
            h->steps++;
        }
    } else {
        // FP64/MIXED. NVE: inside a call the back half of step t and the front
        // half of step t + 1 run as one kernel. Plain steps (no re-sort, not
        // the last of the call) replay a captured graph of their kernels.
        auto nvt_front = [&]() {
            // the kinetic energy the last call's final chain formed, or summed
            // anew when the momenta were set since
            if (h->ke_slot) {
                mt_nh_chain(d, h->M, h->nresp, h->Tred, h->p.dt, h->g, h->w, nblocks, h->stream, 1, 1);
            } else {
                if (!h->ke_valid) mt_reduce_ke(d, h->stream);
                mt_nh_chain(d, h->M, h->nresp, h->Tred, h->p.dt, h->g, h->w, nblocks, h->stream, 1, 0);
            }
            mt_fast_kick_drift(d, h->p, h->nh_scale_slot, mixed, h->stream);
        };
        auto nvt_back = [&](bool kicked) {
            if (!kicked) mt_cap_kick(d, h->p, 1, h->stream);
            mt_nh_chain(d, h->M, h->nresp, h->Tred, h->p.dt, h->g, h->w, nblocks, h->stream, h->fast ? 1 : 0);
            mt_scale_wrap(d, h->p, h->nh_scale_slot, 1, h->stream);
        };
        // NVT between two steps of the call: the back half of step t and the
        // front half of t + 1, both chain propagations in one launch and the
        // scale, wrap, kick and drift in one pass (a call's first front half
        // and last back half run as above)
        auto nvt_mid = [&](bool kicked) {
            if (!kicked) mt_cap_kick(d, h->p, 1, h->stream);
            mt_nh_chain2(d, h->M, h->nresp, h->Tred, h->p.dt, h->g, h->w, nblocks, h->stream);
            mt_fast_scale_kick_drift(d, h->p, h->nh_scale_slot, h->nh_scale_slot + 1, mixed, h->stream);
        };
        // NVE in the MIXED tile mode with every force inside the pair pass:
        // the pair pass also integrates (back half of step t, front half of
        // t + 1) into pos2/posf2, which then become pos/posf
        const bool fuse = !nvt && h->use_fuse && mt_fast_can_fuse(d, h->p, mixed);
        auto fused_forces = [&]() {
            mt_fast_forces_fused(d, h->p, h->stream);
            std::swap(d.pos, d.pos2);
            std::swap(d.posf, d.posf2);
        };
        // the kernels of one plain step after its front half, through the
        // front half of the next
        auto plain_step = [&]() -> int {
            int r = fast_rebuild(h, false, false);
            if (r) return r;
            if (fuse) {
                fused_forces();
                return MTMD_OK;
            }
            bool kicked = false;
            if ((r = fast_forces(h, false, nvt ? &kicked : nullptr))) return r;
            if (nvt) nvt_mid(kicked);
            else mt_fast_nve_fused(d, h->p, mixed, h->stream);
            return MTMD_OK;
        };
        auto replay = [&](int nstep) -> int {
            constexpr int G = mtmd_handle::kGraphsPerEnsemble;
            const int base = nvt ? G : 0;
            const mtmd_handle::StepState from = h->step_state();
            mtmd_handle::StepGraph* sg = nullptr;
            for (int i = 0; i < G; i++)
                if (h->graphs[base + i].valid && h->graphs[base + i].nstep == nstep && h->graphs[base + i].from == from)
                    sg = &h->graphs[base + i];
            if (!sg) {
                // capture into the slot used least recently
                int& nx = h->graph_next[nvt ? 1 : 0];
                sg = &h->graphs[base + nx];
                nx = (nx + 1) % G;
                if (sg->x) (void)hipGraphExecDestroy(sg->x);
                if (sg->g) (void)hipGraphDestroy(sg->g);
                *sg = mtmd_handle::StepGraph();
                HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                int r = MTMD_OK;
                for (int i = 0; i < nstep && r == MTMD_OK; i++) r = plain_step();
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(h->stream, &g);
                // the capture ran no kernel: the step is launched below
                const mtmd_handle::StepState to = h->step_state();
                h->set_step_state(from);
                if (r) { if (g) (void)hipGraphDestroy(g); return r; }
                if (e != hipSuccess) return h->fail(MTMD_EHIP, "step capture: %s", hipGetErrorString(e));
                sg->g = g;
                HIPCHK(h, hipGraphInstantiate(&sg->x, g, nullptr, nullptr, 0));
                sg->from = from;
                sg->to = to;
                sg->nstep = nstep;
                sg->valid = true;
            }
            HIPCHK(h, hipGraphLaunch(sg->x, h->stream));
            h->set_step_state(sg->to);
            return MTMD_OK;
        };
        // this call's first front half finds both rebuild-flag slots clear
        mt_fast_clear_flags(d, h->stream);
        if (nvt) nvt_front();
        else mt_fast_kick_drift(d, h->p, -1, mixed, h->stream);
        constexpr int KB = mtmd_handle::kBlockSteps;
        for (int t = 0; t < nsteps; t++) {
            const bool last = (t == nsteps - 1);
            // a block of KB plain steps, none of them a re-sort or the last
            if (h->use_graphs && t + KB < nsteps && h->steps_since_sort + KB < h->sort_every) {
                if ((rc = replay(KB))) return rc;
                h->steps_since_sort += KB;
                h->ke_valid = nvt;
                h->steps += KB;
                t += KB - 1;
                continue;
            }
            const bool sort_now = (++h->steps_since_sort >= h->sort_every);
            if (!sort_now && !last && h->use_graphs) {
                if ((rc = replay(1))) return rc;
                h->ke_valid = nvt;
                h->steps++;
                continue;
            }
            if ((rc = fast_rebuild(h, sort_now, sort_now))) return rc;
            if (fuse && !last) {
                fused_forces();
                h->ke_valid = false;
                h->steps++;
                continue;
            }
            bool kicked = false;
            if ((rc = fast_forces(h, last && energy_last, nvt ? &kicked : nullptr))) return rc;
            if (nvt) {
                if (last) nvt_back(kicked);
                else nvt_mid(kicked);
            } else if (last) {
                mt_cap_kick(d, h->p, 0, h->stream);
                mt_scale_wrap(d, h->p, -1, 0, h->stream);
            } else {
                mt_fast_nve_fused(d, h->p, mixed, h->stream);
            }
            h->ke_valid = nvt;
            h->steps++;
        }
        // the last chain of the call left p.p/m after its scale in nh[3M + 1]
        h->ke_slot = nvt;
    }
    HIPCHK(h, hipGetLastError());
    if (h->fast) {
        bool ov = false;
        if ((rc = fast_check_capacity(h, &ov))) return rc;
        if (ov) {
            // the steps after that build ran on truncated lists: grow them for
            // md_steps to re-run the call from its starting state
            HIPCHK(h, hipMemcpy(h->h_status, d.nreq, sizeof(int), hipMemcpyDeviceToHost));
            h->have_forces = false;
            if ((rc = grow_neighbors(h, h->h_status[0]))) return rc;
            HIPCHK(h, hipMemsetAsync(d.flags + 2, 0, sizeof(unsigned int), h->stream));
            return MT_LIST_RERUN;
        }
    }
    return MTMD_OK;
}

static int snapshot_take(mtmd_handle* h)
{
    mtmd_handle::Snapshot& sn = h->snap;
    MtDev& d = h->d;
    const int nh = 3 * std::max(h->M, 1) + 2;
    if (!sn.pos) {
        sn.pos = h->mem.alloc<double4>(h->n);
        sn.mom = h->mem.alloc<double4>(h->n);
        sn.nh = h->mem.alloc<double>(nh);
        if (!sn.pos || !sn.mom || !sn.nh) return h->fail(MTMD_ENOMEM, "state snapshot allocation failed");
    }
    mt_fast_snapshot(d, sn.pos, sn.mom, h->stream);
    HIPCHK(h, hipMemcpyAsync(sn.nh, d.nh, nh * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    sn.eps = h->eps;
    sn.peps = h->peps;
    sn.etaP = h->etaP;
    sn.petaP = h->petaP;
    for (int i = 0; i < 3; i++) sn.L[i] = h->p.L[i];
    sn.steps = h->steps;
    sn.steps_since_sort = h->steps_since_sort;
    sn.ke_slot = h->ke_slot;
    return MTMD_OK;
}

static int snapshot_restore(mtmd_handle* h)
{
    const mtmd_handle::Snapshot& sn = h->snap;
    MtDev& d = h->d;
    mt_fast_restore(d, sn.pos, sn.mom, h->stream);
    HIPCHK(h, hipMemcpyAsync(d.nh, sn.nh, (3 * std::max(h->M, 1) + 2) * sizeof(double), hipMemcpyDeviceToDevice,
                             h->stream));
    HIPCHK(h, hipGetLastError());
    h->eps = sn.eps;
    h->peps = sn.peps;
    h->etaP = sn.etaP;
    h->petaP = sn.petaP;
    set_box(h, sn.L);
    h->steps = sn.steps;
    h->steps_since_sort = sn.steps_since_sort;
    h->ke_slot = sn.ke_slot;                          // (nh[3M + 1] is restored with the chain)
    // forces, lists and the kinetic sums of the restored state are rebuilt
    h->have_forces = false;
    h->need_sort = true;
    h->ke_valid = false;
    return MTMD_OK;
}

// The reference grows the list and builds it again when a row overflows
// (MDFFMatrices.cpp:163-178). Inside a multi-step call the FP64/MIXED engines
// learn of an overflow at the call's end: they grow the lists, restore the
// state the call started from and run it again.
static int md_steps(mtmd_handle* h, int nsteps, bool energy_last)
{
    if (!h->fast) return md_steps_once(h, nsteps, energy_last);
    int rc = snapshot_take(h);
    if (rc) return rc;
    for (int attempt = 0;; attempt++) {
        rc = md_steps_once(h, nsteps, energy_last);
        if (rc != MT_LIST_RERUN) return rc;
        if (attempt == 3)
            return h->fail(MTMD_ENEIGH, "neighbour list overflowed in %d runs of these steps (capacity now %d)",
                           attempt + 1, h->d.ncap);
        if ((rc = snapshot_restore(h))) return rc;
    }
}

extern "C" int mtmd_step(mtmd_handle* h, int32_t nsteps)
{
    if (!h || nsteps < 0) return MTMD_EINVAL;
    if (nsteps == 0) return MTMD_OK;
    return md_steps(h, nsteps, true);
}

// calcTemp/calcPress/calcV/calcUtot/calcCurrentKineticEnergy from the sums
// sum p.p/m, sum F.r, sum U and U_cut
static void thermo_from_sums(const mtmd_handle* h, double sp, double sf, double su, double uc, mtmd_thermo* out)
{
    // calcV / getV (SimulationBox.h:44, VelVerlet.cpp:192-197)
    const double V = h->npt ? npt_volume(h) : h->p.L[0] * h->p.L[1] * h->p.L[2];
    out->sum_p2m = sp;
    out->sum_fr = sf;
    out->T = sp / (3 * (double)h->n) * MT_CONV_T;
    out->P = (1.0 / (3.0 * V)) * (sp + sf) * MT_CONV_PRESS;
    out->V = V;
    out->U = su;
    out->K = 0.5 * sp;
    out->Etot = su + 0.5 * sp;
    out->U_cut = uc;
}

extern "C" int mtmd_get_thermo(mtmd_handle* h, mtmd_thermo* out)
{
    if (!h || !out) return MTMD_EINVAL;
    if (!h->have_forces) {
        int rc = compute_forces_now(h);
        if (rc) return rc;
    }
    MtDev& d = h->d;
    mt_reduce_thermo(d, h->stream);
    const int nb = mt_reduce_blocks(h->n);
    std::vector<double> part(3 * (size_t)nb);
    HIPCHK(h, hipMemcpyAsync(part.data(), d.red, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->ke_valid = false;   // d.red now holds thermo partials
    double sp = 0, sf = 0, su = 0;
    for (int b = 0; b < nb; b++) { sp += part[3 * b]; sf += part[3 * b + 1]; su += part[3 * b + 2]; }
    // U_cut over the lists of the last force evaluation (EXACT: rebuilt at rc
    // for those positions; FP64/MIXED: the skin list, valid until the next step)
    const int nbu = mt_ucut(d, h->p, h->fast ? 0 : 1, h->stream);
    HIPCHK(h, hipMemcpyAsync(part.data(), d.red, nbu * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double uc = 0;
    for (int b = 0; b < nbu; b++) uc += part[b];
    thermo_from_sums(h, sp, sf, su, uc, out);
    return MTMD_OK;
}

// ---------------------------------------------------------------------------
// mtmd_get_thermo_noalloc: the thermo reductions with their block partials in
// a file-scope device array instead of the handle's buffer, and no allocation
// of any kind (device or host). Every GPU the library runs on holds its own
// copy of `scratch` until the process ends; host code reaches it only through
// the address hipGetSymbolAddress returns for the device (the array's name is
// a host-side stand-in, not a device address). Calls that overlap on one GPU
// share that array and race.
// ---------------------------------------------------------------------------
#define MT_NOALLOC_THERMO 3072        // 3 partials (p.p/m, F.r, U) of up to 1024 workgroups
#define MT_NOALLOC_UCUT 1024          // U_cut partials of up to 1024 workgroups
#define MT_NOALLOC_DEVICES 64
static __device__ __align__(256) double scratch[MT_NOALLOC_THERMO + MT_NOALLOC_UCUT];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[MT_NOALLOC_DEVICES] = {};

// the device address of `scratch` on device dev, looked up once per device with
// that device current
static int scratch_address(mtmd_handle* h, double** out)
{
    const int dev = h->device;
    if (dev < 0 || dev >= MT_NOALLOC_DEVICES)
        return h->fail(MTMD_EINVAL, "device %d is beyond the %d devices the scratch table holds", dev,
                       MT_NOALLOC_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        HIPCHK(h, hipGetDevice(&cur));
        if (cur != dev) HIPCHK(h, hipSetDevice(dev));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr)
            return h->fail(MTMD_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return MTMD_OK;
}

extern "C" int mtmd_get_thermo_noalloc(mtmd_handle* h, mtmd_thermo* out)
{
    if (!h || !out) return MTMD_EINVAL;
    if (!h->have_forces)
        return h->fail(MTMD_EINVAL, "no forces of the current positions (call mtmd_compute_forces or mtmd_step "
                                    "first: computing them may grow the neighbour lists)");
    const int nb = mt_reduce_blocks(h->n);
    const int nbu = mt_ucut_blocks(h->d, h->fast ? 0 : 1);
    if (3 * nb > MT_NOALLOC_THERMO || nbu > MT_NOALLOC_UCUT)
        return h->fail(MTMD_EINVAL, "reduction of %d + %d workgroups exceeds the scratch of %d + %d partials", nb, nbu,
                       MT_NOALLOC_THERMO / 3, MT_NOALLOC_UCUT);
    double* sc = nullptr;
    int rc = scratch_address(h, &sc);
    if (rc) return rc;
    // every workgroup writes its own partials, so nothing needs clearing first
    mt_reduce_thermo_to(h->d, sc, h->stream);
    mt_ucut_to(h->d, h->p, h->fast ? 0 : 1, sc + MT_NOALLOC_THERMO, h->stream);
    HIPCHK(h, hipGetLastError());
    double part[MT_NOALLOC_THERMO + MT_NOALLOC_UCUT];
    HIPCHK(h, hipMemcpyAsync(part, sc, 3 * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(part + MT_NOALLOC_THERMO, sc + MT_NOALLOC_THERMO, (size_t)nbu * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // the sums in the order of mtmd_get_thermo
    double sp = 0, sf = 0, su = 0, uc = 0;
    for (int b = 0; b < nb; b++) { sp += part[3 * b]; sf += part[3 * b + 1]; su += part[3 * b + 2]; }
    for (int b = 0; b < nbu; b++) uc += part[MT_NOALLOC_THERMO + b];
    thermo_from_sums(h, sp, sf, su, uc, out);
    return MTMD_OK;
}

// CMDLoop::runOptimization (MDLoop.cpp:143-173): per step, forces at the current
// positions, r += lr F (CSimulationBox::gradientDescentStep), U of those forces;
// stop when |dU| < accuracy after more than out_every steps. per_step(i, seconds)
// runs after step i (the reference's updateOutput); a non-zero return ends the loop.
template <class PerStep>
static int optimize_loop(mtmd_handle* h, int max_steps, double lr, double accuracy, int out_every,
                         int32_t* steps_done, PerStep&& per_step)
{
    int rc;
    auto utot = [&](double& u) -> int {
        mt_reduce_thermo(h->d, h->stream);
        const int nb = mt_reduce_blocks(h->n);
        std::vector<double> part(3 * (size_t)nb);
        HIPCHK(h, hipMemcpyAsync(part.data(), h->d.red, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        u = 0;
        for (int b = 0; b < nb; b++) u += part[3 * b + 2];
        return MTMD_OK;
    };
    // FP64/MIXED (outside NPT): after the first full build the skin lists are
    // rebuilt only when an atom moved more than half the skin, as in an MD
    // step (the lists then hold every pair inside rc, as the reference's
    // per-step rebuild at rc does); EXACT rebuilds at rc every step
    const bool skin_lists = h->fast && !h->npt;
    if ((rc = compute_forces_now(h))) return rc;
    double Un = 0.0;
    if ((rc = utot(Un))) return rc;
    int i = 0;
    for (i = 0; i < max_steps; i++) {
        const auto t0 = std::chrono::steady_clock::now();
        if (!skin_lists) {
            if ((rc = compute_forces_now(h))) return rc;
        } else if (i > 0) {
            // positions moved by the last gradient step: displacement test (and
            // MIXED cell-local coordinates), conditional rebuild, forces
            mt_fast_locate(h->d, h->p, h->engine == MTMD_ENGINE_MIXED, h->stream);
            if ((rc = fast_rebuild(h, false, false))) return rc;
            if ((rc = fast_forces(h, true))) return rc;
            bool ov = false;
            if ((rc = fast_check_capacity(h, &ov))) return rc;
            if (ov && (rc = compute_forces_now(h))) return rc;       // regrow and rebuild in full
            h->have_forces = true;
        }
        mt_gradient_step(h->d, lr, h->stream);
        double Unp1 = 0.0;
        if ((rc = utot(Unp1))) return rc;
        const double dU = std::fabs(Unp1 - Un);
        Un = Unp1;
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        // the forces stay those of the positions before the step, as the
        // reference's F (its thermo line and the next U use them)
        h->ke_valid = false;
        if ((rc = per_step(i, sec))) return rc;
        if (dU < accuracy && i > out_every) { i++; break; }
    }
    h->have_forces = false;
    h->ke_valid = false;
    if (steps_done) *steps_done = i;
    return MTMD_OK;
}

extern "C" int mtmd_optimize(mtmd_handle* h, int32_t max_steps, double lr, double accuracy, int32_t* steps_done)
{
    if (!h) return MTMD_EINVAL;
    return optimize_loop(h, max_steps, lr, accuracy, h->cfg.out_stride, steps_done, [](int, double) { return 0; });
}

extern "C" int mtmd_get_stats(mtmd_handle* h, int64_t* rebuilds, int64_t* steps, int32_t* max_neighbors)
{
    if (!h) return MTMD_EINVAL;
    if (h->fast) {
        bool ov;
        int rc = fast_check_capacity(h, &ov);
        if (rc) return rc;
    }
    if (rebuilds) *rebuilds = h->rebuilds;
    if (steps) *steps = h->steps;
    if (max_neighbors) *max_neighbors = h->d.ncap;
    return MTMD_OK;
}

extern "C" int mtmd_get_skin_lists(mtmd_handle* h, int32_t max_per_atom, int32_t* count, int32_t* list,
                                   int32_t* required)
{
    if (!h || max_per_atom < 0 || (list && max_per_atom == 0)) return MTMD_EINVAL;
    if (!h->fast) return h->fail(MTMD_EINVAL, "skin lists belong to the fp64/mixed engines");
    int rc;
    if (!h->have_forces || h->need_sort) {
        if ((rc = compute_forces_now(h))) return rc;
    }
    const int n = h->n;
    const int maxp = list ? max_per_atom : 0;
    int* d_out = nullptr;
    int* d_cnt = nullptr;
    if (hipMalloc((void**)&d_cnt, (size_t)n * sizeof(int)) != hipSuccess)
        return h->fail(MTMD_ENOMEM, "list counts of %d atoms", n);
    if (maxp && hipMalloc((void**)&d_out, (size_t)n * maxp * sizeof(int)) != hipSuccess) {
        (void)hipFree(d_cnt);
        return h->fail(MTMD_ENOMEM, "lists of %d x %d entries", n, maxp);
    }
    mt_fast_decode_lists(h->d, d_out, d_cnt, maxp, h->stream);
    std::vector<int32_t> cnt(n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), d_cnt, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && maxp)
        e = hipMemcpyAsync(list, d_out, (size_t)n * maxp * sizeof(int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d_cnt);
    if (d_out) (void)hipFree(d_out);
    if (e != hipSuccess) return h->fail(MTMD_EHIP, "list decode: %s", hipGetErrorString(e));
    int mx = 0;
    for (int v : cnt) mx = std::max(mx, v);
    if (required) *required = mx;
    if (count) std::memcpy(count, cnt.data(), (size_t)n * sizeof(int32_t));
    return (maxp && mx > maxp) ? h->fail(MTMD_ENEIGH, "row of %d neighbours exceeds %d", mx, maxp) : MTMD_OK;
}

extern "C" int mtmd_synchronize(mtmd_handle* h)
{
    if (!h) return MTMD_EINVAL;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return MTMD_OK;
}

extern "C" int mtmd_bench_kernel(mtmd_handle* h, int which, int reps, double* avg_ms)
{
    if (!h || reps < 1 || !avg_ms || which < 0 || which > 4) return MTMD_EINVAL;
    if (which == 3 && !h->fast) return h->fail(MTMD_EINVAL, "the copy reference needs the fp64/mixed engine");
    if (which == 4 && !(h->fast && h->engine == MTMD_ENGINE_MIXED && h->d.tile16 && !h->p.coulomb_on))
        return h->fail(MTMD_EINVAL, "the FP64 tiled pass needs the mixed engine in tile mode, without charges");
    int rc;
    if (!h->have_forces || (h->fast && h->need_sort)) {
        if ((rc = compute_forces_now(h))) return rc;
    }
    hipEvent_t e0, e1;
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    HIPCHK(h, hipEventRecord(e0, h->stream));
    for (int r = 0; r < reps; r++) {
        if (which == 0) {
            if (h->fast) rc = fast_forces(h, false);
            else { mt_exact_forces(h->d, h->p, h->stream); rc = MTMD_OK; }
        } else if (which == 1) {
            if (h->fast) rc = fast_rebuild(h, true, false);
            else { mt_exact_cells(h->d, h->p, h->stream); mt_exact_neighbors(h->d, h->p, h->stream); rc = MTMD_OK; }
        } else if (which == 2) {
            rc = md_steps(h, 1, false);
        } else if (which == 4) {
            rc = mt_fast_force_tile64(h->d, h->p, h->stream) ? h->fail(MTMD_EHIP, "FP64 tiled pass launch failed")
                                                             : MTMD_OK;
        } else {
            mt_copy_lists(h->d, h->stream);
            rc = MTMD_OK;
        }
        if (rc) break;
    }
    HIPCHK(h, hipEventRecord(e1, h->stream));
    HIPCHK(h, hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (which == 4) {
        unsigned int skipped = 0;
        HIPCHK(h, hipMemcpy(&skipped, h->d.flags + 6, sizeof skipped, hipMemcpyDeviceToHost));
        if (skipped) {
            HIPCHK(h, hipMemset(h->d.flags + 6, 0, sizeof skipped));
            return h->fail(MTMD_EINVAL, "the FP64 tiled pass left out %u blocks of 32-bit lists", skipped / reps);
        }
    }
    *avg_ms = (double)ms / reps;
    return MTMD_OK;
}

// ---------------------------------------------------------------------------
// outer seam: CMDSimulator::run / optimize (Simulator/MDSimulator.cpp:35-66)
// with the thermo table of CMDLoop::printHeading/updateOutput (MDLoop.cpp:194-227,370-405)
// ---------------------------------------------------------------------------
static void tee(FILE* a, FILE* b, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (a) { std::fputs(buf, a); std::fflush(a); }
    if (b) std::fputs(buf, b);
}

// CMDLoop::appendToXYZFile (MDLoop.cpp:229-246)
extern "C" int mtmd_append_xyz(mtmd_handle* h, const char* path, int32_t t)
{
    if (!h || !path) return MTMD_EINVAL;
    std::vector<double> xyz(3 * (size_t)h->n);
    int rc = mtmd_get_state(h, xyz.data(), nullptr);
    if (rc) return rc;
    std::vector<const char*> names(h->T);
    for (int i = 0; i < h->T; i++) names[i] = h->type_names[i].c_str();
    if (mt_xyz_append(path, h->n, t, xyz.data(), h->type_host.data(), names.data(), h->T))
        return h->fail(MTMD_EINVAL, "cannot open %s", path);
    return MTMD_OK;
}

// CMDLoop::appendToDCDFile (MDLoop.cpp:248-258), DCDFile.cpp:670-729
extern "C" int mtmd_append_dcd(mtmd_handle* h, const char* path, int32_t n_steps, int32_t stride)
{
    if (!h || !path) return MTMD_EINVAL;
    std::vector<double> xyz(3 * (size_t)h->n);
    int rc = mtmd_get_state(h, xyz.data(), nullptr);
    if (rc) return rc;
    // the record's box is the current simulation box narrowed to int
    // (std::tuple<int, int, int> parameter of CDCDFile::appendToDCDFile)
    const int box[3] = {(int)h->p.L[0], (int)h->p.L[1], (int)h->p.L[2]};
    if (mt_dcd_create_if_missing(path, n_steps, stride, h->cfg.timestep, h->n) ||
        mt_dcd_append(path, h->n, box, xyz.data()))
        return h->fail(MTMD_EINVAL, "cannot write %s", path);
    return MTMD_OK;
}

// Trajectory frames of CMDLoop::updateOutput (MDLoop.cpp:384-386). Without the
// Gromacs XTC library the reference prints an error and writes the frame to
// the DCD file instead (MDLoop.cpp:288-292), which is what happens here.
static int write_frames(mtmd_handle* h, const mtmd_config* cfg, int t, int n_steps, int stride)
{
    int rc = MTMD_OK;
    if (cfg->include_xyz && (rc = mtmd_append_xyz(h, cfg->xyz_file, t))) return rc;
    if (cfg->include_dcd && (rc = mtmd_append_dcd(h, cfg->dcd_file, n_steps, stride))) return rc;
    if (cfg->include_xtc) {
        std::printf("Error, missing library. MolTwister was not built with the Gromacs XTC library, likely because "
                    "it was not available during the MolTwister build. Using DCD instead!\r\n");
        if ((rc = mtmd_append_dcd(h, cfg->dcd_file, n_steps, stride))) return rc;
    }
    return rc;
}

// Histograms of CMDLoop (MDLoop.cpp:104-140,296-366,407-432): momenta of every
// atom per axis and the first thermostat momentum, and the volume, every step
// after the equilibration steps; written when the run ends.
struct Distributions {
    static constexpr int kBins = 150;          // numPDistrBins (MDLoop.cpp:111)
    static constexpr int kEquilib = 1000;      // equilibSteps (MDLoop.cpp:112)
    bool p_on = false, v_on = false;
    double maxP = 0, maxPeta = 0, maxV = 0;
    int* dev = nullptr;                        // [4 * kBins] on the device
    std::vector<int> vol;

    int init(mtmd_handle* h, const mtmd_config* cfg)
    {
        p_on = cfg->include_pdistr != 0;
        v_on = cfg->include_vdistr != 0;
        maxP = cfg->max_pdistr / 1.0E-2;                     // Conv_P (Constants.h:28)
        // calcMaxPetaDistrOutput (MDLoop.cpp:429-432): 2 Q_0 / tau of the
        // temperature chain, Q_0 = dim N T tau^2 (NHChain.cpp:115-124)
        const double T = cfg->temperature / MT_CONV_T, tau = cfg->temperature_relax / MT_CONV_t;
        maxPeta = 2.0 * ((double)(3 * h->n) * T * tau * tau) / tau;
        maxV = cfg->max_vdistr;
        vol.assign(kBins, 0);
        if (p_on) {
            dev = h->mem.alloc<int>(4 * kBins);
            if (!dev) return h->fail(MTMD_ENOMEM, "histogram allocation failed");
            HIPCHK(h, hipMemsetAsync(dev, 0, 4 * kBins * sizeof(int), h->stream));
        }
        return MTMD_OK;
    }
    bool active(int t) const { return (p_on || v_on) && t > kEquilib; }
    int accumulate(mtmd_handle* h)
    {
        if (p_on) {
            // NVE has no active chain; its thermostat momentum stays at rest (0)
            const double* peta = (h->M > 0) ? h->d.nh + h->M : nullptr;
            if (mt_pdistr(h->d, maxP, kBins, peta, maxPeta, dev, h->stream))
                return h->fail(MTMD_EINVAL, "histogram bins");
            if (!peta) {
                const int b = (int)(((0.0 + maxPeta) / (2.0 * maxPeta)) * (double)kBins);
                peta_rest_bin_hits += (b >= 0 && b < kBins) ? 1 : 0;
                peta_rest_bin = b;
            }
        }
        if (v_on) {
            // CMDLoop::appendToVolumeDistribution (MDLoop.cpp:324-328)
            const double V = h->npt ? npt_volume(h) : h->p.L[0] * h->p.L[1] * h->p.L[2];
            const int i = (int)((V / maxV) * (double)kBins);
            if (i < kBins) vol[i]++;
        }
        return MTMD_OK;
    }
    int finish(mtmd_handle* h, const mtmd_config* cfg)
    {
        if (p_on) {
            std::vector<int> hp(4 * kBins);
            HIPCHK(h, hipMemcpyAsync(hp.data(), dev, hp.size() * sizeof(int), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            if (peta_rest_bin_hits) hp[3 * kBins + peta_rest_bin] += peta_rest_bin_hits;
            const char* suffix[4] = {"_p_x.dat", "_p_y.dat", "_p_z.dat", "_p_eta.dat"};
            for (int a = 0; a < 4; a++) {
                const std::string f = std::string(cfg->pdistr_file) + suffix[a];
                (void)mt_store_pdistr(f.c_str(), hp.data() + a * kBins, kBins, a < 3 ? maxP : maxPeta, a);
            }
        }
        if (v_on) {
            const std::string f = std::string(cfg->vdistr_file) + "_v.dat";
            (void)mt_store_vdistr(f.c_str(), vol.data(), kBins, maxV);
        }
        return MTMD_OK;
    }
    int peta_rest_bin = 0, peta_rest_bin_hits = 0;
};

static int run_common(const mtmd_config* cfg, const mtmd_system* sys, void* log_file, bool optimize)
{
    mtmd_handle* h = nullptr;
    int rc = mtmd_create(sys, cfg, &h);
    if (rc) return rc;
    // COut::setOutputFile prints a warning when the info file cannot be opened (Printf.cpp:46-50)
    FILE* info = cfg->info_file[0] ? std::fopen(cfg->info_file, "w") : nullptr;
    FILE* log = (FILE*)log_file;
    if (!info) tee(nullptr, log, "Warning! Could not open output file!\r\n");
    const int stride = std::max(1, cfg->out_stride);
    const char* ens = (cfg->ensemble == MTMD_NPT) ? "NPT" : ((cfg->ensemble == MTMD_NVT) ? "NVT" : "NVE");
    const bool thermostat = (cfg->ensemble == MTMD_NVT || cfg->ensemble == MTMD_NPT);
    tee(info, log, "\r\n\t----------------------------\r\n\t%s\r\n\t----------------------------\r\n",
        optimize ? "Optimization" : "Molecular dynamics simulation");
    if (!optimize) {
        tee(info, log, "\t Ensemble = %s\r\n", ens);
        if (thermostat) tee(info, log, "\t Temperature, T = %g K\r\n", (cfg->temperature / MT_CONV_T) * MT_CONV_T);
        if (cfg->ensemble == MTMD_NPT)
            tee(info, log, "\t Pressure, P = %g atm\r\n", (cfg->pressure / MT_CONV_PRESS) * MT_CONV_PRESS);
        if (thermostat)
            tee(info, log, "\t NH T relax, tau = %g fs\r\n", (cfg->temperature_relax / MT_CONV_t) * MT_CONV_t);
        if (cfg->ensemble == MTMD_NPT)
            tee(info, log, "\t NH P relax, tau = %g fs\r\n", (cfg->pressure_relax / MT_CONV_t) * MT_CONV_t);
        tee(info, log, "\t Timestep, dt = %g fs\r\n", (cfg->timestep / MT_CONV_t) * MT_CONV_t);
        tee(info, log, "\t Particles, N = %i\r\n", sys->n_atoms);
        tee(info, log, "\t NH RESPA, n = %i\r\n", cfg->temperature_respa);
        tee(info, log, "\t NH Chain length, M = %i\r\n", cfg->temperature_nh_len);
        tee(info, log, "\t Init. box length, Lx = %g, Ly = %g, Lz = %g AA\r\n", sys->box[0], sys->box[1], sys->box[2]);
        tee(info, log, "\t Dimension, d = %i\r\n", 3);
    } else {
        tee(info, log, "\t Learning rate, L = %g\r\n", cfg->opt_learning_rate);
    }
    tee(info, log, "\t----------------------------\r\n\r\n");
    tee(info, log, "\t%-15s%-17s%-15s%-15s%-20s%-15s%-15s%-15s\r\n", optimize ? "Step" : "Timestep", "Avg.steptime[s]",
        "Temp[K]", "Press[atm]", "Vol[AA^3]", "U[kJ/mol]", "K[kJ/mol]", "Etot[kJ/mol]");
    // one output of CMDLoop::updateOutput (MDLoop.cpp:380-404): frames, then the thermo line
    // Avg.steptime: mean wall time of the steps since the last line (CMDLoop::readTimerAverage)
    double acc_s = 0.0;
    int acc_n = 0;
    auto output = [&](int t, int total) -> int {
        const double avg = acc_n ? acc_s / acc_n : -1.0;
        acc_s = 0.0;
        acc_n = 0;
        int r = write_frames(h, cfg, t, total, stride);
        if (r) return r;
        mtmd_thermo th{};
        if ((r = mtmd_get_thermo(h, &th))) return r;
        // the reference passes (avg, t) to a "%-15i%-17.8f" format (MDLoop.cpp:402),
        // which the x86-64 calling convention prints as these columns
        tee(info, log, "\t%-15i%-17.8f%-15g%-15g%-20g%-15g%-15g%-15g\r\n", t, avg, th.T, th.P, th.V, th.U, th.K, th.Etot);
        return 0;
    };
    if (!optimize) {
        Distributions dist;
        rc = dist.init(h, cfg);
        const int total = cfg->n_steps;
        int done = 0;
        while (done < total && !rc) {
            // steps t = done .. done+chunk-1; output after step t when t % stride == 0;
            // past the equilibration steps with histograms on, one step at a time
            int chunk;
            if (dist.active(done)) {
                chunk = 1;
            } else {
                int next = (done % stride == 0) ? done : done + (stride - done % stride);
                if ((dist.p_on || dist.v_on) && done <= Distributions::kEquilib)
                    next = std::min(next, Distributions::kEquilib);
                chunk = std::min(next, total - 1) - done + 1;
            }
            auto t0 = std::chrono::steady_clock::now();
            rc = md_steps(h, chunk, true);
            if (rc) break;
            (void)mtmd_synchronize(h);
            acc_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            acc_n += chunk;
            done += chunk;
            const int t = done - 1;
            if (dist.active(t) && (rc = dist.accumulate(h))) break;
            if (t % stride == 0) rc = output(t, total);
        }
        if (!rc) rc = dist.finish(h, cfg);
    } else {
        rc = optimize_loop(h, cfg->opt_max_steps, cfg->opt_learning_rate, cfg->opt_accuracy, stride, nullptr,
                           [&](int i, double sec) {
                               acc_s += sec;
                               acc_n++;
                               return (i % stride == 0) ? output(i, cfg->opt_max_steps) : 0;
                           });
    }
    if (rc && log) std::fprintf(log, "Error: %s\r\n", h->err.c_str());
    if (info) std::fclose(info);
    mtmd_destroy(h);
    return rc;
}

extern "C" int mtmd_run(const mtmd_config* cfg, const mtmd_system* sys, void* log_file)
{
    if (!cfg || !sys) return MTMD_EINVAL;
    return run_common(cfg, sys, log_file, false);
}

extern "C" int mtmd_run_optimize(const mtmd_config* cfg, const mtmd_system* sys, void* log_file)
{
    if (!cfg || !sys) return MTMD_EINVAL;
    return run_common(cfg, sys, log_file, true);
}
