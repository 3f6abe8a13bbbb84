// solver.h - host-side state of one IBMSolver instance (opaque to callers).
#pragma once

#include <stdexcept>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/ibm3d.h"
#include "common.h"
#include "solver_kernels.h"
#include "step_kernels.h"

namespace ibm {

// Device buffer owned by the solver.
// Process-wide lock around HIP calls that touch the legacy default stream or
// the allocator (hipMemset / hipMemcpy / hipMalloc / hipFree, stream and
// event creation) and around stream capture: HIP rejects legacy-stream work
// issued by one thread while another thread captures a graph, and several
// solvers may run on different threads (the serving loop).
inline std::recursive_mutex &hip_global_mutex() {
    static std::recursive_mutex m;
    return m;
}
using HipGlobalLock = std::lock_guard<std::recursive_mutex>;

// Device bytes allocated by the solver whose API call is running on this
// thread (set by a scope guard in the setup entry points); null elsewhere.
inline thread_local size_t *tl_dev_bytes = nullptr;

struct DevBytesScope {
    size_t *prev;
    explicit DevBytesScope(size_t *c) : prev(tl_dev_bytes) { tl_dev_bytes = c; }
    ~DevBytesScope() { tl_dev_bytes = prev; }
};

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) { free(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { free(); }
    void alloc(size_t cnt) {
        HipGlobalLock lk(hip_global_mutex());
        free();
        n = cnt;
        if (cnt) {
            HIP_CHECK(hipMalloc(&p, cnt * sizeof(T)));
            // the memset runs on the legacy stream, which the solvers'
            // non-blocking streams do not wait for: complete it here, or a
            // busy device could apply it after kernels have written the buffer
            HIP_CHECK(hipMemset(p, 0, cnt * sizeof(T)));
            HIP_CHECK(hipStreamSynchronize(nullptr));
            if (tl_dev_bytes) *tl_dev_bytes += cnt * sizeof(T);
        }
    }
    void upload(const T *h, size_t cnt) {
        HipGlobalLock lk(hip_global_mutex());
        if (cnt > n) alloc(cnt);
        if (cnt) {
            // from pageable memory the call may return before the DMA lands;
            // the solvers' non-blocking streams do not wait for it
            HIP_CHECK(hipMemcpy(p, h, cnt * sizeof(T), hipMemcpyHostToDevice));
            HIP_CHECK(hipStreamSynchronize(nullptr));
        }
    }
    void upload(const std::vector<T> &h) { upload(h.data(), h.size()); }
    void free() {
        if (!p) return;
        HipGlobalLock lk(hip_global_mutex());
        (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct HostBC {
    IBMSolverBCType type = BC_STATIONARY_WALL;
    IBMSolverValType val_type = VAL_CONST;
    double cu[3] = {0, 0, 0}, cp = 0;
    IBMSolverBCValFunc fu[3] = {nullptr, nullptr, nullptr}, fp = nullptr;
    double uc = 0;  // BC_CONVECTIVE_OUTFLOW speed (<= 0: mean outward velocity)
    bool set = false;
};

// Per-axis folded coefficients of one system, both precisions.
struct SysData {
    DBuf<double> c64[3][4];   // cm, cp, dm, dp per axis
    DBuf<float> c32[3][4];
    std::vector<double> h[3][4];  // host copies of c64
    DBuf<double> p64[3];          // packed [n][4] per axis
    DBuf<float> p32[3];
    SysCoef<double> S64{};
    SysCoef<float> S32{};
    double gamma[3][2] = {};  // RHS fold factor per face
    double bc_norm2 = 0;      // squared norm of the BC rows' RHS (constant data)
};

struct GhostData {
    DBuf<long> cell, src;
    DBuf<int> srcg;
    DBuf<double> w64, ws64, v64[2];
    DBuf<float> w32, ws32, v32[2];
    GhostRows<double> G64{};
    GhostRows<float> G32{};
};

// One multigrid level of the pressure system.
struct MGLevel {
    Dims d{};
    std::vector<double> dx, dy, dz, xc, yc, zc;  // padded (+2)
    std::vector<uint8_t> hflag;
    DBuf<uint8_t> flag;
    DBuf<double> dxyz;                 // dx | dy | dz padded
    SysData sys;
    DBuf<double> x64, b64, r64, xa64;  // xa: second buffer of the out-of-place smoother
    DBuf<float> x32, b32, r32, xa32;
    // transfer to this level from the finer one (indices over the finer grid)
    DBuf<int> ci, cj, ck, si, sj, sk, pi, pj, pk;
    DBuf<double> wi, wj, wk;
    DBuf<uint8_t> fmask;               // faces to inactive neighbours (SysCoef::fmask)
    DBuf<uint8_t> pmask;               // MGTransfer::pm (over the finer level's cells)
    MGTransfer tr{};
    // exact solve of the coarsest level: dense inverse (row-major, rows in
    // x-fastest cell order), padded index and volume of each row's cell
    int dn = 0, dpin = -1;
    DBuf<double> dinv64, dvol;
    DBuf<float> dinv32;
    DBuf<long> drow;
};

struct KTimer {
    std::string name;
    double bytes = 0;       // algorithmic bytes summed over the launches begun
    double total_ms = 0;
    long count = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

struct VGraph {
    const void *in;
    void *out;
    hipGraphExec_t exec;
};

struct Krylov {
    DBuf<double> v64[8];
    DBuf<float> v32[8];
};

// Restarted GMRES storage (pressure, SOLVER_GMRES): the Arnoldi basis V,
// the preconditioned directions Z, one unscaled vector, device scalars and
// the block partial sums of the multi-dot kernels.
constexpr int kGmresRestart = 10;  // Krylov dimension: src/ibm3d_setup.c:1474
struct GmresData {
    DBuf<double> V[kGmresRestart + 1], Z[kGmresRestart], t;
    DBuf<double> dev, partial;  // dev: 64 scalars
};

// Krylov scalars published by the finalize kernel into device-mapped,
// coherent host memory with an epoch counter (the host waits on it).
struct KPub {
    KScal ks;
    unsigned epoch;
    unsigned pad;
};

// One lane of the concurrent momentum solves (components 1 and 2; component
// 0 uses the solver's own stream and buffers): a stream with its own Krylov
// vectors, device scalars, block partials and publication slot.
struct MomLane {
    hipStream_t st = nullptr;
    DBuf<double> v[8];
    DBuf<float> d32[2];         // FP32 preconditioned directions (mixed-Krylov mode)
    DBuf<KScal> ks;
    DBuf<double> partial;
    KPub *pub_h = nullptr, *pub_d = nullptr;
    unsigned epoch = 0;
    hipEvent_t done = nullptr;  // recorded on st after the lane's last kernel
};

}  // namespace ibm

struct _ibm_solver {
    // ---- configuration (reference fields, src/ibm3d.h:100-213) ----
    int Nx = 0, Ny = 0, Nz = 0;
    double Re = 1, dt = 1;
    int iter = 0;
    double time = 0;
    const char *autosave_filename = nullptr;
    int autosave_period = 0;
    std::vector<double> xf, yf, zf;
    std::vector<double> dx, dy, dz, xc, yc, zc;           // padded (+2)
    std::vector<double> kW, kE, kS, kN, kD, kU;           // padded (+2)
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0, zmin = 0, zmax = 0;
    ibm::HostBC bc[6];
    IBMSolverValType force_type = VAL_CONST;
    double const_f[3] = {0, 0, 0};
    IBMSolverForceFunc func_f[3] = {nullptr, nullptr, nullptr};
    IBMSolverTurbModelType turb_type = TURBMODEL_NONE;
    double turb_coef = 0;
    Polyhedron *poly = nullptr;
    int has_sphere = 0;
    double sphere[4] = {0, 0, 0, 0};
    IBMSolverLinearSolverType solver_type = SOLVER_BiCGSTAB;
    IBMSolverPrecondType precond_type = PRECOND_AMG;
    double tol = 1e-6;
    int precision = IBM3D_PREC_FP64;
    bool assembled = false;
    bool p_singular = true;
    bool p_mg64 = false;         // mixed-Krylov mode: the pressure preconditioner fell back to the FP64
                                 // multigrid cycle after a solve with repeated stalls
    size_t dev_bytes = 0;        // device memory allocated during setup
    double vd_aplus = 26, vd_utau = 1;  // van Driest damping (TURBMODEL_SMAGORINSKY_VD)
    int vd_walls = 0;
    double mg_gamma = 3.5;       // coarsening threshold (multiples of the level's smallest width)
    int mg_pre = 2, mg_post = 2; // smoothing sweeps per level
    int mg_smoother = 1;         // 0 fused red-black sweeps, 1 one launch per colour
    int mg_light_from = 1;       // levels >= this (if > 0) smooth with one sweep before and after
    double mg_omega = 1.3;       // over-relaxation factor of the pressure multigrid smoother
    int mom_solver = 0;          // IBMSolver_set_momentum_solver: 0 automatic, 1 BiCGSTAB
    int mom_sweeps = -1;         // FP32 sweeps per refinement round of the momentum solves (0: not used; -1: unset)
    double mom_rate = 0;         // estimated red-black sweep contraction of the momentum operators

    // ---- host mirrors ----
    std::vector<double> lvset_h;     // reference layout (k fastest)
    std::vector<int> flag_ref;       // reference layout
    std::vector<uint8_t> flag_dev_h; // device layout

    // ---- device state ----
    hipStream_t stream = nullptr;
    ibm::Dims d{};
    ibm::Geo geo{};
    ibm::DBuf<double> g1d[12];       // dx dy dz xc yc zc kW kE kS kN kD kU
    ibm::DBuf<double> ks1d[3];       // kW+kE, kS+kN, kD+kU (interior, unpadded)
    ibm::DBuf<uint8_t> flag;
    ibm::DBuf<double> lvset;
    ibm::DBuf<double> u[3], un[3], us[3], F[3], Fn[3], Fs[3], N[3], Np[3], b[3];
    ibm::DBuf<double> p, pn, pp, bp;
    ibm::DBuf<double> tau[6];
    ibm::DBuf<double> force_arr[3];
    ibm::DBuf<double> bc_arr[3][2];  // host-evaluated BC data of VAL_FUNC faces
    ibm::DBuf<double> conv_g[6];     // BC_CONVECTIVE_OUTFLOW face values (4 comps, bc[] order)
    bool conv_ready = false;         // conv_g hold the current face values
    ibm::DBuf<double> partial, sums;  // sums: kSumSlots doubles
    // sums_h: device-mapped coherent host memory, kSumSlots doubles then an
    // epoch; sums_d / sums_d_epoch: the same memory's device addresses
    double *sums_d = nullptr;
    unsigned *sums_h_epoch = nullptr, *sums_d_epoch = nullptr;
    unsigned sums_epoch = 0;
    ibm::DBuf<ibm::KScal> ks;
    ibm::KScal *ks_h = nullptr;      // pinned
    ibm::KPub *pub_h = nullptr;      // host address (see KPub)
    ibm::KPub *pub_d = nullptr;      // device address of the same memory
    unsigned pub_epoch = 0;
    // concurrent momentum solves (IBM3D_PREC_FP64 and _MIXED_KRYLOV)
    ibm::MomLane mlane[2];
    hipEvent_t ev_rhs = nullptr;     // momentum right-hand sides formed (solver's stream)
    bool mom_lanes = false;
    double *sums_h = nullptr;        // pinned
    double *gm_h = nullptr;          // pinned, 64 scalars of the GMRES / PCG solvers

    // ghost cells
    int ng = 0;
    std::vector<ibm::GhostCell> ghosts;   // level-sorted
    std::vector<int> ghost_lex;           // lexicographic order -> slot
    std::vector<int> level_start;
    ibm::DBuf<ibm::GhostCell> gc;
    ibm::DBuf<int> ghost_ijk;             // i, j, k of each ghost cell, lexicographic order
    ibm::DBuf<double> ghost_old;
    // body force after each step of IBMSolver_iterate (IBMSolver_set_force_history)
    ibm::DBuf<double> force_hist;         // [force_hist_max][3]
    int force_hist_max = 0, force_hist_n = 0;
    ibm::GhostData gmom, gp;              // solver ghost rows (momentum / pressure)
    int ghost_prev = -1;                  // where the pressure ghost values are (launch_ghost_relax)

    ibm::SysData sys[4];                  // u1, u2, u3, p
    long pin = -1;
    std::vector<ibm::MGLevel> mg;         // mg[0] = fine pressure level
    std::vector<ibm::VGraph> vgraphs;     // captured V-cycles per (input, output)
    ibm::Krylov kry;
    ibm::GmresData gm;                    // SOLVER_GMRES / SOLVER_PCG / SOLVER_AMG scratch (pressure)
    ibm::DBuf<double> ref_r64;            // mixed mode: FP64 defect
    ibm::DBuf<float> ref_rhs32, ref_e32;  // mixed mode: FP32 defect and correction
    ibm::DBuf<float> ref_e32b;            // mixed-Krylov mode: the second cycle output of an iteration

    // diagnostics
    double final_norm[4] = {0, 0, 0, 0};
    int final_iter[4] = {0, 0, 0, 0};
    int timing = 0;             // 0 off, 1 every kernel (eager), 2 kernels + whole V-cycle graphs
    bool graphs = true;         // replay multigrid cycles as HIP graphs
    std::vector<double> step_ms;          // each step of the last iterate call: device time between
                                          // the ends of consecutive steps (ms)
    std::vector<ibm::KTimer> timers;
    std::vector<hipEvent_t> event_pool;
};

namespace ibm {

// slots of IBMSolver::sums: 0..11 reductions of the step kernels, 12..29 the
// convective outflow faces' BC-row norms (3 per face, bc[] order)
constexpr int kSumSlots = 32;
constexpr int kConvNormSlot = 12;
constexpr int kConvSpeedSlot = 31;  // convection speed of the face being advanced

// Errors of one case (rejected input, a non-finite residual, a lost Krylov
// publication, a failed device allocation). Outside a serving job they print
// and abort, as the reference does; on a serving worker thread, while the
// worker runs a job (JobErrorScope), they print and throw JobError, which the
// worker catches: the job fails, the other cases and the process go on.
struct JobError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
inline thread_local bool tl_job_errors_throw = false;
struct JobErrorScope {
    bool prev;
    JobErrorScope() : prev(tl_job_errors_throw) { tl_job_errors_throw = true; }
    ~JobErrorScope() { tl_job_errors_throw = prev; }
};
// prints msg (to stdout when on_stdout, else stderr), then throws or aborts
[[noreturn]] void case_error(const std::string &msg, bool on_stdout = false);

// api.cpp
long host_ref_index(const IBMSolver *s, int i, int j, int k);
// case_error("error: " + msg)
[[noreturn]] void fatal(const char *msg);
void build_system(const IBMSolver *s, const std::vector<double> *D[3], const std::vector<double> *C[3],
                  const std::vector<double> *Km[3], const std::vector<double> *Kp[3], const int n[3],
                  const IBMSolverBCType bct[3][2], int sk, SysData &out);

// step.cpp
void iterate_impl(IBMSolver *s, int steps, bool verbose);
void build_mg(IBMSolver *s);
double max_div_impl(IBMSolver *s);
void body_force_impl(IBMSolver *s, double *dst);
void body_force_read(IBMSolver *s, double out[3]);

// timing helpers (step.cpp)
int timer_begin(IBMSolver *s, const char *name, double bytes);
void timer_end(IBMSolver *s, int id);
void timers_flush(IBMSolver *s);

// geo3d.cpp
void polyhedron_cpt_grid(const Polyhedron *poly, int nx, int ny, int nz, const double *x, const double *y,
                         const double *z, double *f, double maxd);
int polyhedron_cpt_grid_noalloc(const Polyhedron *poly, int nx, int ny, int nz, const double *x, const double *y,
                                const double *z, double *f, double maxd);

// export.cpp
void export_result_impl(IBMSolver *s, const char *filename);
void export_lvset_flag_impl(IBMSolver *s, const char *filename);
void export_intermediate_impl(IBMSolver *s, const char *filename);
void init_flow_file_impl(IBMSolver *s, const char *filename);

}  // namespace ibm
