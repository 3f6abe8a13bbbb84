// api.cpp - the IBMSolver C API: construction, setters, assembly, field I/O.
//
// Function names, arguments and ownership follow the reference
// (src/ibm3d_setup.c:48-639, src/ibm3d_init.c:15-262, src/utils.c:111-171).
// Assembly differs by design: instead of building four Hypre ParCSR matrices
// (src/ibm3d_setup.c:1566-2908) it folds the outer boundary rows into
// separable per-axis stencil coefficients, precomputes the immersed-boundary
// ghost stencils once (they only depend on the static level set), and builds
// the geometric multigrid hierarchy of the pressure system.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "solver.h"

using namespace ibm;

namespace ibm {

void case_error(const std::string &msg, bool on_stdout) {
    FILE *f = on_stdout ? stdout : stderr;
    fputs(msg.c_str(), f);
    if (msg.empty() || msg.back() != '\n') fputc('\n', f);
    fflush(stdout);
    fflush(stderr);
    if (tl_job_errors_throw) throw JobError(msg);
    abort();
}

void fatal(const char *msg) { case_error(std::string("error: ") + msg); }

void hip_failure(hipError_t e, const char *expr, const char *file, int line) {
    char buf[512];
    snprintf(buf, sizeof buf, "error: HIP call %s failed at %s:%d: %s", expr, file, line, hipGetErrorString(e));
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) case_error(buf);
    fprintf(stderr, "%s\n", buf);
    fflush(stdout);
    fflush(stderr);
    abort();
}

long host_ref_index(const IBMSolver *s, int i, int j, int k) {
    return ((long)(i + 2) * (s->Ny + 4) + j + 2) * (s->Nz + 4) + k + 2;
}

static bool isperiodic(IBMSolverBCType t) { return t == BC_ALL_PERIODIC || t == BC_VELOCITY_PERIODIC; }

// bc[] index of face (axis, high side): order N, E, S, W, D, U
static int face_bc(int ax, int hi) {
    static const int t[3][2] = {{3, 1}, {2, 0}, {4, 5}};
    return t[ax][hi];
}

}  // namespace ibm

#define H1(a, i) ((a)[(i) + 2])

extern "C" {

IBMSolver *IBMSolver_new(const int num_process, const int rank, const int Px, const int Py, const int Pz) {
    if (num_process != 1 || rank != 0 || Px * Py * Pz != 1) {
        char buf[160];
        snprintf(buf, sizeof buf, "error: one GPU holds the whole domain: num_process must be 1 (got %d, %dx%dx%d)",
                 num_process, Px, Py, Pz);
        case_error(buf);
    }
    // the device resources are created by IBMSolver_assemble once its input
    // checks pass (init_device), so a rejected setup never touches the GPU
    return new IBMSolver();
}

void IBMSolver_destroy(IBMSolver *s) {
    if (!s) return;
    if (!s->stream) {  // never assembled: no device resources
        delete s;
        return;
    }
    // a job that failed part-way may have left work queued on any stream
    (void)hipStreamSynchronize(s->stream);
    for (auto &L : s->mlane)
        if (L.st) (void)hipStreamSynchronize(L.st);
    HipGlobalLock lk(hip_global_mutex());
    for (auto &t : s->timers)
        for (auto &e : t.pending) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    for (auto e : s->event_pool) (void)hipEventDestroy(e);
    for (auto &g : s->vgraphs) (void)hipGraphExecDestroy(g.exec);
    (void)hipHostFree(s->ks_h);
    (void)hipHostFree(s->pub_h);
    for (auto &L : s->mlane) {
        if (L.pub_h) (void)hipHostFree(L.pub_h);
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.st) (void)hipStreamDestroy(L.st);
    }
    if (s->ev_rhs) (void)hipEventDestroy(s->ev_rhs);
    (void)hipHostFree(s->sums_h);
    (void)hipHostFree(s->gm_h);
    (void)hipStreamDestroy(s->stream);
    delete s;  // device buffers are released by their owners' destructors
}

// <- src/ibm3d_setup.c:203-284 (one process: local grid = global grid)
void IBMSolver_set_grid(IBMSolver *s, const int Nx, const int Ny, const int Nz, const double *xf, const double *yf,
                        const double *zf) {
    if (Nx < 2 || Ny < 2 || Nz < 2) fatal("grid needs at least 2 cells per axis");
    s->Nx = Nx; s->Ny = Ny; s->Nz = Nz;
    s->xf.assign(xf, xf + Nx + 1);
    s->yf.assign(yf, yf + Ny + 1);
    s->zf.assign(zf, zf + Nz + 1);
    s->dx.assign(Nx + 4, 0); s->xc.assign(Nx + 4, 0);
    s->dy.assign(Ny + 4, 0); s->yc.assign(Ny + 4, 0);
    s->dz.assign(Nz + 4, 0); s->zc.assign(Nz + 4, 0);
    for (int i = 0; i < Nx; i++) { H1(s->dx, i) = xf[i + 1] - xf[i]; H1(s->xc, i) = (xf[i + 1] + xf[i]) / 2; }
    for (int j = 0; j < Ny; j++) { H1(s->dy, j) = yf[j + 1] - yf[j]; H1(s->yc, j) = (yf[j + 1] + yf[j]) / 2; }
    for (int k = 0; k < Nz; k++) { H1(s->dz, k) = zf[k + 1] - zf[k]; H1(s->zc, k) = (zf[k + 1] + zf[k]) / 2; }
    s->xmin = xf[0]; s->xmax = xf[Nx];
    s->ymin = yf[0]; s->ymax = yf[Ny];
    s->zmin = zf[0]; s->zmax = zf[Nz];
}

void IBMSolver_set_params(IBMSolver *s, const double Re, const double dt) {
    s->Re = Re;
    s->dt = dt;
}

// <- src/ibm3d_setup.c:306-327
void IBMSolver_set_ext_force(IBMSolver *s, IBMSolverValType t, ...) {
    va_list ap;
    va_start(ap, t);
    s->force_type = t;
    for (int l = 0; l < 3; l++) {
        if (t == VAL_CONST) s->const_f[l] = va_arg(ap, double);
        else s->func_f[l] = va_arg(ap, IBMSolverForceFunc);
    }
    va_end(ap);
}

// <- src/ibm3d_setup.c:336-356 (TURBMODEL_VREMAN is an extension)
void IBMSolver_set_turb_model(IBMSolver *s, IBMSolverTurbModelType t, ...) {
    va_list ap;
    va_start(ap, t);
    s->turb_type = t;
    switch (t) {
    case TURBMODEL_NONE: break;
    case TURBMODEL_SMAGORINSKY:
    case TURBMODEL_VREMAN: s->turb_coef = va_arg(ap, double); break;
    case TURBMODEL_SMAGORINSKY_VD:
        s->turb_coef = va_arg(ap, double);
        s->vd_aplus = va_arg(ap, double);
        s->vd_utau = va_arg(ap, double);
        s->vd_walls = va_arg(ap, int);
        break;
    default:
        va_end(ap);
        fatal("unsupported turbulence model");
    }
    va_end(ap);
}

// <- src/ibm3d_setup.c:367-439
void IBMSolver_set_bc(IBMSolver *s, IBMSolverDirection dir, IBMSolverBCType type, ...) {
    HostBC v;
    va_list ap;
    va_start(ap, type);
    switch (type) {
    case BC_VELOCITY_COMPONENT:
        v.val_type = (IBMSolverValType)va_arg(ap, int);
        if (v.val_type == VAL_CONST) for (int l = 0; l < 3; l++) v.cu[l] = va_arg(ap, double);
        else for (int l = 0; l < 3; l++) v.fu[l] = va_arg(ap, IBMSolverBCValFunc);
        break;
    case BC_PRESSURE:
    case BC_VELOCITY_PERIODIC:
        v.val_type = (IBMSolverValType)va_arg(ap, int);
        if (v.val_type == VAL_CONST) v.cp = va_arg(ap, double);
        else v.fp = va_arg(ap, IBMSolverBCValFunc);
        for (int l = 0; l < 3; l++) v.cu[l] = NAN;
        break;
    case BC_STATIONARY_WALL:
        type = BC_VELOCITY_COMPONENT;
        v.val_type = VAL_CONST;
        break;
    case BC_SYMMETRIC:
        type = BC_FREE_SLIP_WALL;
        break;
    case BC_CONVECTIVE_OUTFLOW:
        v.uc = va_arg(ap, double);
        v.val_type = VAL_CONST;
        break;
    default:;
    }
    va_end(ap);
    if (v.val_type != VAL_CONST && v.val_type != VAL_FUNC)
        fprintf(stderr, "error: invalid boundary condition value type\n");
    v.type = type;
    v.set = true;
    for (int i = 0; i < 6; i++)
        if (dir & (1 << i)) s->bc[i] = v;
}

void IBMSolver_set_obstacle(IBMSolver *s, Polyhedron *poly) {
    s->poly = poly;
    s->has_sphere = 0;
}

void IBMSolver_set_obstacle_sphere(IBMSolver *s, double cx, double cy, double cz, double r) {
    s->poly = nullptr;
    s->has_sphere = 1;
    s->sphere[0] = cx; s->sphere[1] = cy; s->sphere[2] = cz; s->sphere[3] = r;
}

void IBMSolver_set_linear_solver(IBMSolver *s, IBMSolverLinearSolverType t, IBMSolverPrecondType pt,
                                 const double tol) {
    s->solver_type = t;
    s->precond_type = pt;
    s->tol = tol;
}

void IBMSolver_set_autosave(IBMSolver *s, const char *filename, int period) {
    s->autosave_filename = filename;
    s->autosave_period = period;
}

void IBMSolver_set_precision(IBMSolver *s, int mode) {
    if (s->assembled) fatal("IBMSolver_set_precision must precede IBMSolver_assemble");
    if (mode != IBM3D_PREC_FP64 && mode != IBM3D_PREC_MIXED && mode != IBM3D_PREC_MIXED_KRYLOV)
        fatal("unknown precision mode");
    s->precision = mode;
}

void IBMSolver_set_mg_params(IBMSolver *s, double gamma, int pre, int post, int smoother) {
    if (s->assembled) fatal("IBMSolver_set_mg_params must precede IBMSolver_assemble");
    if (pre < 0 || post < 0 || pre + post == 0 || pre > 16 || post > 16) fatal("invalid multigrid sweep counts");
    if (smoother != 0 && smoother != 1) fatal("multigrid smoother must be 0 or 1");
    if (!std::isfinite(gamma)) fatal("multigrid coarsening threshold gamma must be finite");
    s->mg_gamma = gamma;
    s->mg_pre = pre;
    s->mg_post = post;
    s->mg_smoother = smoother;
}

void IBMSolver_set_mg_relaxation(IBMSolver *s, double omega) {
    if (s->assembled) fatal("IBMSolver_set_mg_relaxation must precede IBMSolver_assemble");
    if (!(omega > 0 && omega < 2)) fatal("multigrid over-relaxation factor must lie in (0, 2)");
    s->mg_omega = omega;
}

int IBMSolver_get_pressure_mg_bits(IBMSolver *s) {
    if (!s->assembled) return 0;
    return s->precision == IBM3D_PREC_MIXED_KRYLOV && !s->p_mg64 ? 32 : 64;
}

void IBMSolver_set_mg_light_levels(IBMSolver *s, int from_level) {
    if (s->assembled) fatal("IBMSolver_set_mg_light_levels must precede IBMSolver_assemble");
    if (from_level < 0) fatal("multigrid light levels: from_level must be 0 or positive");
    s->mg_light_from = from_level;
}

// Mirror point and trilinear weights  <- src/utils.c:111-171 (base indices
// clamped to the padded grid, where the reference would read out of bounds)
void IBMSolver_ghost_interp(IBMSolver *s, const int i, const int j, const int k, int idx[][3], double w[]) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    auto L = [&](int a, int b, int c) { return s->lvset_h[host_ref_index(s, a, b, c)]; };
    const double nx = (L(i + 1, j, k) - L(i - 1, j, k)) / (H1(s->xc, i + 1) - H1(s->xc, i - 1));
    const double ny = (L(i, j + 1, k) - L(i, j - 1, k)) / (H1(s->yc, j + 1) - H1(s->yc, j - 1));
    const double nz = (L(i, j, k + 1) - L(i, j, k - 1)) / (H1(s->zc, k + 1) - H1(s->zc, k - 1));
    const double phi = L(i, j, k);
    const double mx = 1 * H1(s->xc, i) + -2 * phi * nx;
    const double my = 1 * H1(s->yc, j) + -2 * phi * ny;
    const double mz = 1 * H1(s->zc, k) + -2 * phi * nz;
    auto ub = [](const std::vector<double> &a, double v) {
        return (int)(std::upper_bound(a.begin(), a.end(), v) - a.begin());
    };
    const int im = std::min(std::max(ub(s->xc, mx) - 3, -2), Nx);
    const int jm = std::min(std::max(ub(s->yc, my) - 3, -2), Ny);
    const int km = std::min(std::max(ub(s->zc, mz) - 3, -2), Nz);
    for (int l = 0; l < 8; l++) {
        idx[l][0] = im + !!(l & 4);
        idx[l][1] = jm + !!(l & 2);
        idx[l][2] = km + !!(l & 1);
    }
    const double xl = H1(s->xc, im), xu = H1(s->xc, im + 1);
    const double yl = H1(s->yc, jm), yu = H1(s->yc, jm + 1);
    const double zl = H1(s->zc, km), zu = H1(s->zc, km + 1);
    const double vol = (xu - xl) * (yu - yl) * (zu - zl);
    w[0] = (xu - mx) * (yu - my) * (zu - mz) / vol;
    w[1] = (xu - mx) * (yu - my) * (mz - zl) / vol;
    w[2] = (xu - mx) * (my - yl) * (zu - mz) / vol;
    w[3] = (xu - mx) * (my - yl) * (mz - zl) / vol;
    w[4] = (mx - xl) * (yu - my) * (zu - mz) / vol;
    w[5] = (mx - xl) * (yu - my) * (mz - zl) / vol;
    w[6] = (mx - xl) * (my - yl) * (zu - mz) / vol;
    w[7] = (mx - xl) * (my - yl) * (mz - zl) / vol;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
namespace ibm {

static void calc_geometry(IBMSolver *s) {
    std::vector<double> *d[3] = {&s->dx, &s->dy, &s->dz};
    std::vector<double> *c[3] = {&s->xc, &s->yc, &s->zc};
    const int n[3] = {s->Nx, s->Ny, s->Nz};
    const int lo[3] = {3, 2, 4}, hi[3] = {1, 0, 5};
    // ghost widths and centres  <- src/ibm3d_setup.c:536-588
    for (int a = 0; a < 3; a++) {
        std::vector<double> &D = *d[a], &C = *c[a];
        const int N = n[a];
        const bool plo = isperiodic(s->bc[lo[a]].type), phi = isperiodic(s->bc[hi[a]].type);
        H1(D, -1) = plo ? H1(D, N - 1) : H1(D, 0);
        H1(D, -2) = plo ? H1(D, N - 2) : H1(D, 1);
        H1(D, N) = phi ? H1(D, 0) : H1(D, N - 1);
        H1(D, N + 1) = phi ? H1(D, 1) : H1(D, N - 2);
        H1(C, -1) = H1(C, 0) - (H1(D, -1) + H1(D, 0)) / 2;
        H1(C, -2) = H1(C, -1) - (H1(D, -2) + H1(D, -1)) / 2;
        H1(C, N) = H1(C, N - 1) + (H1(D, N - 1) + H1(D, N)) / 2;
        H1(C, N + 1) = H1(C, N) + (H1(D, N) + H1(D, N + 1)) / 2;
    }
    // second-derivative coefficients  <- src/ibm3d_setup.c:620-632
    auto kc = [&](std::vector<double> &km, std::vector<double> &kp, const std::vector<double> &D,
                  const std::vector<double> &C, int N) {
        km.assign(N + 4, 0);
        kp.assign(N + 4, 0);
        for (int i = 0; i < N; i++) {
            H1(km, i) = s->dt / (2 * s->Re * (H1(C, i) - H1(C, i - 1)) * H1(D, i));
            H1(kp, i) = s->dt / (2 * s->Re * (H1(C, i + 1) - H1(C, i)) * H1(D, i));
        }
    };
    kc(s->kW, s->kE, s->dx, s->xc, s->Nx);
    kc(s->kS, s->kN, s->dy, s->yc, s->Ny);
    kc(s->kD, s->kU, s->dz, s->zc, s->Nz);
}

// level set and flags  <- src/ibm3d_setup.c:1081-1254
static void calc_lvset_flag(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const long P = (long)(Nx + 4) * (Ny + 4) * (Nz + 4);
    s->lvset_h.assign(P, 0.5);
    s->flag_ref.assign(P, FLAG_FLUID);
    if (!s->poly && !s->has_sphere) return;
    if (s->poly) {
        polyhedron_cpt_grid(s->poly, Nx + 4, Ny + 4, Nz + 4, s->xc.data(), s->yc.data(), s->zc.data(),
                            s->lvset_h.data(), 15);
    } else {
        const double *c = s->sphere;
        for (int i = -2; i < Nx + 2; i++)
            for (int j = -2; j < Ny + 2; j++)
                for (int k = -2; k < Nz + 2; k++) {
                    const double dx = H1(s->xc, i) - c[0], dy = H1(s->yc, j) - c[1], dz = H1(s->zc, k) - c[2];
                    s->lvset_h[host_ref_index(s, i, j, k)] = std::sqrt(dx * dx + dy * dy + dz * dz) - c[3];
                }
    }
    static const int adj[6][3] = {{0, 1, 0}, {1, 0, 0}, {0, -1, 0}, {-1, 0, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int i = -2; i < Nx + 2; i++)
        for (int j = -2; j < Ny + 2; j++)
            for (int k = -2; k < Nz + 2; k++) {
                const long q = host_ref_index(s, i, j, k);
                if (s->lvset_h[q] > 0) { s->flag_ref[q] = FLAG_FLUID; continue; }
                bool ghost = false;
                for (int l = 0; l < 6; l++) {
                    const int a = i + adj[l][0], b = j + adj[l][1], c = k + adj[l][2];
                    if (a < -2 || a >= Nx + 2 || b < -2 || b >= Ny + 2 || c < -2 || c >= Nz + 2) continue;
                    ghost = ghost || s->lvset_h[host_ref_index(s, a, b, c)] > 0;
                }
                s->flag_ref[q] = ghost ? FLAG_GHOST : FLAG_SOLID;
            }
}

// Ghost stencils: weights exactly as update_ghost / create_matrix form them
// (src/ibm3d_fracstep.c:1937-1956, src/ibm3d_setup.c:1656-1698), dependency
// levels for the ordered in-place update.
static void build_ghosts(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const Dims &d = s->d;
    std::vector<GhostCell> lex;
    std::vector<long> slot_of(s->flag_ref.size(), -1);
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++)
                if (s->flag_ref[host_ref_index(s, i, j, k)] == FLAG_GHOST) {
                    slot_of[host_ref_index(s, i, j, k)] = (long)lex.size();
                    GhostCell g{};
                    g.cell = cidx(d, i, j, k);
                    lex.push_back(g);
                }
    const int ng = (int)lex.size();
    {
        std::vector<int> ijk;
        ijk.reserve(3 * (size_t)ng + 3);
        for (int i = 0; i < Nx; i++)
            for (int j = 0; j < Ny; j++)
                for (int k = 0; k < Nz; k++)
                    if (s->flag_ref[host_ref_index(s, i, j, k)] == FLAG_GHOST) ijk.insert(ijk.end(), {i, j, k});
        if (ijk.empty()) ijk.assign(3, 0);
        s->ghost_ijk.upload(ijk);
    }
    int gi = 0;
    std::vector<int> level(ng, 0);
    bool touches_outer = false;
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++) {
                if (s->flag_ref[host_ref_index(s, i, j, k)] != FLAG_GHOST) continue;
                GhostCell &g = lex[gi];
                int idx[8][3];
                double w[8];
                IBMSolver_ghost_interp(s, i, j, k, idx, w);
                double sum = 0;
                for (int l = 0; l < 8; l++) {
                    if (s->flag_ref[host_ref_index(s, idx[l][0], idx[l][1], idx[l][2])] == FLAG_SOLID) w[l] = 0;
                    else sum += w[l];
                }
                g.wself = 0;
                int lev = 0;
                for (int l = 0; l < 8; l++) {
                    w[l] /= sum;
                    const int a = idx[l][0], b = idx[l][1], c = idx[l][2];
                    g.src[l] = -1;
                    g.w[l] = 0;
                    g.later[l] = -1;
                    if (a == i && b == j && c == k) {
                        g.wself = w[l];
                    } else if (s->flag_ref[host_ref_index(s, a, b, c)] != FLAG_SOLID) {
                        if (a < 0 || a >= Nx || b < 0 || b >= Ny || c < 0 || c >= Nz) touches_outer = true;
                        g.src[l] = cidx(d, a, b, c);
                        g.w[l] = w[l];
                        const long hs = slot_of[host_ref_index(s, a, b, c)];
                        if (hs >= 0) {
                            if (hs > gi) g.later[l] = (int)hs;           // not yet visited: old value
                            else lev = std::max(lev, level[hs] + 1);     // visited: new value
                        }
                    }
                }
                level[gi] = lev;
                g.level = lev;
                gi++;
            }
    if (touches_outer)
        fatal("an immersed-boundary ghost stencil reaches the outer ghost layers (body touching the domain "
              "boundary); not supported");
    // stable sort by level; `later` keeps lexicographic slots into the saved-value buffer
    std::vector<int> order(ng);
    for (int q = 0; q < ng; q++) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return level[a] < level[b]; });
    s->ghosts.resize(ng);
    s->ghost_lex.assign(ng, 0);
    int nlev = 0;
    for (int q = 0; q < ng; q++) nlev = std::max(nlev, level[q] + 1);
    s->level_start.assign(nlev + 1, 0);
    // saved values are indexed by level-sorted slot: remap `later`
    std::vector<int> pos(ng);
    for (int q = 0; q < ng; q++) pos[order[q]] = q;
    for (int q = 0; q < ng; q++) {
        GhostCell g = lex[order[q]];
        for (int l = 0; l < 8; l++)
            if (g.later[l] >= 0) g.later[l] = pos[g.later[l]];
        s->ghosts[q] = g;
        s->ghost_lex[order[q]] = q;
        s->level_start[g.level + 1]++;
    }
    for (int l = 0; l < nlev; l++) s->level_start[l + 1] += s->level_start[l];
    s->ng = ng;
    s->gc.upload(s->ghosts);
    s->ghost_old.alloc(4 * (size_t)std::max(ng, 1));

    // solver ghost rows (same stencils for every system)
    std::vector<long> cell(ng), src(8 * (size_t)ng);
    std::vector<int> srcg(8 * (size_t)ng, -1);
    std::vector<double> w64(8 * (size_t)ng), ws64(ng);
    std::unordered_map<long, int> slot;  // ghost cell -> ghost row
    for (int q = 0; q < ng; q++) slot[s->ghosts[q].cell] = q;
    for (int q = 0; q < ng; q++) {
        cell[q] = s->ghosts[q].cell;
        ws64[q] = s->ghosts[q].wself;
        for (int l = 0; l < 8; l++) {
            src[8 * q + l] = s->ghosts[q].src[l];
            w64[8 * q + l] = s->ghosts[q].w[l];
            const auto it = slot.find(s->ghosts[q].src[l]);
            if (s->ghosts[q].src[l] >= 0 && it != slot.end()) srcg[8 * q + l] = it->second;
        }
    }
    std::vector<float> w32(w64.begin(), w64.end()), ws32(ws64.begin(), ws64.end());
    for (GhostData *G : {&s->gmom, &s->gp}) {
        G->cell.upload(cell); G->src.upload(src); G->srcg.upload(srcg);
        G->w64.upload(w64); G->ws64.upload(ws64);
        G->w32.upload(w32); G->ws32.upload(ws32);
        for (int k = 0; k < 2; k++) {
            G->v64[k].alloc(std::max(ng, 1));
            G->v32[k].alloc(std::max(ng, 1));
        }
        G->G64 = {ng, G->cell.p, G->src.p, G->w64.p, G->ws64.p, G->srcg.p, {G->v64[0].p, G->v64[1].p},
                  G->w64.p, G->ws64.p};
        G->G32 = {ng, G->cell.p, G->src.p, G->w32.p, G->ws32.p, G->srcg.p, {G->v32[0].p, G->v32[1].p},
                  G->w64.p, G->ws64.p};
    }
}

// Folded per-axis coefficients of system sk (0..2 momentum component, 3
// pressure) from the BC rows of src/ibm3d_setup.c:1723-2903.
void build_system(const IBMSolver *s, const std::vector<double> *D[3], const std::vector<double> *C[3],
                  const std::vector<double> *Km[3], const std::vector<double> *Kp[3], const int n[3],
                  const IBMSolverBCType bct[3][2], int sk, SysData &out) {
    for (int a = 0; a < 3; a++) {
        const int N = n[a];
        std::vector<double> cm(N), cp(N), dm(N), dp(N);
        for (int i = 0; i < N; i++) {
            cm[i] = dm[i] = H1((*Km[a]), i);
            cp[i] = dp[i] = H1((*Kp[a]), i);
        }
        const bool per_lo = sk < 3 ? isperiodic(bct[a][0]) : bct[a][0] == BC_ALL_PERIODIC;
        const bool per_hi = sk < 3 ? isperiodic(bct[a][1]) : bct[a][1] == BC_ALL_PERIODIC;
        const int periodic = per_lo && per_hi;
        for (int hi = 0; hi < 2; hi++) {
            out.gamma[a][hi] = 0;
            if (periodic) continue;
            const IBMSolverBCType t = bct[a][hi];
            const std::vector<double> &dd = *D[a], &cc = *C[a];
            const int o1 = hi ? N : -1, q0 = hi ? N - 1 : 0, q1 = hi ? N - 2 : 1;
            const double ao = H1(dd, o1) / 2, bi = H1(dd, q0) / 2;
            double alpha = 1, beta = 0, gamma = 0;
            const bool dirichlet_u = t == BC_VELOCITY_COMPONENT || t == BC_STATIONARY_WALL
                                  || t == BC_CONVECTIVE_OUTFLOW;
            if (sk < 3) {
                if (dirichlet_u) { alpha = -ao / bi; gamma = (ao + bi) / bi; }
                else if (t == BC_PRESSURE) {
                    const double so = hi ? H1(cc, o1) - H1(cc, q0) : H1(cc, q0) - H1(cc, o1);
                    const double si = hi ? H1(cc, q0) - H1(cc, q1) : H1(cc, q1) - H1(cc, q0);
                    alpha = (so + si) / si;
                    beta = -so / si;
                } else if (t == BC_FREE_SLIP_WALL) {
                    alpha = (sk == a) ? -ao / bi : 1;
                }
            } else {
                if (t == BC_PRESSURE || t == BC_VELOCITY_PERIODIC || t == BC_CONVECTIVE_OUTFLOW) {
                    alpha = -ao / bi;
                    gamma = (ao + bi) / bi;
                }
                else alpha = 1;  // wall, velocity, free slip: Neumann (1, -1) rows
            }
            out.gamma[a][hi] = gamma;
            if (!hi) {
                const double k0 = H1((*Km[a]), 0);
                dm[0] = k0 * (1 - alpha);
                cm[0] = 0;
                cp[0] += k0 * beta;
            } else {
                const double k1 = H1((*Kp[a]), N - 1);
                dp[N - 1] = k1 * (1 - alpha);
                cp[N - 1] = 0;
                cm[N - 1] += k1 * beta;
            }
        }
        const std::vector<double> *arr[4] = {&cm, &cp, &dm, &dp};
        for (int c = 0; c < 4; c++) {
            out.h[a][c] = *arr[c];
            out.c64[a][c].upload(*arr[c]);
            std::vector<float> f(arr[c]->begin(), arr[c]->end());
            out.c32[a][c].upload(f);
        }
        {
            std::vector<double> pk(4 * (size_t)N);
            for (int i = 0; i < N; i++)
                for (int c = 0; c < 4; c++) pk[4 * (size_t)i + c] = (*arr[c])[i];
            out.p64[a].upload(pk);
            std::vector<float> pf(pk.begin(), pk.end());
            out.p32[a].upload(pf);
        }
        out.S64.ax[a] = {out.c64[a][0].p, out.c64[a][1].p, out.c64[a][2].p, out.c64[a][3].p, N, periodic,
                         out.p64[a].p};
        out.S32.ax[a] = {out.c32[a][0].p, out.c32[a][1].p, out.c32[a][2].p, out.c32[a][3].p, N, periodic,
                         out.p32[a].p};
    }
    out.S64.c0 = sk < 3 ? 1.0 : 0.0;
    out.S32.c0 = sk < 3 ? 1.0f : 0.0f;
    out.S64.pin = out.S32.pin = -1;
    out.S64.is_p = out.S32.is_p = sk == 3;
    out.S64.skip = out.S32.skip = 0;
}

// squared norm of the reference's BC rows' right-hand side for constant data
static void bc_rows_norm(IBMSolver *s) {
    const int n1[3] = {s->Ny, s->Nx + 4, s->Nx + 4}, n2[3] = {s->Nz, s->Nz, s->Ny + 4};
    for (int sk = 0; sk < 4; sk++) {
        double acc = 0;
        for (int a = 0; a < 3; a++)
            for (int hi = 0; hi < 2; hi++) {
                const HostBC &b = s->bc[face_bc(a, hi)];
                const double cnt = 2.0 * n1[a] * n2[a];
                if (b.val_type != VAL_CONST) continue;  // VAL_FUNC faces are added per step
                if (sk < 3 && b.type == BC_VELOCITY_COMPONENT) acc += cnt * b.cu[sk] * b.cu[sk];
                // pressure rows carry p_bc(t + dt) - p_bc(t) = 0 for constant data
            }
        s->sys[sk].bc_norm2 = acc;
    }
}

static void alloc_fields(IBMSolver *s) {
    const size_t n = s->d.ncell;
    for (int l = 0; l < 3; l++) {
        s->u[l].alloc(n); s->un[l].alloc(n); s->us[l].alloc(n);
        s->F[l].alloc(n); s->Fn[l].alloc(n); s->Fs[l].alloc(n);
        s->N[l].alloc(n); s->Np[l].alloc(n); s->b[l].alloc(n);
    }
    s->p.alloc(n); s->pn.alloc(n); s->pp.alloc(n); s->bp.alloc(n);
    if (s->turb_type != TURBMODEL_NONE)
        for (int a = 0; a < 6; a++) s->tau[a].alloc(n);
    for (int v = 0; v < 8; v++) s->kry.v64[v].alloc(n);
    if (s->precision != IBM3D_PREC_MIXED) {
        // lanes of the concurrent momentum solves (components 1 and 2)
        HipGlobalLock lk(hip_global_mutex());
        const long np = std::max({conv_rhs_blocks(s->d), apply_partials(s->d, s->ng), vec_partials(s->d)}) + 64;
        for (auto &L : s->mlane) {
            for (int v = 0; v < 8; v++) L.v[v].alloc(n);
            if (s->precision == IBM3D_PREC_MIXED_KRYLOV)
                for (int v = 0; v < 2; v++) L.d32[v].alloc(n);
            L.ks.alloc(1);
            L.partial.alloc((size_t)np * 4);
            HIP_CHECK(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
            HIP_CHECK(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
            HIP_CHECK(hipHostMalloc((void **)&L.pub_h, sizeof(KPub), hipHostMallocCoherent | hipHostMallocMapped));
            memset((void *)L.pub_h, 0, sizeof(KPub));
            HIP_CHECK(hipHostGetDevicePointer((void **)&L.pub_d, L.pub_h, 0));
        }
        HIP_CHECK(hipEventCreateWithFlags(&s->ev_rhs, hipEventDisableTiming));
        s->mom_lanes = true;
    }
    if (s->precision != IBM3D_PREC_FP64) {
        if (s->precision == IBM3D_PREC_MIXED)
            for (int v = 0; v < 8; v++) s->kry.v32[v].alloc(n);
        s->ref_r64.alloc(n);
        s->ref_rhs32.alloc(n);
        s->ref_e32.alloc(n);
        if (s->precision == IBM3D_PREC_MIXED_KRYLOV) s->ref_e32b.alloc(n);
    }
    if (s->solver_type != SOLVER_BiCGSTAB) {
        // pressure solver scratch: GMRES basis, or the PCG / stationary
        // multigrid vectors (Krylov vectors above) and scalars
        GmresData &G = s->gm;
        if (s->solver_type == SOLVER_GMRES) {
            for (int v = 0; v <= kGmresRestart; v++) G.V[v].alloc(n);
            for (int v = 0; v < kGmresRestart; v++) G.Z[v].alloc(n);
            G.t.alloc(n);
        }
        G.dev.alloc(64);
        G.partial.alloc((size_t)krylov_blocks(s->d) * (kVecSetMax + 1));
    }
    long nb = conv_rhs_blocks(s->d) + 64;
    nb = std::max(nb, apply_partials(s->d, s->ng) + 64);
    nb = std::max(nb, vec_partials(s->d) + 64);
    s->partial.alloc((size_t)nb * 4);
}

}  // namespace ibm

// Stream, mapped host slots for Krylov scalars and step sums, and the device
// scalars of one solver.
static void init_device(IBMSolver *s) {
    HipGlobalLock lk(hip_global_mutex());
    HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIP_CHECK(hipHostMalloc((void **)&s->ks_h, sizeof(KScal), hipHostMallocDefault));
    HIP_CHECK(hipHostMalloc((void **)&s->pub_h, sizeof(KPub), hipHostMallocCoherent | hipHostMallocMapped));
    memset((void *)s->pub_h, 0, sizeof(KPub));
    HIP_CHECK(hipHostGetDevicePointer((void **)&s->pub_d, s->pub_h, 0));
    HIP_CHECK(hipHostMalloc((void **)&s->sums_h, (kSumSlots + 1) * sizeof(double),
                            hipHostMallocCoherent | hipHostMallocMapped));
    memset((void *)s->sums_h, 0, (kSumSlots + 1) * sizeof(double));
    HIP_CHECK(hipHostGetDevicePointer((void **)&s->sums_d, s->sums_h, 0));
    s->sums_h_epoch = reinterpret_cast<unsigned *>(s->sums_h + kSumSlots);
    s->sums_d_epoch = reinterpret_cast<unsigned *>(s->sums_d + kSumSlots);
    HIP_CHECK(hipHostMalloc((void **)&s->gm_h, 64 * sizeof(double), hipHostMallocDefault));
    s->ks.alloc(1);
    s->sums.alloc(kSumSlots);
}

extern "C" void IBMSolver_assemble(IBMSolver *s) {
    if (s->Nx == 0) fatal("IBMSolver_set_grid must precede IBMSolver_assemble");
    if (s->assembled) fatal("IBMSolver_assemble called twice");
    // <- src/ibm3d_setup.c:523-533
    if (isperiodic(s->bc[0].type) != isperiodic(s->bc[2].type) || isperiodic(s->bc[1].type) != isperiodic(s->bc[3].type)
        || isperiodic(s->bc[4].type) != isperiodic(s->bc[5].type)) {
        case_error("Inconsistent periodic boundary condition\n", true);
    }
    // pressure solver and preconditioner types  <- src/ibm3d_setup.c:1435-1536
    if (s->solver_type != SOLVER_AMG && s->solver_type != SOLVER_PCG && s->solver_type != SOLVER_BiCGSTAB &&
        s->solver_type != SOLVER_GMRES) {
        case_error("\nUnknown linear solver type\n", true);
    }
    if (s->precond_type != PRECOND_NONE && s->precond_type != PRECOND_AMG) {
        case_error("\nUnknown preconditioner type\n", true);
    }
    if (s->solver_type == SOLVER_AMG && s->precond_type != PRECOND_NONE) {
        case_error("\nCannot use preconditioner with BoomerAMG solver\n", true);
    }
    // IBM3D_PREC_MIXED runs BiCGSTAB defect correction for every system
    if (s->precision == IBM3D_PREC_MIXED && s->solver_type != SOLVER_BiCGSTAB)
        fatal("IBM3D_PREC_MIXED supports only SOLVER_BiCGSTAB for the pressure (use IBM3D_PREC_MIXED_KRYLOV "
              "or IBM3D_PREC_FP64 for GMRES, PCG and AMG)");
    // Red-black colouring of the multigrid smoother: on a periodic axis with
    // an odd cell count the two cells beside the wrap share a colour, and a
    // colour launch would read a neighbour it is writing.
    if (s->precond_type == PRECOND_AMG || s->solver_type == SOLVER_AMG) {
        const int n[3] = {s->Nx, s->Ny, s->Nz};
        const char *ax = "xyz";
        for (int a = 0; a < 3; a++)
            if (isperiodic(s->bc[face_bc(a, 0)].type) && n[a] % 2) {
                char buf[200];
                snprintf(buf, sizeof buf, "error: periodic %c axis with an odd cell count (%d): the multigrid "
                         "smoother's red-black colouring needs an even count (or use PRECOND_NONE)", ax[a], n[a]);
                case_error(buf);
            }
    }
    s->p_mg64 = false;
    s->p_singular = true;  // <- src/ibm3d_setup.c:1598-1603
    for (int i = 0; i < 6; i++)
        if (s->bc[i].type == BC_PRESSURE || s->bc[i].type == BC_VELOCITY_PERIODIC
            || s->bc[i].type == BC_CONVECTIVE_OUTFLOW)
            s->p_singular = false;
    // The inner FP32 solves of IBM3D_PREC_MIXED stall on a pinned (otherwise
    // singular) pressure system, whose near-null mode they cannot resolve in
    // FP32; IBM3D_PREC_MIXED_KRYLOV restarts its FP64 iteration instead.
    if (s->precision == IBM3D_PREC_MIXED && s->p_singular)
        fatal("IBM3D_PREC_MIXED needs a pressure-type boundary (BC_PRESSURE, BC_VELOCITY_PERIODIC or "
              "BC_CONVECTIVE_OUTFLOW): without one the pressure system is pinned and its FP32 inner solves "
              "stall (use IBM3D_PREC_MIXED_KRYLOV or IBM3D_PREC_FP64)");
    DevBytesScope account(&s->dev_bytes);
    init_device(s);

    calc_geometry(s);
    calc_lvset_flag(s);
    s->d = make_dims(s->Nx, s->Ny, s->Nz);
    const Dims &d = s->d;

    // device 1-D geometry
    std::vector<double> *g[12] = {&s->dx, &s->dy, &s->dz, &s->xc, &s->yc, &s->zc,
                                  &s->kW, &s->kE, &s->kS, &s->kN, &s->kD, &s->kU};
    for (int a = 0; a < 12; a++) s->g1d[a].upload(*g[a]);
    s->geo = {d, s->g1d[0].p, s->g1d[1].p, s->g1d[2].p, s->g1d[3].p, s->g1d[4].p, s->g1d[5].p,
              s->g1d[6].p, s->g1d[7].p, s->g1d[8].p, s->g1d[9].p, s->g1d[10].p, s->g1d[11].p};
    {
        std::vector<double> kx(s->Nx), ky(s->Ny), kz(s->Nz);
        for (int i = 0; i < s->Nx; i++) kx[i] = H1(s->kW, i) + H1(s->kE, i);
        for (int j = 0; j < s->Ny; j++) ky[j] = H1(s->kS, j) + H1(s->kN, j);
        for (int k = 0; k < s->Nz; k++) kz[k] = H1(s->kD, k) + H1(s->kU, k);
        s->ks1d[0].upload(kx); s->ks1d[1].upload(ky); s->ks1d[2].upload(kz);
    }

    // flags and level set in device layout (all padded cells)
    s->flag_dev_h.assign(d.ncell, F_FLUID);
    std::vector<double> lv(d.ncell, 0.5);
    for (int i = -2; i < s->Nx + 2; i++)
        for (int j = -2; j < s->Ny + 2; j++)
            for (int k = -2; k < s->Nz + 2; k++) {
                const long r = host_ref_index(s, i, j, k), q = cidx(d, i, j, k);
                s->flag_dev_h[q] = (uint8_t)s->flag_ref[r];
                lv[q] = s->lvset_h[r];
            }
    s->flag.upload(s->flag_dev_h);
    s->lvset.upload(lv);

    build_ghosts(s);

    // folded systems
    const std::vector<double> *D[3] = {&s->dx, &s->dy, &s->dz}, *C[3] = {&s->xc, &s->yc, &s->zc};
    const std::vector<double> *Km[3] = {&s->kW, &s->kS, &s->kD}, *Kp[3] = {&s->kE, &s->kN, &s->kU};
    const int n[3] = {s->Nx, s->Ny, s->Nz};
    IBMSolverBCType bct[3][2];
    for (int a = 0; a < 3; a++)
        for (int hi = 0; hi < 2; hi++) bct[a][hi] = s->bc[face_bc(a, hi)].type;
    for (int sk = 0; sk < 4; sk++) build_system(s, D, C, Km, Kp, n, bct, sk, s->sys[sk]);
    s->pin = s->p_singular ? cidx(d, 0, 0, 0) : -1;
    s->sys[3].S64.pin = s->sys[3].S32.pin = s->pin;
    bc_rows_norm(s);

    alloc_fields(s);
    {
        const long ext[3] = {(long)s->Ny * s->Nz, (long)(s->Nx + 4) * s->Nz, (long)(s->Nx + 4) * (s->Ny + 4)};
        for (int a = 0; a < 3; a++)
            for (int hi = 0; hi < 2; hi++)
                if (s->bc[face_bc(a, hi)].type == BC_CONVECTIVE_OUTFLOW) s->conv_g[face_bc(a, hi)].alloc(4 * ext[a]);
        s->conv_ready = false;
    }
    build_mg(s);
    s->assembled = true;
    HIP_CHECK(hipStreamSynchronize(s->stream));
}

// ---------------------------------------------------------------------------
// initialisation  <- src/ibm3d_init.c:15-37, 237-262
// ---------------------------------------------------------------------------
static void upload_cell_field(IBMSolver *s, DBuf<double> &dst, const double *ref) {
    std::vector<double> h(s->d.ncell, 0.0);
    for (int i = -2; i < s->Nx + 2; i++)
        for (int j = -2; j < s->Ny + 2; j++)
            for (int k = -2; k < s->Nz + 2; k++) h[cidx(s->d, i, j, k)] = ref[host_ref_index(s, i, j, k)];
    dst.upload(h);
}

extern "C" void IBMSolver_init_flow_const(IBMSolver *s, const double u1, const double u2, const double u3,
                                          const double p) {
    if (!s->assembled) fatal("IBMSolver_assemble must precede IBMSolver_init_flow_const");
    s->iter = 0;
    s->time = 0;
    s->conv_ready = false;
    const double v[4] = {u1, u2, u3, p};
    DBuf<double> *dst[4] = {&s->u[0], &s->u[1], &s->u[2], &s->p};
    std::vector<double> h(s->d.ncell);
    for (int a = 0; a < 4; a++) {
        std::fill(h.begin(), h.end(), v[a]);
        dst[a]->upload(h);
    }
}

extern "C" void IBMSolver_init_flow_func(IBMSolver *s, IBMSolverInitFunc f1, IBMSolverInitFunc f2,
                                         IBMSolverInitFunc f3, IBMSolverInitFunc fp) {
    if (!s->assembled) fatal("IBMSolver_assemble must precede IBMSolver_init_flow_func");
    s->conv_ready = false;
    IBMSolverInitFunc f[4] = {f1, f2, f3, fp};
    DBuf<double> *dst[4] = {&s->u[0], &s->u[1], &s->u[2], &s->p};
    const long P = (long)(s->Nx + 4) * (s->Ny + 4) * (s->Nz + 4);
    std::vector<double> ref(P);
    for (int a = 0; a < 4; a++) {
        for (int i = -2; i < s->Nx + 2; i++)
            for (int j = -2; j < s->Ny + 2; j++)
                for (int k = -2; k < s->Nz + 2; k++)
                    ref[host_ref_index(s, i, j, k)] = f[a](H1(s->xc, i), H1(s->yc, j), H1(s->zc, k));
        upload_cell_field(s, *dst[a], ref.data());
    }
}

extern "C" void IBMSolver_init_flow_file(IBMSolver *s, const char *filename) { init_flow_file_impl(s, filename); }

// ---------------------------------------------------------------------------
// field access (extensions)
// ---------------------------------------------------------------------------
static DBuf<double> *cell_field(IBMSolver *s, int field) {
    switch (field) {
    case IBM3D_FIELD_U1: return &s->u[0];
    case IBM3D_FIELD_U2: return &s->u[1];
    case IBM3D_FIELD_U3: return &s->u[2];
    case IBM3D_FIELD_P: return &s->p;
    case IBM3D_FIELD_U1_STAR: return &s->us[0];
    case IBM3D_FIELD_U2_STAR: return &s->us[1];
    case IBM3D_FIELD_U3_STAR: return &s->us[2];
    case IBM3D_FIELD_P_PRIME: return &s->pp;
    case IBM3D_FIELD_N1: return &s->Np[0];
    case IBM3D_FIELD_N2: return &s->Np[1];
    case IBM3D_FIELD_N3: return &s->Np[2];
    case IBM3D_FIELD_LVSET: return &s->lvset;
    default:
        if (field >= IBM3D_FIELD_TAU11 && field <= IBM3D_FIELD_TAU33) return &s->tau[field - IBM3D_FIELD_TAU11];
        return nullptr;
    }
}

extern "C" void IBMSolver_get_field(IBMSolver *s, int field, double *out) {
    if (!s->assembled) fatal("IBMSolver_get_field before IBMSolver_assemble");
    HIP_CHECK(hipStreamSynchronize(s->stream));
    const Dims &d = s->d;
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    std::vector<double> h(d.ncell);
    if (field >= IBM3D_FIELD_FACE_U1 && field <= IBM3D_FIELD_FACE_U3) {
        const int a = field - IBM3D_FIELD_FACE_U1;
        HIP_CHECK(hipMemcpyAsync(h.data(), s->F[a].p, d.ncell * sizeof(double), hipMemcpyDeviceToHost, s->stream));
        HIP_CHECK(hipStreamSynchronize(s->stream));
        if (a == 0) {
            for (int i = 0; i <= Nx; i++)
                for (int j = -2; j < Ny + 2; j++)
                    for (int k = -2; k < Nz + 2; k++)
                        out[((long)i * (Ny + 4) + j + 2) * (Nz + 4) + k + 2] = h[cidx(d, i, j, k)];
        } else if (a == 1) {
            for (int i = -2; i < Nx + 2; i++)
                for (int j = 0; j <= Ny; j++)
                    for (int k = -2; k < Nz + 2; k++)
                        out[((long)(i + 2) * (Ny + 1) + j) * (Nz + 4) + k + 2] = h[cidx(d, i, j, k)];
        } else {
            for (int i = -2; i < Nx + 2; i++)
                for (int j = -2; j < Ny + 2; j++)
                    for (int k = 0; k <= Nz; k++)
                        out[((long)(i + 2) * (Ny + 4) + j + 2) * (Nz + 1) + k] = h[cidx(d, i, j, k)];
        }
        return;
    }
    DBuf<double> *src = cell_field(s, field);
    if (!src) fatal("unknown field");
    if (!src->p) {
        const long P = (long)(Nx + 4) * (Ny + 4) * (Nz + 4);
        for (long q = 0; q < P; q++) out[q] = 0;
        return;
    }
    HIP_CHECK(hipMemcpyAsync(h.data(), src->p, d.ncell * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    for (int i = -2; i < Nx + 2; i++)
        for (int j = -2; j < Ny + 2; j++)
            for (int k = -2; k < Nz + 2; k++) out[host_ref_index(s, i, j, k)] = h[cidx(d, i, j, k)];
}

extern "C" void IBMSolver_set_field(IBMSolver *s, int field, const double *in) {
    if (!s->assembled) fatal("IBMSolver_set_field before IBMSolver_assemble");
    if (field != IBM3D_FIELD_U1 && field != IBM3D_FIELD_U2 && field != IBM3D_FIELD_U3 && field != IBM3D_FIELD_P)
        fatal("set_field supports u1, u2, u3, p");
    HIP_CHECK(hipStreamSynchronize(s->stream));
    upload_cell_field(s, *cell_field(s, field), in);
    s->conv_ready = false;
}

extern "C" void IBMSolver_get_flag(IBMSolver *s, int *out) {
    std::copy(s->flag_ref.begin(), s->flag_ref.end(), out);
}

extern "C" int IBMSolver_get_iter(IBMSolver *s) { return s->iter; }
extern "C" double IBMSolver_get_time(IBMSolver *s) { return s->time; }
extern "C" void IBMSolver_get_last_norms(IBMSolver *s, double out[4]) { std::copy(s->final_norm, s->final_norm + 4, out); }
extern "C" void IBMSolver_get_last_iters(IBMSolver *s, int out[4]) { std::copy(s->final_iter, s->final_iter + 4, out); }

extern "C" int IBMSolver_get_ghost_cells(IBMSolver *s, int *out) {
    int n = 0;
    for (int i = 0; i < s->Nx; i++)
        for (int j = 0; j < s->Ny; j++)
            for (int k = 0; k < s->Nz; k++)
                if (s->flag_ref[host_ref_index(s, i, j, k)] == FLAG_GHOST) {
                    if (out) { out[3 * n] = i; out[3 * n + 1] = j; out[3 * n + 2] = k; }
                    n++;
                }
    return n;
}

extern "C" int IBMSolver_get_mg_levels(IBMSolver *s, int *out, int max) {
    const int n = (int)s->mg.size();
    for (int l = 0; l < n && l < max && out; l++) {
        out[3 * l] = s->mg[l].d.nx;
        out[3 * l + 1] = s->mg[l].d.ny;
        out[3 * l + 2] = s->mg[l].d.nz;
    }
    return n;
}

extern "C" double IBMSolver_max_div(IBMSolver *s) { return max_div_impl(s); }

// Force on the immersed body (extension; the reference has no drag/lift):
// staircase surface of faces between a fluid and a ghost cell; pressure
// interpolated to the face, wall shear mu du_t/dn across the face, area of
// the face. Outward normal of the body points from the ghost into the fluid.
extern "C" void IBMSolver_calc_body_force(IBMSolver *s, double out[3]) {
    out[0] = out[1] = out[2] = 0;
    if (!s->assembled) return;
    body_force_read(s, out);
}

extern "C" void IBMSolver_iterate(IBMSolver *s, int steps, bool verbose) {
    if (!s->assembled) fatal("IBMSolver_assemble must precede IBMSolver_iterate");
    iterate_impl(s, steps, verbose);
}

extern "C" void IBMSolver_export_result(IBMSolver *s, const char *f) { export_result_impl(s, f); }
extern "C" void IBMSolver_export_lvset_flag(IBMSolver *s, const char *f) { export_lvset_flag_impl(s, f); }
extern "C" void IBMSolver_export_intermediate(IBMSolver *s, const char *f) { export_intermediate_impl(s, f); }
