// step.cpp - the time loop (IBMSolver_iterate) and the linear solves.
//
// Stage order is the reference's (src/ibm3d_fracstep.c:38-115). The Hypre
// BiCGSTAB / BoomerAMG solves (:354-379, :649-703) are replaced by
//   FP64 mode : BiCGSTAB with a Jacobi preconditioner (momentum) or a
//               geometric-multigrid V-cycle (pressure), FP64 throughout;
//   mixed mode: FP64 iterative refinement (defect correction on the full
//               operator, FP64 residuals and solution) around FP32 BiCGSTAB
//               inner solves preconditioned by Jacobi / an FP32 V-cycle.
// Both stop on the reference's criterion ||b - A x||_2 <= tol ||b||_2 of the
// full padded system (pressure rows diagonal-normalised), from x0 = 0.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <sched.h>

#include "solver.h"

#include <type_traits>

namespace ibm {

#define H1(a, i) ((a)[(i) + 2])

static bool isperiodic(IBMSolverBCType t) { return t == BC_ALL_PERIODIC || t == BC_VELOCITY_PERIODIC; }
static int face_bc(int ax, int hi) {
    static const int t[3][2] = {{3, 1}, {2, 0}, {4, 5}};
    return t[ax][hi];
}

// ---------------------------------------------------------------------------
// timing
// ---------------------------------------------------------------------------
static hipEvent_t get_event(IBMSolver *s) {
    if (!s->event_pool.empty()) {
        hipEvent_t e = s->event_pool.back();
        s->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HipGlobalLock lk(hip_global_mutex());
    // timing-only events: no system-scope fence when they complete, so the
    // interval around a kernel does not include a cache write-back
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
        (void)hipGetLastError();
        HIP_CHECK(hipEventCreate(&e));
    }
    return e;
}

int timer_begin(IBMSolver *s, const char *name, double bytes) {
    if (!s->timing) return -1;
    int id = -1;
    for (size_t t = 0; t < s->timers.size(); t++)
        if (s->timers[t].name == name) { id = (int)t; break; }
    if (id < 0) {
        s->timers.emplace_back();
        id = (int)s->timers.size() - 1;
        s->timers[id].name = name;
    }
    s->timers[id].bytes += bytes;
    hipEvent_t a = get_event(s), b = get_event(s);
    HIP_CHECK(hipEventRecord(a, s->stream));
    s->timers[id].pending.push_back({a, b});
    return id;
}

void timer_end(IBMSolver *s, int id) {
    if (id < 0) return;
    HIP_CHECK(hipEventRecord(s->timers[id].pending.back().second, s->stream));
}

void timers_flush(IBMSolver *s) {
    for (auto &t : s->timers) {
        for (auto &e : t.pending) {
            HIP_CHECK(hipEventSynchronize(e.second));
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
            t.total_ms += ms;
            t.count++;
            s->event_pool.push_back(e.first);
            s->event_pool.push_back(e.second);
        }
        t.pending.clear();
    }
}

#define TIMED(name, bytes, ...)                       \
    do {                                              \
        const int tid_ = timer_begin(s, name, bytes); \
        __VA_ARGS__;                                  \
        timer_end(s, tid_);                           \
    } while (0)

static double ncells(const IBMSolver *s) { return (double)s->Nx * s->Ny * s->Nz; }

// ---------------------------------------------------------------------------
// boundary data
// ---------------------------------------------------------------------------
enum { BCD_U = 0, BCD_P = 1, BCD_USOL = 2, BCD_PSOL = 3 };

// Builds the face data for one use: velocity at time t (BCD_U, BCD_USOL),
// pressure at t (BCD_P) or the increment p_bc(t + dt) - p_bc(t) (BCD_PSOL).
// VAL_FUNC faces are evaluated on the host over the reference's tangential
// ranges and uploaded; norm2[c] returns 2 sum g^2 of their BC rows.
static OuterParams make_outer(IBMSolver *s, int what, double t, double norm2[4]) {
    OuterParams op{};
    if (norm2) for (int c = 0; c < 4; c++) norm2[c] = 0;
    for (int ax = 0; ax < 3; ax++)
        for (int hi = 0; hi < 2; hi++) {
            const HostBC &B = s->bc[face_bc(ax, hi)];
            FaceBC &F = op.face[ax][hi];
            F.type = (int)B.type;
            int n1, n2, o1, o2;
            if (ax == 0) { n1 = s->Ny; n2 = s->Nz; o1 = 0; o2 = 0; }
            else if (ax == 1) { n1 = s->Nx + 4; n2 = s->Nz; o1 = 2; o2 = 0; }
            else { n1 = s->Nx + 4; n2 = s->Ny + 4; o1 = 2; o2 = 2; }
            F.n1 = n1; F.t1off = o1; F.t2off = o2;
            for (int c = 0; c < 3; c++) F.g[c] = B.cu[c];
            F.g[3] = what == BCD_PSOL ? 0.0 : B.cp;
            F.garr = nullptr;
            if (B.type == BC_CONVECTIVE_OUTFLOW) {
                // device-resident face values (advanced by conv_outflow), p = 0
                for (int c = 0; c < 4; c++) F.g[c] = 0;
                F.garr = s->conv_g[face_bc(ax, hi)].p;
                continue;
            }
            if (B.val_type != VAL_FUNC) continue;
            std::vector<double> h(4 * (size_t)n1 * n2, 0.0);
            for (int a = 0; a < n1; a++)
                for (int b = 0; b < n2; b++) {
                    const int t1 = a - o1, t2 = b - o2;
                    double x, y, z;
                    if (ax == 0) { x = hi ? s->xmax : s->xmin; y = H1(s->yc, t1); z = H1(s->zc, t2); }
                    else if (ax == 1) { x = H1(s->xc, t1); y = hi ? s->ymax : s->ymin; z = H1(s->zc, t2); }
                    else { x = H1(s->xc, t1); y = H1(s->yc, t2); z = hi ? s->zmax : s->zmin; }
                    const size_t o = (size_t)a * n2 + b;
                    if (B.type == BC_VELOCITY_COMPONENT)
                        for (int c = 0; c < 3; c++) h[c * (size_t)n1 * n2 + o] = B.fu[c](t, x, y, z);
                    if (B.type == BC_PRESSURE || B.type == BC_VELOCITY_PERIODIC) {
                        double v = B.fp(t, x, y, z);
                        if (what == BCD_PSOL) v = B.fp(t + s->dt, x, y, z) - v;
                        h[3 * (size_t)n1 * n2 + o] = v;
                    }
                    if (norm2)
                        for (int c = 0; c < 4; c++) norm2[c] += 2 * h[c * (size_t)n1 * n2 + o] * h[c * (size_t)n1 * n2 + o];
                }
            // kernels queued earlier in this step may still read the buffer
            HIP_CHECK(hipStreamSynchronize(s->stream));
            s->bc_arr[ax][hi].upload(h);
            F.garr = s->bc_arr[ax][hi].p;
        }
    return op;
}

// Convective outflow faces: init copies the first interior cell's velocity,
// otherwise one upwind step of the face values from the current u; each face
// leaves its BC-row norms in sums[kConvNormSlot + 3 bc + c]. Returns whether
// any face has this condition.
static bool conv_outflow(IBMSolver *s, int init) {
    bool any = false;
    for (int ax = 0; ax < 3; ax++)
        for (int hi = 0; hi < 2; hi++) {
            const int bci = face_bc(ax, hi);
            if (s->bc[bci].type != BC_CONVECTIVE_OUTFLOW) continue;
            any = true;
            if ((size_t)conv_outflow_blocks(s->d, ax) * 3 > s->partial.n)
                fatal("convective outflow: partial-sum buffer too small");
            TIMED("conv_outflow", 0, {
                const long nb = launch_conv_outflow(s->stream, s->geo, ax, hi, s->bc[bci].uc, s->dt, init, s->u[0].p,
                                                    s->u[1].p, s->u[2].p, s->conv_g[bci].p, s->sums.p + kConvSpeedSlot,
                                                    s->partial.p);
                launch_sum_partials(s->stream, s->partial.p, nb, 3, s->sums.p + kConvNormSlot + 3 * bci, 0);
            });
        }
    return any;
}

static void update_outer(IBMSolver *s) {
    OuterParams ou = make_outer(s, BCD_U, s->time, nullptr);
    TIMED("outer_u", 0, launch_outer(s->stream, s->geo, ou, s->u[0].p, s->u[1].p, s->u[2].p, 0, 0));
    OuterParams opp = make_outer(s, BCD_P, s->time, nullptr);
    TIMED("outer_p", 0, launch_outer(s->stream, s->geo, opp, s->p.p, nullptr, nullptr, 1, 0));
}

static void update_ghost(IBMSolver *s) {
    TIMED("ghost_update", 0,
          launch_ghost_update(s->stream, s->gc.p, s->ng, s->level_start.data(), (int)s->level_start.size() - 1,
                              s->ghost_old.p, s->u[0].p, s->u[1].p, s->u[2].p, s->p.p));
}

// One poll of a host wait on device-published values: a pause for the first
// ~2000 polls (about 0.1 ms), then a yield of the core, so that waits of
// many solvers on different threads (the serving loop) leave the cores to
// the threads that have work.
static inline void wait_pause(long n) {
    if (n < 2048) __builtin_ia32_pause();
    else sched_yield();
}

// Waits until a kernel enqueued on `st` has published epoch e at *ep (host
// memory), polling the stream now and then so that an error or a lost
// publication cannot make the wait endless.
static void wait_epoch(hipStream_t st, unsigned *ep, unsigned e) {
    for (long n = 1; __atomic_load_n(ep, __ATOMIC_ACQUIRE) != e; n++) {
        if ((n & 1023) == 0) {
            const hipError_t q = hipStreamQuery(st);
            if (q == hipSuccess) {
                if (__atomic_load_n(ep, __ATOMIC_ACQUIRE) == e) break;
                char buf[120];
                snprintf(buf, sizeof buf, "error: Krylov scalars were not published (epoch %u, expected %u)",
                         __atomic_load_n(ep, __ATOMIC_ACQUIRE), e);
                case_error(buf);
            }
            if (q != hipErrorNotReady) HIP_CHECK(q);
        }
        wait_pause(n);
    }
}
static void wait_published(hipStream_t st, KPub *pub_h, unsigned e) { wait_epoch(st, &pub_h->epoch, e); }

// The first nv sums, copied by a small kernel into device-mapped host memory
// (sums_h) and awaited there
static void read_sums(IBMSolver *s, int nv) {
    const unsigned e = ++s->sums_epoch;
    launch_publish(s->stream, s->sums.p, nv, s->sums_d, s->sums_d_epoch, e);
    wait_epoch(s->stream, s->sums_h_epoch, e);
}

// Finalize kernel `op` over the block partials, with the resulting scalars
// published into host memory, then a wait for them there: no copy and no
// stream synchronisation on the Krylov loop's path (each end of iteration
// otherwise costs a copy and a wake-up of the host thread).
static void finalize_read(IBMSolver *s, int op, long nblk, int nv) {
    const unsigned e = ++s->pub_epoch;
    launch_finalize(s->stream, op, s->partial.p, nblk, nv, s->ks.p, &s->pub_d->ks, &s->pub_d->epoch, e);
    wait_published(s->stream, s->pub_h, e);
    memcpy((void *)s->ks_h, (const void *)&s->pub_h->ks, sizeof(KScal));
}

// ---------------------------------------------------------------------------
// multigrid hierarchy of the pressure system
// ---------------------------------------------------------------------------
// coarsening stops at the first level below the fine grid with at most this
// many cells; that level is solved exactly through a dense inverse
static const long kDenseCells = 3000;
// levels with at least this many cells restrict their residual with two
// launches (residual, then restriction), smaller ones with one
static const long kSplitRestrictCells = 1L << 20;

static void level_geometry(IBMSolver *s, MGLevel &L, const std::vector<double> &xf, const std::vector<double> &yf,
                           const std::vector<double> &zf) {
    const std::vector<double> *f[3] = {&xf, &yf, &zf};
    std::vector<double> *D[3] = {&L.dx, &L.dy, &L.dz}, *C[3] = {&L.xc, &L.yc, &L.zc};
    const int lo[3] = {3, 2, 4}, hi[3] = {1, 0, 5};
    for (int a = 0; a < 3; a++) {
        const std::vector<double> &F = *f[a];
        const int N = (int)F.size() - 1;
        std::vector<double> &dd = *D[a], &cc = *C[a];
        dd.assign(N + 4, 0);
        cc.assign(N + 4, 0);
        for (int i = 0; i < N; i++) { H1(dd, i) = F[i + 1] - F[i]; H1(cc, i) = (F[i + 1] + F[i]) / 2; }
        const bool plo = isperiodic(s->bc[lo[a]].type), phi = isperiodic(s->bc[hi[a]].type);
        H1(dd, -1) = plo ? H1(dd, N - 1) : H1(dd, 0);
        H1(dd, -2) = plo ? H1(dd, N - 2) : H1(dd, 1);
        H1(dd, N) = phi ? H1(dd, 0) : H1(dd, N - 1);
        H1(dd, N + 1) = phi ? H1(dd, 1) : H1(dd, N - 2);
        H1(cc, -1) = H1(cc, 0) - (H1(dd, -1) + H1(dd, 0)) / 2;
        H1(cc, -2) = H1(cc, -1) - (H1(dd, -2) + H1(dd, -1)) / 2;
        H1(cc, N) = H1(cc, N - 1) + (H1(dd, N - 1) + H1(dd, N)) / 2;
        H1(cc, N + 1) = H1(cc, N) + (H1(dd, N) + H1(dd, N + 1)) / 2;
    }
}

static void kcoef(const IBMSolver *s, const std::vector<double> &D, const std::vector<double> &C, int N,
                  std::vector<double> &km, std::vector<double> &kp) {
    km.assign(N + 4, 0);
    kp.assign(N + 4, 0);
    for (int i = 0; i < N; i++) {
        H1(km, i) = s->dt / (2 * s->Re * (H1(C, i) - H1(C, i - 1)) * H1(D, i));
        H1(kp, i) = s->dt / (2 * s->Re * (H1(C, i + 1) - H1(C, i)) * H1(D, i));
    }
}

// Dense matrix of a coarse level's pressure operator (rows as the matrix-free
// stencil with solid faces skipped; inactive cells and, for a singular
// problem, the first fluid cell get identity rows), inverted on the device.
static void build_dense(IBMSolver *s, MGLevel &L) {
    const Dims &d = L.d;
    const int n = d.nx * d.ny * d.nz;
    const SysData &S = L.sys;
    std::vector<double> A((size_t)n * n, 0.0), vol(n);
    std::vector<long> rows(n);
    int pin = -1;
    for (int k = 0, c = 0; k < d.nz; k++)
        for (int j = 0; j < d.ny; j++)
            for (int i = 0; i < d.nx; i++, c++) {
                rows[c] = cidx(d, i, j, k);
                vol[c] = H1(L.dx, i) * H1(L.dy, j) * H1(L.dz, k);
                if (s->p_singular && pin < 0 && L.hflag[rows[c]] == F_FLUID) pin = c;
            }
    const int nn[3] = {d.nx, d.ny, d.nz};
    const int stride[3] = {1, d.nx, d.nx * d.ny};
    for (int c = 0; c < n; c++) {
        const long q = rows[c];
        double *row = &A[(size_t)c * n];
        if (L.hflag[q] != F_FLUID || c == pin) { row[c] = 1; continue; }
        const int id[3] = {c % d.nx, (c / d.nx) % d.ny, c / (d.nx * d.ny)};
        double diag = S.S64.c0;
        for (int a = 0; a < 3; a++) {
            const int ii = id[a], na = nn[a];
            const bool per = S.S64.ax[a].periodic != 0;
            double cm = S.h[a][0][ii], cp = S.h[a][1][ii], dm = S.h[a][2][ii], dp = S.h[a][3][ii];
            const bool hm = ii > 0 || per, hp = ii < na - 1 || per;
            const int cmi = c + ((ii == 0 ? na - 1 : -1)) * stride[a];
            const int cpi = c + ((ii == na - 1 ? -(na - 1) : 1)) * stride[a];
            if (hm && L.hflag[rows[cmi]] == F_SOLID) { cm = 0; dm = 0; }
            if (hp && L.hflag[rows[cpi]] == F_SOLID) { cp = 0; dp = 0; }
            diag += dm + dp;
            if (hm) row[cmi] -= cm;
            if (hp) row[cpi] -= cp;
        }
        row[c] += diag;
    }
    L.dn = n;
    L.dpin = pin;
    L.dinv64.upload(A);
    L.drow.upload(rows);
    L.dvol.upload(vol);
    DBuf<double> rowp, colp;
    rowp.alloc(n);
    colp.alloc(n);
    gauss_jordan_invert(s->stream, L.dinv64.p, n, rowp.p, colp.p);
    if (s->precision != IBM3D_PREC_FP64) {
        L.dinv32.alloc((size_t)n * n);
        launch_d2f_flat(s->stream, L.dinv64.p, L.dinv32.p, (long)n * n);
    }
    HIP_CHECK(hipMemcpyAsync(A.data(), L.dinv64.p, A.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    for (double v : A)
        if (!std::isfinite(v)) fatal("coarsest multigrid level: singular dense operator");
}

void build_mg(IBMSolver *s) {
    s->mg.clear();
    s->mg.emplace_back();
    {
        MGLevel &L0 = s->mg[0];
        L0.d = s->d;
        L0.dx = s->dx; L0.dy = s->dy; L0.dz = s->dz;
        L0.xc = s->xc; L0.yc = s->yc; L0.zc = s->zc;
        L0.hflag = s->flag_dev_h;
        std::vector<double> pk;
        pk.insert(pk.end(), s->dx.begin(), s->dx.end());
        pk.insert(pk.end(), s->dy.begin(), s->dy.end());
        pk.insert(pk.end(), s->dz.begin(), s->dz.end());
        L0.dxyz.upload(pk);
        // the smoother's over-relaxation factor (the operator kernels do not
        // read it)
        s->sys[3].S64.omega = s->mg_omega;
        s->sys[3].S32.omega = (float)s->mg_omega;
    }
    std::vector<double> xf = s->xf, yf = s->yf, zf = s->zf;
    const int lo[3] = {3, 2, 4}, hi[3] = {1, 0, 5};
    double h0 = 1e300;
    for (const std::vector<double> *f : {&xf, &yf, &zf})
        for (size_t i = 0; i + 1 < f->size(); i++) h0 = std::min(h0, (*f)[i + 1] - (*f)[i]);
    while (true) {
        const MGLevel &P = s->mg.back();
        const int n[3] = {P.d.nx, P.d.ny, P.d.nz};
        if (s->mg.size() > 1 && (long)n[0] * n[1] * n[2] <= kDenseCells) break;
        // Spacing-aware coarsening: along each axis, two neighbouring cells
        // merge when their combined width is at most 2 * gamma * h_l, where
        // h_l = 2^l * (smallest fine-grid width) is the target width of level
        // l. Where one axis is much finer than the others only that axis
        // coarsens (semi-coarsening in the strongly coupled direction), so the
        // point smoother stays effective on stretched grids; gamma <= 0
        // selects plain full coarsening of even-sized axes.
        std::vector<double> *F[3] = {&xf, &yf, &zf};
        const double thr = 2 * s->mg_gamma * h0 * std::ldexp(1.0, (int)s->mg.size() - 1) * (1 + 1e-9);
        std::vector<int> parent[3], start[3];
        bool merged = false;
        for (int a = 0; a < 3; a++) {
            const std::vector<double> &f = *F[a];
            const bool full = s->mg_gamma <= 0 && n[a] % 2 == 0 && n[a] >= 2;
            // a periodic axis keeps an even cell count, so the two neighbours
            // across the wrap have different red-black colours
            const bool per = s->bc[lo[a]].type == BC_ALL_PERIODIC && s->bc[hi[a]].type == BC_ALL_PERIODIC;
            for (int pass = 0; pass < 2; pass++) {
                const bool allow = pass == 0;
                std::vector<double> g{f[0]};
                parent[a].assign(n[a], 0);
                start[a].clear();
                bool any = false;
                for (int i = 0; i < n[a];) {
                    const bool pair = allow && i + 1 < n[a] &&
                                      (full || (s->mg_gamma > 0 && f[i + 2] - f[i] <= thr));
                    start[a].push_back(i);
                    parent[a][i] = (int)start[a].size() - 1;
                    if (pair) { parent[a][i + 1] = parent[a][i]; any = true; }
                    i += pair ? 2 : 1;
                    g.push_back(f[i]);
                }
                start[a].push_back(n[a]);
                if (pass == 0 && per && any && (g.size() - 1) % 2 == 1) continue;  // odd: leave this axis
                merged = merged || any;
                *F[a] = g;
                break;
            }
        }
        if (!merged) break;
        MGLevel L;
        const int m[3] = {(int)xf.size() - 1, (int)yf.size() - 1, (int)zf.size() - 1};
        L.d = make_dims(m[0], m[1], m[2]);
        level_geometry(s, L, xf, yf, zf);
        // activity: any fluid child (outer layers active)
        L.hflag.assign(L.d.ncell, F_FLUID);
        for (int K = 0; K < m[2]; K++)
            for (int J = 0; J < m[1]; J++)
                for (int I = 0; I < m[0]; I++) {
                    bool act = false;
                    for (int k = start[2][K]; k < start[2][K + 1]; k++)
                        for (int j = start[1][J]; j < start[1][J + 1]; j++)
                            for (int i = start[0][I]; i < start[0][I + 1]; i++)
                                if (P.hflag[cidx(P.d, i, j, k)] == F_FLUID) act = true;
                    L.hflag[cidx(L.d, I, J, K)] = act ? F_FLUID : F_SOLID;
                }
        L.flag.upload(L.hflag);
        std::vector<double> pk;
        pk.insert(pk.end(), L.dx.begin(), L.dx.end());
        pk.insert(pk.end(), L.dy.begin(), L.dy.end());
        pk.insert(pk.end(), L.dz.begin(), L.dz.end());
        L.dxyz.upload(pk);
        // rediscretised operator with the same boundary folds
        std::vector<double> kW, kE, kS, kN, kD, kU;
        kcoef(s, L.dx, L.xc, m[0], kW, kE);
        kcoef(s, L.dy, L.yc, m[1], kS, kN);
        kcoef(s, L.dz, L.zc, m[2], kD, kU);
        const std::vector<double> *D[3] = {&L.dx, &L.dy, &L.dz}, *C[3] = {&L.xc, &L.yc, &L.zc};
        const std::vector<double> *Km[3] = {&kW, &kS, &kD}, *Kp[3] = {&kE, &kN, &kU};
        IBMSolverBCType bct[3][2];
        for (int a = 0; a < 3; a++)
            for (int h = 0; h < 2; h++) bct[a][h] = s->bc[face_bc(a, h)].type;
        build_system(s, D, C, Km, Kp, m, bct, 3, L.sys);
        L.sys.S64.skip = L.sys.S32.skip = 1;
        L.sys.S64.omega = s->mg_omega;
        L.sys.S32.omega = (float)s->mg_omega;
        // faces to inactive (solid) neighbours, per cell
        {
            std::vector<uint8_t> fm(L.d.ncell, 0);
            for (int K = 0; K < m[2]; K++)
                for (int J = 0; J < m[1]; J++)
                    for (int I = 0; I < m[0]; I++) {
                        const int id[3] = {I, J, K};
                        uint8_t bits = 0;
                        for (int a = 0; a < 3; a++) {
                            const bool per = L.sys.S64.ax[a].periodic != 0;
                            int lo_[3] = {I, J, K}, hi_[3] = {I, J, K};
                            lo_[a] = id[a] == 0 ? (per ? m[a] - 1 : -1) : id[a] - 1;
                            hi_[a] = id[a] == m[a] - 1 ? (per ? 0 : -1) : id[a] + 1;
                            if (lo_[a] >= 0 && L.hflag[cidx(L.d, lo_[0], lo_[1], lo_[2])] == F_SOLID) bits |= 1u << (2 * a);
                            if (hi_[a] >= 0 && L.hflag[cidx(L.d, hi_[0], hi_[1], hi_[2])] == F_SOLID)
                                bits |= 1u << (2 * a + 1);
                        }
                        fm[cidx(L.d, I, J, K)] = bits;
                    }
            L.fmask.upload(fm);
            L.sys.S64.fmask = L.fmask.p;
            L.sys.S32.fmask = L.fmask.p;
        }
        // trilinear prolongation over the finer grid's cell centres: the
        // parent and the coarse neighbour on the far side of the fine centre
        const std::vector<double> *cf[3] = {&P.xc, &P.yc, &P.zc}, *cc[3] = {&L.xc, &L.yc, &L.zc};
        DBuf<int> *pidx[3] = {&L.pi, &L.pj, &L.pk}, *cpar[3] = {&L.ci, &L.cj, &L.ck}, *cst[3] = {&L.si, &L.sj, &L.sk};
        DBuf<double> *pw[3] = {&L.wi, &L.wj, &L.wk};
        std::vector<int> hfar[3];
        for (int a = 0; a < 3; a++) {
            std::vector<int> &far = hfar[a];
            far.resize(n[a]);
            std::vector<double> w(n[a]);
            const bool per = s->bc[lo[a]].type == BC_ALL_PERIODIC && s->bc[hi[a]].type == BC_ALL_PERIODIC;
            for (int i = 0; i < n[a]; i++) {
                const int I = parent[a][i];
                const double xf_ = H1((*cf[a]), i), Xn = H1((*cc[a]), I);
                if (start[a][I + 1] - start[a][I] == 1) { far[i] = I; w[i] = 1; continue; }
                const int J = xf_ > Xn ? I + 1 : I - 1;
                const double Xf = H1((*cc[a]), J);  // outer-layer centre beyond the ends
                const double wn = (Xf - xf_) / (Xf - Xn);
                if (J < 0 || J >= m[a]) {
                    if (per) {
                        w[i] = wn;
                        far[i] = (J + m[a]) % m[a];
                    } else {
                        const IBMSolverBCType t = s->bc[J < 0 ? lo[a] : hi[a]].type;
                        const bool dir = t == BC_PRESSURE || t == BC_VELOCITY_PERIODIC || t == BC_CONVECTIVE_OUTFLOW;
                        w[i] = dir ? wn - (1 - wn) : 1.0;  // Dirichlet: odd mirror, Neumann: even
                        far[i] = I;
                    }
                } else {
                    w[i] = wn;
                    far[i] = J;
                }
            }
            pidx[a]->upload(far);
            pw[a]->upload(w);
            cpar[a]->upload(parent[a]);
            cst[a]->upload(start[a]);
        }
        {
            // fine cells whose interpolation stencil touches an inactive coarse cell
            std::vector<uint8_t> pm(P.d.ncell, 0);
            for (int k = 0; k < n[2]; k++)
                for (int j = 0; j < n[1]; j++)
                    for (int i = 0; i < n[0]; i++) {
                        const int I2[2] = {parent[0][i], hfar[0][i]}, J2[2] = {parent[1][j], hfar[1][j]},
                                  K2[2] = {parent[2][k], hfar[2][k]};
                        bool solid = false;
                        for (int c = 0; c < 8; c++)
                            solid |= L.hflag[cidx(L.d, I2[c & 1], J2[(c >> 1) & 1], K2[c >> 2])] == F_SOLID;
                        pm[cidx(P.d, i, j, k)] = solid ? 1 : 0;
                    }
            L.pmask.upload(pm);
        }
        L.tr = {L.ci.p, L.cj.p, L.ck.p, L.si.p, L.sj.p, L.sk.p, L.pi.p, L.pj.p, L.pk.p, L.wi.p, L.wj.p, L.wk.p,
                L.pmask.p};
        L.x64.alloc(L.d.ncell); L.b64.alloc(L.d.ncell); L.r64.alloc(L.d.ncell); L.xa64.alloc(L.d.ncell);
        if (s->precision != IBM3D_PREC_FP64) {
            L.x32.alloc(L.d.ncell); L.b32.alloc(L.d.ncell); L.r32.alloc(L.d.ncell); L.xa32.alloc(L.d.ncell);
        }
        s->mg.push_back(std::move(L));
    }
    if (s->mg.size() > 1) {
        MGLevel &Lc = s->mg.back();
        if ((long)Lc.d.nx * Lc.d.ny * Lc.d.nz <= kDenseCells) build_dense(s, Lc);
    }
    // fine-level residual scratch
    s->mg[0].r64.alloc(s->d.ncell);
    s->mg[0].xa64.alloc(s->d.ncell);
    if (s->precision != IBM3D_PREC_FP64) {
        s->mg[0].r32.alloc(s->d.ncell);
        s->mg[0].xa32.alloc(s->d.ncell);
    }
}

template <typename T> static const T *dinv(const MGLevel &L);
template <> const double *dinv<double>(const MGLevel &L) { return L.dinv64.p; }
template <> const float *dinv<float>(const MGLevel &L) { return L.dinv32.p; }
template <typename T> static DBuf<T> &lev_xa(MGLevel &L);
template <> DBuf<double> &lev_xa<double>(MGLevel &L) { return L.xa64; }
template <> DBuf<float> &lev_xa<float>(MGLevel &L) { return L.xa32; }
template <typename T> static DBuf<T> &lev_r(MGLevel &L);
template <> DBuf<double> &lev_r<double>(MGLevel &L) { return L.r64; }
template <> DBuf<float> &lev_r<float>(MGLevel &L) { return L.r32; }
template <typename T> static DBuf<T> &lev_x(MGLevel &L);
template <> DBuf<double> &lev_x<double>(MGLevel &L) { return L.x64; }
template <> DBuf<float> &lev_x<float>(MGLevel &L) { return L.x32; }
template <typename T> static DBuf<T> &lev_b(MGLevel &L);
template <> DBuf<double> &lev_b<double>(MGLevel &L) { return L.b64; }
template <> DBuf<float> &lev_b<float>(MGLevel &L) { return L.b32; }
template <typename T> static const SysCoef<T> &sysc(const SysData &S);
template <> const SysCoef<double> &sysc<double>(const SysData &S) { return S.S64; }
template <> const SysCoef<float> &sysc<float>(const SysData &S) { return S.S32; }
template <typename T> static const GhostRows<T> &grows(const GhostData &G);
template <> const GhostRows<double> &grows<double>(const GhostData &G) { return G.G64; }
template <> const GhostRows<float> &grows<float>(const GhostData &G) { return G.G32; }
template <typename T> static DBuf<T> *kvec(Krylov &K);
template <> DBuf<double> *kvec<double>(Krylov &K) { return K.v64; }
template <> DBuf<float> *kvec<float>(Krylov &K) { return K.v32; }

static const int kCoarse = 40;

// x = V-cycle approximation of A^-1 b on level l (x0 = 0)

// Red-black Gauss-Seidel sweeps on x (zero_init: the initial x is zero and
// is not read). Smoother 0: one fused out-of-place launch per sweep,
// alternating between x and xalt, then a Jacobi pass over the ghost rows;
// algorithmic bytes per sweep: x and b read, x written, one flag byte per
// cell. Smoother 1: one launch per colour (x read once, b read and x
// written on half of the cells), ghost rows relaxed after each colour.
template <typename T>
static void smooth(IBMSolver *s, const Dims &d, const SysCoef<T> &S, const GhostRows<T> &G, const uint8_t *flag,
                   const T *b, T *x, T *xalt, int sweeps, bool zero_init, int level, bool first_done = false) {
    const double n = (double)d.nx * d.ny * d.nz;
    char nm[32];
    // a zero start also zeroes the ghost values (s->ghost_prev = -1)
    if (sweeps == 0) {
        if (zero_init) {
            TIMED("mg_zero", n * sizeof(T), launch_zero<T>(s->stream, d, x));
            if (G.ng) s->ghost_prev = -1;
        }
        return;
    }
    if (s->mg_smoother == 1) {
        for (int sw = 0; sw < sweeps; sw++)
            for (int c = 0; c < 2; c++) {
                // the first colour from a zero start writes every cell;
                // first_done: the kernel that wrote b already wrote it
                const bool z = zero_init && sw == 0 && c == 0;
                if (level >= 0) snprintf(nm, sizeof nm, "rbgs_L%d", level);
                else snprintf(nm, sizeof nm, "mom_rbgs");
                // Ghost rows are relaxed after each colour, those of the
                // colour just updated (the others keep their values); past a
                // zero start that happens inside the colour launch itself.
                if (G.ng && !z && s->ghost_prev >= 0) {
                    TIMED(nm, n * (2 * sizeof(T) + 1),
                          s->ghost_prev = launch_rbgs_color_ghost<T>(s->stream, d, S, flag, b, x, c, G, s->ghost_prev));
                    continue;
                }
                if (!(z && first_done))
                    TIMED(nm, z ? n * (1.5 * sizeof(T) + 1) : n * (2 * sizeof(T) + 1),
                          launch_rbgs_color<T>(s->stream, d, S, flag, b, x, c, z ? 1 : 0));
                if (!G.ng) continue;
                if (z) s->ghost_prev = -1;
                TIMED("ghost_relax", 0,
                      s->ghost_prev = launch_ghost_relax<T>(s->stream, G, S.is_p, b, x, s->ghost_prev, d, c));
            }
        return;
    }
    T *cur = x, *nxt = xalt;
    for (int sw = 0; sw < sweeps; sw++) {
        const bool z = zero_init && sw == 0;
        if (level >= 0) snprintf(nm, sizeof nm, "rbgs_fused_L%d", level);
        else snprintf(nm, sizeof nm, "mom_rbgs_fused");
        TIMED(nm, n * ((z ? 2 : 3) * sizeof(T) + 1),
              launch_rbgs_fused<T>(s->stream, d, S, flag, b, cur, nxt, z ? 1 : 0));
        if (G.ng) {
            if (z) s->ghost_prev = -1;
            TIMED("ghost_relax", 0, s->ghost_prev = launch_ghost_relax<T>(s->stream, G, S.is_p, b, nxt, s->ghost_prev));
        }
        std::swap(cur, nxt);
    }
    if (cur != x) TIMED("mg_copy", 2 * n * sizeof(T), launch_copy<T>(s->stream, d, cur, x));
}

// Whether the kernel that produces level m's right-hand side also writes the
// first colour of its pre-smoothing (one-launch-per-colour smoother, a
// smoothed level, at least one pre-sweep).
static bool fuse_first_colour(const IBMSolver *s, int m) {
    if (s->mg_smoother != 1 || s->mg_pre < 1) return false;
    if (m >= 1 && s->mg[m].dn > 0) return false;       // dense coarsest solve
    if (m + 1 == (int)s->mg.size()) return false;      // iterative coarsest solve
    return true;
}

// first_done: x already holds the first colour of the pre-smoothing
template <typename T>
static void vcycle(IBMSolver *s, int l, const T *b, T *x, bool first_done = false) {
    MGLevel &L = s->mg[l];
    const bool fine = l == 0;
    const SysCoef<T> &S = fine ? sysc<T>(s->sys[3]) : sysc<T>(L.sys);
    const GhostRows<T> G0 = fine ? grows<T>(s->gp) : GhostRows<T>{};
    const uint8_t *flag = fine ? s->flag.p : L.flag.p;
    const double cell_b = (double)L.d.nx * L.d.ny * L.d.nz * sizeof(T);
    if (l >= 1 && L.dn > 0) {
        TIMED("mg_dense", (double)L.dn * L.dn * sizeof(T),
              launch_dense_solve<T>(s->stream, dinv<T>(L), L.dn, L.drow.p, L.dvol.p, L.flag.p, s->p_singular ? 1 : 0,
                                    L.dpin, b, x));
        return;
    }
    if (l + 1 == (int)s->mg.size()) {
        TIMED("mg_coarse", 0, launch_coarse_solve<T>(s->stream, L.d, S, flag, b, x, kCoarse, s->p_singular ? 1 : 0,
                                                     L.dxyz.p));
        return;
    }
    const bool light = s->mg_light_from > 0 && l >= s->mg_light_from;
    smooth<T>(s, L.d, S, G0, flag, b, x, lev_xa<T>(L).p, light ? 1 : s->mg_pre, true, l, first_done);
    MGLevel &Cl = s->mg[l + 1];
    const bool fuse = fuse_first_colour(s, l + 1);
    const SysCoef<T> *Sc = fuse ? &sysc<T>(Cl.sys) : nullptr;
    const uint8_t *flag_c = fuse ? Cl.flag.p : nullptr;
    T *xc = fuse ? lev_x<T>(Cl).p : nullptr;
    // coarse-level bytes of a restriction: b written, and with the fused first
    // colour the coarse flags read and every coarse x written
    const double cn = (double)Cl.d.nx * Cl.d.ny * Cl.d.nz;
    const double coarse_b = cn * (sizeof(T) + (fuse ? 1 + sizeof(T) : 0));
    if ((long)L.d.nx * L.d.ny * L.d.nz >= kSplitRestrictCells) {
        // large levels: residual, then restriction (measured faster than the
        // one-pass kernel, whose per-coarse-cell child loops run short here)
        char rn[32], tn[32];
        snprintf(rn, sizeof rn, "mg_resid_L%d", l);
        snprintf(tn, sizeof tn, "mg_restrict_L%d", l);
        bool rz = false;
        TIMED(rn, 3 * cell_b + cell_b / sizeof(T),
              rz = launch_mg_residual<T>(s->stream, L.d, S, flag, b, x, lev_r<T>(L).p));
        // fine r and flags read
        TIMED(tn, cell_b + cell_b / sizeof(T) + coarse_b,
              launch_restrict<T>(s->stream, L.d, Cl.d, Cl.tr, flag, lev_r<T>(L).p, L.dxyz.p, L.dxyz.p + (L.d.nx + 4),
                                 L.dxyz.p + (L.d.nx + 4) + (L.d.ny + 4), Cl.dxyz.p, Cl.dxyz.p + (Cl.d.nx + 4),
                                 Cl.dxyz.p + (Cl.d.nx + 4) + (Cl.d.ny + 4), lev_b<T>(Cl).p, rz, Sc, flag_c, xc));
    } else {
        // residual restricted in one pass: x and b read, one flag byte per cell
        char nm[32];
        snprintf(nm, sizeof nm, "mg_resid_restrict_L%d", l);
        TIMED(nm, 2 * cell_b + cell_b / sizeof(T) + coarse_b,
              launch_resid_restrict<T>(s->stream, L.d, S, flag, b, x, Cl.d, Cl.tr, L.dxyz.p, Cl.dxyz.p,
                                       lev_b<T>(Cl).p, Sc, flag_c, xc));
    }
    vcycle<T>(s, l + 1, lev_b<T>(Cl).p, lev_x<T>(Cl).p, fuse);
    char pnm[32];
    snprintf(pnm, sizeof pnm, "mg_prolong_L%d", l);
    // fine x read and written, a flag and an interpolation-mask byte per fine
    // cell, each coarse correction read once
    TIMED(pnm, 2 * cell_b + 2 * cell_b / sizeof(T) + cn * sizeof(T),
          launch_prolong<T>(s->stream, L.d, Cl.d, Cl.tr, flag, Cl.flag.p, lev_x<T>(Cl).p, x));
    smooth<T>(s, L.d, S, G0, flag, b, x, lev_xa<T>(L).p, light ? 1 : s->mg_post, false, l);
}

// V-cycle as a preconditioner. Without per-kernel timing the cycle for a
// given (input, output) pair is captured once into a HIP graph and replayed:
// the coarse levels are launch-latency bound, and a graph replay dispatches
// its kernels back to back without host launch overhead.
template <typename F>
static void replay_or_capture(IBMSolver *s, const void *in, void *out, F body) {
    if (s->timing == 1 || !s->graphs) {
        body();
        return;
    }
    for (auto &g : s->vgraphs)
        if (g.in == in && g.out == out) {
            TIMED("vcycle_graph", 0, HIP_CHECK(hipGraphLaunch(g.exec, s->stream)));
            return;
        }
    hipGraph_t graph;
    hipGraphExec_t exec;
    {
        HipGlobalLock lk(hip_global_mutex());
        const int timing = s->timing;
        s->timing = 0;  // no event records inside the capture
        HIP_CHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        body();
        HIP_CHECK(hipStreamEndCapture(s->stream, &graph));
        s->timing = timing;
        HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        HIP_CHECK(hipGraphDestroy(graph));
    }
    s->vgraphs.push_back({in, out, exec});
    TIMED("vcycle_graph", 0, HIP_CHECK(hipGraphLaunch(exec, s->stream)));
}

template <typename T>
static void mg_precond(IBMSolver *s, const T *in, T *out) {
    replay_or_capture(s, in, out, [&]() { vcycle<T>(s, 0, in, out); });
}

// FP32 V-cycle applied to an FP64 vector (the preconditioner of the FP64
// Krylov iteration in mixed precision)
static void mg_precond_f32(IBMSolver *s, const double *in, double *out) {
    const double cb = ncells(s);
    replay_or_capture(s, in, out, [&]() {
        // the conversion also writes the first smoother colour of the fine level
        const bool fuse = fuse_first_colour(s, 0);
        if (fuse)
            TIMED("d2f", cb * 17, launch_d2f_first(s->stream, s->d, s->sys[3].S32, s->flag.p, in, s->ref_rhs32.p,
                                                   s->ref_e32.p));
        else
            TIMED("d2f", cb * 12, launch_d2f(s->stream, s->d, in, s->ref_rhs32.p));
        vcycle<float>(s, 0, s->ref_rhs32.p, s->ref_e32.p, fuse);
        TIMED("f2d", cb * 12, launch_f2d(s->stream, s->d, s->ref_e32.p, out));
    });
}

// The FP32 V-cycle on ref_rhs32 into `out` alone (its conversions fused
// into the neighbouring Krylov kernels); first_done as for vcycle.
static void mg_vcycle_f32(IBMSolver *s, bool first_done, float *out) {
    replay_or_capture(s, s->ref_rhs32.p, out, [&]() { vcycle<float>(s, 0, s->ref_rhs32.p, out, first_done); });
}

// ---------------------------------------------------------------------------
// BiCGSTAB (van der Vorst 1992), right-preconditioned, device scalars.
// prec: 0 none, 1 Jacobi, 2 multigrid (pressure only).
// bnorm > 0: stopping norm of the reference system's right-hand side;
// bnorm <= 0: use ||b|| of this system (inner solves).
// ---------------------------------------------------------------------------
// Stagnation test of the restarted (mixed-precision) BiCGSTAB: from iteration
// kStallFrom after a (re)start, a recurrence residual above kStallFactor times
// the one kStallWindow iterations earlier restarts the iteration (measured on
// the pinned cavity and channel pressure systems: 40-50 iterations per step
// without it, 8-20 with it, as FP64; unchanged where the FP32 cycle does not
// stall, the sphere and urban cases).
constexpr int kStallWindow = 3, kStallFrom = 4;
constexpr double kStallFactor = 0.1;
// A mixed-Krylov pressure solve that stalls this many times switches the
// solver's pressure preconditioner to the FP64 cycle for the following steps
// (the channel at tol 1e-11: 14-39 iterations per step with FP32 cycles and
// restarts, 10 with the FP64 cycle; the cavity stalls at most once a solve).
constexpr int kStallFallback = 2;

template <typename T>
static int bicgstab(IBMSolver *s, const SysCoef<T> &S, const GhostRows<T> &G, const T *b, T *x, int prec,
                    const NormSpec &ns, double tol, int maxit, double bnorm, double *final_rel, const char *tag,
                    int restart = 0, int *stalls = nullptr) {
    const Dims &d = s->d;
    DBuf<T> *V = kvec<T>(s->kry);
    T *r = V[0].p, *rh = V[1].p, *p = V[2].p, *v = V[3].p, *ph = V[4].p, *sv = V[5].p, *sh = V[6].p, *t = V[7].p;
    const uint8_t *flag = s->flag.p;
    const long nv = vec_partials(d), na = apply_partials(d, G.ng);
    const double cb = ncells(s) * sizeof(T);
    char nm[64];

    // the fused direction updates (Jacobi, FP32 multigrid) form the first
    // direction as p = r, so p and v need no clearing
    const bool fused_dir = prec == 1 || (prec == 3 && std::is_same<T, double>::value);
    // x0 = 0: r0 = b, whose norms the initialisation kernel forms
    TIMED("bicg_init", (fused_dir ? 4 : 6) * cb,
          launch_bicg_init<T>(s->stream, d, r, rh, fused_dir ? nullptr : p, v, x, b, ns, s->partial.p));
    finalize_read(s, 3, nv, 2);
    bool first_dir = fused_dir;
    const double bn = bnorm > 0 ? bnorm : std::sqrt(ns.on ? s->ks_h->bbn : s->ks_h->bb);
    if (bn == 0) { *final_rel = 0; return 0; }
    double rn = std::sqrt(ns.on ? s->ks_h->rrn : s->ks_h->rr);
    auto precond = [&](const T *in, T *out) {
        if (prec == 1) {
            TIMED("jacobi", 2 * cb, launch_jacobi<T>(s->stream, d, S, G, flag, in, out));
        } else if (prec == 2) {
            mg_precond<T>(s, in, out);
        } else if (prec == 3) {
            if constexpr (std::is_same<T, double>::value) mg_precond_f32(s, in, out);
            else abort();
        } else {
            launch_copy<T>(s->stream, d, in, out);
        }
    };
    // FP32 multigrid preconditioner of the FP64 iteration: conversions fused
    const bool f32mg = prec == 3 && std::is_same<T, double>::value;
    const bool fuse1 = f32mg && fuse_first_colour(s, 0);
    const SysCoef<float> *S32f = fuse1 ? &s->sys[3].S32 : nullptr;
    // Mixed-Krylov mode with Jacobi (momentum): the preconditioned directions
    // are stored in FP32 as well (rounded once, then used unchanged by the
    // operator and the update).
    const bool f32mom = prec == 1 && std::is_same<T, double>::value && s->precision == IBM3D_PREC_MIXED_KRYLOV;
    // the iteration's two FP32 directions (M^-1 p, M^-1 s): read by the
    // operator and the update in FP32, no FP64 copies
    float *e32a = f32mg || f32mom ? s->ref_e32.p : nullptr, *e32b = f32mg || f32mom ? s->ref_e32b.p : nullptr;
    int it = 0;
    // recurrence residuals since the last (re)start, for the stagnation test
    std::vector<double> hist{rn};
    int it_rs = 0;
    for (;;) {
    while (it < maxit && rn > tol * bn) {
        it++;
        // which direction update: 2 forms the first direction p = r
        const int wp = first_dir ? 2 : 0;
        first_dir = false;
        snprintf(nm, sizeof nm, "spmv_%s", tag);
        // direction p and ph = M^-1 p, then v = A ph. FP32 multigrid: the
        // update writes the FP32 right-hand side (and the cycle's first
        // colour), the operator reads the FP32 result and writes ph.
        if (prec == 1) {
            if constexpr (std::is_same<T, double>::value) {
                if (f32mom)
                    TIMED("bicg_jacobi", 4.5 * cb, launch_bicg_jacobi_f32(s->stream, d, S, G, flag, s->ks.p, wp, r, v, p,
                                                                        e32a));
                else
                    TIMED("bicg_jacobi", 5 * cb, launch_bicg_jacobi<T>(s->stream, d, S, G, flag, s->ks.p, wp, r, v, p, ph));
            } else {
                TIMED("bicg_jacobi", 5 * cb, launch_bicg_jacobi<T>(s->stream, d, S, G, flag, s->ks.p, wp, r, v, p, ph));
            }
        } else if (f32mg) {
            if constexpr (std::is_same<T, double>::value) {
                TIMED("bicg_d2f", (fuse1 ? 5.5 : 5) * cb, launch_bicg_d2f(s->stream, d, S32f, flag, s->ks.p, wp, r, v, p,
                                                                     s->ref_rhs32.p, fuse1 ? e32a : nullptr));
                mg_vcycle_f32(s, fuse1, e32a);
            }
        } else {
            TIMED("bicg_vec", 4 * cb, launch_bicg_p<T>(s->stream, d, s->ks.p, r, v, p));
            precond(p, ph);
        }
        if constexpr (std::is_same<T, double>::value) {
            if (f32mg || f32mom)
                TIMED(nm, 2.5 * cb, launch_apply_f32in(s->stream, d, S, G, flag, e32a, nullptr, v, 1, rh,
                                                       s->partial.p));
            else
                TIMED(nm, 3 * cb, launch_apply<T>(s->stream, d, S, G, flag, ph, v, 1, rh, s->partial.p));
        } else {
            TIMED(nm, 3 * cb, launch_apply<T>(s->stream, d, S, G, flag, ph, v, 1, rh, s->partial.p));
        }
        launch_finalize(s->stream, 0, s->partial.p, na, 1, s->ks.p);
        if (prec == 1) {
            if constexpr (std::is_same<T, double>::value) {
                if (f32mom)
                    TIMED("bicg_jacobi", 3.5 * cb, launch_bicg_jacobi_f32(s->stream, d, S, G, flag, s->ks.p, 1, r, v, sv,
                                                                        e32b));
                else
                    TIMED("bicg_jacobi", 4 * cb, launch_bicg_jacobi<T>(s->stream, d, S, G, flag, s->ks.p, 1, r, v, sv, sh));
            } else {
                TIMED("bicg_jacobi", 4 * cb, launch_bicg_jacobi<T>(s->stream, d, S, G, flag, s->ks.p, 1, r, v, sv, sh));
            }
        } else if (f32mg) {
            if constexpr (std::is_same<T, double>::value) {
                TIMED("bicg_d2f", (fuse1 ? 4.5 : 4) * cb, launch_bicg_d2f(s->stream, d, S32f, flag, s->ks.p, 1, r, v, sv,
                                                                     s->ref_rhs32.p, fuse1 ? e32b : nullptr));
                mg_vcycle_f32(s, fuse1, e32b);
            }
        } else {
            TIMED("bicg_vec", 3 * cb, launch_bicg_s<T>(s->stream, d, s->ks.p, r, v, sv));
            precond(sv, sh);
        }
        if constexpr (std::is_same<T, double>::value) {
            if (f32mg || f32mom)
                TIMED(nm, 2.5 * cb, launch_apply_f32in(s->stream, d, S, G, flag, e32b, nullptr, t, 2, sv,
                                                       s->partial.p));
            else
                TIMED(nm, 3 * cb, launch_apply<T>(s->stream, d, S, G, flag, sh, t, 2, sv, s->partial.p));
        } else {
            TIMED(nm, 3 * cb, launch_apply<T>(s->stream, d, S, G, flag, sh, t, 2, sv, s->partial.p));
        }
        launch_finalize(s->stream, 1, s->partial.p, na, 2, s->ks.p);
        if constexpr (std::is_same<T, double>::value) {
            if (f32mg || f32mom)
                TIMED("bicg_xr", 7 * cb, launch_bicg_xr_f32dir(s->stream, d, s->ks.p, e32a, e32b, sv, t, rh, x, r, ns,
                                                               s->partial.p));
            else
                TIMED("bicg_xr", 8 * cb, launch_bicg_xr<T>(s->stream, d, s->ks.p, ph, sh, sv, t, rh, x, r, ns,
                                                           s->partial.p));
        } else {
            TIMED("bicg_xr", 8 * cb, launch_bicg_xr<T>(s->stream, d, s->ks.p, ph, sh, sv, t, rh, x, r, ns, s->partial.p));
        }
        finalize_read(s, 2, nv, 3);
        rn = std::sqrt(ns.on ? s->ks_h->rrn : s->ks_h->rr);
        const bool bad = s->ks_h->breakdown || !std::isfinite(rn);
        hist.push_back(rn);
        const int since = it - it_rs;
        // stagnation: from the fourth iteration after a (re)start, less
        // than a tenfold reduction over the last three iterations
        const bool stalled = since >= kStallFrom && rn > kStallFactor * hist[since - kStallWindow];
        if (restart > 0 && it < maxit && (bad || it % restart == 0 || stalled) && !(rn <= tol * bn)) {
            if (stalled && stalls) ++*stalls;
            // Restart from the true residual of the current iterate. With a
            // preconditioner that is not exactly the same linear map every
            // time (an FP32 multigrid cycle), BiCGSTAB's recurrences lose
            // their bi-orthogonality; on pinned (otherwise singular)
            // pressure systems it then stalls near 1e-6 of its starting
            // residual and a few iterations later may carry the iterate
            // away. Restarting resets the recurrences, keeping x: at the
            // first stall, and every `restart` iterations in any case.
            TIMED("residual", 3 * cb, launch_residual<T>(s->stream, d, S, G, flag, b, x, r, ns, s->partial.p));
            launch_bicg_restart<T>(s->stream, d, r, rh, p, v);
            finalize_read(s, 3, na, 2);
            rn = std::sqrt(ns.on ? s->ks_h->rrn : s->ks_h->rr);
            it_rs = it;
            hist.assign(1, rn);
            if (!std::isfinite(rn)) break;
            continue;
        }
        if (bad) break;
    }
    // true residual of the final iterate
    TIMED("residual", 3 * cb, launch_residual<T>(s->stream, d, S, G, flag, b, x, r, ns, s->partial.p));
    finalize_read(s, 4, na, 2);
    const double rt = std::sqrt(ns.on ? s->ks_h->rrn : s->ks_h->rr);
    if (restart > 0 && it < maxit && std::isfinite(rt) && rt > tol * bn) {
        // the recurrence met the tolerance but the true residual did not:
        // restart from it (r already holds b - A x)
        launch_bicg_restart<T>(s->stream, d, r, rh, p, v);
        finalize_read(s, 3, na, 2);
        rn = rt;
        it_rs = it;
        hist.assign(1, rn);
        continue;
    }
    *final_rel = rt / bn;
    return it;
    }
}

// ---------------------------------------------------------------------------
// The other pressure solver types of src/ibm3d_setup.c:1434-1536 (FP64
// iterates; the multigrid cycle runs in the precision the mode selects).
// `prec` as for bicgstab: 0 none, 1 Jacobi, 2 multigrid, 3 FP32 multigrid.
// ---------------------------------------------------------------------------
static void read_gm(IBMSolver *s, int n) {
    HIP_CHECK(hipMemcpyAsync(s->gm_h, s->gm.dev.p, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
}

static void precond64(IBMSolver *s, int prec, const double *in, double *out) {
    const double cb = ncells(s) * sizeof(double);
    if (prec == 1) TIMED("jacobi", 2 * cb, launch_jacobi<double>(s->stream, s->d, s->sys[3].S64, s->gp.G64, s->flag.p, in, out));
    else if (prec == 2) mg_precond<double>(s, in, out);
    else if (prec == 3) mg_precond_f32(s, in, out);
    else launch_copy<double>(s->stream, s->d, in, out);
}

// true residual r = b - A x; returns its norm (normalised rows when ns.on)
static double true_residual(IBMSolver *s, const double *b, const double *x, double *r, const NormSpec &ns) {
    const Dims &d = s->d;
    TIMED("residual", 3 * ncells(s) * 8,
          launch_residual<double>(s->stream, d, s->sys[3].S64, s->gp.G64, s->flag.p, b, x, r, ns, s->partial.p));
    finalize_read(s, 4, apply_partials(d, s->gp.G64.ng), 2);
    return std::sqrt(ns.on ? s->ks_h->rrn : s->ks_h->rr);
}

// Restarted GMRES(10) (Saad & Schultz 1986) on the system with normalised
// rows (each fluid row divided by its sum of K, the system whose residual
// the reference's stopping test measures), right-preconditioned by the
// multigrid cycle applied after undoing that scaling, so that the
// preconditioned operator stays close to the identity. Arnoldi with
// classical Gram-Schmidt applied twice (one host read per step), Givens
// rotations on the host; x0 = 0 as in src/ibm3d_fracstep.c:650.
static int gmres_p(IBMSolver *s, const double *b, double *x, int prec, const NormSpec &ns, double tol, int maxit,
                   double bn, double *final_rel) {
    const Dims &d = s->d;
    GmresData &W = s->gm;
    const int m = kGmresRestart;
    const double cb = ncells(s) * sizeof(double);
    double *dev = W.dev.p, *part = W.partial.p, *h = s->gm_h;
    launch_zero<double>(s->stream, d, x);
    auto residual_to_v0 = [&]() {
        // V0 = N (b - A x), returns |V0|
        true_residual(s, b, x, W.V[0].p, ns);
        VecSet<double> none{};
        TIMED("gmres_dots", 2 * cb, launch_gm_dots<double>(s->stream, d, none, 0, W.V[0].p, ns, 1, part, dev));
        read_gm(s, kVecSetMax + 1);
        return std::sqrt(h[kVecSetMax]);
    };
    double beta = residual_to_v0(), rn = beta;
    if (bn == 0) { *final_rel = 0; return 0; }
    int it = 0;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    auto Hc = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
    while (it < maxit && rn > tol * bn && std::isfinite(rn)) {
        launch_gm_scale<double>(s->stream, d, W.V[0].p, 1.0 / beta, W.V[0].p, ns, 0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        while (j < m && it < maxit) {
            it++;
            launch_gm_scale<double>(s->stream, d, W.V[j].p, 1.0, W.t.p, ns, 1);
            precond64(s, prec, W.t.p, W.Z[j].p);
            TIMED("spmv_p64", 3 * cb,
                  launch_apply<double>(s->stream, d, s->sys[3].S64, s->gp.G64, s->flag.p, W.Z[j].p, W.V[j + 1].p, 0,
                                       nullptr, s->partial.p));
            VecSet<double> Vs{};
            for (int i = 0; i <= j; i++) Vs.v[i] = W.V[i].p;
            TIMED("gmres_dots", (j + 3) * cb,
                  launch_gm_dots<double>(s->stream, d, Vs, j + 1, W.V[j + 1].p, ns, 1, part, dev));
            TIMED("gmres_update", (j + 3) * cb,
                  launch_gm_update<double>(s->stream, d, Vs, j + 1, dev, W.V[j + 1].p, part, dev + 16));
            TIMED("gmres_dots", (j + 2) * cb,
                  launch_gm_dots<double>(s->stream, d, Vs, j + 1, W.V[j + 1].p, ns, 0, part, dev + 32));
            TIMED("gmres_update", (j + 3) * cb,
                  launch_gm_update<double>(s->stream, d, Vs, j + 1, dev + 32, W.V[j + 1].p, part, dev + 48));
            read_gm(s, 49);
            const double hn = std::sqrt(std::max(h[48], 0.0));
            for (int i = 0; i <= j; i++) Hc(i, j) = h[i] + h[32 + i];
            Hc(j + 1, j) = hn;
            if (hn > 0) launch_gm_scale<double>(s->stream, d, W.V[j + 1].p, 1.0 / hn, W.V[j + 1].p, ns, 0);
            for (int i = 0; i < j; i++) {
                const double t0 = cs[i] * Hc(i, j) + sn[i] * Hc(i + 1, j);
                Hc(i + 1, j) = -sn[i] * Hc(i, j) + cs[i] * Hc(i + 1, j);
                Hc(i, j) = t0;
            }
            const double den = std::hypot(Hc(j, j), Hc(j + 1, j));
            cs[j] = den > 0 ? Hc(j, j) / den : 1.0;
            sn[j] = den > 0 ? Hc(j + 1, j) / den : 0.0;
            Hc(j, j) = den;
            Hc(j + 1, j) = 0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            rn = std::fabs(g[j + 1]);
            j++;
            if (!std::isfinite(rn) || rn <= tol * bn || hn == 0) break;
        }
        // x += Z y, H(0:j, 0:j) y = g(0:j)
        for (int i = j - 1; i >= 0; i--) {
            double acc = g[i];
            for (int c = i + 1; c < j; c++) acc -= Hc(i, c) * y[c];
            y[i] = Hc(i, i) != 0 ? acc / Hc(i, i) : 0.0;
        }
        for (int i = 0; i < j; i++) h[i] = -y[i];
        HIP_CHECK(hipMemcpyAsync(dev, h, j * sizeof(double), hipMemcpyHostToDevice, s->stream));
        VecSet<double> Zs{};
        for (int i = 0; i < j; i++) Zs.v[i] = W.Z[i].p;
        TIMED("gmres_update", (j + 2) * cb, launch_gm_update<double>(s->stream, d, Zs, j, dev, x, part, dev + 16));
        beta = residual_to_v0();
        rn = beta;
    }
    *final_rel = rn / bn;
    return it;
}

// Preconditioned conjugate gradients (src/ibm3d_setup.c:1459-1464): the
// reference offers it for the pressure although the ghost rows make the
// operator nonsymmetric. Stops on the true residual's normalised norm.
static int pcg_p(IBMSolver *s, const double *b, double *x, int prec, const NormSpec &ns, double tol, int maxit,
                 double bn, double *final_rel) {
    const Dims &d = s->d;
    DBuf<double> *K = s->kry.v64;
    double *r = K[0].p, *z = K[1].p, *pv = K[2].p, *q = K[3].p;
    double *dev = s->gm.dev.p, *part = s->gm.partial.p, *h = s->gm_h;
    const double cb = ncells(s) * sizeof(double);
    launch_zero<double>(s->stream, d, x);
    double rn = true_residual(s, b, x, r, ns);
    if (bn == 0) { *final_rel = 0; return 0; }
    precond64(s, prec, r, z);
    launch_dot<double>(s->stream, d, r, z, part, dev);
    read_gm(s, 1);
    double rho = h[0];
    launch_copy<double>(s->stream, d, z, pv);
    int it = 0;
    while (it < maxit && rn > tol * bn && std::isfinite(rn)) {
        it++;
        TIMED("spmv_p64", 3 * cb,
              launch_apply<double>(s->stream, d, s->sys[3].S64, s->gp.G64, s->flag.p, pv, q, 0, nullptr, s->partial.p));
        launch_dot<double>(s->stream, d, pv, q, part, dev);
        read_gm(s, 1);
        if (h[0] == 0 || !std::isfinite(h[0])) break;
        const double alpha = rho / h[0];
        TIMED("pcg_update", 5 * cb, launch_pcg_update<double>(s->stream, d, alpha, pv, q, x, r, ns, part, dev));
        read_gm(s, 2);
        rn = std::sqrt(ns.on ? h[1] : h[0]);
        if (rn <= tol * bn) break;
        precond64(s, prec, r, z);
        launch_dot<double>(s->stream, d, r, z, part, dev);
        read_gm(s, 1);
        const double rho1 = h[0];
        if (rho == 0 || !std::isfinite(rho1)) break;
        launch_lincomb<double>(s->stream, d, z, rho1 / rho, pv, 0);
        rho = rho1;
    }
    rn = true_residual(s, b, x, r, ns);
    *final_rel = rn / bn;
    return it;
}

// The multigrid cycle as a stationary solver (BoomerAMG as the solver,
// src/ibm3d_setup.c:1435-1457): x += M (b - A x) until the residual test.
static int mg_solve_p(IBMSolver *s, const double *b, double *x, int prec, const NormSpec &ns, double tol, int maxit,
                      double bn, double *final_rel) {
    const Dims &d = s->d;
    DBuf<double> *K = s->kry.v64;
    double *r = K[0].p, *z = K[1].p;
    launch_zero<double>(s->stream, d, x);
    double rn = true_residual(s, b, x, r, ns);
    if (bn == 0) { *final_rel = 0; return 0; }
    int it = 0;
    while (it < maxit && rn > tol * bn && std::isfinite(rn)) {
        it++;
        precond64(s, prec, r, z);
        launch_lincomb<double>(s->stream, d, z, 1.0, x, 1);
        rn = true_residual(s, b, x, r, ns);
    }
    *final_rel = rn / bn;
    return it;
}

// FP64 iterative refinement around FP32 BiCGSTAB inner solves.
// BiCGSTAB iterations between restarts when the preconditioner runs in FP32
static const int kMixedRestart = 40;

static int refine_mixed(IBMSolver *s, int sk, const double *b, double *x, int prec, const NormSpec &ns, double tol,
                        int maxit, double bnorm, double *final_rel) {
    const Dims &d = s->d;
    const SysCoef<double> &S64 = s->sys[sk].S64;
    const SysCoef<float> &S32 = s->sys[sk].S32;
    const GhostData &GD = sk == 3 ? s->gp : s->gmom;
    const uint8_t *flag = s->flag.p;
    const long na = apply_partials(d, GD.G64.ng);
    double *r64 = s->ref_r64.p;
    float *rhs32 = s->ref_rhs32.p, *e32 = s->ref_e32.p;
    const double cb = ncells(s);
    launch_zero<double>(s->stream, d, x);
    TIMED("residual64", 3 * cb * 8, launch_residual<double>(s->stream, d, S64, GD.G64, flag, b, x, r64, ns, s->partial.p));
    finalize_read(s, 4, na, 2);
    const double bn = bnorm > 0 ? bnorm : std::sqrt(ns.on ? s->ks_h->rrn : s->ks_h->rr);
    if (bn == 0) { *final_rel = 0; return 0; }
    double rn = std::sqrt(ns.on ? s->ks_h->rrn : s->ks_h->rr);
    int total = 0, outer = 0;
    const NormSpec off{0, nullptr, nullptr, nullptr, nullptr, -1};
    const double inner_tol = 1e-4;
    while (rn > tol * bn && outer < 20 && total < maxit) {
        outer++;
        TIMED("d2f", cb * 12, launch_d2f(s->stream, d, r64, rhs32));
        double rel = 0;
        const int it = bicgstab<float>(s, S32, GD.G32, rhs32, e32, prec, off, inner_tol, 200, -1.0, &rel,
                                       sk == 3 ? "p32" : "u32");
        total += std::max(it, 1);
        TIMED("add_f2d", cb * 20, launch_add_f2d(s->stream, d, e32, x));
        TIMED("residual64", 3 * cb * 8,
              launch_residual<double>(s->stream, d, S64, GD.G64, flag, b, x, r64, ns, s->partial.p));
        finalize_read(s, 4, na, 2);
        const double rn_new = std::sqrt(ns.on ? s->ks_h->rrn : s->ks_h->rr);
        if (!std::isfinite(rn_new)) { rn = rn_new; break; }
        rn = rn_new;
    }
    *final_rel = rn / bn;
    return total;
}

// Records the outcome of the solve of system sk and reacts like the reference
// (src/ibm3d_fracstep.c:363-369, 673-679): a non-finite residual aborts,
// one above the tolerance warns.
static void report_solve(IBMSolver *s, int sk, double rel, int it) {
    s->final_norm[sk] = rel;
    s->final_iter[sk] = it;
    if (!std::isfinite(rel)) {
        if (sk < 3) case_error("error: floating pointer error raised in u" + std::to_string(sk + 1) + "_star");
        case_error("error: floating pointer error raised in p_prime");
    }
    if (rel > s->tol) {
        if (sk < 3) fprintf(stderr, "warning: u%d_star did not converge\n", sk + 1);
        else fprintf(stderr, "warning: p_prime did not converge\n");
    }
}

// ---------------------------------------------------------------------------
// The three momentum solves at once (IBM3D_PREC_FP64 and _MIXED_KRYLOV):
// BiCGSTAB with Jacobi preconditioning from x0 = 0, component 0 on the
// solver's stream and buffers, components 1 and 2 on lanes of their own.
// The host polls the lanes' published scalars and enqueues a component's next
// iteration as soon as its own are in, so the components' kernels, reductions
// and host reads overlap on the device and no lane waits for another. Per component: the kernels and decisions
// of bicgstab<double> with prec 1 and no restarts.
// ---------------------------------------------------------------------------
// norms(bnorm) gives the three stopping norms (a host read of the step's
// sums), called once the initialisations are enqueued so that the read
// overlaps them.
template <typename F>
static void momentum_concurrent(IBMSolver *s, F &&norms) {
    const Dims &d = s->d;
    const GhostRows<double> &G = s->gmom.G64;
    const uint8_t *flag = s->flag.p;
    const long nv = vec_partials(d), na = apply_partials(d, G.ng);
    const NormSpec ns{0, nullptr, nullptr, nullptr, nullptr, -1};
    const int maxit = 200;  // src/ibm3d_setup.c:1413
    struct Lane {
        hipStream_t st;
        DBuf<double> *V;
        KScal *ks;
        double *partial;
        KPub *pub_h, *pub_d;
        unsigned *epoch;
        float *da, *db;  // FP32 preconditioned directions (mixed-Krylov mode)
    };
    // mixed-Krylov mode: preconditioned directions stored in FP32 (as in bicgstab)
    const bool f32 = s->precision == IBM3D_PREC_MIXED_KRYLOV;
    Lane ln[3] = {{s->stream, s->kry.v64, s->ks.p, s->partial.p, s->pub_h, s->pub_d, &s->pub_epoch,
                   f32 ? s->ref_e32.p : nullptr, f32 ? s->ref_e32b.p : nullptr}, {}, {}};
    for (int c = 1; c < 3; c++) {
        MomLane &L = s->mlane[c - 1];
        ln[c] = {L.st, L.v, L.ks.p, L.partial.p, L.pub_h, L.pub_d, &L.epoch, f32 ? L.d32[0].p : nullptr,
                 f32 ? L.d32[1].p : nullptr};
    }
    // the right-hand sides were formed on the solver's stream
    HIP_CHECK(hipEventRecord(s->ev_rhs, s->stream));
    for (int c = 1; c < 3; c++) HIP_CHECK(hipStreamWaitEvent(ln[c].st, s->ev_rhs, 0));
    KScal kh[3];
    unsigned ep[3];
    auto publish = [&](int c, int op, long nblk, int nvals) {
        ep[c] = ++*ln[c].epoch;
        launch_finalize(ln[c].st, op, ln[c].partial, nblk, nvals, ln[c].ks, &ln[c].pub_d->ks, &ln[c].pub_d->epoch,
                        ep[c]);
    };
    auto collect = [&](int c) {
        wait_published(ln[c].st, ln[c].pub_h, ep[c]);
        memcpy((void *)&kh[c], (const void *)&ln[c].pub_h->ks, sizeof(KScal));
    };
    double bn[3], rn[3];
    int it[3] = {0, 0, 0};
    bool act[3];
    for (int c = 0; c < 3; c++) {
        DBuf<double> *V = ln[c].V;
        // p and v need no clearing: the first direction is formed as p = r
        launch_bicg_init<double>(ln[c].st, d, V[0].p, V[1].p, nullptr, V[3].p, s->us[c].p, s->b[c].p, ns,
                                 ln[c].partial);
        publish(c, 3, nv, 2);
    }
    double bnorm[3];
    norms(bnorm);
    for (int c = 0; c < 3; c++) {
        collect(c);
        bn[c] = bnorm[c] > 0 ? bnorm[c] : std::sqrt(kh[c].bb);
        rn[c] = std::sqrt(kh[c].rr);
        act[c] = bn[c] != 0;
    }
    const bool zero[3] = {!act[0], !act[1], !act[2]};
    // one BiCGSTAB iteration of component c, enqueued on its lane
    auto enqueue_iteration = [&](int c) {
        it[c]++;
        const SysCoef<double> &S = s->sys[c].S64;
        DBuf<double> *V = ln[c].V;
        double *r = V[0].p, *rh = V[1].p, *p = V[2].p, *v = V[3].p, *ph = V[4].p, *sv = V[5].p, *sh = V[6].p,
               *t = V[7].p;
        const hipStream_t st = ln[c].st;
        const int wp = it[c] == 1 ? 2 : 0;  // the first direction: p = r
        if (f32) {
            launch_bicg_jacobi_f32(st, d, S, G, flag, ln[c].ks, wp, r, v, p, ln[c].da);
            launch_apply_f32in(st, d, S, G, flag, ln[c].da, nullptr, v, 1, rh, ln[c].partial);
        } else {
            launch_bicg_jacobi<double>(st, d, S, G, flag, ln[c].ks, wp, r, v, p, ph);
            launch_apply<double>(st, d, S, G, flag, ph, v, 1, rh, ln[c].partial);
        }
        launch_finalize(st, 0, ln[c].partial, na, 1, ln[c].ks);
        if (f32) {
            launch_bicg_jacobi_f32(st, d, S, G, flag, ln[c].ks, 1, r, v, sv, ln[c].db);
            launch_apply_f32in(st, d, S, G, flag, ln[c].db, nullptr, t, 2, sv, ln[c].partial);
        } else {
            launch_bicg_jacobi<double>(st, d, S, G, flag, ln[c].ks, 1, r, v, sv, sh);
            launch_apply<double>(st, d, S, G, flag, sh, t, 2, sv, ln[c].partial);
        }
        launch_finalize(st, 1, ln[c].partial, na, 2, ln[c].ks);
        if (f32)
            launch_bicg_xr_f32dir(st, d, ln[c].ks, ln[c].da, ln[c].db, sv, t, rh, s->us[c].p, r, ns, ln[c].partial);
        else
            launch_bicg_xr<double>(st, d, ln[c].ks, ph, sh, sv, t, rh, s->us[c].p, r, ns, ln[c].partial);
        publish(c, 2, nv, 3);
    };
    // a component goes on while it is not converged, not broken down and
    // under maxit
    auto goes_on = [&](int c) { return act[c] && it[c] < maxit && rn[c] > s->tol * bn[c]; };
    // Each lane's next iteration is enqueued as soon as that lane's scalars
    // are in, so the lanes do not wait for one another; a component's own
    // kernels and decisions stay those of the one-at-a-time loop.
    bool fly[3] = {false, false, false};
    for (int c = 0; c < 3; c++) {
        act[c] = goes_on(c);
        if (act[c]) { enqueue_iteration(c); fly[c] = true; }
    }
    for (long n = 1;; n++) {
        bool any = false;
        for (int c = 0; c < 3; c++) {
            if (!fly[c]) continue;
            any = true;
            if (__atomic_load_n(&ln[c].pub_h->epoch, __ATOMIC_ACQUIRE) != ep[c]) {
                // now and then: an idle lane stream with nothing published
                // (or an error) ends the wait instead of making it endless
                if ((n & 1023) == 0) {
                    const hipError_t q = hipStreamQuery(ln[c].st);
                    if (q == hipSuccess && __atomic_load_n(&ln[c].pub_h->epoch, __ATOMIC_ACQUIRE) != ep[c]) {
                        char buf[120];
                        snprintf(buf, sizeof buf, "error: Krylov scalars were not published (epoch %u, expected %u)",
                                 __atomic_load_n(&ln[c].pub_h->epoch, __ATOMIC_ACQUIRE), ep[c]);
                        case_error(buf);
                    }
                    if (q != hipSuccess && q != hipErrorNotReady) HIP_CHECK(q);
                }
                continue;
            }
            collect(c);
            fly[c] = false;
            rn[c] = std::sqrt(kh[c].rr);
            if (kh[c].breakdown || !std::isfinite(rn[c])) act[c] = false;
            act[c] = goes_on(c);
            if (act[c]) { enqueue_iteration(c); fly[c] = true; }
        }
        if (!any) break;
        wait_pause(n);
    }
    // true residuals of the final iterates
    for (int c = 0; c < 3; c++) {
        if (zero[c]) continue;
        launch_residual<double>(ln[c].st, d, s->sys[c].S64, G, flag, s->b[c].p, s->us[c].p, ln[c].V[0].p, ns,
                                ln[c].partial);
        publish(c, 4, na, 2);
    }
    double rel[3] = {0, 0, 0};
    for (int c = 0; c < 3; c++) {
        if (zero[c]) continue;
        collect(c);
        rel[c] = std::sqrt(kh[c].rr) / bn[c];
    }
    // the solver's stream continues after the lanes' work
    for (int c = 1; c < 3; c++) {
        HIP_CHECK(hipEventRecord(s->mlane[c - 1].done, ln[c].st));
        HIP_CHECK(hipStreamWaitEvent(s->stream, s->mlane[c - 1].done, 0));
    }
    for (int c = 0; c < 3; c++) report_solve(s, c, rel[c], it[c]);
}

static void solve_system(IBMSolver *s, int sk, const double *b, double *x, double bnorm) {
    const int prec = sk == 3 ? (s->precond_type == PRECOND_AMG ? 2 : 1) : 1;
    NormSpec ns{0, nullptr, nullptr, nullptr, nullptr, -1};
    if (sk == 3) ns = {1, s->ks1d[0].p, s->ks1d[1].p, s->ks1d[2].p, s->flag.p, s->pin};
    const int maxit = sk == 3 ? 1000 : 200;  // src/ibm3d_setup.c:1413, 1467
    double rel = 0;
    int it;
    if (sk == 3 && s->solver_type != SOLVER_BiCGSTAB && s->precision != IBM3D_PREC_MIXED) {
        // multigrid cycles in FP32 in the mixed-Krylov mode; SOLVER_AMG is the
        // cycle itself, with no separate preconditioner
        const int pp = s->solver_type == SOLVER_AMG ? 2 : prec;
        const int p2 = pp == 2 && s->precision == IBM3D_PREC_MIXED_KRYLOV ? 3 : pp;
        if (s->solver_type == SOLVER_GMRES) it = gmres_p(s, b, x, p2, ns, s->tol, maxit, bnorm, &rel);
        else if (s->solver_type == SOLVER_PCG) it = pcg_p(s, b, x, p2, ns, s->tol, maxit, bnorm, &rel);
        else it = mg_solve_p(s, b, x, p2, ns, s->tol, maxit, bnorm, &rel);
    } else if (s->precision == IBM3D_PREC_MIXED) {
        it = refine_mixed(s, sk, b, x, prec, ns, s->tol, maxit, bnorm, &rel);
    } else if (s->precision == IBM3D_PREC_MIXED_KRYLOV) {
        const GhostData &GD = sk == 3 ? s->gp : s->gmom;
        // pressure with the multigrid preconditioner: FP32 cycles (prec 3),
        // or the FP64 cycle (prec 2) once a solve has stalled kStallFallback
        // times; restarts only with the FP32 cycles
        const bool mg32 = prec == 2 && !s->p_mg64;
        int stalls = 0;
        it = bicgstab<double>(s, s->sys[sk].S64, GD.G64, b, x, mg32 ? 3 : prec, ns, s->tol, maxit, bnorm, &rel,
                              sk == 3 ? "p64" : "u64", mg32 ? kMixedRestart : 0, &stalls);
        if (mg32 && stalls >= kStallFallback) s->p_mg64 = true;
    } else {
        const GhostData &GD = sk == 3 ? s->gp : s->gmom;
        it = bicgstab<double>(s, s->sys[sk].S64, GD.G64, b, x, prec, ns, s->tol, maxit, bnorm, &rel,
                              sk == 3 ? "p64" : "u64");
    }
    report_solve(s, sk, rel, it);
}

// ---------------------------------------------------------------------------
// Momentum solves of the mixed-Krylov mode: FP32 red-black Gauss-Seidel
// sweeps inside FP64 iterative refinement. The Helmholtz operator
// (1 + sum K) I - sum K_nb is strongly diagonally dominant wherever
// dt / (Re h^2) is small: a Jacobi sweep then contracts fluid rows by
// rho_J = max sum K / (1 + sum K) and a red-black sweep by about rho_J^2, so
// m sweeps per round reduce the correction's residual by rho_J^(2m) until FP32
// rounding (~1e-7) limits it. Each round: FP32 right-hand side (the FP64
// residual; the first colour written in the same launch), m sweeps from zero,
// x += e in FP64, FP64 residual and its norm. The stopping test is the
// reference's ||b - A x|| <= tol ||b|| in FP64. Where rho_J^2 > 0.03 (few
// sweeps would not do) the BiCGSTAB path solves instead, as it does for a
// component whose residual stops falling by at least 10x per round.
// ---------------------------------------------------------------------------
static const double kMomRoundTarget = 1e-6;  // residual reduction aimed at per round
static const double kMomMaxRate = 0.03;      // largest rho_J^2 for which sweeps are used
static const int kMomMaxRounds = 8;

static int momentum_sweeps(IBMSolver *s) {
    if (s->mom_sweeps >= 0) return s->mom_sweeps;
    double mx[3] = {0, 0, 0};
    for (int i = 0; i < s->Nx; i++) mx[0] = std::max(mx[0], H1(s->kW, i) + H1(s->kE, i));
    for (int j = 0; j < s->Ny; j++) mx[1] = std::max(mx[1], H1(s->kS, j) + H1(s->kN, j));
    for (int k = 0; k < s->Nz; k++) mx[2] = std::max(mx[2], H1(s->kD, k) + H1(s->kU, k));
    const double sk = mx[0] + mx[1] + mx[2], rj = sk / (1 + sk), rate = rj * rj;
    s->mom_rate = rate;
    s->mom_sweeps = (rate > 0 && rate <= kMomMaxRate)
                        ? std::max(2, (int)std::ceil(std::log(kMomRoundTarget) / std::log(rate))) : 0;
    return s->mom_sweeps;
}

// Sweeps for a round that must reduce the residual by `need` (< 1): enough
// at the estimated rate with a factor 2 to spare, at least one, at most the
// full round's count.
static int momentum_round_sweeps(const IBMSolver *s, double need) {
    const double f = std::max(need, kMomRoundTarget) * 0.5;
    return std::min(s->mom_sweeps, std::max(1, (int)std::ceil(std::log(f) / std::log(s->mom_rate))));
}

// The three momentum solves by refinement; components that do not converge
// this way are solved by solve_system. bn: the stopping norms.
static void momentum_refine(IBMSolver *s, const double bn[3]) {
    const Dims &d = s->d;
    const int m = momentum_sweeps(s);
    const double cb = ncells(s);
    const long na = apply_partials(d, s->gmom.G64.ng);
    const NormSpec off{0, nullptr, nullptr, nullptr, nullptr, -1};
    double *r64[3] = {s->ref_r64.p, s->mlane[0].v[0].p, s->mlane[1].v[0].p};
    float *rhs32[3] = {s->ref_rhs32.p, s->mlane[0].d32[0].p, s->mlane[1].d32[0].p};
    float *e32[3] = {s->ref_e32.p, s->mlane[0].d32[1].p, s->mlane[1].d32[1].p};
    double rn[3] = {0, 0, 0};
    int sweeps[3] = {0, 0, 0};
    bool active[3], fallback[3] = {false, false, false};
    for (int l = 0; l < 3; l++) {
        // a zero right-hand side has the zero solution; a non-finite one
        // stays active, so its residual reports the error
        active[l] = bn[l] != 0;
        if (!active[l]) {
            launch_zero<double>(s->stream, d, s->us[l].p);
            report_solve(s, l, 0.0, 0);
        }
    }
    for (int round = 0; round < kMomMaxRounds; round++) {
        bool any = false;
        for (int l = 0; l < 3; l++) {
            if (!active[l]) continue;
            any = true;
            const SysCoef<double> &S64 = s->sys[l].S64;
            const SysCoef<float> &S32 = s->sys[l].S32;
            // right-hand side of the correction (b itself from x = 0) in FP32,
            // with the first smoother colour from a zero start
            // the first round aims at the FP32 limit, later ones at what is left
            const int mr = round ? momentum_round_sweeps(s, s->tol * bn[l] / rn[l]) : m;
            TIMED("mom_d2f", cb * 17, launch_d2f_first(s->stream, d, S32, s->flag.p, round ? r64[l] : s->b[l].p,
                                                       rhs32[l], e32[l]));
            smooth<float>(s, d, S32, s->gmom.G32, s->flag.p, rhs32[l], e32[l], s->ref_e32b.p, mr, true, -1, true);
            sweeps[l] += mr;
            if (round) TIMED("mom_add_f2d", cb * 20, launch_add_f2d(s->stream, d, e32[l], s->us[l].p));
            else TIMED("mom_f2d", cb * 12, launch_f2d(s->stream, d, e32[l], s->us[l].p));
            TIMED("mom_residual64", 3 * cb * 8,
                  launch_residual<double>(s->stream, d, S64, s->gmom.G64, s->flag.p, s->b[l].p, s->us[l].p, r64[l],
                                          off, s->partial.p));
            launch_sum_partials(s->stream, s->partial.p, na, 2, s->sums.p + 3 + 2 * l, 0);
        }
        if (!any) break;
        read_sums(s, 9);
        for (int l = 0; l < 3; l++) {
            if (!active[l]) continue;
            const double r = std::sqrt(s->sums_h[3 + 2 * l]);
            const bool stalled = round > 0 && !(r <= 0.1 * rn[l]);
            rn[l] = r;
            if (!std::isfinite(r) || r <= s->tol * bn[l]) {
                active[l] = false;
                report_solve(s, l, r / bn[l], sweeps[l]);
            } else if (stalled) {
                active[l] = false;
                fallback[l] = true;
            }
        }
    }
    for (int l = 0; l < 3; l++)
        if (active[l] || fallback[l]) solve_system(s, l, s->b[l].p, s->us[l].p, bn[l]);
}

// ---------------------------------------------------------------------------
// one fractional step
// ---------------------------------------------------------------------------
static StepParams step_params(IBMSolver *s) {
    StepParams sp{};
    sp.dt = s->dt;
    sp.Re = s->Re;
    for (int l = 0; l < 3; l++) { sp.force[l] = s->const_f[l]; sp.fa[l] = nullptr; }
    if (s->force_type == VAL_FUNC) {
        // host-evaluated force over the interior cells (src/ibm3d_fracstep.c:266)
        for (int l = 0; l < 3; l++) {
            std::vector<double> h(s->d.ncell, 0.0);
            for (int k = 0; k < s->Nz; k++)
                for (int j = 0; j < s->Ny; j++)
                    for (int i = 0; i < s->Nx; i++)
                        h[cidx(s->d, i, j, k)] = s->func_f[l](s->time, H1(s->xc, i), H1(s->yc, j), H1(s->zc, k));
            HIP_CHECK(hipStreamSynchronize(s->stream));  // previous readers of the buffer
            s->force_arr[l].upload(h);
            sp.fa[l] = s->force_arr[l].p;
        }
    }
    return sp;
}

static TauPtrs tau_ptrs(IBMSolver *s) {
    TauPtrs tp{};
    tp.on = s->turb_type != TURBMODEL_NONE;
    static const int slot[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int a = 0; a < 6; a++) tp.t[a] = s->tau[a].p;
    for (int l = 0; l < 3; l++)
        for (int m = 0; m < 3; m++) tp.slot[l][m] = slot[l][m];
    return tp;
}

static void calc_tau(IBMSolver *s) {
    if (s->turb_type == TURBMODEL_NONE) return;
    const double *u3[3] = {s->u[0].p, s->u[1].p, s->u[2].p};
    VanDriest vd{};
    const bool smag = s->turb_type == TURBMODEL_SMAGORINSKY || s->turb_type == TURBMODEL_SMAGORINSKY_VD;
    if (s->turb_type == TURBMODEL_SMAGORINSKY_VD) {
        vd.walls = s->vd_walls;
        vd.aplus = s->vd_aplus;
        vd.ure = s->vd_utau * s->Re;
        vd.lo[0] = s->xmin; vd.lo[1] = s->ymin; vd.lo[2] = s->zmin;
        vd.hi[0] = s->xmax; vd.hi[1] = s->ymax; vd.hi[2] = s->zmax;
    }
    TIMED("sgs", ncells(s) * (3 + 6) * 8 + ncells(s),
          launch_sgs(s->stream, s->geo, smag ? 1 : 2, s->turb_coef, vd, s->flag.p, s->lvset.p, u3, tau_ptrs(s)));
}

static void do_step(IBMSolver *s) {
    const StepParams sp = step_params(s);
    const double cn = ncells(s);
    calc_tau(s);
    // K1: convection + momentum RHS
    {
        const double *u3[3] = {s->u[0].p, s->u[1].p, s->u[2].p};
        const double *U3[3] = {s->F[0].p, s->F[1].p, s->F[2].p};
        const double *Np[3] = {s->Np[0].p, s->Np[1].p, s->Np[2].p};
        double *N[3] = {s->N[0].p, s->N[1].p, s->N[2].p};
        double *b[3] = {s->b[0].p, s->b[1].p, s->b[2].p};
        TIMED("k1_conv_rhs", cn * (16 * 8 + 1),
              launch_conv_rhs(s->stream, s->geo, sp, s->flag.p, u3, U3, s->p.p, Np, N, b, tau_ptrs(s),
                              s->partial.p));
        launch_sum_partials(s->stream, s->partial.p, conv_rhs_blocks(s->d), 3, s->sums.p, 0);
    }
    const bool conv = conv_outflow(s, 0);
    double fnorm[4];
    const OuterParams ob = make_outer(s, BCD_USOL, s->time, fnorm);
    for (int l = 0; l < 3; l++)
        for (int ax = 0; ax < 3; ax++)
            for (int hi = 0; hi < 2; hi++) {
                const double gm = s->sys[l].gamma[ax][hi];
                const IBMSolverBCType t = s->bc[face_bc(ax, hi)].type;
                if (gm == 0 || (t != BC_VELOCITY_COMPONENT && t != BC_CONVECTIVE_OUTFLOW)) continue;
                const double *Kf = ax == 0 ? (hi ? s->g1d[7].p : s->g1d[6].p)
                                 : ax == 1 ? (hi ? s->g1d[9].p : s->g1d[8].p) : (hi ? s->g1d[11].p : s->g1d[10].p);
                launch_bc_rhs(s->stream, s->geo, ob, ax, hi, l, Kf, gm, s->b[l].p);
            }
    // stopping norms of the momentum solves: the right-hand sides' sums
    // (host read) plus the BC rows' and forcing parts
    auto mom_norms = [&](double bn[3]) {
        read_sums(s, conv ? kConvNormSlot + 18 : 3);
        for (int l = 0; l < 3; l++) {
            double bn2 = s->sums_h[l] + s->sys[l].bc_norm2 + fnorm[l];
            for (int f = 0; conv && f < 6; f++)
                if (s->bc[f].type == BC_CONVECTIVE_OUTFLOW) bn2 += s->sums_h[kConvNormSlot + 3 * f + l];
            bn[l] = std::sqrt(bn2);
        }
    };
    // momentum Helmholtz solves: the three at once on their lanes, or one
    // after the other (mixed defect-correction mode, and under the per-kernel
    // timers, whose events are recorded on the solver's stream)
    if (s->precision == IBM3D_PREC_MIXED_KRYLOV && s->mom_solver == 0 && momentum_sweeps(s) > 0) {
        double bn[3];
        mom_norms(bn);
        momentum_refine(s, bn);
    } else if (s->mom_lanes && s->timing == 0 && s->precision != IBM3D_PREC_MIXED) {
        momentum_concurrent(s, mom_norms);
    } else {
        double bn[3];
        mom_norms(bn);
        for (int l = 0; l < 3; l++) solve_system(s, l, s->b[l].p, s->us[l].p, bn[l]);
    }
    // outer layers of the three solutions from their BC rows (the per-component
    // fill of each u* with the solve's boundary data, all three in one pass)
    TIMED("outer_sol", 0, launch_outer(s->stream, s->geo, ob, s->us[0].p, s->us[1].p, s->us[2].p, 0, 0));
    // K2: U*, Poisson RHS
    {
        const double *us3[3] = {s->us[0].p, s->us[1].p, s->us[2].p};
        double *Fs[3] = {s->Fs[0].p, s->Fs[1].p, s->Fs[2].p};
        TIMED("k2_face_star", cn * (4 * 8 + 3 * 8 + 1), launch_face_star(s->stream, s->geo, sp, s->flag.p, us3, s->p.p, Fs));
        const double *Fc[3] = {s->Fs[0].p, s->Fs[1].p, s->Fs[2].p};
        TIMED("k2_poisson_rhs", cn * (4 * 8 + 1),
              launch_poisson_rhs(s->stream, s->geo, sp, s->flag.p, s->pin, Fc, s->bp.p, s->partial.p));
        launch_sum_partials(s->stream, s->partial.p, conv_rhs_blocks(s->d), 1, s->sums.p, 0);
    }
    double pnorm[4];
    const OuterParams opp = make_outer(s, BCD_PSOL, s->time, pnorm);
    for (int ax = 0; ax < 3; ax++)
        for (int hi = 0; hi < 2; hi++) {
            const IBMSolverBCType t = s->bc[face_bc(ax, hi)].type;
            if ((t != BC_PRESSURE && t != BC_VELOCITY_PERIODIC) || s->bc[face_bc(ax, hi)].val_type != VAL_FUNC) continue;
            const double *Kf = ax == 0 ? (hi ? s->g1d[7].p : s->g1d[6].p)
                             : ax == 1 ? (hi ? s->g1d[9].p : s->g1d[8].p) : (hi ? s->g1d[11].p : s->g1d[10].p);
            launch_bc_rhs(s->stream, s->geo, opp, ax, hi, 3, Kf, s->sys[3].gamma[ax][hi], s->bp.p);
        }
    read_sums(s, 1);
    const double bnp = std::sqrt(s->sums_h[0] + s->sys[3].bc_norm2 + pnorm[3]);
    solve_system(s, 3, s->bp.p, s->pp.p, bnp);
    TIMED("outer_sol", 0, launch_outer(s->stream, s->geo, opp, s->pp.p, nullptr, nullptr, 3, 0));
    // K3: projection
    {
        const double *us3[3] = {s->us[0].p, s->us[1].p, s->us[2].p};
        const double *Fs[3] = {s->Fs[0].p, s->Fs[1].p, s->Fs[2].p};
        double *un[3] = {s->un[0].p, s->un[1].p, s->un[2].p};
        double *Fn[3] = {s->Fn[0].p, s->Fn[1].p, s->Fn[2].p};
        TIMED("k3_update", cn * (15 * 8 + 1),
              launch_update(s->stream, s->geo, sp, s->flag.p, s->p.p, s->pp.p, us3, Fs, s->pn.p, un, Fn));
    }
    // pointer swaps  <- src/ibm3d_fracstep.c:795-804
    for (int l = 0; l < 3; l++) {
        std::swap(s->u[l], s->un[l]);
        std::swap(s->F[l], s->Fn[l]);
        std::swap(s->Np[l], s->N[l]);
    }
    std::swap(s->p, s->pn);
}

void iterate_impl(IBMSolver *s, int steps, bool verbose) {
    const double t0 = s->time;
    if (!s->conv_ready) conv_outflow(s, 1);
    s->conv_ready = true;
    // prologue  <- src/ibm3d_fracstep.c:46-54
    update_outer(s);
    launch_interp_stag(s->stream, s->geo, s->u[0].p, s->u[1].p, s->u[2].p, s->F[0].p, s->F[1].p, s->F[2].p);
    launch_solid_nan(s->stream, s->d, s->flag.p, s->u[0].p, s->u[1].p, s->u[2].p, s->p.p);
    update_ghost(s);
    {
        const StepParams sp = step_params(s);
        const double *u3[3] = {s->u[0].p, s->u[1].p, s->u[2].p};
        const double *U3[3] = {s->F[0].p, s->F[1].p, s->F[2].p};
        const double *Np[3] = {s->Np[0].p, s->Np[1].p, s->Np[2].p};
        double *N[3] = {s->N[0].p, s->N[1].p, s->N[2].p};
        double *b[3] = {s->b[0].p, s->b[1].p, s->b[2].p};
        TauPtrs tp{};
        tp.on = 0;
        launch_conv_rhs(s->stream, s->geo, sp, s->flag.p, u3, U3, s->p.p, Np, N, b, tp, s->partial.p);
        for (int l = 0; l < 3; l++) std::swap(s->N[l], s->Np[l]);
    }
    // step times: events recorded on the stream at each step's end (the host
    // runs ahead into the next step instead of draining the stream)
    s->step_ms.clear();
    s->force_hist_n = 0;
    std::vector<hipEvent_t> ends;
    hipEvent_t start = get_event(s);
    HIP_CHECK(hipEventRecord(start, s->stream));
    for (int i = 0; i < steps;) {
        do_step(s);
        i++;
        s->iter++;
        s->time = t0 + s->dt * i;
        update_outer(s);
        update_ghost(s);
        if (i <= s->force_hist_max) {  // body force of the step's fields, kept on the device
            body_force_impl(s, s->force_hist.p + 3 * (i - 1));
            s->force_hist_n = i;
        }
        ends.push_back(get_event(s));
        HIP_CHECK(hipEventRecord(ends.back(), s->stream));
        if (verbose) {
            if (s->iter % 10 == 1 || i == 1) {
                printf("\n  iter       u1 res       u2 res       u3 res        p res   its(u1 u2 u3 p)\n");
                printf("-------------------------------------------------------------------------------\n");
            }
            printf("%6d   %10.4e   %10.4e   %10.4e   %10.4e   %d %d %d %d\n", s->iter, s->final_norm[0],
                   s->final_norm[1], s->final_norm[2], s->final_norm[3], s->final_iter[0], s->final_iter[1],
                   s->final_iter[2], s->final_iter[3]);
            fflush(stdout);
        }
        if (s->autosave_period > 0 && s->iter % s->autosave_period == 0 && s->autosave_filename) {
            char name[256];
            snprintf(name, sizeof name, "%s-%05d", s->autosave_filename, s->iter);  // :847-853
            export_result_impl(s, name);
        }
    }
    HIP_CHECK(hipStreamSynchronize(s->stream));
    hipEvent_t prev = start;
    for (hipEvent_t e : ends) {
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, prev, e));
        s->step_ms.push_back(ms);
        s->event_pool.push_back(prev);
        prev = e;
    }
    s->event_pool.push_back(prev);
}

// Body force from the current u and p into dst[0..3) on the device (two
// fixed-order reduction stages).
void body_force_impl(IBMSolver *s, double *dst) {
    const double *u[3] = {s->u[0].p, s->u[1].p, s->u[2].p};
    launch_body_force(s->stream, s->geo, s->flag.p, s->ghost_ijk.p, s->ng, 1.0 / s->Re, s->p.p, u, s->partial.p);
    launch_sum_partials(s->stream, s->partial.p, body_force_blocks(s->ng), 3, dst, 0);
}

void body_force_read(IBMSolver *s, double out[3]) {
    body_force_impl(s, s->sums.p);
    read_sums(s, 3);
    for (int a = 0; a < 3; a++) out[a] = s->sums_h[a];
}

double max_div_impl(IBMSolver *s) {
    const double *U3[3] = {s->F[0].p, s->F[1].p, s->F[2].p};
    launch_maxdiv(s->stream, s->geo, s->flag.p, s->pin, U3, s->partial.p);
    launch_sum_partials(s->stream, s->partial.p, maxdiv_blocks(s->d), 1, s->sums.p, 1);
    read_sums(s, 1);
    return s->sums_h[0];
}

}  // namespace ibm

using namespace ibm;

extern "C" {

void IBMSolver_enable_timing(IBMSolver *s, int mode) {
    if (mode < 0 || mode > 2) fatal("timing mode must be 0, 1 or 2");
    s->timing = mode;
}

void IBMSolver_set_graph_replay(IBMSolver *s, int on) { s->graphs = on != 0; }

void IBMSolver_set_momentum_solver(IBMSolver *s, int mode) {
    if (mode < 0 || mode > 1) fatal("momentum solver mode must be 0 or 1");
    s->mom_solver = mode;
}

void IBMSolver_reset_timing(IBMSolver *s) {
    timers_flush(s);
    s->timers.clear();
}

int IBMSolver_get_kernel_stat(IBMSolver *s, int idx, char *name, int namelen, double *total_ms, long *count,
                              double *bytes) {
    timers_flush(s);
    if (idx >= 0 && idx < (int)s->timers.size()) {
        const KTimer &t = s->timers[idx];
        if (name && namelen > 0) { strncpy(name, t.name.c_str(), namelen - 1); name[namelen - 1] = 0; }
        if (total_ms) *total_ms = t.total_ms;
        if (count) *count = t.count;
        if (bytes) *bytes = t.count > 0 ? t.bytes / t.count : 0.0;  // average per launch
    }
    return (int)s->timers.size();
}

void IBMSolver_calc_tau_r(IBMSolver *s) {
    if (!s->assembled) fatal("IBMSolver_assemble must precede IBMSolver_calc_tau_r");
    calc_tau(s);
}

void IBMSolver_set_force_history(IBMSolver *s, int max_steps) {
    if (!s->assembled) fatal("IBMSolver_assemble must precede IBMSolver_set_force_history");
    if (max_steps < 0) fatal("IBMSolver_set_force_history: max_steps must be >= 0");
    HIP_CHECK(hipStreamSynchronize(s->stream));
    s->force_hist = DBuf<double>();
    if (max_steps > 0) s->force_hist.alloc(3 * (size_t)max_steps);
    s->force_hist_max = max_steps;
    s->force_hist_n = 0;
}

int IBMSolver_get_force_history(IBMSolver *s, double *out, int max) {
    const int n = s->force_hist_n;
    if (out && n > 0 && max > 0) {
        HIP_CHECK(hipStreamSynchronize(s->stream));
        HIP_CHECK(hipMemcpy(out, s->force_hist.p, 3 * (size_t)std::min(n, max) * sizeof(double),
                            hipMemcpyDeviceToHost));
    }
    return n;
}

int IBMSolver_get_step_times(IBMSolver *s, double *out, int max) {
    const int n = (int)s->step_ms.size();
    for (int i = 0; out && i < n && i < max; i++) out[i] = s->step_ms[i];
    return n;
}

}  // extern "C"
