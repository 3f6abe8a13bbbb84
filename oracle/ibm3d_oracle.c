/*
 * ibm3d_oracle.c - FP64 CPU oracle of the fractional-step IBM time step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Clean-room restatement of the
 * reference solver for one process (Px = Py = Pz = 1). Each function cites
 * the reference lines it restates. The reference's Hypre solves are replaced
 * by a restatement of their arithmetic contract: the full padded linear system
 * of src/ibm3d_setup.c:1566-2908 is assembled row by row and solved by
 * BiCGSTAB from x0 = 0 until ||b - A x||_2 <= tol ||b||_2, with ILU(0) standing
 * in for BoomerAMG as the preconditioner when PRECOND_AMG is requested (only
 * iteration counts depend on the preconditioner, not the converged answer).
 *
 * Deviations, both unobservable in u, p and U:
 *  - the East free-slip second ghost layer is written only for interior (j,k)
 *    (src/ibm3d_fracstep.c:1045-1046 also writes the y/z ghost corners, which
 *    the South/North/Down/Up passes overwrite right after);
 *  - mirror-point base indices are clamped to the padded grid (the reference
 *    reads out of bounds there, src/utils.c:137-139).
 */
#include "oracle.h"

#include <stdarg.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const int ADJ[6][3] = {
    {0, 1, 0}, {1, 0, 0}, {0, -1, 0}, {-1, 0, 0}, {0, 0, -1}, {0, 0, 1}
};  /* src/ibm3d_setup.c:12-14 */

static void fatal(const char *msg) {
    fprintf(stderr, "error: %s\n", msg);
    fflush(stdout);
    abort();
}

static double *dalloc(size_t n) {
    double *p = calloc(n ? n : 1, sizeof(double));
    if (!p) fatal("out of memory");
    return p;
}

static size_t ncell(const IBMSolver *s) {
    return (size_t)(s->Nx + 4) * (s->Ny + 4) * (s->Nz + 4);
}

/* ------------------------------------------------------------------------ */
/* construction and setters  <- src/ibm3d_setup.c:48-491                     */
/* ------------------------------------------------------------------------ */

IBMSolver *IBMSolver_new(const int num_process, const int rank, const int Px,
                         const int Py, const int Pz) {
    if (num_process != 1 || rank != 0 || Px * Py * Pz != 1)
        fatal("the oracle runs one process (Px = Py = Pz = 1)");
    IBMSolver *s = calloc(1, sizeof(IBMSolver));
    s->force_type = VAL_CONST;  /* src/ibm3d_setup.c:96-101 */
    s->turb_type = TURBMODEL_NONE;
    s->solver_type = SOLVER_BiCGSTAB;
    s->precond_type = PRECOND_NONE;
    s->tol = 1e-6;
    return s;
}

static void free_mat(EllMat *A) {
    free(A->nnz); free(A->col); free(A->val);
    free(A->scol); free(A->lu); free(A->diagpos);
    memset(A, 0, sizeof(*A));
}

void IBMSolver_destroy(IBMSolver *s) {
    if (!s) return;
    for (int i = 0; i < 6; i++) free(s->conv_g[i]);
    double **arrs[] = {
        &s->dx, &s->dy, &s->dz, &s->xc, &s->yc, &s->zc,
        &s->kx_W, &s->kx_E, &s->ky_S, &s->ky_N, &s->kz_D, &s->kz_U,
        &s->lvset,
        &s->u1, &s->u1_next, &s->u1_star, &s->u1_tilde,
        &s->u2, &s->u2_next, &s->u2_star, &s->u2_tilde,
        &s->u3, &s->u3_next, &s->u3_star, &s->u3_tilde,
        &s->U1, &s->U1_next, &s->U1_star, &s->U2, &s->U2_next, &s->U2_star,
        &s->U3, &s->U3_next, &s->U3_star,
        &s->p, &s->p_next, &s->p_prime,
        &s->N1, &s->N1_prev, &s->N2, &s->N2_prev, &s->N3, &s->N3_prev,
        &s->p_coeffsum, &s->b, &s->x, &s->work,
    };
    for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); i++) free(*arrs[i]);
    for (int i = 0; i < 6; i++) { free(s->S[i]); free(s->tau[i]); }
    free(s->flag);
    free(s->cell_idx_periodic);
    for (int i = 0; i < 4; i++) free_mat(&s->A[i]);
    free(s);
}

/* <- src/ibm3d_setup.c:203-284 (single process: local == global) and
   alloc_arrays :641-729 */
void IBMSolver_set_grid(IBMSolver *s, const int Nx, const int Ny, const int Nz,
                        const double *xf, const double *yf, const double *zf) {
    if (Nx < 2 || Ny < 2 || Nz < 2) fatal("grid needs at least 2 cells per axis");
    s->Nx = Nx; s->Ny = Ny; s->Nz = Nz;
    const size_t P = ncell(s);
    s->dx = dalloc(Nx + 4); s->dy = dalloc(Ny + 4); s->dz = dalloc(Nz + 4);
    s->xc = dalloc(Nx + 4); s->yc = dalloc(Ny + 4); s->zc = dalloc(Nz + 4);
    s->kx_W = dalloc(Nx + 4); s->kx_E = dalloc(Nx + 4);
    s->ky_S = dalloc(Ny + 4); s->ky_N = dalloc(Ny + 4);
    s->kz_D = dalloc(Nz + 4); s->kz_U = dalloc(Nz + 4);
    s->flag = calloc(P, sizeof(int));
    s->cell_idx_periodic = calloc(P, sizeof(int));
    s->lvset = dalloc(P);
    double **cells[] = {
        &s->u1, &s->u1_next, &s->u1_star, &s->u1_tilde,
        &s->u2, &s->u2_next, &s->u2_star, &s->u2_tilde,
        &s->u3, &s->u3_next, &s->u3_star, &s->u3_tilde,
        &s->p, &s->p_next, &s->p_prime,
        &s->N1, &s->N1_prev, &s->N2, &s->N2_prev, &s->N3, &s->N3_prev,
        &s->p_coeffsum, &s->b, &s->x,
    };
    for (size_t i = 0; i < sizeof(cells) / sizeof(cells[0]); i++) *cells[i] = dalloc(P);
    for (int i = 0; i < 6; i++) { s->S[i] = dalloc(P); s->tau[i] = dalloc(P); }
    s->work = dalloc(8 * P);
    const size_t PX = (size_t)(Nx + 1) * (Ny + 4) * (Nz + 4);
    const size_t PY = (size_t)(Nx + 4) * (Ny + 1) * (Nz + 4);
    const size_t PZ = (size_t)(Nx + 4) * (Ny + 4) * (Nz + 1);
    s->U1 = dalloc(PX); s->U1_next = dalloc(PX); s->U1_star = dalloc(PX);
    s->U2 = dalloc(PY); s->U2_next = dalloc(PY); s->U2_star = dalloc(PY);
    s->U3 = dalloc(PZ); s->U3_next = dalloc(PZ); s->U3_star = dalloc(PZ);

    for (int i = 0; i < Nx; i++) {
        C1(s->dx, i) = xf[i + 1] - xf[i];
        C1(s->xc, i) = (xf[i + 1] + xf[i]) / 2;
    }
    for (int j = 0; j < Ny; j++) {
        C1(s->dy, j) = yf[j + 1] - yf[j];
        C1(s->yc, j) = (yf[j + 1] + yf[j]) / 2;
    }
    for (int k = 0; k < Nz; k++) {
        C1(s->dz, k) = zf[k + 1] - zf[k];
        C1(s->zc, k) = (zf[k + 1] + zf[k]) / 2;
    }
    s->xmin = xf[0]; s->xmax = xf[Nx];
    s->ymin = yf[0]; s->ymax = yf[Ny];
    s->zmin = zf[0]; s->zmax = zf[Nz];
}

void IBMSolver_set_params(IBMSolver *s, const double Re, const double dt) {
    s->Re = Re;  /* src/ibm3d_setup.c:294-297 */
    s->dt = dt;
}

/* <- src/ibm3d_setup.c:306-327 */
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

/* <- src/ibm3d_setup.c:336-356; TURBMODEL_VREMAN is an extension. */
void IBMSolver_set_turb_model(IBMSolver *s, IBMSolverTurbModelType t, ...) {
    va_list ap;
    va_start(ap, t);
    s->turb_type = t;
    switch (t) {
    case TURBMODEL_NONE: break;
    case TURBMODEL_SMAGORINSKY:
    case TURBMODEL_VREMAN: s->Cs = va_arg(ap, double); break;
    case TURBMODEL_SMAGORINSKY_VD:
        s->Cs = va_arg(ap, double);
        s->vd_aplus = va_arg(ap, double);
        s->vd_utau = va_arg(ap, double);
        s->vd_walls = va_arg(ap, int);
        break;
    default: va_end(ap); fatal("unsupported turbulence model");
    }
    va_end(ap);
}

/* <- src/ibm3d_setup.c:367-439 */
void IBMSolver_set_bc(IBMSolver *s, IBMSolverDirection dir, IBMSolverBCType type, ...) {
    IBMSolverValType vt = VAL_CONST;
    double cu1 = NAN, cu2 = NAN, cu3 = NAN, cp = NAN;
    IBMSolverBCValFunc fu1 = NULL, fu2 = NULL, fu3 = NULL, fp = NULL;
    double uc = 0;
    va_list ap;
    va_start(ap, type);
    switch (type) {
    case BC_VELOCITY_COMPONENT:
        vt = va_arg(ap, IBMSolverValType);
        if (vt == VAL_CONST) {
            cu1 = va_arg(ap, double); cu2 = va_arg(ap, double); cu3 = va_arg(ap, double);
        } else {
            fu1 = va_arg(ap, IBMSolverBCValFunc);
            fu2 = va_arg(ap, IBMSolverBCValFunc);
            fu3 = va_arg(ap, IBMSolverBCValFunc);
        }
        break;
    case BC_PRESSURE:
    case BC_VELOCITY_PERIODIC:
        vt = va_arg(ap, IBMSolverValType);
        if (vt == VAL_CONST) cp = va_arg(ap, double);
        else fp = va_arg(ap, IBMSolverBCValFunc);
        break;
    case BC_STATIONARY_WALL:
        type = BC_VELOCITY_COMPONENT;
        vt = VAL_CONST; cu1 = cu2 = cu3 = 0;
        break;
    case BC_SYMMETRIC:
        type = BC_FREE_SLIP_WALL;
        break;
    case BC_CONVECTIVE_OUTFLOW:
        uc = va_arg(ap, double);
        cp = 0;
        break;
    default:;
    }
    va_end(ap);
    for (int i = 0; i < 6; i++) {
        if (dir & (1 << i)) {
            OBC *b = &s->bc[i];
            s->conv_uc[i] = uc;
            b->type = type; b->val_type = vt;
            b->const_u1 = cu1; b->const_u2 = cu2; b->const_u3 = cu3; b->const_p = cp;
            b->func_u1 = fu1; b->func_u2 = fu2; b->func_u3 = fu3; b->func_p = fp;
        }
    }
}

void IBMSolver_set_obstacle(IBMSolver *s, Polyhedron *poly) {
    s->poly = poly;  /* src/ibm3d_setup.c:448-450 */
    s->has_sphere = 0;
}

void IBMSolver_set_obstacle_sphere(IBMSolver *s, double cx, double cy, double cz, double r) {
    s->poly = NULL;
    s->has_sphere = 1;
    s->sphere[0] = cx; s->sphere[1] = cy; s->sphere[2] = cz; s->sphere[3] = r;
}

void IBMSolver_set_precision(IBMSolver *s, int mode) {
    (void)s; (void)mode;  /* the oracle is FP64 only */
}

void IBMSolver_set_mg_params(IBMSolver *s, double gamma, int pre, int post, int smoother) {
    (void)s; (void)gamma; (void)pre; (void)post; (void)smoother;  /* the oracle has no multigrid */
}

void IBMSolver_set_linear_solver(IBMSolver *s, IBMSolverLinearSolverType t,
                                 IBMSolverPrecondType pt, const double tol) {
    s->solver_type = t;  /* src/ibm3d_setup.c:461-470 */
    s->precond_type = pt;
    s->tol = tol;
}

void IBMSolver_set_autosave(IBMSolver *s, const char *filename, int period) {
    s->autosave_filename = filename;  /* src/ibm3d_setup.c:484-491 */
    s->autosave_period = period;
}

static bool isperiodic(IBMSolverBCType t) {
    return t == BC_ALL_PERIODIC || t == BC_VELOCITY_PERIODIC;  /* :2910-2918 */
}

/* ------------------------------------------------------------------------ */
/* geometry helpers for the six outer faces                                  */
/* ------------------------------------------------------------------------ */

/* bc[] index of face (axis, high side); order N, E, S, W, D, U. */
static int face_bc(int ax, int hi) {
    static const int t[3][2] = {{3, 1}, {2, 0}, {4, 5}};
    return t[ax][hi];
}

static int axis_n(const IBMSolver *s, int ax) {
    return ax == 0 ? s->Nx : ax == 1 ? s->Ny : s->Nz;
}
static const double *axis_d(const IBMSolver *s, int ax) {
    return ax == 0 ? s->dx : ax == 1 ? s->dy : s->dz;
}
static const double *axis_c(const IBMSolver *s, int ax) {
    return ax == 0 ? s->xc : ax == 1 ? s->yc : s->zc;
}

/* Tangential loop ranges of the face passes: x faces (j, k) interior; y faces
   i over the padded extent, k interior; z faces i and j padded
   (src/ibm3d_fracstep.c:878-881 and the loops of update_outer). */
static void face_ranges(const IBMSolver *s, int ax, int *a0, int *a1, int *b0, int *b1) {
    if (ax == 0) { *a0 = 0; *a1 = s->Ny; *b0 = 0; *b1 = s->Nz; }
    else if (ax == 1) { *a0 = -2; *a1 = s->Nx + 2; *b0 = 0; *b1 = s->Nz; }
    else { *a0 = -2; *a1 = s->Nx + 2; *b0 = -2; *b1 = s->Ny + 2; }
}

/* Padded linear index of the cell at normal position n on face axis ax with
   tangential coordinates (t1, t2). */
static size_t face_cell(const IBMSolver *s, int ax, int n, int t1, int t2) {
    if (ax == 0) return CI(s, n, t1, t2);
    if (ax == 1) return CI(s, t1, n, t2);
    return CI(s, t1, t2, n);
}

static void face_coord(const IBMSolver *s, int ax, int hi, int t1, int t2,
                       double *x, double *y, double *z) {
    if (ax == 0) { *x = hi ? s->xmax : s->xmin; *y = C1(s->yc, t1); *z = C1(s->zc, t2); }
    else if (ax == 1) { *x = C1(s->xc, t1); *y = hi ? s->ymax : s->ymin; *z = C1(s->zc, t2); }
    else { *x = C1(s->xc, t1); *y = C1(s->yc, t2); *z = hi ? s->zmax : s->zmin; }
}

/* <- src/ibm3d_fracstep.c:2001-2035 */
static double bc_val_u(const IBMSolver *s, int l, int bci, double t, double x, double y, double z) {
    const OBC *b = &s->bc[bci];
    if (b->val_type == VAL_CONST)
        return l == 1 ? b->const_u1 : l == 2 ? b->const_u2 : b->const_u3;
    IBMSolverBCValFunc f = l == 1 ? b->func_u1 : l == 2 ? b->func_u2 : b->func_u3;
    return f(t, x, y, z);
}
static double bc_val_p(const IBMSolver *s, int bci, double t, double x, double y, double z) {
    const OBC *b = &s->bc[bci];
    return b->val_type == VAL_CONST ? b->const_p : b->func_p(t, x, y, z);
}
/* <- src/ibm3d_fracstep.c:1978-1999 */
static double ext_force(const IBMSolver *s, int l, double t, double x, double y, double z) {
    return s->force_type == VAL_CONST ? s->const_f[l - 1] : s->func_f[l - 1](t, x, y, z);
}

/* Positions along the normal of a face: first and second outer layers and
   first and second interior cells. */
typedef struct { int o1, o2, q0, q1; } FacePos;
static FacePos face_pos(int n, int hi) {
    FacePos f;
    if (hi) { f.o1 = n; f.o2 = n + 1; f.q0 = n - 1; f.q1 = n - 2; }
    else { f.o1 = -1; f.o2 = -2; f.q0 = 0; f.q1 = 1; }
    return f;
}

/* Distances of the linear interpolation to the face for ghost layers 1 and 2
   (src/ibm3d_fracstep.c:889-899 and mirrors): ao outer, bi inner. */
static void dirichlet_dist(const double *d, FacePos f, int layer, double *ao, double *bi) {
    if (layer == 1) { *ao = C1(d, f.o1) / 2; *bi = C1(d, f.q0) / 2; }
    else { *ao = C1(d, f.o2) / 2 + C1(d, f.o1); *bi = C1(d, f.q0) + C1(d, f.q1) / 2; }
}

/* ------------------------------------------------------------------------ */
/* Convective outflow (extension; no counterpart in the reference)           */
/* ------------------------------------------------------------------------ */

static size_t conv_extent(const IBMSolver *s, int ax) {
    int a0, a1, b0, b1;
    face_ranges(s, ax, &a0, &a1, &b0, &b1);
    return (size_t)(a1 - a0) * (b1 - b0);
}

/* face value of component l (1..3) at tangential position (t1, t2) */
static double conv_val(const IBMSolver *s, int bci, int l, int t1, int t2) {
    const int ax = bci == 1 || bci == 3 ? 0 : bci == 0 || bci == 2 ? 1 : 2;
    int a0, a1, b0, b1;
    face_ranges(s, ax, &a0, &a1, &b0, &b1);
    return s->conv_g[bci][(size_t)(l - 1) * conv_extent(s, ax) + (size_t)(t1 - a0) * (b1 - b0) + (t2 - b0)];
}

/* init: g <- u at the first interior cell (zero normal gradient). Otherwise
   one explicit upwind step of dg/dt + Uc dg/dn = 0 between the first
   interior cell centre and the face: g <- g - dt Uc (g - u_q0) / (d_q0 / 2),
   with Uc the set speed, or (Uc <= 0) the area-weighted mean outward normal
   component of g over the face's interior cells, clamped at 0. */
static void conv_outflow(IBMSolver *s, int init) {
    const double *u[4] = {NULL, s->u1, s->u2, s->u3};
    for (int bci = 0; bci < 6; bci++) {
        if (s->bc[bci].type != BC_CONVECTIVE_OUTFLOW) continue;
        const int ax = bci == 1 || bci == 3 ? 0 : bci == 0 || bci == 2 ? 1 : 2;
        const int hi = bci == 0 || bci == 1 || bci == 5;
        const int N = axis_n(s, ax);
        const FacePos f = face_pos(N, hi);
        const double bi = C1(axis_d(s, ax), f.q0) / 2;
        int a0, a1, b0, b1;
        face_ranges(s, ax, &a0, &a1, &b0, &b1);
        const size_t ext = conv_extent(s, ax);
        if (!s->conv_g[bci]) s->conv_g[bci] = calloc(3 * ext, sizeof(double));
        double *g = s->conv_g[bci];
        double uc = s->conv_uc[bci];
        if (!init && uc <= 0) {
            /* tangential widths of the face's interior cells */
            const double *w1 = ax == 0 ? s->dy : s->dx, *w2 = ax == 2 ? s->dy : s->dz;
            const int n1 = ax == 0 ? s->Ny : s->Nx, n2 = ax == 2 ? s->Ny : s->Nz;
            double num = 0, den = 0;
            for (int t1 = 0; t1 < n1; t1++)
                for (int t2 = 0; t2 < n2; t2++) {
                    const double A = C1(w1, t1) * C1(w2, t2);
                    num += (hi ? 1 : -1) * g[(size_t)ax * ext + (size_t)(t1 - a0) * (b1 - b0) + (t2 - b0)] * A;
                    den += A;
                }
            uc = fmax(num / den, 0.0);
        }
        for (int l = 0; l < 3; l++)
            for (int t1 = a0; t1 < a1; t1++)
                for (int t2 = b0; t2 < b1; t2++) {
                    double *gv = &g[(size_t)l * ext + (size_t)(t1 - a0) * (b1 - b0) + (t2 - b0)];
                    const double uq = u[l + 1][face_cell(s, ax, f.q0, t1, t2)];
                    *gv = init ? uq : *gv - s->dt * uc * (*gv - uq) / bi;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* assemble  <- src/ibm3d_setup.c:499-639                                   */
/* ------------------------------------------------------------------------ */

static void calc_lvset_flag(IBMSolver *s);
static void calc_flag(IBMSolver *s);
static void build_matrices(IBMSolver *s);

void IBMSolver_assemble(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    if (isperiodic(s->bc[0].type) != isperiodic(s->bc[2].type) ||
        isperiodic(s->bc[1].type) != isperiodic(s->bc[3].type) ||
        isperiodic(s->bc[4].type) != isperiodic(s->bc[5].type)) {
        printf("Inconsistent periodic boundary condition\n");  /* :523-533 */
        fflush(stdout);
        abort();
    }
    /* solver / preconditioner type checks, src/ibm3d_setup.c:1435-1536 (the
       oracle itself always solves with BiCGSTAB) */
    if (s->solver_type != SOLVER_AMG && s->solver_type != SOLVER_PCG && s->solver_type != SOLVER_BiCGSTAB
        && s->solver_type != SOLVER_GMRES) {
        printf("\nUnknown linear solver type\n");
        fflush(stdout);
        abort();
    }
    if (s->precond_type != PRECOND_NONE && s->precond_type != PRECOND_AMG) {
        printf("\nUnknown preconditioner type\n");
        fflush(stdout);
        abort();
    }
    if (s->solver_type == SOLVER_AMG && s->precond_type != PRECOND_NONE) {
        printf("\nCannot use preconditioner with BoomerAMG solver\n");
        fflush(stdout);
        abort();
    }
    /* ghost widths and centres, :536-588 */
    double *d[3] = {s->dx, s->dy, s->dz};
    double *c[3] = {s->xc, s->yc, s->zc};
    const int n[3] = {Nx, Ny, Nz};
    const int lo_bc[3] = {3, 2, 4}, hi_bc[3] = {1, 0, 5};
    for (int a = 0; a < 3; a++) {
        const int N = n[a];
        const bool plo = isperiodic(s->bc[lo_bc[a]].type), phi = isperiodic(s->bc[hi_bc[a]].type);
        C1(d[a], -1) = plo ? C1(d[a], N - 1) : C1(d[a], 0);
        C1(d[a], -2) = plo ? C1(d[a], N - 2) : C1(d[a], 1);
        C1(d[a], N) = phi ? C1(d[a], 0) : C1(d[a], N - 1);
        C1(d[a], N + 1) = phi ? C1(d[a], 1) : C1(d[a], N - 2);
        C1(c[a], -1) = C1(c[a], 0) - (C1(d[a], -1) + C1(d[a], 0)) / 2;
        C1(c[a], -2) = C1(c[a], -1) - (C1(d[a], -2) + C1(d[a], -1)) / 2;
        C1(c[a], N) = C1(c[a], N - 1) + (C1(d[a], N - 1) + C1(d[a], N)) / 2;
        C1(c[a], N + 1) = C1(c[a], N) + (C1(d[a], N) + C1(d[a], N + 1)) / 2;
    }
    /* second-derivative coefficients, :620-632 */
    for (int i = 0; i < Nx; i++) {
        C1(s->kx_W, i) = s->dt / (2 * s->Re * (C1(s->xc, i) - C1(s->xc, i - 1)) * C1(s->dx, i));
        C1(s->kx_E, i) = s->dt / (2 * s->Re * (C1(s->xc, i + 1) - C1(s->xc, i)) * C1(s->dx, i));
    }
    for (int j = 0; j < Ny; j++) {
        C1(s->ky_S, j) = s->dt / (2 * s->Re * (C1(s->yc, j) - C1(s->yc, j - 1)) * C1(s->dy, j));
        C1(s->ky_N, j) = s->dt / (2 * s->Re * (C1(s->yc, j + 1) - C1(s->yc, j)) * C1(s->dy, j));
    }
    for (int k = 0; k < Nz; k++) {
        C1(s->kz_D, k) = s->dt / (2 * s->Re * (C1(s->zc, k) - C1(s->zc, k - 1)) * C1(s->dz, k));
        C1(s->kz_U, k) = s->dt / (2 * s->Re * (C1(s->zc, k + 1) - C1(s->zc, k)) * C1(s->dz, k));
    }

    /* periodic image of every padded cell, src/ibm3d_setup.c:893-1074
       (applied on all three axes in turn, x then y then z) */
    const size_t P = ncell(s);
    for (size_t q = 0; q < P; q++) s->cell_idx_periodic[q] = (int)q;
    for (int i = -2; i < Nx + 2; i++)
        for (int j = -2; j < Ny + 2; j++)
            for (int k = -2; k < Nz + 2; k++) {
                if (i < 0) C3(s, s->cell_idx_periodic, i, j, k) = C3(s, s->cell_idx_periodic, i + Nx, j, k);
            }
    for (int i = Nx; i <= Nx + 1; i++)
        for (int j = -2; j < Ny + 2; j++)
            for (int k = -2; k < Nz + 2; k++)
                C3(s, s->cell_idx_periodic, i, j, k) = C3(s, s->cell_idx_periodic, i - Nx, j, k);
    for (int i = -2; i < Nx + 2; i++)
        for (int k = -2; k < Nz + 2; k++) {
            for (int j = -2; j <= -1; j++)
                C3(s, s->cell_idx_periodic, i, j, k) = C3(s, s->cell_idx_periodic, i, j + Ny, k);
            for (int j = Ny; j <= Ny + 1; j++)
                C3(s, s->cell_idx_periodic, i, j, k) = C3(s, s->cell_idx_periodic, i, j - Ny, k);
        }
    for (int i = -2; i < Nx + 2; i++)
        for (int j = -2; j < Ny + 2; j++) {
            for (int k = -2; k <= -1; k++)
                C3(s, s->cell_idx_periodic, i, j, k) = C3(s, s->cell_idx_periodic, i, j, k + Nz);
            for (int k = Nz; k <= Nz + 1; k++)
                C3(s, s->cell_idx_periodic, i, j, k) = C3(s, s->cell_idx_periodic, i, j, k - Nz);
        }

    calc_lvset_flag(s);
    build_matrices(s);
}

/* <- src/ibm3d_setup.c:1081-1254 (single process: no halo exchanges, so the
   post-CPT sweeps of :1125-1227 never run) */
static void calc_lvset_flag(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    if (!s->poly && !s->has_sphere) {
        for (size_t q = 0; q < ncell(s); q++) { s->lvset[q] = .5; s->flag[q] = FLAG_FLUID; }
        return;
    }
    if (s->poly) {
        Polyhedron_cpt(s->poly, Nx + 4, Ny + 4, Nz + 4, s->xc, s->yc, s->zc, s->lvset, 15);
    } else {
        for (int i = -2; i < Nx + 2; i++)
            for (int j = -2; j < Ny + 2; j++)
                for (int k = -2; k < Nz + 2; k++)
                    C3(s, s->lvset, i, j, k) = oracle_lvset_sphere(
                        s->sphere, C1(s->xc, i), C1(s->yc, j), C1(s->zc, k));
    }
    calc_flag(s);
}

/* flags from the level set  <- src/ibm3d_setup.c:1235-1254 */
static void calc_flag(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    for (int i = -2; i < Nx + 2; i++)
        for (int j = -2; j < Ny + 2; j++)
            for (int k = -2; k < Nz + 2; k++) {
                if (C3(s, s->lvset, i, j, k) > 0) {
                    C3(s, s->flag, i, j, k) = FLAG_FLUID;
                    continue;
                }
                bool ghost = false;
                for (int l = 0; l < 6; l++) {
                    const int ni = i + ADJ[l][0], nj = j + ADJ[l][1], nk = k + ADJ[l][2];
                    if (ni < -2 || ni >= Nx + 2 || nj < -2 || nj >= Ny + 2 || nk < -2 || nk >= Nz + 2)
                        continue;
                    ghost = ghost || C3(s, s->lvset, ni, nj, nk) > 0;
                }
                C3(s, s->flag, i, j, k) = ghost ? FLAG_GHOST : FLAG_SOLID;
            }
}

/* upper_bound of src/utils.c:66-81 */
static int upper_bound(int len, const double *arr, double val) {
    int l = 0, h = len;
    while (l < h) {
        int mid = l + (h - l) / 2;
        if (val >= arr[mid]) l = mid + 1;
        else h = mid;
    }
    return l;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

/* Mirror point and trilinear weights  <- src/utils.c:111-171 */
void IBMSolver_ghost_interp(IBMSolver *s, const int i, const int j, const int k,
                            int idx[][3], double w[]) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const double nx = (C3(s, s->lvset, i + 1, j, k) - C3(s, s->lvset, i - 1, j, k))
        / (C1(s->xc, i + 1) - C1(s->xc, i - 1));
    const double ny = (C3(s, s->lvset, i, j + 1, k) - C3(s, s->lvset, i, j - 1, k))
        / (C1(s->yc, j + 1) - C1(s->yc, j - 1));
    const double nz = (C3(s, s->lvset, i, j, k + 1) - C3(s, s->lvset, i, j, k - 1))
        / (C1(s->zc, k + 1) - C1(s->zc, k - 1));
    const double phi = C3(s, s->lvset, i, j, k);
    /* Vector_lincom(1, xc, -2 phi, n) */
    const double mx = 1 * C1(s->xc, i) + -2 * phi * nx;
    const double my = 1 * C1(s->yc, j) + -2 * phi * ny;
    const double mz = 1 * C1(s->zc, k) + -2 * phi * nz;
    const int im = clampi(upper_bound(Nx + 4, s->xc, mx) - 3, -2, Nx);
    const int jm = clampi(upper_bound(Ny + 4, s->yc, my) - 3, -2, Ny);
    const int km = clampi(upper_bound(Nz + 4, s->zc, mz) - 3, -2, Nz);
    for (int l = 0; l < 8; l++) {
        idx[l][0] = im + !!(l & 4);
        idx[l][1] = jm + !!(l & 2);
        idx[l][2] = km + !!(l & 1);
    }
    const double xl = C1(s->xc, im), xu = C1(s->xc, im + 1);
    const double yl = C1(s->yc, jm), yu = C1(s->yc, jm + 1);
    const double zl = C1(s->zc, km), zu = C1(s->zc, km + 1);
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

/* ------------------------------------------------------------------------ */
/* operator assembly  <- src/ibm3d_setup.c:1566-2908                        */
/* ------------------------------------------------------------------------ */

static void mat_alloc(EllMat *A, int nrows) {
    free_mat(A);
    A->nrows = nrows;
    A->nnz = calloc(nrows, sizeof(int));
    A->col = calloc((size_t)9 * nrows, sizeof(int));
    A->val = calloc((size_t)9 * nrows, sizeof(double));
}

static void mat_set_row(EllMat *A, int row, int n, const int *cols, const double *vals) {
    A->nnz[row] = n;
    for (int l = 0; l < n; l++) {
        A->col[(size_t)9 * row + l] = cols[l];
        A->val[(size_t)9 * row + l] = vals[l];
    }
}

static bool p_singular(const IBMSolver *s) {
    for (int i = 0; i < 6; i++)
        if (s->bc[i].type == BC_PRESSURE || s->bc[i].type == BC_VELOCITY_PERIODIC
            || s->bc[i].type == BC_CONVECTIVE_OUTFLOW) return false;
    return true;  /* src/ibm3d_setup.c:1598-1603 */
}

/* Two-entry Dirichlet interpolation row of ghost layer `layer` (velocity
   rows of :1726-1762 and pressure rows of :1795-1811). */
static void row_dirichlet(EllMat *A, const IBMSolver *s, int ax, int hi, int t1, int t2, int layer) {
    const FacePos f = face_pos(axis_n(s, ax), hi);
    double ao, bi;
    dirichlet_dist(axis_d(s, ax), f, layer, &ao, &bi);
    int cols[2] = {(int)face_cell(s, ax, layer == 1 ? f.o1 : f.o2, t1, t2),
                   (int)face_cell(s, ax, layer == 1 ? f.q0 : f.q1, t1, t2)};
    double vals[2] = {bi / (ao + bi), ao / (ao + bi)};
    mat_set_row(A, cols[0], 2, cols, vals);
}

/* (1, -1) Neumann row pairing the ghost with its mirror interior cell. */
static void row_neumann(EllMat *A, const IBMSolver *s, int ax, int hi, int t1, int t2, int layer) {
    const FacePos f = face_pos(axis_n(s, ax), hi);
    int cols[2] = {(int)face_cell(s, ax, layer == 1 ? f.o1 : f.o2, t1, t2),
                   (int)face_cell(s, ax, layer == 1 ? f.q0 : f.q1, t1, t2)};
    double vals[2] = {1, -1};
    mat_set_row(A, cols[0], 2, cols, vals);
}

/* Three-point linear extrapolation row for velocity at a pressure boundary
   (:1764-1791 and mirrors). */
static void row_extrap(EllMat *A, const IBMSolver *s, int ax, int hi, int t1, int t2, int layer) {
    const FacePos f = face_pos(axis_n(s, ax), hi);
    const double *c = axis_c(s, ax);
    int o, n1, n2;
    double so, si;
    if (layer == 2) {
        o = f.o2; n1 = f.o1; n2 = f.q0;
        so = hi ? C1(c, f.o2) - C1(c, f.o1) : C1(c, f.o1) - C1(c, f.o2);
        si = hi ? C1(c, f.o1) - C1(c, f.q0) : C1(c, f.q0) - C1(c, f.o1);
    } else {
        o = f.o1; n1 = f.q0; n2 = f.q1;
        so = hi ? C1(c, f.o1) - C1(c, f.q0) : C1(c, f.q0) - C1(c, f.o1);
        si = hi ? C1(c, f.q0) - C1(c, f.q1) : C1(c, f.q1) - C1(c, f.q0);
    }
    int cols[3] = {(int)face_cell(s, ax, o, t1, t2), (int)face_cell(s, ax, n1, t1, t2),
                   (int)face_cell(s, ax, n2, t1, t2)};
    double vals[3] = {si / (so + si), -1, so / (so + si)};
    mat_set_row(A, cols[0], 3, cols, vals);
}

static void row_periodic(EllMat *A, const IBMSolver *s, int ax, int hi, int t1, int t2, int layer) {
    const FacePos f = face_pos(axis_n(s, ax), hi);
    const int c0 = (int)face_cell(s, ax, layer == 1 ? f.o1 : f.o2, t1, t2);
    int cols[2] = {c0, s->cell_idx_periodic[c0]};
    double vals[2] = {1, -1};
    mat_set_row(A, c0, 2, cols, vals);
}

/* type 1..3 = momentum component, 4 = pressure */
static void build_matrix(IBMSolver *s, int type) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    EllMat *A = &s->A[type - 1];
    mat_alloc(A, (int)ncell(s));
    const bool sing = p_singular(s);
    int iidx[8][3];
    double w[8];

    /* interior rows, :1615-1721 */
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++) {
                int cols[9] = {0};
                double vals[9] = {0};
                cols[0] = (int)CI(s, i, j, k);
                const int fl = C3(s, s->flag, i, j, k);
                if (type == 4 && sing && i == 0 && j == 0 && k == 0) {
                    vals[0] = 1;
                } else if (fl == FLAG_FLUID) {
                    for (int l = 0; l < 6; l++)
                        cols[l + 1] = (int)CI(s, i + ADJ[l][0], j + ADJ[l][1], k + ADJ[l][2]);
                    vals[1] = -C1(s->ky_N, j);
                    vals[2] = -C1(s->kx_E, i);
                    vals[3] = -C1(s->ky_S, j);
                    vals[4] = -C1(s->kx_W, i);
                    vals[5] = -C1(s->kz_D, k);
                    vals[6] = -C1(s->kz_U, k);
                    vals[0] = type != 4
                        ? 1 - vals[1] - vals[2] - vals[3] - vals[4] - vals[5] - vals[6]
                        : -vals[1] - vals[2] - vals[3] - vals[4] - vals[5] - vals[6];
                    if (type == 4) {
                        C3(s, s->p_coeffsum, i, j, k) = vals[0];
                        for (int l = 1; l < 7; l++) vals[l] /= vals[0];
                        vals[0] = 1;
                    }
                } else if (fl == FLAG_GHOST) {
                    vals[0] = 1;
                    IBMSolver_ghost_interp(s, i, j, k, iidx, w);
                    for (int l = 0; l < 8; l++)
                        if (C3(s, s->flag, iidx[l][0], iidx[l][1], iidx[l][2]) == FLAG_SOLID) w[l] = 0;
                    double sum = 0;
                    for (int l = 0; l < 8; l++) sum += w[l];
                    for (int l = 0; l < 8; l++) w[l] /= sum;
                    for (int l = 0; l < 8; l++) {
                        if ((int)CI(s, iidx[l][0], iidx[l][1], iidx[l][2]) == cols[0]) {
                            if (type != 4) vals[0] += w[l];
                            else vals[0] -= w[l];
                            w[l] = 0;
                            break;
                        }
                    }
                    for (int l = 0; l < 8; l++) {
                        cols[l + 1] = (int)CI(s, iidx[l][0], iidx[l][1], iidx[l][2]);
                        vals[l + 1] = type != 4 ? w[l] : -w[l];
                    }
                } else if (fl == FLAG_SOLID) {
                    vals[0] = 1;
                } else {
                    fatal("wrong flag");
                }
                int nc = 1;
                for (int l = 1; l < 9; l++) {
                    if (vals[l] != 0) { cols[nc] = cols[l]; vals[nc] = vals[l]; nc++; }
                }
                mat_set_row(A, cols[0], nc, cols, vals);
            }

    /* outer rows, :1723-2903 */
    for (int ax = 0; ax < 3; ax++)
        for (int hi = 0; hi < 2; hi++) {
            const IBMSolverBCType bt = s->bc[face_bc(ax, hi)].type;
            int a0, a1, b0, b1;
            face_ranges(s, ax, &a0, &a1, &b0, &b1);
            for (int t1 = a0; t1 < a1; t1++)
                for (int t2 = b0; t2 < b1; t2++)
                    for (int layer = 2; layer >= 1; layer--) {
                        switch (bt) {
                        case BC_VELOCITY_COMPONENT:
                        case BC_STATIONARY_WALL:
                            if (type != 4) row_dirichlet(A, s, ax, hi, t1, t2, layer);
                            else row_neumann(A, s, ax, hi, t1, t2, layer);
                            break;
                        case BC_PRESSURE:
                            if (type != 4) row_extrap(A, s, ax, hi, t1, t2, layer);
                            else row_dirichlet(A, s, ax, hi, t1, t2, layer);
                            break;
                        case BC_FREE_SLIP_WALL:
                            if (type == ax + 1) row_dirichlet(A, s, ax, hi, t1, t2, layer);
                            else row_neumann(A, s, ax, hi, t1, t2, layer);
                            break;
                        case BC_ALL_PERIODIC:
                            row_periodic(A, s, ax, hi, t1, t2, layer);
                            break;
                        case BC_VELOCITY_PERIODIC:
                            if (type != 4) row_periodic(A, s, ax, hi, t1, t2, layer);
                            else row_dirichlet(A, s, ax, hi, t1, t2, layer);
                            break;
                        case BC_CONVECTIVE_OUTFLOW:
                            row_dirichlet(A, s, ax, hi, t1, t2, layer);
                            break;
                        default:;
                        }
                    }
        }
}

static void build_matrices(IBMSolver *s) {
    for (int t = 1; t <= 4; t++) build_matrix(s, t);
    if (s->precond_type == PRECOND_AMG) oracle_ilu0_build(&s->A[3]);
}

/* ------------------------------------------------------------------------ */
/* Krylov solver (stands in for Hypre BiCGSTAB)                              */
/* ------------------------------------------------------------------------ */

#define NCHUNK 256
/* Deterministic dot product: fixed chunking independent of thread count. */
static double ddot(const double *a, const double *b, int n) {
    double part[NCHUNK];
    #pragma omp parallel for schedule(static)
    for (int c = 0; c < NCHUNK; c++) {
        const long lo = (long)n * c / NCHUNK, hi = (long)n * (c + 1) / NCHUNK;
        double acc = 0;
        for (long q = lo; q < hi; q++) acc += a[q] * b[q];
        part[c] = acc;
    }
    double acc = 0;
    for (int c = 0; c < NCHUNK; c++) acc += part[c];
    return acc;
}

static void spmv(const EllMat *A, const double *x, double *y) {
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < A->nrows; r++) {
        double acc = 0;
        const size_t o = (size_t)9 * r;
        for (int l = 0; l < A->nnz[r]; l++) acc += A->val[o + l] * x[A->col[o + l]];
        y[r] = acc;
    }
}

/* Block-Jacobi ILU(0) for large systems: the rows split into contiguous
   blocks and each block's ILU(0) keeps only the couplings inside the block,
   so the blocks factor and solve independently (in parallel). Systems below
   ILU_BLOCK_ROWS rows keep one block (plain ILU(0)). The block count depends
   only on the row count, so the preconditioner, and with it every iterate,
   does not depend on the number of threads. Only iteration counts depend on
   the preconditioner (it stands in for the reference's BoomerAMG). */
#define ILU_MAX_BLOCKS 16
#define ILU_BLOCK_ROWS 262144

static int ilu_blocks(int n) {
    const int b = n / ILU_BLOCK_ROWS;
    return b < 1 ? 1 : b > ILU_MAX_BLOCKS ? ILU_MAX_BLOCKS : b;
}

static int ilu_block_len(int n) { return (n + ilu_blocks(n) - 1) / ilu_blocks(n); }

void oracle_ilu0_build(EllMat *A) {
    const int n = A->nrows;
    const int bl = ilu_block_len(n);
    free(A->scol); free(A->lu); free(A->diagpos);
    A->scol = malloc((size_t)9 * n * sizeof(int));
    A->lu = malloc((size_t)9 * n * sizeof(double));
    A->diagpos = malloc((size_t)n * sizeof(int));
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < n; r++) {
        const size_t o = (size_t)9 * r;
        const int m = A->nnz[r];
        for (int l = 0; l < m; l++) { A->scol[o + l] = A->col[o + l]; A->lu[o + l] = A->val[o + l]; }
        for (int a = 1; a < m; a++)  /* insertion sort by column */
            for (int b = a; b > 0 && A->scol[o + b - 1] > A->scol[o + b]; b--) {
                int tc = A->scol[o + b]; A->scol[o + b] = A->scol[o + b - 1]; A->scol[o + b - 1] = tc;
                double tv = A->lu[o + b]; A->lu[o + b] = A->lu[o + b - 1]; A->lu[o + b - 1] = tv;
            }
        A->diagpos[r] = -1;
        for (int l = 0; l < m; l++) if (A->scol[o + l] == r) A->diagpos[r] = l;
    }
    #pragma omp parallel for schedule(dynamic, 1)
    for (int blk = 0; blk < ilu_blocks(n); blk++) {
        const int lo = blk * bl, hi = lo + bl < n ? lo + bl : n;
        for (int r = lo; r < hi; r++) {
            const size_t o = (size_t)9 * r;
            const int m = A->nnz[r];
            for (int a = 0; a < m && A->scol[o + a] < r; a++) {
                const int kk = A->scol[o + a];
                if (kk < lo) continue;  /* coupling to another block: dropped */
                const size_t ok = (size_t)9 * kk;
                double piv = A->diagpos[kk] >= 0 ? A->lu[ok + A->diagpos[kk]] : 0;
                if (fabs(piv) < 1e-300) piv = 1;
                A->lu[o + a] /= piv;
                for (int b = a + 1; b < m; b++) {
                    const int jc = A->scol[o + b];
                    for (int c = 0; c < A->nnz[kk]; c++)
                        if (A->scol[ok + c] == jc) { A->lu[o + b] -= A->lu[o + a] * A->lu[ok + c]; break; }
                }
            }
        }
    }
    A->has_ilu = true;
}

static void ilu_apply(const EllMat *A, const double *rhs, double *z) {
    const int n = A->nrows;
    const int bl = ilu_block_len(n);
    #pragma omp parallel for schedule(dynamic, 1)
    for (int blk = 0; blk < ilu_blocks(n); blk++) {
        const int lo = blk * bl, hi = lo + bl < n ? lo + bl : n;
        for (int r = lo; r < hi; r++) {  /* L z = rhs (unit lower) */
            const size_t o = (size_t)9 * r;
            double acc = rhs[r];
            for (int a = 0; a < A->nnz[r] && A->scol[o + a] < r; a++)
                if (A->scol[o + a] >= lo) acc -= A->lu[o + a] * z[A->scol[o + a]];
            z[r] = acc;
        }
        for (int r = hi - 1; r >= lo; r--) {  /* U z = z */
            const size_t o = (size_t)9 * r;
            double acc = z[r];
            for (int a = 0; a < A->nnz[r]; a++)
                if (A->scol[o + a] > r && A->scol[o + a] < hi) acc -= A->lu[o + a] * z[A->scol[o + a]];
            double piv = A->diagpos[r] >= 0 ? A->lu[o + A->diagpos[r]] : 0;
            if (fabs(piv) < 1e-300) piv = 1;
            z[r] = acc / piv;
        }
    }
}

/* BiCGSTAB (van der Vorst 1992) with optional right ILU(0) preconditioning;
   x0 = 0, stop when ||r||_2 <= tol ||b||_2. Returns iterations; sets the final
   true relative residual. */
int oracle_bicgstab(const EllMat *A, const double *b, double *x, double tol, int maxit,
                    int use_ilu, double *work, double *final_norm) {
    const int n = A->nrows;
    double *r = work, *rh = work + n, *p = work + 2 * (size_t)n, *v = work + 3 * (size_t)n;
    double *ph = work + 4 * (size_t)n, *sv = work + 5 * (size_t)n, *sh = work + 6 * (size_t)n;
    double *t = work + 7 * (size_t)n;
    memset(x, 0, sizeof(double) * n);
    const double bnorm = sqrt(ddot(b, b, n));
    if (bnorm == 0) { *final_norm = 0; return 0; }
    #pragma omp parallel for
    for (int q = 0; q < n; q++) { r[q] = b[q]; rh[q] = b[q]; p[q] = 0; v[q] = 0; }
    double rho = 1, alpha = 1, omega = 1, rnorm = bnorm;
    int it = 0;
    while (it < maxit && rnorm > tol * bnorm) {
        it++;
        const double rho1 = ddot(rh, r, n);
        if (rho1 == 0) break;
        const double beta = (rho1 / rho) * (alpha / omega);
        #pragma omp parallel for
        for (int q = 0; q < n; q++) p[q] = r[q] + beta * (p[q] - omega * v[q]);
        if (use_ilu) ilu_apply(A, p, ph); else memcpy(ph, p, sizeof(double) * n);
        spmv(A, ph, v);
        alpha = rho1 / ddot(rh, v, n);
        #pragma omp parallel for
        for (int q = 0; q < n; q++) sv[q] = r[q] - alpha * v[q];
        const double snorm = sqrt(ddot(sv, sv, n));
        if (snorm <= tol * bnorm) {
            #pragma omp parallel for
            for (int q = 0; q < n; q++) { x[q] += alpha * ph[q]; r[q] = sv[q]; }
            rnorm = snorm;
            break;
        }
        if (use_ilu) ilu_apply(A, sv, sh); else memcpy(sh, sv, sizeof(double) * n);
        spmv(A, sh, t);
        const double tt = ddot(t, t, n);
        omega = tt > 0 ? ddot(t, sv, n) / tt : 0;
        #pragma omp parallel for
        for (int q = 0; q < n; q++) {
            x[q] += alpha * ph[q] + omega * sh[q];
            r[q] = sv[q] - omega * t[q];
        }
        rnorm = sqrt(ddot(r, r, n));
        rho = rho1;
        if (omega == 0) break;
    }
    spmv(A, x, t);  /* true residual */
    #pragma omp parallel for
    for (int q = 0; q < n; q++) t[q] = b[q] - t[q];
    *final_norm = sqrt(ddot(t, t, n)) / bnorm;
    return it;
}

/* ------------------------------------------------------------------------ */
/* time step  <- src/ibm3d_fracstep.c                                        */
/* ------------------------------------------------------------------------ */

/* <- src/ibm3d_fracstep.c:117-170 */
static void calc_N(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const double *dx = s->dx, *dy = s->dy, *dz = s->dz;
    double *u[4] = {NULL, s->u1, s->u2, s->u3};
    double *N[4] = {NULL, s->N1, s->N2, s->N3};
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++) {
                if (C3(s, s->flag, i, j, k) != FLAG_FLUID) {
                    for (int l = 1; l <= 3; l++) C3(s, N[l], i, j, k) = NAN;
                    continue;
                }
                for (int l = 1; l <= 3; l++) {
                    const double *a = u[l];
                    const double uw = (C3(s, a, i - 1, j, k) * C1(dx, i) + C3(s, a, i, j, k) * C1(dx, i - 1))
                        / (C1(dx, i - 1) + C1(dx, i));
                    const double ue = (C3(s, a, i, j, k) * C1(dx, i + 1) + C3(s, a, i + 1, j, k) * C1(dx, i))
                        / (C1(dx, i) + C1(dx, i + 1));
                    const double us = (C3(s, a, i, j - 1, k) * C1(dy, j) + C3(s, a, i, j, k) * C1(dy, j - 1))
                        / (C1(dy, j - 1) + C1(dy, j));
                    const double un = (C3(s, a, i, j, k) * C1(dy, j + 1) + C3(s, a, i, j + 1, k) * C1(dy, j))
                        / (C1(dy, j) + C1(dy, j + 1));
                    const double ud = (C3(s, a, i, j, k - 1) * C1(dz, k) + C3(s, a, i, j, k) * C1(dz, k - 1))
                        / (C1(dz, k - 1) + C1(dz, k));
                    const double uu = (C3(s, a, i, j, k) * C1(dz, k + 1) + C3(s, a, i, j, k + 1) * C1(dz, k))
                        / (C1(dz, k) + C1(dz, k + 1));
                    C3(s, N[l], i, j, k)
                        = (XS(s, s->U1, i + 1, j, k) * ue - XS(s, s->U1, i, j, k) * uw) / C1(dx, i)
                        + (YS(s, s->U2, i, j + 1, k) * un - YS(s, s->U2, i, j, k) * us) / C1(dy, j)
                        + (ZS(s, s->U3, i, j, k + 1) * uu - ZS(s, s->U3, i, j, k) * ud) / C1(dz, k);
                }
            }
}

/* <- src/ibm3d_turb.c:47-82 */
static void calc_S(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const double *xc = s->xc, *yc = s->yc, *zc = s->zc;
    const double *u1 = s->u1, *u2 = s->u2, *u3 = s->u3;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++) {
                if (C3(s, s->flag, i, j, k) != FLAG_FLUID) continue;
                const double ddx = C1(xc, i + 1) - C1(xc, i - 1);
                const double ddy = C1(yc, j + 1) - C1(yc, j - 1);
                const double ddz = C1(zc, k + 1) - C1(zc, k - 1);
                C3(s, s->S[0], i, j, k) = (C3(s, u1, i + 1, j, k) - C3(s, u1, i - 1, j, k)) / ddx;
                C3(s, s->S[1], i, j, k) = 0.5 * ((C3(s, u1, i, j + 1, k) - C3(s, u1, i, j - 1, k)) / ddy
                                               + (C3(s, u2, i + 1, j, k) - C3(s, u2, i - 1, j, k)) / ddx);
                C3(s, s->S[2], i, j, k) = 0.5 * ((C3(s, u1, i, j, k + 1) - C3(s, u1, i, j, k - 1)) / ddz
                                               + (C3(s, u3, i + 1, j, k) - C3(s, u3, i - 1, j, k)) / ddx);
                C3(s, s->S[3], i, j, k) = (C3(s, u2, i, j + 1, k) - C3(s, u2, i, j - 1, k)) / ddy;
                C3(s, s->S[4], i, j, k) = 0.5 * ((C3(s, u2, i, j, k + 1) - C3(s, u2, i, j, k - 1)) / ddz
                                               + (C3(s, u3, i, j + 1, k) - C3(s, u3, i, j - 1, k)) / ddy);
                C3(s, s->S[5], i, j, k) = (C3(s, u3, i, j, k + 1) - C3(s, u3, i, j, k - 1)) / ddz;
            }
}

static inline double sqd(double x) { return x * x; }

/* tensor slot of symmetric pair (a, b), a, b in 1..3 */
static int tslot(int a, int b) {
    static const int t[4][4] = {{0}, {0, 0, 1, 2}, {0, 1, 3, 4}, {0, 2, 4, 5}};
    return t[a][b];
}

/* van Driest damping factor 1 - exp(-y+/A+), y+ = y_w u_tau Re, y_w the
   distance of the cell centre to the nearest outer face in vd_walls (an
   extension with no reference counterpart; van Driest, J. Aero. Sci. 23,
   1956). */
static double van_driest(const IBMSolver *s, int i, int j, int k) {
    const double c[3] = {C1(s->xc, i), C1(s->yc, j), C1(s->zc, k)};
    const double lo[3] = {s->xmin, s->ymin, s->zmin}, hi[3] = {s->xmax, s->ymax, s->zmax};
    const int bit_lo[3] = {DIR_WEST, DIR_SOUTH, DIR_DOWN}, bit_hi[3] = {DIR_EAST, DIR_NORTH, DIR_UP};
    double yw = 1e300;
    for (int a = 0; a < 3; a++) {
        if (s->vd_walls & bit_lo[a]) yw = fmin(yw, c[a] - lo[a]);
        if (s->vd_walls & bit_hi[a]) yw = fmin(yw, hi[a] - c[a]);
    }
    return 1 - exp(-(yw * (s->vd_utau * s->Re)) / s->vd_aplus);
}

/* Smagorinsky  <- src/ibm3d_turb.c:8-45. Vreman (extension): nu_t =
   c sqrt(B_beta / (alpha_ij alpha_ij)) with alpha_ij = du_j/dx_i from the same
   central differences and Delta_i the cell widths (Vreman, Phys. Fluids 16,
   3670, 2004); tau_ij = -2 nu_t S_ij. */
void IBMSolver_calc_tau_r(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    if (s->turb_type == TURBMODEL_NONE) return;
    calc_S(s);
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++) {
                if (C3(s, s->flag, i, j, k) != FLAG_FLUID) continue;
                double nut;
                if (s->turb_type == TURBMODEL_SMAGORINSKY || s->turb_type == TURBMODEL_SMAGORINSKY_VD) {
                    const double delta = 2 * cbrt(C1(s->dx, i) * C1(s->dy, j) * C1(s->dz, k));
                    double smag = 0;
                    for (int a = 1; a <= 3; a++)
                        for (int b = 1; b <= 3; b++) smag += 2 * sqd(C3(s, s->S[tslot(a, b)], i, j, k));
                    smag = sqrt(smag);
                    double l = s->Cs * delta;
                    if (s->turb_type == TURBMODEL_SMAGORINSKY_VD && s->vd_walls) l *= van_driest(s, i, j, k);
                    const double lw = 0.41 * C3(s, s->lvset, i, j, k);
                    if (lw < l) l = lw;
                    for (int a = 1; a <= 3; a++)
                        for (int b = a; b <= 3; b++)
                            C3(s, s->tau[tslot(a, b)], i, j, k) = -2 * sqd(l) * smag * C3(s, s->S[tslot(a, b)], i, j, k);
                    continue;
                }
                /* Vreman */
                const double *uu[3] = {s->u1, s->u2, s->u3};
                const double dd[3] = {C1(s->xc, i + 1) - C1(s->xc, i - 1), C1(s->yc, j + 1) - C1(s->yc, j - 1),
                                      C1(s->zc, k + 1) - C1(s->zc, k - 1)};
                const double del[3] = {C1(s->dx, i), C1(s->dy, j), C1(s->dz, k)};
                double al[3][3];  /* al[m][n] = d u_n / d x_m */
                for (int n = 0; n < 3; n++) {
                    al[0][n] = (C3(s, uu[n], i + 1, j, k) - C3(s, uu[n], i - 1, j, k)) / dd[0];
                    al[1][n] = (C3(s, uu[n], i, j + 1, k) - C3(s, uu[n], i, j - 1, k)) / dd[1];
                    al[2][n] = (C3(s, uu[n], i, j, k + 1) - C3(s, uu[n], i, j, k - 1)) / dd[2];
                }
                double aa = 0, bt[3][3];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) {
                        aa += al[a][b] * al[a][b];
                        double acc = 0;
                        for (int m = 0; m < 3; m++) acc += sqd(del[m]) * al[m][a] * al[m][b];
                        bt[a][b] = acc;
                    }
                const double Bb = bt[0][0] * bt[1][1] - bt[0][1] * bt[0][1] + bt[0][0] * bt[2][2]
                    - bt[0][2] * bt[0][2] + bt[1][1] * bt[2][2] - bt[1][2] * bt[1][2];
                nut = (aa > 1e-300 && Bb > 0) ? s->Cs * sqrt(Bb / aa) : 0;
                for (int a = 1; a <= 3; a++)
                    for (int b = a; b <= 3; b++)
                        C3(s, s->tau[tslot(a, b)], i, j, k) = -2 * nut * C3(s, s->S[tslot(a, b)], i, j, k);
            }
}

/* <- src/ibm3d_fracstep.c:172-386 */
static void calc_u_star(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const double dt = s->dt;
    const double *xc = s->xc, *yc = s->yc, *zc = s->zc;
    double *u[4] = {NULL, s->u1, s->u2, s->u3};
    double *us[4] = {NULL, s->u1_star, s->u2_star, s->u3_star};
    double *N[4] = {NULL, s->N1, s->N2, s->N3};
    double *Np[4] = {NULL, s->N1_prev, s->N2_prev, s->N3_prev};
    const size_t P = ncell(s);

    IBMSolver_calc_tau_r(s);

    for (int l = 1; l <= 3; l++) {
        memset(s->b, 0, sizeof(double) * P);
        #pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < Nx; i++)
            for (int j = 0; j < Ny; j++)
                for (int k = 0; k < Nz; k++) {
                    if (C3(s, s->flag, i, j, k) != FLAG_FLUID) continue;
                    const double *p = s->p;
                    double dpdxl;
                    if (l == 1) dpdxl = (C3(s, p, i + 1, j, k) - C3(s, p, i - 1, j, k)) / (C1(xc, i + 1) - C1(xc, i - 1));
                    else if (l == 2) dpdxl = (C3(s, p, i, j + 1, k) - C3(s, p, i, j - 1, k)) / (C1(yc, j + 1) - C1(yc, j - 1));
                    else dpdxl = (C3(s, p, i, j, k + 1) - C3(s, p, i, j, k - 1)) / (C1(zc, k + 1) - C1(zc, k - 1));
                    const double *t1 = s->tau[tslot(l, 1)], *t2 = s->tau[tslot(l, 2)], *t3 = s->tau[tslot(l, 3)];
                    const double div_tau
                        = (C3(s, t1, i + 1, j, k) - C3(s, t1, i - 1, j, k)) / (C1(xc, i + 1) - C1(xc, i - 1))
                        + (C3(s, t2, i, j + 1, k) - C3(s, t2, i, j - 1, k)) / (C1(yc, j + 1) - C1(yc, j - 1))
                        + (C3(s, t3, i, j, k + 1) - C3(s, t3, i, j, k - 1)) / (C1(zc, k + 1) - C1(zc, k - 1));
                    const double *a = u[l];
                    s->b[CI(s, i, j, k)]
                        = -dt / 2 * (3 * C3(s, N[l], i, j, k) - C3(s, Np[l], i, j, k))
                        - dt * dpdxl
                        + (1 - C1(s->kx_W, i) - C1(s->kx_E, i) - C1(s->ky_S, j) - C1(s->ky_N, j)
                           - C1(s->kz_D, k) - C1(s->kz_U, k)) * C3(s, a, i, j, k)
                        + C1(s->kx_W, i) * C3(s, a, i - 1, j, k) + C1(s->kx_E, i) * C3(s, a, i + 1, j, k)
                        + C1(s->ky_S, j) * C3(s, a, i, j - 1, k) + C1(s->ky_N, j) * C3(s, a, i, j + 1, k)
                        + C1(s->kz_D, k) * C3(s, a, i, j, k - 1) + C1(s->kz_U, k) * C3(s, a, i, j, k + 1)
                        + dt * ext_force(s, l, s->time, C1(xc, i), C1(yc, j), C1(zc, k))
                        - dt * div_tau;
                }
        /* velocity BC rows, :271-352 */
        for (int ax = 0; ax < 3; ax++)
            for (int hi = 0; hi < 2; hi++) {
                const int bci = face_bc(ax, hi);
                const bool conv = s->bc[bci].type == BC_CONVECTIVE_OUTFLOW;
                if (s->bc[bci].type != BC_VELOCITY_COMPONENT && !conv) continue;
                const FacePos f = face_pos(axis_n(s, ax), hi);
                int a0, a1, b0, b1;
                face_ranges(s, ax, &a0, &a1, &b0, &b1);
                for (int t1 = a0; t1 < a1; t1++)
                    for (int t2 = b0; t2 < b1; t2++) {
                        double x, y, z;
                        face_coord(s, ax, hi, t1, t2, &x, &y, &z);
                        const double g = conv ? conv_val(s, bci, l, t1, t2)
                                              : bc_val_u(s, l, bci, s->time, x, y, z);
                        s->b[face_cell(s, ax, f.o1, t1, t2)] = g;
                        s->b[face_cell(s, ax, f.o2, t1, t2)] = g;
                    }
            }
        double norm;
        const int it = oracle_bicgstab(&s->A[l - 1], s->b, s->x, s->tol, 200, 0, s->work, &norm);
        if (!isfinite(norm)) {  /* src/ibm3d_fracstep.c:363-366 */
            fprintf(stderr, "error: floating pointer error raised in u%d_star\n", l);
            abort();
        }
        if (norm > s->tol) fprintf(stderr, "warning: u%d_star did not converge\n", l);
        memcpy(us[l], s->x, sizeof(double) * P);
        s->final_norm[l - 1] = norm;
        s->final_iter[l - 1] = it;
    }
}

/* <- src/ibm3d_fracstep.c:388-412 */
static void calc_u_tilde(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const double dt = s->dt;
    const double *p = s->p;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = -1; i < Nx + 1; i++)
        for (int j = -1; j < Ny + 1; j++)
            for (int k = -1; k < Nz + 1; k++) {
                C3(s, s->u1_tilde, i, j, k) = C3(s, s->u1_star, i, j, k)
                    + dt * (C3(s, p, i + 1, j, k) - C3(s, p, i - 1, j, k)) / (C1(s->xc, i + 1) - C1(s->xc, i - 1));
                C3(s, s->u2_tilde, i, j, k) = C3(s, s->u2_star, i, j, k)
                    + dt * (C3(s, p, i, j + 1, k) - C3(s, p, i, j - 1, k)) / (C1(s->yc, j + 1) - C1(s->yc, j - 1));
                C3(s, s->u3_tilde, i, j, k) = C3(s, s->u3_star, i, j, k)
                    + dt * (C3(s, p, i, j, k + 1) - C3(s, p, i, j, k - 1)) / (C1(s->zc, k + 1) - C1(s->zc, k - 1));
            }
}

static bool fg_pair(int a, int b) {
    return (a == FLAG_FLUID && b == FLAG_GHOST) || (a == FLAG_GHOST && b == FLAG_FLUID);
}

/* <- src/ibm3d_fracstep.c:414-487 */
static void calc_U_star(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const double dt = s->dt;
    const double *dx = s->dx, *dy = s->dy, *dz = s->dz, *xc = s->xc, *yc = s->yc, *zc = s->zc;
    const double *p = s->p;
    const int *fl = s->flag;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i <= Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++) {
                const double *a = fg_pair(C3(s, fl, i - 1, j, k), C3(s, fl, i, j, k)) ? s->u1_star : s->u1_tilde;
                double v = (C3(s, a, i - 1, j, k) * C1(dx, i) + C3(s, a, i, j, k) * C1(dx, i - 1))
                    / (C1(dx, i - 1) + C1(dx, i));
                if (a == s->u1_tilde)
                    v = v - dt * (C3(s, p, i, j, k) - C3(s, p, i - 1, j, k)) / (C1(xc, i) - C1(xc, i - 1));
                XS(s, s->U1_star, i, j, k) = v;
            }
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j <= Ny; j++)
            for (int k = 0; k < Nz; k++) {
                const double *a = fg_pair(C3(s, fl, i, j - 1, k), C3(s, fl, i, j, k)) ? s->u2_star : s->u2_tilde;
                double v = (C3(s, a, i, j - 1, k) * C1(dy, j) + C3(s, a, i, j, k) * C1(dy, j - 1))
                    / (C1(dy, j - 1) + C1(dy, j));
                if (a == s->u2_tilde)
                    v = v - dt * (C3(s, p, i, j, k) - C3(s, p, i, j - 1, k)) / (C1(yc, j) - C1(yc, j - 1));
                YS(s, s->U2_star, i, j, k) = v;
            }
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k <= Nz; k++) {
                const double *a = fg_pair(C3(s, fl, i, j, k - 1), C3(s, fl, i, j, k)) ? s->u3_star : s->u3_tilde;
                double v = (C3(s, a, i, j, k - 1) * C1(dz, k) + C3(s, a, i, j, k) * C1(dz, k - 1))
                    / (C1(dz, k - 1) + C1(dz, k));
                if (a == s->u3_tilde)
                    v = v - dt * (C3(s, p, i, j, k) - C3(s, p, i, j, k - 1)) / (C1(zc, k) - C1(zc, k - 1));
                ZS(s, s->U3_star, i, j, k) = v;
            }
}

/* <- src/ibm3d_fracstep.c:489-707 */
static void calc_p_prime(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const size_t P = ncell(s);
    const bool sing = p_singular(s);
    memset(s->b, 0, sizeof(double) * P);
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++) {
                const size_t q = CI(s, i, j, k);
                if (sing && i == 0 && j == 0 && k == 0) { s->b[q] = 0; continue; }
                if (C3(s, s->flag, i, j, k) != FLAG_FLUID) continue;
                double v = -1 / (2 * s->Re) * (
                    (XS(s, s->U1_star, i + 1, j, k) - XS(s, s->U1_star, i, j, k)) / C1(s->dx, i)
                    + (YS(s, s->U2_star, i, j + 1, k) - YS(s, s->U2_star, i, j, k)) / C1(s->dy, j)
                    + (ZS(s, s->U3_star, i, j, k + 1) - ZS(s, s->U3_star, i, j, k)) / C1(s->dz, k));
                s->b[q] = v / C3(s, s->p_coeffsum, i, j, k);
            }
    for (int ax = 0; ax < 3; ax++)
        for (int hi = 0; hi < 2; hi++) {
            const int bci = face_bc(ax, hi);
            if (s->bc[bci].type != BC_PRESSURE && s->bc[bci].type != BC_VELOCITY_PERIODIC) continue;
            const FacePos f = face_pos(axis_n(s, ax), hi);
            int a0, a1, b0, b1;
            face_ranges(s, ax, &a0, &a1, &b0, &b1);
            for (int t1 = a0; t1 < a1; t1++)
                for (int t2 = b0; t2 < b1; t2++) {
                    double x, y, z;
                    face_coord(s, ax, hi, t1, t2, &x, &y, &z);
                    const double g = bc_val_p(s, bci, s->time + s->dt, x, y, z) - bc_val_p(s, bci, s->time, x, y, z);
                    s->b[face_cell(s, ax, f.o1, t1, t2)] = g;
                    s->b[face_cell(s, ax, f.o2, t1, t2)] = g;
                }
        }
    double norm;
    const int use_ilu = s->precond_type == PRECOND_AMG && s->A[3].has_ilu;
    const int it = oracle_bicgstab(&s->A[3], s->b, s->x, s->tol, 1000, use_ilu, s->work, &norm);
    if (!isfinite(norm)) {  /* src/ibm3d_fracstep.c:673-676 */
        fprintf(stderr, "error: floating pointer error raised in p_prime\n");
        abort();
    }
    if (norm > s->tol) fprintf(stderr, "warning: p_prime did not converge\n");
    memcpy(s->p_prime, s->x, sizeof(double) * P);
    s->final_norm[3] = norm;
    s->final_iter[3] = it;
}

#define SWAPP(a, b) do { double *t_ = (a); (a) = (b); (b) = t_; } while (0)

/* <- src/ibm3d_fracstep.c:709-805 */
static void update_next(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const double dt = s->dt;
    const double *pp = s->p_prime;
    const int *fl = s->flag;
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++) {
                const int f = C3(s, fl, i, j, k);
                C3(s, s->p_next, i, j, k) = (f == FLAG_FLUID || f == FLAG_GHOST)
                    ? C3(s, s->p, i, j, k) + C3(s, pp, i, j, k) : NAN;
                if (f == FLAG_FLUID) {
                    C3(s, s->u1_next, i, j, k) = C3(s, s->u1_star, i, j, k)
                        - dt * (C3(s, pp, i + 1, j, k) - C3(s, pp, i - 1, j, k)) / (C1(s->xc, i + 1) - C1(s->xc, i - 1));
                    C3(s, s->u2_next, i, j, k) = C3(s, s->u2_star, i, j, k)
                        - dt * (C3(s, pp, i, j + 1, k) - C3(s, pp, i, j - 1, k)) / (C1(s->yc, j + 1) - C1(s->yc, j - 1));
                    C3(s, s->u3_next, i, j, k) = C3(s, s->u3_star, i, j, k)
                        - dt * (C3(s, pp, i, j, k + 1) - C3(s, pp, i, j, k - 1)) / (C1(s->zc, k + 1) - C1(s->zc, k - 1));
                } else if (f == FLAG_SOLID) {
                    C3(s, s->u1_next, i, j, k) = C3(s, s->u2_next, i, j, k) = C3(s, s->u3_next, i, j, k) = NAN;
                }
            }
    #pragma omp parallel for schedule(static)
    for (int i = 0; i <= Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++)
                XS(s, s->U1_next, i, j, k) = (C3(s, fl, i - 1, j, k) != FLAG_SOLID && C3(s, fl, i, j, k) != FLAG_SOLID)
                    ? XS(s, s->U1_star, i, j, k) - dt * (C3(s, pp, i, j, k) - C3(s, pp, i - 1, j, k)) / (C1(s->xc, i) - C1(s->xc, i - 1))
                    : NAN;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j <= Ny; j++)
            for (int k = 0; k < Nz; k++)
                YS(s, s->U2_next, i, j, k) = (C3(s, fl, i, j - 1, k) != FLAG_SOLID && C3(s, fl, i, j, k) != FLAG_SOLID)
                    ? YS(s, s->U2_star, i, j, k) - dt * (C3(s, pp, i, j, k) - C3(s, pp, i, j - 1, k)) / (C1(s->yc, j) - C1(s->yc, j - 1))
                    : NAN;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k <= Nz; k++)
                ZS(s, s->U3_next, i, j, k) = (C3(s, fl, i, j, k - 1) != FLAG_SOLID && C3(s, fl, i, j, k) != FLAG_SOLID)
                    ? ZS(s, s->U3_star, i, j, k) - dt * (C3(s, pp, i, j, k) - C3(s, pp, i, j, k - 1)) / (C1(s->zc, k) - C1(s->zc, k - 1))
                    : NAN;
    SWAPP(s->u1, s->u1_next); SWAPP(s->u2, s->u2_next); SWAPP(s->u3, s->u3_next);
    SWAPP(s->U1, s->U1_next); SWAPP(s->U2, s->U2_next); SWAPP(s->U3, s->U3_next);
    SWAPP(s->p, s->p_next);
    SWAPP(s->N1_prev, s->N1); SWAPP(s->N2_prev, s->N2); SWAPP(s->N3_prev, s->N3);
}

/* <- src/ibm3d_fracstep.c:807-845 */
static void interp_stag_vel(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const double *dx = s->dx, *dy = s->dy, *dz = s->dz;
    for (int i = 0; i <= Nx; i++)
        for (int j = -2; j < Ny + 2; j++)
            for (int k = -2; k < Nz + 2; k++)
                XS(s, s->U1, i, j, k) = (C3(s, s->u1, i - 1, j, k) * C1(dx, i) + C3(s, s->u1, i, j, k) * C1(dx, i - 1))
                    / (C1(dx, i - 1) + C1(dx, i));
    for (int i = -2; i < Nx + 2; i++)
        for (int j = 0; j <= Ny; j++)
            for (int k = -2; k < Nz + 2; k++)
                YS(s, s->U2, i, j, k) = (C3(s, s->u2, i, j - 1, k) * C1(dy, j) + C3(s, s->u2, i, j, k) * C1(dy, j - 1))
                    / (C1(dy, j - 1) + C1(dy, j));
    for (int i = -2; i < Nx + 2; i++)
        for (int j = -2; j < Ny + 2; j++)
            for (int k = 0; k <= Nz; k++)
                ZS(s, s->U3, i, j, k) = (C3(s, s->u3, i, j, k - 1) * C1(dz, k) + C3(s, s->u3, i, j, k) * C1(dz, k - 1))
                    / (C1(dz, k - 1) + C1(dz, k));
}

/* Outer ghost layers of u  <- src/ibm3d_fracstep.c:883-1513 */
static void outer_u_face(IBMSolver *s, int ax, int hi) {
    const int bci = face_bc(ax, hi);
    const IBMSolverBCType bt = s->bc[bci].type;
    const int N = axis_n(s, ax);
    const double *d = axis_d(s, ax), *c = axis_c(s, ax);
    const FacePos f = face_pos(N, hi);
    double *u[4] = {NULL, s->u1, s->u2, s->u3};
    int a0, a1, b0, b1;
    face_ranges(s, ax, &a0, &a1, &b0, &b1);
    for (int layer = 1; layer <= 2; layer++) {
        const int o = layer == 1 ? f.o1 : f.o2;
        const int q = layer == 1 ? f.q0 : f.q1;
        double ao, bi;
        dirichlet_dist(d, f, layer, &ao, &bi);
        /* extrapolation spacings and anchors of BC_PRESSURE */
        int n1, n2;
        double so, si;
        if (layer == 1) {
            n1 = f.q0; n2 = f.q1;
            so = hi ? C1(c, f.o1) - C1(c, f.q0) : C1(c, f.q0) - C1(c, f.o1);
            si = hi ? C1(c, f.q0) - C1(c, f.q1) : C1(c, f.q1) - C1(c, f.q0);
        } else {
            n1 = f.o1; n2 = f.q0;
            so = hi ? C1(c, f.o2) - C1(c, f.o1) : C1(c, f.o1) - C1(c, f.o2);
            si = hi ? C1(c, f.o1) - C1(c, f.q0) : C1(c, f.q0) - C1(c, f.o1);
        }
        for (int t1 = a0; t1 < a1; t1++)
            for (int t2 = b0; t2 < b1; t2++) {
                const size_t co = face_cell(s, ax, o, t1, t2);
                double x, y, z;
                face_coord(s, ax, hi, t1, t2, &x, &y, &z);
                for (int l = 1; l <= 3; l++) {
                    double *a = u[l];
                    switch (bt) {
                    case BC_VELOCITY_COMPONENT:
                    case BC_CONVECTIVE_OUTFLOW: {
                        const double g = bt == BC_CONVECTIVE_OUTFLOW ? conv_val(s, bci, l, t1, t2)
                                                                     : bc_val_u(s, l, bci, s->time, x, y, z);
                        a[co] = (ao + bi) / bi * g - ao / bi * a[face_cell(s, ax, q, t1, t2)];
                        break;
                    }
                    case BC_PRESSURE:
                        a[co] = (so + si) / si * a[face_cell(s, ax, n1, t1, t2)]
                              - so / si * a[face_cell(s, ax, n2, t1, t2)];
                        break;
                    case BC_FREE_SLIP_WALL:
                        if (l == ax + 1) a[co] = -ao / bi * a[face_cell(s, ax, q, t1, t2)];
                        else a[co] = a[face_cell(s, ax, q, t1, t2)];
                        break;
                    case BC_ALL_PERIODIC:
                    case BC_VELOCITY_PERIODIC:
                        a[co] = a[face_cell(s, ax, hi ? o - N : o + N, t1, t2)];
                        break;
                    default:;
                    }
                }
            }
    }
}

/* Outer ghost layers of p  <- src/ibm3d_fracstep.c:1515-1905, including the
   Down pressure branch that spaces with dx instead of dz (:1792-1800). */
static void outer_p_face(IBMSolver *s, int ax, int hi) {
    const int bci = face_bc(ax, hi);
    const IBMSolverBCType bt = s->bc[bci].type;
    const int N = axis_n(s, ax);
    const double *d = (ax == 2 && !hi && bt != BC_CONVECTIVE_OUTFLOW) ? s->dx : axis_d(s, ax);
    const FacePos f = face_pos(N, hi);
    double *p = s->p;
    int a0, a1, b0, b1;
    face_ranges(s, ax, &a0, &a1, &b0, &b1);
    for (int layer = 1; layer <= 2; layer++) {
        const int o = layer == 1 ? f.o1 : f.o2;
        const int q = layer == 1 ? f.q0 : f.q1;
        double ao, bi;
        dirichlet_dist(d, f, layer, &ao, &bi);
        for (int t1 = a0; t1 < a1; t1++)
            for (int t2 = b0; t2 < b1; t2++) {
                const size_t co = face_cell(s, ax, o, t1, t2);
                switch (bt) {
                case BC_VELOCITY_COMPONENT:
                case BC_STATIONARY_WALL:
                case BC_FREE_SLIP_WALL:
                    p[co] = p[face_cell(s, ax, q, t1, t2)];
                    break;
                case BC_PRESSURE:
                case BC_VELOCITY_PERIODIC: {
                    double x, y, z;
                    face_coord(s, ax, hi, t1, t2, &x, &y, &z);
                    const double g = bc_val_p(s, bci, s->time, x, y, z);
                    p[co] = (ao + bi) / bi * g - ao / bi * p[face_cell(s, ax, q, t1, t2)];
                    break;
                }
                case BC_ALL_PERIODIC:
                    p[co] = p[face_cell(s, ax, hi ? o - N : o + N, t1, t2)];
                    break;
                case BC_CONVECTIVE_OUTFLOW:  /* p = 0 on the face */
                    p[co] = -ao / bi * p[face_cell(s, ax, q, t1, t2)];
                    break;
                default:;
                }
            }
    }
}

/* <- src/ibm3d_fracstep.c:855-1906 */
static void update_outer(IBMSolver *s) {
    for (int ax = 0; ax < 3; ax++)
        for (int hi = 0; hi < 2; hi++) outer_u_face(s, ax, hi);
    for (int ax = 0; ax < 3; ax++)
        for (int hi = 0; hi < 2; hi++) outer_p_face(s, ax, hi);
}

/* In-place lexicographic ghost update  <- src/ibm3d_fracstep.c:1915-1976 */
static void update_ghost(IBMSolver *s) {
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    int idx[8][3];
    double w[8];
    for (int i = 0; i < Nx; i++)
        for (int j = 0; j < Ny; j++)
            for (int k = 0; k < Nz; k++) {
                const int f = C3(s, s->flag, i, j, k);
                if (f == FLAG_GHOST) {
                    IBMSolver_ghost_interp(s, i, j, k, idx, w);
                    double sum = 0, lhs_u = 1, lhs_p = 1;
                    double s1 = 0, s2 = 0, s3 = 0, sp = 0;
                    for (int l = 0; l < 8; l++) {
                        if (C3(s, s->flag, idx[l][0], idx[l][1], idx[l][2]) == FLAG_SOLID) w[l] = 0;
                        else sum += w[l];
                    }
                    for (int l = 0; l < 8; l++) {
                        w[l] /= sum;
                        const int a = idx[l][0], b = idx[l][1], c = idx[l][2];
                        if (a == i && b == j && c == k) {
                            lhs_u += w[l];
                            lhs_p -= w[l];
                        } else if (C3(s, s->flag, a, b, c) != FLAG_SOLID) {
                            s1 += w[l] * C3(s, s->u1, a, b, c);
                            s2 += w[l] * C3(s, s->u2, a, b, c);
                            s3 += w[l] * C3(s, s->u3, a, b, c);
                            sp += w[l] * C3(s, s->p, a, b, c);
                        }
                    }
                    C3(s, s->u1, i, j, k) = -s1 / lhs_u;
                    C3(s, s->u2, i, j, k) = -s2 / lhs_u;
                    C3(s, s->u3, i, j, k) = -s3 / lhs_u;
                    if (lhs_p != 0) C3(s, s->p, i, j, k) = sp / lhs_p;
                } else if (f == FLAG_SOLID) {
                    C3(s, s->u1, i, j, k) = C3(s, s->u2, i, j, k) = C3(s, s->u3, i, j, k) = C3(s, s->p, i, j, k) = NAN;
                }
            }
}

/* <- src/ibm3d_fracstep.c:38-115 */
void IBMSolver_iterate(IBMSolver *s, int num_time_steps, bool verbose) {
    const double start_time = s->time;
    if (!s->conv_ready) conv_outflow(s, 1);
    s->conv_ready = true;
    update_outer(s);
    interp_stag_vel(s);
    update_ghost(s);
    calc_N(s);
    SWAPP(s->N1, s->N1_prev); SWAPP(s->N2, s->N2_prev); SWAPP(s->N3, s->N3_prev);
    for (int i = 0; i < num_time_steps;) {
        calc_N(s);
        conv_outflow(s, 0);
        calc_u_star(s);
        calc_u_tilde(s);
        calc_U_star(s);
        calc_p_prime(s);
        update_next(s);
        i++;
        s->iter++;
        s->time = start_time + s->dt * i;
        update_outer(s);
        update_ghost(s);
        if (verbose) {
            if (s->iter % 10 == 1 || i == 1) {
                printf("\n  iter       u1 res       u2 res       u3 res        p res\n");
                printf("---------------------------------------------------------\n");
            }
            printf("%6d   %10.4e   %10.4e   %10.4e   %10.4e\n", s->iter, s->final_norm[0],
                   s->final_norm[1], s->final_norm[2], s->final_norm[3]);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* initialisation  <- src/ibm3d_init.c:15-37, 237-262                        */
/* ------------------------------------------------------------------------ */

void IBMSolver_init_flow_const(IBMSolver *s, const double u1, const double u2,
                               const double u3, const double p) {
    s->iter = 0;
    s->time = 0;
    s->conv_ready = false;
    for (size_t q = 0; q < ncell(s); q++) { s->u1[q] = u1; s->u2[q] = u2; s->u3[q] = u3; s->p[q] = p; }
}

void IBMSolver_init_flow_func(IBMSolver *s, IBMSolverInitFunc f1, IBMSolverInitFunc f2,
                              IBMSolverInitFunc f3, IBMSolverInitFunc fp) {
    s->conv_ready = false;
    for (int i = -2; i < s->Nx + 2; i++)
        for (int j = -2; j < s->Ny + 2; j++)
            for (int k = -2; k < s->Nz + 2; k++) {
                const double x = C1(s->xc, i), y = C1(s->yc, j), z = C1(s->zc, k);
                C3(s, s->u1, i, j, k) = f1(x, y, z);
                C3(s, s->u2, i, j, k) = f2(x, y, z);
                C3(s, s->u3, i, j, k) = f3(x, y, z);
                C3(s, s->p, i, j, k) = fp(x, y, z);
            }
}

void IBMSolver_init_flow_file(IBMSolver *s, const char *filename) {
    (void)s; (void)filename;
    fatal("the oracle does not read netCDF restarts");
}

void IBMSolver_export_result(IBMSolver *s, const char *f) { (void)s; (void)f; }
void IBMSolver_export_lvset_flag(IBMSolver *s, const char *f) { (void)s; (void)f; }
void IBMSolver_export_intermediate(IBMSolver *s, const char *f) { (void)s; (void)f; }

/* ------------------------------------------------------------------------ */
/* extensions: field access for the tests                                    */
/* ------------------------------------------------------------------------ */

void IBMSolver_get_field(IBMSolver *s, int field, double *out) {
    const size_t P = ncell(s);
    const size_t PX = (size_t)(s->Nx + 1) * (s->Ny + 4) * (s->Nz + 4);
    const size_t PY = (size_t)(s->Nx + 4) * (s->Ny + 1) * (s->Nz + 4);
    const size_t PZ = (size_t)(s->Nx + 4) * (s->Ny + 4) * (s->Nz + 1);
    const double *src = NULL;
    size_t n = P;
    switch (field) {
    case IBM3D_FIELD_U1: src = s->u1; break;
    case IBM3D_FIELD_U2: src = s->u2; break;
    case IBM3D_FIELD_U3: src = s->u3; break;
    case IBM3D_FIELD_P: src = s->p; break;
    case IBM3D_FIELD_U1_STAR: src = s->u1_star; break;
    case IBM3D_FIELD_U2_STAR: src = s->u2_star; break;
    case IBM3D_FIELD_U3_STAR: src = s->u3_star; break;
    case IBM3D_FIELD_P_PRIME: src = s->p_prime; break;
    /* after a step the last N lives in N_prev (src/ibm3d_fracstep.c:802-804) */
    case IBM3D_FIELD_N1: src = s->N1_prev; break;
    case IBM3D_FIELD_N2: src = s->N2_prev; break;
    case IBM3D_FIELD_N3: src = s->N3_prev; break;
    case IBM3D_FIELD_LVSET: src = s->lvset; break;
    case IBM3D_FIELD_FACE_U1: src = s->U1; n = PX; break;
    case IBM3D_FIELD_FACE_U2: src = s->U2; n = PY; break;
    case IBM3D_FIELD_FACE_U3: src = s->U3; n = PZ; break;
    default:
        if (field >= IBM3D_FIELD_TAU11 && field <= IBM3D_FIELD_TAU33) { src = s->tau[field - IBM3D_FIELD_TAU11]; break; }
        fatal("unknown field");
    }
    memcpy(out, src, n * sizeof(double));
}

void IBMSolver_set_field(IBMSolver *s, int field, const double *in) {
    double *dst = field == IBM3D_FIELD_U1 ? s->u1 : field == IBM3D_FIELD_U2 ? s->u2
                : field == IBM3D_FIELD_U3 ? s->u3 : field == IBM3D_FIELD_P ? s->p : NULL;
    if (!dst) fatal("set_field supports u1, u2, u3, p");
    memcpy(dst, in, ncell(s) * sizeof(double));
    s->conv_ready = false;
}

void IBMSolver_get_flag(IBMSolver *s, int *out) { memcpy(out, s->flag, ncell(s) * sizeof(int)); }
int IBMSolver_get_iter(IBMSolver *s) { return s->iter; }
double IBMSolver_get_time(IBMSolver *s) { return s->time; }
void IBMSolver_get_last_norms(IBMSolver *s, double out[4]) { memcpy(out, s->final_norm, sizeof(s->final_norm)); }
void IBMSolver_get_last_iters(IBMSolver *s, int out[4]) { memcpy(out, s->final_iter, sizeof(s->final_iter)); }

int IBMSolver_get_ghost_cells(IBMSolver *s, int *out) {
    int n = 0;
    for (int i = 0; i < s->Nx; i++)
        for (int j = 0; j < s->Ny; j++)
            for (int k = 0; k < s->Nz; k++)
                if (C3(s, s->flag, i, j, k) == FLAG_GHOST) {
                    if (out) { out[3 * n] = i; out[3 * n + 1] = j; out[3 * n + 2] = k; }
                    n++;
                }
    return n;
}

/* max |div U| over the fluid cells whose row of the pressure system is the
   continuity equation: where no pressure-type boundary exists, the row of
   cell (0,0,0) pins p' instead (src/ibm3d_setup.c:1598-1624) and that cell
   takes up the residual of every other row, so it is left out. */
double IBMSolver_max_div(IBMSolver *s) {
    double m = 0;
    const bool sing = p_singular(s);
    for (int i = 0; i < s->Nx; i++)
        for (int j = 0; j < s->Ny; j++)
            for (int k = 0; k < s->Nz; k++) {
                if (C3(s, s->flag, i, j, k) != FLAG_FLUID) continue;
                if (sing && i == 0 && j == 0 && k == 0) continue;
                const double d = (XS(s, s->U1, i + 1, j, k) - XS(s, s->U1, i, j, k)) / C1(s->dx, i)
                    + (YS(s, s->U2, i, j + 1, k) - YS(s, s->U2, i, j, k)) / C1(s->dy, j)
                    + (ZS(s, s->U3, i, j, k + 1) - ZS(s, s->U3, i, j, k)) / C1(s->dz, k);
                if (fabs(d) > m) m = fabs(d);
            }
    return m;
}

/* Force on the immersed body per unit density, from the staircase surface
   (faces between an interior ghost cell and an interior fluid 6-neighbour):
   -p_f n A with p_f interpolated linearly to the face position, plus
   (1/Re) (u_b(fluid) - u_b(ghost)) / |x_fluid - x_ghost| A for the tangential
   components (n: outward body normal, ghost -> fluid). An extension with no
   reference counterpart; ghost cells in lexicographic order. */
void IBMSolver_calc_body_force(IBMSolver *s, double out[3]) {
    const int n[3] = {s->Nx, s->Ny, s->Nz};
    const double *C[3] = {s->xc, s->yc, s->zc}, *D[3] = {s->dx, s->dy, s->dz};
    const double *u[3] = {s->u1, s->u2, s->u3};
    const double mu = 1 / s->Re;
    out[0] = out[1] = out[2] = 0;
    for (int i = 0; i < n[0]; i++)
        for (int j = 0; j < n[1]; j++)
            for (int k = 0; k < n[2]; k++) {
                if (C3(s, s->flag, i, j, k) != FLAG_GHOST) continue;
                const int id[3] = {i, j, k};
                for (int a = 0; a < 3; a++)
                    for (int sd = -1; sd <= 1; sd += 2) {
                        int f[3] = {i, j, k};
                        f[a] += sd;
                        if (f[a] < 0 || f[a] >= n[a]) continue;
                        if (C3(s, s->flag, f[0], f[1], f[2]) != FLAG_FLUID) continue;
                        double area = 1;
                        for (int b = 0; b < 3; b++)
                            if (b != a) area *= C1(D[b], id[b]);
                        const double xg = C1(C[a], id[a]), xf = C1(C[a], f[a]);
                        const double xface = xg + sd * C1(D[a], id[a]) / 2;
                        const double pg = C3(s, s->p, i, j, k), pf = C3(s, s->p, f[0], f[1], f[2]);
                        out[a] -= (pg + (pf - pg) * (xface - xg) / (xf - xg)) * sd * area;
                        for (int b = 0; b < 3; b++)
                            if (b != a)
                                out[b] += mu * (C3(s, u[b], f[0], f[1], f[2]) - C3(s, u[b], i, j, k))
                                    / fabs(xf - xg) * area;
                    }
            }
}

/* Applies system `type` (1..4) of the assembled reference matrix to x. Test
   hook for operator identities (A_u 1 = 1, A_p 1 = 0 on fluid rows). */
void oracle_apply_matrix(IBMSolver *s, int type, const double *x, double *y) {
    spmv(&s->A[type - 1], x, y);
}

/* Copies row `row` of system `type`; returns its entry count. */
int oracle_matrix_row(IBMSolver *s, int type, int row, int *cols, double *vals) {
    const EllMat *A = &s->A[type - 1];
    for (int l = 0; l < A->nnz[row]; l++) {
        cols[l] = A->col[(size_t)9 * row + l];
        vals[l] = A->val[(size_t)9 * row + l];
    }
    return A->nnz[row];
}

void oracle_get_coords(IBMSolver *s, double *xc, double *yc, double *zc, double *dx, double *dy, double *dz) {
    memcpy(xc, s->xc, (s->Nx + 4) * sizeof(double)); memcpy(dx, s->dx, (s->Nx + 4) * sizeof(double));
    memcpy(yc, s->yc, (s->Ny + 4) * sizeof(double)); memcpy(dy, s->dy, (s->Ny + 4) * sizeof(double));
    memcpy(zc, s->zc, (s->Nz + 4) * sizeof(double)); memcpy(dz, s->dz, (s->Nz + 4) * sizeof(double));
}

void oracle_get_kcoef(IBMSolver *s, double *out6) {
    /* kx_W, kx_E (Nx+4 each), ky_S, ky_N (Ny+4), kz_D, kz_U (Nz+4), packed */
    double *o = out6;
    memcpy(o, s->kx_W, (s->Nx + 4) * sizeof(double)); o += s->Nx + 4;
    memcpy(o, s->kx_E, (s->Nx + 4) * sizeof(double)); o += s->Nx + 4;
    memcpy(o, s->ky_S, (s->Ny + 4) * sizeof(double)); o += s->Ny + 4;
    memcpy(o, s->ky_N, (s->Ny + 4) * sizeof(double)); o += s->Ny + 4;
    memcpy(o, s->kz_D, (s->Nz + 4) * sizeof(double)); o += s->Nz + 4;
    memcpy(o, s->kz_U, (s->Nz + 4) * sizeof(double));
}

/* Runs only the boundary-layer pass (update_outer) on the current fields. */
void oracle_update_outer(IBMSolver *s) { update_outer(s); }
/* Runs only the ghost-cell pass (update_ghost) on the current fields. */
void oracle_update_ghost(IBMSolver *s) { update_ghost(s); }
/* Replaces the level set (reference layout, P values) and recomputes the
   flags from it by the rule of calc_lvset_flag; the assembled matrices are
   left as they were. Test hook for the ghost-cell pass on bodies that have no
   STL (e.g. a plane). */
void oracle_set_lvset(IBMSolver *s, const double *phi) {
    memcpy(s->lvset, phi, ncell(s) * sizeof(double));
    calc_flag(s);
}
/* Runs calc_N on the current fields after interp_stag_vel; the result is
   returned through IBM3D_FIELD_N* (stored in the N_prev slot). */
void oracle_calc_N(IBMSolver *s) {
    interp_stag_vel(s);
    calc_N(s);
    SWAPP(s->N1, s->N1_prev); SWAPP(s->N2, s->N2_prev); SWAPP(s->N3, s->N3_prev);
}
