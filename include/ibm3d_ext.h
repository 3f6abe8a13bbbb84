/*
 * ibm3d_ext.h - extensions of the solver API. None of these has a counterpart
 * in the reference; they exist for callers that need what the reference gets
 * by reading the public IBMSolver struct (src/ibm3d.h:100-213: fields, flags,
 * residual norms) and for the GPU-only controls (precision mode, timing).
 *
 * Array layouts returned by IBMSolver_get_field are the reference's host
 * layouts (src/utils.h:8-16), so a caller can compare against the reference
 * struct fields element by element:
 *   cell fields   (Nx+4)(Ny+4)(Nz+4), index ((i+2)*(Ny+4) + j+2)*(Nz+4) + k+2
 *   x-face field  (Nx+1)(Ny+4)(Nz+4), index (i*(Ny+4) + j+2)*(Nz+4) + k+2
 *   y-face field  (Nx+4)(Ny+1)(Nz+4), index ((i+2)*(Ny+1) + j)*(Nz+4) + k+2
 *   z-face field  (Nx+4)(Ny+4)(Nz+1), index ((i+2)*(Ny+4) + j+2)*(Nz+1) + k
 */
#ifndef IBM3D_EXT_H
#define IBM3D_EXT_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum _ibm3d_field {
    IBM3D_FIELD_U1 = 0, IBM3D_FIELD_U2, IBM3D_FIELD_U3, IBM3D_FIELD_P,
    IBM3D_FIELD_U1_STAR, IBM3D_FIELD_U2_STAR, IBM3D_FIELD_U3_STAR,
    IBM3D_FIELD_P_PRIME,
    IBM3D_FIELD_N1, IBM3D_FIELD_N2, IBM3D_FIELD_N3,
    IBM3D_FIELD_LVSET,
    IBM3D_FIELD_FACE_U1, IBM3D_FIELD_FACE_U2, IBM3D_FIELD_FACE_U3,
    IBM3D_FIELD_TAU11, IBM3D_FIELD_TAU12, IBM3D_FIELD_TAU13,
    IBM3D_FIELD_TAU22, IBM3D_FIELD_TAU23, IBM3D_FIELD_TAU33,
    IBM3D_FIELD_COUNT
} IBM3DField;

typedef enum _ibm3d_precision {
    IBM3D_PREC_FP64 = 0,   /* every array and solve in FP64 */
    IBM3D_PREC_MIXED = 1,  /* FP32 smoother / multigrid inside FP64 refinement */
    IBM3D_PREC_MIXED_KRYLOV = 2, /* FP64 BiCGSTAB, FP32 multigrid preconditioner */
} IBM3DPrecision;

/* Analytic obstacle phi = |x - c| - r (positive in the fluid). Replaces
   IBMSolver_set_obstacle's STL level set for the sphere case. */
void IBMSolver_set_obstacle_sphere(IBMSolver *, double cx, double cy, double cz,
                                   double r);

/* Precision mode of the linear solves (default IBM3D_PREC_FP64). Must be
   called before IBMSolver_assemble. IBM3D_PREC_MIXED needs a pressure-type
   boundary (BC_PRESSURE, BC_VELOCITY_PERIODIC or BC_CONVECTIVE_OUTFLOW) and
   SOLVER_BiCGSTAB; IBMSolver_assemble refuses it otherwise. */
void IBMSolver_set_precision(IBMSolver *, int mode);

/* Multigrid parameters of the pressure preconditioner (no counterpart in
   the reference, whose pressure preconditioner is BoomerAMG). gamma: along
   each axis two neighbouring cells merge on the next level when their
   combined width is at most 2 * gamma times the smallest cell width of the
   level (gamma <= 0: merge all pairs of even-sized axes); pre / post:
   red-black Gauss-Seidel sweeps before / after the coarse correction;
   smoother 0: each sweep is one fused launch (both colours), ghost rows
   relaxed after each sweep; 1: one launch per colour, ghost rows relaxed
   after each colour. Defaults 3.5, 2, 2, 1. gamma must be finite; pre and
   post 0..16, not both 0. Must be called before IBMSolver_assemble. */
void IBMSolver_set_mg_params(IBMSolver *, double gamma, int pre, int post, int smoother);
/* Levels from `from_level` down (when from_level > 0) smooth with one sweep
   before and one after the coarse correction instead of pre/post; 0 smooths
   every level with pre/post. Default 1: the finest level takes pre/post
   sweeps, every coarser level one (cheaper cycles at the same iteration
   count on the sphere and cavity cases); from_level >= 0. Before
   IBMSolver_assemble. */
void IBMSolver_set_mg_light_levels(IBMSolver *, int from_level);
/* Over-relaxation factor omega of the red-black smoother of the pressure
   multigrid (smoother 1): a fluid row takes x + omega (x_GS - x) instead of
   its Gauss-Seidel value x_GS, on every level and in both precisions (the
   ghost rows relaxed inside a colour launch take the Gauss-Seidel values of
   their fluid sources of that colour). 0 < omega < 2; default 1.3 (8 -> 6
   pressure iterations per step on the 256 x 128 x 128 sphere case); 1 gives
   plain Gauss-Seidel. The fused smoother 0 ignores it. Only the
   preconditioner changes: the stopping test is the same. Before
   IBMSolver_assemble. No counterpart in the reference. */
void IBMSolver_set_mg_relaxation(IBMSolver *, double omega);
/* Precision of the pressure multigrid cycle the solver uses now: 32 or 64
   (0 before IBMSolver_assemble). In IBM3D_PREC_MIXED_KRYLOV it is 32 until a
   pressure solve stalls twice (a residual falling less than tenfold over
   three BiCGSTAB iterations, each time followed by a restart from the true
   residual); the solver then keeps the FP64 cycle for its remaining steps.
   No counterpart in the reference. */
int IBMSolver_get_pressure_mg_bits(IBMSolver *);

/* Momentum Helmholtz solver of the IBM3D_PREC_MIXED_KRYLOV mode. 0
   (default): FP32 red-black Gauss-Seidel sweeps inside FP64 iterative
   refinement where the operator is diagonally dominant enough for a few
   sweeps per round (max over cells of sum K / (1 + sum K) squared <= 0.03),
   BiCGSTAB elsewhere and for any component whose residual stops falling;
   1: always BiCGSTAB (FP64, Jacobi preconditioner, FP32 stored directions).
   Both stop on the reference's test ||b - A x|| <= tol ||b|| in FP64. */
void IBMSolver_set_momentum_solver(IBMSolver *, int mode);

/* Copies a field to host memory in the reference layout (see above). Solid
   cells hold NaN where the reference stores NaN. */
void IBMSolver_get_field(IBMSolver *, int field, double *out);
/* Overwrites u1, u2, u3 or p (cell fields only, reference layout). */
void IBMSolver_set_field(IBMSolver *, int field, const double *in);
/* Cell flags FLAG_FLUID/GHOST/SOLID, cell layout. */
void IBMSolver_get_flag(IBMSolver *, int *out);

int IBMSolver_get_iter(IBMSolver *);
double IBMSolver_get_time(IBMSolver *);

/* Final relative residual norms (u1, u2, u3, p) and iteration counts of the
   last time step, as printed by the reference's verbose table. */
void IBMSolver_get_last_norms(IBMSolver *, double out[4]);
void IBMSolver_get_last_iters(IBMSolver *, int out[4]);

/* Interior ghost cells in the reference's lexicographic order (i, j, k; k
   fastest): writes 3*count ints when out != NULL, returns count. */
int IBMSolver_get_ghost_cells(IBMSolver *, int *out);

/* Shapes of the pressure multigrid levels built by IBMSolver_assemble
   (level 0 = the fine grid): writes nx, ny, nz of up to max levels into out
   when out != NULL, returns the number of levels. */
int IBMSolver_get_mg_levels(IBMSolver *, int *out, int max);

/* max |div U| over fluid cells, U the face velocity after projection. Where
   no pressure-type boundary exists, the pressure row of cell (0,0,0) is the
   pin p' = 0 (src/ibm3d_setup.c:1598-1624), not continuity; that cell takes
   up the residual of all other rows and is left out. */
double IBMSolver_max_div(IBMSolver *);

/* Force on the immersed body per unit density (Fx, Fy, Fz), pressure plus
   viscous part, summed over the staircase surface: the faces between an
   interior ghost cell and an interior fluid 6-neighbour. With n the outward
   body normal (ghost -> fluid) and A the face area, each face adds
   -p_f n A (p_f: p linearly interpolated between the two cell centres to the
   face position) and, for the two tangential components b,
   (1/Re) (u_b(fluid) - u_b(ghost)) / |x_fluid - x_ghost| A. Computed on the
   GPU (two fixed-order reduction stages). Drag and lift coefficients follow
   by scaling with 1 / (U^2 A_ref / 2). */
void IBMSolver_calc_body_force(IBMSolver *, double out[3]);

/* Records the body force (as IBMSolver_calc_body_force) after each of the
   first max_steps steps of every IBMSolver_iterate call, on the device with
   no host synchronisation; 0 switches it off. After IBMSolver_assemble. */
void IBMSolver_set_force_history(IBMSolver *, int max_steps);
/* Copies the recorded forces of the last IBMSolver_iterate call, 3 values per
   step, up to max steps into out (when out != NULL); returns the number of
   steps recorded. */
int IBMSolver_get_force_history(IBMSolver *, double *out, int max);

/* Milliseconds of each time step of the last IBMSolver_iterate call: the
   time between the completions of consecutive steps on the device (HIP
   events recorded on the solver's stream at each step's end; the first step
   is measured from the start of the call's device work). Writes up to max
   values, returns the number of steps. */
int IBMSolver_get_step_times(IBMSolver *, double *out, int max);

/* Timing with HIP events on the solver's stream. mode 0: off; 1: every
   kernel launch, multigrid cycles launched eagerly; 2: every launch outside
   the multigrid cycles plus each replay of a cycle's HIP graph as a whole
   ("vcycle_graph"). */
void IBMSolver_enable_timing(IBMSolver *, int mode);
/* Multigrid cycles are captured once per (input, output) pair into a HIP
   graph and replayed (default on). Off: every cycle is launched kernel by
   kernel, e.g. for tracing tools that do not follow graph launches. */
void IBMSolver_set_graph_replay(IBMSolver *, int on);
void IBMSolver_reset_timing(IBMSolver *);
/* Returns the number of timed kernels; fills entry idx when idx is valid:
   name, accumulated milliseconds, launch count, algorithmic bytes per launch (averaged over the launches). */
int IBMSolver_get_kernel_stat(IBMSolver *, int idx, char *name, int namelen,
                              double *total_ms, long *count, double *bytes);

#ifdef __cplusplus
}
#endif

#endif
