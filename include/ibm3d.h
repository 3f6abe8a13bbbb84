/*
 * ibm3d.h - public C API of the MI355X-native fractional-step IBM solver.
 *
 * Drop-in for the reference's solver API (src/ibm3d.h, src/ibm3d_setup.h,
 * src/ibm3d_init.h, src/ibm3d_fracstep.h, src/ibm3d_turb.h,
 * src/ibm3d_export.h). Every type keeps its reference name and numeric values;
 * every function keeps its reference name, argument order and meaning. The
 * solver struct is opaque: the reference exposes its fields (src/ibm3d.h:
 * 100-213) but none of its drivers read them, so callers compile unchanged.
 *
 * One GPU holds the whole domain. IBMSolver_new still takes the MPI process
 * grid so existing drivers compile; it requires num_process == 1 (the
 * reference's Px = Py = Pz = 1 branch) and aborts otherwise.
 *
 * Error behaviour follows the reference: functions return void, fatal errors
 * print to stderr and abort(), solver non-convergence prints a warning and
 * continues (src/ibm3d_fracstep.c:363-369, 673-679).
 */
#ifndef IBM3D_H
#define IBM3D_H

#include <stdbool.h>
#include "geo3d.h"

#ifdef __cplusplus
extern "C" {
#endif

/* (x, y, z) => init value.  <- src/ibm3d.h:14 */
typedef double (*IBMSolverInitFunc)(double, double, double);

/* <- src/ibm3d.h:16-21 */
typedef enum _linear_solver_type {
    SOLVER_AMG,
    SOLVER_PCG,
    SOLVER_BiCGSTAB,
    SOLVER_GMRES,
} IBMSolverLinearSolverType;

/* <- src/ibm3d.h:23-26 */
typedef enum _precond_type {
    PRECOND_NONE,
    PRECOND_AMG,
} IBMSolverPrecondType;

/* Direction bits; bc[] index order is N, E, S, W, D, U.
   <- src/ibm3d.h:28-35, src/utils.c:92-109 */
typedef enum _direction {
    DIR_NORTH = 1 << 0,
    DIR_EAST  = 1 << 1,
    DIR_SOUTH = 1 << 2,
    DIR_WEST  = 1 << 3,
    DIR_DOWN  = 1 << 4,
    DIR_UP    = 1 << 5
} IBMSolverDirection;

/* <- src/ibm3d.h:37-45 */
typedef enum _bc_type {
    BC_VELOCITY_COMPONENT,
    BC_PRESSURE,
    BC_STATIONARY_WALL,
    BC_FREE_SLIP_WALL,
    BC_SYMMETRIC,
    BC_ALL_PERIODIC,
    BC_VELOCITY_PERIODIC,
    /* Extension (no counterpart in the reference): convective outflow. The
       velocity on the face obeys du/dt + Uc du/dn = 0 (first-order upwind
       between the face and the first interior cell, advanced once per step
       before the momentum solve); the pressure on the face is held at 0
       (Dirichlet rows, as BC_PRESSURE with value 0). */
    BC_CONVECTIVE_OUTFLOW,
} IBMSolverBCType;

/* <- src/ibm3d.h:47-50 */
typedef enum _val_type {
    VAL_CONST,
    VAL_FUNC,
} IBMSolverValType;

/* (t, x, y, z) => boundary value.  <- src/ibm3d.h:53 */
typedef double (*IBMSolverBCValFunc)(double, double, double, double);
/* (t, x, y, z) => force value.  <- src/ibm3d.h:69 */
typedef double (*IBMSolverForceFunc)(double, double, double, double);

/* <- src/ibm3d.h:83-87. TURBMODEL_VREMAN and TURBMODEL_SMAGORINSKY_VD are new
   (no reference counterpart); TURBMODEL_WALE aborts in set_turb_model, as in
   the reference (src/ibm3d_setup.c:352-353). */
typedef enum _turb_model_type {
    TURBMODEL_NONE,
    TURBMODEL_SMAGORINSKY,
    TURBMODEL_WALE,
    TURBMODEL_VREMAN,
    TURBMODEL_SMAGORINSKY_VD,
} IBMSolverTurbModelType;

/* <- src/ibm3d.h:94-98 */
enum {
    FLAG_FLUID = 1,
    FLAG_GHOST,
    FLAG_SOLID
};

typedef struct _ibm_solver IBMSolver;

/* ---- setup  <- src/ibm3d_setup.h:6-31 ---------------------------------- */
IBMSolver *IBMSolver_new(const int num_process, const int rank,
                         const int Px, const int Py, const int Pz);
void IBMSolver_destroy(IBMSolver *);

/* Face coordinates, lengths Nx+1, Ny+1, Nz+1; copied. */
void IBMSolver_set_grid(IBMSolver *, const int Nx, const int Ny, const int Nz,
                        const double *xf, const double *yf, const double *zf);
void IBMSolver_set_params(IBMSolver *, const double Re, const double dt);

/* VAL_CONST: 3 doubles; VAL_FUNC: 3 IBMSolverForceFunc. */
void IBMSolver_set_ext_force(IBMSolver *, IBMSolverValType, ...);
/* TURBMODEL_SMAGORINSKY: double Cs; TURBMODEL_VREMAN: double c;
   TURBMODEL_SMAGORINSKY_VD (Smagorinsky with van Driest damping, l =
   min(Cs Delta (1 - exp(-y+/A+)), 0.41 phi), y+ = y_w u_tau Re with y_w the
   distance to the nearest outer face in `walls`): double Cs, double A+,
   double u_tau, int walls (IBMSolverDirection bits). */
void IBMSolver_set_turb_model(IBMSolver *, IBMSolverTurbModelType, ...);
/* Varargs as in src/ibm3d_setup.c:378-415; values must be doubles.
   BC_CONVECTIVE_OUTFLOW: double Uc, the convection speed; Uc <= 0 takes the
   mean outward normal velocity over the face at each step. The face values
   start from the first interior cell's velocity (zero gradient) at the first
   IBMSolver_iterate after an init_flow_* or set_field call. */
void IBMSolver_set_bc(IBMSolver *, IBMSolverDirection, IBMSolverBCType, ...);
/* Borrowed; used only inside IBMSolver_assemble. NULL = no body. */
void IBMSolver_set_obstacle(IBMSolver *, Polyhedron *);
void IBMSolver_set_linear_solver(IBMSolver *, IBMSolverLinearSolverType,
                                 IBMSolverPrecondType, const double tol);
/* Borrowed filename (no extension); period <= 0 disables autosave. */
void IBMSolver_set_autosave(IBMSolver *, const char *, int);

void IBMSolver_assemble(IBMSolver *);

/* Mirror-point trilinear stencil of ghost cell (i,j,k): 8 cell indices in bit
   order (x=4, y=2, z=1) and their weights.  <- src/utils.c:111-171 */
void IBMSolver_ghost_interp(IBMSolver *, const int i, const int j, const int k,
                            int interp_idx[][3], double interp_coeff[]);

/* ---- initialisation  <- src/ibm3d_init.h:6-8 ---------------------------- */
void IBMSolver_init_flow_const(IBMSolver *, const double u1, const double u2,
                               const double u3, const double p);
/* Restart from "<filename>.nc" (float32 fields, lossy as in the reference). */
void IBMSolver_init_flow_file(IBMSolver *, const char *filename);
void IBMSolver_init_flow_func(IBMSolver *, IBMSolverInitFunc, IBMSolverInitFunc,
                              IBMSolverInitFunc, IBMSolverInitFunc);

/* ---- time stepping  <- src/ibm3d_fracstep.h:6 --------------------------- */
void IBMSolver_iterate(IBMSolver *, int num_time_steps, bool verbose);

/* ---- LES  <- src/ibm3d_turb.h:6 ------------------------------------------ */
void IBMSolver_calc_tau_r(IBMSolver *);

/* ---- export  <- src/ibm3d_export.h:6-9 ---------------------------------- */
void IBMSolver_export_result(IBMSolver *, const char *filename);
void IBMSolver_export_lvset_flag(IBMSolver *, const char *filename);
void IBMSolver_export_intermediate(IBMSolver *, const char *filename);

#ifdef __cplusplus
}
#endif

#include "ibm3d_ext.h"

#endif
