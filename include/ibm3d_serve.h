/*
 * ibm3d_serve.h - request loop over independent solver cases on one GPU.
 *
 * No counterpart in the reference (its drivers run one case per MPI job).
 * Cases arrive over time; each is admitted when a slot is free and its
 * device-memory footprint fits the budget next to the cases already running
 * (first fit in arrival order), then runs to completion on its own HIP
 * stream, so the kernels of concurrent cases overlap on the device. A case's
 * fields and multigrid hierarchy stay resident in HBM from admission to
 * completion; the multigrid cycles of each case replay as HIP graphs.
 */
#ifndef IBM3D_SERVE_H
#define IBM3D_SERVE_H

#include <stddef.h>

#include "ibm3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct _ibm_server IBMServer;

/* Initial flow of a job submitted with IBMServer_submit_solver: the constant
   flow (u1, u2, u3, p) of IBMSolver_init_flow_const and, when amp > 0, on each
   interior cell and velocity component c = 1, 2, 3 the seeded perturbation
   amp * (2 r - 1) added to the constant, with r in [0, 1) a hash of
   (seed, c, cell centre): splitmix64 finalisers chained over the IEEE bits
   of z, y, x, then c, then seed; r = (h >> 11) * 2^-53. */
typedef struct {
    double u1, u2, u3, p;
    double amp;
    unsigned long long seed;
} IBMServeInit;

/* Outcome of a job (IBMServer_job_result). state: 0 queued, 1 running,
   2 done, 3 failed; t: seconds since IBMServer_new of submission, admission
   and completion (negative while not reached); iters: iterations of the last
   step's u1, u2, u3 and p' solves; max_div: IBMSolver_max_div after the last
   step; device_bytes: IBMSolver_get_device_bytes; error: the failed job's
   message (empty otherwise). */
typedef struct {
    int state;
    double t[3];
    int iters[4];
    double max_div;
    double device_bytes;
    char error[256];
} IBMServeResult;

/* Builds and assembles the case's solver (called on a worker thread after
   admission); may return NULL for a job without a solver. */
typedef IBMSolver *(*IBMServeSetupFunc)(void *user);
/* Called after the last step, before the solver is destroyed. */
typedef void (*IBMServeDoneFunc)(IBMSolver *solver, void *user);

/* budget_bytes: device memory the admitted cases may hold together;
   max_active: cases running at once (worker threads). */
IBMServer *IBMServer_new(double budget_bytes, int max_active);
/* Queues a case needing `bytes` of device memory that runs `steps` time steps
   (IBMSolver_iterate). The callbacks run with case errors aborting the
   process, as outside the server; an error inside the time steps fails the
   job only. Returns the job id, or -1 if bytes exceeds the budget. */
int IBMServer_submit(IBMServer *, double bytes, int steps, IBMServeSetupFunc setup, IBMServeDoneFunc done,
                     void *user);
/* Queues a case whose solver the caller has configured but not assembled
   (IBMSolver_new and the IBMSolver_set_* calls; an obstacle polyhedron stays
   the caller's until the job has finished). The server owns the solver from
   here on: a worker assembles it, applies `init` (NULL: fluid at rest), runs
   `steps` time steps, copies U1, U2, U3 and P (padded reference layout, see
   IBMSolver_get_field) into fields[0..3] where those are not NULL (buffers
   the caller keeps valid until the job has finished), and destroys it. All of
   this is native code on the worker thread. An error of the case (input that
   IBMSolver_assemble rejects, a non-finite residual, a failed device
   allocation) fails this job only: its state becomes 3 with the message, its
   budget and solver are released, and the server goes on. HIP errors other
   than a failed allocation still abort the process. Returns the job id, or
   -1 if bytes exceeds the budget or solver is NULL. */
int IBMServer_submit_solver(IBMServer *, IBMSolver *solver, double bytes, int steps, const IBMServeInit *init,
                            double *fields[4]);
/* Blocks until every submitted case has finished (done or failed). */
void IBMServer_wait(IBMServer *);
/* Seconds since IBMServer_new of the job's submission, admission and
   completion (negative while not reached); returns 0 queued, 1 running,
   2 done, 3 failed, -1 unknown job. */
int IBMServer_job_times(IBMServer *, int job, double out[3]);
/* Fills *out for the job and returns its state (-1: unknown job). */
int IBMServer_job_result(IBMServer *, int job, IBMServeResult *out);
/* Jobs that have failed so far. */
int IBMServer_failed_jobs(IBMServer *);
/* Largest sum of admitted footprints seen so far. */
double IBMServer_peak_bytes(IBMServer *);
/* Waits for the queued cases, then joins the workers. */
void IBMServer_destroy(IBMServer *);

/* Upper estimate of the device memory a case allocates (fields, Krylov and
   refinement vectors, the concurrent momentum lanes, multigrid hierarchy,
   coarsest dense inverse) for an Nx x Ny x Nz grid with the BiCGSTAB, PCG or
   AMG pressure solver; les: 1 if a turbulence model is on. Not counted:
   SOLVER_GMRES's basis (21 more FP64 vectors per cell) and function-valued
   forcing (3 FP64 arrays per cell). */
double IBMSolver_estimate_device_bytes(int Nx, int Ny, int Nz, int precision, int les);
/* Device memory the solver allocated during IBMSolver_new / assemble. */
double IBMSolver_get_device_bytes(IBMSolver *);

#ifdef __cplusplus
}
#endif

#endif
