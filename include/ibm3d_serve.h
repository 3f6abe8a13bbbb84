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

/* Builds and assembles the case's solver (called on a worker thread after
   admission); may return NULL for a job without a solver. */
typedef IBMSolver *(*IBMServeSetupFunc)(void *user);
/* Called after the last step, before the solver is destroyed. */
typedef void (*IBMServeDoneFunc)(IBMSolver *solver, void *user);

/* budget_bytes: device memory the admitted cases may hold together;
   max_active: cases running at once (worker threads). */
IBMServer *IBMServer_new(double budget_bytes, int max_active);
/* Queues a case needing `bytes` of device memory that runs `steps` time steps
   (IBMSolver_iterate). Returns the job id, or -1 if bytes exceeds the budget. */
int IBMServer_submit(IBMServer *, double bytes, int steps, IBMServeSetupFunc setup, IBMServeDoneFunc done,
                     void *user);
/* Blocks until every submitted case has finished. */
void IBMServer_wait(IBMServer *);
/* Seconds since IBMServer_new of the job's submission, admission and
   completion (negative while not reached); returns 0 queued, 1 running,
   2 done, -1 unknown job. */
int IBMServer_job_times(IBMServer *, int job, double out[3]);
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
