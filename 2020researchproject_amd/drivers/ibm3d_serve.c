/*
 * ibm3d_serve - the request loop of include/ibm3d_serve.h driven from C, with
 * no Python anywhere: each case is configured with the IBMSolver_set_* calls
 * (host work only) and handed to IBMServer_submit_solver, which assembles,
 * initialises and runs it on a worker thread.
 *
 *   ibm3d_serve BUDGET_GIB SLOTS CASEFILE
 *
 * CASEFILE holds one case per line:
 *   cavity N STEPS RE        lid-driven cavity on the unit cube, N^3 cells,
 *                            rest plus the seeded perturbation (amplitude 1e-3,
 *                            seed 2020) of the cavity workload
 *   sphere NX NY NZ STEPS RE analytic sphere of radius 0.5 at the origin in
 *                            [-5,15]x[-5,5]^2 on a uniform grid, uniform inflow,
 *                            pressure outlet, free-slip lateral walls
 *   badperiodic N STEPS      a cavity whose periodic x pair is inconsistent:
 *                            IBMSolver_assemble rejects it, the job fails alone
 * Every case runs in IBM3D_PREC_MIXED_KRYLOV with BiCGSTAB + multigrid, tol
 * 1e-10, dt 0.005 (cavity) or 0.002 (sphere). One line per job is printed
 * when all have finished: id, kind, state, times, iterations, max|div U|,
 * device bytes and the error of a failed job. A malformed or unknown case
 * line ends the reading of CASEFILE with an error on stderr: the jobs read
 * before it still run and are reported, and the exit status is 2. No
 * counterpart in the reference, whose drivers run one case per process.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ibm3d.h"
#include "ibm3d_serve.h"

#define MAX_JOBS 256

static double *uniform(double a, double b, int n) {
    double *f = malloc((size_t)(n + 1) * sizeof *f);
    for (int i = 0; i <= n; i++) f[i] = a + (b - a) * i / n;
    return f;
}

static IBMSolver *configure(int nx, int ny, int nz, const double *box, double Re, double dt) {
    IBMSolver *s = IBMSolver_new(1, 0, 1, 1, 1);
    double *xf = uniform(box[0], box[1], nx), *yf = uniform(box[2], box[3], ny), *zf = uniform(box[4], box[5], nz);
    IBMSolver_set_grid(s, nx, ny, nz, xf, yf, zf);
    free(xf);
    free(yf);
    free(zf);
    IBMSolver_set_params(s, Re, dt);
    IBMSolver_set_linear_solver(s, SOLVER_BiCGSTAB, PRECOND_AMG, 1e-10);
    IBMSolver_set_precision(s, IBM3D_PREC_MIXED_KRYLOV);
    return s;
}

int main(int argc, char **argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s BUDGET_GIB SLOTS CASEFILE\n", argv[0]);
        return 2;
    }
    const double budget = atof(argv[1]) * 1073741824.0;
    const int slots = atoi(argv[2]);
    FILE *in = fopen(argv[3], "r");
    if (!in) {
        fprintf(stderr, "error: cannot open %s\n", argv[3]);
        return 1;
    }
    if (!(budget > 0) || slots < 1 || slots > 64) {
        fprintf(stderr, "error: budget must be positive and slots 1..64\n");
        fclose(in);
        return 2;
    }
    IBMServer *srv = IBMServer_new(budget, slots);
    int ids[MAX_JOBS];
    char kinds[MAX_JOBS][16];
    int njobs = 0;
    char kind[16];
    int bad_input = 0;
    while (njobs < MAX_JOBS && fscanf(in, "%15s", kind) == 1) {
        IBMSolver *s = NULL;
        IBMServeInit init = {0, 0, 0, 0, 0, 0};
        int n[3], steps = 0;
        double Re = 100;
        if (!strcmp(kind, "cavity") || !strcmp(kind, "badperiodic")) {
            const int bad = !strcmp(kind, "badperiodic");
            if (fscanf(in, "%d %d", &n[0], &steps) != 2 || (!bad && fscanf(in, "%lf", &Re) != 1)) {
                fprintf(stderr, "error: malformed %s case (expected %s)\n", kind, bad ? "N STEPS" : "N STEPS RE");
                bad_input = 1;
                break;
            }
            n[1] = n[2] = n[0];
            const double box[6] = {0, 1, 0, 1, 0, 1};
            s = configure(n[0], n[1], n[2], box, Re, 0.005);
            IBMSolver_set_bc(s, DIR_NORTH, BC_VELOCITY_COMPONENT, VAL_CONST, 1.0, 0.0, 0.0);
            IBMSolver_set_bc(s, DIR_EAST | DIR_WEST | DIR_SOUTH | DIR_DOWN | DIR_UP, BC_STATIONARY_WALL);
            if (bad) IBMSolver_set_bc(s, DIR_WEST, BC_ALL_PERIODIC);
            IBMSolver_set_obstacle(s, NULL);
            init.amp = 1e-3;
            init.seed = 2020;
        } else if (!strcmp(kind, "sphere")) {
            if (fscanf(in, "%d %d %d %d %lf", &n[0], &n[1], &n[2], &steps, &Re) != 5) {
                fprintf(stderr, "error: malformed sphere case (expected NX NY NZ STEPS RE)\n");
                bad_input = 1;
                break;
            }
            const double box[6] = {-5, 15, -5, 5, -5, 5};
            s = configure(n[0], n[1], n[2], box, Re, 0.002);
            IBMSolver_set_bc(s, DIR_WEST, BC_VELOCITY_COMPONENT, VAL_CONST, 1.0, 0.0, 0.0);
            IBMSolver_set_bc(s, DIR_EAST, BC_PRESSURE, VAL_CONST, 0.0);
            IBMSolver_set_bc(s, DIR_NORTH | DIR_SOUTH | DIR_DOWN | DIR_UP, BC_FREE_SLIP_WALL);
            IBMSolver_set_obstacle_sphere(s, 0, 0, 0, 0.5);
            init.u1 = 1.0;
        } else {
            fprintf(stderr, "error: unknown case kind '%s'\n", kind);
            bad_input = 1;
            break;
        }
        const double bytes = IBMSolver_estimate_device_bytes(n[0], n[1], n[2], IBM3D_PREC_MIXED_KRYLOV, 0);
        const int id = IBMServer_submit_solver(srv, s, bytes, steps, &init, NULL);
        if (id < 0) {
            fprintf(stderr, "error: a %s case of %.3g bytes does not fit the budget\n", kind, bytes);
            IBMSolver_destroy(s);
            continue;
        }
        ids[njobs] = id;
        snprintf(kinds[njobs], sizeof kinds[njobs], "%s", kind);
        njobs++;
    }
    fclose(in);
    IBMServer_wait(srv);
    int failed = 0;
    for (int j = 0; j < njobs; j++) {
        IBMServeResult r;
        IBMServer_job_result(srv, ids[j], &r);
        failed += r.state == 3;
        printf("job %d %s state %s wait %.3f s run %.3f s iters %d %d %d %d max_div %.3e device_bytes %.0f error \"%s\"\n",
               ids[j], kinds[j], r.state == 2 ? "done" : r.state == 3 ? "failed" : "unfinished", r.t[1] - r.t[0],
               r.t[2] - r.t[1], r.iters[0], r.iters[1], r.iters[2], r.iters[3], r.max_div, r.device_bytes,
               r.state == 3 ? r.error : "");
    }
    printf("jobs %d failed %d peak_admitted_bytes %.0f\n", njobs, failed, IBMServer_peak_bytes(srv));
    IBMServer_destroy(srv);
    return bad_input ? 2 : 0;
}
