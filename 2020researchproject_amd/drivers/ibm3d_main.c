/*
 * ibm3d_main.c - the sphere driver of the reference (src/main.c:61-173) on the
 * MI355X library: reads the same ibm3d.in format, sets up the same case and
 * writes the same netCDF files, with one GPU in place of MPI + Hypre.
 *
 * ibm3d.in (src/main.c:70-107): STLFile <path>; Nx <n> then n+1 faces; the
 * same for Ny and Nz; Re <f>; dt <f> numtstep <n>; InitUsingFile <0|1>;
 * InitFile <path without .nc>; OutputFile <path without .nc>.
 *
 * Usage: ibm3d_main [input] [-steps N] [-precision fp64|mixed|mixed-krylov]
 *   input defaults to ./ibm3d.in; -steps overrides numtstep; the STL path is
 *   taken as written, or relative to the input file's folder if it is not
 *   found as written.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ibm3d.h"

static FILE *open_or_die(const char *path, const char *mode) {
    FILE *fp = fopen(path, mode);
    if (!fp) {
        fprintf(stderr, "error: cannot open %s\n", path);
        exit(1);
    }
    return fp;
}

static void read_or_die(int got, int want, const char *what) {
    if (got != want) {
        fprintf(stderr, "error: malformed input file near %s\n", what);
        exit(1);
    }
}

static double *read_faces(FILE *fp, int *n, const char *name) {
    read_or_die(fscanf(fp, "%*s %d", n), 1, name);
    if (*n < 2) {
        fprintf(stderr, "error: %s must be at least 2\n", name);
        exit(1);
    }
    double *f = calloc((size_t)*n + 1, sizeof(double));
    for (int i = 0; i <= *n; i++) read_or_die(fscanf(fp, "%lf", &f[i]), 1, name);
    return f;
}

int main(int argc, char **argv) {
    const char *input = "ibm3d.in";
    int steps_override = -1;
    int precision = IBM3D_PREC_FP64;
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "-steps") && a + 1 < argc) steps_override = atoi(argv[++a]);
        else if (!strcmp(argv[a], "-precision") && a + 1 < argc) {
            const char *p = argv[++a];
            if (!strcmp(p, "fp64")) precision = IBM3D_PREC_FP64;
            else if (!strcmp(p, "mixed")) precision = IBM3D_PREC_MIXED;
            else if (!strcmp(p, "mixed-krylov")) precision = IBM3D_PREC_MIXED_KRYLOV;
            else {
                fprintf(stderr, "error: unknown precision '%s' (fp64, mixed or mixed-krylov)\n", p);
                return 2;
            }
        } else input = argv[a];
    }

    /* ----- read the input file (src/main.c:61-110) ----------------------- */
    printf("\nRead input file\n");
    FILE *fp_in = open_or_die(input, "r");
    char stl_file[1024], init_file[1024], output_file[1024];
    read_or_die(fscanf(fp_in, "%*s %1023s", stl_file), 1, "STLFile");
    FILE *fp_poly = fopen(stl_file, "rb");
    if (!fp_poly) {
        /* relative to the input file's folder */
        char alt[2048];
        const char *slash = strrchr(input, '/');
        if (slash) {
            snprintf(alt, sizeof alt, "%.*s/%s", (int)(slash - input), input, stl_file);
            fp_poly = fopen(alt, "rb");
        }
    }
    if (!fp_poly) {
        fprintf(stderr, "error: cannot open %s\n", stl_file);
        return 1;
    }
    printf("Read polyhedron file: %s\n", stl_file);
    Polyhedron *poly = Polyhedron_new();
    Polyhedron_read_stl(poly, fp_poly);
    fclose(fp_poly);

    int Nx, Ny, Nz, num_time_steps, init_using_file;
    double Re, dt;
    double *xf = read_faces(fp_in, &Nx, "Nx");
    double *yf = read_faces(fp_in, &Ny, "Ny");
    double *zf = read_faces(fp_in, &Nz, "Nz");
    read_or_die(fscanf(fp_in, "%*s %lf", &Re), 1, "Re");
    read_or_die(fscanf(fp_in, "%*s %lf %*s %d", &dt, &num_time_steps), 2, "dt / numtstep");
    read_or_die(fscanf(fp_in, "%*s %d", &init_using_file), 1, "InitUsingFile");
    read_or_die(fscanf(fp_in, "%*s %1023s", init_file), 1, "InitFile");
    read_or_die(fscanf(fp_in, "%*s %1023s", output_file), 1, "OutputFile");
    fclose(fp_in);
    if (steps_override >= 0) num_time_steps = steps_override;

    printf("\nInput mesh size: %d x %d x %d\n", Nx, Ny, Nz);
    printf("  xmin: %10.4lf, xmax: %10.4lf\n", xf[0], xf[Nx]);
    printf("  ymin: %10.4lf, ymax: %10.4lf\n", yf[0], yf[Ny]);
    printf("  zmin: %10.4lf, zmax: %10.4lf\n", zf[0], zf[Nz]);
    Polyhedron_print_stats(poly);
    printf("\nReynolds no.: %.6lf\n", Re);
    printf("delta t     : %.6lf\n", dt);

    /* ----- solver (src/main.c:133-148), one GPU ------------------------ */
    IBMSolver *solver = IBMSolver_new(1, 0, 1, 1, 1);
    IBMSolver_set_grid(solver, Nx, Ny, Nz, xf, yf, zf);
    IBMSolver_set_params(solver, Re, dt);
    IBMSolver_set_bc(solver, DIR_WEST, BC_VELOCITY_COMPONENT, VAL_CONST, 1., 0., 0.);
    IBMSolver_set_bc(solver, DIR_EAST, BC_PRESSURE, VAL_CONST, 0.);
    IBMSolver_set_bc(solver, DIR_NORTH | DIR_SOUTH | DIR_DOWN | DIR_UP, BC_FREE_SLIP_WALL);
    IBMSolver_set_obstacle(solver, poly);
    IBMSolver_set_linear_solver(solver, SOLVER_BiCGSTAB, PRECOND_AMG, 1e-6);
    IBMSolver_set_autosave(solver, output_file, 0);
    IBMSolver_set_precision(solver, precision);
    IBMSolver_assemble(solver);
    Polyhedron_destroy(poly);

    /* ----- initial field (src/main.c:153-159) -------------------------- */
    if (!init_using_file) IBMSolver_init_flow_const(solver, 1, 0, 0, 0);
    else IBMSolver_init_flow_file(solver, init_file);

    char lf[1100];
    snprintf(lf, sizeof lf, "%s_lvset_flag", output_file);
    IBMSolver_export_lvset_flag(solver, lf);
    printf("\nInitialization done\n");

    /* ----- run and export (src/main.c:169-173) ------------------------- */
    IBMSolver_iterate(solver, num_time_steps, true);
    IBMSolver_export_result(solver, output_file);

    double force[3];
    IBMSolver_calc_body_force(solver, force);
    printf("\nbody force: %.6e %.6e %.6e  max|div U|: %.3e\n", force[0], force[1], force[2],
           IBMSolver_max_div(solver));

    free(xf);
    free(yf);
    free(zf);
    IBMSolver_destroy(solver);
    return 0;
}
