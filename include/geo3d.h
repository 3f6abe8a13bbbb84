/*
 * geo3d.h - geometry API of the MI355X IBM solver (drop-in for the reference's
 * src/geo3d.h:1-46).
 *
 * Vectors are plain value types; Polyhedron is opaque. The functions keep the
 * reference's names, argument meaning and ownership:
 *   Vector_*                      <- src/geo3d.h:21-32 (src/geo3d.c:120-206)
 *   Polyhedron_new/destroy        <- src/geo3d.h:34-35 (src/geo3d.c:209-226)
 *   Polyhedron_read_stl(FILE *)   <- src/geo3d.h:36    (src/geo3d.c:229-401)
 *   Polyhedron_print_stats        <- src/geo3d.h:37    (src/geo3d.c:403-426)
 *   Polyhedron_scale              <- src/geo3d.h:39    (src/geo3d.c:428-434)
 *   Polyhedron_cpt                <- src/geo3d.h:41-49 (src/geo3d.c:436-704)
 *
 * Polyhedron_cpt writes f[i][j][k] (k fastest), exactly like the reference's
 * variably-modified array parameter; here it is spelled as a flat pointer so
 * that the header is valid C and C++.
 */
#ifndef IBM3D_GEO3D_H
#define IBM3D_GEO3D_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCALE_INCH_TO_MM 25.4
#define SCALE_MM_TO_INCH (1 / 25.4)

typedef struct _vector {
    double x, y, z;
} Vector;

typedef struct _polyhedron Polyhedron;

Vector Vector_add(const Vector, const Vector);
Vector Vector_sub(const Vector, const Vector);
double Vector_dot(const Vector, const Vector);
Vector Vector_crs(const Vector, const Vector);

double Vector_norm(const Vector);
double Vector_dist(const Vector, const Vector);
double Vector_angle(const Vector, const Vector);

Vector Vector_normalize(const Vector);
Vector Vector_lincom(const double, const Vector, const double, const Vector);
Vector Vector_rot(const Vector, const Vector, const double);

Polyhedron *Polyhedron_new(void);
void Polyhedron_destroy(Polyhedron *);
void Polyhedron_read_stl(Polyhedron *, FILE *);
void Polyhedron_print_stats(Polyhedron *);

void Polyhedron_scale(Polyhedron *, double);

/* f has nx*ny*nz doubles, index (i*ny + j)*nz + k. Signed distance (positive
   outside the body) within maxd; points no extrusion reaches get +-maxd. */
void Polyhedron_cpt(
    const Polyhedron *poly,
    const int nx, const int ny, const int nz,
    const double *x, const double *y, const double *z,
    double *f,
    const double maxd
);

/* Device scratch of Polyhedron_cpt_noalloc, per GPU: 64 MiB. */
#define IBM3D_CPT_SCRATCH_BYTES (64u << 20)

/* Polyhedron_cpt without any device-memory allocation (no counterpart in the
   reference). For callers that cannot allocate device memory, e.g. a
   long-running serving loop that keeps the allocator out of its path. Same
   arguments and the same values of f, bit for bit, as Polyhedron_cpt. Every
   device buffer of the transform (coordinates, extrusion features, the
   distances) lives in one fixed device array of IBM3D_CPT_SCRATCH_BYTES per
   GPU, held until the process ends; it keeps the last call's contents.
   Returns 0, or -1 with a message on stderr (and f untouched) when the grid
   and the polyhedron's features need more than that, e.g. more than about
   8 million points. Calls from several threads of one process are
   serialised; calls that overlap on one GPU from other processes use their
   own copy. */
int Polyhedron_cpt_noalloc(
    const Polyhedron *poly,
    const int nx, const int ny, const int nz,
    const double *x, const double *y, const double *z,
    double *f,
    const double maxd
);

#ifdef __cplusplus
}
#endif

#endif
