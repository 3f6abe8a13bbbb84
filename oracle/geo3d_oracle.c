/*
 * geo3d_oracle.c - FP64 CPU oracle of the geometry setup.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Clean-room restatement of the
 * reference's src/geo3d.c: vector algebra (:120-206), binary STL reading with
 * vertex/edge de-duplication in first-appearance order (:229-401; the
 * reference uses GLib trees, this uses hash tables that assign the same
 * indices), and the closest-point transform (:436-938).
 */
#include "oracle.h"

#include <stdint.h>

static const double PI = 3.141592653589793;

/* <- src/geo3d.c:120-206 */
Vector Vector_add(const Vector a, const Vector b) { return (Vector){a.x + b.x, a.y + b.y, a.z + b.z}; }
Vector Vector_sub(const Vector a, const Vector b) { return (Vector){a.x - b.x, a.y - b.y, a.z - b.z}; }
double Vector_dot(const Vector a, const Vector b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vector Vector_crs(const Vector a, const Vector b) {
    return (Vector){a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
double Vector_norm(const Vector v) { return sqrt(Vector_dot(v, v)); }
double Vector_dist(const Vector a, const Vector b) { return Vector_norm(Vector_sub(a, b)); }
double Vector_angle(const Vector a, const Vector b) {
    const double la = Vector_norm(a), lb = Vector_norm(b);
    double c = Vector_dot(a, b) / (la * lb);
    if (c < -1) c = -1;
    else if (c > 1) c = 1;
    return acos(c);
}
Vector Vector_normalize(const Vector v) {
    const double len = Vector_norm(v);
    return (Vector){v.x / len, v.y / len, v.z / len};
}
Vector Vector_lincom(const double a, const Vector u, const double b, const Vector v) {
    return (Vector){a * u.x + b * v.x, a * u.y + b * v.y, a * u.z + b * v.z};
}
Vector Vector_rot(const Vector v, const Vector axis, const double angle) {
    const Vector n = Vector_normalize(axis);
    Vector r;
    r.x = (cos(angle) + n.x * n.x * (1 - cos(angle))) * v.x
        + (n.x * n.y * (1 - cos(angle)) - n.z * sin(angle)) * v.y
        + (n.x * n.z * (1 - cos(angle)) + n.y * sin(angle)) * v.z;
    r.y = (n.x * n.y * (1 - cos(angle)) + n.z * sin(angle)) * v.x
        + (cos(angle) + n.y * n.y * (1 - cos(angle))) * v.y
        + (n.y * n.z * (1 - cos(angle)) - n.x * sin(angle)) * v.z;
    r.z = (n.x * n.z * (1 - cos(angle)) - n.y * sin(angle)) * v.x
        + (n.y * n.z * (1 - cos(angle)) + n.x * sin(angle)) * v.y
        + (cos(angle) + n.z * n.z * (1 - cos(angle))) * v.z;
    return r;
}

typedef struct { Vector coord; int num_faces; int *face_idx; } OVertex;
typedef struct { int vertex_idx[2]; int face_idx[2]; } OEdge;
typedef struct { int vertex_idx[3]; int edge_idx[3]; Vector n; } OFace;

struct _polyhedron {
    size_t num_vertices, num_edges, num_faces;
    OVertex *vertex_list;
    OEdge *edge_list;
    OFace *face_list;
};

Polyhedron *Polyhedron_new(void) { return calloc(1, sizeof(Polyhedron)); }

void Polyhedron_destroy(Polyhedron *poly) {
    for (size_t i = 0; i < poly->num_vertices; i++) free(poly->vertex_list[i].face_idx);
    free(poly->vertex_list); free(poly->edge_list); free(poly->face_list);
    free(poly);
}

static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static uint64_t dbits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

/* <- src/geo3d.c:229-401 */
void Polyhedron_read_stl(Polyhedron *poly, FILE *f) {
    char header[81] = "";
    uint32_t nf = 0;
    if (fread(header, 1, 80, f) != 80 || fread(&nf, 4, 1, f) != 1) { fprintf(stderr, "error: bad STL\n"); abort(); }
    poly->num_faces = nf;
    poly->face_list = calloc(nf ? nf : 1, sizeof(OFace));

    /* vertex table: first-appearance index of each distinct coordinate */
    size_t cap = 16;
    while (cap < 6 * (size_t)nf + 16) cap <<= 1;
    int *vslot = malloc(cap * sizeof(int));
    for (size_t q = 0; q < cap; q++) vslot[q] = -1;
    Vector *vc = malloc((3 * (size_t)nf + 1) * sizeof(Vector));
    int *vnf = calloc(3 * (size_t)nf + 1, sizeof(int));
    int *vfirst = calloc(3 * (size_t)nf + 1, sizeof(int));
    int nv = 0;
    for (uint32_t i = 0; i < nf; i++) {
        float c[3];
        Vector vt[3];
        uint16_t attr;
        if (fread(c, 4, 3, f) != 3) abort();
        poly->face_list[i].n = (Vector){c[0], c[1], c[2]};
        for (int j = 0; j < 3; j++) {
            if (fread(c, 4, 3, f) != 3) abort();
            vt[j] = (Vector){c[0], c[1], c[2]};
        }
        if (fread(&attr, 2, 1, f) != 1) abort();
        for (int j = 0; j < 3; j++) {
            const uint64_t h = mix64(dbits(vt[j].x) ^ mix64(dbits(vt[j].y) ^ mix64(dbits(vt[j].z))));
            size_t q = h & (cap - 1);
            while (vslot[q] >= 0) {
                const Vector o = vc[vslot[q]];
                if (o.x == vt[j].x && o.y == vt[j].y && o.z == vt[j].z) break;
                q = (q + 1) & (cap - 1);
            }
            if (vslot[q] < 0) {
                vslot[q] = nv;
                vc[nv] = vt[j];
                vfirst[nv] = (int)i;
                nv++;
            }
            const int v = vslot[q];
            poly->face_list[i].vertex_idx[j] = v;
            vnf[v]++;
        }
    }
    poly->num_vertices = nv;
    poly->vertex_list = calloc(nv ? nv : 1, sizeof(OVertex));
    for (int v = 0; v < nv; v++) {
        poly->vertex_list[v].coord = vc[v];
        poly->vertex_list[v].num_faces = vnf[v];
        poly->vertex_list[v].face_idx = malloc(vnf[v] * sizeof(int));
        poly->vertex_list[v].face_idx[0] = vfirst[v];
        for (int q = 1; q < vnf[v]; q++) poly->vertex_list[v].face_idx[q] = -1;
    }
    free(vslot); free(vc); free(vnf); free(vfirst);

    /* edge table keyed on the unordered vertex pair */
    int *eslot = malloc(cap * sizeof(int));
    for (size_t q = 0; q < cap; q++) eslot[q] = -1;
    OEdge *el = calloc(3 * (size_t)nf + 1, sizeof(OEdge));
    int ne = 0;
    for (uint32_t i = 0; i < nf; i++)
        for (int j = 0; j < 3; j++) {
            const int a = poly->face_list[i].vertex_idx[j];
            const int b = poly->face_list[i].vertex_idx[(j + 1) % 3];
            const int lo = a < b ? a : b, hi = a < b ? b : a;
            size_t q = mix64(((uint64_t)lo << 32) | (uint32_t)hi) & (cap - 1);
            while (eslot[q] >= 0) {
                const OEdge *e = &el[eslot[q]];
                const int elo = e->vertex_idx[0] < e->vertex_idx[1] ? e->vertex_idx[0] : e->vertex_idx[1];
                const int ehi = e->vertex_idx[0] < e->vertex_idx[1] ? e->vertex_idx[1] : e->vertex_idx[0];
                if (elo == lo && ehi == hi) break;
                q = (q + 1) & (cap - 1);
            }
            if (eslot[q] < 0) {
                eslot[q] = ne;
                el[ne].vertex_idx[0] = a;
                el[ne].vertex_idx[1] = b;
                el[ne].face_idx[0] = (int)i;
                el[ne].face_idx[1] = -1;
                ne++;
            } else {
                el[eslot[q]].face_idx[1] = (int)i;
            }
            poly->face_list[i].edge_idx[j] = eslot[q];
        }
    free(eslot);
    poly->num_edges = ne;
    poly->edge_list = el;

    /* faces around each vertex, walking across edges (:344-368) */
    for (int v = 0; v < nv; v++) {
        int cur = poly->vertex_list[v].face_idx[0];
        for (int q = 1; q < poly->vertex_list[v].num_faces; q++) {
            const OFace *fc = &poly->face_list[cur];
            int e;
            if (fc->vertex_idx[0] == v) e = fc->edge_idx[2];
            else if (fc->vertex_idx[1] == v) e = fc->edge_idx[0];
            else e = fc->edge_idx[1];
            cur = poly->edge_list[e].face_idx[0] == cur ? poly->edge_list[e].face_idx[1] : poly->edge_list[e].face_idx[0];
            poly->vertex_list[v].face_idx[q] = cur;
        }
    }
}

void Polyhedron_print_stats(Polyhedron *poly) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    printf("Input polyhedron size: %zu faces, %zu edges, %zu vertices\n",
           poly->num_faces, poly->num_edges, poly->num_vertices);
    for (size_t i = 0; i < poly->num_vertices; i++) {
        const double c[3] = {poly->vertex_list[i].coord.x, poly->vertex_list[i].coord.y, poly->vertex_list[i].coord.z};
        for (int a = 0; a < 3; a++) { if (c[a] < mn[a]) mn[a] = c[a]; if (c[a] > mx[a]) mx[a] = c[a]; }
    }
    printf("  xmin: %10.4lf, xmax: %10.4lf\n", mn[0], mx[0]);
    printf("  ymin: %10.4lf, ymax: %10.4lf\n", mn[1], mx[1]);
    printf("  zmin: %10.4lf, zmax: %10.4lf\n", mn[2], mx[2]);
}

void Polyhedron_scale(Polyhedron *poly, double scale) {
    for (size_t i = 0; i < poly->num_vertices; i++) {
        poly->vertex_list[i].coord.x *= scale;
        poly->vertex_list[i].coord.y *= scale;
        poly->vertex_list[i].coord.z *= scale;
    }
}

/* ---- closest-point transform  <- src/geo3d.c:436-938 -------------------- */

typedef struct { double a, b, c, d; } Plane;

static Plane plane3(Vector a, Vector b, Vector c) {  /* :706-716 */
    const Vector n = Vector_crs(Vector_sub(b, a), Vector_sub(c, a));
    return (Plane){n.x, n.y, n.z, -Vector_dot(n, a)};
}
static double plane_dist(Plane p, Vector v) {  /* :718-720 */
    return fabs(p.a * v.x + p.b * v.y + p.c * v.z + p.d) / Vector_norm((Vector){p.a, p.b, p.c});
}
static double segment_dist(Vector a, Vector b, Vector v) {  /* :722-728 */
    const Vector ab = Vector_sub(b, a), av = Vector_sub(v, a);
    const double t = Vector_dot(ab, av) / Vector_dot(ab, ab);
    return Vector_dist(v, Vector_lincom(1 - t, a, t, b));
}

static int lower_bound(int len, const double *arr, double val) {  /* src/utils.c:44-59 */
    int l = 0, h = len;
    while (l < h) {
        int mid = l + (h - l) / 2;
        if (val <= arr[mid]) h = mid;
        else l = mid + 1;
    }
    return l;
}
static int upper_bound_g(int len, const double *arr, double val) {  /* src/utils.c:66-81 */
    int l = 0, h = len;
    while (l < h) {
        int mid = l + (h - l) / 2;
        if (val >= arr[mid]) l = mid + 1;
        else h = mid;
    }
    return l;
}

typedef struct {
    int nx, ny, nz;
    const double *x, *y, *z;
    double *f;
} Grid;

static void update_point(double *fv, double dist, int sgn) {
    if (isnan(*fv) || dist < fabs(*fv)) *fv = sgn * dist;
}

/* prism extrusion of a face or an edge wedge: 6 points, 5 bounding planes;
   face = 1 uses plane distance, else segment distance (:820-938) */
static void prism_cpt(const Vector pts[6], int sgn, int face, const Grid *g) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int q = 0; q < 6; q++) {
        const double c[3] = {pts[q].x, pts[q].y, pts[q].z};
        for (int a = 0; a < 3; a++) {
            if (c[a] > mx[a]) mx[a] = c[a];
            if (c[a] < mn[a]) mn[a] = c[a];
        }
    }
    const int i1 = upper_bound_g(g->nx, g->x, mx[0]), i0 = lower_bound(g->nx, g->x, mn[0]);
    const int j1 = upper_bound_g(g->ny, g->y, mx[1]), j0 = lower_bound(g->ny, g->y, mn[1]);
    const int k1 = upper_bound_g(g->nz, g->z, mx[2]), k0 = lower_bound(g->nz, g->z, mn[2]);
    Plane pl[5];
    if (face) {
        pl[0] = plane3(pts[0], pts[2], pts[1]);
        pl[1] = plane3(pts[3], pts[4], pts[5]);
        pl[2] = plane3(pts[0], pts[1], pts[4]);
        pl[3] = plane3(pts[1], pts[2], pts[5]);
        pl[4] = plane3(pts[0], pts[3], pts[5]);
    } else {
        pl[0] = plane3(pts[0], pts[4], pts[2]);
        pl[1] = plane3(pts[1], pts[3], pts[5]);
        pl[2] = plane3(pts[0], pts[2], pts[3]);
        pl[3] = plane3(pts[2], pts[4], pts[5]);
        pl[4] = plane3(pts[0], pts[1], pts[5]);
    }
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = i0; i < i1; i++)
        for (int j = j0; j < j1; j++)
            for (int k = k0; k < k1; k++) {
                const Vector cur = {g->x[i], g->y[j], g->z[k]};
                bool in = true;
                for (int l = 0; l < 5; l++)
                    if (Vector_dot((Vector){pl[l].a, pl[l].b, pl[l].c}, cur) + pl[l].d > 1e-12) in = false;
                if (!in) continue;
                const double dist = face ? plane_dist(pl[0], cur) : segment_dist(pts[0], pts[1], cur);
                update_point(&g->f[((size_t)i * g->ny + j) * g->nz + k], dist, sgn);
            }
}

/* :784-818 */
static void cone_cpt(Vector vertex, Vector axis, double radius, double angle, int sgn, const Grid *g) {
    const int i1 = upper_bound_g(g->nx, g->x, vertex.x + radius), i0 = lower_bound(g->nx, g->x, vertex.x - radius);
    const int j1 = upper_bound_g(g->ny, g->y, vertex.y + radius), j0 = lower_bound(g->ny, g->y, vertex.y - radius);
    const int k1 = upper_bound_g(g->nz, g->z, vertex.z + radius), k0 = lower_bound(g->nz, g->z, vertex.z - radius);
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = i0; i < i1; i++)
        for (int j = j0; j < j1; j++)
            for (int k = k0; k < k1; k++) {
                const Vector d = Vector_sub((Vector){g->x[i], g->y[j], g->z[k]}, vertex);
                const double dist = Vector_norm(d);
                if (Vector_angle(axis, d) <= angle + 1e-12 && dist <= radius + 1e-12)
                    update_point(&g->f[((size_t)i * g->ny + j) * g->nz + k], dist, sgn);
            }
}

static Vector offset(Vector p, double t, Vector n) { return Vector_lincom(1, p, t, n); }

void Polyhedron_cpt(const Polyhedron *poly, const int nx, const int ny, const int nz,
                    const double *x, const double *y, const double *z, double *f, const double maxd) {
    const Grid g = {nx, ny, nz, x, y, z, f};
    const size_t n = (size_t)nx * ny * nz;
    for (size_t q = 0; q < n; q++) f[q] = NAN;

    /* face prisms, outward then inward (:458-475) */
    for (size_t i = 0; i < poly->num_faces; i++) {
        const OFace *fc = &poly->face_list[i];
        Vector pts[6];
        for (int j = 0; j < 3; j++) {
            pts[j] = poly->vertex_list[fc->vertex_idx[j]].coord;
            pts[j + 3] = offset(pts[j], maxd, fc->n);
        }
        prism_cpt(pts, 1, 1, &g);
        for (int j = 0; j < 3; j++) {
            pts[j] = poly->vertex_list[fc->vertex_idx[(3 - j) % 3]].coord;
            pts[j + 3] = offset(pts[j], -maxd, fc->n);
        }
        prism_cpt(pts, -1, 1, &g);
    }

    /* edge wedges (:477-570) */
    for (size_t i = 0; i < poly->num_edges; i++) {
        const OEdge *e = &poly->edge_list[i];
        Vector pts[6];
        pts[0] = poly->vertex_list[e->vertex_idx[0]].coord;
        pts[1] = poly->vertex_list[e->vertex_idx[1]].coord;
        const Vector v = Vector_sub(pts[1], pts[0]);
        const Vector nl = poly->face_list[e->face_idx[0]].n;
        const Vector nr = poly->face_list[e->face_idx[1]].n;
        const int sgn = Vector_dot(Vector_crs(v, nl), nr) >= 0 ? 1 : -1;
        const double ang = Vector_angle(nl, nr);
        Vector dirs[4];
        int nd;
        if (ang <= PI / 3) {
            nd = 2;
            if (sgn == 1) { dirs[0] = nl; dirs[1] = nr; }
            else { dirs[0] = nr; dirs[1] = nl; }
        } else if (ang <= 2 * PI / 3) {
            nd = 3;
            const Vector nm = Vector_rot(nl, v, ang / 2);
            if (sgn == 1) { dirs[0] = nl; dirs[1] = nm; dirs[2] = nr; }
            else { dirs[0] = nr; dirs[1] = nm; dirs[2] = nl; }
        } else {
            nd = 4;
            if (sgn == 1) {
                const Vector n1 = Vector_rot(nl, v, ang / 3), n2 = Vector_rot(nl, v, 2 * ang / 3);
                dirs[0] = nl; dirs[1] = n1; dirs[2] = n2; dirs[3] = nr;
            } else {
                const Vector n1 = Vector_rot(nl, v, -ang / 3), n2 = Vector_rot(nl, v, -2 * ang / 3);
                dirs[0] = nr; dirs[1] = n2; dirs[2] = n1; dirs[3] = nl;
            }
        }
        const double t = sgn == 1 ? maxd : -maxd;
        for (int q = 0; q + 1 < nd; q++) {
            pts[2] = offset(pts[0], t, dirs[q]);
            pts[3] = offset(pts[1], t, dirs[q]);
            pts[4] = offset(pts[0], t, dirs[q + 1]);
            pts[5] = offset(pts[1], t, dirs[q + 1]);
            prism_cpt(pts, sgn, 0, &g);
        }
    }

    /* vertex cones (:572-641) */
    for (size_t i = 0; i < poly->num_vertices; i++) {
        const OVertex *vx = &poly->vertex_list[i];
        const int m = vx->num_faces;
        int *adjv = malloc((m ? m : 1) * sizeof(int));
        double adj_angle = 0;
        Vector axis = {0, 0, 0};
        for (int j = 0; j < m; j++) {
            const OFace *fc = &poly->face_list[vx->face_idx[j]];
            for (int k = 0; k < 3; k++) {
                if (fc->vertex_idx[k] == (int)i) {
                    adj_angle = Vector_angle(
                        Vector_sub(poly->vertex_list[fc->vertex_idx[(k + 1) % 3]].coord, vx->coord),
                        Vector_sub(poly->vertex_list[fc->vertex_idx[(k + 2) % 3]].coord, vx->coord));
                    adjv[j] = fc->vertex_idx[(k + 2) % 3];
                }
            }
            axis = Vector_lincom(1, axis, adj_angle, fc->n);
        }
        axis = Vector_normalize(axis);
        double cone = 0;
        for (int j = 0; j < m; j++) {
            const Vector nrm = poly->face_list[vx->face_idx[j]].n;
            if (Vector_dot(axis, nrm) > 0) {
                const double a = Vector_angle(axis, nrm);
                if (a > cone) cone = a;
            }
        }
        bool convex = true, concave = true;
        for (int j = 0; j < m; j++) {
            const double d = Vector_dot(Vector_sub(poly->vertex_list[adjv[j]].coord, vx->coord), axis);
            if (d > 0) convex = false;
            if (d < 0) concave = false;
        }
        if (!concave) cone_cpt(vx->coord, axis, maxd, cone, 1, &g);
        if (!convex) cone_cpt(vx->coord, (Vector){-axis.x, -axis.y, -axis.z}, maxd, cone, -1, &g);
        free(adjv);
    }

    /* fill unreached points by sweeps along z, y, x (:643-703) */
    #define FV(i, j, k) f[((size_t)(i) * ny + (j)) * nz + (k)]
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            for (int k = 1; k < nz; k++) {
                if (isnan(FV(i, j, k)) && FV(i, j, k - 1) < 0) FV(i, j, k) = -maxd;
                else if (isnan(FV(i, j, k)) && FV(i, j, k - 1) > 0) FV(i, j, k) = maxd;
            }
            for (int k = nz - 2; k >= 0; k--) {
                if (isnan(FV(i, j, k)) && FV(i, j, k + 1) < 0) FV(i, j, k) = -maxd;
                else if (isnan(FV(i, j, k)) && FV(i, j, k + 1) > 0) FV(i, j, k) = maxd;
            }
        }
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) {
            for (int j = 1; j < ny; j++) {
                if (isnan(FV(i, j, k)) && FV(i, j - 1, k) < 0) FV(i, j, k) = -maxd;
                else if (isnan(FV(i, j, k)) && FV(i, j - 1, k) > 0) FV(i, j, k) = maxd;
            }
            for (int j = ny - 2; j >= 0; j--) {
                if (isnan(FV(i, j, k)) && FV(i, j + 1, k) < 0) FV(i, j, k) = -maxd;
                else if (isnan(FV(i, j, k)) && FV(i, j + 1, k) > 0) FV(i, j, k) = maxd;
            }
        }
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
            for (int i = 1; i < nx; i++) {
                if (isnan(FV(i, j, k)) && FV(i - 1, j, k) < 0) FV(i, j, k) = -maxd;
                else if (isnan(FV(i, j, k)) && FV(i - 1, j, k) > 0) FV(i, j, k) = maxd;
            }
            for (int i = nx - 2; i >= 0; i--) {
                if (isnan(FV(i, j, k)) && FV(i + 1, j, k) < 0) FV(i, j, k) = -maxd;
                else if (isnan(FV(i, j, k)) && FV(i + 1, j, k) > 0) FV(i, j, k) = maxd;
            }
        }
    #undef FV
}

/* Analytic sphere level set (extension; no reference counterpart). */
double oracle_lvset_sphere(const double *c, double x, double y, double z) {
    const double dx = x - c[0], dy = y - c[1], dz = z - c[2];
    return sqrt(dx * dx + dy * dy + dz * dz) - c[3];
}

/* Accessors for the tests. */
size_t oracle_poly_counts(const Polyhedron *p, size_t *ne, size_t *nv) {
    *ne = p->num_edges;
    *nv = p->num_vertices;
    return p->num_faces;
}
