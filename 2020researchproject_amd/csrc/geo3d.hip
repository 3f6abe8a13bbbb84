// geo3d.hip - polyhedron geometry and the closest-point transform on the GPU.
//
// API of the reference's src/geo3d.h. The STL reader de-duplicates vertices
// and edges in first-appearance order and walks the faces around each vertex
// exactly as src/geo3d.c:229-401 does (there with GLib trees, here with hash
// maps). Polyhedron_cpt evaluates the reference's extrusion scan
// (src/geo3d.c:436-704: face prisms, edge wedges split by dihedral angle,
// vertex cones, then z/y/x sweeps filling unreached points with +-maxd), but
// inverts the loops: one GPU thread per grid point walks the feature list in
// the reference's order, so "first minimum wins" ties resolve identically.
// Feature geometry (planes, axes, angles) is built on the host in FP64; the
// per-point plane tests run with floating-point contraction disabled so they
// round like the reference's C code.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "solver.h"

static const double kPi = 3.141592653589793;

extern "C" {

// <- src/geo3d.c:120-206
Vector Vector_add(const Vector a, const Vector b) { return Vector{a.x + b.x, a.y + b.y, a.z + b.z}; }
Vector Vector_sub(const Vector a, const Vector b) { return Vector{a.x - b.x, a.y - b.y, a.z - b.z}; }
double Vector_dot(const Vector a, const Vector b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vector Vector_crs(const Vector a, const Vector b) {
    return Vector{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
double Vector_norm(const Vector v) { return std::sqrt(Vector_dot(v, v)); }
double Vector_dist(const Vector a, const Vector b) { return Vector_norm(Vector_sub(a, b)); }
double Vector_angle(const Vector a, const Vector b) {
    const double la = Vector_norm(a), lb = Vector_norm(b);
    double c = Vector_dot(a, b) / (la * lb);
    c = c < -1 ? -1 : c > 1 ? 1 : c;
    return std::acos(c);
}
Vector Vector_normalize(const Vector v) {
    const double len = Vector_norm(v);
    return Vector{v.x / len, v.y / len, v.z / len};
}
Vector Vector_lincom(const double a, const Vector u, const double b, const Vector v) {
    return Vector{a * u.x + b * v.x, a * u.y + b * v.y, a * u.z + b * v.z};
}
Vector Vector_rot(const Vector v, const Vector axis, const double ang) {
    const Vector n = Vector_normalize(axis);
    const double c = std::cos(ang), s = std::sin(ang), t = 1 - std::cos(ang);
    Vector r;
    r.x = (c + n.x * n.x * t) * v.x + (n.x * n.y * t - n.z * s) * v.y + (n.x * n.z * t + n.y * s) * v.z;
    r.y = (n.x * n.y * t + n.z * s) * v.x + (c + n.y * n.y * t) * v.y + (n.y * n.z * t - n.x * s) * v.z;
    r.z = (n.x * n.z * t - n.y * s) * v.x + (n.y * n.z * t + n.x * s) * v.y + (c + n.z * n.z * t) * v.z;
    return r;
}

}  // extern "C"

struct PVertex { Vector c; std::vector<int> faces; };
struct PEdge { int v[2]; int f[2]; };
struct PFace { int v[3]; int e[3]; Vector n; };

struct _polyhedron {
    std::vector<PVertex> verts;
    std::vector<PEdge> edges;
    std::vector<PFace> faces;
};

namespace {
struct VKey {
    double x, y, z;
    bool operator==(const VKey &o) const { return x == o.x && y == o.y && z == o.z; }
};
struct VHash {
    size_t operator()(const VKey &k) const {
        uint64_t a, b, c;
        memcpy(&a, &k.x, 8); memcpy(&b, &k.y, 8); memcpy(&c, &k.z, 8);
        uint64_t h = a * 0x9E3779B97F4A7C15ULL ^ (b + 0x632BE59BD9B4E019ULL) * 0xC2B2AE3D27D4EB4FULL ^ c;
        return (size_t)(h ^ (h >> 29));
    }
};
}  // namespace

extern "C" {

Polyhedron *Polyhedron_new(void) { return new _polyhedron(); }
void Polyhedron_destroy(Polyhedron *p) { delete p; }

// <- src/geo3d.c:229-401
void Polyhedron_read_stl(Polyhedron *poly, FILE *f) {
    char header[80];
    uint32_t nf = 0;
    if (fread(header, 1, 80, f) != 80 || fread(&nf, 4, 1, f) != 1) ibm::fatal("cannot read STL header");
    poly->faces.assign(nf, PFace{});
    poly->verts.clear();
    poly->edges.clear();
    std::unordered_map<VKey, int, VHash> vmap;
    vmap.reserve(3 * (size_t)nf);
    std::vector<int> vfirst;
    for (uint32_t i = 0; i < nf; i++) {
        float c[12];
        uint16_t attr;
        if (fread(c, 4, 12, f) != 12 || fread(&attr, 2, 1, f) != 1) ibm::fatal("truncated STL file");
        PFace &F = poly->faces[i];
        F.n = Vector{c[0], c[1], c[2]};
        for (int j = 0; j < 3; j++) {
            const VKey key{(double)c[3 + 3 * j], (double)c[4 + 3 * j], (double)c[5 + 3 * j]};
            auto it = vmap.find(key);
            int v;
            if (it == vmap.end()) {
                v = (int)poly->verts.size();
                vmap.emplace(key, v);
                poly->verts.push_back(PVertex{Vector{key.x, key.y, key.z}, {}});
                vfirst.push_back((int)i);
            } else {
                v = it->second;
            }
            F.v[j] = v;
            poly->verts[v].faces.push_back(-1);  // counts faces; filled by the walk below
        }
    }
    std::unordered_map<uint64_t, int> emap;
    emap.reserve(3 * (size_t)nf);
    for (uint32_t i = 0; i < nf; i++)
        for (int j = 0; j < 3; j++) {
            const int a = poly->faces[i].v[j], b = poly->faces[i].v[(j + 1) % 3];
            const uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
            auto it = emap.find(key);
            int e;
            if (it == emap.end()) {
                e = (int)poly->edges.size();
                emap.emplace(key, e);
                poly->edges.push_back(PEdge{{a, b}, {(int)i, -1}});
            } else {
                e = it->second;
                poly->edges[e].f[1] = (int)i;
            }
            poly->faces[i].e[j] = e;
        }
    // faces around each vertex: start at the first face, cross the edge that
    // precedes the vertex in that face (:344-368)
    for (size_t v = 0; v < poly->verts.size(); v++) {
        std::vector<int> &fl = poly->verts[v].faces;
        int cur = vfirst[v];
        fl[0] = cur;
        for (size_t q = 1; q < fl.size(); q++) {
            const PFace &F = poly->faces[cur];
            int e;
            if (F.v[0] == (int)v) e = F.e[2];
            else if (F.v[1] == (int)v) e = F.e[0];
            else e = F.e[1];
            cur = poly->edges[e].f[0] == cur ? poly->edges[e].f[1] : poly->edges[e].f[0];
            if (cur < 0) ibm::fatal("STL surface is not closed (an edge has one face)");
            fl[q] = cur;
        }
    }
}

void Polyhedron_print_stats(Polyhedron *poly) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    printf("Input polyhedron size: %zu faces, %zu edges, %zu vertices\n", poly->faces.size(), poly->edges.size(),
           poly->verts.size());
    for (const PVertex &v : poly->verts) {
        const double c[3] = {v.c.x, v.c.y, v.c.z};
        for (int a = 0; a < 3; a++) { mn[a] = std::min(mn[a], c[a]); mx[a] = std::max(mx[a], c[a]); }
    }
    printf("  xmin: %10.4lf, xmax: %10.4lf\n", mn[0], mx[0]);
    printf("  ymin: %10.4lf, ymax: %10.4lf\n", mn[1], mx[1]);
    printf("  zmin: %10.4lf, zmax: %10.4lf\n", mn[2], mx[2]);
}

void Polyhedron_scale(Polyhedron *poly, double s) {
    for (PVertex &v : poly->verts) { v.c.x *= s; v.c.y *= s; v.c.z *= s; }
}

void Polyhedron_cpt(const Polyhedron *poly, const int nx, const int ny, const int nz, const double *x,
                    const double *y, const double *z, double *f, const double maxd) {
    ibm::polyhedron_cpt_grid(poly, nx, ny, nz, x, y, z, f, maxd);
}

int Polyhedron_cpt_noalloc(const Polyhedron *poly, const int nx, const int ny, const int nz, const double *x,
                           const double *y, const double *z, double *f, const double maxd) {
    return ibm::polyhedron_cpt_grid_noalloc(poly, nx, ny, nz, x, y, z, f, maxd);
}

}  // extern "C"

namespace ibm {

struct CptFeature {
    int kind;            // 0 face prism, 1 edge prism, 2 vertex cone
    int sgn;
    double lo[3], hi[3]; // point bounding box (inclusive)
    double pl[5][4];     // bounding planes a, b, c, d (prisms)
    double p0[3], p1[3]; // edge segment (edge) / vertex (cone)
    double nrm0;         // |n| of plane 0 (face distance)
    double axis[3], axis_norm, radius, angle;
};

static void plane3(Vector a, Vector b, Vector c, double out[4]) {  // src/geo3d.c:706-716
    const Vector n = Vector_crs(Vector_sub(b, a), Vector_sub(c, a));
    out[0] = n.x; out[1] = n.y; out[2] = n.z; out[3] = -Vector_dot(n, a);
}

static CptFeature prism(const Vector pts[6], int sgn, bool face) {
    CptFeature F{};
    F.kind = face ? 0 : 1;
    F.sgn = sgn;
    for (int a = 0; a < 3; a++) { F.lo[a] = INFINITY; F.hi[a] = -INFINITY; }
    for (int q = 0; q < 6; q++) {
        const double c[3] = {pts[q].x, pts[q].y, pts[q].z};
        for (int a = 0; a < 3; a++) { F.lo[a] = std::min(F.lo[a], c[a]); F.hi[a] = std::max(F.hi[a], c[a]); }
    }
    if (face) {
        plane3(pts[0], pts[2], pts[1], F.pl[0]); plane3(pts[3], pts[4], pts[5], F.pl[1]);
        plane3(pts[0], pts[1], pts[4], F.pl[2]); plane3(pts[1], pts[2], pts[5], F.pl[3]);
        plane3(pts[0], pts[3], pts[5], F.pl[4]);
    } else {
        plane3(pts[0], pts[4], pts[2], F.pl[0]); plane3(pts[1], pts[3], pts[5], F.pl[1]);
        plane3(pts[0], pts[2], pts[3], F.pl[2]); plane3(pts[2], pts[4], pts[5], F.pl[3]);
        plane3(pts[0], pts[1], pts[5], F.pl[4]);
    }
    F.nrm0 = Vector_norm(Vector{F.pl[0][0], F.pl[0][1], F.pl[0][2]});
    F.p0[0] = pts[0].x; F.p0[1] = pts[0].y; F.p0[2] = pts[0].z;
    F.p1[0] = pts[1].x; F.p1[1] = pts[1].y; F.p1[2] = pts[1].z;
    return F;
}

// Feature list in the reference's processing order (src/geo3d.c:458-641).
static std::vector<CptFeature> build_features(const Polyhedron *poly, double maxd) {
    std::vector<CptFeature> out;
    for (const PFace &F : poly->faces) {
        Vector pts[6];
        for (int j = 0; j < 3; j++) {
            pts[j] = poly->verts[F.v[j]].c;
            pts[j + 3] = Vector_lincom(1, pts[j], maxd, F.n);
        }
        out.push_back(prism(pts, 1, true));
        for (int j = 0; j < 3; j++) {
            pts[j] = poly->verts[F.v[(3 - j) % 3]].c;
            pts[j + 3] = Vector_lincom(1, pts[j], -maxd, F.n);
        }
        out.push_back(prism(pts, -1, true));
    }
    for (const PEdge &E : poly->edges) {
        Vector pts[6];
        pts[0] = poly->verts[E.v[0]].c;
        pts[1] = poly->verts[E.v[1]].c;
        const Vector v = Vector_sub(pts[1], pts[0]);
        const Vector nl = poly->faces[E.f[0]].n, nr = poly->faces[E.f[1]].n;
        const int sgn = Vector_dot(Vector_crs(v, nl), nr) >= 0 ? 1 : -1;
        const double ang = Vector_angle(nl, nr);
        std::vector<Vector> dirs;
        if (ang <= kPi / 3) {
            dirs = sgn == 1 ? std::vector<Vector>{nl, nr} : std::vector<Vector>{nr, nl};
        } else if (ang <= 2 * kPi / 3) {
            const Vector nm = Vector_rot(nl, v, ang / 2);
            dirs = sgn == 1 ? std::vector<Vector>{nl, nm, nr} : std::vector<Vector>{nr, nm, nl};
        } else if (sgn == 1) {
            dirs = {nl, Vector_rot(nl, v, ang / 3), Vector_rot(nl, v, 2 * ang / 3), nr};
        } else {
            dirs = {nr, Vector_rot(nl, v, -2 * ang / 3), Vector_rot(nl, v, -ang / 3), nl};
        }
        const double t = sgn == 1 ? maxd : -maxd;
        for (size_t q = 0; q + 1 < dirs.size(); q++) {
            pts[2] = Vector_lincom(1, pts[0], t, dirs[q]);
            pts[3] = Vector_lincom(1, pts[1], t, dirs[q]);
            pts[4] = Vector_lincom(1, pts[0], t, dirs[q + 1]);
            pts[5] = Vector_lincom(1, pts[1], t, dirs[q + 1]);
            out.push_back(prism(pts, sgn, false));
        }
    }
    for (size_t i = 0; i < poly->verts.size(); i++) {
        const PVertex &V = poly->verts[i];
        const int m = (int)V.faces.size();
        std::vector<int> adjv(m);
        double adj_angle = 0;
        Vector axis{0, 0, 0};
        for (int j = 0; j < m; j++) {
            const PFace &F = poly->faces[V.faces[j]];
            for (int k = 0; k < 3; k++)
                if (F.v[k] == (int)i) {
                    adj_angle = Vector_angle(Vector_sub(poly->verts[F.v[(k + 1) % 3]].c, V.c),
                                             Vector_sub(poly->verts[F.v[(k + 2) % 3]].c, V.c));
                    adjv[j] = F.v[(k + 2) % 3];
                }
            axis = Vector_lincom(1, axis, adj_angle, F.n);
        }
        axis = Vector_normalize(axis);
        double cone = 0;
        for (int j = 0; j < m; j++) {
            const Vector nrm = poly->faces[V.faces[j]].n;
            if (Vector_dot(axis, nrm) > 0) cone = std::max(cone, Vector_angle(axis, nrm));
        }
        bool convex = true, concave = true;
        for (int j = 0; j < m; j++) {
            const double d = Vector_dot(Vector_sub(poly->verts[adjv[j]].c, V.c), axis);
            if (d > 0) convex = false;
            if (d < 0) concave = false;
        }
        for (int sgn : {1, -1}) {
            if (sgn == 1 && concave) continue;
            if (sgn == -1 && convex) continue;
            CptFeature F{};
            F.kind = 2;
            F.sgn = sgn;
            const Vector ax = sgn == 1 ? axis : Vector{-axis.x, -axis.y, -axis.z};
            F.axis[0] = ax.x; F.axis[1] = ax.y; F.axis[2] = ax.z;
            F.axis_norm = Vector_norm(ax);
            F.radius = maxd;
            F.angle = cone;
            F.p0[0] = V.c.x; F.p0[1] = V.c.y; F.p0[2] = V.c.z;
            F.lo[0] = V.c.x - maxd; F.hi[0] = V.c.x + maxd;
            F.lo[1] = V.c.y - maxd; F.hi[1] = V.c.y + maxd;
            F.lo[2] = V.c.z - maxd; F.hi[2] = V.c.z + maxd;
            out.push_back(F);
        }
    }
    return out;
}

// one thread per grid point (layout [i][j][k], k fastest, as the reference)
__global__ void __launch_bounds__(256) k_cpt(int nx, int ny, int nz, const double *x, const double *y,
                                             const double *z, const CptFeature *feat, int nfeat, double *f) {
#pragma clang fp contract(off)
    const long q = blockIdx.x * 256L + threadIdx.x;
    const long n = (long)nx * ny * nz;
    if (q >= n) return;
    const int k = q % nz, j = (q / nz) % ny, i = q / ((long)nz * ny);
    const double cx = x[i], cy = y[j], cz = z[k];
    double fv = __builtin_nan("");
    for (int t = 0; t < nfeat; t++) {
        const CptFeature &F = feat[t];
        if (cx < F.lo[0] || cx > F.hi[0] || cy < F.lo[1] || cy > F.hi[1] || cz < F.lo[2] || cz > F.hi[2]) continue;
        double dist;
        if (F.kind == 2) {
            const double dx = cx - F.p0[0], dy = cy - F.p0[1], dz = cz - F.p0[2];
            const double dd = dx * dx + dy * dy + dz * dz;
            dist = sqrt(dd);
            double c = (F.axis[0] * dx + F.axis[1] * dy + F.axis[2] * dz) / (F.axis_norm * dist);
            c = c < -1 ? -1 : c > 1 ? 1 : c;
            const double ang = acos(c);
            if (!(ang <= F.angle + 1e-12 && dist <= F.radius + 1e-12)) continue;
        } else {
            bool in = true;
            for (int l = 0; l < 5; l++)
                if (F.pl[l][0] * cx + F.pl[l][1] * cy + F.pl[l][2] * cz + F.pl[l][3] > 1e-12) in = false;
            if (!in) continue;
            if (F.kind == 0) {
                dist = fabs(F.pl[0][0] * cx + F.pl[0][1] * cy + F.pl[0][2] * cz + F.pl[0][3]) / F.nrm0;
            } else {
                const double abx = F.p1[0] - F.p0[0], aby = F.p1[1] - F.p0[1], abz = F.p1[2] - F.p0[2];
                const double avx = cx - F.p0[0], avy = cy - F.p0[1], avz = cz - F.p0[2];
                const double tt = (abx * avx + aby * avy + abz * avz) / (abx * abx + aby * aby + abz * abz);
                const double hx = (1 - tt) * F.p0[0] + tt * F.p1[0];
                const double hy = (1 - tt) * F.p0[1] + tt * F.p1[1];
                const double hz = (1 - tt) * F.p0[2] + tt * F.p1[2];
                const double ex = cx - hx, ey = cy - hy, ez = cz - hz;
                dist = sqrt(ex * ex + ey * ey + ez * ez);
            }
        }
        if (isnan(fv) || dist < fabs(fv)) fv = F.sgn * dist;
    }
    f[q] = fv;
}

// sequential fills of unreached points along lines (src/geo3d.c:643-703)
__global__ void k_cpt_fill(int nx, int ny, int nz, double *f, double maxd, int axis) {
    const long t = blockIdx.x * 256L + threadIdx.x;
    int n, a, b;
    long st, base;
    if (axis == 2) {  // lines along k: (i, j)
        if (t >= (long)nx * ny) return;
        a = t / ny; b = t % ny; n = nz; st = 1; base = ((long)a * ny + b) * nz;
    } else if (axis == 1) {  // along j: (i, k)
        if (t >= (long)nx * nz) return;
        a = t / nz; b = t % nz; n = ny; st = nz; base = (long)a * ny * nz + b;
    } else {  // along i: (j, k)
        if (t >= (long)ny * nz) return;
        a = t / nz; b = t % nz; n = nx; st = (long)ny * nz; base = (long)a * nz + b;
    }
    for (int c = 1; c < n; c++) {
        double &v = f[base + c * st];
        const double pv = f[base + (c - 1) * st];
        if (isnan(v) && pv < 0) v = -maxd;
        else if (isnan(v) && pv > 0) v = maxd;
    }
    for (int c = n - 2; c >= 0; c--) {
        double &v = f[base + c * st];
        const double nv = f[base + (c + 1) * st];
        if (isnan(v) && nv < 0) v = -maxd;
        else if (isnan(v) && nv > 0) v = maxd;
    }
}

void polyhedron_cpt_grid(const Polyhedron *poly, int nx, int ny, int nz, const double *x, const double *y,
                         const double *z, double *f, double maxd) {
    const std::vector<CptFeature> feat = build_features(poly, maxd);
    const long n = (long)nx * ny * nz;
    // legacy-stream launches and copies: keep other threads' graph captures out
    HipGlobalLock lk(hip_global_mutex());
    DBuf<double> dx, dy, dz, df;
    DBuf<CptFeature> dfeat;
    dx.upload(x, nx); dy.upload(y, ny); dz.upload(z, nz);
    dfeat.upload(feat);
    df.alloc(n);
    hipLaunchKernelGGL(k_cpt, dim3((n + 255) / 256), dim3(256), 0, 0, nx, ny, nz, dx.p, dy.p, dz.p, dfeat.p,
                       (int)feat.size(), df.p);
    const long lines[3] = {(long)ny * nz, (long)nx * nz, (long)nx * ny};
    for (int ax : {2, 1, 0})
        hipLaunchKernelGGL(k_cpt_fill, dim3((lines[ax] + 255) / 256), dim3(256), 0, 0, nx, ny, nz, df.p, maxd, ax);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(f, df.p, n * sizeof(double), hipMemcpyDeviceToHost));
}

// ---------------------------------------------------------------------------
// Polyhedron_cpt_noalloc: the same transform with every device buffer (the
// coordinates, the feature list, the grid of distances) placed in one
// file-scope device array instead of allocations, for callers that cannot
// allocate device memory. The array lives as long as the process on every GPU
// the library runs on; each GPU has its own copy at its own address, found
// once per device with hipGetSymbolAddress (the host-side name `scratch` is
// not a device address and is never passed to a kernel or a copy).
// ---------------------------------------------------------------------------
constexpr size_t kCptScratchBytes = IBM3D_CPT_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kCptScratchBytes];

constexpr int kMaxDevices = 64;
static unsigned char *g_scratch_dev[kMaxDevices];  // per device, set on first use there

// Device address of `scratch` on the current device (looked up once per
// device, with that device current), or null with a message on failure.
static unsigned char *scratch_on_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        fprintf(stderr, "error: Polyhedron_cpt_noalloc: no usable current device\n");
        return nullptr;
    }
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            fprintf(stderr, "error: Polyhedron_cpt_noalloc: hipGetSymbolAddress failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_dev[dev] = static_cast<unsigned char *>(ptr);
    }
    return g_scratch_dev[dev];
}

static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

int polyhedron_cpt_grid_noalloc(const Polyhedron *poly, int nx, int ny, int nz, const double *x, const double *y,
                                const double *z, double *f, double maxd) {
    if (nx <= 0 || ny <= 0 || nz <= 0) {
        fprintf(stderr, "error: Polyhedron_cpt_noalloc: empty grid %d x %d x %d\n", nx, ny, nz);
        return -1;
    }
    const std::vector<CptFeature> feat = build_features(poly, maxd);
    const long n = (long)nx * ny * nz;
    const size_t ox = 0, oy = ox + round256(nx * sizeof(double)), oz = oy + round256(ny * sizeof(double));
    const size_t of = oz + round256(nz * sizeof(double));
    const size_t op = of + round256(std::max<size_t>(feat.size(), 1) * sizeof(CptFeature));
    const size_t need = op + round256(n * sizeof(double));
    if (need > kCptScratchBytes) {
        fprintf(stderr,
                "error: Polyhedron_cpt_noalloc: %ld points and %zu extrusion features need %zu bytes of device "
                "scratch, more than its %zu\n",
                n, feat.size(), need, kCptScratchBytes);
        return -1;
    }
    // one caller at a time in this process: the scratch array is shared
    HipGlobalLock lk(hip_global_mutex());
    unsigned char *base = scratch_on_current_device();
    if (!base) return -1;
    double *dx = reinterpret_cast<double *>(base + ox), *dy = reinterpret_cast<double *>(base + oy);
    double *dz = reinterpret_cast<double *>(base + oz), *df = reinterpret_cast<double *>(base + op);
    CptFeature *dfeat = reinterpret_cast<CptFeature *>(base + of);
    HIP_CHECK(hipMemcpy(dx, x, nx * sizeof(double), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dy, y, ny * sizeof(double), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dz, z, nz * sizeof(double), hipMemcpyHostToDevice));
    if (!feat.empty()) HIP_CHECK(hipMemcpy(dfeat, feat.data(), feat.size() * sizeof(CptFeature), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_cpt, dim3((n + 255) / 256), dim3(256), 0, 0, nx, ny, nz, dx, dy, dz, dfeat,
                       (int)feat.size(), df);
    const long lines[3] = {(long)ny * nz, (long)nx * nz, (long)nx * ny};
    for (int ax : {2, 1, 0})
        hipLaunchKernelGGL(k_cpt_fill, dim3((lines[ax] + 255) / 256), dim3(256), 0, 0, nx, ny, nz, df, maxd, ax);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(f, df, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

}  // namespace ibm

extern "C" size_t ibm3d_polyhedron_counts(const Polyhedron *p, size_t *ne, size_t *nv) {
    *ne = p->edges.size();
    *nv = p->verts.size();
    return p->faces.size();
}
