// step_kernels.hip - stencil kernels of one fractional step (FP64).
//
// Each kernel restates one stage of the reference's IBMSolver_iterate
// (src/ibm3d_fracstep.c) for the x-contiguous device layout of common.h:
//   k_interp_stag    interp_stag_vel            :807-845
//   k_conv_rhs       calc_N + momentum RHS      :117-170, :225-269
//   k_face_star      calc_u_tilde + calc_U_star :388-487 (u~ formed on the fly)
//   k_poisson_rhs    pressure RHS               :539-552
//   k_update         update_next                :709-792
//   k_outer_*        update_outer               :855-1906
//   k_ghost_*        update_ghost               :1915-1976
//   k_sgs            Smagorinsky / Vreman stress src/ibm3d_turb.c:8-82
// One thread per cell; blocks of 64 (x) by 4 (y) threads so a wavefront reads
// one 512-byte x-row segment per array.
#include "step_kernels.h"

namespace ibm {

#define C1(a, i) ((a)[(i) + 2])

__device__ inline double nan_d() { return __builtin_nan(""); }

// ---------------------------------------------------------------------------
// block reduction of NV doubles into partial[blockLinear * NV + v]
// ---------------------------------------------------------------------------
template <int NV>
__device__ inline void block_sum_store(double (&v)[NV], double *partial, long blk) {
    __shared__ double sh[NV][16];
    const int tid = threadIdx.x + threadIdx.y * blockDim.x + threadIdx.z * blockDim.x * blockDim.y;
    const int nth = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int a = 0; a < NV; a++) {
        double x = v[a];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) sh[a][wave] = x;
    }
    __syncthreads();
    if (tid == 0) {
        const int nw = (nth + 63) >> 6;
#pragma unroll
        for (int a = 0; a < NV; a++) {
            double acc = 0;
            for (int w = 0; w < nw; w++) acc += sh[a][w];
            partial[blk * NV + a] = acc;
        }
    }
}

__device__ inline long block_linear() {
    return blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
}

// ---------------------------------------------------------------------------
// interp_stag_vel: face velocities from cell velocities over the padded
// tangential extent.  <- src/ibm3d_fracstep.c:807-845
// ---------------------------------------------------------------------------
__global__ void k_interp_stag(Geo g, const double *u1, const double *u2, const double *u3,
                              double *U1, double *U2, double *U3) {
    const Dims &d = g.d;
    const int i = blockIdx.x * blockDim.x + threadIdx.x - 2;
    const int j = blockIdx.y * blockDim.y + threadIdx.y - 2;
    const int k = blockIdx.z - 2;
    if (i > d.nx + 1 || j > d.ny + 1) return;
    const long q = cidx(d, i, j, k);
    if (i >= 0 && i <= d.nx)
        U1[q] = (u1[q - 1] * C1(g.dx, i) + u1[q] * C1(g.dx, i - 1)) / (C1(g.dx, i - 1) + C1(g.dx, i));
    if (j >= 0 && j <= d.ny)
        U2[q] = (u2[q - d.sy] * C1(g.dy, j) + u2[q] * C1(g.dy, j - 1)) / (C1(g.dy, j - 1) + C1(g.dy, j));
    if (k >= 0 && k <= d.nz)
        U3[q] = (u3[q - d.sz] * C1(g.dz, k) + u3[q] * C1(g.dz, k - 1)) / (C1(g.dz, k - 1) + C1(g.dz, k));
}

// ---------------------------------------------------------------------------
// K1: convection N (AB2 history) and the Crank-Nicolson momentum RHS of the
// three components, fused.  <- src/ibm3d_fracstep.c:117-170 and :237-269
// Writes N (non-fluid: NaN) and b_l (non-fluid rows: 0); accumulates the
// squared norms of b_l over the interior rows for the stopping test.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_conv_rhs(Geo g, StepParams sp, const uint8_t *flag,
                                                  const double *u1, const double *u2, const double *u3,
                                                  const double *U1, const double *U2, const double *U3,
                                                  const double *p, const double *Np1, const double *Np2,
                                                  const double *Np3, double *N1, double *N2, double *N3,
                                                  double *b1, double *b2, double *b3, TauPtrs tau,
                                                  double *partial) {
    const Dims &d = g.d;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    double nrm[3] = {0, 0, 0};
    if (i < d.nx && j < d.ny) {
        // every cell is computed (the neighbours of interior cells are in the
        // padded range) and non-fluid results are replaced at the end, so all
        // loads issue before the first wait
        const long q = cidx(d, i, j, k);
        const long sy = d.sy, sz = d.sz;
        const bool fl = flag[q] == F_FLUID;
        const double dxm = C1(g.dx, i - 1), dx0 = C1(g.dx, i), dxp = C1(g.dx, i + 1);
        const double dym = C1(g.dy, j - 1), dy0 = C1(g.dy, j), dyp = C1(g.dy, j + 1);
        const double dzm = C1(g.dz, k - 1), dz0 = C1(g.dz, k), dzp = C1(g.dz, k + 1);
        const double Uw = U1[q], Ue = U1[q + 1], Vs = U2[q], Vn = U2[q + sy], Wd = U3[q], Wu = U3[q + sz];
        const double kW = C1(g.kW, i), kE = C1(g.kE, i), kS = C1(g.kS, j), kN = C1(g.kN, j);
        const double kD = C1(g.kD, k), kU = C1(g.kU, k);
        const double kc = 1 - kW - kE - kS - kN - kD - kU;
        const double ddx = C1(g.xc, i + 1) - C1(g.xc, i - 1);
        const double ddy = C1(g.yc, j + 1) - C1(g.yc, j - 1);
        const double ddz = C1(g.zc, k + 1) - C1(g.zc, k - 1);
        const double *uu[3] = {u1, u2, u3};
        const double *Np[3] = {Np1, Np2, Np3};
        double *NN[3] = {N1, N2, N3};
        double *bb[3] = {b1, b2, b3};
        const double dp[3] = {(p[q + 1] - p[q - 1]) / ddx, (p[q + sy] - p[q - sy]) / ddy,
                              (p[q + sz] - p[q - sz]) / ddz};
#pragma unroll
        for (int l = 0; l < 3; l++) {
            const double *a = uu[l];
            const double c = a[q], xm = a[q - 1], xp = a[q + 1], ym = a[q - sy], yp = a[q + sy];
            const double zm = a[q - sz], zp = a[q + sz];
            const double npl = Np[l][q];
            const double uw = (xm * dx0 + c * dxm) / (dxm + dx0);
            const double ue = (c * dxp + xp * dx0) / (dx0 + dxp);
            const double us = (ym * dy0 + c * dym) / (dym + dy0);
            const double un = (c * dyp + yp * dy0) / (dy0 + dyp);
            const double ud = (zm * dz0 + c * dzm) / (dzm + dz0);
            const double uu_ = (c * dzp + zp * dz0) / (dz0 + dzp);
            const double n = (Ue * ue - Uw * uw) / dx0 + (Vn * un - Vs * us) / dy0 + (Wu * uu_ - Wd * ud) / dz0;
            double divt = 0;
            if (tau.on) {
                const double *t1 = tau.t[tau.slot[l][0]], *t2 = tau.t[tau.slot[l][1]], *t3 = tau.t[tau.slot[l][2]];
                divt = (t1[q + 1] - t1[q - 1]) / ddx + (t2[q + sy] - t2[q - sy]) / ddy + (t3[q + sz] - t3[q - sz]) / ddz;
            }
            const double b = -sp.dt / 2 * (3 * n - npl) - sp.dt * dp[l] + kc * c
                + kW * xm + kE * xp + kS * ym + kN * yp + kD * zm + kU * zp
                + sp.dt * (sp.fa[l] ? sp.fa[l][q] : sp.force[l]) - sp.dt * divt;
            NN[l][q] = fl ? n : nan_d();
            bb[l][q] = fl ? b : 0.0;
            nrm[l] = fl ? b * b : 0.0;
        }
    }
    block_sum_store<3>(nrm, partial, block_linear());
}

// ---------------------------------------------------------------------------
// K2a: predicted face velocities U*.  <- src/ibm3d_fracstep.c:388-487
// U* = interp(u~) - dt dp/dn with u~ = u* + dt grad p (central), except on
// fluid|ghost faces where U* = interp(u*). Threads cover (nx+1)(ny+1)(nz+1)
// cells; cell (i,j,k) owns its three low faces.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_face_star(Geo g, StepParams sp, const uint8_t *flag,
                                                   const double *us1, const double *us2, const double *us3,
                                                   const double *p, double *F1, double *F2, double *F3) {
    const Dims &d = g.d;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > d.nx || j > d.ny) return;
    const long q = cidx(d, i, j, k);
    const long sy = d.sy, sz = d.sz;
    const double dt = sp.dt;
    const uint8_t f0 = flag[q];
    if (j < d.ny && k < d.nz) {  // x face i-1/2
        const uint8_t fm = flag[q - 1];
        const double dm = C1(g.dx, i - 1), d0 = C1(g.dx, i);
        double v;
        if ((fm == F_FLUID && f0 == F_GHOST) || (fm == F_GHOST && f0 == F_FLUID)) {
            v = (us1[q - 1] * d0 + us1[q] * dm) / (dm + d0);
        } else {
            const double tm = us1[q - 1] + dt * (p[q] - p[q - 2]) / (C1(g.xc, i) - C1(g.xc, i - 2));
            const double t0 = us1[q] + dt * (p[q + 1] - p[q - 1]) / (C1(g.xc, i + 1) - C1(g.xc, i - 1));
            v = (tm * d0 + t0 * dm) / (dm + d0) - dt * (p[q] - p[q - 1]) / (C1(g.xc, i) - C1(g.xc, i - 1));
        }
        F1[q] = v;
    }
    if (i < d.nx && k < d.nz) {  // y face j-1/2
        const uint8_t fm = flag[q - sy];
        const double dm = C1(g.dy, j - 1), d0 = C1(g.dy, j);
        double v;
        if ((fm == F_FLUID && f0 == F_GHOST) || (fm == F_GHOST && f0 == F_FLUID)) {
            v = (us2[q - sy] * d0 + us2[q] * dm) / (dm + d0);
        } else {
            const double tm = us2[q - sy] + dt * (p[q] - p[q - 2 * sy]) / (C1(g.yc, j) - C1(g.yc, j - 2));
            const double t0 = us2[q] + dt * (p[q + sy] - p[q - sy]) / (C1(g.yc, j + 1) - C1(g.yc, j - 1));
            v = (tm * d0 + t0 * dm) / (dm + d0) - dt * (p[q] - p[q - sy]) / (C1(g.yc, j) - C1(g.yc, j - 1));
        }
        F2[q] = v;
    }
    if (i < d.nx && j < d.ny) {  // z face k-1/2
        const uint8_t fm = flag[q - sz];
        const double dm = C1(g.dz, k - 1), d0 = C1(g.dz, k);
        double v;
        if ((fm == F_FLUID && f0 == F_GHOST) || (fm == F_GHOST && f0 == F_FLUID)) {
            v = (us3[q - sz] * d0 + us3[q] * dm) / (dm + d0);
        } else {
            const double tm = us3[q - sz] + dt * (p[q] - p[q - 2 * sz]) / (C1(g.zc, k) - C1(g.zc, k - 2));
            const double t0 = us3[q] + dt * (p[q + sz] - p[q - sz]) / (C1(g.zc, k + 1) - C1(g.zc, k - 1));
            v = (tm * d0 + t0 * dm) / (dm + d0) - dt * (p[q] - p[q - sz]) / (C1(g.zc, k) - C1(g.zc, k - 1));
        }
        F3[q] = v;
    }
}

// ---------------------------------------------------------------------------
// K2b: pressure-correction RHS b = -(1/2Re) div U* on fluid rows (the
// reference divides it by the diagonal sum K, :545-550; the solver here keeps
// the unnormalised row and applies that scaling only in the stopping test).
// Accumulates ||b / sum K||^2.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_poisson_rhs(Geo g, StepParams sp, const uint8_t *flag, long pin,
                                                     const double *F1, const double *F2, const double *F3,
                                                     double *b, double *partial) {
    const Dims &d = g.d;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    double nrm[1] = {0};
    if (i < d.nx && j < d.ny) {
        const long q = cidx(d, i, j, k);
        double v = 0;
        if (q != pin && flag[q] == F_FLUID) {
            v = -1 / (2 * sp.Re) * ((F1[q + 1] - F1[q]) / C1(g.dx, i) + (F2[q + d.sy] - F2[q]) / C1(g.dy, j)
                                    + (F3[q + d.sz] - F3[q]) / C1(g.dz, k));
            const double ks = C1(g.kN, j) + C1(g.kE, i) + C1(g.kS, j) + C1(g.kW, i) + C1(g.kD, k) + C1(g.kU, k);
            const double vn = v / ks;
            nrm[0] = vn * vn;
        }
        b[q] = v;
    }
    block_sum_store<1>(nrm, partial, block_linear());
}

// ---------------------------------------------------------------------------
// K3: projection.  <- src/ibm3d_fracstep.c:740-792
// p_next = p + p' (fluid, ghost) / NaN (solid); u_next = u* - dt grad p'
// (fluid) / NaN (solid); ghost entries of u_next are left untouched, as in
// the reference. U_next = U* - dt dp'/dn on faces with no solid side, else NaN.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_update(Geo g, StepParams sp, const uint8_t *flag,
                                                const double *p, const double *pp,
                                                const double *us1, const double *us2, const double *us3,
                                                const double *F1, const double *F2, const double *F3,
                                                double *pn, double *un1, double *un2, double *un3,
                                                double *G1, double *G2, double *G3) {
    const Dims &d = g.d;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > d.nx || j > d.ny) return;
    const long q = cidx(d, i, j, k);
    const long sy = d.sy, sz = d.sz;
    const double dt = sp.dt;
    const uint8_t f0 = flag[q];
    if (i < d.nx && j < d.ny && k < d.nz) {
        pn[q] = (f0 == F_FLUID || f0 == F_GHOST) ? p[q] + pp[q] : nan_d();
        if (f0 == F_FLUID) {
            un1[q] = us1[q] - dt * (pp[q + 1] - pp[q - 1]) / (C1(g.xc, i + 1) - C1(g.xc, i - 1));
            un2[q] = us2[q] - dt * (pp[q + sy] - pp[q - sy]) / (C1(g.yc, j + 1) - C1(g.yc, j - 1));
            un3[q] = us3[q] - dt * (pp[q + sz] - pp[q - sz]) / (C1(g.zc, k + 1) - C1(g.zc, k - 1));
        } else if (f0 == F_SOLID) {
            un1[q] = un2[q] = un3[q] = nan_d();
        }
    }
    if (j < d.ny && k < d.nz)
        G1[q] = (flag[q - 1] != F_SOLID && f0 != F_SOLID)
            ? F1[q] - dt * (pp[q] - pp[q - 1]) / (C1(g.xc, i) - C1(g.xc, i - 1)) : nan_d();
    if (i < d.nx && k < d.nz)
        G2[q] = (flag[q - sy] != F_SOLID && f0 != F_SOLID)
            ? F2[q] - dt * (pp[q] - pp[q - sy]) / (C1(g.yc, j) - C1(g.yc, j - 1)) : nan_d();
    if (i < d.nx && j < d.ny)
        G3[q] = (flag[q - sz] != F_SOLID && f0 != F_SOLID)
            ? F3[q] - dt * (pp[q] - pp[q - sz]) / (C1(g.zc, k) - C1(g.zc, k - 1)) : nan_d();
}

// ---------------------------------------------------------------------------
// Outer boundary layers.  <- src/ibm3d_fracstep.c:855-1906
// One launch per axis (x, then y, then z: the y and z passes read corner
// values the previous passes wrote). Tangential ranges as in the reference:
// x faces (j,k) interior; y faces i padded, k interior; z faces i, j padded.
// mode: 0 = velocity fields (3), 1 = pressure, 2 = solution fill of the
// momentum system (BC rows, component `comp`), 3 = solution fill of the
// pressure-correction system.
// ---------------------------------------------------------------------------
__device__ inline void face_pos(int n, int hi, int &o1, int &o2, int &q0, int &q1) {
    if (hi) { o1 = n; o2 = n + 1; q0 = n - 1; q1 = n - 2; }
    else { o1 = -1; o2 = -2; q0 = 0; q1 = 1; }
}

__device__ inline double bcval(const FaceBC &f, int comp, int t1, int t2, int n2) {
    if (f.garr) return f.garr[((long)comp * f.n1 + (t1 + f.t1off)) * n2 + (t2 + f.t2off)];
    return f.g[comp];
}

__global__ void k_outer(Geo g, OuterParams op, int ax, double *f1, double *f2, double *f3, int mode, int comp) {
    const Dims &d = g.d;
    const int hi = blockIdx.z;
    int a0, a1, b0, b1;
    if (ax == 0) { a0 = 0; a1 = d.ny; b0 = 0; b1 = d.nz; }
    else if (ax == 1) { a0 = -2; a1 = d.nx + 2; b0 = 0; b1 = d.nz; }
    else { a0 = -2; a1 = d.nx + 2; b0 = -2; b1 = d.ny + 2; }
    const int t1 = a0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int t2 = b0 + blockIdx.y * blockDim.y + threadIdx.y;
    if (t1 >= a1 || t2 >= b1) return;
    const FaceBC &F = op.face[ax][hi];
    const int N = ax == 0 ? d.nx : ax == 1 ? d.ny : d.nz;
    const long st = ax == 0 ? 1 : ax == 1 ? d.sy : d.sz;
    // linear index of position n along the normal
    long base;
    if (ax == 0) base = cidx(d, 0, t1, t2);
    else if (ax == 1) base = cidx(d, t1, 0, t2);
    else base = cidx(d, t1, t2, 0);
    int o1, o2, q0, q1;
    face_pos(N, hi, o1, o2, q0, q1);
    const double *dd = ax == 0 ? g.dx : ax == 1 ? g.dy : g.dz;
    const double *cc = ax == 0 ? g.xc : ax == 1 ? g.yc : g.zc;
    const int n2 = ax == 0 ? d.nz : ax == 1 ? d.nz : d.ny + 4;

    for (int layer = 1; layer <= 2; layer++) {
        const int o = layer == 1 ? o1 : o2, qq = layer == 1 ? q0 : q1;
        const long co = base + o * st, cq = base + qq * st;
        // Dirichlet interpolation distances (ao outer, bi inner)
        const double *dl = (mode == 1 && ax == 2 && !hi && F.type != BCT_CONVECTIVE) ? g.dx : dd;  // :1792-1800 spacing quirk
        double ao, bi;
        if (layer == 1) { ao = C1(dl, o1) / 2; bi = C1(dl, q0) / 2; }
        else { ao = C1(dl, o2) / 2 + C1(dl, o1); bi = C1(dl, q0) + C1(dl, q1) / 2; }
        // extrapolation spacings of the pressure outlet for velocity
        int n1, nn2;
        double so, si;
        if (layer == 1) {
            n1 = q0; nn2 = q1;
            so = hi ? C1(cc, o1) - C1(cc, q0) : C1(cc, q0) - C1(cc, o1);
            si = hi ? C1(cc, q0) - C1(cc, q1) : C1(cc, q1) - C1(cc, q0);
        } else {
            n1 = o1; nn2 = q0;
            so = hi ? C1(cc, o2) - C1(cc, o1) : C1(cc, o1) - C1(cc, o2);
            si = hi ? C1(cc, o1) - C1(cc, q0) : C1(cc, q0) - C1(cc, o1);
        }
        const long cn1 = base + n1 * st, cn2 = base + nn2 * st;
        const long cw = base + (hi ? o - N : o + N) * st;  // periodic image
        if (mode == 0 || mode == 2) {
            double *ff[3] = {f1, f2, f3};
            const int l0 = mode == 0 ? 0 : comp, l1 = mode == 0 ? 3 : comp + 1;
            for (int l = l0; l < l1; l++) {
                double *a = ff[mode == 0 ? l : 0];
                switch (F.type) {
                case BCT_VELOCITY:
                case BCT_CONVECTIVE: {
                    const double gv = bcval(F, l, t1, t2, n2);
                    a[co] = (ao + bi) / bi * gv - ao / bi * a[cq];
                    break;
                }
                case BCT_PRESSURE:
                    a[co] = (so + si) / si * a[cn1] - so / si * a[cn2];
                    break;
                case BCT_FREE_SLIP:
                    a[co] = (l == ax) ? -ao / bi * a[cq] : a[cq];
                    break;
                case BCT_ALL_PERIODIC:
                case BCT_VELOCITY_PERIODIC:
                    a[co] = a[cw];
                    break;
                default:;
                }
            }
        } else {
            double *a = f1;
            switch (F.type) {
            case BCT_VELOCITY:
            case BCT_FREE_SLIP:
                a[co] = a[cq];
                break;
            case BCT_PRESSURE:
            case BCT_VELOCITY_PERIODIC:
            case BCT_CONVECTIVE: {
                const double gv = bcval(F, 3, t1, t2, n2);
                a[co] = (ao + bi) / bi * gv - ao / bi * a[cq];
                break;
            }
            case BCT_ALL_PERIODIC:
                a[co] = a[cw];
                break;
            default:;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Convective outflow (extension; no counterpart in the reference). The
// face's convection speed comes from one workgroup (fixed-order sums); the
// update of the face values runs one thread per face entry over many
// workgroups, each leaving fixed-order partial sums of its BC-row norms.
// ---------------------------------------------------------------------------
static constexpr int kConvThreads = 1024;

// fixed-order tree sum over the workgroup; every thread gets the result
__device__ inline double conv_block_sum(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = kConvThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

struct ConvFace {
    int a0, a1, b0, b1, n2, q0;
    long ext, st;
    double bi;
};

__device__ inline ConvFace conv_face(const Geo &g, int ax, int hi) {
    const Dims &d = g.d;
    ConvFace f;
    if (ax == 0) { f.a0 = 0; f.a1 = d.ny; f.b0 = 0; f.b1 = d.nz; }
    else if (ax == 1) { f.a0 = -2; f.a1 = d.nx + 2; f.b0 = 0; f.b1 = d.nz; }
    else { f.a0 = -2; f.a1 = d.nx + 2; f.b0 = -2; f.b1 = d.ny + 2; }
    f.n2 = f.b1 - f.b0;
    f.ext = (long)(f.a1 - f.a0) * f.n2;
    const int N = ax == 0 ? d.nx : ax == 1 ? d.ny : d.nz;
    f.q0 = hi ? N - 1 : 0;
    f.st = ax == 0 ? 1 : ax == 1 ? d.sy : d.sz;
    f.bi = C1(ax == 0 ? g.dx : ax == 1 ? g.dy : g.dz, f.q0) / 2;
    return f;
}

// uc_out = the set speed uc, or for uc <= 0 the area-weighted mean outward
// normal component of the face values over the face's interior cells,
// clamped at 0
__global__ void __launch_bounds__(kConvThreads)
k_conv_speed(Geo g, int ax, int hi, double uc, const double *garr, double *uc_out) {
    __shared__ double red[kConvThreads];
    const Dims &d = g.d;
    if (uc <= 0) {
        const ConvFace f = conv_face(g, ax, hi);
        const int m1 = ax == 0 ? d.ny : d.nx, m2 = ax == 2 ? d.ny : d.nz;
        const double *w1 = ax == 0 ? g.dy : g.dx, *w2 = ax == 2 ? g.dy : g.dz;
        double num = 0, den = 0;
        for (long e = threadIdx.x; e < (long)m1 * m2; e += kConvThreads) {
            const int t1 = (int)(e / m2), t2 = (int)(e % m2);
            const double A = C1(w1, t1) * C1(w2, t2);
            num += garr[ax * f.ext + (long)(t1 - f.a0) * f.n2 + (t2 - f.b0)] * A;
            den += A;
        }
        num = conv_block_sum(num, red);
        den = conv_block_sum(den, red);
        uc = fmax((hi ? num : -num) / den, 0.0);
    }
    if (threadIdx.x == 0) *uc_out = uc;
}

// init: g = the first interior cell's velocity; otherwise one explicit
// upwind step g <- g - dt uc (g - u_q0) / (d_q0 / 2). One thread per face
// entry; partial[blk * 3 + c] = the block's sum of 2 g_c^2.
__global__ void __launch_bounds__(256)
k_conv_update(Geo g, int ax, int hi, const double *uc_in, double dt, int init, const double *u1, const double *u2,
              const double *u3, double *garr, double *partial) {
    const Dims &d = g.d;
    const ConvFace f = conv_face(g, ax, hi);
    const long e = blockIdx.x * 256L + threadIdx.x;
    double nrm[3] = {0, 0, 0};
    if (e < f.ext) {
        const double uc = init ? 0.0 : *uc_in;
        const int t1 = f.a0 + (int)(e / f.n2), t2 = f.b0 + (int)(e % f.n2);
        long base;
        if (ax == 0) base = cidx(d, 0, t1, t2);
        else if (ax == 1) base = cidx(d, t1, 0, t2);
        else base = cidx(d, t1, t2, 0);
        const long cq = base + f.q0 * f.st;
        const double *u[3] = {u1, u2, u3};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double uq = u[c][cq];
            double gv = garr[c * f.ext + e];
            gv = init ? uq : gv - dt * uc * (gv - uq) / f.bi;
            garr[c * f.ext + e] = gv;
            nrm[c] = 2 * gv * gv;
        }
    }
    block_sum_store<3>(nrm, partial, blockIdx.x);
}

long conv_outflow_blocks(const Dims &d, int ax) {
    const long ext = ax == 0 ? (long)d.ny * d.nz : ax == 1 ? (long)(d.nx + 4) * d.nz : (long)(d.nx + 4) * (d.ny + 4);
    return (ext + 255) / 256;
}

long launch_conv_outflow(hipStream_t st, const Geo &g, int ax, int hi, double uc, double dt, int init,
                         const double *u1, const double *u2, const double *u3, double *garr, double *uc_slot,
                         double *partial) {
    if (!init) hipLaunchKernelGGL(k_conv_speed, dim3(1), dim3(kConvThreads), 0, st, g, ax, hi, uc, garr, uc_slot);
    const long nblk = conv_outflow_blocks(g.d, ax);
    hipLaunchKernelGGL(k_conv_update, dim3(nblk), dim3(256), 0, st, g, ax, hi, uc_slot, dt, init, u1, u2, u3, garr,
                       partial);
    return nblk;
}

// ---------------------------------------------------------------------------
// Inhomogeneous boundary data folded into the right-hand side of the first
// interior cell: b += K_face * gamma * g (the outer row x_o = alpha x_q +
// gamma g eliminated from the fluid row). The BC rows' own right-hand side
// enters the stopping norm on the host.
// ---------------------------------------------------------------------------
__global__ void k_bc_rhs(Geo g, OuterParams op, int ax, int hi, int comp, const double *Kf, double gamma,
                         double *b) {
    const Dims &d = g.d;
    int a1, b1;
    if (ax == 0) { a1 = d.ny; b1 = d.nz; }
    else if (ax == 1) { a1 = d.nx; b1 = d.nz; }
    else { a1 = d.nx; b1 = d.ny; }
    const int t1 = blockIdx.x * blockDim.x + threadIdx.x;
    const int t2 = blockIdx.y * blockDim.y + threadIdx.y;
    if (t1 >= a1 || t2 >= b1) return;
    const FaceBC &F = op.face[ax][hi];
    const int N = ax == 0 ? d.nx : ax == 1 ? d.ny : d.nz;
    const int n = hi ? N - 1 : 0;
    long q;
    if (ax == 0) q = cidx(d, n, t1, t2);
    else if (ax == 1) q = cidx(d, t1, n, t2);
    else q = cidx(d, t1, t2, n);
    const int n2 = ax == 0 ? d.nz : ax == 1 ? d.nz : d.ny + 4;
    b[q] += C1(Kf, n) * gamma * bcval(F, comp, t1, t2, n2);
}

// ---------------------------------------------------------------------------
// Immersed-boundary ghost cells, ordered in-place update.
// <- src/ibm3d_fracstep.c:1933-1975. Ghost cells are grouped by dependency
// level; a source that is a ghost cell visited later in the reference's
// lexicographic order reads the value saved before the pass (old[]), one
// visited earlier reads the updated value (its level is lower).
// ---------------------------------------------------------------------------
__global__ void k_ghost_save(const GhostCell *gc, int ng, const double *u1, const double *u2,
                             const double *u3, const double *p, double *old) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ng) return;
    const long c = gc[s].cell;
    old[4 * (long)s + 0] = u1[c];
    old[4 * (long)s + 1] = u2[c];
    old[4 * (long)s + 2] = u3[c];
    old[4 * (long)s + 3] = p[c];
}

__global__ void k_ghost_level(const GhostCell *gc, int s0, int s1, const double *old,
                              double *u1, double *u2, double *u3, double *p) {
    const int s = s0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= s1) return;
    const GhostCell &G = gc[s];
    double a1 = 0, a2 = 0, a3 = 0, ap = 0;
#pragma unroll
    for (int l = 0; l < 8; l++) {
        const double w = G.w[l];
        if (G.src[l] < 0) continue;
        double v1, v2, v3, vp;
        if (G.later[l] >= 0) {
            const long o = 4 * (long)G.later[l];
            v1 = old[o]; v2 = old[o + 1]; v3 = old[o + 2]; vp = old[o + 3];
        } else {
            const long c = G.src[l];
            v1 = u1[c]; v2 = u2[c]; v3 = u3[c]; vp = p[c];
        }
        a1 += w * v1; a2 += w * v2; a3 += w * v3; ap += w * vp;
    }
    const double lu = 1 + G.wself, lp = 1 - G.wself;
    u1[G.cell] = -a1 / lu;
    u2[G.cell] = -a2 / lu;
    u3[G.cell] = -a3 / lu;
    if (lp != 0) p[G.cell] = ap / lp;
}

// Interior solid cells hold NaN in u and p (:1972-1973).
__global__ void k_solid_nan(Dims d, const uint8_t *flag, double *u1, double *u2, double *u3, double *p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    if (flag[q] == F_SOLID) u1[q] = u2[q] = u3[q] = p[q] = nan_d();
}

// ---------------------------------------------------------------------------
// SGS stress.  <- src/ibm3d_turb.c:8-82 (Smagorinsky, l = min(Cs Delta,
// 0.41 phi), Delta = 2 cbrt(dx dy dz)); Vreman (extension) with
// nu_t = c sqrt(B / (alpha:alpha)). tau = -2 nu_t S on fluid cells only.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_sgs(Geo g, int model, double coef, VanDriest vd, const uint8_t *flag,
                                             const double *lvset, const double *u1, const double *u2,
                                             const double *u3, TauPtrs tau) {
    const Dims &d = g.d;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    if (flag[q] != F_FLUID) return;
    const long sy = d.sy, sz = d.sz;
    const double ddx = C1(g.xc, i + 1) - C1(g.xc, i - 1);
    const double ddy = C1(g.yc, j + 1) - C1(g.yc, j - 1);
    const double ddz = C1(g.zc, k + 1) - C1(g.zc, k - 1);
    const double *uu[3] = {u1, u2, u3};
    double al[3][3];  // al[m][n] = d u_n / d x_m
#pragma unroll
    for (int n = 0; n < 3; n++) {
        al[0][n] = (uu[n][q + 1] - uu[n][q - 1]) / ddx;
        al[1][n] = (uu[n][q + sy] - uu[n][q - sy]) / ddy;
        al[2][n] = (uu[n][q + sz] - uu[n][q - sz]) / ddz;
    }
    // S in the reference's slot order 11, 12, 13, 22, 23, 33 (src/ibm3d_turb.c:60-81)
    double S[6];
    S[0] = al[0][0];
    S[1] = 0.5 * (al[1][0] + al[0][1]);
    S[2] = 0.5 * (al[2][0] + al[0][2]);
    S[3] = al[1][1];
    S[4] = 0.5 * (al[2][1] + al[1][2]);
    S[5] = al[2][2];
    double nut;
    if (model == 1) {
        const double delta = 2 * cbrt(C1(g.dx, i) * C1(g.dy, j) * C1(g.dz, k));
        // sum over all (a, b) of 2 S_ab^2, in the reference's loop order
        const double s2 = 2 * S[0] * S[0] + 2 * S[1] * S[1] + 2 * S[2] * S[2] + 2 * S[1] * S[1]
            + 2 * S[3] * S[3] + 2 * S[4] * S[4] + 2 * S[2] * S[2] + 2 * S[4] * S[4] + 2 * S[5] * S[5];
        const double smag = sqrt(s2);
        double l = coef * delta;
        if (vd.walls) {
            // van Driest: l *= 1 - exp(-y+ / A+), y+ = y_w u_tau / nu
            const double c[3] = {C1(g.xc, i), C1(g.yc, j), C1(g.zc, k)};
            const int bit_lo[3] = {8, 4, 16}, bit_hi[3] = {2, 1, 32};  // W S D, E N U
            double yw = 1e300;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if (vd.walls & bit_lo[a]) yw = fmin(yw, c[a] - vd.lo[a]);
                if (vd.walls & bit_hi[a]) yw = fmin(yw, vd.hi[a] - c[a]);
            }
            l *= 1 - exp(-(yw * vd.ure) / vd.aplus);
        }
        const double lw = 0.41 * lvset[q];
        if (lw < l) l = lw;
        const double f = -2 * (l * l) * smag;
#pragma unroll
        for (int a = 0; a < 6; a++) tau.t[a][q] = f * S[a];
        return;
    }
    const double del[3] = {C1(g.dx, i), C1(g.dy, j), C1(g.dz, k)};
    double aa = 0, bt[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            aa += al[a][b] * al[a][b];
            double acc = 0;
#pragma unroll
            for (int m = 0; m < 3; m++) acc += del[m] * del[m] * al[m][a] * al[m][b];
            bt[a][b] = acc;
        }
    const double Bb = bt[0][0] * bt[1][1] - bt[0][1] * bt[0][1] + bt[0][0] * bt[2][2] - bt[0][2] * bt[0][2]
        + bt[1][1] * bt[2][2] - bt[1][2] * bt[1][2];
    nut = (aa > 1e-300 && Bb > 0) ? coef * sqrt(Bb / aa) : 0;
#pragma unroll
    for (int a = 0; a < 6; a++) tau.t[a][q] = -2 * nut * S[a];
}

// max |div U| over fluid cells (partial maxima per block), the pinned cell
// (whose pressure row is the pin, not continuity) left out
__global__ void __launch_bounds__(256) k_maxdiv(Geo g, const uint8_t *flag, long pin, const double *U1,
                                                const double *U2, const double *U3, double *partial) {
    const Dims &d = g.d;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    double m = 0;
    if (i < d.nx && j < d.ny) {
        const long q = cidx(d, i, j, k);
        if (flag[q] == F_FLUID && q != pin) {
            const double v = (U1[q + 1] - U1[q]) / C1(g.dx, i) + (U2[q + d.sy] - U2[q]) / C1(g.dy, j)
                + (U3[q + d.sz] - U3[q]) / C1(g.dz, k);
            m = fabs(v);
        }
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    __shared__ double sh[16];
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if ((tid & 63) == 0) sh[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
        double r = 0;
        for (int w = 0; w < (int)((blockDim.x * blockDim.y + 63) / 64); w++) r = fmax(r, sh[w]);
        partial[block_linear()] = r;
    }
}

// Force on the immersed body (pressure and viscous parts, per unit density)
// from the staircase surface: the faces between an interior ghost cell and an
// interior fluid 6-neighbour, one thread per ghost cell. With n the outward
// body normal (from the ghost into the fluid cell) and A the face area:
//   pressure: F -= p_f n A, p_f linearly interpolated to the face position;
//   viscous:  F_b += mu (u_b(fluid) - u_b(ghost)) / |x_fluid - x_ghost| A for
//             the two tangential components b.
// No reference counterpart (the reference has no force diagnostic).
__global__ void __launch_bounds__(256) k_body_force(Geo g, const uint8_t *flag, const int *gijk, int ng, double mu,
                                                    const double *p, const double *u1, const double *u2,
                                                    const double *u3, double *partial) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    double f[3] = {0, 0, 0};
    if (t < ng) {
        const Dims &d = g.d;
        const int id[3] = {gijk[3 * t], gijk[3 * t + 1], gijk[3 * t + 2]};
        const long q = cidx(d, id[0], id[1], id[2]);
        const int n[3] = {d.nx, d.ny, d.nz};
        const long st[3] = {1, d.sy, d.sz};
        const double *C[3] = {g.xc, g.yc, g.zc}, *D[3] = {g.dx, g.dy, g.dz};
        const double *u[3] = {u1, u2, u3};
        const double h[3] = {C1(g.dx, id[0]), C1(g.dy, id[1]), C1(g.dz, id[2])};
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int sd = -1; sd <= 1; sd += 2) {
                const int nb = id[a] + sd;
                if (nb < 0 || nb >= n[a]) continue;
                const long r = q + sd * st[a];
                if (flag[r] != F_FLUID) continue;
                const double area = h[0] * h[1] * h[2] / h[a];
                const double xg = C1(C[a], id[a]), xf = C1(C[a], nb);
                const double xface = xg + sd * 0.5 * C1(D[a], id[a]);
                const double pf = p[q] + (p[r] - p[q]) * ((xface - xg) / (xf - xg));
                f[a] -= pf * sd * area;
                const double inv = 1.0 / fabs(xf - xg);
#pragma unroll
                for (int b = 0; b < 3; b++)
                    if (b != a) f[b] += mu * (u[b][r] - u[b][q]) * inv * area;
            }
    }
    block_sum_store<3>(f, partial, blockIdx.x);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static dim3 cell_grid(const Dims &d, int ex) {
    return dim3((d.nx + ex + 63) / 64, (d.ny + ex + 3) / 4, d.nz + ex);
}

void launch_interp_stag(hipStream_t st, const Geo &g, const double *u1, const double *u2, const double *u3,
                        double *U1, double *U2, double *U3) {
    dim3 grid((g.d.nx + 4 + 63) / 64, (g.d.ny + 4 + 3) / 4, g.d.nz + 4);
    hipLaunchKernelGGL(k_interp_stag, grid, dim3(64, 4), 0, st, g, u1, u2, u3, U1, U2, U3);
}

long conv_rhs_blocks(const Dims &d) {
    dim3 gr = cell_grid(d, 0);
    return (long)gr.x * gr.y * gr.z;
}

void launch_conv_rhs(hipStream_t st, const Geo &g, const StepParams &sp, const uint8_t *flag,
                     const double *const u[3], const double *const U[3], const double *p,
                     const double *const Np[3], double *const N[3], double *const b[3], const TauPtrs &tau,
                     double *partial) {
    hipLaunchKernelGGL(k_conv_rhs, cell_grid(g.d, 0), dim3(64, 4), 0, st, g, sp, flag, u[0], u[1], u[2],
                       U[0], U[1], U[2], p, Np[0], Np[1], Np[2], N[0], N[1], N[2], b[0], b[1], b[2], tau,
                       partial);
}

void launch_face_star(hipStream_t st, const Geo &g, const StepParams &sp, const uint8_t *flag,
                      const double *const us[3], const double *p, double *const F[3]) {
    hipLaunchKernelGGL(k_face_star, cell_grid(g.d, 1), dim3(64, 4), 0, st, g, sp, flag, us[0], us[1], us[2],
                       p, F[0], F[1], F[2]);
}

void launch_poisson_rhs(hipStream_t st, const Geo &g, const StepParams &sp, const uint8_t *flag, long pin,
                        const double *const F[3], double *b, double *partial) {
    hipLaunchKernelGGL(k_poisson_rhs, cell_grid(g.d, 0), dim3(64, 4), 0, st, g, sp, flag, pin, F[0], F[1], F[2],
                       b, partial);
}

void launch_update(hipStream_t st, const Geo &g, const StepParams &sp, const uint8_t *flag, const double *p,
                   const double *pp, const double *const us[3], const double *const F[3], double *pn,
                   double *const un[3], double *const G[3]) {
    hipLaunchKernelGGL(k_update, cell_grid(g.d, 1), dim3(64, 4), 0, st, g, sp, flag, p, pp, us[0], us[1], us[2],
                       F[0], F[1], F[2], pn, un[0], un[1], un[2], G[0], G[1], G[2]);
}

void launch_outer(hipStream_t st, const Geo &g, const OuterParams &op, double *f1, double *f2, double *f3,
                  int mode, int comp) {
    const Dims &d = g.d;
    for (int ax = 0; ax < 3; ax++) {
        int n1, n2;
        if (ax == 0) { n1 = d.ny; n2 = d.nz; }
        else if (ax == 1) { n1 = d.nx + 4; n2 = d.nz; }
        else { n1 = d.nx + 4; n2 = d.ny + 4; }
        dim3 grid((n1 + 15) / 16, (n2 + 15) / 16, 2);
        hipLaunchKernelGGL(k_outer, grid, dim3(16, 16), 0, st, g, op, ax, f1, f2, f3, mode, comp);
    }
}

void launch_bc_rhs(hipStream_t st, const Geo &g, const OuterParams &op, int ax, int hi, int comp,
                   const double *Kf, double gamma, double *b) {
    const Dims &d = g.d;
    int a1, b1;
    if (ax == 0) { a1 = d.ny; b1 = d.nz; }
    else if (ax == 1) { a1 = d.nx; b1 = d.nz; }
    else { a1 = d.nx; b1 = d.ny; }
    hipLaunchKernelGGL(k_bc_rhs, dim3((a1 + 15) / 16, (b1 + 15) / 16), dim3(16, 16), 0, st, g, op, ax, hi, comp,
                       Kf, gamma, b);
}

void launch_ghost_update(hipStream_t st, const GhostCell *gc, int ng, const int *level_start, int nlevels,
                         double *old, double *u1, double *u2, double *u3, double *p) {
    if (ng == 0) return;
    hipLaunchKernelGGL(k_ghost_save, dim3((ng + 255) / 256), dim3(256), 0, st, gc, ng, u1, u2, u3, p, old);
    for (int l = 0; l < nlevels; l++) {
        const int s0 = level_start[l], s1 = level_start[l + 1];
        if (s1 > s0)
            hipLaunchKernelGGL(k_ghost_level, dim3((s1 - s0 + 255) / 256), dim3(256), 0, st, gc, s0, s1, old,
                               u1, u2, u3, p);
    }
}

void launch_solid_nan(hipStream_t st, const Dims &d, const uint8_t *flag, double *u1, double *u2, double *u3,
                      double *p) {
    hipLaunchKernelGGL(k_solid_nan, cell_grid(d, 0), dim3(64, 4), 0, st, d, flag, u1, u2, u3, p);
}

void launch_sgs(hipStream_t st, const Geo &g, int model, double coef, const VanDriest &vd, const uint8_t *flag,
                const double *lvset, const double *const u[3], const TauPtrs &tau) {
    hipLaunchKernelGGL(k_sgs, cell_grid(g.d, 0), dim3(64, 4), 0, st, g, model, coef, vd, flag, lvset, u[0], u[1],
                       u[2], tau);
}

long maxdiv_blocks(const Dims &d) { return conv_rhs_blocks(d); }

long body_force_blocks(int ng) { return (ng + 255) / 256 > 0 ? (ng + 255) / 256 : 1; }

void launch_body_force(hipStream_t st, const Geo &g, const uint8_t *flag, const int *gijk, int ng, double mu,
                       const double *p, const double *const u[3], double *partial) {
    hipLaunchKernelGGL(k_body_force, dim3(body_force_blocks(ng)), dim3(256), 0, st, g, flag, gijk, ng, mu, p, u[0],
                       u[1], u[2], partial);
}

void launch_maxdiv(hipStream_t st, const Geo &g, const uint8_t *flag, long pin, const double *const U[3],
                   double *partial) {
    hipLaunchKernelGGL(k_maxdiv, cell_grid(g.d, 0), dim3(64, 4), 0, st, g, flag, pin, U[0], U[1], U[2], partial);
}

}  // namespace ibm
