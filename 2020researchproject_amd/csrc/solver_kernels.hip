// solver_kernels.hip - linear-algebra kernels of the momentum and pressure
// solves (the part the reference delegates to Hypre: src/ibm3d_fracstep.c:
// 354-379 and 649-703, matrices of src/ibm3d_setup.c:1566-2908).
//
// The operator is matrix-free: fluid rows are the 7-point stencil with
// separable per-axis coefficients (the reference's kx_W .. kz_U, with the
// outer boundary rows folded into the first/last interior cells), ghost rows
// come from a compact list, solid and pinned rows are the identity. Vectors
// live in the padded cell layout of common.h and only interior cells are
// touched. Reductions are deterministic: per-block partial sums in a fixed
// order, then a one-block finalize kernel.
#include "solver_kernels.h"

#include <algorithm>

#include "rbgs_fused.h"
#include "rbgs_quad.h"

namespace ibm {

#define C1(a, i) ((a)[(i) + 2])

// Workgroups of a launch go to the 8 XCDs round-robin by linear workgroup id
// (x fastest), and each XCD has its own L2. xcd_block() maps the hardware id
// to a logical block so that each XCD runs one contiguous range of logical
// blocks (x fastest, then y, then z): the planes k-1, k, k+1 a stencil sweep
// reads are then fetched by one XCD and reused from its L2, whatever the
// number of blocks per plane.
constexpr unsigned kNumXcd = 8;
// Logical block (x fastest, then y, then z, on a gx x gy x - grid of n
// blocks) of the block with linear id w, where XCD w % 8 runs it.
__device__ inline uint3 xcd_map(unsigned w, unsigned n, unsigned gx, unsigned gy) {
    const unsigned x = w % kNumXcd, s = w / kNumXcd, q = n / kNumXcd, r = n % kNumXcd;
    const unsigned L = x * q + (x < r ? x : r) + s;
    return make_uint3(L % gx, (L / gx) % gy, L / (gx * gy));
}
__device__ inline uint3 xcd_block(int on) {
    if (!on) return make_uint3(blockIdx.x, blockIdx.y, blockIdx.z);
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned w = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    return xcd_map(w, gx * gy * gridDim.z, gx, gy);
}

__device__ inline long blin() {
    return blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
}

template <int NV>
__device__ inline void bsum(const double *v, double *partial, long blk) {
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

static inline dim3 igrid(const Dims &d) { return dim3((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz); }
// Kernels that reduce walk kRedZ planes per block, so there are kRedZ times
// fewer block partials for the finalize kernel to sum.
constexpr int kRedZ = 4;
static inline dim3 rgrid(const Dims &d) { return dim3((d.nx + 63) / 64, (d.ny + 3) / 4, (d.nz + kRedZ - 1) / kRedZ); }
// Operator application and residual: two cells of a row per lane and kRedZP
// planes per block (as many block partials as kRedZ planes of one cell per
// lane; measured 25 us against 29 us per FP64 launch at 256 x 128 x 128
// with four planes, whose registers halve the occupancy).
constexpr int kRedZP = 2;
static inline long pblocks(const Dims &d) {
    return (long)(((d.nx + 1) / 2 + 63) / 64) * ((d.ny + 3) / 4) * ((d.nz + kRedZP - 1) / kRedZP);
}

template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
    using type = float2;
};
template <>
struct Vec2<double> {
    using type = double2;
};

// A x on cells (i0, j, k) and (i0 + 1, j, k), i0 even: loads of both cells'
// neighbours as 2-vectors (x row, the rows j -/+ 1 and the planes k -/+ 1),
// one cell-sized load on each side of the pair, flags and face masks as one
// 16-bit load. Writes the rows' diagonal-weighted terms: yv[o] = diag x -
// off, xq[o] = x, for o = 0, 1. Same arithmetic per cell as stencil(). x may
// be stored in a narrower type TX (values converted exactly on load).
template <typename T, typename TX>
__device__ inline typename Vec2<T>::type load_pair(const TX *p) {
    const typename Vec2<TX>::type v = *reinterpret_cast<const typename Vec2<TX>::type *>(p);
    typename Vec2<T>::type o;
    o.x = (T)v.x;
    o.y = (T)v.y;
    return o;
}

template <typename T, typename TX = T>
__device__ inline void pair_apply(const Dims &d, const SysCoef<T> &S, const TX *x, int i0, int j, int k, long q0,
                                  T yv[2], T xq[2]) {
    using T2 = typename Vec2<T>::type;
    const AxisCoef<T> &AX = S.ax[0], &AY = S.ax[1], &AZ = S.ax[2];
    const int jm = (j == 0 && AY.periodic) ? AY.n - 1 : j - 1, jp = (j == AY.n - 1 && AY.periodic) ? 0 : j + 1;
    const int km = (k == 0 && AZ.periodic) ? AZ.n - 1 : k - 1, kp = (k == AZ.n - 1 && AZ.periodic) ? 0 : k + 1;
    const T2 xc = load_pair<T>(x + q0);
    const T2 ym = load_pair<T>(x + q0 + (long)(jm - j) * d.sy);
    const T2 yp = load_pair<T>(x + q0 + (long)(jp - j) * d.sy);
    const T2 zm = load_pair<T>(x + q0 + (long)(km - k) * d.sz);
    const T2 zp = load_pair<T>(x + q0 + (long)(kp - k) * d.sz);
    T xl = (T)x[q0 - 1], xr = (T)x[q0 + 2];
    if (AX.periodic) {
        if (i0 == 0) xl = (T)x[q0 + AX.n - 1];
        if (i0 + 1 == AX.n - 1) xr = (T)x[q0 - i0];
    }
    const unsigned mm = S.skip ? *reinterpret_cast<const unsigned short *>(S.fmask + q0) : 0u;
    using P4 = typename Packed4<T>::type;
    const P4 cx0 = reinterpret_cast<const P4 *>(AX.c4)[i0];
    const P4 cx1 = reinterpret_cast<const P4 *>(AX.c4)[i0 + 1 < d.nx ? i0 + 1 : i0];
    const P4 cy = reinterpret_cast<const P4 *>(AY.c4)[j];
    const P4 cz = reinterpret_cast<const P4 *>(AZ.c4)[k];
#pragma unroll
    for (int o = 0; o < 2; o++) {
        const unsigned m = (mm >> (8 * o)) & 0xFFu;
        // periodic x, cell i0 + 1 the last of its row: the wrap neighbour
        T xp = o ? xr : xc.y;
        if (o == 0 && AX.periodic && i0 == AX.n - 1) xp = (T)x[q0 - i0];
        const T xm = o ? xc.x : xl;
        const P4 cx = o ? cx1 : cx0;
        T diag = S.c0, off = 0;
        axis_term4(cx.x, cx.y, cx.z, cx.w, (m & 1u) != 0, (m & 2u) != 0, xm, xp, diag, off);
        axis_term4(cy.x, cy.y, cy.z, cy.w, (m & 4u) != 0, (m & 8u) != 0, o ? ym.y : ym.x, o ? yp.y : yp.x, diag, off);
        axis_term4(cz.x, cz.y, cz.z, cz.w, (m & 16u) != 0, (m & 32u) != 0, o ? zm.y : zm.x, o ? zp.y : zp.x, diag, off);
        xq[o] = o ? xc.y : xc.x;
        yv[o] = diag * xq[o] - off;
    }
}
static inline long iblocks(const Dims &d) {
    return (long)((d.nx + 63) / 64) * ((d.ny + 3) / 4) * ((d.nz + kRedZ - 1) / kRedZ);
}

// Fluid-row stencil: returns diag and the off-diagonal sum sum(c x_nb).
template <typename T>
__device__ inline void stencil(const Dims &d, const SysCoef<T> &S, const uint8_t *flag, int i, int j, int k,
                               long q, const T *x, T &diag, T &off) {
    diag = S.c0;
    off = 0;
    const int id[3] = {i, j, k};
    const long st[3] = {1, (long)d.sy, d.sz};
    // neighbour loads first: the face-mask byte (coarse levels) is the last
    // load issued, so waiting for it does not hold back the others
    T xm[3], xp[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const AxisCoef<T> &A = S.ax[a];
        const int n = A.n, ii = id[a];
        const long qm = (ii == 0 && A.periodic) ? q + (long)(n - 1) * st[a] : q - st[a];
        const long qp = (ii == n - 1 && A.periodic) ? q - (long)(n - 1) * st[a] : q + st[a];
        xm[a] = x[qm];
        xp[a] = x[qp];
    }
    const unsigned m = S.skip ? S.fmask[q] : 0u;
#pragma unroll
    for (int a = 0; a < 3; a++)
        axis_term(S.ax[a], id[a], ((m >> (2 * a)) & 1u) != 0, ((m >> (2 * a + 1)) & 1u) != 0, xm[a], xp[a], diag, off);
}

// diagonal of a fluid row (the stencil with every neighbour value zero)
template <typename T>
__device__ inline T stencil_diag(const SysCoef<T> &S, int i, int j, int k, long q) {
    T diag = S.c0, off = 0;
    const int id[3] = {i, j, k};
    const unsigned m = S.skip ? S.fmask[q] : 0u;
#pragma unroll
    for (int a = 0; a < 3; a++)
        axis_term(S.ax[a], id[a], ((m >> (2 * a)) & 1u) != 0, ((m >> (2 * a + 1)) & 1u) != 0, T(0), T(0), diag, off);
    return diag;
}

// First colour of a red-black sweep from a zero start (colour 0: cells with
// i + j + k even), the value the one-colour smoother writes at (i, j, k):
// same arithmetic (the neighbour sum is 0), so a kernel that produces a
// level's right-hand side can also write this colour and save that launch.
// The diagonal (stencil_diag) is formed beforehand, so its loads can issue
// ahead of those that produce bv.
template <typename T>
__device__ inline T first_colour_of(const SysCoef<T> &S, unsigned f, int i, int j, int k, long q, T bv, T diag) {
    if ((i + j + k) & 1) return T(0);
    if (q == S.pin || f == F_SOLID) return bv;
    if (f != F_FLUID) return T(0);
    return sor_value(S.omega, T(0), diag != T(0) ? (bv + T(0)) / diag : T(0));
}

// ---------------------------------------------------------------------------
// y = A x (interior non-ghost rows); mode 1 adds dot(w, y), mode 2 adds
// dot(y, w) and dot(y, y).
// ---------------------------------------------------------------------------
// The kernels below compute the stencil of every cell before looking at its
// flag (solid and ghost cells have in-bounds neighbours too; their results
// are discarded), so all loads of a thread issue before the first wait.
// Ghost rows run inside the launches of the operator, residual and Jacobi
// kernels: the first blocks of a 1-D grid (a multiple of 8, so they spread
// over the XCDs and the other blocks keep their XCD assignment) each take 64
// rows in their first wave; the other blocks do the fluid, solid and pinned
// rows and leave the ghost cells' outputs to them.
static inline unsigned ghost_blocks(int ng) { return ng > 0 ? (unsigned)(((ng + 63) / 64 + 7) / 8 * 8) : 0u; }

// stores of a cell pair that skip ghost cells (g0, g1)
template <typename T>
__device__ inline void store_pair_skip(const Dims &d, int i0, long q0, T *a, T v0, T v1, bool g0, bool g1) {
    using T2 = typename Vec2<T>::type;
    if (i0 + 1 < d.nx && !g0 && !g1) {
        T2 o;
        o.x = v0;
        o.y = v1;
        *reinterpret_cast<T2 *>(a + q0) = o;
    } else {
        if (!g0) a[q0] = v0;
        if (i0 + 1 < d.nx && !g1) a[q0 + 1] = v1;
    }
}

// y at ghost row s: is_p ? (1 - ws) x - sum w x : (1 + ws) x + sum w x
template <typename T, typename TX>
__device__ inline T ghost_row_apply(const GhostRows<T> &G, int is_p, const TX *x, int s, long &c) {
    T a = 0;
#pragma unroll
    for (int l = 0; l < 8; l++) {
        const long q = G.src[8 * (long)s + l];
        if (q >= 0) a += G.w[8 * (long)s + l] * (T)x[q];
    }
    c = G.cell[s];
    const T ws = G.wself[s], xc = (T)x[c];
    return is_p ? (T(1) - ws) * xc - a : (T(1) + ws) * xc + a;
}

// y = A x; TX: storage type of x; xo, when not null, receives x in type T
// (the operand of a later vector update).
template <typename T, int MODE, typename TX = T>
__global__ void __launch_bounds__(256) k_apply(Dims d, SysCoef<T> S, GhostRows<T> G, const uint8_t *__restrict__ flag,
                                               const TX *__restrict__ x, T *__restrict__ y, const T *__restrict__ w,
                                               double *partial, T *__restrict__ xo, unsigned ngb, unsigned gx,
                                               unsigned gy) {
    using T2 = typename Vec2<T>::type;
    const unsigned wb = blockIdx.x;
    double acc[2] = {0, 0};
    if (wb < ngb) {
        const int s = (int)wb * 64 + threadIdx.x;
        if (threadIdx.y == 0 && s < G.ng) {
            long c;
            const T yv = ghost_row_apply<T>(G, S.is_p, x, s, c);
            y[c] = yv;
            if (MODE == 1) acc[0] = (double)w[c] * (double)yv;
            if (MODE == 2) { acc[0] = (double)yv * (double)w[c]; acc[1] = (double)yv * (double)yv; }
        }
    } else {
        const uint3 bk = xcd_map(wb - ngb, gridDim.x - ngb, gx, gy);
        const int i0 = 2 * (bk.x * 64 + threadIdx.x), j = bk.y * 4 + threadIdx.y;
        const int k0 = bk.z * kRedZP, k1 = min(k0 + kRedZP, d.nz);
        if (i0 < d.nx && j < d.ny) {
            T yv[kRedZP][2], xq[kRedZP][2];
            T2 wv[kRedZP];
            unsigned fv[kRedZP];
#pragma unroll
            for (int kk = 0; kk < kRedZP; kk++) {
                const int k = min(k0 + kk, k1 - 1);  // past the last plane: a repeat, discarded
                const long q0 = cidx(d, i0, j, k);
                fv[kk] = *reinterpret_cast<const unsigned short *>(flag + q0);
                if (MODE > 0) wv[kk] = *reinterpret_cast<const T2 *>(w + q0);
                pair_apply(d, S, x, i0, j, k, q0, yv[kk], xq[kk]);
            }
#pragma unroll
            for (int kk = 0; kk < kRedZP; kk++) {
                if (k0 + kk >= k1) break;
                const long q0 = cidx(d, i0, j, k0 + kk);
                T v[2];
                bool on[2], gh[2];
#pragma unroll
                for (int o = 0; o < 2; o++) {
                    const unsigned f = (fv[kk] >> (8 * o)) & 0xFFu;
                    const bool ident = q0 + o == S.pin || f == F_SOLID;
                    v[o] = ident ? xq[kk][o] : yv[kk][o];
                    gh[o] = !ident && f == F_GHOST;  // written by the ghost blocks
                    on[o] = i0 + o < d.nx && (ident || f == F_FLUID);
                }
                store_pair_skip(d, i0, q0, y, v[0], v[1], gh[0], gh[1]);
                if (xo) store_pair_skip(d, i0, q0, xo, xq[kk][0], xq[kk][1], false, false);
#pragma unroll
                for (int o = 0; o < 2; o++) {
                    if (!on[o]) continue;
                    const double wo = o ? (double)wv[kk].y : (double)wv[kk].x;
                    if (MODE == 1) acc[0] += wo * (double)v[o];
                    if (MODE == 2) { acc[0] += (double)v[o] * wo; acc[1] += (double)v[o] * (double)v[o]; }
                }
            }
        }
    }
    if constexpr (MODE > 0) bsum<MODE>(acc, partial, wb);
}

long apply_partials(const Dims &d, int ng) { return pblocks(d) + ghost_blocks(ng); }
// 1-D grid of the operator / residual kernels: ghost blocks + pair blocks
static inline unsigned agrid_blocks(const Dims &d, int ng) { return (unsigned)apply_partials(d, ng); }
static inline unsigned pgx(const Dims &d) { return ((d.nx + 1) / 2 + 63) / 64; }
static inline unsigned pgy(const Dims &d) { return (d.ny + 3) / 4; }

template <typename T>
void launch_apply(hipStream_t st, const Dims &d, const SysCoef<T> &S, const GhostRows<T> &G, const uint8_t *flag,
                  const T *x, T *y, int mode, const T *w, double *partial) {
    const unsigned ngb = ghost_blocks(G.ng), nb = agrid_blocks(d, G.ng);
    if (mode == 0)
        hipLaunchKernelGGL((k_apply<T, 0>), dim3(nb), dim3(64, 4), 0, st, d, S, G, flag, x, y, w, partial, (T *)nullptr,
                           ngb, pgx(d), pgy(d));
    else if (mode == 1)
        hipLaunchKernelGGL((k_apply<T, 1>), dim3(nb), dim3(64, 4), 0, st, d, S, G, flag, x, y, w, partial, (T *)nullptr,
                           ngb, pgx(d), pgy(d));
    else
        hipLaunchKernelGGL((k_apply<T, 2>), dim3(nb), dim3(64, 4), 0, st, d, S, G, flag, x, y, w, partial, (T *)nullptr,
                           ngb, pgx(d), pgy(d));
}

void launch_apply_f32in(hipStream_t st, const Dims &d, const SysCoef<double> &S, const GhostRows<double> &G,
                        const uint8_t *flag, const float *x32, double *x, double *y, int mode, const double *w,
                        double *partial) {
    const unsigned ngb = ghost_blocks(G.ng), nb = agrid_blocks(d, G.ng);
    if (mode == 1)
        hipLaunchKernelGGL((k_apply<double, 1, float>), dim3(nb), dim3(64, 4), 0, st, d, S, G, flag, x32, y, w, partial,
                           x, ngb, pgx(d), pgy(d));
    else
        hipLaunchKernelGGL((k_apply<double, 2, float>), dim3(nb), dim3(64, 4), 0, st, d, S, G, flag, x32, y, w, partial,
                           x, ngb, pgx(d), pgy(d));
}

// ---------------------------------------------------------------------------
// r = b - A x with partial norms ||r||^2 and ||D^-1 r||^2 (D = sum K on fluid
// rows when ns.on, else identity).
// ---------------------------------------------------------------------------
__device__ inline double nscale(const NormSpec &ns, const Dims &d, int i, int j, int k, long q) {
    if (!ns.on || q == ns.pin || ns.flag[q] != F_FLUID) return 1.0;
    return 1.0 / (ns.ksx[i] + ns.ksy[j] + ns.ksz[k]);
}

template <typename T>
__global__ void __launch_bounds__(256) k_residual(Dims d, SysCoef<T> S, GhostRows<T> G,
                                                  const uint8_t *__restrict__ flag, const T *__restrict__ b,
                                                  const T *__restrict__ x, T *__restrict__ r, NormSpec ns,
                                                  double *partial, unsigned ngb, unsigned gx, unsigned gy) {
    using T2 = typename Vec2<T>::type;
    const unsigned wb = blockIdx.x;
    double acc[2] = {0, 0};
    if (wb < ngb) {
        const int s = (int)wb * 64 + threadIdx.x;
        if (threadIdx.y == 0 && s < G.ng) {
            long c;
            const T yv = ghost_row_apply<T>(G, S.is_p, x, s, c);
            const T rv = b[c] - yv;
            r[c] = rv;
            acc[0] = acc[1] = (double)rv * (double)rv;
        }
    } else {
        const uint3 bk = xcd_map(wb - ngb, gridDim.x - ngb, gx, gy);
        const int i0 = 2 * (bk.x * 64 + threadIdx.x), j = bk.y * 4 + threadIdx.y;
        const int k0 = bk.z * kRedZP, k1 = min(k0 + kRedZP, d.nz);
        if (i0 < d.nx && j < d.ny) {
            T yv[kRedZP][2], xq[kRedZP][2];
            T2 bv[kRedZP];
            unsigned fv[kRedZP];
#pragma unroll
            for (int kk = 0; kk < kRedZP; kk++) {
                const int k = min(k0 + kk, k1 - 1);  // past the last plane: a repeat, discarded
                const long q0 = cidx(d, i0, j, k);
                fv[kk] = *reinterpret_cast<const unsigned short *>(flag + q0);
                bv[kk] = *reinterpret_cast<const T2 *>(b + q0);
                pair_apply(d, S, x, i0, j, k, q0, yv[kk], xq[kk]);
            }
#pragma unroll
            for (int kk = 0; kk < kRedZP; kk++) {
                const int k = k0 + kk;
                if (k >= k1) break;
                const long q0 = cidx(d, i0, j, k);
                T v[2];
                bool on[2], gh[2];
#pragma unroll
                for (int o = 0; o < 2; o++) {
                    const unsigned f = (fv[kk] >> (8 * o)) & 0xFFu;
                    const bool ident = q0 + o == S.pin || f == F_SOLID;
                    const T bo = o ? bv[kk].y : bv[kk].x;
                    v[o] = ident ? bo - xq[kk][o] : bo - yv[kk][o];
                    gh[o] = !ident && f == F_GHOST;  // written by the ghost blocks
                    on[o] = i0 + o < d.nx && (ident || f == F_FLUID);
                }
                store_pair_skip(d, i0, q0, r, v[0], v[1], gh[0], gh[1]);
#pragma unroll
                for (int o = 0; o < 2; o++) {
                    if (!on[o]) continue;
                    const double rd = (double)v[o];
                    acc[0] += rd * rd;
                    const double rsc = rd * nscale(ns, d, i0 + o, j, k, q0 + o);
                    acc[1] += rsc * rsc;
                }
            }
        }
    }
    bsum<2>(acc, partial, wb);
}

template <typename T>
void launch_residual(hipStream_t st, const Dims &d, const SysCoef<T> &S, const GhostRows<T> &G,
                     const uint8_t *flag, const T *b, const T *x, T *r, const NormSpec &ns, double *partial) {
    hipLaunchKernelGGL((k_residual<T>), dim3(agrid_blocks(d, G.ng)), dim3(64, 4), 0, st, d, S, G, flag, b, x, r, ns,
                       partial, ghost_blocks(G.ng), pgx(d), pgy(d));
}

// ---------------------------------------------------------------------------
// Jacobi preconditioner z = D^-1 r
// ---------------------------------------------------------------------------
// Two cells of a row per lane. Fluid rows: r / diag; solid and pinned rows:
// r (identity); ghost rows: done by the ghost blocks of the launch.
template <typename T>
__device__ inline T jacobi_row(const Dims &d, const SysCoef<T> &S, unsigned f, int i, int j, int k, long q, T rv) {
    const T diag = stencil_diag(S, min(i, d.nx - 1), j, k, q);
    if (q == S.pin || f == F_SOLID) return rv;
    return f == F_FLUID && diag != T(0) ? rv / diag : T(0);
}

// diagonal of ghost row s
template <typename T>
__device__ inline T ghost_diag(const GhostRows<T> &G, int is_p, int s) {
    return is_p ? T(1) - G.wself[s] : T(1) + G.wself[s];
}

template <typename T>
__global__ void __launch_bounds__(256) k_jacobi(Dims d, SysCoef<T> S, GhostRows<T> G, const uint8_t *flag,
                                                const T *r, T *z, unsigned ngb, unsigned gx, unsigned gy) {
    using T2 = typename Vec2<T>::type;
    const unsigned wb = blockIdx.x;
    if (wb < ngb) {
        const int s = (int)wb * 64 + threadIdx.x;
        if (threadIdx.y == 0 && s < G.ng) {
            const long c = G.cell[s];
            const T dg = ghost_diag(G, S.is_p, s);
            z[c] = dg != T(0) ? r[c] / dg : r[c];
        }
        return;
    }
    const uint3 bk = xcd_map(wb - ngb, gridDim.x - ngb, gx, gy);
    const int i0 = 2 * (bk.x * 64 + threadIdx.x), j = bk.y * 4 + threadIdx.y, k = bk.z;
    if (i0 >= d.nx || j >= d.ny) return;
    const long q0 = cidx(d, i0, j, k);
    const unsigned fv = *reinterpret_cast<const unsigned short *>(flag + q0);
    const T2 rv = *reinterpret_cast<const T2 *>(r + q0);
    const unsigned f0 = fv & 0xFFu, f1 = fv >> 8;
    store_pair_skip(d, i0, q0, z, jacobi_row(d, S, f0, i0, j, k, q0, rv.x),
                    jacobi_row(d, S, f1, i0 + 1, j, k, q0 + 1, rv.y), f0 == F_GHOST && q0 != S.pin,
                    f1 == F_GHOST && q0 + 1 != S.pin);
}

// BiCGSTAB direction updates fused with the Jacobi preconditioner:
//   SJ = 0: p = r + beta (p - omega v), ph = D^-1 p
//   SJ = 1: s = r - alpha v,            sh = D^-1 s
//   SJ = 2: p = r (the first direction), ph = D^-1 p
// Same arithmetic as k_bicg_p / k_bicg_s followed by k_jacobi; ghost cells
// (both outputs) are done by the ghost blocks. TPH: storage type of ph.
template <typename T, int SJ, typename TPH = T>
__global__ void __launch_bounds__(256) k_bicg_jacobi(Dims d, SysCoef<T> S, GhostRows<T> G, const uint8_t *flag,
                                                     const KScal *ks, const T *r, const T *v, T *p, TPH *ph,
                                                     unsigned ngb, unsigned gx, unsigned gy) {
    using T2 = typename Vec2<T>::type;
    const unsigned wb = blockIdx.x;
    if (wb < ngb) {
        const int s = (int)wb * 64 + threadIdx.x;
        if (threadIdx.y == 0 && s < G.ng) {
            const long c = G.cell[s];
            T n;
            if (SJ == 0) n = r[c] + (T)ks->beta * (p[c] - (T)ks->omega * v[c]);
            else if (SJ == 1) n = r[c] - (T)ks->alpha * v[c];
            else n = r[c];
            p[c] = n;
            const T dg = ghost_diag(G, S.is_p, s);
            ph[c] = (TPH)(dg != T(0) ? n / dg : n);
        }
        return;
    }
    const uint3 bk = xcd_map(wb - ngb, gridDim.x - ngb, gx, gy);
    const int i0 = 2 * (bk.x * 64 + threadIdx.x), j = bk.y * 4 + threadIdx.y, k = bk.z;
    if (i0 >= d.nx || j >= d.ny) return;
    const long q0 = cidx(d, i0, j, k);
    const unsigned fv = *reinterpret_cast<const unsigned short *>(flag + q0);
    const T2 rv = *reinterpret_cast<const T2 *>(r + q0);
    T n0 = rv.x, n1 = rv.y;
    if (SJ == 0) {
        const T2 vv = *reinterpret_cast<const T2 *>(v + q0);
        const T2 pv = *reinterpret_cast<const T2 *>(p + q0);
        const T beta = (T)ks->beta, omega = (T)ks->omega;
        n0 = rv.x + beta * (pv.x - omega * vv.x);
        n1 = rv.y + beta * (pv.y - omega * vv.y);
    } else if (SJ == 1) {
        const T2 vv = *reinterpret_cast<const T2 *>(v + q0);
        const T alpha = (T)ks->alpha;
        n0 = rv.x - alpha * vv.x;
        n1 = rv.y - alpha * vv.y;
    }
    const unsigned f0 = fv & 0xFFu, f1 = fv >> 8;
    const bool g0 = f0 == F_GHOST && q0 != S.pin, g1 = f1 == F_GHOST && q0 + 1 != S.pin;
    store_pair_skip(d, i0, q0, p, n0, n1, g0, g1);
    store_pair_skip(d, i0, q0, ph, (TPH)jacobi_row(d, S, f0, i0, j, k, q0, n0),
                    (TPH)jacobi_row(d, S, f1, i0 + 1, j, k, q0 + 1, n1), g0, g1);
}

// one cell pair per lane, one plane per block
static inline dim3 jgrid(const Dims &d) { return dim3(((d.nx + 1) / 2 + 63) / 64, (d.ny + 3) / 4, d.nz); }
// the same blocks on a 1-D grid behind the ghost blocks
static inline unsigned jblocks(const Dims &d, int ng) { return ghost_blocks(ng) + pgx(d) * pgy(d) * d.nz; }

template <typename T>
void launch_jacobi(hipStream_t st, const Dims &d, const SysCoef<T> &S, const GhostRows<T> &G, const uint8_t *flag,
                   const T *r, T *z) {
    hipLaunchKernelGGL((k_jacobi<T>), dim3(jblocks(d, G.ng)), dim3(64, 4), 0, st, d, S, G, flag, r, z,
                       ghost_blocks(G.ng), pgx(d), pgy(d));
}

// BiCGSTAB direction update in FP64 (SJ 0: p = r + beta (p - omega v); SJ 1:
// s = r - alpha v, into p) that also writes the FP32 right-hand side of the
// multigrid preconditioner (b32 = (float) p) and, FIRST, the first smoother
// colour on it (x32). Same arithmetic as k_bicg_p / k_bicg_s + k_d2f_first.
template <int SJ, bool FIRST>
__global__ void __launch_bounds__(256) k_bicg_d2f(Dims d, SysCoef<float> S, const uint8_t *flag, const KScal *ks,
                                                  const double *r, const double *v, double *p, float *b32,
                                                  float *x32) {
    const int i0 = 2 * (blockIdx.x * 64 + threadIdx.x), j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i0 >= d.nx || j >= d.ny) return;
    const long q0 = cidx(d, i0, j, k);
    const double2 rv = *reinterpret_cast<const double2 *>(r + q0);
    const unsigned fv = FIRST ? *reinterpret_cast<const unsigned short *>(flag + q0) : 0u;
    // diagonals of the pair (loaded ahead of the update)
    const float dg0 = FIRST ? stencil_diag(S, i0, j, k, q0) : 0.f;
    const float dg1 = FIRST ? stencil_diag(S, min(i0 + 1, d.nx - 1), j, k, q0 + 1) : 0.f;
    double n0 = rv.x, n1 = rv.y;
    if (SJ == 0) {
        const double2 vv = *reinterpret_cast<const double2 *>(v + q0);
        const double2 pv = *reinterpret_cast<const double2 *>(p + q0);
        const double beta = ks->beta, omega = ks->omega;
        n0 = rv.x + beta * (pv.x - omega * vv.x);
        n1 = rv.y + beta * (pv.y - omega * vv.y);
    } else if (SJ == 1) {
        const double2 vv = *reinterpret_cast<const double2 *>(v + q0);
        const double alpha = ks->alpha;
        n0 = rv.x - alpha * vv.x;
        n1 = rv.y - alpha * vv.y;
    }
    const float b0 = (float)n0, b1 = (float)n1;
    if (i0 + 1 < d.nx) {
        *reinterpret_cast<double2 *>(p + q0) = make_double2(n0, n1);
        *reinterpret_cast<float2 *>(b32 + q0) = make_float2(b0, b1);
        if (FIRST)
            *reinterpret_cast<float2 *>(x32 + q0) =
                make_float2(first_colour_of(S, fv & 0xFFu, i0, j, k, q0, b0, dg0),
                            first_colour_of(S, fv >> 8, i0 + 1, j, k, q0 + 1, b1, dg1));
    } else {
        p[q0] = n0;
        b32[q0] = b0;
        if (FIRST) x32[q0] = first_colour_of(S, fv & 0xFFu, i0, j, k, q0, b0, dg0);
    }
}

void launch_bicg_d2f(hipStream_t st, const Dims &d, const SysCoef<float> *S, const uint8_t *flag, const KScal *ks,
                     int which, const double *r, const double *v, double *p, float *b32, float *x32) {
    const SysCoef<float> S0 = S ? *S : SysCoef<float>{};
    if (which == 0) {
        if (x32) hipLaunchKernelGGL((k_bicg_d2f<0, true>), jgrid(d), dim3(64, 4), 0, st, d, S0, flag, ks, r, v, p, b32, x32);
        else hipLaunchKernelGGL((k_bicg_d2f<0, false>), jgrid(d), dim3(64, 4), 0, st, d, S0, flag, ks, r, v, p, b32, x32);
    } else if (which == 2) {
        if (x32) hipLaunchKernelGGL((k_bicg_d2f<2, true>), jgrid(d), dim3(64, 4), 0, st, d, S0, flag, ks, r, v, p, b32, x32);
        else hipLaunchKernelGGL((k_bicg_d2f<2, false>), jgrid(d), dim3(64, 4), 0, st, d, S0, flag, ks, r, v, p, b32, x32);
    } else {
        if (x32) hipLaunchKernelGGL((k_bicg_d2f<1, true>), jgrid(d), dim3(64, 4), 0, st, d, S0, flag, ks, r, v, p, b32, x32);
        else hipLaunchKernelGGL((k_bicg_d2f<1, false>), jgrid(d), dim3(64, 4), 0, st, d, S0, flag, ks, r, v, p, b32, x32);
    }
}

template <typename T>
void launch_bicg_jacobi(hipStream_t st, const Dims &d, const SysCoef<T> &S, const GhostRows<T> &G,
                        const uint8_t *flag, const KScal *ks, int which, const T *r, const T *v, T *p, T *ph) {
    const unsigned ngb = ghost_blocks(G.ng), nb = jblocks(d, G.ng);
    if (which == 0)
        hipLaunchKernelGGL((k_bicg_jacobi<T, 0>), dim3(nb), dim3(64, 4), 0, st, d, S, G, flag, ks, r, v, p, ph, ngb,
                           pgx(d), pgy(d));
    else if (which == 1)
        hipLaunchKernelGGL((k_bicg_jacobi<T, 1>), dim3(nb), dim3(64, 4), 0, st, d, S, G, flag, ks, r, v, p, ph, ngb,
                           pgx(d), pgy(d));
    else
        hipLaunchKernelGGL((k_bicg_jacobi<T, 2>), dim3(nb), dim3(64, 4), 0, st, d, S, G, flag, ks, r, v, p, ph, ngb,
                           pgx(d), pgy(d));
}

void launch_bicg_jacobi_f32(hipStream_t st, const Dims &d, const SysCoef<double> &S, const GhostRows<double> &G,
                            const uint8_t *flag, const KScal *ks, int which, const double *r, const double *v,
                            double *p, float *ph) {
    const unsigned ngb = ghost_blocks(G.ng), nb = jblocks(d, G.ng);
    if (which == 0)
        hipLaunchKernelGGL((k_bicg_jacobi<double, 0, float>), dim3(nb), dim3(64, 4), 0, st, d, S, G, flag, ks, r, v, p,
                           ph, ngb, pgx(d), pgy(d));
    else if (which == 1)
        hipLaunchKernelGGL((k_bicg_jacobi<double, 1, float>), dim3(nb), dim3(64, 4), 0, st, d, S, G, flag, ks, r, v, p,
                           ph, ngb, pgx(d), pgy(d));
    else
        hipLaunchKernelGGL((k_bicg_jacobi<double, 2, float>), dim3(nb), dim3(64, 4), 0, st, d, S, G, flag, ks, r, v, p,
                           ph, ngb, pgx(d), pgy(d));
}

// ---------------------------------------------------------------------------
// Red-black Gauss-Seidel (over-relaxed by SysCoef::omega) on fluid rows,
// Jacobi on ghost rows after each colour.
// Thread t of a row handles x = 2t + parity so every lane owns one cell.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline void rbgs_cell(const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b, T *x,
                                 int color, int zero_in, uint3 bk) {
    const int j = bk.y * 4 + threadIdx.y, k = bk.z;
    const int i = 2 * (bk.x * 64 + threadIdx.x) + ((j + k + color) & 1);
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    if (zero_in) {
        // the initial x is zero: this lane also clears the other cell of its
        // pair (2t, 2t + 1), so every interior cell is written once
        const int ip = i ^ 1;
        if (ip < d.nx) x[cidx(d, ip, j, k)] = T(0);
    }
    const uint8_t f = flag[q];
    const T bv = b[q];
    const T xo = zero_in ? T(0) : x[q];
    T diag, off;
    if (zero_in) {
        diag = stencil_diag(S, i, j, k, q);
        off = 0;
    } else {
        stencil(d, S, flag, i, j, k, q, x, diag, off);
    }
    if (q == S.pin || f == F_SOLID) { x[q] = bv; return; }
    if (f != F_FLUID) {
        if (zero_in) x[q] = T(0);
        return;
    }
    x[q] = sor_value(S.omega, xo, diag != T(0) ? (bv + off) / diag : T(0));
}

template <typename T>
__global__ void __launch_bounds__(256) k_rbgs(Dims d, SysCoef<T> S, const uint8_t *flag, const T *b, T *x,
                                              int color, int zero_in, int xcd) {
    rbgs_cell(d, S, flag, b, x, color, zero_in, xcd_block(xcd));
}

// (i, j, k) of a linear index of the padded layout (any cell, ghost layers
// included; 32-bit arithmetic: padded grids stay below 2^31 cells)
__device__ inline void cell_ijk(const Dims &d, long q, int &i, int &j, int &k) {
    const unsigned r = (unsigned)(q - (d.off - 2 - 2 * (long)d.sy - 2 * d.sz));  // cell (-2, -2, -2) -> 0
    const unsigned sy = (unsigned)d.sy, sz = (unsigned)d.sz;
    const unsigned kz = r / sz, rem = r - kz * sz, jy = rem / sy;
    k = (int)kz - 2;
    j = (int)jy - 2;
    i = (int)(rem - jy * sy) - 2;
}

// Ghost rows relaxed inside the smoother launch of colour `color` (not a
// zero start), eight lanes per row: lane l of the row's group evaluates
// source l, and the row's weighted sum is formed in source order from the
// group's values (exchanged within the wave), so the result is that of one
// lane summing its eight sources. Rows whose cell has that colour take the
// Jacobi value of their row from the current sources, the rest carry their
// value over (next = prev). A fluid source of the launch's colour is updated
// by the same launch, so its new value is evaluated here with the
// Gauss-Seidel formula of the smoother (its neighbours are of the other
// colour, which the launch does not change); fluid cells never read a ghost
// cell of the launch's colour. With SysCoef::omega = 1 the result equals a
// colour launch followed by a relaxation of that colour's ghost rows; an
// over-relaxed launch writes xo + omega (gs - xo) at such a source while the
// ghost row takes its Gauss-Seidel value gs (the old value xo is being
// overwritten by the same launch and cannot be read), a fixed linear map all
// the same. Spreading a row's gathers (a source that needs no
// stencil evaluates the ghost cell's, whose neighbours are in range) over
// eight lanes keeps the ghost blocks shorter than the sweep they run beside.
// t: lane in the wave; s: the row of this lane's group (may be >= G.ng).
template <typename T>
__device__ inline void ghost_rows_in_colour(const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b,
                                            T *x, const GhostRows<T> &G, int is_p, const T *prev, T *next,
                                            int color, int t, int s) {
    const int l = t & 7;
    const bool in = s < G.ng;
    const long c = in ? G.cell[s] : 0;
    int ic = 0, jc = 0, kc = 0;
    bool act = false;
    if (in) {
        cell_ijk(d, c, ic, jc, kc);
        act = ((ic + jc + kc) & 1) == color;
    }
    double wl = 0;
    T vl = 0;
    int used = 0;
    if (act) {
        const long q0 = G.src[8 * (long)s + l];
        const int g = G.srcg[8 * (long)s + l];
        wl = G.w64[8 * (long)s + l];
        used = q0 >= 0;
        const long q = q0 >= 0 ? q0 : c;
        int i, j, k;
        cell_ijk(d, q, i, j, k);
        const bool same = g < 0 && i >= 0 && i < d.nx && j >= 0 && j < d.ny && k >= 0 && k < d.nz &&
                          ((i + j + k) & 1) == color;
        const long qs = same ? q : c;
        const uint8_t f = flag[q];
        const T bq = b[q], xq = x[q], pg = prev[g >= 0 ? g : 0];
        T diag, off;
        stencil(d, S, flag, same ? i : ic, same ? j : jc, same ? k : kc, qs, x, diag, off);
        T v = xq;
        if (g >= 0) v = pg;
        else if (same && (q == S.pin || f == F_SOLID)) v = bq;
        else if (same && f == F_FLUID) v = diag != T(0) ? (bq + off) / diag : T(0);
        vl = v;
    }
    // every lane of the wave takes part in the exchange
    const int base = t & ~7;
    double a = 0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const double wm = __shfl(wl, base + m, 64);
        const T vm = __shfl(vl, base + m, 64);
        const int um = __shfl(used, base + m, 64);
        if (um) a += wm * (double)vm;
    }
    if (!in || l != 0) return;
    if (!act) {
        next[s] = prev[s];
        return;
    }
    const double ws = G.ws64[s];
    T v;
    if (is_p) {
        const double dg = 1.0 - ws;
        v = dg != 0.0 ? (T)(((double)b[c] + a) / dg) : x[c];
    } else {
        v = (T)(((double)b[c] - a) / (1.0 + ws));
    }
    x[c] = v;
    next[s] = v;
}

// ghost rows per block of the colour launches and of k_ghost_relax: eight
// lanes per row
__device__ inline int ghost_row_of_lane(unsigned w) {
    return (int)(w * (blockDim.x * blockDim.y / 8) + threadIdx.y * (blockDim.x / 8) + (threadIdx.x >> 3));
}
static inline unsigned colour_ghost_blocks(int ng, int threads) {
    const int rows = threads / 8;
    return ng > 0 ? (unsigned)(((ng + rows - 1) / rows + 7) / 8 * 8) : 0u;
}

// One colour (not a zero start) with that colour's ghost rows, on a 1-D grid:
// the first ngb blocks (a multiple of 8, dispatched first and spread over
// the XCDs) relax ghost rows, eight lanes per row; the other blocks run the
// one-colour sweep on a gx x gy x nz block grid, XCD-contiguous.
template <typename T>
__global__ void __launch_bounds__(256) k_rbgs_g(Dims d, SysCoef<T> S, const uint8_t *flag, const T *b, T *x,
                                                int color, GhostRows<T> G, int is_p, const T *prev, T *next,
                                                unsigned ngb, unsigned gx, unsigned gy) {
    const unsigned w = blockIdx.x;
    if (w < ngb) {
        ghost_rows_in_colour(d, S, flag, b, x, G, is_p, prev, next, color, threadIdx.x, ghost_row_of_lane(w));
        return;
    }
    rbgs_cell(d, S, flag, b, x, color, 0, xcd_map(w - ngb, gridDim.x - ngb, gx, gy));
}

// four cells per lane (rbgs_quad.h) on levels with rows of at least this
// many cells; shorter rows leave most lanes of a 64-lane row idle
constexpr int kQuadMinRow = 128;

template <int BY>
__global__ void __launch_bounds__(64 * BY) k_rbgs_quad(Dims d, SysCoef<float> S, const uint8_t *flag, const float *b,
                                                       float *x, int color, int zero_in) {
    rbgs_quad<BY>(d, S, flag, b, x, color, zero_in, xcd_block(1));
}

// k_rbgs_g for the four-cells-per-lane sweep
template <int BY>
__global__ void __launch_bounds__(64 * BY) k_rbgs_quad_g(Dims d, SysCoef<float> S, const uint8_t *flag,
                                                         const float *b, float *x, int color, GhostRows<float> G,
                                                         int is_p, const float *prev, float *next, unsigned ngb,
                                                         unsigned gx, unsigned gy) {
    const unsigned w = blockIdx.x;
    if (w < ngb) {
        ghost_rows_in_colour(d, S, flag, b, x, G, is_p, prev, next, color, threadIdx.x, ghost_row_of_lane(w));
        return;
    }
    rbgs_quad<BY>(d, S, flag, b, x, color, 0, xcd_map(w - ngb, gridDim.x - ngb, gx, gy));
}

// r = b - A x on the fluid rows (pinned row: b - x); no norms
template <typename T>
__global__ void __launch_bounds__(256) k_mg_residual(Dims d, SysCoef<T> S, const uint8_t *__restrict__ flag,
                                                     const T *__restrict__ b, const T *__restrict__ x,
                                                     T *__restrict__ r) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    const uint8_t f = flag[q];
    const T xv = x[q], bv = b[q];
    T diag, off;
    stencil(d, S, flag, i, j, k, q, x, diag, off);
    if (f != F_FLUID) return;
    r[q] = q == S.pin ? bv - xv : bv - (diag * xv - off);
}

template <int BY>
__global__ void __launch_bounds__(64 * BY) k_resid_quad(Dims d, SysCoef<float> S, const uint8_t *flag, const float *b,
                                                        const float *x, float *r) {
    resid_quad<BY>(d, S, flag, b, x, r, xcd_block(1));
}

template <typename T>
bool launch_mg_residual(hipStream_t st, const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b,
                        const T *x, T *r) {
    if constexpr (std::is_same<T, float>::value) {
        if (d.nx >= kQuadMinRow) {
            dim3 grid(((d.nx + 3) / 4 + 63) / 64, (d.ny + 7) / 8, d.nz);
            hipLaunchKernelGGL((k_resid_quad<8>), grid, dim3(64, 8), 0, st, d, S, flag, b, x, r);
            return true;
        }
    }
    hipLaunchKernelGGL((k_mg_residual<T>), igrid(d), dim3(64, 4), 0, st, d, S, flag, b, x, r);
    return false;
}

// Residual of the fluid rows restricted to the coarse grid in one pass:
// bc = (sum over fluid children of r * V_child) / V_coarse, r = b - A x
// (pinned row: b - x). Equal to the residual kernel followed by the
// restriction kernel, without storing r.
template <typename T, bool FIRST>
__global__ void __launch_bounds__(256) k_resid_restrict(Dims df, SysCoef<T> S, const uint8_t *flag_f, const T *b,
                                                        const T *x, Dims dc, MGTransfer tr, const double *dxf,
                                                        const double *dyf, const double *dzf, const double *dxc,
                                                        const double *dyc, const double *dzc, T *bc,
                                                        SysCoef<T> Sc, const uint8_t *flag_c, T *xc) {
    const int I = blockIdx.x * 64 + threadIdx.x, J = blockIdx.y * 4 + threadIdx.y, K = blockIdx.z;
    if (I >= dc.nx || J >= dc.ny) return;
    // one or two children per axis; a single child is evaluated twice with
    // the second copy weighted 0, and non-fluid children are weighted 0, so
    // the loads of all eight children issue together (these levels are
    // latency bound)
    const int i0 = tr.si[I], ni = tr.si[I + 1] - i0, j0 = tr.sj[J], nj = tr.sj[J + 1] - j0;
    const int k0 = tr.sk[K], nk = tr.sk[K + 1] - k0;
    const int ii[2] = {i0, i0 + ni - 1}, jj[2] = {j0, j0 + nj - 1}, kk[2] = {k0, k0 + nk - 1};
    const double vx[2] = {C1(dxf, ii[0]), ni > 1 ? C1(dxf, ii[1]) : 0.0};
    const double vy[2] = {C1(dyf, jj[0]), nj > 1 ? C1(dyf, jj[1]) : 0.0};
    const double vz[2] = {C1(dzf, kk[0]), nk > 1 ? C1(dzf, kk[1]) : 0.0};
    // the coarse cell's diagonal and flag (first colour), loaded ahead
    const long qc = cidx(dc, I, J, K);
    const T cdiag = FIRST ? stencil_diag(Sc, I, J, K, qc) : T(0);
    const uint8_t cflag = FIRST ? flag_c[qc] : (uint8_t)0;
    T rv[8];
    uint8_t fv[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int i = ii[c & 1], j = jj[(c >> 1) & 1], k = kk[c >> 2];
        const long q = cidx(df, i, j, k);
        fv[c] = flag_f[q];
        const T bq = b[q], xq = x[q];
        T diag, off;
        stencil(df, S, flag_f, i, j, k, q, x, diag, off);
        rv[c] = q == S.pin ? bq - xq : bq - (diag * xq - off);
    }
    double acc = 0;
#pragma unroll
    for (int c = 0; c < 8; c++)
        if (fv[c] == F_FLUID) acc += (double)rv[c] * vx[c & 1] * vy[(c >> 1) & 1] * vz[c >> 2];
    const T bv = (T)(acc / (C1(dxc, I) * C1(dyc, J) * C1(dzc, K)));
    bc[qc] = bv;
    if (FIRST) xc[qc] = first_colour_of(Sc, cflag, I, J, K, qc, bv, cdiag);
}

template <typename T>
void launch_resid_restrict(hipStream_t st, const Dims &df, const SysCoef<T> &S, const uint8_t *flag_f, const T *b,
                           const T *x, const Dims &dc, const MGTransfer &tr, const double *dxyz_f,
                           const double *dxyz_c, T *bc, const SysCoef<T> *Sc, const uint8_t *flag_c, T *xc) {
    const double *dxf = dxyz_f, *dyf = dxf + df.nx + 4, *dzf = dyf + df.ny + 4;
    const double *dxc = dxyz_c, *dyc = dxc + dc.nx + 4, *dzc = dyc + dc.ny + 4;
    if (xc)
        hipLaunchKernelGGL((k_resid_restrict<T, true>), igrid(dc), dim3(64, 4), 0, st, df, S, flag_f, b, x, dc, tr, dxf,
                           dyf, dzf, dxc, dyc, dzc, bc, *Sc, flag_c, xc);
    else
        hipLaunchKernelGGL((k_resid_restrict<T, false>), igrid(dc), dim3(64, 4), 0, st, df, S, flag_f, b, x, dc, tr,
                           dxf, dyf, dzf, dxc, dyc, dzc, bc, S, flag_c, xc);
}

// planes per workgroup: enough workgroups to fill the chip
static int rb_kchunk(const Dims &d) {
    const long tiles = (long)((d.nx + kRbTX - 1) / kRbTX) * ((d.ny + kRbTY - 1) / kRbTY);
    long kc = (tiles * d.nz + 1023) / 1024;
    kc = std::max(4L, std::min(std::min((long)d.nz, (long)kRbMaxChunk), kc));
    return (int)kc;
}

template <typename T>
void launch_rbgs_fused(hipStream_t st, const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b,
                       const T *xa, T *xb, int zero_in) {
    const int kc = rb_kchunk(d);
    dim3 grid((d.nx + kRbTX - 1) / kRbTX, (d.ny + kRbTY - 1) / kRbTY, (d.nz + kc - 1) / kc);
    hipLaunchKernelGGL((k_rbgs_fused<T>), grid, dim3(256), 0, st, d, S, flag, b, xa, xb, zero_in, kc);
}

// Jacobi relaxation of the ghost rows. A ghost stencil may read other ghost
// cells, which this launch also writes: those sources are read from the
// values the previous relaxation left in a separate buffer (no other kernel
// writes x at ghost cells in between), so the result does not depend on the
// order in which waves run. Eight lanes per row, one per source; the row's
// sum is formed in source order from the values exchanged within the wave.
template <typename T>
__global__ void __launch_bounds__(256) k_ghost_relax(GhostRows<T> G, int is_p, const T *b, T *x, const T *prev,
                                                     T *next, Dims d, int only) {
    const int s = ghost_row_of_lane(blockIdx.x), t = threadIdx.x & 63, l = t & 7;
    const bool in = s < G.ng;
    bool act = in;
    if (in && only >= 0) {
        // rows of the other colour carry their value over
        int iq, jq, kq;
        cell_ijk(d, G.cell[s], iq, jq, kq);
        act = ((iq + jq + kq) & 1) == only;
    }
    double wl = 0;
    T vl = 0;
    int used = 0;
    if (act) {
        const long c = G.src[8 * (long)s + l];
        if (c >= 0) {
            const int g = G.srcg[8 * (long)s + l];
            vl = g < 0 ? x[c] : prev ? prev[g] : T(0);
            wl = G.w64[8 * (long)s + l];
            used = 1;
        }
    }
    // every lane of the wave takes part in the exchange
    const int base = t & ~7;
    double a = 0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const double wm = __shfl(wl, base + m, 64);
        const T vm = __shfl(vl, base + m, 64);
        const int um = __shfl(used, base + m, 64);
        if (um) a += wm * (double)vm;
    }
    if (!in || l != 0) return;
    if (!act) {
        next[s] = prev ? prev[s] : T(0);
        return;
    }
    const long c = G.cell[s];
    const double ws = G.ws64[s];
    T v;
    if (is_p) {
        const double dg = 1.0 - ws;
        v = dg != 0.0 ? (T)(((double)b[c] + a) / dg) : x[c];
    } else {
        v = (T)(((double)b[c] - a) / (1.0 + ws));
    }
    x[c] = v;
    next[s] = v;
}

template <typename T>
void launch_rbgs_color(hipStream_t st, const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b, T *x,
                       int color, int zero_in) {
    if constexpr (std::is_same<T, float>::value) {
        if (d.nx >= kQuadMinRow) {
            dim3 grid(((d.nx + 3) / 4 + 63) / 64, (d.ny + 7) / 8, d.nz);
            hipLaunchKernelGGL((k_rbgs_quad<8>), grid, dim3(64, 8), 0, st, d, S, flag, b, x, color, zero_in);
            return;
        }
    }
    dim3 grid(((d.nx + 1) / 2 + 63) / 64, (d.ny + 3) / 4, d.nz);
    // XCD-contiguous block order: measured 16.4 us against 19.7 us per FP32
    // launch on a 231 x 115 x 115 level (58 blocks per plane, so neighbouring
    // planes otherwise land on different XCDs), no change at 256 x 128 x 128
    hipLaunchKernelGGL((k_rbgs<T>), grid, dim3(64, 4), 0, st, d, S, flag, b, x, color, zero_in, 1);
}

template <typename T>
int launch_rbgs_color_ghost(hipStream_t st, const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b,
                            T *x, int color, const GhostRows<T> &G, int prev) {
    if (prev < 0 || G.ng <= 0) abort();  // needs relaxed ghost values (not a zero start)
    const int out = prev == 0 ? 1 : 0;
    // ghost blocks: a row per eight lanes, rounded up to a multiple of the
    // XCD count so that the sweep blocks keep their XCD assignment
    if constexpr (std::is_same<T, float>::value) {
        if (d.nx >= kQuadMinRow) {
            const unsigned ngb = colour_ghost_blocks(G.ng, 64 * 8);
            const unsigned gx = ((d.nx + 3) / 4 + 63) / 64, gy = (d.ny + 7) / 8;
            hipLaunchKernelGGL((k_rbgs_quad_g<8>), dim3(ngb + gx * gy * d.nz), dim3(64, 8), 0, st, d, S, flag, b, x,
                               color, G, S.is_p, (const float *)G.vals[prev], G.vals[out], ngb, gx, gy);
            return out;
        }
    }
    const unsigned ngb = colour_ghost_blocks(G.ng, 64 * 4);
    const unsigned gx = ((d.nx + 1) / 2 + 63) / 64, gy = (d.ny + 3) / 4;
    hipLaunchKernelGGL((k_rbgs_g<T>), dim3(ngb + gx * gy * d.nz), dim3(64, 4), 0, st, d, S, flag, b, x, color, G,
                       S.is_p, (const T *)G.vals[prev], G.vals[out], ngb, gx, gy);
    return out;
}

template <typename T>
int launch_ghost_relax(hipStream_t st, const GhostRows<T> &G, int is_p, const T *b, T *x, int prev, Dims d, int only) {
    const int out = prev == 0 ? 1 : 0;
    if (G.ng <= 0) return out;
    const int gb = (G.ng + 31) / 32;  // 32 rows per block of 256 lanes
    hipLaunchKernelGGL((k_ghost_relax<T>), dim3(gb), dim3(256), 0, st, G, is_p, b, x,
                       prev < 0 ? (const T *)nullptr : (const T *)G.vals[prev], G.vals[out], d, only);
    return out;
}

// ---------------------------------------------------------------------------
// BiCGSTAB vector kernels (van der Vorst 1992) with device-resident scalars.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_bicg_init(Dims d, T *r, T *rh, T *p, T *v, T *x, const T *b, NormSpec ns,
                                                   double *partial) {
    // x0 = 0, so r0 = b on every row (fluid, solid, pinned and ghost): the
    // norms of r0 are those of b (||b||^2, ||N b||^2), kRedZ planes per block
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int k0 = blockIdx.z * kRedZ, k1 = min(k0 + kRedZ, d.nz);
    double acc[2] = {0, 0};
    if (i < d.nx && j < d.ny) {
        for (int k = k0; k < k1; k++) {
            const long q = cidx(d, i, j, k);
            const T bv = b[q];
            r[q] = bv; rh[q] = bv; x[q] = 0;
            if (p) { p[q] = 0; v[q] = 0; }
            const double bd = (double)bv, bs = bd * nscale(ns, d, i, j, k, q);
            acc[0] += bd * bd;
            acc[1] += bs * bs;
        }
    }
    bsum<2>(acc, partial, blin());
}

// restart from the current iterate: r holds b - A x; the shadow residual
// becomes r, the search directions are cleared
template <typename T>
__global__ void __launch_bounds__(256) k_bicg_restart(Dims d, const T *r, T *rh, T *p, T *v) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    rh[q] = r[q]; p[q] = 0; v[q] = 0;
}

template <typename T>
__global__ void __launch_bounds__(256) k_bicg_p(Dims d, const KScal *ks, const T *r, const T *v, T *p) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    const T beta = (T)ks->beta, omega = (T)ks->omega;
    p[q] = r[q] + beta * (p[q] - omega * v[q]);
}

template <typename T>
__global__ void __launch_bounds__(256) k_bicg_s(Dims d, const KScal *ks, const T *r, const T *v, T *s) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    s[q] = r[q] - (T)ks->alpha * v[q];
}

// TP: storage type of the preconditioned directions ph, sh
template <typename T, typename TP = T>
__global__ void __launch_bounds__(256) k_bicg_xr(Dims d, const KScal *ks, const TP *ph, const TP *sh, const T *s,
                                                 const T *t, const T *rh, T *x, T *r, NormSpec ns, double *partial) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int k0 = blockIdx.z * kRedZ, k1 = min(k0 + kRedZ, d.nz);
    double acc[3] = {0, 0, 0};
    if (i < d.nx && j < d.ny) {
        const T alpha = (T)ks->alpha, omega = (T)ks->omega;
#pragma unroll
        for (int kk = 0; kk < kRedZ; kk++) {
            const int k = k0 + kk;
            if (k >= k1) break;
            const long q = cidx(d, i, j, k);
            x[q] = x[q] + alpha * (T)ph[q] + omega * (T)sh[q];
            const T rv = s[q] - omega * t[q];
            r[q] = rv;
            const double rd = (double)rv;
            acc[0] += rd * rd;
            const double rs = rd * nscale(ns, d, i, j, k, q);
            acc[1] += rs * rs;
            acc[2] += (double)rh[q] * rd;
        }
    }
    bsum<3>(acc, partial, blin());
}

long vec_partials(const Dims &d) { return iblocks(d); }

template <typename T>
void launch_bicg_init(hipStream_t st, const Dims &d, T *r, T *rh, T *p, T *v, T *x, const T *b, const NormSpec &ns,
                      double *partial) {
    hipLaunchKernelGGL((k_bicg_init<T>), rgrid(d), dim3(64, 4), 0, st, d, r, rh, p, v, x, b, ns, partial);
}
template <typename T>
void launch_bicg_restart(hipStream_t st, const Dims &d, const T *r, T *rh, T *p, T *v) {
    hipLaunchKernelGGL((k_bicg_restart<T>), igrid(d), dim3(64, 4), 0, st, d, r, rh, p, v);
}
template <typename T>
void launch_bicg_p(hipStream_t st, const Dims &d, const KScal *ks, const T *r, const T *v, T *p) {
    hipLaunchKernelGGL((k_bicg_p<T>), igrid(d), dim3(64, 4), 0, st, d, ks, r, v, p);
}
template <typename T>
void launch_bicg_s(hipStream_t st, const Dims &d, const KScal *ks, const T *r, const T *v, T *s) {
    hipLaunchKernelGGL((k_bicg_s<T>), igrid(d), dim3(64, 4), 0, st, d, ks, r, v, s);
}
template <typename T>
void launch_bicg_xr(hipStream_t st, const Dims &d, const KScal *ks, const T *ph, const T *sh, const T *s,
                    const T *t, const T *rh, T *x, T *r, const NormSpec &ns, double *partial) {
    hipLaunchKernelGGL((k_bicg_xr<T>), rgrid(d), dim3(64, 4), 0, st, d, ks, ph, sh, s, t, rh, x, r, ns, partial);
}
void launch_bicg_xr_f32dir(hipStream_t st, const Dims &d, const KScal *ks, const float *ph, const float *sh,
                           const double *s, const double *t, const double *rh, double *x, double *r,
                           const NormSpec &ns, double *partial) {
    hipLaunchKernelGGL((k_bicg_xr<double, float>), rgrid(d), dim3(64, 4), 0, st, d, ks, ph, sh, s, t, rh, x, r, ns,
                       partial);
}

// One block of kFinThreads threads sums the block partials: thread t adds
// partials t, t + kFinThreads, ... in order, then a fixed-shape tree, so the
// result is bitwise reproducible for a given grid.
constexpr int kFinThreads = 1024;

__device__ inline void sum_columns(const double *partial, long nblk, int nv, int reduce_max,
                                   double (&sh)[4][kFinThreads]) {
    const int t = threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    // four rows per thread in flight at once, the rows past nblk predicated
    // off (a grid of up to 4 kFinThreads blocks takes one round of loads)
    for (long b = t; b < nblk; b += 4 * kFinThreads) {
        double v[4][4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const long bu = b + u * kFinThreads;
            for (int a = 0; a < nv; a++) v[u][a] = bu < nblk ? partial[bu * nv + a] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (b + u * kFinThreads >= nblk) break;
            for (int a = 0; a < nv; a++) acc[a] = reduce_max ? fmax(acc[a], v[u][a]) : acc[a] + v[u][a];
        }
    }
    // fixed-shape tree: within each wave by shuffles, then over the waves
    const int lane = t & 63, wave = t >> 6;
    for (int a = 0; a < nv; a++) {
        double x = acc[a];
        for (int o = 32; o > 0; o >>= 1) {
            const double y = __shfl_down(x, o, 64);
            x = reduce_max ? fmax(x, y) : x + y;
        }
        if (lane == 0) sh[a][wave] = x;
    }
    __syncthreads();
    if (t == 0)
        for (int a = 0; a < nv; a++) {
            double x = sh[a][0];
            for (int w = 1; w < kFinThreads / 64; w++) x = reduce_max ? fmax(x, sh[a][w]) : x + sh[a][w];
            sh[a][0] = x;
        }
    __syncthreads();
}

__global__ void __launch_bounds__(kFinThreads) k_finalize(int op, const double *partial, long nblk, int nv,
                                                          KScal *ks, KScal *pub, unsigned *pub_epoch,
                                                          unsigned epoch) {
    __shared__ double sh[4][kFinThreads];
    const int t = threadIdx.x;
    sum_columns(partial, nblk, nv, 0, sh);
    if (t != 0) return;
    const double s0 = sh[0][0], s1 = nv > 1 ? sh[1][0] : 0, s2 = nv > 2 ? sh[2][0] : 0;
    switch (op) {
    case 0:  // alpha = rho1 / (rh . v)
        ks->rv = s0;
        if (s0 != 0) ks->alpha = ks->rho1 / s0;
        else { ks->alpha = 0; ks->breakdown = 1; }
        break;
    case 1:  // omega = (t . s) / (t . t)
        ks->ts = s0; ks->tt = s1;
        ks->omega = s1 > 0 ? s0 / s1 : 0;
        break;
    case 2: {  // after x, r update
        ks->rr = s0; ks->rrn = s1;
        const double rho1 = s2;
        ks->rho = ks->rho1;
        ks->rho1 = rho1;
        if (ks->rho != 0 && ks->omega != 0) ks->beta = (rho1 / ks->rho) * (ks->alpha / ks->omega);
        else { ks->beta = 0; ks->breakdown = 1; }
        break;
    }
    case 3:  // start: norms of b (= r0 = rhat), rho1 = (rhat, r0)
        ks->bb = s0; ks->bbn = s1; ks->rr = s0; ks->rrn = s1;
        ks->rho = 1; ks->alpha = 1; ks->omega = 1;
        ks->rho1 = s0;
        ks->beta = 0;
        ks->breakdown = 0;
        break;
    case 4:  // residual norms only
        ks->rr = s0; ks->rrn = s1;
        break;
    default:;
    }
    if (pub) {
        // host copy of the scalars, then the epoch that tells the host they
        // are complete (system-scope fence between the two)
        *pub = *ks;
        __threadfence_system();
        *reinterpret_cast<volatile unsigned *>(pub_epoch) = epoch;
    }
}

// sums nv partial columns into out[0..nv) (same fixed order as k_finalize)
// out[0 .. nv) = in[0 .. nv), then *epoch_out = epoch (system-scope fence
// between): a publication into device-mapped host memory
__global__ void k_publish(const double *in, int nv, double *out, unsigned *epoch_out, unsigned epoch) {
    if (threadIdx.x != 0) return;
    for (int a = 0; a < nv; a++) out[a] = in[a];
    __threadfence_system();
    *reinterpret_cast<volatile unsigned *>(epoch_out) = epoch;
}
void launch_publish(hipStream_t st, const double *in, int nv, double *out, unsigned *epoch_out, unsigned epoch) {
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, st, in, nv, out, epoch_out, epoch);
}

__global__ void __launch_bounds__(kFinThreads) k_sum_partials(const double *partial, long nblk, int nv, double *out,
                                                              int reduce_max) {
    __shared__ double sh[4][kFinThreads];
    sum_columns(partial, nblk, nv, reduce_max, sh);
    if (threadIdx.x == 0)
        for (int a = 0; a < nv; a++) out[a] = sh[a][0];
}

void launch_sum_partials(hipStream_t st, const double *partial, long nblk, int nv, double *out, int reduce_max) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kFinThreads), 0, st, partial, nblk, nv, out, reduce_max);
}

void launch_finalize(hipStream_t st, int op, const double *partial, long nblk, int nv, KScal *ks, KScal *pub,
                     unsigned *pub_epoch, unsigned epoch) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinThreads), 0, st, op, partial, nblk, nv, ks, pub, pub_epoch,
                       epoch);
}

// ---------------------------------------------------------------------------
// Vector kernels of the restarted GMRES (Saad & Schultz 1986, classical
// Gram-Schmidt applied twice), preconditioned CG and stationary multigrid
// pressure solvers. Sums are block partials added in a fixed order.
// ---------------------------------------------------------------------------
// 1 / nscale: the row's sum of K for fluid rows, else 1
__device__ inline double nfactor(const NormSpec &ns, int i, int j, int k, long q) {
    if (!ns.on || q == ns.pin || ns.flag[q] != F_FLUID) return 1.0;
    return ns.ksx[i] + ns.ksy[j] + ns.ksz[k];
}

template <typename T>
__global__ void __launch_bounds__(256) k_gm_dots(Dims d, VecSet<T> V, int nv, T *w, NormSpec ns, int scale,
                                                 double *partial) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int k0 = blockIdx.z * kRedZ, k1 = min(k0 + kRedZ, d.nz);
    double acc[kVecSetMax + 1];
#pragma unroll
    for (int a = 0; a <= kVecSetMax; a++) acc[a] = 0;
    if (i < d.nx && j < d.ny)
        for (int k = k0; k < k1; k++) {
            const long q = cidx(d, i, j, k);
            T wv = w[q];
            if (scale) {
                wv = (T)((double)wv * nscale(ns, d, i, j, k, q));
                w[q] = wv;
            }
#pragma unroll
            for (int a = 0; a < kVecSetMax; a++)
                if (a < nv) acc[a] += (double)wv * (double)V.v[a][q];
            acc[kVecSetMax] += (double)wv * (double)wv;
        }
    bsum<kVecSetMax + 1>(acc, partial, blin());
}

template <typename T>
__global__ void __launch_bounds__(256) k_gm_update(Dims d, VecSet<T> V, int nv, const double *c, T *w,
                                                   double *partial) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int k0 = blockIdx.z * kRedZ, k1 = min(k0 + kRedZ, d.nz);
    double cv[kVecSetMax];
#pragma unroll
    for (int a = 0; a < kVecSetMax; a++) cv[a] = a < nv ? c[a] : 0.0;
    double acc[1] = {0};
    if (i < d.nx && j < d.ny)
        for (int k = k0; k < k1; k++) {
            const long q = cidx(d, i, j, k);
            double sum = 0;
#pragma unroll
            for (int a = 0; a < kVecSetMax; a++)
                if (a < nv) sum += cv[a] * (double)V.v[a][q];
            const T wv = (T)((double)w[q] - sum);
            w[q] = wv;
            acc[0] += (double)wv * (double)wv;
        }
    bsum<1>(acc, partial, blin());
}

// v = f w (f on the host); unscale: t = v / nscale (the normalised rows' scale undone)
template <typename T>
__global__ void __launch_bounds__(256) k_gm_scale(Dims d, const T *w, double f, T *v, NormSpec ns, int unscale) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    v[q] = (T)((double)w[q] * (unscale ? nfactor(ns, i, j, k, q) : f));
}

// x += a p, r -= a q; sums of r^2 and (N r)^2
template <typename T>
__global__ void __launch_bounds__(256) k_pcg_update(Dims d, double a, const T *p, const T *qv, T *x, T *r,
                                                    NormSpec ns, double *partial) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int k0 = blockIdx.z * kRedZ, k1 = min(k0 + kRedZ, d.nz);
    double acc[2] = {0, 0};
    if (i < d.nx && j < d.ny)
        for (int k = k0; k < k1; k++) {
            const long q = cidx(d, i, j, k);
            x[q] = (T)((double)x[q] + a * (double)p[q]);
            const T rv = (T)((double)r[q] - a * (double)qv[q]);
            r[q] = rv;
            acc[0] += (double)rv * (double)rv;
            const double rs = (double)rv * nscale(ns, d, i, j, k, q);
            acc[1] += rs * rs;
        }
    bsum<2>(acc, partial, blin());
}

// (a, b) per block
template <typename T>
__global__ void __launch_bounds__(256) k_dot(Dims d, const T *a, const T *b, double *partial) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int k0 = blockIdx.z * kRedZ, k1 = min(k0 + kRedZ, d.nz);
    double acc[1] = {0};
    if (i < d.nx && j < d.ny)
        for (int k = k0; k < k1; k++) {
            const long q = cidx(d, i, j, k);
            acc[0] += (double)a[q] * (double)b[q];
        }
    bsum<1>(acc, partial, blin());
}

// y = z + f y (xpby) or y = y + f z (axpy)
template <typename T>
__global__ void __launch_bounds__(256) k_lincomb(Dims d, const T *z, double f, T *y, int axpy) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    y[q] = axpy ? (T)((double)y[q] + f * (double)z[q]) : (T)((double)z[q] + f * (double)y[q]);
}

// out[a] = sum over blocks of partial[b * nv + a], one workgroup per column
__global__ void __launch_bounds__(kFinThreads) k_sum_cols(const double *partial, long nblk, int nv, double *out) {
    __shared__ double sh[kFinThreads];
    const int a = blockIdx.x, t = threadIdx.x;
    double acc = 0;
    for (long b = t; b < nblk; b += kFinThreads) acc += partial[b * nv + a];
    sh[t] = acc;
    __syncthreads();
    for (int w = kFinThreads / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) out[a] = sh[0];
}

long krylov_blocks(const Dims &d) { return iblocks(d); }

template <typename T>
void launch_gm_dots(hipStream_t st, const Dims &d, const VecSet<T> &V, int nv, T *w, const NormSpec &ns, int scale,
                    double *partial, double *out) {
    hipLaunchKernelGGL((k_gm_dots<T>), rgrid(d), dim3(64, 4), 0, st, d, V, nv, w, ns, scale, partial);
    hipLaunchKernelGGL(k_sum_cols, dim3(kVecSetMax + 1), dim3(kFinThreads), 0, st, partial, iblocks(d), kVecSetMax + 1,
                       out);
}
template <typename T>
void launch_gm_update(hipStream_t st, const Dims &d, const VecSet<T> &V, int nv, const double *c, T *w,
                      double *partial, double *out) {
    hipLaunchKernelGGL((k_gm_update<T>), rgrid(d), dim3(64, 4), 0, st, d, V, nv, c, w, partial);
    hipLaunchKernelGGL(k_sum_cols, dim3(1), dim3(kFinThreads), 0, st, partial, iblocks(d), 1, out);
}
template <typename T>
void launch_gm_scale(hipStream_t st, const Dims &d, const T *w, double f, T *v, const NormSpec &ns, int unscale) {
    hipLaunchKernelGGL((k_gm_scale<T>), igrid(d), dim3(64, 4), 0, st, d, w, f, v, ns, unscale);
}
template <typename T>
void launch_pcg_update(hipStream_t st, const Dims &d, double a, const T *p, const T *q, T *x, T *r,
                       const NormSpec &ns, double *partial, double *out) {
    hipLaunchKernelGGL((k_pcg_update<T>), rgrid(d), dim3(64, 4), 0, st, d, a, p, q, x, r, ns, partial);
    hipLaunchKernelGGL(k_sum_cols, dim3(2), dim3(kFinThreads), 0, st, partial, iblocks(d), 2, out);
}
template <typename T>
void launch_dot(hipStream_t st, const Dims &d, const T *a, const T *b, double *partial, double *out) {
    hipLaunchKernelGGL((k_dot<T>), rgrid(d), dim3(64, 4), 0, st, d, a, b, partial);
    hipLaunchKernelGGL(k_sum_cols, dim3(1), dim3(kFinThreads), 0, st, partial, iblocks(d), 1, out);
}
template <typename T>
void launch_lincomb(hipStream_t st, const Dims &d, const T *z, double f, T *y, int axpy) {
    hipLaunchKernelGGL((k_lincomb<T>), igrid(d), dim3(64, 4), 0, st, d, z, f, y, axpy);
}
template void launch_gm_dots<double>(hipStream_t, const Dims &, const VecSet<double> &, int, double *,
                                     const NormSpec &, int, double *, double *);
template void launch_gm_update<double>(hipStream_t, const Dims &, const VecSet<double> &, int, const double *,
                                       double *, double *, double *);
template void launch_gm_scale<double>(hipStream_t, const Dims &, const double *, double, double *, const NormSpec &,
                                      int);
template void launch_pcg_update<double>(hipStream_t, const Dims &, double, const double *, const double *, double *,
                                        double *, const NormSpec &, double *, double *);
template void launch_dot<double>(hipStream_t, const Dims &, const double *, const double *, double *, double *);
template void launch_lincomb<double>(hipStream_t, const Dims &, const double *, double, double *, int);

// ---------------------------------------------------------------------------
// precision conversion and copies over interior cells
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_d2f(Dims d, const double *a, float *b) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    b[q] = (float)a[q];
}
// b = (float) a, and x = the first smoother colour from a zero start; two
// cells of a row per lane
__global__ void __launch_bounds__(256) k_d2f_first(Dims d, SysCoef<float> S, const uint8_t *flag, const double *a,
                                                   float *b, float *x) {
    const int i0 = 2 * (blockIdx.x * 64 + threadIdx.x), j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i0 >= d.nx || j >= d.ny) return;
    const long q0 = cidx(d, i0, j, k);
    const double2 av = *reinterpret_cast<const double2 *>(a + q0);
    const unsigned fv = *reinterpret_cast<const unsigned short *>(flag + q0);
    const float dg0 = stencil_diag(S, i0, j, k, q0), dg1 = stencil_diag(S, min(i0 + 1, d.nx - 1), j, k, q0 + 1);
    const float b0 = (float)av.x, b1 = (float)av.y;
    const float x0 = first_colour_of(S, fv & 0xFFu, i0, j, k, q0, b0, dg0);
    if (i0 + 1 < d.nx) {
        const float x1 = first_colour_of(S, fv >> 8, i0 + 1, j, k, q0 + 1, b1, dg1);
        *reinterpret_cast<float2 *>(b + q0) = make_float2(b0, b1);
        *reinterpret_cast<float2 *>(x + q0) = make_float2(x0, x1);
    } else {
        b[q0] = b0;
        x[q0] = x0;
    }
}
__global__ void __launch_bounds__(256) k_add_f2d(Dims d, const float *e, double *x) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    x[q] += (double)e[q];
}
__global__ void __launch_bounds__(256) k_f2d(Dims d, const float *e, double *x) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    x[q] = (double)e[q];
}
template <typename T>
__global__ void __launch_bounds__(256) k_zero(Dims d, T *a) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    a[cidx(d, i, j, k)] = 0;
}
template <typename T>
__global__ void __launch_bounds__(256) k_copy(Dims d, const T *a, T *b) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = cidx(d, i, j, k);
    b[q] = a[q];
}
void launch_d2f(hipStream_t st, const Dims &d, const double *a, float *b) {
    hipLaunchKernelGGL(k_d2f, igrid(d), dim3(64, 4), 0, st, d, a, b);
}
void launch_d2f_first(hipStream_t st, const Dims &d, const SysCoef<float> &S, const uint8_t *flag, const double *a,
                      float *b, float *x) {
    hipLaunchKernelGGL(k_d2f_first, jgrid(d), dim3(64, 4), 0, st, d, S, flag, a, b, x);
}
void launch_add_f2d(hipStream_t st, const Dims &d, const float *e, double *x) {
    hipLaunchKernelGGL(k_add_f2d, igrid(d), dim3(64, 4), 0, st, d, e, x);
}
void launch_f2d(hipStream_t st, const Dims &d, const float *e, double *x) {
    hipLaunchKernelGGL(k_f2d, igrid(d), dim3(64, 4), 0, st, d, e, x);
}
template <typename T>
void launch_zero(hipStream_t st, const Dims &d, T *a) {
    hipLaunchKernelGGL((k_zero<T>), igrid(d), dim3(64, 4), 0, st, d, a);
}
template <typename T>
void launch_copy(hipStream_t st, const Dims &d, const T *a, T *b) {
    hipLaunchKernelGGL((k_copy<T>), igrid(d), dim3(64, 4), 0, st, d, a, b);
}

// ---------------------------------------------------------------------------
// multigrid transfers
// ---------------------------------------------------------------------------
// coarse rhs = volume-weighted sum of the fluid children's residuals / coarse volume
constexpr int kRestZ = 2;
template <typename T, bool FLAGS, bool FIRST>
__global__ void __launch_bounds__(256) k_restrict(Dims df, Dims dc, MGTransfer tr, const uint8_t *flag_f,
                                                  const T *rf, const double *dxf, const double *dyf,
                                                  const double *dzf, const double *dxc, const double *dyc,
                                                  const double *dzc, T *bc, SysCoef<T> Sc, const uint8_t *flag_c,
                                                  T *xc) {
    // kRestZ coarse planes per lane (shared x / y child maps, all loads
    // issued together)
    const int I = blockIdx.x * 64 + threadIdx.x, J = blockIdx.y * 4 + threadIdx.y;
    if (I >= dc.nx || J >= dc.ny) return;
    const int K0 = blockIdx.z * kRestZ;
    // one or two children per axis: a single child is read twice with the
    // second copy weighted 0, so all eight loads issue without branches
    const int i0 = tr.si[I], ni = tr.si[I + 1] - i0, j0 = tr.sj[J], nj = tr.sj[J + 1] - j0;
    const int ii[2] = {i0, i0 + ni - 1}, jj[2] = {j0, j0 + nj - 1};
    const double vx[2] = {C1(dxf, ii[0]), ni > 1 ? C1(dxf, ii[1]) : 0.0};
    const double vy[2] = {C1(dyf, jj[0]), nj > 1 ? C1(dyf, jj[1]) : 0.0};
    T rv[kRestZ][8];
    uint8_t fv[kRestZ][8];
    double vz[kRestZ][2];
    T cdiag[kRestZ];
    uint8_t cflag[kRestZ];
#pragma unroll
    for (int z = 0; z < kRestZ; z++) {
        const int K = min(K0 + z, dc.nz - 1);  // past the last plane: a repeat, not written
        if (FIRST) {
            // the coarse cell's diagonal and flag, loaded ahead of the children
            const long qc = cidx(dc, I, J, K);
            cdiag[z] = stencil_diag(Sc, I, J, K, qc);
            cflag[z] = flag_c[qc];
        }
        const int k0 = tr.sk[K], nk = tr.sk[K + 1] - k0;
        const int kk[2] = {k0, k0 + nk - 1};
        vz[z][0] = C1(dzf, kk[0]);
        vz[z][1] = nk > 1 ? C1(dzf, kk[1]) : 0.0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const long q = cidx(df, ii[c & 1], jj[(c >> 1) & 1], kk[c >> 2]);
            rv[z][c] = rf[q];
            fv[z][c] = FLAGS ? flag_f[q] : (uint8_t)F_FLUID;  // !FLAGS: rf is 0 at non-fluid cells
        }
    }
#pragma unroll
    for (int z = 0; z < kRestZ; z++) {
        const int K = K0 + z;
        if (K >= dc.nz) break;
        double acc = 0;
#pragma unroll
        for (int c = 0; c < 8; c++)
            if (fv[z][c] == F_FLUID) acc += (double)rv[z][c] * vx[c & 1] * vy[(c >> 1) & 1] * vz[z][c >> 2];
        const long qc = cidx(dc, I, J, K);
        const T bv = (T)(acc / (C1(dxc, I) * C1(dyc, J) * C1(dzc, K)));
        bc[qc] = bv;
        if (FIRST) xc[qc] = first_colour_of(Sc, cflag[z], I, J, K, qc, bv, cdiag[z]);
    }
}

// x_fine += trilinear interpolation of the coarse correction (fluid cells)
// Each lane corrects KZ planes of one (i, j) column: the index and weight
// loads are shared and all planes' loads issue together (on the large
// levels, one plane per lane left the dependent index -> coarse-value loads
// latency bound; small levels keep one plane per lane for parallelism).
constexpr long kProlManyCells = 1L << 20;
template <typename T, int KZ>
__global__ void __launch_bounds__(256) k_prolong(Dims df, Dims dc, MGTransfer tr, const uint8_t *flag_f,
                                                 const uint8_t *flag_c, const T *ec, T *xf) {
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    if (i >= df.nx || j >= df.ny) return;
    const int k0 = blockIdx.z * KZ;
    const int I = tr.ci[i], J = tr.cj[j];
    const int Is[2] = {I, tr.pi[i]}, Js[2] = {J, tr.pj[j]};
    const double wi[2] = {tr.wi[i], 1 - tr.wi[i]}, wj[2] = {tr.wj[j], 1 - tr.wj[j]};
    bool fluid[KZ];
    T xv[KZ], ev[KZ][8];
    uint8_t fv[KZ][8];
    double wk[KZ][2];
#pragma unroll
    for (int z = 0; z < KZ; z++) {
        const int k = min(k0 + z, df.nz - 1);  // past the last plane: a repeat, not written
        const long q = cidx(df, i, j, k);
        // the loads of every cell issue before the fluid test (non-fluid
        // cells are not written); the far neighbours are interior coarse
        // cells (the parent itself where the weight is 0 or at
        // non-periodic ends), so all eight loads issue without branches;
        // coarse flags are read only where the stencil touches an inactive
        // cell
        fluid[z] = flag_f[q] == F_FLUID && k0 + z < df.nz;
        xv[z] = xf[q];
        const int Ks[2] = {tr.ck[k], tr.pk[k]};
        wk[z][0] = tr.wk[k];
        wk[z][1] = 1 - wk[z][0];
        const bool chk = tr.pm[q] != 0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const long qc = cidx(dc, Is[c & 1], Js[(c >> 1) & 1], Ks[c >> 2]);
            ev[z][c] = ec[qc];
            fv[z][c] = chk ? flag_c[qc] : (uint8_t)F_FLUID;
        }
    }
#pragma unroll
    for (int z = 0; z < KZ; z++) {
        const double e0 = (double)ev[z][0];
        double e = 0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const double w = wi[c & 1] * wj[(c >> 1) & 1] * wk[z][c >> 2];
            if (w != 0) e += w * (fv[z][c] == F_SOLID ? e0 : (double)ev[z][c]);
        }
        if (fluid[z]) xf[cidx(df, i, j, k0 + z)] = xv[z] + (T)e;
    }
}

// single-workgroup red-black Gauss-Seidel on the coarsest grid; for a
// singular (all-Neumann) problem the volume-weighted mean of b is removed
// first and of x afterwards.
template <typename T>
__global__ void __launch_bounds__(1024) k_coarse(Dims d, SysCoef<T> S, const uint8_t *flag, const T *b, T *x,
                                                 int sweeps, int singular, const double *dx, const double *dy,
                                                 const double *dz) {
    __shared__ double red[2][1024];
    const int t = threadIdx.x;
    const long n = (long)d.nx * d.ny * d.nz;
    double bmean = 0;
    if (singular) {
        double sv = 0, sb = 0;
        for (long c = t; c < n; c += 1024) {
            const int i = c % d.nx, j = (c / d.nx) % d.ny, k = c / ((long)d.nx * d.ny);
            const long q = cidx(d, i, j, k);
            if (flag[q] != F_FLUID) continue;
            const double v = C1(dx, i) * C1(dy, j) * C1(dz, k);
            sv += v; sb += v * (double)b[q];
        }
        red[0][t] = sv; red[1][t] = sb;
        __syncthreads();
        for (int s = 512; s > 0; s >>= 1) {
            if (t < s) { red[0][t] += red[0][t + s]; red[1][t] += red[1][t + s]; }
            __syncthreads();
        }
        bmean = red[0][0] > 0 ? red[1][0] / red[0][0] : 0;
        __syncthreads();
    }
    for (long c = t; c < n; c += 1024) {
        const int i = c % d.nx, j = (c / d.nx) % d.ny, k = c / ((long)d.nx * d.ny);
        x[cidx(d, i, j, k)] = 0;
    }
    __syncthreads();
    for (int s = 0; s < sweeps; s++)
        for (int color = 0; color < 2; color++) {
            for (long c = t; c < n; c += 1024) {
                const int i = c % d.nx, j = (c / d.nx) % d.ny, k = c / ((long)d.nx * d.ny);
                if (((i + j + k) & 1) != color) continue;
                const long q = cidx(d, i, j, k);
                const uint8_t f = flag[q];
                if (q == S.pin || f == F_SOLID) { x[q] = b[q]; continue; }
                T diag, off;
                stencil(d, S, flag, i, j, k, q, x, diag, off);
                const T rhs = b[q] - (T)bmean;
                x[q] = diag != T(0) ? (rhs + off) / diag : T(0);
            }
            __syncthreads();
        }
    if (singular) {
        double sv = 0, sx = 0;
        for (long c = t; c < n; c += 1024) {
            const int i = c % d.nx, j = (c / d.nx) % d.ny, k = c / ((long)d.nx * d.ny);
            const long q = cidx(d, i, j, k);
            if (flag[q] != F_FLUID) continue;
            const double v = C1(dx, i) * C1(dy, j) * C1(dz, k);
            sv += v; sx += v * (double)x[q];
        }
        red[0][t] = sv; red[1][t] = sx;
        __syncthreads();
        for (int s = 512; s > 0; s >>= 1) {
            if (t < s) { red[0][t] += red[0][t + s]; red[1][t] += red[1][t + s]; }
            __syncthreads();
        }
        const double xm = red[0][0] > 0 ? red[1][0] / red[0][0] : 0;
        for (long c = t; c < n; c += 1024) {
            const int i = c % d.nx, j = (c / d.nx) % d.ny, k = c / ((long)d.nx * d.ny);
            const long q = cidx(d, i, j, k);
            if (flag[q] == F_FLUID) x[q] -= (T)xm;
        }
    }
}

template <typename T>
void launch_restrict(hipStream_t st, const Dims &df, const Dims &dc, const MGTransfer &tr, const uint8_t *flag_f,
                     const T *rf, const double *dxf, const double *dyf, const double *dzf, const double *dxc,
                     const double *dyc, const double *dzc, T *bc, bool rf_zero_off_fluid, const SysCoef<T> *Sc,
                     const uint8_t *flag_c, T *xc) {
    const dim3 grid((dc.nx + 63) / 64, (dc.ny + 3) / 4, (dc.nz + kRestZ - 1) / kRestZ);
    const SysCoef<T> S0 = Sc ? *Sc : SysCoef<T>{};
#define RESTRICT_LAUNCH(FL, FI)                                                                                 \
    hipLaunchKernelGGL((k_restrict<T, FL, FI>), grid, dim3(64, 4), 0, st, df, dc, tr, flag_f, rf, dxf, dyf, dzf, \
                       dxc, dyc, dzc, bc, S0, flag_c, xc)
    if (rf_zero_off_fluid) {
        if (xc) RESTRICT_LAUNCH(false, true);
        else RESTRICT_LAUNCH(false, false);
    } else {
        if (xc) RESTRICT_LAUNCH(true, true);
        else RESTRICT_LAUNCH(true, false);
    }
#undef RESTRICT_LAUNCH
}
template <typename T>
void launch_prolong(hipStream_t st, const Dims &df, const Dims &dc, const MGTransfer &tr, const uint8_t *flag_f,
                    const uint8_t *flag_c, const T *ec, T *xf) {
    if ((long)df.nx * df.ny * df.nz >= kProlManyCells) {
        const dim3 grid((df.nx + 63) / 64, (df.ny + 3) / 4, (df.nz + 3) / 4);
        hipLaunchKernelGGL((k_prolong<T, 4>), grid, dim3(64, 4), 0, st, df, dc, tr, flag_f, flag_c, ec, xf);
    } else {
        hipLaunchKernelGGL((k_prolong<T, 1>), igrid(df), dim3(64, 4), 0, st, df, dc, tr, flag_f, flag_c, ec, xf);
    }
}
template <typename T>
void launch_coarse_solve(hipStream_t st, const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b, T *x,
                         int sweeps, int singular, const double *dxyz) {
    const double *dx = dxyz, *dy = dxyz + d.nx + 4, *dz = dxyz + d.nx + 4 + d.ny + 4;
    hipLaunchKernelGGL((k_coarse<T>), dim3(1), dim3(1024), 0, st, d, S, flag, b, x, sweeps, singular, dx, dy, dz);
}

// ---------------------------------------------------------------------------
// exact coarsest-level solve through a precomputed dense inverse
// ---------------------------------------------------------------------------
__global__ void k_gj_prep(const double *A, int n, int p, double *rowp, double *colp) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const double piv = A[(long)p * n + p];
    rowp[t] = (t == p ? 1.0 : A[(long)p * n + t]) / piv;
    colp[t] = t == p ? 0.0 : A[(long)t * n + p];
}

__global__ void k_gj_update(double *A, int n, int p, const double *rowp, const double *colp) {
    const long e = blockIdx.x * 256L + threadIdx.x;
    if (e >= (long)n * n) return;
    const int r = (int)(e / n), c = (int)(e % n);
    if (r == p) A[e] = rowp[c];
    else A[e] = (c == p ? 0.0 : A[e]) - colp[r] * rowp[c];
}

void gauss_jordan_invert(hipStream_t st, double *A, int n, double *rowp, double *colp) {
    const int gp = (n + 255) / 256;
    const long gu = ((long)n * n + 255) / 256;
    for (int p = 0; p < n; p++) {
        hipLaunchKernelGGL(k_gj_prep, dim3(gp), dim3(256), 0, st, A, n, p, rowp, colp);
        hipLaunchKernelGGL(k_gj_update, dim3(gu), dim3(256), 0, st, A, n, p, rowp, colp);
    }
}

__global__ void k_d2f_flat(const double *a, float *b, long n) {
    const long e = blockIdx.x * 256L + threadIdx.x;
    if (e < n) b[e] = (float)a[e];
}

void launch_d2f_flat(hipStream_t st, const double *a, float *b, long n) {
    hipLaunchKernelGGL(k_d2f_flat, dim3((n + 255) / 256), dim3(256), 0, st, a, b, n);
}

constexpr int kDenseMaxRows = 4096;  // right-hand side staged in LDS

// One wavefront per row of the inverse; four independent partial sums per
// lane keep several row loads in flight.
// rows per workgroup (one wavefront each): every workgroup first gathers the
// whole right-hand side into LDS, so fewer, larger workgroups repeat that less
constexpr int kDenseWaves = 16;

template <typename T>
__global__ void __launch_bounds__(64 * kDenseWaves) k_dense_solve(const T *Ainv, int n, const long *rows,
                                                                   const double *vol, const uint8_t *flag,
                                                                   int singular, int pinrow, const T *b, T *x) {
    __shared__ double rhs[kDenseMaxRows];
    __shared__ double red[2][kDenseWaves];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    double bmean = 0;
    if (singular) {
        double sv = 0, sb = 0;
        for (int c = t; c < n; c += 64 * kDenseWaves) {
            const long q = rows[c];
            if (flag[q] != F_FLUID) continue;
            sv += vol[c];
            sb += vol[c] * (double)b[q];
        }
        for (int o = 32; o > 0; o >>= 1) { sv += __shfl_down(sv, o, 64); sb += __shfl_down(sb, o, 64); }
        if (lane == 0) { red[0][wave] = sv; red[1][wave] = sb; }
        __syncthreads();
        double V = 0, B = 0;
        for (int w = 0; w < kDenseWaves; w++) { V += red[0][w]; B += red[1][w]; }
        bmean = V > 0 ? B / V : 0;
    }
    for (int c = t; c < n; c += 64 * kDenseWaves) {
        const long q = rows[c];
        double v = (double)b[q];
        if (singular) {
            if (c == pinrow) v = 0;
            else if (flag[q] == F_FLUID) v -= bmean;
        }
        rhs[c] = v;
    }
    __syncthreads();
    const int r = blockIdx.x * kDenseWaves + wave;
    if (r >= n) return;
    const T *row = Ainv + (long)r * n;
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int c = lane;
    for (; c + 192 < n; c += 256) {
        a0 += (double)row[c] * rhs[c];
        a1 += (double)row[c + 64] * rhs[c + 64];
        a2 += (double)row[c + 128] * rhs[c + 128];
        a3 += (double)row[c + 192] * rhs[c + 192];
    }
    for (; c < n; c += 64) a0 += (double)row[c] * rhs[c];
    double acc = (a0 + a1) + (a2 + a3);
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if (lane == 0) x[rows[r]] = (T)acc;
}

template <typename T>
void launch_dense_solve(hipStream_t st, const T *Ainv, int n, const long *rows, const double *vol,
                        const uint8_t *flag, int singular, int pinrow, const T *b, T *x) {
    if (n > kDenseMaxRows) abort();
    hipLaunchKernelGGL((k_dense_solve<T>), dim3((n + kDenseWaves - 1) / kDenseWaves), dim3(64 * kDenseWaves), 0, st,
                       Ainv, n, rows, vol, flag, singular, pinrow, b, x);
}

// explicit instantiations
#define INST(T)                                                                                                   \
    template void launch_dense_solve<T>(hipStream_t, const T *, int, const long *, const double *, const uint8_t *, \
                                        int, int, const T *, T *);                                                \
    template void launch_rbgs_fused<T>(hipStream_t, const Dims &, const SysCoef<T> &, const uint8_t *, const T *,  \
                                       const T *, T *, int);                                                      \
    template void launch_rbgs_color<T>(hipStream_t, const Dims &, const SysCoef<T> &, const uint8_t *, const T *, T *, \
                                       int, int);                                                                 \
    template bool launch_mg_residual<T>(hipStream_t, const Dims &, const SysCoef<T> &, const uint8_t *, const T *, \
                                        const T *, T *);                                                          \
    template void launch_resid_restrict<T>(hipStream_t, const Dims &, const SysCoef<T> &, const uint8_t *,         \
                                           const T *, const T *, const Dims &, const MGTransfer &, const double *, \
                                           const double *, T *, const SysCoef<T> *, const uint8_t *, T *);        \
    template int launch_ghost_relax<T>(hipStream_t, const GhostRows<T> &, int, const T *, T *, int, Dims, int);    \
    template int launch_rbgs_color_ghost<T>(hipStream_t, const Dims &, const SysCoef<T> &, const uint8_t *,        \
                                            const T *, T *, int, const GhostRows<T> &, int);                      \
    template void launch_apply<T>(hipStream_t, const Dims &, const SysCoef<T> &, const GhostRows<T> &,             \
                                  const uint8_t *, const T *, T *, int, const T *, double *);                      \
    template void launch_residual<T>(hipStream_t, const Dims &, const SysCoef<T> &, const GhostRows<T> &,          \
                                     const uint8_t *, const T *, const T *, T *, const NormSpec &, double *);       \
    template void launch_jacobi<T>(hipStream_t, const Dims &, const SysCoef<T> &, const GhostRows<T> &,            \
                                   const uint8_t *, const T *, T *);                                               \
    template void launch_bicg_restart<T>(hipStream_t, const Dims &, const T *, T *, T *, T *);                \
    template void launch_bicg_init<T>(hipStream_t, const Dims &, T *, T *, T *, T *, T *, const T *,               \
                                      const NormSpec &, double *);                                                \
    template void launch_bicg_p<T>(hipStream_t, const Dims &, const KScal *, const T *, const T *, T *);           \
    template void launch_bicg_s<T>(hipStream_t, const Dims &, const KScal *, const T *, const T *, T *);           \
    template void launch_bicg_jacobi<T>(hipStream_t, const Dims &, const SysCoef<T> &, const GhostRows<T> &,       \
                                        const uint8_t *, const KScal *, int, const T *, const T *, T *, T *);      \
    template void launch_bicg_xr<T>(hipStream_t, const Dims &, const KScal *, const T *, const T *, const T *,     \
                                    const T *, const T *, T *, T *, const NormSpec &, double *);                   \
    template void launch_zero<T>(hipStream_t, const Dims &, T *);                                                  \
    template void launch_copy<T>(hipStream_t, const Dims &, const T *, T *);                                       \
    template void launch_restrict<T>(hipStream_t, const Dims &, const Dims &, const MGTransfer &, const uint8_t *, \
                                     const T *, const double *, const double *, const double *, const double *,    \
                                     const double *, const double *, T *, bool, const SysCoef<T> *,                \
                                     const uint8_t *, T *);                                                        \
    template void launch_prolong<T>(hipStream_t, const Dims &, const Dims &, const MGTransfer &, const uint8_t *,  \
                                    const uint8_t *, const T *, T *);                                              \
    template void launch_coarse_solve<T>(hipStream_t, const Dims &, const SysCoef<T> &, const uint8_t *,           \
                                         const T *, T *, int, int, const double *);
INST(double)
INST(float)

}  // namespace ibm
