// solver_kernels.h - matrix-free operators, smoothers, Krylov vector kernels
// and multigrid transfer kernels (templated on FP64 / FP32).
#pragma once

#include "common.h"

namespace ibm {

// Ghost rows of one system in compact form (struct of arrays).
template <typename T>
struct GhostRows {
    int ng;
    const long *cell;   // ng
    const long *src;    // 8 * ng (-1 = unused)
    const T *w;         // 8 * ng
    const T *wself;     // ng
    const int *srcg;    // 8 * ng: ghost row of the source cell, -1 for a fluid source
    T *vals[2];         // ng each: ghost values of the last two relaxations
    // the same weights in FP64: the smoothers' ghost-row relaxation forms its
    // weighted sum and divides by 1 -/+ w_self in FP64 whatever T is, since a
    // nearly degenerate stencil (w_self close to 1) would amplify the
    // rounding of FP32 weights
    const double *w64, *ws64;
};

// Device-resident scalars of BiCGSTAB.
struct KScal {
    double rho, rho1, alpha, omega, beta;
    double rr, rrn, bb, bbn;
    double ts, tt, rv;
    int breakdown;
    int pad;
};

// Which per-row scale the normalised residual uses: none, or 1 / sum K on
// fluid rows (the reference's diagonal-normalised pressure rows).
struct NormSpec {
    int on;
    const double *ksx, *ksy, *ksz;  // [n] sums of the unfolded K per axis
    const uint8_t *flag;
    long pin;
};

template <typename T>
void launch_apply(hipStream_t st, const Dims &d, const SysCoef<T> &S, const GhostRows<T> &G, const uint8_t *flag,
                  const T *x, T *y, int mode, const T *w, double *partial);
long apply_partials(const Dims &d, int ng);

template <typename T>
void launch_residual(hipStream_t st, const Dims &d, const SysCoef<T> &S, const GhostRows<T> &G,
                     const uint8_t *flag, const T *b, const T *x, T *r, const NormSpec &ns, double *partial);

template <typename T>
void launch_jacobi(hipStream_t st, const Dims &d, const SysCoef<T> &S, const GhostRows<T> &G, const uint8_t *flag,
                   const T *r, T *z);

// BiCGSTAB direction update fused with the Jacobi preconditioner:
// which 0: p = r + beta (p - omega v), ph = D^-1 p; which 1: p = r - alpha v
// (the s vector), ph = D^-1 p; which 2: p = r (the first direction, equal to
// which 0 with beta = 0), ph = D^-1 p. Results equal launch_bicg_p / _s +
// launch_jacobi.
template <typename T>
void launch_bicg_jacobi(hipStream_t st, const Dims &d, const SysCoef<T> &S, const GhostRows<T> &G,
                        const uint8_t *flag, const KScal *ks, int which, const T *r, const T *v, T *p, T *ph);

// launch_bicg_jacobi<double> with ph stored in FP32 (rounded once: the
// operator and the update then use exactly these values, so the iteration's
// residual recurrence stays that of the iterate)
void launch_bicg_jacobi_f32(hipStream_t st, const Dims &d, const SysCoef<double> &S, const GhostRows<double> &G,
                            const uint8_t *flag, const KScal *ks, int which, const double *r, const double *v,
                            double *p, float *ph);

// BiCGSTAB direction update in FP64 (which 0: p = r + beta (p - omega v);
// 1: p = r - alpha v, the s vector; 2: p = r) that also writes b32 = (float) p and,
// when x32 is not null, the first smoother colour on b32 with system *S
// (the values launch_d2f_first writes).
void launch_bicg_d2f(hipStream_t st, const Dims &d, const SysCoef<float> *S, const uint8_t *flag, const KScal *ks,
                     int which, const double *r, const double *v, double *p, float *b32, float *x32);
// y = A x for x stored in FP32 (x32): as launch_apply (modes 1 and 2) on x =
// (double) x32, which is also written to x unless x is null.
void launch_apply_f32in(hipStream_t st, const Dims &d, const SysCoef<double> &S, const GhostRows<double> &G,
                        const uint8_t *flag, const float *x32, double *x, double *y, int mode, const double *w,
                        double *partial);

// one colour of red-black Gauss-Seidel on the fluid rows (cells with
// (i + j + k) % 2 == color)
template <typename T>
void launch_rbgs_color(hipStream_t st, const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b, T *x,
                       int color, int zero_in);
// one full red-black sweep out of place: xb = sweep(xa) (zero_in: xa = 0);
// cells outside the fluid/solid rows are copied
template <typename T>
void launch_rbgs_fused(hipStream_t st, const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b,
                       const T *xa, T *xb, int zero_in);
// Jacobi relaxation of the ghost rows in one launch. `prev` names where the
// current ghost values are: -1 all zero, 0 or 1 G.vals[prev] (they equal x
// at the ghost cells). The new values go to x and to the other buffer, whose
// index is returned: the next relaxation's `prev`. only >= 0: only the rows
// whose cell (i, j, k) has i + j + k % 2 == only are relaxed (the others
// carry their values over); d is then the grid of the cells.
template <typename T>
int launch_ghost_relax(hipStream_t st, const GhostRows<T> &G, int is_p, const T *b, T *x, int prev,
                       Dims d = Dims{}, int only = -1);
// One colour (not from a zero start) followed by the relaxation of that
// colour's ghost rows (launch_ghost_relax with only = color), in one launch;
// prev >= 0 as for launch_ghost_relax, whose return value it shares.
template <typename T>
int launch_rbgs_color_ghost(hipStream_t st, const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b,
                            T *x, int color, const GhostRows<T> &G, int prev);

// BiCGSTAB pieces
template <typename T>
// x = 0, r = rh = b, p = v = 0 (p and v left alone when p is null: the
// first direction is then formed with which = 2 below); partial
// (vec_partials blocks x 2): the norms of r = b, ||r||^2 and ||N r||^2, as
// launch_residual gives for x = 0
void launch_bicg_init(hipStream_t st, const Dims &d, T *r, T *rh, T *p, T *v, T *x, const T *b, const NormSpec &ns,
                      double *partial);
template <typename T>
void launch_bicg_restart(hipStream_t st, const Dims &d, const T *r, T *rh, T *p, T *v);
template <typename T>
void launch_bicg_p(hipStream_t st, const Dims &d, const KScal *ks, const T *r, const T *v, T *p);
template <typename T>
void launch_bicg_s(hipStream_t st, const Dims &d, const KScal *ks, const T *r, const T *v, T *s);
template <typename T>
void launch_bicg_xr(hipStream_t st, const Dims &d, const KScal *ks, const T *ph, const T *sh, const T *s,
                    const T *t, const T *rh, T *x, T *r, const NormSpec &ns, double *partial);
long vec_partials(const Dims &d);
// launch_bicg_xr<double> with the preconditioned directions stored in FP32
// (the FP32 multigrid outputs; converted exactly on load)
void launch_bicg_xr_f32dir(hipStream_t st, const Dims &d, const KScal *ks, const float *ph, const float *sh,
                           const double *s, const double *t, const double *rh, double *x, double *r,
                           const NormSpec &ns, double *partial);

// Restarted GMRES, PCG and stationary multigrid: up to kVecSetMax vectors
// per multi-dot / multi-update launch.
constexpr int kVecSetMax = 12;
template <typename T>
struct VecSet {
    const T *v[kVecSetMax];
};
// partial-sum blocks of the kernels below (scratch of kVecSetMax + 1 values each)
long krylov_blocks(const Dims &d);
// out[a] = (w, V_a) for a < nv, out[kVecSetMax] = (w, w); scale: w <- N w
// first (N: the normalised-row scale of ns)
template <typename T>
void launch_gm_dots(hipStream_t st, const Dims &d, const VecSet<T> &V, int nv, T *w, const NormSpec &ns, int scale,
                    double *partial, double *out);
// w <- w - sum_{a < nv} c_a V_a (c on the device); out[0] = (w, w)
template <typename T>
void launch_gm_update(hipStream_t st, const Dims &d, const VecSet<T> &V, int nv, const double *c, T *w,
                      double *partial, double *out);
// v = f w, or (unscale) v = w / nscale
template <typename T>
void launch_gm_scale(hipStream_t st, const Dims &d, const T *w, double f, T *v, const NormSpec &ns, int unscale);
// x += a p, r -= a q; out = (r, r), (N r, N r)
template <typename T>
void launch_pcg_update(hipStream_t st, const Dims &d, double a, const T *p, const T *q, T *x, T *r,
                       const NormSpec &ns, double *partial, double *out);
template <typename T>
void launch_dot(hipStream_t st, const Dims &d, const T *a, const T *b, double *partial, double *out);
// y = z + f y, or (axpy) y = y + f z
template <typename T>
void launch_lincomb(hipStream_t st, const Dims &d, const T *z, double f, T *y, int axpy);

// finalize ops: 0 = alpha from dot(rh, v); 1 = omega from (t.s, t.t);
// 2 = rho/beta/norms after the x, r update; 3 = store b norms and rho1 at
// start; 4 = store residual norms only. pub (device-mapped host memory, or
// null): the scalars are also copied there, then *pub_epoch = epoch.
void launch_finalize(hipStream_t st, int op, const double *partial, long nblk, int nv, KScal *ks,
                     KScal *pub = nullptr, unsigned *pub_epoch = nullptr, unsigned epoch = 0);
// out[0 .. nv) = in[0 .. nv) into device-mapped host memory, then *epoch_out
// = epoch (after a system-scope fence), for a host thread waiting on it
void launch_publish(hipStream_t st, const double *in, int nv, double *out, unsigned *epoch_out, unsigned epoch);
// out[a] = sum (or max) over blocks of partial[b * nv + a]
void launch_sum_partials(hipStream_t st, const double *partial, long nblk, int nv, double *out, int reduce_max);

// precision conversion / accumulation over interior cells
void launch_d2f(hipStream_t st, const Dims &d, const double *a, float *b);
// b = (float) a and x = the first smoother colour from a zero start on b
void launch_d2f_first(hipStream_t st, const Dims &d, const SysCoef<float> &S, const uint8_t *flag, const double *a,
                      float *b, float *x);
void launch_add_f2d(hipStream_t st, const Dims &d, const float *e, double *x);
void launch_f2d(hipStream_t st, const Dims &d, const float *e, double *x);
template <typename T>
void launch_zero(hipStream_t st, const Dims &d, T *a);
template <typename T>
void launch_copy(hipStream_t st, const Dims &d, const T *a, T *b);

// Multigrid transfers between a fine grid df and a coarse grid dc. Each axis
// is coarsened independently: a coarse cell covers one or two consecutive
// fine cells along every axis.
struct MGTransfer {
    const int *ci, *cj, *ck;      // fine index -> coarse (parent) index
    const int *si, *sj, *sk;      // coarse index -> first fine child (n_coarse + 1 entries)
    const int *pi, *pj, *pk;      // fine index -> far coarse neighbour index (or the parent)
    const double *wi, *wj, *wk;   // weight of the parent (1 - weight of the far neighbour)
    const uint8_t *pm;            // per fine cell: 1 when one of its eight coarse cells is inactive
};

// r = b - A x on the fluid rows (no norms)
// returns whether r was set to 0 at the non-fluid cells (else left as it was)
template <typename T>
bool launch_mg_residual(hipStream_t st, const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b,
                        const T *x, T *r);
// bc = restriction of the fluid rows' residual b - A x (one pass, no r).
// xc (with the coarse system Sc and flags flag_c), when not null, receives
// the first colour of the coarse smoother from a zero start (the values
// launch_rbgs_color(..., color 0, zero_in 1) would write).
template <typename T>
void launch_resid_restrict(hipStream_t st, const Dims &df, const SysCoef<T> &S, const uint8_t *flag_f, const T *b,
                           const T *x, const Dims &dc, const MGTransfer &tr, const double *dxyz_f,
                           const double *dxyz_c, T *bc, const SysCoef<T> *Sc = nullptr,
                           const uint8_t *flag_c = nullptr, T *xc = nullptr);
// rf_zero_off_fluid: rf is 0 at non-fluid cells, so no flags are read;
// Sc, flag_c, xc as for launch_resid_restrict
template <typename T>
void launch_restrict(hipStream_t st, const Dims &df, const Dims &dc, const MGTransfer &tr, const uint8_t *flag_f,
                     const T *rf, const double *dxf, const double *dyf, const double *dzf, const double *dxc,
                     const double *dyc, const double *dzc, T *bc,
                     bool rf_zero_off_fluid, const SysCoef<T> *Sc = nullptr, const uint8_t *flag_c = nullptr,
                     T *xc = nullptr);
template <typename T>
void launch_prolong(hipStream_t st, const Dims &df, const Dims &dc, const MGTransfer &tr, const uint8_t *flag_f,
                    const uint8_t *flag_c, const T *ec, T *xf);
template <typename T>
void launch_coarse_solve(hipStream_t st, const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b, T *x,
                         int sweeps, int singular, const double *dxyz);

// In-place Gauss-Jordan inversion without pivoting of the n x n row-major
// matrix A (diagonally dominant M-matrix rows); rowp, colp: n doubles of
// scratch.
void gauss_jordan_invert(hipStream_t st, double *A, int n, double *rowp, double *colp);
void launch_d2f_flat(hipStream_t st, const double *a, float *b, long n);
// x[rows[r]] = sum_c Ainv[r][c] rhs[c], rhs[c] = b[rows[c]]; singular: the
// volume-weighted mean of b over fluid rows is removed first and the pinned
// row's right-hand side is zero.
template <typename T>
void launch_dense_solve(hipStream_t st, const T *Ainv, int n, const long *rows, const double *vol,
                        const uint8_t *flag, int singular, int pinrow, const T *b, T *x);

}  // namespace ibm
