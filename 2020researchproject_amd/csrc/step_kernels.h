// step_kernels.h - launchers of the stencil kernels of one fractional step.
#pragma once

#include "common.h"

namespace ibm {

// Geometry of one grid: padded 1-D arrays indexed [i + 2] (i = -2 .. n+1).
struct Geo {
    Dims d;
    const double *dx, *dy, *dz, *xc, *yc, *zc;
    const double *kW, *kE, *kS, *kN, *kD, *kU;  // dt / (2 Re h dx), interior only
};

struct StepParams {
    double dt, Re;
    double force[3];       // constant body force
    const double *fa[3];   // per-cell body force (host-evaluated), or null
};

// van Driest damping of the Smagorinsky length: walls = direction bits
// (N 1, E 2, S 4, W 8, D 16, U 32) of the outer faces that count as walls;
// lo / hi: domain bounds; ure = u_tau * Re (1 / nu = Re).
struct VanDriest {
    int walls;
    double aplus, ure;
    double lo[3], hi[3];
};

struct TauPtrs {
    int on;
    double *t[6];          // 11, 12, 13, 22, 23, 33
    int slot[3][3];        // tau_{l,m} slot for row l, column m
};

enum BCType : int {
    BCT_VELOCITY = 0, BCT_PRESSURE = 1, BCT_STATIONARY = 2, BCT_FREE_SLIP = 3,
    BCT_SYMMETRIC = 4, BCT_ALL_PERIODIC = 5, BCT_VELOCITY_PERIODIC = 6, BCT_CONVECTIVE = 7
};

// Boundary data of one outer face. g[0..2] velocity components, g[3]
// pressure value (or pressure increment for the p' fill); garr, when set,
// holds [4][n1][n2] values over the face's padded tangential range.
struct FaceBC {
    int type;
    int n1, t1off, t2off;
    double g[4];
    const double *garr;
};

struct OuterParams {
    FaceBC face[3][2];  // [axis][low/high]
};

// Convective outflow face values of one face (bci = N,E,S,W,D,U index):
// init != 0 copies the first interior cell's velocity; otherwise one upwind
// step of dg/dt + Uc dg/dn = 0, with Uc = uc or (uc <= 0) the face's mean
// outward velocity, written to *uc_slot. garr holds 4 components x the
// face's tangential extent (the 4th, the pressure value, stays 0). Leaves
// per-block partial sums of 2 g_c^2 (the BC rows' share of |b|^2, 3 per
// block) in partial; returns the block count.
long conv_outflow_blocks(const Dims &d, int ax);
long launch_conv_outflow(hipStream_t st, const Geo &g, int ax, int hi, double uc, double dt, int init,
                         const double *u1, const double *u2, const double *u3, double *garr, double *uc_slot,
                         double *partial);

void launch_interp_stag(hipStream_t st, const Geo &g, const double *u1, const double *u2, const double *u3,
                        double *U1, double *U2, double *U3);
long conv_rhs_blocks(const Dims &d);
void launch_conv_rhs(hipStream_t st, const Geo &g, const StepParams &sp, const uint8_t *flag,
                     const double *const u[3], const double *const U[3], const double *p,
                     const double *const Np[3], double *const N[3], double *const b[3], const TauPtrs &tau,
                     double *partial);
void launch_face_star(hipStream_t st, const Geo &g, const StepParams &sp, const uint8_t *flag,
                      const double *const us[3], const double *p, double *const F[3]);
void launch_poisson_rhs(hipStream_t st, const Geo &g, const StepParams &sp, const uint8_t *flag, long pin,
                        const double *const F[3], double *b, double *partial);
void launch_update(hipStream_t st, const Geo &g, const StepParams &sp, const uint8_t *flag, const double *p,
                   const double *pp, const double *const us[3], const double *const F[3], double *pn,
                   double *const un[3], double *const G[3]);
void launch_outer(hipStream_t st, const Geo &g, const OuterParams &op, double *f1, double *f2, double *f3,
                  int mode, int comp);
void launch_bc_rhs(hipStream_t st, const Geo &g, const OuterParams &op, int ax, int hi, int comp,
                   const double *Kf, double gamma, double *b);
void launch_ghost_update(hipStream_t st, const GhostCell *gc, int ng, const int *level_start, int nlevels,
                         double *old, double *u1, double *u2, double *u3, double *p);
void launch_solid_nan(hipStream_t st, const Dims &d, const uint8_t *flag, double *u1, double *u2, double *u3,
                      double *p);
void launch_sgs(hipStream_t st, const Geo &g, int model, double coef, const VanDriest &vd, const uint8_t *flag,
                const double *lvset, const double *const u[3], const TauPtrs &tau);
long maxdiv_blocks(const Dims &d);
// body force from the faces between interior ghost cells (gijk: i, j, k of
// each, ng of them) and their interior fluid neighbours: partial[blk * 3 + a]
long body_force_blocks(int ng);
void launch_body_force(hipStream_t st, const Geo &g, const uint8_t *flag, const int *gijk, int ng, double mu,
                       const double *p, const double *const u[3], double *partial);
void launch_maxdiv(hipStream_t st, const Geo &g, const uint8_t *flag, long pin, const double *const U[3],
                   double *partial);

}  // namespace ibm
