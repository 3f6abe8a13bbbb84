// common.h - device layout and shared declarations of the MI355X IBM solver.
//
// Cell arrays are x-contiguous with two ghost layers per side. Row pitch `sy`
// is rounded up to 32 elements and the origin is shifted so that interior cell
// (0, j, k) starts a 128-byte line for both FP64 and FP32 arrays; a wavefront
// of 64 lanes then reads whole lines along x. Face velocities are stored in
// the same cell layout: FX(i,j,k) holds the face x_{i-1/2}, FY(i,j,k) the face
// y_{j-1/2}, FZ(i,j,k) the face z_{k-1/2}.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define HIP_CHECK(expr)                                          \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) ibm::hip_failure(e_, #expr, __FILE__, __LINE__); \
    } while (0)

namespace ibm {

// Prints the failed call. A failed allocation inside a serving job fails that
// job (it throws); every other HIP error aborts the process, since the device
// state of all cases is then in doubt.
[[noreturn]] void hip_failure(hipError_t e, const char *expr, const char *file, int line);

enum : uint8_t { F_FLUID = 1, F_GHOST = 2, F_SOLID = 3 };

constexpr int kAlign = 32;  // elements

struct Dims {
    int nx, ny, nz;   // interior cells
    int sy;           // row pitch (elements)
    long sz;          // plane pitch
    long off;         // linear index of cell (0,0,0)
    long ncell;       // allocation length (elements)
};

inline Dims make_dims(int nx, int ny, int nz) {
    Dims d;
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.sy = ((nx + 4 + kAlign - 1) / kAlign) * kAlign;  // i = -2..nx+1 fit in one pitch
    d.sz = (long)d.sy * (ny + 4);
    d.off = kAlign + 2L * d.sy + 2L * d.sz;
    d.ncell = d.off + (long)(nz + 2) * d.sz + 2L * kAlign;
    return d;
}

__host__ __device__ inline long cidx(const Dims &d, int i, int j, int k) {
    return d.off + i + (long)j * d.sy + (long)k * d.sz;
}

// One immersed-boundary ghost cell: mirror-point stencil with solid points
// dropped and weights renormalised (src/utils.c:111-171 and
// src/ibm3d_fracstep.c:1937-1963). Entries with w = 0 are padding.
struct GhostCell {
    long cell;         // linear index of the ghost cell
    long src[8];       // linear indices of the stencil points (self removed)
    double w[8];       // normalised weights
    double wself;      // normalised weight of the ghost cell itself
    int later[8];      // source is a ghost visited later in lexicographic order: slot of it, else -1
    int level;         // dependency level of the ordered in-place update
    int pad;
};

// Per-axis coefficients of a 7-point system after folding the outer boundary
// rows into the first/last interior cells. Fluid row:
//   (c0 + sum_axes (dm[i] + dp[i])) x_P - sum_axes (cm[i] x_minus + cp[i] x_plus) = b
// dm/dp are the diagonal shares of the minus/plus faces (they differ from
// cm/cp only at folded boundary faces). On coarse multigrid levels a face to
// an inactive (solid) neighbour is dropped from both sums (zero flux).
template <typename T>
struct AxisCoef {
    const T *cm, *cp, *dm, *dp;
    int n;
    int periodic;
    const T *c4;  // [n][4]: cm, cp, dm, dp packed (one vector load per axis)
};

template <typename T>
struct Packed4;
template <>
struct Packed4<float> {
    using type = float4;
};
template <>
struct Packed4<double> {
    using type = double4;
};

template <typename T>
struct SysCoef {
    AxisCoef<T> ax[3];
    T c0;           // 1 for momentum, 0 for pressure
    long pin;       // linear index of a pinned cell (identity row) or -1
    int is_p;       // pressure-type ghost rows ((1 - wself) x - sum w x = 0)
    int skip;       // drop faces to F_SOLID interior neighbours (coarse levels)
    const uint8_t *fmask;  // skip: per cell, bit 2a / 2a+1 set when the minus / plus
                           // neighbour along axis a is an interior F_SOLID cell
    T omega = T(1);        // red-black smoother: over-relaxation factor of the
                           // fluid rows (1: Gauss-Seidel); read by the smoother
                           // kernels only
};

}  // namespace ibm
