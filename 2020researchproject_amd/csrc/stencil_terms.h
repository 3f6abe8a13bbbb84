// stencil_terms.h - per-axis terms of the 7-point operator (device code
// shared by the HIP kernels and the host emulation tests).
#pragma once

#include "common.h"

namespace ibm {

// One axis of a fluid row: diag += dm + dp, off += cm x_m + cp x_p, with the
// faces to solid neighbours dropped (sm / sp: neighbour is F_SOLID and the
// system skips such faces).
template <typename T>
__device__ inline void axis_term4(T cm, T cp, T dm, T dp, bool sm, bool sp, T xm, T xp, T &diag, T &off) {
    if (sm) { cm = 0; dm = 0; }
    if (sp) { cp = 0; dp = 0; }
    diag += dm + dp;
    off += cm * xm + cp * xp;
}

template <typename T>
__device__ inline void axis_term(const AxisCoef<T> &A, int ii, bool sm, bool sp, T xm, T xp, T &diag, T &off) {
    const typename Packed4<T>::type c = reinterpret_cast<const typename Packed4<T>::type *>(A.c4)[ii];
    axis_term4(c.x, c.y, c.z, c.w, sm, sp, xm, xp, diag, off);
}

// New value of a fluid row from its old value xo and its Gauss-Seidel value
// gs: successive over-relaxation xo + om (gs - xo), and gs itself (bit for
// bit) for om = 1.
template <typename T>
__device__ inline T sor_value(T om, T xo, T gs) {
    return om == T(1) ? gs : xo + om * (gs - xo);
}

}  // namespace ibm
