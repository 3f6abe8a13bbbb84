// rbgs_quad.h - FP32 one-colour red-black Gauss-Seidel sweep, four cells
// per lane (device code shared by the HIP build and the host emulation test,
// which compiles this file with stand-ins for the thread indices).
#pragma once

#include "stencil_terms.h"

namespace ibm {

// component p + 2h of a quad: selects on a value (a runtime-indexed array,
// or a float4 whose address is taken, is placed in scratch or LDS memory)
__device__ inline float quad_pick(float4 v, int h, int p) { return h ? (p ? v.w : v.z) : (p ? v.y : v.x); }

// FP32 one-colour sweep with each lane owning four consecutive cells of a
// row (two of them of the colour): one 16-byte load per neighbouring row and
// plane, b and flags serve four cells per load, so a cell costs about half
// the load instructions of k_rbgs, which bounds that kernel (measured on a
// 256 x 128 x 128 level: 13.2 us against 15.4 us per launch). Same
// arithmetic as k_rbgs, so the results are identical. Loads past the last
// cell of a row stay inside the padded allocation and are not used. Body of
// one lane of a 64 x BY workgroup at block position bk.
template <int BY>
__device__ inline void rbgs_quad(const Dims &d, const SysCoef<float> &S, const uint8_t *flag,
                                 const float *b, float *x, int color, int zero_in, uint3 bk) {
    const int j = bk.y * BY + threadIdx.y, k = bk.z;
    const int i0 = 4 * (bk.x * 64 + threadIdx.x);
    if (i0 >= d.nx || j >= d.ny) return;
    const long q0 = cidx(d, i0, j, k);
    const int p = (j + k + color) & 1;  // i0 is even: the colour's cells are i0 + p and i0 + p + 2
    const AxisCoef<float> &AX = S.ax[0], &AY = S.ax[1], &AZ = S.ax[2];
    const float4 bb = *reinterpret_cast<const float4 *>(b + q0);
    const unsigned ff = *reinterpret_cast<const unsigned *>(flag + q0);
    const unsigned mm = S.skip ? *reinterpret_cast<const unsigned *>(S.fmask + q0) : 0u;
    const float4 cA = reinterpret_cast<const float4 *>(AX.c4)[i0 + p < d.nx ? i0 + p : d.nx - 1];
    const float4 cB = reinterpret_cast<const float4 *>(AX.c4)[i0 + p + 2 < d.nx ? i0 + p + 2 : d.nx - 1];
    const float4 cy = reinterpret_cast<const float4 *>(AY.c4)[j];
    const float4 cz = reinterpret_cast<const float4 *>(AZ.c4)[k];
    float4 xc = make_float4(0.f, 0.f, 0.f, 0.f), ym = xc, yp = xc, zm = xc, zp = xc;
    float xl = 0.f, xr = 0.f;
    if (!zero_in) {
        const int jm = (j == 0 && AY.periodic) ? AY.n - 1 : j - 1, jp = (j == AY.n - 1 && AY.periodic) ? 0 : j + 1;
        const int km = (k == 0 && AZ.periodic) ? AZ.n - 1 : k - 1, kp = (k == AZ.n - 1 && AZ.periodic) ? 0 : k + 1;
        xc = *reinterpret_cast<const float4 *>(x + q0);
        ym = *reinterpret_cast<const float4 *>(x + q0 + (long)(jm - j) * d.sy);
        yp = *reinterpret_cast<const float4 *>(x + q0 + (long)(jp - j) * d.sy);
        zm = *reinterpret_cast<const float4 *>(x + q0 + (long)(km - k) * d.sz);
        zp = *reinterpret_cast<const float4 *>(x + q0 + (long)(kp - k) * d.sz);
        xl = x[q0 - 1];
        xr = x[q0 + 4];
    }
    // periodic x: the neighbours across the wrap, from the row's other end
    float wrap_m = 0.f, wrap_p = 0.f;
    if (!zero_in && AX.periodic) {
        if (i0 == 0) wrap_m = x[q0 + AX.n - 1];
        if (i0 + 4 >= AX.n) wrap_p = x[q0 - i0];
    }
    // all loads are issued above; the two cells' values are formed, then stored
    float res[2] = {0.f, 0.f};
    bool put[2] = {false, false};
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int o = p + 2 * h, i = i0 + o;
        const unsigned f = (ff >> (8 * o)) & 0xFFu, m = (mm >> (8 * o)) & 0xFFu;
        const float4 cx = h ? cB : cA;
        float diag = S.c0, off = 0;
        if (zero_in) {
            axis_term4(cx.x, cx.y, cx.z, cx.w, (m & 1u) != 0, (m & 2u) != 0, 0.f, 0.f, diag, off);
            axis_term4(cy.x, cy.y, cy.z, cy.w, (m & 4u) != 0, (m & 8u) != 0, 0.f, 0.f, diag, off);
            axis_term4(cz.x, cz.y, cz.z, cz.w, (m & 16u) != 0, (m & 32u) != 0, 0.f, 0.f, diag, off);
            off = 0;
        } else {
            // x neighbours of cell i0 + o: quad entries o - 1 and o + 1 (the
            // cells beside the quad for o = 0 and 3)
            float xm = h ? (p ? xc.z : xc.y) : (p ? xc.x : xl);
            float xp = h ? (p ? xr : xc.w) : (p ? xc.z : xc.y);
            if (AX.periodic && i == 0) xm = wrap_m;
            if (AX.periodic && i == AX.n - 1) xp = wrap_p;
            axis_term4(cx.x, cx.y, cx.z, cx.w, (m & 1u) != 0, (m & 2u) != 0, xm, xp, diag, off);
            axis_term4(cy.x, cy.y, cy.z, cy.w, (m & 4u) != 0, (m & 8u) != 0, quad_pick(ym, h, p), quad_pick(yp, h, p), diag, off);
            axis_term4(cz.x, cz.y, cz.z, cz.w, (m & 16u) != 0, (m & 32u) != 0, quad_pick(zm, h, p), quad_pick(zp, h, p), diag, off);
        }
        const float bv = quad_pick(bb, h, p);
        const bool inside = i < d.nx;
        if (q0 + o == S.pin || f == F_SOLID) {
            res[h] = bv;
            put[h] = inside;
        } else if (f != F_FLUID) {
            put[h] = inside && zero_in;
        } else {
            // over-relaxed from the cell's old value (0 from a zero start)
            res[h] = sor_value(S.omega, quad_pick(xc, h, p), diag != 0.f ? (bv + off) / diag : 0.f);
            put[h] = inside;
        }
    }
    if (zero_in) {
        // the initial x is zero: clear the quad's cells of the other colour
        if (i0 + 1 - p < d.nx) x[q0 + 1 - p] = 0.f;
        if (i0 + 3 - p < d.nx) x[q0 + 3 - p] = 0.f;
    }
    if (put[0]) x[q0 + p] = res[0];
    if (put[1]) x[q0 + p + 2] = res[1];
}


// FP32 residual r = b - A x on the fluid rows (pinned row: b - x) with four
// consecutive cells of a row per lane, the loads of rbgs_quad, written as
// one 16-byte store; non-fluid cells get r = 0, so a restriction may sum
// every child without reading flags. Same arithmetic per fluid cell as the
// one-cell residual kernel.
template <int BY>
__device__ inline void resid_quad(const Dims &d, const SysCoef<float> &S, const uint8_t *flag, const float *b,
                                  const float *x, float *r, uint3 bk) {
    const int j = bk.y * BY + threadIdx.y, k = bk.z;
    const int i0 = 4 * (bk.x * 64 + threadIdx.x);
    if (i0 >= d.nx || j >= d.ny) return;
    const long q0 = cidx(d, i0, j, k);
    const AxisCoef<float> &AX = S.ax[0], &AY = S.ax[1], &AZ = S.ax[2];
    const int jm = (j == 0 && AY.periodic) ? AY.n - 1 : j - 1, jp = (j == AY.n - 1 && AY.periodic) ? 0 : j + 1;
    const int km = (k == 0 && AZ.periodic) ? AZ.n - 1 : k - 1, kp = (k == AZ.n - 1 && AZ.periodic) ? 0 : k + 1;
    const float4 bb = *reinterpret_cast<const float4 *>(b + q0);
    const unsigned ff = *reinterpret_cast<const unsigned *>(flag + q0);
    const unsigned mm = S.skip ? *reinterpret_cast<const unsigned *>(S.fmask + q0) : 0u;
    const float4 xc = *reinterpret_cast<const float4 *>(x + q0);
    const float4 ym = *reinterpret_cast<const float4 *>(x + q0 + (long)(jm - j) * d.sy);
    const float4 yp = *reinterpret_cast<const float4 *>(x + q0 + (long)(jp - j) * d.sy);
    const float4 zm = *reinterpret_cast<const float4 *>(x + q0 + (long)(km - k) * d.sz);
    const float4 zp = *reinterpret_cast<const float4 *>(x + q0 + (long)(kp - k) * d.sz);
    const float xl = x[q0 - 1], xr = x[q0 + 4];
    const float4 cy = reinterpret_cast<const float4 *>(AY.c4)[j];
    const float4 cz = reinterpret_cast<const float4 *>(AZ.c4)[k];
    float4 cxv[4];
#pragma unroll
    for (int o = 0; o < 4; o++) cxv[o] = reinterpret_cast<const float4 *>(AX.c4)[i0 + o < d.nx ? i0 + o : d.nx - 1];
    float wrap_m = 0.f, wrap_p = 0.f;
    if (AX.periodic) {
        if (i0 == 0) wrap_m = x[q0 + AX.n - 1];
        if (i0 + 4 >= AX.n) wrap_p = x[q0 - i0];
    }
    const float xs[6] = {xl, xc.x, xc.y, xc.z, xc.w, xr};
    const float ymv[4] = {ym.x, ym.y, ym.z, ym.w}, ypv[4] = {yp.x, yp.y, yp.z, yp.w};
    const float zmv[4] = {zm.x, zm.y, zm.z, zm.w}, zpv[4] = {zp.x, zp.y, zp.z, zp.w};
    const float bv[4] = {bb.x, bb.y, bb.z, bb.w};
    float rv[4];
#pragma unroll
    for (int o = 0; o < 4; o++) {
        // o is a compile-time index after unrolling (no runtime-indexed arrays)
        const int i = i0 + o;
        const unsigned m = (mm >> (8 * o)) & 0xFFu;
        float xm = xs[o], xp = xs[o + 2];
        if (AX.periodic && i == 0) xm = wrap_m;
        if (AX.periodic && i == AX.n - 1) xp = wrap_p;
        float diag = S.c0, off = 0;
        axis_term4(cxv[o].x, cxv[o].y, cxv[o].z, cxv[o].w, (m & 1u) != 0, (m & 2u) != 0, xm, xp, diag, off);
        axis_term4(cy.x, cy.y, cy.z, cy.w, (m & 4u) != 0, (m & 8u) != 0, ymv[o], ypv[o], diag, off);
        axis_term4(cz.x, cz.y, cz.z, cz.w, (m & 16u) != 0, (m & 32u) != 0, zmv[o], zpv[o], diag, off);
        const float xo = xs[o + 1];
        rv[o] = ((ff >> (8 * o)) & 0xFFu) != F_FLUID ? 0.f : q0 + o == S.pin ? bv[o] - xo : bv[o] - (diag * xo - off);
    }
    if (i0 + 3 < d.nx) {
        *reinterpret_cast<float4 *>(r + q0) = make_float4(rv[0], rv[1], rv[2], rv[3]);
    } else {
#pragma unroll
        for (int o = 0; o < 4; o++)
            if (i0 + o < d.nx) r[q0 + o] = rv[o];
    }
}
}  // namespace ibm
