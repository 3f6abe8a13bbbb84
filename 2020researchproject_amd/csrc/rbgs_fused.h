// rbgs_fused.h - fused out-of-place red-black Gauss-Seidel sweep (device
// code shared by the HIP build and the host emulation test, which compiles
// this file with stand-ins for the thread indices, LDS and barriers).
#pragma once

#include "common.h"
#include "stencil_terms.h"

namespace ibm {

// ---------------------------------------------------------------------------
// Fused red-black Gauss-Seidel sweep, out of place: xb = one full sweep
// (colour 0, then colour 1) applied to xa; identical to the two one-colour
// launches on the same input. Each workgroup owns a kRbTX x kRbTY column
// tile and marches along z over its chunk of planes. Three planes of the old
// field (tile + 2-cell halo) and three planes of the half-updated field
// (colour 0 done; tile + 1-cell halo) live in LDS: colour 0 of plane k-1 is
// computed once the old plane k is loaded, colour 1 of plane k-2 once the
// half-updated plane k-1 exists. Reading only xa and writing only xb keeps
// the tiles independent (halo cells are recomputed, never exchanged).
// Cells outside the interior are copied from xa (zero_in: xa is taken as 0).
// ---------------------------------------------------------------------------
constexpr int kRbTX = 64, kRbTY = 8;
constexpr int kRbPX = kRbTX + 4, kRbPY = kRbTY + 4;  // old planes, halo 2
constexpr int kRbHX = kRbTX + 2, kRbHY = kRbTY + 2;  // half-updated planes, halo 1

__device__ inline int wrapi(int g, int n) {
    const int r = g % n;
    return r < 0 ? r + n : r;
}

constexpr int kRbPN = kRbPY * kRbPX, kRbHN = kRbHY * kRbHX, kRbTN = kRbTY * kRbTX;
constexpr int kRbMaxChunk = 64;  // planes per workgroup (z coefficients staged in LDS)
// lanes work on pairs of x-neighbours (2t, 2t+1): one cell of each pair has
// the colour being updated, the other is copied, so no lane idles
constexpr int kRbHPX = kRbHX / 2, kRbTPX = kRbTX / 2;
constexpr int kRbHP = kRbHY * kRbHPX, kRbTP = kRbTY * kRbTPX;
constexpr int kRbNP = (kRbPN + 255) / 256, kRbNH = (kRbHP + 255) / 256, kRbNF = (kRbTP + 255) / 256;
static_assert(kRbHX % 2 == 0 && kRbTX % 2 == 0, "pair mapping needs even tile widths");

// Software pipelining: while planes k-1 and k-2 are computed from LDS, the
// next old plane (x, flags) and the b values of the next step's two phases
// are already in flight into registers, so the HBM latency of one plane
// overlaps the arithmetic of the previous one. The compute phases issue no
// global loads (the stencil coefficients of the tile are staged in LDS):
// the in-order vector-memory counter would otherwise make them wait for the
// prefetches too. Each lane handles fixed in-plane positions (tid + 256 m)
// of every phase; their addresses are computed once.
template <typename T>
__global__ void __launch_bounds__(256) k_rbgs_fused(Dims d, SysCoef<T> S, const uint8_t *flag, const T *b,
                                                    const T *xa, T *xb, int zero_in, int kchunk) {
    // old values of planes k-2..k (slot p mod 3), flags of planes k-3..k
    // (slot p mod 4), half-updated values of planes k-3..k-1 (slot p mod 3);
    // plane indices p >= -3, so the slot formulas add 3 or 4 first
    __shared__ T xs[3][kRbPN];
    __shared__ uint8_t fs[4][kRbPN];
    __shared__ T hs[3][kRbHN];
    // stencil coefficients (cm, cp, dm, dp) of the tile's halo-1 columns,
    // rows and planes kb-1..ke
    __shared__ T cxs[kRbHX][4], cys[kRbHY][4], czs[kRbMaxChunk + 2][4];
    const int tid = threadIdx.x;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int i0 = blockIdx.x * kRbTX, j0 = blockIdx.y * kRbTY;
    const int kb = blockIdx.z * kchunk, ke = kb + kchunk < nz ? kb + kchunk : nz;
    const bool px = S.ax[0].periodic, py = S.ax[1].periodic, pz = S.ax[2].periodic;
    const bool sk = S.skip;

    {
        auto stage = [&](const AxisCoef<T> &A, int g, int n, bool per, T(&dst)[4]) {
            int c = g;
            bool ok = true;
            if (per) c = wrapi(g, n); else ok = g >= 0 && g < n;
            for (int v = 0; v < 4; v++) dst[v] = ok ? A.c4[4 * (long)c + v] : T(0);
        };
        for (int e = tid; e < kRbHX + kRbHY + kchunk + 2; e += 256) {
            if (e < kRbHX) stage(S.ax[0], i0 - 1 + e, nx, px, cxs[e]);
            else if (e < kRbHX + kRbHY) stage(S.ax[1], j0 - 1 + (e - kRbHX), ny, py, cys[e - kRbHX]);
            else stage(S.ax[2], kb - 1 + (e - kRbHX - kRbHY), nz, pz, czs[e - kRbHX - kRbHY]);
        }
    }

    // old-plane positions of this lane (tile + 2-cell halo)
    long poff[kRbNP];
    bool pok[kRbNP];
#pragma unroll
    for (int m = 0; m < kRbNP; m++) {
        const int e = tid + 256 * m;
        const int li = e % kRbPX, lj = e / kRbPX;
        int gi = i0 - 2 + li, gj = j0 - 2 + lj;
        bool ok = e < kRbPN;
        if (px) gi = wrapi(gi, nx); else ok = ok && gi <= nx + 1;
        if (py) gj = wrapi(gj, ny); else ok = ok && gj <= ny + 1;
        pok[m] = ok;
        poff[m] = ok ? gi + (long)gj * d.sy : 0;
    }
    // colour-0 pairs of this lane (tile + 1-cell halo): first cell hx0 of the
    // pair in the halo-1 frame, row hy; interior coordinates of both cells
    int hx0[kRbNH], hy[kRbNH], hci[kRbNH][2], hcj[kRbNH];
    bool hok[kRbNH], hin[kRbNH][2];
#pragma unroll
    for (int m = 0; m < kRbNH; m++) {
        const int p = tid + 256 * m;
        hok[m] = p < kRbHP;
        hy[m] = p / kRbHPX;
        hx0[m] = 2 * (p % kRbHPX);
        const int gj = j0 - 1 + hy[m];
        bool yin = true;
        int cj = gj;
        if (py) cj = wrapi(gj, ny); else yin = gj >= 0 && gj < ny;
        hcj[m] = cj;
        for (int c = 0; c < 2; c++) {
            const int gi = i0 - 1 + hx0[m] + c;
            bool in = hok[m] && yin;
            int ci = gi;
            if (px) ci = wrapi(gi, nx); else in = in && gi >= 0 && gi < nx;
            hci[m][c] = ci;
            hin[m][c] = in;
        }
    }
    // colour-1 pairs of this lane (tile): first cell (fx0, fy) in the tile
    int fx0[kRbNF], fy[kRbNF];
    bool fin[kRbNF][2];
#pragma unroll
    for (int m = 0; m < kRbNF; m++) {
        const int p = tid + 256 * m;
        fy[m] = p / kRbTPX;
        fx0[m] = 2 * (p % kRbTPX);
        for (int c = 0; c < 2; c++) fin[m][c] = p < kRbTP && i0 + fx0[m] + c < nx && j0 + fy[m] < ny;
    }

    T xr[kRbNP], bh[kRbNH], bf[kRbNF];
    uint8_t fr[kRbNP];
    auto load_plane = [&](int k) {
        const int wk = pz ? wrapi(k, nz) : k;  // non-periodic: outer planes -2..nz+1 exist
        const long base = d.off + (long)wk * d.sz;
#pragma unroll
        for (int m = 0; m < kRbNP; m++) {
            T v = 0;
            uint8_t f = 0;
            if (pok[m]) {
                f = flag[base + poff[m]];
                if (!zero_in) v = xa[base + poff[m]];
            }
            xr[m] = v;
            fr[m] = f;
        }
    };
    // colour-0 cell of pair m on plane kh: 0 or 1 (parity of the unwrapped
    // coordinates; equal to the wrapped parity on even periodic axes)
    auto c0_of = [&](int m, int kh) { return (i0 - 1 + hx0[m] + j0 - 1 + hy[m] + kh) & 1; };
    // colour-1 cell of pair m on plane kf
    auto c1_of = [&](int m, int kf) { return ((i0 + fx0[m] + j0 + fy[m] + kf) & 1) ^ 1; };
    // b at the colour-0 cells of plane kh (tile + 1-cell halo)
    auto load_bh = [&](int kh, T (&out)[kRbNH]) {
        const bool kin = pz || (kh >= 0 && kh < nz);
        const int ck = pz ? wrapi(kh, nz) : kh;
#pragma unroll
        for (int m = 0; m < kRbNH; m++) {
            const int c = c0_of(m, kh);
            out[m] = (hin[m][c] && kin) ? b[cidx(d, hci[m][c], hcj[m], ck)] : T(0);
        }
    };
    // b at the colour-1 cells of plane kf of the tile
    auto load_bf = [&](int kf, T (&out)[kRbNF]) {
#pragma unroll
        for (int m = 0; m < kRbNF; m++) {
            const int c = c1_of(m, kf);
            out[m] = (fin[m][c] && kf >= kb && kf < ke) ? b[cidx(d, i0 + fx0[m] + c, j0 + fy[m], kf)] : T(0);
        }
    };

    load_plane(kb - 2);
#pragma unroll
    for (int m = 0; m < kRbNH; m++) bh[m] = 0;
#pragma unroll
    for (int m = 0; m < kRbNF; m++) bf[m] = 0;

    for (int k = kb - 2; k <= ke + 1; k++) {
        // (1) old plane k (2-cell halo) from registers into LDS; then start
        // the loads of the next step
        {
            const int xsl = (k + 3) % 3, fsl = (k + 4) % 4;
#pragma unroll
            for (int m = 0; m < kRbNP; m++) {
                const int e = tid + 256 * m;
                if (e < kRbPN) {
                    xs[xsl][e] = xr[m];
                    fs[fsl][e] = fr[m];
                }
            }
        }
        T bhn[kRbNH], bfn[kRbNF];
        if (k + 1 <= ke + 1) load_plane(k + 1);
        load_bh(k, bhn);      // phase (2) of the next step works on plane k
        load_bf(k - 1, bfn);  // phase (3) of the next step works on plane k - 1
        __syncthreads();
        // (2) colour 0 of plane kh = k - 1 over the tile and a 1-cell halo
        const int kh = k - 1;
        if (kh >= kb - 1) {
            const int x0 = (kh + 3) % 3, xm = (kh + 2) % 3, xp = (kh + 4) % 3;
            const int f0 = (kh + 4) % 4, fm = (kh + 3) % 4, fp = (kh + 5) % 4;
            const bool kin = pz || (kh >= 0 && kh < nz);
            const int ck = pz ? wrapi(kh, nz) : kh;
            const T *cz = czs[kh - kb + 1];
#pragma unroll
            for (int m = 0; m < kRbNH; m++) {
                if (!hok[m]) continue;
                const int cc = c0_of(m, kh);
                const int hi = hx0[m] + cc, hj = hy[m];
                const int c = (hj + 1) * kRbPX + hi + 1;
                const int e = hj * kRbHX + hi;
                // the other cell of the pair keeps its old value
                hs[x0][e + 1 - 2 * cc] = xs[x0][c + 1 - 2 * cc];
                T v = xs[x0][c];
                if (hin[m][cc] && kin) {
                    const int ci = hci[m][cc], cj = hcj[m];
                    const long q = cidx(d, ci, cj, ck);
                    const uint8_t f = fs[f0][c];
                    if (q == S.pin || f == F_SOLID) {
                        v = bh[m];
                    } else if (f == F_FLUID) {
                        T diag = S.c0, off = 0;
                        const T *cx = cxs[hi], *cy = cys[hj];
                        axis_term4(cx[0], cx[1], cx[2], cx[3], sk && (ci > 0 || px) && fs[f0][c - 1] == F_SOLID,
                                   sk && (ci < nx - 1 || px) && fs[f0][c + 1] == F_SOLID, xs[x0][c - 1],
                                   xs[x0][c + 1], diag, off);
                        axis_term4(cy[0], cy[1], cy[2], cy[3], sk && (cj > 0 || py) && fs[f0][c - kRbPX] == F_SOLID,
                                   sk && (cj < ny - 1 || py) && fs[f0][c + kRbPX] == F_SOLID, xs[x0][c - kRbPX],
                                   xs[x0][c + kRbPX], diag, off);
                        axis_term4(cz[0], cz[1], cz[2], cz[3], sk && (ck > 0 || pz) && fs[fm][c] == F_SOLID,
                                   sk && (ck < nz - 1 || pz) && fs[fp][c] == F_SOLID, xs[xm][c], xs[xp][c], diag,
                                   off);
                        v = diag != T(0) ? (bh[m] + off) / diag : T(0);
                    }
                }
                hs[x0][e] = v;
            }
        }
        __syncthreads();
        // (3) colour 1 of plane kf = k - 2 over the tile, written to xb
        const int kf = k - 2;
        if (kf >= kb && kf < ke) {
            const int h0 = (kf + 3) % 3, hm = (kf + 2) % 3, hp = (kf + 4) % 3;
            const int f0 = (kf + 4) % 4, fm = (kf + 3) % 4, fp = (kf + 5) % 4;
            const T *cz = czs[kf - kb + 1];
#pragma unroll
            for (int m = 0; m < kRbNF; m++) {
                const int c1 = c1_of(m, kf);
                const int tj = fy[m], gj = j0 + tj;
                // the colour-0 cell of the pair: its half-updated value is final
                if (fin[m][1 - c1]) {
                    const int ti = fx0[m] + 1 - c1;
                    xb[cidx(d, i0 + ti, gj, kf)] = hs[h0][(tj + 1) * kRbHX + ti + 1];
                }
                if (!fin[m][c1]) continue;
                const int ti = fx0[m] + c1, gi = i0 + ti;
                const int h = (tj + 1) * kRbHX + ti + 1;
                T v = hs[h0][h];
                const long q = cidx(d, gi, gj, kf);
                const int c = (tj + 2) * kRbPX + ti + 2;
                const uint8_t f = fs[f0][c];
                if (q == S.pin || f == F_SOLID) {
                    v = bf[m];
                } else if (f == F_FLUID) {
                    T diag = S.c0, off = 0;
                    const T *cx = cxs[ti + 1], *cy = cys[tj + 1];
                    axis_term4(cx[0], cx[1], cx[2], cx[3], sk && (gi > 0 || px) && fs[f0][c - 1] == F_SOLID,
                               sk && (gi < nx - 1 || px) && fs[f0][c + 1] == F_SOLID, hs[h0][h - 1], hs[h0][h + 1],
                               diag, off);
                    axis_term4(cy[0], cy[1], cy[2], cy[3], sk && (gj > 0 || py) && fs[f0][c - kRbPX] == F_SOLID,
                               sk && (gj < ny - 1 || py) && fs[f0][c + kRbPX] == F_SOLID, hs[h0][h - kRbHX],
                               hs[h0][h + kRbHX], diag, off);
                    axis_term4(cz[0], cz[1], cz[2], cz[3], sk && (kf > 0 || pz) && fs[fm][c] == F_SOLID,
                               sk && (kf < nz - 1 || pz) && fs[fp][c] == F_SOLID, hs[hm][h], hs[hp][h], diag, off);
                    v = diag != T(0) ? (bf[m] + off) / diag : T(0);
                }
                xb[q] = v;
            }
        }
        __syncthreads();  // the slots read above are refilled next step
#pragma unroll
        for (int m = 0; m < kRbNH; m++) bh[m] = bhn[m];
#pragma unroll
        for (int m = 0; m < kRbNF; m++) bf[m] = bfn[m];
    }
}

}  // namespace ibm
