// rbgs_fused.h - fused out-of-place red-black Gauss-Seidel sweep (device
// code shared by the HIP build and the host emulation test, which compiles
// this file with stand-ins for the thread indices, LDS and barriers).
#pragma once

#include "common.h"

namespace ibm {

// One axis of a fluid row: diag += dm + dp, off += cm x_m + cp x_p, with the
// faces to solid neighbours dropped (sm / sp: neighbour is F_SOLID and the
// system skips such faces).
template <typename T>
__device__ inline void axis_term(const AxisCoef<T> &A, int ii, bool sm, bool sp, T xm, T xp, T &diag, T &off) {
    const typename Packed4<T>::type c = reinterpret_cast<const typename Packed4<T>::type *>(A.c4)[ii];
    T cm = c.x, cp = c.y, dm = c.z, dp = c.w;
    if (sm) { cm = 0; dm = 0; }
    if (sp) { cp = 0; dp = 0; }
    diag += dm + dp;
    off += cm * xm + cp * xp;
}

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

template <typename T>
__global__ void __launch_bounds__(256) k_rbgs_fused(Dims d, SysCoef<T> S, const uint8_t *flag, const T *b,
                                                    const T *xa, T *xb, int zero_in, int kchunk) {
    // old values of planes k-2..k (slot p mod 3), flags of planes k-3..k
    // (slot p mod 4), half-updated values of planes k-3..k-1 (slot p mod 3);
    // plane indices p >= -3, so the slot formulas add 3 or 4 first
    __shared__ T xs[3][kRbPY * kRbPX];
    __shared__ uint8_t fs[4][kRbPY * kRbPX];
    __shared__ T hs[3][kRbHY * kRbHX];
    const int tid = threadIdx.x;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int i0 = blockIdx.x * kRbTX, j0 = blockIdx.y * kRbTY;
    const int kb = blockIdx.z * kchunk, ke = kb + kchunk < nz ? kb + kchunk : nz;
    const bool px = S.ax[0].periodic, py = S.ax[1].periodic, pz = S.ax[2].periodic;
    const bool sk = S.skip;

    for (int k = kb - 2; k <= ke + 1; k++) {
        // (1) old plane k with a 2-cell halo
        {
            const int xsl = (k + 3) % 3, fsl = (k + 4) % 4;
            const int wk = pz ? wrapi(k, nz) : k;  // non-periodic: outer planes -2..nz+1 exist
            for (int e = tid; e < kRbPY * kRbPX; e += 256) {
                const int li = e % kRbPX, lj = e / kRbPX;
                int gi = i0 - 2 + li, gj = j0 - 2 + lj;
                bool ok = true;
                if (px) gi = wrapi(gi, nx); else ok = gi <= nx + 1;
                if (py) gj = wrapi(gj, ny); else ok = ok && gj <= ny + 1;
                T v = 0;
                uint8_t f = 0;
                if (ok) {
                    const long q = cidx(d, gi, gj, wk);
                    f = flag[q];
                    if (!zero_in) v = xa[q];
                }
                xs[xsl][e] = v;
                fs[fsl][e] = f;
            }
        }
        __syncthreads();
        // (2) colour 0 of plane kh = k - 1 over the tile and a 1-cell halo
        const int kh = k - 1;
        if (kh >= kb - 1) {
            const int x0 = (kh + 3) % 3, xm = (kh + 2) % 3, xp = (kh + 4) % 3;
            const int f0 = (kh + 4) % 4, fm = (kh + 3) % 4, fp = (kh + 5) % 4;
            const bool kin = pz || (kh >= 0 && kh < nz);
            const int ck = pz ? wrapi(kh, nz) : kh;
            for (int e = tid; e < kRbHY * kRbHX; e += 256) {
                const int hi = e % kRbHX, hj = e / kRbHX;
                const int c = (hj + 1) * kRbPX + hi + 1;
                T v = xs[x0][c];
                const int gi = i0 - 1 + hi, gj = j0 - 1 + hj;
                int ci = gi, cj = gj;
                bool in = kin;
                if (px) ci = wrapi(gi, nx); else in = in && gi >= 0 && gi < nx;
                if (py) cj = wrapi(gj, ny); else in = in && gj >= 0 && gj < ny;
                if (in && ((ci + cj + ck) & 1) == 0) {
                    const long q = cidx(d, ci, cj, ck);
                    const uint8_t f = fs[f0][c];
                    if (q == S.pin || f == F_SOLID) {
                        v = b[q];
                    } else if (f == F_FLUID) {
                        T diag = S.c0, off = 0;
                        axis_term(S.ax[0], ci, sk && (ci > 0 || px) && fs[f0][c - 1] == F_SOLID,
                                  sk && (ci < nx - 1 || px) && fs[f0][c + 1] == F_SOLID, xs[x0][c - 1], xs[x0][c + 1],
                                  diag, off);
                        axis_term(S.ax[1], cj, sk && (cj > 0 || py) && fs[f0][c - kRbPX] == F_SOLID,
                                  sk && (cj < ny - 1 || py) && fs[f0][c + kRbPX] == F_SOLID, xs[x0][c - kRbPX],
                                  xs[x0][c + kRbPX], diag, off);
                        axis_term(S.ax[2], ck, sk && (ck > 0 || pz) && fs[fm][c] == F_SOLID,
                                  sk && (ck < nz - 1 || pz) && fs[fp][c] == F_SOLID, xs[xm][c], xs[xp][c], diag, off);
                        v = diag != T(0) ? (b[q] + off) / diag : T(0);
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
            for (int e = tid; e < kRbTY * kRbTX; e += 256) {
                const int ti = e % kRbTX, tj = e / kRbTX;
                const int gi = i0 + ti, gj = j0 + tj;
                if (gi >= nx || gj >= ny) continue;
                const int h = (tj + 1) * kRbHX + ti + 1;
                T v = hs[h0][h];
                const long q = cidx(d, gi, gj, kf);
                if (((gi + gj + kf) & 1) == 1) {
                    const int c = (tj + 2) * kRbPX + ti + 2;
                    const uint8_t f = fs[f0][c];
                    if (q == S.pin || f == F_SOLID) {
                        v = b[q];
                    } else if (f == F_FLUID) {
                        T diag = S.c0, off = 0;
                        axis_term(S.ax[0], gi, sk && (gi > 0 || px) && fs[f0][c - 1] == F_SOLID,
                                  sk && (gi < nx - 1 || px) && fs[f0][c + 1] == F_SOLID, hs[h0][h - 1], hs[h0][h + 1],
                                  diag, off);
                        axis_term(S.ax[1], gj, sk && (gj > 0 || py) && fs[f0][c - kRbPX] == F_SOLID,
                                  sk && (gj < ny - 1 || py) && fs[f0][c + kRbPX] == F_SOLID, hs[h0][h - kRbHX],
                                  hs[h0][h + kRbHX], diag, off);
                        axis_term(S.ax[2], kf, sk && (kf > 0 || pz) && fs[fm][c] == F_SOLID,
                                  sk && (kf < nz - 1 || pz) && fs[fp][c] == F_SOLID, hs[hm][h], hs[hp][h], diag, off);
                        v = diag != T(0) ? (b[q] + off) / diag : T(0);
                    }
                }
                xb[q] = v;
            }
        }
        __syncthreads();  // the slots read above are refilled next iteration
    }
}

}  // namespace ibm
