// Host emulation of the four-cells-per-lane FP32 red-black kernel
// (csrc/rbgs_quad.h).
//
// The kernel source is compiled for the CPU with stand-ins for the thread and
// block indices and run lane by lane over the launch grid of
// launch_rbgs_color<float>. Within one colour launch a lane reads only cells
// of the other colour and writes only cells of its own quad, so running the
// lanes one after another gives what the device computes. Built with
// AddressSanitizer, every load of the kernel (including the 16-byte loads
// that run past the last cell of a row) is checked against the allocation.
// Two launches (colour 0, then colour 1) are compared bit for bit with a
// direct sweep that applies the stencil cell by cell, as k_rbgs does
// (Gauss-Seidel, and over-relaxed), and
// the four-cells-per-lane residual with the one-cell residual at every fluid
// cell (and with 0 at the others).
//
// Usage: rbgs_quad_emul  (prints "ok <cases>" and exits 0 on success)
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "common.h"

#ifndef __launch_bounds__
#define __launch_bounds__(n)
#endif

struct EmuIdx {
    int x = 0, y = 0, z = 0;
};
static EmuIdx emu_thread;
#define threadIdx emu_thread

#include "rbgs_quad.h"

#undef threadIdx

using namespace ibm;

static void run_quad(const Dims &d, const SysCoef<float> &S, const uint8_t *flag, const float *b, float *x, int color,
                     int zero_in) {
    constexpr int BY = 8;  // as launch_rbgs_color<float>
    const int gx = ((d.nx + 3) / 4 + 63) / 64, gy = (d.ny + BY - 1) / BY, gz = d.nz;
    for (int bz = 0; bz < gz; bz++)
        for (int by = 0; by < gy; by++)
            for (int bx = 0; bx < gx; bx++)
                for (int ty = 0; ty < BY; ty++)
                    for (int tx = 0; tx < 64; tx++) {
                        emu_thread = {tx, ty, 0};
                        rbgs_quad<BY>(d, S, flag, b, x, color, zero_in, make_uint3(bx, by, bz));
                    }
}

// direct sweep of one colour, in place (the k_rbgs arithmetic)
static void run_direct(const Dims &d, const SysCoef<float> &S, const uint8_t *flag, const float *b, float *x,
                       int color) {
    const int n[3] = {d.nx, d.ny, d.nz};
    const long st[3] = {1, (long)d.sy, d.sz};
    for (int k = 0; k < d.nz; k++)
        for (int j = 0; j < d.ny; j++)
            for (int i = 0; i < d.nx; i++) {
                if (((i + j + k) & 1) != color) continue;
                const long q = cidx(d, i, j, k);
                if (q == S.pin || flag[q] == F_SOLID) { x[q] = b[q]; continue; }
                if (flag[q] != F_FLUID) continue;
                float diag = S.c0, off = 0;
                const int id[3] = {i, j, k};
                for (int a = 0; a < 3; a++) {
                    const AxisCoef<float> &A = S.ax[a];
                    const int ii = id[a];
                    const long qm = (ii == 0 && A.periodic) ? q + (long)(n[a] - 1) * st[a] : q - st[a];
                    const long qp = (ii == n[a] - 1 && A.periodic) ? q - (long)(n[a] - 1) * st[a] : q + st[a];
                    const bool sm = S.skip && (ii > 0 || A.periodic) && flag[qm] == F_SOLID;
                    const bool sp = S.skip && (ii < n[a] - 1 || A.periodic) && flag[qp] == F_SOLID;
                    axis_term(A, ii, sm, sp, x[qm], x[qp], diag, off);
                }
                const float gs = diag != 0.f ? (b[q] + off) / diag : 0.f;
                // over-relaxation: xo + om (gs - xo), gs itself for om = 1
                const float om = S.omega;
                x[q] = om == 1.f ? gs : x[q] + om * (gs - x[q]);
            }
}

static void run_resid_quad(const Dims &d, const SysCoef<float> &S, const uint8_t *flag, const float *b, const float *x,
                           float *r) {
    constexpr int BY = 8;  // as launch_mg_residual<float>
    const int gx = ((d.nx + 3) / 4 + 63) / 64, gy = (d.ny + BY - 1) / BY, gz = d.nz;
    for (int bz = 0; bz < gz; bz++)
        for (int by = 0; by < gy; by++)
            for (int bx = 0; bx < gx; bx++)
                for (int ty = 0; ty < BY; ty++)
                    for (int tx = 0; tx < 64; tx++) {
                        emu_thread = {tx, ty, 0};
                        resid_quad<BY>(d, S, flag, b, x, r, make_uint3(bx, by, bz));
                    }
}

// residual of one fluid cell as the one-cell kernel computes it
static float direct_resid(const Dims &d, const SysCoef<float> &S, const uint8_t *flag, const float *b, const float *x,
                          int i, int j, int k) {
    const int n[3] = {d.nx, d.ny, d.nz};
    const long st[3] = {1, (long)d.sy, d.sz};
    const long q = cidx(d, i, j, k);
    if (q == S.pin) return b[q] - x[q];
    float diag = S.c0, off = 0;
    const int id[3] = {i, j, k};
    for (int a = 0; a < 3; a++) {
        const AxisCoef<float> &A = S.ax[a];
        const int ii = id[a];
        const long qm = (ii == 0 && A.periodic) ? q + (long)(n[a] - 1) * st[a] : q - st[a];
        const long qp = (ii == n[a] - 1 && A.periodic) ? q - (long)(n[a] - 1) * st[a] : q + st[a];
        const bool sm = S.skip && (ii > 0 || A.periodic) && flag[qm] == F_SOLID;
        const bool sp = S.skip && (ii < n[a] - 1 || A.periodic) && flag[qp] == F_SOLID;
        axis_term(A, ii, sm, sp, x[qm], x[qp], diag, off);
    }
    return b[q] - (diag * x[q] - off);
}

static bool one_case(std::mt19937_64 &rng, int nx, int ny, int nz, int per, int skip, int pin, int zero_in,
                     float omega) {
    const Dims d = make_dims(nx, ny, nz);
    std::uniform_real_distribution<double> U(0.1, 1.0), V(-1.0, 1.0);
    std::vector<std::vector<float>> coef(12), packed(3);
    const int n[3] = {nx, ny, nz};
    SysCoef<float> S{};
    for (int a = 0; a < 3; a++) {
        for (int c = 0; c < 4; c++) {
            coef[4 * a + c].resize(n[a]);
            for (auto &v : coef[4 * a + c]) v = (float)U(rng);
        }
        packed[a].resize(4 * (size_t)n[a]);
        for (int i = 0; i < n[a]; i++)
            for (int c = 0; c < 4; c++) packed[a][4 * (size_t)i + c] = coef[4 * a + c][i];
        S.ax[a] = {coef[4 * a].data(), coef[4 * a + 1].data(), coef[4 * a + 2].data(), coef[4 * a + 3].data(), n[a],
                   (per >> a) & 1, packed[a].data()};
    }
    S.c0 = 0.25f;
    S.is_p = 1;
    std::vector<uint8_t> flag(d.ncell, 0), fmask(d.ncell, 0);
    std::vector<float> b(d.ncell, 0), x0(d.ncell, 0);
    std::uniform_int_distribution<int> F(0, 9);
    for (long q = 0; q < d.ncell; q++) {
        x0[q] = (float)V(rng);
        b[q] = (float)V(rng);
        const int r = F(rng);
        flag[q] = r < 7 ? F_FLUID : r < 9 ? F_SOLID : F_GHOST;
    }
    // face masks as the coarse levels store them: bit 2a / 2a+1 when the
    // minus / plus interior neighbour along axis a is solid
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const int id[3] = {i, j, k};
                unsigned bits = 0;
                for (int a = 0; a < 3; a++) {
                    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
                    const bool p = (per >> a) & 1;
                    lo[a] = id[a] == 0 ? (p ? n[a] - 1 : -1) : id[a] - 1;
                    hi[a] = id[a] == n[a] - 1 ? (p ? 0 : -1) : id[a] + 1;
                    if (lo[a] >= 0 && flag[cidx(d, lo[0], lo[1], lo[2])] == F_SOLID) bits |= 1u << (2 * a);
                    if (hi[a] >= 0 && flag[cidx(d, hi[0], hi[1], hi[2])] == F_SOLID) bits |= 1u << (2 * a + 1);
                }
                fmask[cidx(d, i, j, k)] = (uint8_t)bits;
            }
    S.omega = omega;
    S.skip = skip;
    S.fmask = skip ? fmask.data() : nullptr;
    S.pin = pin ? cidx(d, 0, 0, 0) : -1;
    if (zero_in) std::fill(x0.begin(), x0.end(), 0.f);
    std::vector<float> xq = x0, xd = x0;
    run_quad(d, S, flag.data(), b.data(), xq.data(), 0, zero_in);
    run_quad(d, S, flag.data(), b.data(), xq.data(), 1, 0);
    run_direct(d, S, flag.data(), b.data(), xd.data(), 0);
    run_direct(d, S, flag.data(), b.data(), xd.data(), 1);
    long bad = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const long q = cidx(d, i, j, k);
                // zero start: the kernel also zeroes ghost cells the direct
                // sweep leaves at their (zero) input
                if (std::memcmp(&xq[q], &xd[q], sizeof(float)) != 0) {
                    if (bad < 5)
                        fprintf(stderr, "mismatch (%d,%d,%d): quad %.9g direct %.9g\n", i, j, k, (double)xq[q],
                                (double)xd[q]);
                    bad++;
                }
            }
    // residual of the swept field at the fluid cells
    std::vector<float> r(d.ncell, 0.f);
    std::fill(r.begin(), r.end(), -3.f);
    run_resid_quad(d, S, flag.data(), b.data(), xq.data(), r.data());
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const long q = cidx(d, i, j, k);
                const float want = flag[q] != F_FLUID ? 0.f : direct_resid(d, S, flag.data(), b.data(), xq.data(), i, j, k);
                if (std::memcmp(&r[q], &want, sizeof(float)) != 0) {
                    if (bad < 5) fprintf(stderr, "residual mismatch (%d,%d,%d): %.9g vs %.9g\n", i, j, k, r[q], want);
                    bad++;
                }
            }
    if (bad)
        fprintf(stderr, "case %dx%dx%d per=%d skip=%d pin=%d zero=%d omega=%g: %ld mismatches\n", nx, ny, nz, per,
                skip, pin, zero_in, (double)omega, bad);
    return bad == 0;
}

int main() {
    std::mt19937_64 rng(20202);
    struct C {
        int nx, ny, nz, per;
    } cs[] = {{70, 13, 11, 0}, {5, 9, 6, 0}, {130, 17, 7, 0}, {16, 12, 8, 5}, {6, 10, 4, 7}, {66, 8, 9, 2},
              {2, 2, 2, 0}, {64, 8, 3, 0}, {20, 4, 10, 4}, {257, 9, 5, 0}, {258, 9, 6, 1}, {259, 3, 4, 0},
              {1, 3, 2, 0}, {3, 1, 1, 0}, {28, 6, 6, 7}};
    // periodic axes have even cell counts: with an odd count the two cells
    // beside the wrap share a colour and a one-colour sweep has no fixed order
    int ncase = 0;
    bool ok = true;
    for (const C &c : cs)
        for (int skip = 0; skip < 2; skip++)
            for (int zero = 0; zero < 2; zero++)
                for (float omega : {1.0f, 1.3f}) {
                    const int pin = (c.nx + skip + zero) & 1;
                    ok &= one_case(rng, c.nx, c.ny, c.nz, c.per, skip, pin, zero, omega);
                    ncase++;
                }
    if (!ok) return 1;
    printf("ok %d\n", ncase);
    return 0;
}
