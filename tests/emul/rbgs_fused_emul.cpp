// Host emulation of the fused red-black Gauss-Seidel kernel (csrc/rbgs_fused.h).
//
// The kernel source is compiled for the CPU with stand-ins for the HIP
// execution model: one std::thread per lane of a 256-lane workgroup,
// workgroups run one after another, __shared__ arrays become function-static
// (shared by the lanes of the running workgroup) and __syncthreads() is a
// std::barrier. Built with AddressSanitizer, every LDS and global access of
// the kernel is bounds-checked. The result of one fused sweep is compared
// bit for bit with a direct two-colour sweep (colour 0 over all cells, then
// colour 1), which is what the two one-colour launches compute.
//
// Usage: rbgs_fused_emul  (prints "ok <cases>" and exits 0 on success)
#include <barrier>
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <thread>
#include <vector>

#include "common.h"

#undef __shared__
#define __shared__ static
#ifndef __launch_bounds__
#define __launch_bounds__(n)
#endif

struct EmuIdx {
    int x = 0, y = 0, z = 0;
};
static thread_local EmuIdx emu_thread, emu_block;
static std::barrier<> *emu_bar = nullptr;
#define threadIdx emu_thread
#define blockIdx emu_block
#define __syncthreads() emu_bar->arrive_and_wait()

#include "rbgs_fused.h"

#undef threadIdx
#undef blockIdx

using namespace ibm;

template <typename T>
static void run_fused(const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b, const T *xa, T *xb,
                      int zero_in) {
    // same grid as launch_rbgs_fused (kchunk forced small to exercise chunk seams)
    const int kc = 3;
    const int gx = (d.nx + kRbTX - 1) / kRbTX, gy = (d.ny + kRbTY - 1) / kRbTY, gz = (d.nz + kc - 1) / kc;
    // one team of 256 lanes runs the workgroups in turn; the kernel ends each
    // plane step with a barrier, so no lane touches the LDS stand-ins of the
    // next workgroup before every lane has finished the previous one
    std::barrier<> bar(256);
    emu_bar = &bar;
    std::vector<std::thread> th;
    for (int t = 0; t < 256; t++)
        th.emplace_back([&, t]() {
            emu_thread = {t, 0, 0};
            for (int bz = 0; bz < gz; bz++)
                for (int by = 0; by < gy; by++)
                    for (int bx = 0; bx < gx; bx++) {
                        emu_block = {bx, by, bz};
                        k_rbgs_fused<T>(d, S, flag, b, xa, xb, zero_in, kc);
                    }
        });
    for (auto &x : th) x.join();
}

// direct sweep: colour 0 then colour 1, in place
template <typename T>
static void run_direct(const Dims &d, const SysCoef<T> &S, const uint8_t *flag, const T *b, T *x) {
    const int n[3] = {d.nx, d.ny, d.nz};
    const long st[3] = {1, (long)d.sy, d.sz};
    for (int color = 0; color < 2; color++)
        for (int k = 0; k < d.nz; k++)
            for (int j = 0; j < d.ny; j++)
                for (int i = 0; i < d.nx; i++) {
                    if (((i + j + k) & 1) != color) continue;
                    const long q = cidx(d, i, j, k);
                    if (q == S.pin || flag[q] == F_SOLID) { x[q] = b[q]; continue; }
                    if (flag[q] != F_FLUID) continue;
                    T diag = S.c0, off = 0;
                    const int id[3] = {i, j, k};
                    for (int a = 0; a < 3; a++) {
                        const AxisCoef<T> &A = S.ax[a];
                        const int ii = id[a];
                        const long qm = (ii == 0 && A.periodic) ? q + (long)(n[a] - 1) * st[a] : q - st[a];
                        const long qp = (ii == n[a] - 1 && A.periodic) ? q - (long)(n[a] - 1) * st[a] : q + st[a];
                        const bool sm = S.skip && (ii > 0 || A.periodic) && flag[qm] == F_SOLID;
                        const bool sp = S.skip && (ii < n[a] - 1 || A.periodic) && flag[qp] == F_SOLID;
                        axis_term(A, ii, sm, sp, x[qm], x[qp], diag, off);
                    }
                    x[q] = diag != T(0) ? (b[q] + off) / diag : T(0);
                }
}

template <typename T>
static bool one_case(std::mt19937_64 &rng, int nx, int ny, int nz, int per, int skip, int pin, int zero_in) {
    const Dims d = make_dims(nx, ny, nz);
    std::uniform_real_distribution<double> U(0.1, 1.0), V(-1.0, 1.0);
    std::vector<std::vector<T>> coef(12), packed(3);
    const int n[3] = {nx, ny, nz};
    SysCoef<T> S{};
    for (int a = 0; a < 3; a++) {
        for (int c = 0; c < 4; c++) {
            coef[4 * a + c].resize(n[a]);
            for (auto &v : coef[4 * a + c]) v = (T)U(rng);
        }
        packed[a].resize(4 * (size_t)n[a]);
        for (int i = 0; i < n[a]; i++)
            for (int c = 0; c < 4; c++) packed[a][4 * (size_t)i + c] = coef[4 * a + c][i];
        S.ax[a] = {coef[4 * a].data(), coef[4 * a + 1].data(), coef[4 * a + 2].data(), coef[4 * a + 3].data(), n[a],
                   (per >> a) & 1, packed[a].data()};
    }
    S.c0 = (T)0.25;
    S.skip = skip;
    S.is_p = 1;
    std::vector<uint8_t> flag(d.ncell, 0);
    std::vector<T> b(d.ncell, 0), x0(d.ncell, 0), xb(d.ncell, (T)-7), xd;
    std::uniform_int_distribution<int> F(0, 9);
    for (long q = 0; q < d.ncell; q++) {
        x0[q] = (T)V(rng);  // outer layers too: the kernel must copy, never update them
        b[q] = (T)V(rng);
        const int r = F(rng);
        flag[q] = r < 7 ? F_FLUID : r < 9 ? F_SOLID : F_GHOST;
    }
    S.pin = pin ? cidx(d, 0, 0, 0) : -1;
    if (zero_in)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) x0[cidx(d, i, j, k)] = 0;
    // the kernel with zero_in treats every value of xa as 0, outer layers
    // included; give the direct sweep the same input
    if (zero_in) std::fill(x0.begin(), x0.end(), (T)0);
    run_fused<T>(d, S, flag.data(), b.data(), x0.data(), xb.data(), zero_in);
    xd = x0;
    run_direct<T>(d, S, flag.data(), b.data(), xd.data());
    long bad = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const long q = cidx(d, i, j, k);
                if (std::memcmp(&xb[q], &xd[q], sizeof(T)) != 0) {
                    if (bad < 5)
                        fprintf(stderr, "mismatch (%d,%d,%d): fused %.17g direct %.17g\n", i, j, k, (double)xb[q],
                                (double)xd[q]);
                    bad++;
                }
            }
    if (bad)
        fprintf(stderr, "case %dx%dx%d per=%d skip=%d pin=%d zero=%d sizeof(T)=%zu: %ld mismatches\n", nx, ny, nz,
                per, skip, pin, zero_in, sizeof(T), bad);
    return bad == 0;
}

int main() {
    std::mt19937_64 rng(20201);
    struct C {
        int nx, ny, nz, per;
    } cs[] = {{70, 13, 11, 0}, {5, 9, 6, 0}, {130, 17, 7, 0}, {16, 12, 8, 5}, {6, 10, 4, 7}, {66, 8, 9, 2},
              {2, 2, 2, 0}, {64, 8, 3, 0}, {20, 4, 10, 4}};
    int ncase = 0;
    bool ok = true;
    for (const C &c : cs)
        for (int skip = 0; skip < 2; skip++)
            for (int zero = 0; zero < 2; zero++) {
                const int pin = (c.nx + skip + zero) & 1;
                ok &= one_case<double>(rng, c.nx, c.ny, c.nz, c.per, skip, pin, zero);
                ok &= one_case<float>(rng, c.nx, c.ny, c.nz, c.per, skip, pin, zero);
                ncase += 2;
            }
    if (!ok) return 1;
    printf("ok %d\n", ncase);
    return 0;
}
