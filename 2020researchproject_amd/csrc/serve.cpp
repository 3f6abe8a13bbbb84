// serve.cpp - request loop over independent cases on one GPU (a stream of
// mixed-size cavity and sphere cases arriving while others run).
//
// A pool of worker threads, one per slot. A worker takes the first queued job
// (in arrival order) whose footprint fits the remaining device-memory budget,
// assembles the job's solver (configured, unassembled, by the submitter, so
// the worker runs native code only), sets the initial flow, runs the time
// steps on the solver's own non-blocking stream, records the results and
// destroys the solver. Concurrent cases therefore overlap on the device;
// admission never lets the admitted footprints exceed the budget. An error of
// one case (rejected input, a non-finite residual, a failed allocation)
// fails that job: the worker catches it, releases the job's budget and
// solver, and takes the next job.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "solver.h"
#include "../../include/ibm3d_ext.h"
#include "../../include/ibm3d_serve.h"

namespace {

// splitmix64 finaliser
uint64_t sm64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t bits(double d) {
    uint64_t b;
    memcpy(&b, &d, 8);
    return b;
}

// Uniform draw in [0, 1) keyed by (seed, component, cell centre): the hash
// chain z, y, x, component, seed of splitmix64 finalisers, top 53 bits.
double uniform01(uint64_t seed, uint64_t comp, double x, double y, double z) {
    uint64_t h = sm64(bits(z));
    h = sm64(bits(y) ^ h);
    h = sm64(bits(x) ^ h);
    h = sm64(comp ^ h);
    h = sm64(seed ^ h);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// Initial flow of a job: constants, plus the seeded perturbation on the
// interior cells of each velocity component.
void apply_init(IBMSolver *s, const IBMServeInit &in) {
    IBMSolver_init_flow_const(s, in.u1, in.u2, in.u3, in.p);
    if (!(in.amp > 0)) return;
    const int X = s->Nx + 4, Y = s->Ny + 4, Z = s->Nz + 4;
    std::vector<double> a((size_t)X * Y * Z);
    const double base[3] = {in.u1, in.u2, in.u3};
    const int field[3] = {IBM3D_FIELD_U1, IBM3D_FIELD_U2, IBM3D_FIELD_U3};
    for (int c = 0; c < 3; c++) {
        std::fill(a.begin(), a.end(), base[c]);
        for (int i = 2; i < X - 2; i++)
            for (int j = 2; j < Y - 2; j++)
                for (int k = 2; k < Z - 2; k++)
                    a[((size_t)i * Y + j) * Z + k] =
                        base[c] + in.amp * (2.0 * uniform01(in.seed, c + 1, s->xc[i], s->yc[j], s->zc[k]) - 1.0);
        IBMSolver_set_field(s, field[c], a.data());
    }
}

}  // namespace

struct _ibm_server {
    struct Job {
        double bytes = 0;
        int steps = 0;
        // callback jobs
        IBMServeSetupFunc setup = nullptr;
        IBMServeDoneFunc done = nullptr;
        void *user = nullptr;
        // native jobs: the configured solver (owned from submission), the
        // initial flow and the caller's field buffers (NULL: not kept)
        IBMSolver *solver = nullptr;
        IBMServeInit init{};
        double *fields[4] = {nullptr, nullptr, nullptr, nullptr};
        double t[3] = {-1, -1, -1};  // submitted, admitted, finished
        int state = 0;               // 0 queued, 1 running, 2 done, 3 failed
        int iters[4] = {0, 0, 0, 0};
        double max_div = -1, dev_bytes = 0;
        std::string error;
    };
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job> jobs;          // deque: references stay valid as jobs are appended
    std::deque<int> queue;         // ids of queued jobs, arrival order
    double budget = 0, used = 0, peak = 0;
    int running = 0, finished = 0, failed = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // Runs one job; returns "" on success or the error text. Case errors
    // throw (JobErrorScope) only inside native code: a setup or done
    // callback runs with them aborting, as outside the server, since an
    // exception cannot cross the caller's frames.
    static std::string run(Job &J) {
        IBMSolver *s = J.solver;
        J.solver = nullptr;
        std::string err;
        try {
            if (J.setup) s = J.setup(J.user);
            if (s) {
                {
                    ibm::JobErrorScope scope;
                    if (!J.setup) {
                        IBMSolver_assemble(s);
                        apply_init(s, J.init);
                    }
                    IBMSolver_iterate(s, J.steps, false);
                    IBMSolver_get_last_iters(s, J.iters);
                    J.max_div = IBMSolver_max_div(s);
                    for (int f = 0; f < 4; f++)
                        if (J.fields[f]) IBMSolver_get_field(s, f, J.fields[f]);
                }
                J.dev_bytes = IBMSolver_get_device_bytes(s);
                if (J.done) J.done(s, J.user);
            } else if (J.done) {
                J.done(nullptr, J.user);
            }
        } catch (const std::exception &e) {
            // the message without the line breaks of its printed form
            err = e.what();
            const size_t a = err.find_first_not_of("\n"), b = err.find_last_not_of("\n");
            err = a == std::string::npos ? std::string("error") : err.substr(a, b - a + 1);
        }
        IBMSolver_destroy(s);
        return err;
    }

    void work() {
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            int id = -1;
            cv.wait(lk, [&] {
                if (stop && queue.empty()) return true;
                for (size_t q = 0; q < queue.size(); q++)
                    if (used + jobs[queue[q]].bytes <= budget) { id = (int)q; return true; }
                return false;
            });
            if (id < 0) return;  // stopping and nothing queued
            const int job = queue[id];
            queue.erase(queue.begin() + id);
            Job &J = jobs[job];
            J.state = 1;
            J.t[1] = now();
            used += J.bytes;
            peak = std::max(peak, used);
            running++;
            lk.unlock();
            std::string err = run(J);
            lk.lock();
            J.error = err;
            J.state = err.empty() ? 2 : 3;
            if (!err.empty()) failed++;
            J.t[2] = now();
            used -= J.bytes;
            running--;
            finished++;
            cv.notify_all();
        }
    }

    int push(Job &&J) {
        std::lock_guard<std::mutex> lk(mu);
        if (stop) return -1;
        J.t[0] = now();
        jobs.push_back(std::move(J));
        const int id = (int)jobs.size() - 1;
        queue.push_back(id);
        cv.notify_all();
        return id;
    }
};

extern "C" {

IBMServer *IBMServer_new(double budget_bytes, int max_active) {
    if (max_active < 1 || max_active > 64) ibm::fatal("IBMServer_new: max_active must be 1..64");
    if (!(budget_bytes > 0)) ibm::fatal("IBMServer_new: budget must be positive");
    IBMServer *sv = new IBMServer();
    sv->budget = budget_bytes;
    for (int w = 0; w < max_active; w++) sv->workers.emplace_back([sv] { sv->work(); });
    return sv;
}

int IBMServer_submit(IBMServer *sv, double bytes, int steps, IBMServeSetupFunc setup, IBMServeDoneFunc done,
                     void *user) {
    if (bytes > sv->budget || bytes < 0 || steps < 0) return -1;
    _ibm_server::Job J;
    J.bytes = bytes;
    J.steps = steps;
    J.setup = setup;
    J.done = done;
    J.user = user;
    return sv->push(std::move(J));
}

int IBMServer_submit_solver(IBMServer *sv, IBMSolver *solver, double bytes, int steps, const IBMServeInit *init,
                            double *fields[4]) {
    if (!solver || bytes > sv->budget || bytes < 0 || steps < 0) return -1;
    _ibm_server::Job J;
    J.bytes = bytes;
    J.steps = steps;
    J.solver = solver;
    if (init) J.init = *init;
    if (fields)
        for (int f = 0; f < 4; f++) J.fields[f] = fields[f];
    return sv->push(std::move(J));
}

int IBMServer_job_result(IBMServer *sv, int job, IBMServeResult *out) {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (job < 0 || job >= (int)sv->jobs.size()) return -1;
    const _ibm_server::Job &J = sv->jobs[job];
    memset(out, 0, sizeof *out);
    out->state = J.state;
    for (int c = 0; c < 3; c++) out->t[c] = J.t[c];
    for (int c = 0; c < 4; c++) out->iters[c] = J.iters[c];
    out->max_div = J.max_div;
    out->device_bytes = J.dev_bytes;
    snprintf(out->error, sizeof out->error, "%s", J.error.c_str());
    return J.state;
}

int IBMServer_failed_jobs(IBMServer *sv) {
    std::lock_guard<std::mutex> lk(sv->mu);
    return sv->failed;
}

void IBMServer_wait(IBMServer *sv) {
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv.wait(lk, [&] { return sv->finished == (int)sv->jobs.size(); });
}

int IBMServer_job_times(IBMServer *sv, int job, double out[3]) {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (job < 0 || job >= (int)sv->jobs.size()) return -1;
    for (int c = 0; c < 3; c++) out[c] = sv->jobs[job].t[c];
    return sv->jobs[job].state;
}

double IBMServer_peak_bytes(IBMServer *sv) {
    std::lock_guard<std::mutex> lk(sv->mu);
    return sv->peak;
}

void IBMServer_destroy(IBMServer *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        sv->cv.notify_all();
    }
    for (auto &t : sv->workers) t.join();
    delete sv;
}

double IBMSolver_estimate_device_bytes(int Nx, int Ny, int Nz, int precision, int les) {
    const ibm::Dims d = ibm::make_dims(Nx, Ny, Nz);
    const double cells = (double)d.ncell;
    const bool mixed = precision != IBM3D_PREC_FP64;
    // per padded cell, as allocated by alloc_fields: step fields (u, u_next,
    // u*, U, U_next, U*, N, N_prev, b per component; p, p_next, p', b_p),
    // level set, 8 Krylov vectors, flags
    double per = 8.0 * (27 + 4 + 1 + 8) + 1;
    if (les) per += 8.0 * 6;  // tau
    // the two extra lanes of the concurrent momentum solves: 8 FP64 vectors
    // each, plus 2 FP32 preconditioned directions each in mode 2
    if (precision != IBM3D_PREC_MIXED) per += 2 * 8.0 * 8;
    if (precision == IBM3D_PREC_MIXED_KRYLOV) per += 2 * 2 * 4.0;
    // refinement vectors: FP64 residual, FP32 right-hand side and correction
    // (+ a second FP32 correction in mode 2); mode 1: 8 FP32 Krylov vectors
    if (mixed) per += 8 + 4 + 4 + (precision == IBM3D_PREC_MIXED_KRYLOV ? 4 : 0);
    if (precision == IBM3D_PREC_MIXED) per += 4 * 8;
    // multigrid (build_mg): fine-level residual and smoother scratch (FP64,
    // + FP32); coarse levels x, b, r, second smoother buffer (FP64, + FP32),
    // flag and face-mask bytes, and one interpolation-mask byte per cell of
    // the next finer level. The coarse levels hold at most 1.5 times the fine
    // cells at the default coarsening threshold.
    const double coarse = 1.5 * cells;
    const double mg = cells * (8.0 * 2 + (mixed ? 4.0 * 2 : 0))
        + coarse * (8.0 * 4 + 2 + (mixed ? 4.0 * 4 : 0)) + (cells + coarse) * 1;
    // dense inverse of the coarsest level (at most 3000 rows; FP64, + FP32)
    // and the pivot scratch of its inversion
    const double n = std::min(3000.0, (double)Nx * Ny * Nz);
    const double dense = n * n * (8 + (mixed ? 4 : 0)) + 2 * 8.0 * n;
    // reduction partials, ghost lists, 1-D metrics and folded coefficients
    return cells * per + mg + dense + 16.0 * (1 << 20);
}

double IBMSolver_get_device_bytes(IBMSolver *s) { return (double)s->dev_bytes; }

}  // extern "C"
