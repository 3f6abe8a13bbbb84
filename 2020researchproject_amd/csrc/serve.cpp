// serve.cpp - request loop over independent cases on one GPU (C5 serving).
//
// A pool of worker threads, one per slot. A worker takes the first queued job
// (in arrival order) whose footprint fits the remaining device-memory budget,
// builds its solver through the job's setup callback, runs the time steps on
// the solver's own non-blocking stream and hands the solver to the done
// callback. Concurrent cases therefore overlap on the device; admission never
// lets the admitted footprints exceed the budget.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "solver.h"
#include "../../include/ibm3d_serve.h"

struct _ibm_server {
    struct Job {
        double bytes = 0;
        int steps = 0;
        IBMServeSetupFunc setup = nullptr;
        IBMServeDoneFunc done = nullptr;
        void *user = nullptr;
        double t[3] = {-1, -1, -1};  // submitted, admitted, finished
        int state = 0;               // 0 queued, 1 running, 2 done
    };
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job> jobs;          // deque: references stay valid as jobs are appended
    std::deque<int> queue;         // ids of queued jobs, arrival order
    double budget = 0, used = 0, peak = 0;
    int running = 0, finished = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
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
            IBMSolver *s = J.setup ? J.setup(J.user) : nullptr;
            if (s) {
                IBMSolver_iterate(s, J.steps, false);
                if (J.done) J.done(s, J.user);
                IBMSolver_destroy(s);
            } else if (J.done) {
                J.done(nullptr, J.user);
            }
            lk.lock();
            J.state = 2;
            J.t[2] = now();
            used -= J.bytes;
            running--;
            finished++;
            cv.notify_all();
        }
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
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return -1;
    _ibm_server::Job J;
    J.bytes = bytes;
    J.steps = steps;
    J.setup = setup;
    J.done = done;
    J.user = user;
    J.t[0] = sv->now();
    sv->jobs.push_back(J);
    const int id = (int)sv->jobs.size() - 1;
    sv->queue.push_back(id);
    sv->cv.notify_all();
    return id;
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
