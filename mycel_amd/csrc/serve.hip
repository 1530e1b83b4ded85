// Job server: continuous batching of FEA solves and SIMP runs on one GPU.
//
// The reference runs one optimisation at a time behind a RAM guard (Diagnostics/MemoryGuard.cs:434-495,
// Optimization/OptimizationCoordinator.cs:214-353). Here jobs are submitted at any time; a job is
// admitted as soon as its HBM footprint fits the budget and a worker is free (FIFO, no head-of-line
// bypass), so work that arrives while others run joins them on the device. Each worker owns one HIP
// stream; a job's data (densities, warm-start displacements, PCG state) stays resident on its worker
// for the job's whole life. Kernels of different jobs overlap on the device through the workers'
// streams. A job's arithmetic does not depend on what runs beside it: every reduction is a fixed
// per-job tree, so a job's result is bit-identical alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mycel_internal.hpp"

namespace mycel {
namespace {

using Clock = std::chrono::steady_clock;

// What a run produces; a worker fills its own copy and publishes it into the job under the server lock.
struct JobResult {
    std::vector<double> u, rho_bar;
    double compliance = 0, max_disp = 0;
    int32_t iterations = 0, simp_iterations = 0, converged = 0;
    int32_t status = MYCEL_OK;
    std::string error;
};

struct Job {
    int64_t id = 0;
    mycel_job_desc desc{};
    std::vector<uint8_t> domain;
    std::vector<int32_t> fixed, patch;
    std::vector<double> forces, density;
    double footprint = 0;
    JobResult res;  // written only under the server lock, when the job finishes
    int32_t state = MYCEL_JOB_QUEUED;
    Clock::time_point t_submit, t_start, t_end;
};

double ms(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

// HBM footprint of one job: operator arrays, PCG workspace (FP64 + FP32), SIMP vectors, filter maps.
double job_footprint(const mycel_job_desc &d) {
    const double ne = static_cast<double>(d.grid.nx) * d.grid.ny * d.grid.nz;
    const double nn = static_cast<double>(d.grid.nx + 1) * (d.grid.ny + 1) * (d.grid.nz + 1);
    const double npad = static_cast<double>(d.grid.nx + 2) * (d.grid.ny + 2) * (d.grid.nz + 2);
    const double ndof = 3 * nn;
    const int prec = d.kind == MYCEL_JOB_SIMP ? d.simp.precision : d.precision;
    // Multigrid hierarchy: six FP32 vectors per level (the levels below the finest add < 1/7 of it), the
    // coarse moduli, the Galerkin blocks and the coarsest inverse.
    const double gmg = prec == MYCEL_PRECISION_MIXED_GMG ? ndof * 6 * 4 * (8.0 / 7.0) + npad * 4 / 7.0 + (1 << 20) : 0.0;
    return 12 * npad + ndof * (1 + 8 + 8 + 4 + 16 + 32 + 20 + 24) + ne * (1 + 8 + 4 + 4 + 8 * 8) + gmg + (8 << 20);
}

}  // namespace mycel

struct mycel_server {
    double budget = 0, in_use = 0;
    int device = 0;
    int n_workers = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<mycel::Job *> queue;
    std::map<int64_t, std::unique_ptr<mycel::Job>> jobs;
    std::vector<std::thread> workers;
    int64_t next_id = 1;
    int active = 0;
    bool stopping = false;
    bool packing = true;                   // the jobs' multigrid solves go to the broker
    mycel::SolveBroker *broker = nullptr;  // created with the server; destroyed after the workers
};

namespace mycel {
namespace {

// Runs a job from its own input copies (read-only while it runs) into r; broker: where its multigrid solves go
// (null: on the worker's own stream).
void run_job(const Job &j, int device, SolveBroker *broker, JobResult &r) {
    MYCEL_HIP(hipSetDevice(device));
    const mycel_job_desc &d = j.desc;
    std::unique_ptr<mycel_op, void (*)(mycel_op *)> op(
        op_create(d.grid, j.domain.data(), j.fixed.data(), static_cast<int>(j.fixed.size()), d.e0, d.nu,
                  d.emin_ratio, d.penal, j.density.empty() ? nullptr : j.density.data()),
        op_destroy);
    op->broker = broker;
    const size_t ndof = static_cast<size_t>(op->d.n_dofs);
    r.u.assign(ndof, 0.0);
    if (d.kind == MYCEL_JOB_FEA) {
        mycel_pcg_result pr{};
        static_analysis_run(op.get(), j.forces.data(), r.u.data(), d.tol, d.max_iter, true, d.precision,
                            &r.compliance, &r.max_disp, nullptr, nullptr, &pr);
        r.iterations = pr.iterations;
        r.converged = pr.converged;
    } else {
        mycel_simp_result res{};
        const int na = static_cast<int>(op->active.size());
        r.rho_bar.assign(std::max(1, na), 0.0);
        std::vector<mycel_snapshot> snaps(std::max(1, d.simp.max_iterations - d.simp.start_iteration + 1));
        simp_run(op.get(), j.forces.data(), 1, d.simp, nullptr, j.patch.data(), static_cast<int>(j.patch.size()),
                 nullptr, 0, nullptr, nullptr, nullptr, &res, snaps.data(), static_cast<int>(snaps.size()), nullptr,
                 r.rho_bar.data(), r.u.data());
        r.rho_bar.resize(na);
        r.compliance = res.final_compliance;
        r.simp_iterations = res.total_iterations;
        r.converged = res.converged;
        for (int i = 0; i < std::min<int>(res.total_iterations, static_cast<int>(snaps.size())); i++)
            r.iterations += snaps[i].solver_iterations;
    }
}

void worker_loop(mycel_server *sv) {
    for (;;) {
        Job *job = nullptr;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_work.wait(lk, [&] {
                if (sv->stopping) return true;
                if (sv->queue.empty()) return false;
                Job *head = sv->queue.front();
                // Admission by HBM footprint; a job larger than the whole budget runs alone.
                return sv->in_use + head->footprint <= sv->budget || sv->active == 0;
            });
            if (sv->stopping) return;
            job = sv->queue.front();
            sv->queue.pop_front();
            sv->in_use += job->footprint;
            sv->active++;
            job->state = MYCEL_JOB_RUNNING;
            job->t_start = Clock::now();
        }
        JobResult r;
        try {
            run_job(*job, sv->device, sv->packing ? sv->broker : nullptr, r);
            r.status = MYCEL_OK;
        } catch (const Error &e) {
            r.status = e.status;
            r.error = e.what();
        } catch (const std::bad_alloc &) {
            r.status = MYCEL_E_OUT_OF_MEMORY;
            r.error = "host allocation failed";
        } catch (const std::exception &e) {
            r.status = MYCEL_E_DEVICE;
            r.error = e.what();
        }
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            job->res = std::move(r);
            job->t_end = Clock::now();
            job->state = job->res.status == MYCEL_OK ? MYCEL_JOB_DONE : MYCEL_JOB_FAILED;
            sv->in_use -= job->footprint;
            sv->active--;
        }
        sv->cv_work.notify_all();
        sv->cv_done.notify_all();
    }
}

void fill_info(const Job &j, mycel_job_info *info) {
    *info = mycel_job_info{};
    info->state = j.state;
    info->status = j.res.status;
    info->compliance = j.res.compliance;
    info->max_displacement = j.res.max_disp;
    info->pcg_iterations = j.res.iterations;
    info->simp_iterations = j.res.simp_iterations;
    info->converged = j.res.converged;
    info->footprint_bytes = j.footprint;
    const auto now = Clock::now();
    const auto start = j.state == MYCEL_JOB_QUEUED ? now : j.t_start;
    const auto end = (j.state == MYCEL_JOB_DONE || j.state == MYCEL_JOB_FAILED) ? j.t_end : now;
    info->queue_ms = ms(j.t_submit, start);
    info->run_ms = j.state == MYCEL_JOB_QUEUED ? 0.0 : ms(j.t_start, end);
    info->latency_ms = ms(j.t_submit, end);
}

}  // namespace
}  // namespace mycel

using namespace mycel;

extern "C" {

int32_t mycel_server_create(double hbm_budget_bytes, int32_t workers, mycel_server **out) {
    try {
        if (!out) throw Error(MYCEL_E_ARGUMENT_NULL, "out");
        if (workers <= 0 || workers > 16) throw Error(MYCEL_E_OUT_OF_RANGE, "workers must be in [1, 16]");
        auto sv = std::make_unique<mycel_server>();
        MYCEL_HIP(hipGetDevice(&sv->device));
        size_t free_b = 0, total_b = 0;
        MYCEL_HIP(hipMemGetInfo(&free_b, &total_b));
        // Budget: the caller's figure, capped by what the device has free (leave 5% headroom).
        const double cap = 0.95 * static_cast<double>(free_b);
        sv->budget = hbm_budget_bytes > 0 ? std::min(hbm_budget_bytes, cap) : cap;
        sv->n_workers = workers;
        sv->broker = broker_create(sv->device, workers);  // at most one solve per resident job
        for (int i = 0; i < workers; i++) sv->workers.emplace_back(worker_loop, sv.get());
        *out = sv.release();
        return MYCEL_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.status;
    }
}

int32_t mycel_server_destroy(mycel_server *sv) {
    if (!sv) return MYCEL_OK;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        sv->cv_done.wait(lk, [&] { return sv->queue.empty() && sv->active == 0; });
        sv->stopping = true;
    }
    sv->cv_work.notify_all();
    for (auto &t : sv->workers) t.join();
    broker_destroy(sv->broker);
    delete sv;
    return MYCEL_OK;
}

int32_t mycel_server_set_packing(mycel_server *sv, int32_t enable) {
    if (!sv) return MYCEL_E_ARGUMENT_NULL;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (!sv->jobs.empty()) {
        set_last_error("packing is set before the first submit");
        return MYCEL_E_INVALID_OPERATION;
    }
    sv->packing = enable != 0;
    return MYCEL_OK;
}

int32_t mycel_server_pack_stats(mycel_server *sv, mycel_pack_stats *out) {
    if (!sv || !out) return MYCEL_E_ARGUMENT_NULL;
    *out = broker_stats(sv->broker);
    return MYCEL_OK;
}

int32_t mycel_server_submit(mycel_server *sv, const mycel_job_desc *d, int64_t *id) {
    try {
        if (!sv || !d || !id || !d->domain || !d->forces) throw Error(MYCEL_E_ARGUMENT_NULL, "null argument");
        if (d->kind != MYCEL_JOB_FEA && d->kind != MYCEL_JOB_SIMP) throw Error(MYCEL_E_ARGUMENT, "unknown job kind");
        if (d->grid.nx <= 0 || d->grid.ny <= 0 || d->grid.nz <= 0 || !(d->grid.h > 0))
            throw Error(MYCEL_E_OUT_OF_RANGE, "grid dimensions and h must be positive");
        if (d->n_fixed < 0 || d->n_patch_nodes < 0) throw Error(MYCEL_E_ARGUMENT, "negative array length");
        if ((d->n_fixed > 0 && !d->fixed_dofs) || (d->n_patch_nodes > 0 && !d->patch_nodes))
            throw Error(MYCEL_E_ARGUMENT_NULL, "fixed_dofs / patch_nodes");
        auto j = std::make_unique<Job>();
        j->desc = *d;
        const size_t ne = static_cast<size_t>(d->grid.nx) * d->grid.ny * d->grid.nz;
        const size_t ndof = 3 * static_cast<size_t>(d->grid.nx + 1) * (d->grid.ny + 1) * (d->grid.nz + 1);
        j->domain.assign(d->domain, d->domain + ne);
        if (d->n_fixed > 0) j->fixed.assign(d->fixed_dofs, d->fixed_dofs + d->n_fixed);
        j->forces.assign(d->forces, d->forces + ndof);
        if (d->density_box) j->density.assign(d->density_box, d->density_box + ne);
        if (d->n_patch_nodes > 0) j->patch.assign(d->patch_nodes, d->patch_nodes + d->n_patch_nodes);
        j->desc.domain = nullptr;  // the copies above are the job's own
        j->desc.fixed_dofs = nullptr;
        j->desc.forces = nullptr;
        j->desc.density_box = nullptr;
        j->desc.patch_nodes = nullptr;
        j->footprint = job_footprint(*d);
        j->t_submit = Clock::now();
        std::lock_guard<std::mutex> lk(sv->mu);
        j->id = sv->next_id++;
        *id = j->id;
        sv->queue.push_back(j.get());
        sv->jobs[j->id] = std::move(j);
        sv->cv_work.notify_all();
        return MYCEL_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return MYCEL_E_OUT_OF_MEMORY;
    }
}

int32_t mycel_server_poll(mycel_server *sv, int64_t id, mycel_job_info *info) {
    if (!sv || !info) return MYCEL_E_ARGUMENT_NULL;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(id);
    if (it == sv->jobs.end()) {
        set_last_error("unknown job id");
        return MYCEL_E_ARGUMENT;
    }
    fill_info(*it->second, info);
    return MYCEL_OK;
}

int32_t mycel_server_wait(mycel_server *sv, int64_t id, double timeout_s, mycel_job_info *info) {
    if (!sv || !info) return MYCEL_E_ARGUMENT_NULL;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(id);
    if (it == sv->jobs.end()) {
        set_last_error("unknown job id");
        return MYCEL_E_ARGUMENT;
    }
    Job *j = it->second.get();
    auto finished = [&] { return j->state == MYCEL_JOB_DONE || j->state == MYCEL_JOB_FAILED; };
    if (timeout_s < 0) sv->cv_done.wait(lk, finished);
    else sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished);
    fill_info(*j, info);
    if (j->state == MYCEL_JOB_FAILED) set_last_error(j->res.error);
    return MYCEL_OK;
}

int32_t mycel_server_fetch(mycel_server *sv, int64_t id, double *u_out, double *rho_bar_out) {
    if (!sv) return MYCEL_E_ARGUMENT_NULL;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(id);
    if (it == sv->jobs.end()) {
        set_last_error("unknown job id");
        return MYCEL_E_ARGUMENT;
    }
    Job &j = *it->second;
    if (j.state != MYCEL_JOB_DONE) {
        set_last_error(j.state == MYCEL_JOB_FAILED ? j.res.error : "job not finished");
        return j.state == MYCEL_JOB_FAILED ? j.res.status : MYCEL_E_INVALID_OPERATION;
    }
    if (u_out) std::copy(j.res.u.begin(), j.res.u.end(), u_out);
    if (rho_bar_out) std::copy(j.res.rho_bar.begin(), j.res.rho_bar.end(), rho_bar_out);
    return MYCEL_OK;
}

int32_t mycel_server_release(mycel_server *sv, int64_t id) {
    if (!sv) return MYCEL_E_ARGUMENT_NULL;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(id);
    if (it == sv->jobs.end()) return MYCEL_E_ARGUMENT;
    if (it->second->state == MYCEL_JOB_QUEUED || it->second->state == MYCEL_JOB_RUNNING) {
        set_last_error("job still in flight");
        return MYCEL_E_INVALID_OPERATION;
    }
    sv->jobs.erase(it);
    return MYCEL_OK;
}

int32_t mycel_server_stats(mycel_server *sv, double *budget_bytes, double *in_use_bytes, int32_t *queued,
                           int32_t *active) {
    if (!sv) return MYCEL_E_ARGUMENT_NULL;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (budget_bytes) *budget_bytes = sv->budget;
    if (in_use_bytes) *in_use_bytes = sv->in_use;
    if (queued) *queued = static_cast<int32_t>(sv->queue.size());
    if (active) *active = sv->active;
    return MYCEL_OK;
}

}  // extern "C"
