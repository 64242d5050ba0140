// serve.cpp -- continuous-batching SCF server (include/hfx_serve.h).
//
// Worker threads (one per in-flight slot) pull admissible jobs from a FIFO
// queue.  A job is admissible when its HBM footprint fits what the running
// jobs leave of the budget; each runs its whole SCF on its own HIP stream
// (every hfx_basis owns a non-blocking stream), so kernels of different jobs
// overlap on the device while the host side of each job (DIIS, eigen-solves)
// runs on its own thread.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/hfx.h"
#include "../../../include/hfx_serve.h"
#include "scf.hpp"

namespace hfx_detail {
void set_error(const std::string& msg);
void disable_graphs_in_this_thread();
}

namespace {

using clk = std::chrono::steady_clock;

struct Job {
    long long id = 0;
    hfxh::Config cfg;
    std::string bset, bpath;
    hfx_scf_opts scf{};
    std::vector<double> P0;
    int K = 0;
    size_t bytes_stored = 0, bytes_direct = 0, charged = 0;
    int mode = HFX_MODE_STORED;
    hfx_job_result res{};
    std::vector<double> P;
    bool done = false;
};

// HBM footprint model of one job: stored ERIs (Q doubles) or the direct
// quartet lists (upper bound: all shell-quartet pairs), plus six K x K device
// matrices and ~1 MiB of shell / primitive-pair tables.
void footprint(const hfxh::FlatBasis& fb, const hfxh::BasisInfo& info, size_t& stored, size_t& direct) {
    const size_t K = fb.K;
    size_t nshell = 0;   // a shell starts where the CSV shell id or the centre changes
    for (size_t i = 0; i < info.shells.size(); ++i)
        if (i == 0 || info.shells[i] != info.shells[i - 1] || info.cx[i] != info.cx[i - 1] ||
            info.cy[i] != info.cy[i - 1] || info.cz[i] != info.cz[i - 1])
            ++nshell;
    const size_t psh = nshell * (nshell + 1) / 2;
    const size_t nq = psh * (psh + 1) / 2;
    const size_t P = K * (K + 1) / 2;
    const size_t Q = P * (P + 1) / 2;
    const size_t fixed = 48 * K * K + (1u << 20);
    stored = 8 * Q + 8 * nq + fixed;
    direct = 8 * nq + fixed;
}

}  // namespace

struct hfx_server {
    hfx_server_opts opts{};
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> queue;
    std::vector<std::unique_ptr<Job>> jobs;
    size_t inflight_bytes = 0;
    int inflight = 0;
    bool stopping = false;
    std::vector<std::thread> workers;
    clk::time_point t0 = clk::now();
    int peak_inflight = 0;
    size_t peak_bytes = 0;
    long long n_done = 0;

    double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }

    // Chooses the mode and charge for a job against the free budget; false
    // when it cannot start now.
    bool admissible(Job* j, size_t free_bytes, bool unlimited) {
        const int forced = j->scf.eri_mode;
        // both forms prefer stored ERIs when they fit (the mixed form then
        // stores its FP32-evaluated small-bound integrals in the FP64 array)
        const bool may_store = forced != HFX_MODE_DIRECT;
        if (may_store && (unlimited || j->bytes_stored <= free_bytes)) {
            j->mode = HFX_MODE_STORED;
            j->charged = j->bytes_stored;
            return true;
        }
        if (forced == HFX_MODE_STORED && may_store) return false;
        if (unlimited || j->bytes_direct <= free_bytes) {
            j->mode = HFX_MODE_DIRECT;
            j->charged = j->bytes_direct;
            return true;
        }
        return false;
    }

    void run_job(Job* j) {
        hfxh::ScfOptions o;
        o.charge = j->scf.charge;
        o.max_it = j->scf.max_iterations;
        o.epsilon = j->scf.epsilon;
        o.damping = j->scf.damping;
        o.screening_threshold = j->scf.screening_threshold;
        o.eri_mode = j->mode;
        o.precision = j->scf.precision;
        o.mixed_split = j->scf.mixed_split;
        o.energy_convergence = j->scf.energy_convergence;
        o.hbm_budget = 0;
        o.polish = j->scf.polish != 0;
        o.verbose = false;
        o.device_iteration = j->scf.device_iteration;
        o.incremental_fock = j->scf.incremental_fock;
        try {
            auto fill = [&](auto& r, int Ne) {
                hfx_scf_result& s = j->res.scf;
                s.E_tot = r.E_tot;
                s.E_ele = r.E_ele;
                s.E_nuc = r.E_nuc;
                s.E_before_polish = r.E_before_polish;
                s.iterations = r.iterations;
                s.polish_iterations = r.polish_iterations;
                s.converged = r.converged;
                s.K = r.K;
                s.n_electrons = Ne;
                s.mode = r.mode;
                s.n_quartets = (long long)r.n_quartets;
                s.t_setup = r.t.setup;
                s.t_eri = r.t.eri;
                s.t_fock = r.t.fock;
                s.t_diag = r.t.diag;
                s.t_total = r.t.total;
            };
            if (j->scf.n_alpha < 0) {
                auto r = hfxh::rhf_diis(j->cfg, j->bset, j->bpath, o, j->P0.empty() ? nullptr : &j->P0);
                fill(r, r.Ne);
                j->P = std::move(r.density);
            } else {
                auto r = hfxh::uhf_diis(j->cfg, j->bset, j->bpath, {j->scf.n_alpha, j->scf.n_beta}, o);
                fill(r, r.Na + r.Nb);
                j->P = r.Pa;
                for (size_t i = 0; i < j->P.size(); ++i) j->P[i] += r.Pb[i];
            }
            j->res.status = 0;
        } catch (const std::exception& e) {
            j->res.status = HFX_ERR_INVALID;
            std::strncpy(j->res.error, e.what(), sizeof(j->res.error) - 1);
        }
    }

    void worker() {
        // jobs run side by side on their own threads: no stream capture here
        hfx_detail::disable_graphs_in_this_thread();
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            Job* pick = nullptr;
            cv_work.wait(lk, [&] {
                if (stopping && queue.empty()) return true;
                const bool unlimited = opts.hbm_budget == 0;
                const size_t free_bytes = unlimited ? 0 : (opts.hbm_budget > inflight_bytes ? opts.hbm_budget - inflight_bytes : 0);
                for (auto it = queue.begin(); it != queue.end(); ++it) {
                    Job* j = *it;
                    if (admissible(j, free_bytes, unlimited)) {
                        pick = j;
                        queue.erase(it);
                        return true;
                    }
                    // a job too large for an empty device can never run: fail it
                    if (inflight == 0 && !unlimited && j->bytes_direct > opts.hbm_budget) {
                        j->res.status = HFX_ERR_NOMEM;
                        std::strncpy(j->res.error, "job does not fit the server's HBM budget", sizeof(j->res.error) - 1);
                        j->res.t_admit = j->res.t_done = now();
                        j->done = true;
                        j->res.done = 1;
                        ++n_done;
                        queue.erase(it);
                        cv_done.notify_all();
                        return false;
                    }
                    if (!opts.allow_skip_ahead) break;
                }
                return false;
            });
            if (!pick) {
                if (stopping && queue.empty()) return;
                continue;
            }
            ++inflight;
            inflight_bytes += pick->charged;
            peak_inflight = std::max(peak_inflight, inflight);
            peak_bytes = std::max(peak_bytes, inflight_bytes);
            pick->res.t_admit = now();
            pick->res.device_bytes = pick->charged;
            lk.unlock();
            run_job(pick);
            lk.lock();
            pick->res.t_done = now();
            pick->res.K = pick->K;
            pick->done = true;
            pick->res.done = 1;
            --inflight;
            inflight_bytes -= pick->charged;
            ++n_done;
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

int hfx_server_create(const hfx_server_opts* opts, hfx_server** out) {
    if (!opts || !out || opts->max_inflight < 1) {
        hfx_detail::set_error("invalid server options");
        return HFX_ERR_INVALID;
    }
    if (hfx_device_count() < 1) {
        hfx_detail::set_error("no GPU visible");
        return HFX_ERR_DEVICE;
    }
    hfx_server* s = new hfx_server();
    s->opts = *opts;
    for (int i = 0; i < opts->max_inflight; ++i) s->workers.emplace_back([s] { s->worker(); });
    *out = s;
    return 0;
}

void hfx_server_destroy(hfx_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    delete s;
}

int hfx_server_submit(hfx_server* s, const hfx_job_desc* d, long long* job_id) {
    if (!s || !d || !d->symbols || !d->xyz || !d->basis_set || !d->basis_path) {
        hfx_detail::set_error("null argument");
        return HFX_ERR_INVALID;
    }
    auto j = std::make_unique<Job>();
    try {
        const char* p = d->symbols;
        for (int a = 0; a < d->natom; ++a) {
            j->cfg.push_back({std::string(p), {d->xyz[3 * a], d->xyz[3 * a + 1], d->xyz[3 * a + 2]}});
            p += std::strlen(p) + 1;
        }
        j->bset = d->basis_set;
        j->bpath = d->basis_path;
        j->scf = d->scf;
        hfxh::BasisInfo info;
        auto fb = hfxh::basis_construction(j->cfg, j->bset, j->bpath, &info);
        j->K = fb.K;
        footprint(fb, info, j->bytes_stored, j->bytes_direct);
        if (d->P0) j->P0.assign(d->P0, d->P0 + (size_t)fb.K * fb.K);
    } catch (const std::exception& e) {
        hfx_detail::set_error(e.what());
        return HFX_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->m);
    if (s->stopping) {
        hfx_detail::set_error("server is stopping");
        return HFX_ERR_INVALID;
    }
    j->id = (long long)s->jobs.size();
    j->res.t_submit = s->now();
    *job_id = j->id;
    s->queue.push_back(j.get());
    s->jobs.push_back(std::move(j));
    s->cv_work.notify_one();
    return 0;
}

int hfx_server_result(hfx_server* s, long long id, int wait, hfx_job_result* out) {
    if (!s) return HFX_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->m);
    if (id < 0 || id >= (long long)s->jobs.size()) return HFX_ERR_INVALID;
    Job* j = s->jobs[id].get();
    if (wait) s->cv_done.wait(lk, [&] { return j->done; });
    if (!j->done) return 0;
    if (out) *out = j->res;
    return 1;
}

int hfx_server_drain(hfx_server* s) {
    if (!s) return HFX_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->m);
    s->cv_done.wait(lk, [&] { return s->n_done == (long long)s->jobs.size(); });
    return 0;
}

double hfx_server_now(hfx_server* s) { return s ? s->now() : 0.0; }

int hfx_server_stats(hfx_server* s, int* peak_inflight, size_t* peak_bytes, long long* n_done) {
    if (!s) return HFX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    if (peak_inflight) *peak_inflight = s->peak_inflight;
    if (peak_bytes) *peak_bytes = s->peak_bytes;
    if (n_done) *n_done = s->n_done;
    return 0;
}

int hfx_server_density(hfx_server* s, long long id, double* P) {
    if (!s || !P) return HFX_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->m);
    if (id < 0 || id >= (long long)s->jobs.size() || !s->jobs[id]->done) return HFX_ERR_INVALID;
    const auto& v = s->jobs[id]->P;
    std::memcpy(P, v.data(), v.size() * sizeof(double));
    return 0;
}

}  // extern "C"
