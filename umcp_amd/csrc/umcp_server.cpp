// Serving loop (umcp_server_* in umcp.h): Brownian-dynamics requests (a mesh,
// parameters, N steps) admitted as they arrive and stepped together.
//
// The reference runs one simulation per process (src/hd.C) and has no request
// loop; this is the production front end of the batched ensemble:
//
//   submit()    copies the request into the FIFO queue (any thread)
//   prep threads take the queue head once a slot is free and its device
//               footprint fits the HBM budget: template (parsed and planned
//               on the host once per mesh file) -> clone -> plan -> setg0 ->
//               start positions; the footprint is reserved before any device
//               allocation
//   worker      adds prepared requests to the ensemble at a chunk boundary,
//               runs min(chunk_steps, fewest remaining steps) batched steps of
//               every active request, retires the finished ones (results to
//               host, device memory freed) and repeats
//
// Preparation of new requests (several at once, one per preparation thread,
// each on the request's own stream) overlaps the worker's chunks.  A
// request's trajectory is the one it has alone: the ensemble runs each
// member's own kernels and per-member (seed, step) noise.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/umcp.h"
#include "umcp_kernels.hpp"
#include "umcp_state.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int id = 0;
    std::string mesh;
    umcp_params prm{};
    double c0 = 0.0;
    int steps = 0;
    unsigned long long seed = 0;
    umcp_bd_params bd{};
    int nv = 0;
    std::vector<double> r0, r;          // host copies (setg0 state, start)
    // device state while admitted
    umcp_surface *s = nullptr;
    double *d_energies = nullptr;
    size_t footprint = 0;
    int remaining = 0;
    // results
    std::vector<double> r_out, energies;
    double t_submit = 0, t_admit = 0, t_start = -1, t_done = 0;
    int status = UMCP_OK;
    std::string err;
    bool done = false;
};

}  // namespace

struct umcp_server {
    umcp_server_config cfg{};
    int device = 0;
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_prep, cv_work, cv_done;
    bool stopping = false;
    std::deque<Request *> queue;              // submitted, not yet taken by prep
    std::vector<Request *> ready;             // prepared, not yet in the ensemble
    std::vector<Request *> active;            // in the ensemble
    std::map<int, std::unique_ptr<Request>> all;
    int next_id = 1;
    int preparing = 0;
    size_t used = 0, peak_used = 0;
    int peak_active = 0;
    std::map<std::string, umcp_surface *> templates;
    std::mutex tmpl_mu;
    umcp_ensemble *ens = nullptr;
    hipStream_t stream = nullptr;
    std::thread worker;
    std::vector<std::thread> prep;
    std::string err;
    // statistics
    long chunks = 0;
    double member_steps = 0, face_steps = 0, busy_s = 0;
    double build_s = 0, retire_s = 0, wait_s = 0, setup_s = 0;   // where the worker's and prep's time goes
    int n_done = 0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

umcp_surface *get_template(umcp_server *sv, const std::string &path, std::string &err)
{
    std::lock_guard<std::mutex> lk(sv->tmpl_mu);
    auto it = sv->templates.find(path);
    if (it != sv->templates.end()) return it->second;
    umcp_surface *t = nullptr;
    int rc = umcp_load_mesh(path.c_str(), &t);
    if (rc == UMCP_OK) rc = umcp_prepare(t);
    if (rc != UMCP_OK) {
        err = std::string("template ") + path + ": " + (t ? umcp_last_error(t) : "load failed");
        umcp_destroy(t);
        return nullptr;
    }
    sv->templates[path] = t;
    return t;
}

void finish_failed(umcp_server *sv, Request *q, int code, const std::string &msg)
{
    if (q->s) umcp_destroy(q->s);
    q->s = nullptr;
    if (q->d_energies) (void)hipFree(q->d_energies);
    q->d_energies = nullptr;
    std::lock_guard<std::mutex> lk(sv->mu);
    q->status = code;
    q->err = msg;
    q->done = true;
    q->t_done = sv->now();
    if (q->footprint) sv->used -= q->footprint;
    q->footprint = 0;
    sv->n_done++;
    sv->cv_done.notify_all();
    sv->cv_prep.notify_all();
    // a failed preparation lowers `preparing`, which the worker's exit
    // condition at shutdown waits on
    sv->cv_work.notify_all();
}

// device set-up of one request (prep thread); footprint already reserved
int setup(umcp_server *sv, Request *q, umcp_surface *tmpl)
{
    umcp_surface *s = nullptr;
    int rc = umcp_clone(tmpl, &s);
    if (rc != UMCP_OK) return rc;
    q->s = s;
    umcp_params p = q->prm;
    p.precision = sv->cfg.precision;
    if ((rc = umcp_set_params(s, &p)) || (rc = umcp_plan(s))) return rc;
    int nreg = 0, nirr = 0;
    const int nf = umcp_nfaces(s, &nreg, &nirr);
    std::vector<double> c0(nf, q->c0);
    if ((rc = umcp_set_face_c0(s, c0.data())) || (rc = umcp_setg0(s, q->r0.data())) || (rc = umcp_set_step(s, 0)))
        return rc;
    hipError_t e = hipMemcpyAsync(s->d_r, q->r.data(), q->r.size() * sizeof(double), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMalloc(&q->d_energies, sizeof(double) * (size_t)std::max(q->steps, 1));
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? UMCP_OK : UMCP_ERR_DEVICE;
}

void prep_loop(umcp_server *sv)
{
    (void)hipSetDevice(sv->device);
    (void)umcp_set_device(sv->device);   // handles created on this thread
    for (;;) {
        Request *q = nullptr;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_prep.wait(lk, [&] {
                const int slots = (int)sv->active.size() + (int)sv->ready.size() + sv->preparing;
                return sv->stopping || (!sv->queue.empty() && slots < sv->cfg.max_active);
            });
            if (sv->queue.empty()) {
                if (sv->stopping) return;
                continue;
            }
            q = sv->queue.front();
            sv->queue.pop_front();
            sv->preparing++;
        }
        std::string err;
        umcp_surface *tmpl = get_template(sv, q->mesh, err);
        if (!tmpl) {
            {
                std::lock_guard<std::mutex> lk(sv->mu);
                sv->preparing--;
            }
            finish_failed(sv, q, UMCP_ERR_BAD_MESH, err);
            continue;
        }
        if (tmpl->T.nv != q->nv) {
            {
                std::lock_guard<std::mutex> lk(sv->mu);
                sv->preparing--;
            }
            finish_failed(sv, q, UMCP_ERR_ARG, "request positions do not match the mesh's vertex count");
            continue;
        }
        const size_t need = umcp::plan_bytes(tmpl) + sizeof(double) * (size_t)std::max(q->steps, 1);
        {
            // admission by footprint: wait until the budget holds the request
            std::unique_lock<std::mutex> lk(sv->mu);
            if ((double)need > sv->cfg.budget_bytes) {
                sv->preparing--;
                lk.unlock();
                finish_failed(sv, q, UMCP_ERR_OOM, "request exceeds the server's device budget");
                continue;
            }
            sv->cv_prep.wait(lk, [&] { return sv->stopping || (double)(sv->used + need) <= sv->cfg.budget_bytes; });
            if ((double)(sv->used + need) > sv->cfg.budget_bytes) {   // stopping
                sv->preparing--;
                lk.unlock();
                finish_failed(sv, q, UMCP_ERR_STATE, "server shut down");
                continue;
            }
            sv->used += need;
            sv->peak_used = std::max(sv->peak_used, sv->used);
            q->footprint = need;
        }
        const double t_setup = sv->now();
        const int rc = setup(sv, q, tmpl);
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->setup_s += sv->now() - t_setup;
        }
        if (rc != UMCP_OK) {
            const std::string msg = q->s ? umcp_last_error(q->s) : "set-up failed";
            {
                std::lock_guard<std::mutex> lk(sv->mu);
                sv->preparing--;
            }
            finish_failed(sv, q, rc, msg);
            continue;
        }
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->preparing--;
        q->remaining = q->steps;
        q->t_admit = sv->now();
        sv->ready.push_back(q);
        sv->cv_work.notify_all();
    }
}

void retire(umcp_server *sv, Request *q)
{
    umcp_surface *s = q->s;
    q->r_out.resize(q->r.size());
    q->energies.resize((size_t)q->steps);
    hipError_t e = hipMemcpyAsync(q->r_out.data(), s->d_r, q->r_out.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  sv->stream);
    if (e == hipSuccess && q->steps > 0)
        e = hipMemcpyAsync(q->energies.data(), q->d_energies, q->energies.size() * sizeof(double),
                           hipMemcpyDeviceToHost, sv->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(sv->stream);
    (void)umcp_ensemble_remove(sv->ens, s);
    umcp_destroy(s);
    q->s = nullptr;
    (void)hipFree(q->d_energies);
    q->d_energies = nullptr;
    std::lock_guard<std::mutex> lk(sv->mu);
    q->status = e == hipSuccess ? UMCP_OK : UMCP_ERR_DEVICE;
    if (e != hipSuccess) q->err = hipGetErrorString(e);
    q->done = true;
    q->t_done = sv->now();
    sv->used -= q->footprint;
    q->footprint = 0;
    sv->n_done++;
    sv->cv_done.notify_all();
    sv->cv_prep.notify_all();
}

void work_loop(umcp_server *sv)
{
    (void)hipSetDevice(sv->device);
    for (;;) {
        std::vector<Request *> joining;
        const double t_w = sv->now();
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_work.wait(lk, [&] {
                return !sv->ready.empty() || !sv->active.empty() ||
                       (sv->stopping && sv->queue.empty() && sv->preparing == 0);
            });
            if (sv->ready.empty() && sv->active.empty()) return;   // stopping and drained
            joining.swap(sv->ready);
            sv->wait_s += sv->now() - t_w;
        }
        for (Request *q : joining) {
            const int rc = umcp_ensemble_add(sv->ens, q->s, q->s->d_r, &q->bd, q->seed, q->d_energies,
                                             std::max(q->steps, 1));
            if (rc != UMCP_OK) {
                finish_failed(sv, q, rc, umcp_ensemble_last_error(sv->ens));
                continue;
            }
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->active.push_back(q);
        }
        std::vector<Request *> act;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            act = sv->active;
            sv->peak_active = std::max(sv->peak_active, (int)act.size());
        }
        if (act.empty()) continue;
        int chunk = sv->cfg.chunk_steps;
        for (Request *q : act) chunk = std::min(chunk, q->remaining);
        const double t_0 = sv->now();
        int rc = umcp_ensemble_bd_steps(sv->ens, 0, 0, sv->stream);   // tables of the new membership
        const double t_a = sv->now();
        for (Request *q : act)
            if (q->t_start < 0) q->t_start = t_a;
        if (rc == UMCP_OK && chunk > 0) {
            const int g = sv->cfg.graph_steps > 0 ? std::min(sv->cfg.graph_steps, chunk) : 0;
            rc = umcp_ensemble_bd_steps(sv->ens, chunk, g, sv->stream);
            hipError_t e = hipStreamSynchronize(sv->stream);
            if (rc == UMCP_OK && e != hipSuccess) rc = UMCP_ERR_DEVICE;
        }
        const double t_b = sv->now();
        size_t faces = 0;
        for (Request *q : act) faces += (size_t)q->s->P.nfaces;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            sv->chunks++;
            sv->build_s += t_a - t_0;
            sv->busy_s += t_b - t_a;
            sv->member_steps += (double)chunk * act.size();
            sv->face_steps += (double)chunk * faces;
        }
        for (Request *q : act) {
            q->remaining -= chunk;
            if (rc != UMCP_OK || q->remaining <= 0) {
                {
                    std::lock_guard<std::mutex> lk(sv->mu);
                    sv->active.erase(std::find(sv->active.begin(), sv->active.end(), q));
                }
                if (rc != UMCP_OK) {
                    (void)umcp_ensemble_remove(sv->ens, q->s);
                    finish_failed(sv, q, rc, umcp_ensemble_last_error(sv->ens));
                } else {
                    const double t_r = sv->now();
                    retire(sv, q);
                    std::lock_guard<std::mutex> lk(sv->mu);
                    sv->retire_s += sv->now() - t_r;
                }
            }
        }
    }
}

}  // namespace

extern "C" {

int umcp_server_create(const umcp_server_config *cfg, umcp_server **out)
{
    if (!cfg || !out) return UMCP_ERR_ARG;
    *out = nullptr;
    if (cfg->max_active < 1 || cfg->chunk_steps < 1 || cfg->graph_steps < 0 ||
        (cfg->precision != UMCP_PREC_FP64 && cfg->precision != UMCP_PREC_MIXED))
        return UMCP_ERR_ARG;
    auto *sv = new umcp_server();
    sv->cfg = *cfg;
    if (!(sv->cfg.budget_bytes > 0)) sv->cfg.budget_bytes = 200e9;
    sv->t0 = Clock::now();
    if (hipGetDevice(&sv->device) != hipSuccess || hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking) != hipSuccess ||
        umcp_ensemble_create(&sv->ens) != UMCP_OK) {
        if (sv->stream) (void)hipStreamDestroy(sv->stream);
        delete sv;
        return UMCP_ERR_DEVICE;
    }
    const int rc = umcp_ensemble_set_launch(sv->ens, cfg->point_split, cfg->flags, cfg->face_grid, cfg->order);
    if (rc != UMCP_OK) {
        umcp_ensemble_destroy(sv->ens);
        (void)hipStreamDestroy(sv->stream);
        delete sv;
        return rc;
    }
    sv->worker = std::thread(work_loop, sv);
    const int np = cfg->prep_threads > 0 ? cfg->prep_threads : 1;
    for (int i = 0; i < np; i++) sv->prep.emplace_back(prep_loop, sv);
    *out = sv;
    return UMCP_OK;
}

int umcp_server_submit(umcp_server *sv, const umcp_request *rq, int *id)
{
    if (!sv || !rq || !rq->mesh_path || !rq->r0 || !rq->r || rq->nv <= 0 || rq->steps < 0) return UMCP_ERR_ARG;
    auto q = std::make_unique<Request>();
    q->mesh = rq->mesh_path;
    q->prm = rq->params;
    q->c0 = rq->c0;
    q->steps = rq->steps;
    q->seed = rq->seed;
    q->bd = rq->bd;
    q->nv = rq->nv;
    const size_t n = 3 * (size_t)rq->nv + 3;
    q->r0.assign(rq->r0, rq->r0 + n);
    q->r.assign(rq->r, rq->r + n);
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stopping) return UMCP_ERR_STATE;
    q->id = sv->next_id++;
    q->t_submit = sv->now();
    if (id) *id = q->id;
    sv->queue.push_back(q.get());
    sv->all[q->id] = std::move(q);
    sv->cv_prep.notify_all();
    return UMCP_OK;
}

int umcp_server_wait(umcp_server *sv, int id, double *r_out, double *energies, double *times)
{
    if (!sv) return UMCP_ERR_ARG;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    if (it == sv->all.end()) return UMCP_ERR_ARG;
    Request *q = it->second.get();
    sv->cv_done.wait(lk, [&] { return q->done; });
    if (times) {
        times[0] = q->t_submit;
        times[1] = q->t_admit;
        times[2] = q->t_start;
        times[3] = q->t_done;
    }
    const int status = q->status;
    if (status != UMCP_OK) {
        sv->err = q->err;
    } else {
        if (r_out) std::copy(q->r_out.begin(), q->r_out.end(), r_out);
        if (energies) std::copy(q->energies.begin(), q->energies.end(), energies);
    }
    sv->all.erase(it);
    return status;
}

int umcp_server_drain(umcp_server *sv)
{
    if (!sv) return UMCP_ERR_ARG;
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_done.wait(lk, [&] {
        return sv->queue.empty() && sv->ready.empty() && sv->active.empty() && sv->preparing == 0;
    });
    return UMCP_OK;
}

int umcp_server_stats(umcp_server *sv, double *stats, int n)
{
    if (!sv || !stats || n < 0) return UMCP_ERR_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    const double v[] = {(double)sv->n_done, (double)sv->chunks, sv->member_steps, sv->face_steps, sv->busy_s,
                        (double)sv->peak_used, (double)sv->peak_active, sv->now(), sv->build_s, sv->retire_s,
                        sv->wait_s, sv->setup_s};
    const int m = std::min(n, (int)(sizeof v / sizeof v[0]));
    for (int i = 0; i < m; i++) stats[i] = v[i];
    return m;
}

const char *umcp_server_last_error(const umcp_server *sv) { return sv ? sv->err.c_str() : "null handle"; }

void umcp_server_destroy(umcp_server *sv)
{
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stopping = true;
        // requests never taken by the prep thread are dropped
        for (Request *q : sv->queue) {
            q->status = UMCP_ERR_STATE;
            q->err = "server shut down";
            q->done = true;
        }
        sv->queue.clear();
        sv->cv_prep.notify_all();
        sv->cv_work.notify_all();
        sv->cv_done.notify_all();
    }
    for (std::thread &t : sv->prep)
        if (t.joinable()) t.join();
    {
        // no preparation is left: wake the worker to check its exit condition
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->cv_work.notify_all();
    }
    if (sv->worker.joinable()) sv->worker.join();
    umcp_ensemble_destroy(sv->ens);
    for (auto &kv : sv->templates) umcp_destroy(kv.second);
    (void)hipStreamDestroy(sv->stream);
    delete sv;
}

}  // extern "C"
