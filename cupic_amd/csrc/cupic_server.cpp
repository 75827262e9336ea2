// cupic_server.cpp -- continuous batching of I-V sweep points.
//
// A request is one bias point: a full parameter set (grid, density, bias,
// seed).  The server keeps every admitted run resident on the device and
// advances all of them together with one batched step per time step (one
// launch per kernel, replayed as a HIP graph), in rounds of round_steps
// steps.  A poll:
//
//   1. enqueues the next round of the current batch (asynchronous);
//   2. admits queued requests, in arrival order or fewest particles first
//      (the policy), while slots, the HBM budget and the resident-particle
//      budget allow, and starts their init_sim on their own streams, so
//      that initialisation overlaps the round on the device;
//   3. waits for the round, updates each run's probe-current windows and
//      retires the runs that converged or reached max_steps;
//   4. when the membership changed, rebuilds the batch (and its graph) from
//      the remaining runs plus the newly initialised ones.
//
// The particle budget bounds the work resident at once: beyond what fills
// the device, another resident point adds no throughput but slows every
// point sharing the device, so a budget near the device's saturation lets
// points finish in turn (lower latency) at the same rate.
//
// Convergence: every window_steps steps of a run, the mean probe currents of
// the window are computed from the tally (electrons and ions absorbed at
// y < 0).  The run has converged when the net current (ions - electrons)
// changed by less than rel_change times the gross current (ions + electrons)
// since the previous window.  The gross current is the scale because the net
// current passes through zero at the floating potential.
//
// A run's arithmetic does not depend on which other runs share its batch
// (per-run state, per-run solver flags, integer charge), so a point's result
// is the same alone and in any mix.
//
// Only the public C-ABI of the library is used here.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "cupic.h"

namespace {

struct Request {
    cupic_point pt{};
    cupic_params p{};
    cupic_run *run = nullptr;
    int64_t footprint = 0;
    int64_t nparticles = 0;          // initial load, both species
    int64_t win_start[2] = {0, 0};   // cumulative probe counts at the window start
    double prev_net = 0.0;
    bool have_prev = false;
    std::string msg;
};

}  // namespace

struct cupic_server {
    cupic_server_config cfg{};
    std::map<int64_t, Request> reqs;
    std::deque<int64_t> queue;
    std::vector<int64_t> active;    // members of the current batch
    std::vector<int64_t> joining;   // initialising, join at the next rebuild
    cupic_batch *batch = nullptr;
    int64_t used = 0;               // HBM bytes of resident runs
    int64_t used_particles = 0;     // their initial loads
    int64_t next_id = 0;
    int64_t rounds = 0, rebuilds = 0;
    std::chrono::steady_clock::time_point t0;
    std::string err;
};

namespace {

double now(const cupic_server *s)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

int fail(cupic_server *s, int code, const std::string &m)
{
    s->err = m;
    return code;
}

void finish(cupic_server *s, Request &q, int status)
{
    q.pt.status = status;
    q.pt.t_done = now(s);
    if (q.run) {
        cupic_run_destroy(q.run);
        q.run = nullptr;
        s->used -= q.footprint;
        s->used_particles -= q.nparticles;
    }
}

// window bookkeeping after a round; returns true when the run is done
bool update_windows(cupic_server *s, Request &q)
{
    int64_t lp[2], lt[2], li[2], tot[2];
    double charge = 0.0;
    int st = cupic_probe_tally(q.run, lp, lt, li, tot, &charge);
    if (st) {
        q.msg = cupic_last_error(q.run);
        q.pt.error = st;
        finish(s, q, CUPIC_POINT_FAILED);
        return true;
    }
    q.pt.total_probe[0] = tot[0];
    q.pt.total_probe[1] = tot[1];
    for (int sp = 0; sp < 2; sp++) {
        int64_t n = 0;
        cupic_number_of_particles(q.run, sp, &n);
        q.pt.particles[sp] = n;
    }
    if (q.pt.steps % s->cfg.window_steps == 0) {
        double span = s->cfg.window_steps * q.p.dt;
        double ie = (double)(tot[0] - q.win_start[0]) / span;
        double ii = (double)(tot[1] - q.win_start[1]) / span;
        q.win_start[0] = tot[0];
        q.win_start[1] = tot[1];
        q.pt.current[0] = ie;
        q.pt.current[1] = ii;
        double net = ii - ie, gross = ii + ie;
        q.pt.windows += 1;
        if (q.have_prev) {
            q.pt.rel_change = gross > 0.0 ? std::fabs(net - q.prev_net) / gross : INFINITY;
            if (gross > 0.0 && q.pt.rel_change < s->cfg.rel_change) {
                finish(s, q, CUPIC_POINT_CONVERGED);
                return true;
            }
        }
        q.prev_net = net;
        q.have_prev = true;
    }
    if (s->cfg.max_steps > 0 && q.pt.steps >= s->cfg.max_steps) {
        finish(s, q, CUPIC_POINT_CAPPED);
        return true;
    }
    return false;
}

// The queued request the policy admits next (the queue is in arrival order).
std::deque<int64_t>::iterator next_queued(cupic_server *s)
{
    if (s->cfg.policy != CUPIC_SERVE_SMALLEST_FIRST) return s->queue.begin();
    return std::min_element(s->queue.begin(), s->queue.end(), [s](int64_t a, int64_t b) {
        return s->reqs[a].nparticles < s->reqs[b].nparticles;  // first of equals kept
    });
}

void admit(cupic_server *s)
{
    while (!s->queue.empty()) {
        auto it = next_queued(s);
        int64_t id = *it;
        Request &q = s->reqs[id];
        int nres = (int)(s->active.size() + s->joining.size());
        if (nres >= s->cfg.max_active || s->used + q.footprint > s->cfg.hbm_budget) break;
        if (nres > 0 && s->cfg.max_particles > 0 &&
            s->used_particles + q.nparticles > s->cfg.max_particles)
            break;
        s->queue.erase(it);
        int st = cupic_run_create(&q.p, 0, &q.run);
        double t = 0.0;
        if (!st) st = cupic_init_sim(q.run, &t);
        if (st) {
            q.msg = cupic_last_error(q.run);
            q.pt.error = st;
            if (q.run) cupic_run_destroy(q.run);
            q.run = nullptr;
            q.pt.status = CUPIC_POINT_FAILED;
            q.pt.t_done = now(s);
            continue;
        }
        s->used += q.footprint;
        s->used_particles += q.nparticles;
        q.pt.status = CUPIC_POINT_RUNNING;
        q.pt.t_admit = now(s);
        s->joining.push_back(id);
    }
}

int rebuild(cupic_server *s)
{
    if (s->batch) {
        cupic_batch_destroy(s->batch);
        s->batch = nullptr;
    }
    s->active.insert(s->active.end(), s->joining.begin(), s->joining.end());
    s->joining.clear();
    if (s->active.empty()) return CUPIC_OK;
    std::vector<cupic_run *> runs;
    for (int64_t id : s->active) runs.push_back(s->reqs[id].run);
    int st = cupic_batch_create(runs.data(), (int32_t)runs.size(), s->cfg.use_graph, &s->batch);
    if (st) return fail(s, st, std::string("batch: ") + cupic_batch_last_error(nullptr));
    s->rebuilds++;
    return CUPIC_OK;
}

}  // namespace

extern "C" {

int cupic_server_config_default(cupic_server_config *c)
{
    if (!c) return CUPIC_ERR_ARG;
    *c = cupic_server_config{};
    c->hbm_budget = (int64_t)200 << 30;
    c->max_active = 256;
    c->round_steps = 50;
    c->window_steps = 500;
    c->max_steps = 20000;
    c->rel_change = 0.005;
    c->use_graph = 1;
    // a point must reach a physically steady state: the capped Jacobi solve
    // cannot keep grids of hundreds of rows quasi-neutral over thousands of
    // steps, the direct solve of the same system can
    c->jacobi_mode = CUPIC_POISSON_DIRECT;
    c->jacobi_max_sweeps = 500;
    c->jacobi_check_every = 50;
    c->jacobi_tol = 1e-6;
    // about what fills one MI355X: a 50-point stream served as fast with
    // this budget as with none (0.508 against 0.502 points/s) and with half
    // the median latency (49 s against 87 s)
    c->max_particles = 300000000;
    c->policy = CUPIC_SERVE_FCFS;
    return CUPIC_OK;
}

int cupic_server_create(const cupic_server_config *cfg, cupic_server **out)
{
    if (!cfg || !out) return CUPIC_ERR_ARG;
    *out = nullptr;
    if (cfg->round_steps < 1 || cfg->window_steps < 1 || cfg->window_steps % cfg->round_steps ||
        cfg->max_active < 1 || cfg->hbm_budget <= 0 || !(cfg->rel_change > 0.0) ||
        cfg->jacobi_max_sweeps < 1 || cfg->jacobi_check_every < 1 || cfg->max_particles < 0 ||
        (cfg->policy != CUPIC_SERVE_FCFS && cfg->policy != CUPIC_SERVE_SMALLEST_FIRST))
        return CUPIC_ERR_ARG;
    cupic_server *s = new cupic_server();
    s->cfg = *cfg;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return CUPIC_OK;
}

int cupic_server_submit(cupic_server *s, const cupic_params *p, int64_t *id)
{
    if (!s || !p) return CUPIC_ERR_ARG;
    Request q;
    q.p = *p;
    q.p.jacobi_mode = s->cfg.jacobi_mode;
    q.p.jacobi_max_sweeps = s->cfg.jacobi_max_sweeps;
    q.p.jacobi_check_every = s->cfg.jacobi_check_every;
    q.p.jacobi_tol = s->cfg.jacobi_tol;
    q.p.jacobi_tol_init = s->cfg.jacobi_tol;
    int st = cupic_params_derive(&q.p);
    if (st) return fail(s, st, std::string("params: ") + cupic_last_error(nullptr));
    st = cupic_run_footprint(&q.p, 0, &q.footprint);
    if (st) return fail(s, st, "footprint");
    if (q.footprint > s->cfg.hbm_budget)
        return fail(s, CUPIC_ERR_ALLOC, "request larger than the HBM budget");
    q.nparticles = 2 * q.p.per_row * (int64_t)q.p.ncy;  // src/init.cu:136, both species
    q.pt.id = s->next_id++;
    q.pt.status = CUPIC_POINT_QUEUED;
    q.pt.phi_p = q.p.phi_p;
    q.pt.footprint = q.footprint;
    q.pt.t_submit = now(s);
    q.pt.t_admit = q.pt.t_done = -1.0;
    s->reqs[q.pt.id] = q;
    s->queue.push_back(q.pt.id);
    if (id) *id = q.pt.id;
    return CUPIC_OK;
}

int cupic_server_poll(cupic_server *s, int32_t *n_active, int32_t *n_queued)
{
    if (!s) return CUPIC_ERR_ARG;
    int st = CUPIC_OK;
    bool stepped = false;
    if (s->batch) {
        st = cupic_batch_launch(s->batch, s->cfg.round_steps);
        if (st) return fail(s, st, std::string("round: ") + cupic_batch_last_error(s->batch));
        stepped = true;
    }
    admit(s);   // overlaps the round
    bool changed = !s->joining.empty();
    if (stepped) {
        st = cupic_batch_wait(s->batch);
        if (st) return fail(s, st, std::string("round: ") + cupic_batch_last_error(s->batch));
        s->rounds++;
        std::vector<int64_t> keep;
        for (int64_t id : s->active) {
            Request &q = s->reqs[id];
            q.pt.steps += s->cfg.round_steps;
            if (update_windows(s, q)) changed = true;
            else keep.push_back(id);
        }
        if (changed && s->batch) {
            // retired runs were destroyed: the batch must go before it is used again
            cupic_batch_destroy(s->batch);
            s->batch = nullptr;
        }
        s->active.swap(keep);
    }
    if (changed || (!s->batch && !s->active.empty())) {
        st = rebuild(s);
        if (st) return st;
    }
    if (n_active) *n_active = (int32_t)(s->active.size() + s->joining.size());
    if (n_queued) *n_queued = (int32_t)s->queue.size();
    return CUPIC_OK;
}

int cupic_server_result(cupic_server *s, int64_t id, cupic_point *out)
{
    if (!s || !out) return CUPIC_ERR_ARG;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(s, CUPIC_ERR_ARG, "no such request");
    *out = it->second.pt;
    return CUPIC_OK;
}

int cupic_server_stats(cupic_server *s, int64_t *rounds, int64_t *rebuilds, int64_t *hbm_used,
                       double *clock)
{
    if (!s) return CUPIC_ERR_ARG;
    if (rounds) *rounds = s->rounds;
    if (rebuilds) *rebuilds = s->rebuilds;
    if (hbm_used) *hbm_used = s->used;
    if (clock) *clock = now(s);
    return CUPIC_OK;
}

int cupic_server_destroy(cupic_server *s)
{
    if (!s) return CUPIC_OK;
    if (s->batch) cupic_batch_destroy(s->batch);
    for (auto &kv : s->reqs)
        if (kv.second.run) cupic_run_destroy(kv.second.run);
    delete s;
    return CUPIC_OK;
}

const char *cupic_server_last_error(const cupic_server *s)
{
    return s ? s->err.c_str() : "";
}

}  // extern "C"
