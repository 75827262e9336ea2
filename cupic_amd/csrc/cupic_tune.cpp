// cupic_tune.cpp -- autotuner for the push passes' tile (cells per wave
// strip).  Each candidate runs the real cycle on a run of its own and is
// timed with the library's per-kernel HIP-event timers; only the two
// particle passes count, since nothing else depends on the tile.  Uses the
// public C-ABI only.
#include <cstring>
#include <string>

#include "cupic.h"

namespace {

const int kCand[] = {1, 2, 3, 4, 6, 8, 10, 12, 16};

// positions of push_rank and scatter among the timing classes
bool timing_classes(int *ipush, int *iscat, int *ncls)
{
    std::string s(cupic_timing_names()), tok;
    *ipush = *iscat = -1;
    *ncls = 0;
    size_t a = 0;
    while (a <= s.size()) {
        size_t b = s.find(',', a);
        if (b == std::string::npos) b = s.size();
        tok = s.substr(a, b - a);
        if (tok == "push_rank") *ipush = *ncls;
        if (tok == "scatter") *iscat = *ncls;
        (*ncls)++;
        a = b + 1;
    }
    return *ipush >= 0 && *iscat >= 0 && *ncls <= 32;
}

// ms per step of the two push passes over `steps` steps of r from *t
int time_passes(cupic_run *r, double dt, double *t, int steps, int ipush, int iscat, int ncls,
                double *per)
{
    double ms[32] = {};
    int64_t n[32] = {};
    int st = cupic_synchronize(r);
    if (!st) st = cupic_timing_enable(r, 1);
    if (!st) st = cupic_timing_read(r, ms, n, ncls);
    for (int k = 0; k < steps && !st; k++) {
        st = cupic_step(r, *t);
        *t += dt;
    }
    if (!st) st = cupic_synchronize(r);
    if (!st) st = cupic_timing_read(r, ms, n, ncls);
    if (!st) st = cupic_timing_enable(r, 0);
    *per = (ms[ipush] + ms[iscat]) / steps;
    return st;
}

}  // namespace

extern "C" int cupic_autotune_run(cupic_run *r, double *t, int32_t steps, int32_t *best_tile,
                                  double *best_ms, double *times)
{
    if (!r || !t || steps < 1) return CUPIC_ERR_ARG;
    int ipush, iscat, ncls;
    if (!timing_classes(&ipush, &iscat, &ncls)) return CUPIC_ERR_STATE;
    if (times)
        for (int c = 0; c <= 16; c++) times[c] = 0.0;
    cupic_params q;
    int st0 = cupic_run_params(r, &q);
    if (st0) return st0;
    const double dt = q.dt;
    int best = 0;
    double bms = 0.0;
    for (int c : kCand) {
        int st = cupic_run_set_tile(r, c);
        if (st) return st;
        st = cupic_step(r, *t);   // settle
        *t += dt;
        if (st) return st;
        double per = 0.0;
        st = time_passes(r, dt, t, steps, ipush, iscat, ncls, &per);
        if (st) return st;
        if (times) times[c] = per;
        if (best == 0 || per < bms) {
            best = c;
            bms = per;
        }
    }
    int st = cupic_run_set_tile(r, best);
    if (st) return st;
    if (best_tile) *best_tile = best;
    if (best_ms) *best_ms = bms;
    return CUPIC_OK;
}

extern "C" int cupic_autotune_tile(const cupic_params *p, int32_t steps, int32_t *best_tile,
                                   double *best_ms, double *times)
{
    if (!p || steps < 1) return CUPIC_ERR_ARG;
    const int *cand = kCand;
    if (times)
        for (int c = 0; c <= 16; c++) times[c] = 0.0;
    int ipush, iscat, ncls;
    if (!timing_classes(&ipush, &iscat, &ncls)) return CUPIC_ERR_STATE;
    int best = 0;
    double bms = 0.0;
    for (int ci = 0; ci < (int)(sizeof(kCand) / sizeof(kCand[0])); ci++) {
        int c = cand[ci];
        cupic_params q = *p;
        q.tile_cells = c;
        cupic_run *r = nullptr;
        int st = cupic_run_create(&q, 0, &r);
        double t = 0.0;
        if (!st) st = cupic_init_sim(r, &t);
        for (int k = 0; k < 2 && !st; k++) {
            t += q.dt;
            st = cupic_step(r, t);
        }
        double per = 0.0;
        t += q.dt;
        if (!st) st = time_passes(r, q.dt, &t, steps, ipush, iscat, ncls, &per);
        if (r) cupic_run_destroy(r);
        if (st) return st;
        if (times) times[c] = per;
        if (best == 0 || per < bms) {
            best = c;
            bms = per;
        }
    }
    if (best_tile) *best_tile = best;
    if (best_ms) *best_ms = bms;
    return CUPIC_OK;
}
