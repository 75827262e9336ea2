// cupic_main.cpp -- C++ host driver over the C-ABI (counterpart of the
// reference's src/main.cu:26-117).
//
//   cupic_main [--input PATH | --config small|prod] [--steps N] [--out DIR]
//              [--save-every K] [--jacobi-mode M] [--restart N]
//              [--restart-dir DIR] [--quiet]
//
// Runs init_sim, then the fixed cycle deposit -> solve -> field -> move ->
// cc(t) with t accumulated as t += dt (src/main.cu:57-73), writes particle
// snapshots in the reference's text format every K steps (src/main.cu:79-83)
// and prints the time, particle counts and the probe tally.  With n_ini > 0
// (from the input file, or --restart) the run resumes at t = n_ini dt from
// DIR/ions.dat and DIR/electrons.dat (src/init.cu:84-95, 338-341; DIR
// defaults to the working directory, as in the reference).  With --input the
// file's n_ini line is set to the last step at the end (src/main.cu:98-113).
// Errors are reported with their status and message; the exit code is
// non-zero.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "cupic.h"

static int fail(cupic_run *r, int st, const char *what)
{
    std::fprintf(stderr, "cupic_main: %s failed (%d): %s\n", what, st, cupic_last_error(r));
    return 1;
}

int main(int argc, char **argv)
{
    std::string input, config = "small", out, restart_dir = ".";
    int steps = -1, save_every = 0, mode = -1, restart = -1;
    bool quiet = false;
    for (int k = 1; k < argc; k++) {
        std::string a = argv[k];
        auto next = [&]() -> const char * { return k + 1 < argc ? argv[++k] : ""; };
        if (a == "--input") input = next();
        else if (a == "--config") config = next();
        else if (a == "--steps") steps = std::atoi(next());
        else if (a == "--out") out = next();
        else if (a == "--save-every") save_every = std::atoi(next());
        else if (a == "--jacobi-mode") mode = std::atoi(next());
        else if (a == "--restart") restart = std::atoi(next());
        else if (a == "--restart-dir") restart_dir = next();
        else if (a == "--quiet") quiet = true;
        else {
            std::fprintf(stderr, "usage: %s [--input PATH | --config small|prod] [--steps N] "
                                 "[--out DIR] [--save-every K] [--jacobi-mode M] [--restart N] "
                                 "[--restart-dir DIR] [--quiet]\n",
                         argv[0]);
            return 2;
        }
    }
    char name[256];
    int cus = 0;
    int st = cupic_init_dev(0, name, sizeof name, &cus);
    if (st) return fail(nullptr, st, "init_dev");
    std::printf("device 0: %s, %d compute units\n", name, cus);

    cupic_params p;
    int32_t n_ini = 0, n_prev = 0, n_save = 0, n_fin = 1000;
    if (!input.empty()) {
        st = cupic_params_from_input_file(input.c_str(), &p, &n_ini, &n_prev, &n_save, &n_fin);
        if (st) return fail(nullptr, st, "params_from_input_file");
    } else {
        cupic_params_default(&p);
        if (config == "prod") {
            p.ncx = 1024; p.ncy = 512; p.n = 200.0 / (p.ds * p.ds * p.ds);
            p.mi = 1836.0; p.kti = 0.1; p.phi_p = -10.0;
        } else if (config != "small") {
            std::fprintf(stderr, "unknown config %s\n", config.c_str());
            return 2;
        }
        st = cupic_params_derive(&p);
        if (st) return fail(nullptr, st, "params_derive");
    }
    if (mode >= 0) p.jacobi_mode = mode;
    if (restart >= 0) n_ini = restart;
    if (steps >= 0) n_fin = steps;
    if (n_ini < 0) {
        std::fprintf(stderr, "Wrong input parameter (n_ini<0)\n");  // src/init.cu:97-99
        return 2;
    }
    cupic_run *r = nullptr;
    st = cupic_run_create(&p, 0, &r);
    if (st) return fail(nullptr, st, "run_create");
    double t = 0.0;
    int64_t ne = 0, ni = 0;
    if (n_ini == 0) {
        st = cupic_init_sim(r, &t);
        if (st) return fail(r, st, "init_sim");
        cupic_number_of_particles(r, 0, &ne);
        cupic_number_of_particles(r, 1, &ni);
        std::printf("Simulation initialized with %lld particles.\n", (long long)(ne + ni));
    } else {
        const std::string ions = restart_dir + "/ions.dat", electrons = restart_dir + "/electrons.dat";
        st = cupic_init_sim_restart(r, n_ini, ions.c_str(), electrons.c_str(), &t);
        if (st) return fail(r, st, "init_sim_restart");
        std::printf("Simulation state loaded from time t = %g\n", t);
    }
    auto snapshot = [&](int i) {
        if (out.empty()) return 0;
        char fn[512];
        std::snprintf(fn, sizeof fn, "%s/electrons_t_%d", out.c_str(), i);
        int s1 = cupic_particles_snapshot(r, 0, fn, t);
        std::snprintf(fn, sizeof fn, "%s/ions_t_%d", out.c_str(), i);
        int s2 = cupic_particles_snapshot(r, 1, fn, t);
        return s1 ? s1 : s2;
    };
    if ((st = snapshot(n_ini))) return fail(r, st, "particles_snapshot");
    t += p.dt;
    for (int i = n_ini + 1; i <= n_fin; i++, t += p.dt) {
        st = cupic_step(r, t);
        if (st) return fail(r, st, "step");
        bool save = save_every > 0 && i % save_every == 0;
        if (save && (st = snapshot(i))) return fail(r, st, "particles_snapshot");
        if (!quiet && (i % 100 == 0 || i == n_fin)) {
            int64_t lp[2], lt[2], li[2], tp[2];
            double q = 0.0;
            st = cupic_probe_tally(r, lp, lt, li, tp, &q);
            if (st) return fail(r, st, "probe_tally");
            cupic_number_of_particles(r, 0, &ne);
            cupic_number_of_particles(r, 1, &ni);
            int32_t sw = 0;
            double res = 0.0;
            cupic_last_solve(r, &sw, &res);
            std::printf("t = %.6g  step %d  Ne %lld  Ni %lld  probe e %lld i %lld  "
                        "charge %.0f  sweeps %d  resid %.3g\n",
                        t, i, (long long)ne, (long long)ni, (long long)tp[0], (long long)tp[1],
                        q, sw, res);
        }
    }
    st = cupic_synchronize(r);
    if (st) return fail(r, st, "synchronize");
    if (!input.empty() && (st = cupic_input_file_set_n_ini(input.c_str(), n_fin)))
        return fail(r, st, "input_file_set_n_ini");
    std::printf("Simulation finished!\n");
    cupic_run_destroy(r);
    return 0;
}
