/*
 * cupic.h -- C-ABI of the MI355X-native electrostatic PIC cycle.
 *
 * Drop-in boundary for cupic's module entry points (init / mesh / particles /
 * cc / diagnostic).  Every function returns an int status (CUPIC_OK == 0,
 * negative on error) and never exits the process; cupic_last_error() gives
 * the message.  Parameters travel in an explicit cupic_params struct instead
 * of the reference's memoised global getters (src/init.h:40-63), so one
 * process can host many runs.
 *
 * Units are normalised as in the reference (src/init.cu:451-789): lengths in
 * Debye lengths, time in 1/omega_pe, mass in electron masses, potential in
 * kTe/e.  Grids are exported in the reference layout: nny rows of nnx nodes,
 * index i + j*nnx, column nnx-1 being the periodic image of column 0.
 *
 * Threading: one host thread per run at a time.  Each run owns a HIP stream;
 * calls enqueue work and return, except those that return host values.
 */
#ifndef CUPIC_H
#define CUPIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* libcupic.so is built with hidden visibility: only the functions declared
 * here are exported, so no kernel handle or internal helper of the library
 * can bind to, or be bound by, another object in the process. */
#pragma GCC visibility push(default)

#define CUPIC_OK 0
#define CUPIC_ERR_ARG (-1)        /* invalid argument or parameter           */
#define CUPIC_ERR_HIP (-2)        /* HIP runtime error                       */
#define CUPIC_ERR_ALLOC (-3)      /* device allocation failed or over budget */
#define CUPIC_ERR_CAPACITY (-4)   /* particle / injection / far-list overflow */
#define CUPIC_ERR_STATE (-5)      /* call out of order (e.g. cc before mover) */
#define CUPIC_ERR_NUMERIC (-6)    /* non-finite values in the particle state  */
#define CUPIC_ERR_IO (-7)         /* file could not be read or written        */

/* Poisson stopping rules */
#define CUPIC_JACOBI_REFERENCE 0  /* max|dphi| < tol after >= min sweeps (src/mesh.cu:94-114) */
#define CUPIC_JACOBI_RESIDUAL 1   /* FP32 sweeps, FP64 residual checked every check_every     */
#define CUPIC_JACOBI_REFINE 2     /* FP64 residual + FP32 correction solve (refinement)        */
#define CUPIC_POISSON_DIRECT 3    /* the system Jacobi converges to, solved directly in FP64:
                                     DFT along periodic x, sine transform along y, by the
                                     method direct_method selects                           */

/* direct_method: how CUPIC_POISSON_DIRECT transforms */
#define CUPIC_DIRECT_AUTO 0   /* FFT kernels when ncx and 2 ncy are powers of two up to 4096,
                                 else the dense FP64 matrix products                    */
#define CUPIC_DIRECT_DENSE 1  /* always the dense FP64 matrix products                  */
#define CUPIC_DIRECT_FFT 2    /* FFT kernels for any ncx and ncy whose FFT lengths are at
                                 most 4096: a power-of-two length n as one FFT, another
                                 length by Bluestein's chirp z-transform through two FFTs
                                 of the power of two P >= 2n - 1; else the dense products */

/* quirk flags */
#define CUPIC_QUIRK_Q7_REFERENCE_KICK 1u /* injection half-kick sign as committed (src/cc.cu:243-244) */
#define CUPIC_QUIRK_Q10_SKIP_FIRST_LINE 2u /* particle files: skip the first line as the reference's
                                               reader does (src/init.cu:372,392), losing a particle */

typedef struct cupic_params {
    int32_t ncx, ncy;        /* cells (nnx = ncx+1, nny = ncy+1 nodes)           */
    double ds, dt;           /* cell size, time step                             */
    double n;                /* density, particles per Debye volume; ppc = n ds^3 */
    double mi, kti;          /* ion mass / m_e, ion temperature / kTe            */
    double me, kte;          /* 1 and 1 (src/init.cu:487-494, 512-519)           */
    double phi_p;            /* probe bias, kTe/e                                */
    double eps0;             /* permittivity; 0 selects the normalised value n   */
    uint64_t seed, run_id;   /* RNG key                                          */
    uint32_t quirks;
    int32_t jacobi_mode;     /* CUPIC_JACOBI_*                                   */
    int32_t jacobi_min_sweeps;   /* mode 0 minimum (reference: max(nnx,nny)+1)   */
    int32_t jacobi_max_sweeps;   /* cap on sweeps per solve                       */
    int32_t jacobi_check_every;  /* modes 1, 2: sweeps between FP64 residuals    */
    int32_t tile_cells;      /* cells per wave strip in the push passes: 0 picks
                                round(2048 / ppc), 1..16 fixes it (results do not
                                depend on it; see cupic_autotune_tile)           */
    int32_t direct_method;   /* CUPIC_DIRECT_*: the transforms of CUPIC_POISSON_DIRECT */
    double jacobi_tol;       /* per-step tolerance (reference 1e-4, src/main.cu:64) */
    double jacobi_tol_init;  /* init tolerance (reference 1e-3, src/init.cu:254)     */
    /* derived by cupic_params_derive */
    double Lx, Ly;
    double dtin_e, dtin_i;   /* injection intervals (src/init.cu:686-714)        */
    int64_t per_row;         /* initial particles per y row (src/init.cu:136)    */
} cupic_params;

typedef struct cupic_run cupic_run;
typedef struct cupic_batch cupic_batch;

/* ---- device and parameters ------------------------------------------ */

/* Replaces init_dev (src/init.cu:15-55): select a device, report it. */
int cupic_init_dev(int device, char *name, int name_len, int *compute_units);

/* Config-1 defaults: 128x64 cells, ds 0.5, dt 0.05, 20 ppc, mi 100, phi_p -5. */
int cupic_params_default(cupic_params *p);

/* Replaces the init_* getters (src/init.cu:451-789): fills the derived
 * fields and checks the parameters. */
int cupic_params_derive(cupic_params *p);

/* Reads the reference's positional input file (src/init.cu:424-447,
 * input/input_data) and converts its SI values (src/init.cu:523-556,
 * 612-630, 718-733) into normalised parameters. */
int cupic_params_from_input_file(const char *path, cupic_params *p, int32_t *n_ini,
                                 int32_t *n_prev, int32_t *n_save, int32_t *n_fin);

/* ---- run lifetime ---------------------------------------------------- */

/* Allocates every device buffer of one run (the reference's main owns them,
 * src/main.cu:42-44, and init_sim allocates, src/init.cu:147-156, 225-232).
 * capacity_hint: particles per species; 0 picks N + headroom. */
int cupic_run_create(const cupic_params *p, int64_t capacity_hint, cupic_run **out);
int cupic_run_destroy(cupic_run *r);
/* the run's parameters (derived fields included) */
int cupic_run_params(const cupic_run *r, cupic_params *out);
const char *cupic_last_error(const cupic_run *r);   /* r may be NULL */
int cupic_synchronize(cupic_run *r);
/* device bytes this run holds */
int cupic_run_footprint(const cupic_params *p, int64_t capacity_hint, int64_t *bytes);

/* ---- init (src/init.cu) ------------------------------------------------ */

/* init_sim with n_ini == 0 (src/init.cu:57-83): create_particles,
 * initialize_mesh, adjust_leap_frog.  *t is set to 0. */
int cupic_init_sim(cupic_run *r, double *t);
/* create_particles (src/init.cu:106-198): device Philox draws, cell-sorted */
int cupic_create_particles(cupic_run *r);
/* initialize_mesh (src/init.cu:202-260): linear phi, deposit, solve, field */
int cupic_initialize_mesh(cupic_run *r);
/* adjust_leap_frog (src/init.cu:264-323): v -= dt/2 * qE/m */
int cupic_adjust_leap_frog(cupic_run *r);
/* init_sim with n_ini > 0 (src/init.cu:84-95): *t = n_ini dt, both species
 * from particle files (the reference reads ./ions.dat and ./electrons.dat,
 * src/init.cu:338-341), then initialize_mesh; no half kick, the files hold
 * staggered velocities.  Like the reference, the injection clocks start
 * afresh at the first cc call and the field is solved from the linear
 * guess. */
int cupic_init_sim_restart(cupic_run *r, int32_t n_ini, const char *ions_path,
                           const char *electrons_path, double *t);
/* read_particle_file (src/init.cu:353-420): one particle per line in the
 * snapshot format " %.17e %.17e %.17e %.17e " (FP64, rounded to FP32),
 * uploaded cell-sorted as `species`.  Every line is read unless the run's
 * quirks hold CUPIC_QUIRK_Q10_SKIP_FIRST_LINE; blank lines are skipped and
 * any other line that is not four numbers is an error (CUPIC_ERR_IO). */
int cupic_read_particle_file(cupic_run *r, int species, const char *path);
/* The end of src/main.cu (98-113): line 2 of the input file becomes
 * "n_ini = <n_ini>;", the other lines are kept (written to path.new, then
 * renamed over path). */
int cupic_input_file_set_n_ini(const char *path, int32_t n_ini);

/* ---- mesh (src/mesh.h:27-29) ----------------------------------------- */

/* charge_deposition (src/mesh.cu:15-51): int64 fixed-point CIC charge of
 * both species, converted to rho. */
int cupic_charge_deposition(cupic_run *r);
/* poisson_solver (src/mesh.cu:55-120).  tol < 0 uses the run's per-step
 * tolerance.  Returns sweeps run and the final residual (max |update|). */
int cupic_poisson_solver(cupic_run *r, double tol, int32_t *sweeps, double *resid);

/* Device scratch of cupic_poisson_solver_noalloc, held by every GPU the
 * library runs on until the process ends. */
#define CUPIC_NOALLOC_SCRATCH_BYTES (16u << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the Jacobi solve of cupic_poisson_solver, modes
 * CUPIC_JACOBI_REFERENCE and CUPIC_JACOBI_RESIDUAL, on caller-owned device
 * buffers, with every temporary in a static device array of
 * CUPIC_NOALLOC_SCRATCH_BYTES per GPU.  No hipMalloc, hipFree or allocating
 * library call is made.
 *   p       grid, ds, eps0, jacobi_mode, sweep limits and check interval
 *           (derived inside; p is not modified)
 *   charge  int64 node charge [ncy+1][ncx] (periodic in x), 2^32 per unit
 *           macro-particle charge, as CUPIC_MESH_CHARGE exports it; read only
 *   phi     FP32 [ncy+1][ncx]: in, the starting iterate with the Dirichlet
 *           rows 0 (probe) and ncy; out, the solution
 *   rho     FP32 [ncy+1][ncx] charge density out, or NULL
 *   tol     < 0 uses p->jacobi_tol
 *   stream  hipStream_t to run on (NULL: the null stream); the call returns
 *           once the solve has finished on it
 * Grids whose temporaries exceed the scratch (about 20 bytes per node plus
 * the per-sweep reduction slots: 1024 x 512 cells fit) are rejected with
 * CUPIC_ERR_ARG.  Calls share their GPU's scratch, which keeps the last
 * call's contents: calls that overlap on one GPU race. */
int cupic_poisson_solver_noalloc(const cupic_params *p, const int64_t *charge, float *phi,
                                 float *rho, double tol, void *stream, int32_t *sweeps,
                                 double *resid);
/* field_solver (src/mesh.cu:124-154): E = -grad phi */
int cupic_field_solver(cupic_run *r);

/* ---- particles (src/particles.h:28) ----------------------------------- */

/* particle_mover (src/particles.cu:15-102): CIC gather + leapfrog of both
 * species.  The pushed state is committed by the following cupic_cc. */
int cupic_particle_mover(cupic_run *r);

/* ---- boundary conditions (src/cc.h:29-32) ------------------------------ */

/* cc (src/cc.cu:15-69): probe / top absorption with probe-current tally,
 * flux-Maxwellian injection at y = Ly, periodic x, cell-key re-sort. */
int cupic_cc(cupic_run *r, double t);

/* one full cycle, src/main.cu:61-73: deposit, solve, field, move, cc(t) */
int cupic_step(cupic_run *r, double t);

/* ---- diagnostics (src/diagnostic.h:27-31) ------------------------------ */

/* number_of_particles (src/diagnostic.cu:15-44); species 0 = e-, 1 = ions */
int cupic_number_of_particles(cupic_run *r, int species, int64_t *n);
/* probe tally of the last cc and running totals (no reference counterpart:
 * src/cc.cu:184 computes the probe count and discards it) */
int cupic_probe_tally(cupic_run *r, int64_t last_probe[2], int64_t last_top[2],
                      int64_t last_injected[2], int64_t total_probe[2],
                      double *probe_charge_total);
int cupic_last_solve(cupic_run *r, int32_t *sweeps, double *resid);
/* Jacobi sweeps and solves since init_sim */
int cupic_solver_totals(cupic_run *r, int64_t *sweeps_total, int64_t *solves);
/* Statistics since init_sim: per species, particles that moved more than
 * one cell in a step (the far list) and particles the counting pass pushed
 * in full because their push hint was missing or the field could overturn
 * it (counted only while cupic_run_set_statistics is on); the steps taken
 * and the particles moved summed over those steps. */
int cupic_run_counters(cupic_run *r, int64_t total_far[2], int64_t total_undecided[2],
                       int64_t *steps, int64_t *particle_steps);
/* Turns the count of fully pushed particles on or off (off after
 * cupic_run_create).  It costs one same-address atomic per wave of the
 * counting pass, about 0.3 ms per step at 1024 x 512 cells; results never
 * depend on it. */
int cupic_run_set_statistics(cupic_run *r, int32_t on);
int cupic_get_particles(cupic_run *r, int species, float *x, float *y, float *vx, float *vy,
                        int64_t capacity, int64_t *n);
/* uploads a particle state (any order); it is cell-sorted on the device */
int cupic_set_particles(cupic_run *r, int species, const float *x, const float *y,
                        const float *vx, const float *vy, int64_t n);
/* cell offsets of the sorted state, ncx*ncy + 1 entries */
int cupic_get_offsets(cupic_run *r, int species, int32_t *offsets);
/* particles per cell row y in [j ds, (j+1) ds), j = 0..ncy-1: the density
 * profile n_s(y) of one species (the reference keeps only the net rho;
 * src/diagnostic.cu:137-215 writes row bookmarks) */
int cupic_density_profile(cupic_run *r, int species, int64_t *per_row);
/* row bookmarks in the reference layout, 2*ncy entries: first and last index
 * of the particles of row j at [2j] and [2j+1], -1 and -1 for an empty row
 * (the d_e_bm / d_i_bm arrays, src/cc.cu:760-763, src/init.cu:401-407) */
int cupic_get_bookmarks(cupic_run *r, int species, int32_t *bm);

#define CUPIC_MESH_RHO 0     /* float, charge density                      */
#define CUPIC_MESH_PHI 1     /* double, potential                          */
#define CUPIC_MESH_EX 2      /* float                                      */
#define CUPIC_MESH_EY 3      /* float                                      */
#define CUPIC_MESH_CHARGE 4  /* int64, fixed-point node charge, 2^-32 e    */
/* copies a mesh quantity into host memory in the reference layout (nnx*nny) */
int cupic_get_mesh(cupic_run *r, int which, void *host, int64_t count);
/* sets the potential from a host array in the reference layout (doubles) */
int cupic_set_phi(cupic_run *r, const double *phi, int64_t count);
int cupic_get_clock(cupic_run *r, double tin[2], uint64_t injected_total[2], int32_t *started);
int cupic_set_clock(cupic_run *r, const double tin[2], const uint64_t injected_total[2],
                    int32_t started);

/* snapshot writers in the reference's text formats (src/diagnostic.cu:48-133) */
int cupic_particles_snapshot(cupic_run *r, int species, const char *path, double t);
int cupic_mesh_snapshot(cupic_run *r, int which, const char *path);
/* debug writers save_bm and save_bins (src/diagnostic.cu:137-215):
 * " %d %d %d \n" per row (2j, first, last) and " %d %d \n" per particle
 * (index, int(y/ds)); ".dat" is appended to the path */
int cupic_bookmarks_snapshot(cupic_run *r, int species, const char *path);
int cupic_bins_snapshot(cupic_run *r, int species, const char *path);

/* ---- batched runs -------------------------------------------------------- */

/* Groups runs (same Jacobi mode and check interval) so that one launch of
 * each kernel advances all of them; the step is captured once in a HIP graph
 * and replayed.  The runs must have been initialised. */
int cupic_batch_create(cupic_run *const *runs, int32_t nruns, int32_t use_graph,
                       cupic_batch **out);
int cupic_batch_step(cupic_batch *b, int32_t nsteps);
/* cupic_batch_step in two halves: enqueue nsteps steps and return at once;
 * wait for them.  Nothing else may touch the batch's runs in between. */
int cupic_batch_launch(cupic_batch *b, int32_t nsteps);
int cupic_batch_wait(cupic_batch *b);
int cupic_batch_destroy(cupic_batch *b);
const char *cupic_batch_last_error(const cupic_batch *b);

/* ---- serving: continuous batching of I-V sweep points --------------------- */

/* No reference counterpart (the reference runs one configuration per
 * process).  Each request is one bias point with its own parameters; admitted
 * runs stay resident and advance together in batched rounds until their
 * probe current is steady.  Convergence: at every window_steps steps of a
 * run, the change of the window-mean net probe current (ions - electrons)
 * since the previous window, relative to the gross current (ions +
 * electrons), is below rel_change. */
typedef struct cupic_server_config {
    int64_t hbm_budget;        /* bytes of device memory for resident runs     */
    int32_t max_active;        /* resident runs at most                        */
    int32_t round_steps;       /* steps per batched round (divides window)     */
    int32_t window_steps;      /* probe-current window                         */
    int32_t max_steps;         /* cap per point (0: none)                      */
    double rel_change;         /* convergence threshold                        */
    int32_t use_graph;
    int32_t jacobi_mode;       /* applied to every request                     */
    int32_t jacobi_max_sweeps;
    int32_t jacobi_check_every;
    double jacobi_tol;
    int64_t max_particles;     /* resident particles (initial load, both
                                  species) at most, 0: no limit; one run is
                                  always admitted into an empty server     */
    int32_t policy;            /* CUPIC_SERVE_*: which queued point next      */
} cupic_server_config;

/* admission order of queued points */
#define CUPIC_SERVE_FCFS 0            /* arrival order                          */
#define CUPIC_SERVE_SMALLEST_FIRST 1  /* fewest particles first (arrival order
                                         among equals); a large point waits
                                         while smaller ones keep arriving     */

#define CUPIC_POINT_QUEUED 0
#define CUPIC_POINT_RUNNING 1
#define CUPIC_POINT_CONVERGED 2
#define CUPIC_POINT_CAPPED 3     /* reached max_steps first                    */
#define CUPIC_POINT_FAILED 4     /* error holds the status                     */

typedef struct cupic_point {
    int64_t id;
    int32_t status;
    int32_t error;
    double phi_p;
    int64_t footprint;         /* device bytes while resident                 */
    double t_submit, t_admit, t_done;   /* server clock, seconds (-1: not yet) */
    int64_t steps;
    int64_t windows;
    double current[2];         /* last window: absorbed per unit time, e-, ions */
    double rel_change;         /* last window-to-window change                 */
    int64_t total_probe[2];
    int64_t particles[2];
} cupic_point;

typedef struct cupic_server cupic_server;

/* defaults: 200 GiB, 256 runs, 300 M resident particles, first come first
 * served, rounds of 50 steps, 500-step windows, cap 20000 steps, rel_change
 * 0.005, graphs on, CUPIC_POISSON_DIRECT */
int cupic_server_config_default(cupic_server_config *cfg);
int cupic_server_create(const cupic_server_config *cfg, cupic_server **out);
/* queues a point; the Jacobi settings of the config replace those of p */
int cupic_server_submit(cupic_server *s, const cupic_params *p, int64_t *id);
/* one round: admit (slots, HBM budget, particle budget, in the policy's
 * order), advance every resident run by round_steps, retire */
int cupic_server_poll(cupic_server *s, int32_t *n_active, int32_t *n_queued);
int cupic_server_result(cupic_server *s, int64_t id, cupic_point *out);
int cupic_server_stats(cupic_server *s, int64_t *rounds, int64_t *rebuilds, int64_t *hbm_used,
                       double *clock);
int cupic_server_destroy(cupic_server *s);
const char *cupic_server_last_error(const cupic_server *s);

/* ---- autotuning ---------------------------------------------------------- */

/* No reference counterpart.  Times the two push passes (k_push_rank +
 * k_scatter, HIP events) of a run with parameters p for each candidate
 * tile_cells in 1..16 that divides the work sensibly (1, 2, 3, 4, 6, 8, 10,
 * 12, 16), `steps` steps after 2 warm-up steps each, and returns the fastest.
 * times[c] (c = 0..16, may be NULL) receives ms per step of candidate c, or
 * 0 where c was not tried.  The particle states do not depend on the tile,
 * so the choice changes speed only. */
int cupic_autotune_tile(const cupic_params *p, int32_t steps, int32_t *best_tile,
                        double *best_ms, double *times);

/* Changes a run's tile between steps (0 = round(2048 / ppc)).  A batch keeps
 * the descriptors it was created with: create batches after this call. */
int cupic_run_set_tile(cupic_run *r, int32_t tile);

/* Push hints of the counting pass (no counterpart in the reference, which
 * does not sort): the scatter stores with each particle one byte naming the
 * cell it reaches if its velocity does not change and how far that point
 * lies from the cell's walls; the next counting pass trusts the byte when
 * the field cannot move the particle that far, instead of reading and
 * pushing it.  Modes: CUPIC_HINTS_OFF (push every particle),
 * CUPIC_HINTS_ON (default), CUPIC_HINTS_CHECKED (push every particle and
 * flag any trusted hint the push contradicts as a device error).  Results
 * do not depend on the mode.  A batch keeps the descriptors it was created
 * with: create batches after this call. */
#define CUPIC_HINTS_OFF 0
#define CUPIC_HINTS_ON 1
#define CUPIC_HINTS_CHECKED 2
int cupic_run_set_hint_mode(cupic_run *r, int32_t mode);

/* In-situ tuning: steps the run itself `steps` steps per candidate tile
 * (after one step to settle each), from *t (advanced as t += dt per step),
 * and leaves the fastest tile set.  The steps are real steps of the run. */
int cupic_autotune_run(cupic_run *r, double *t, int32_t steps, int32_t *best_tile,
                       double *best_ms, double *times);

/* ---- timing ------------------------------------------------------------- */

/* Per-kernel-class device time (ms) accumulated with HIP events on the run's
 * stream while timing is on; classes are listed by cupic_timing_names. */
int cupic_timing_enable(cupic_run *r, int32_t on);
/* Turns timing on for the classes in class_mask only (bit k: class k of
 * cupic_timing_names), so that a timed run carries the events of the kernels
 * it reports and no others; cupic_timing_enable(r, 1) times every class. */
int cupic_timing_select(cupic_run *r, uint32_t class_mask);
int cupic_timing_read(cupic_run *r, double *ms, int64_t *launches, int32_t n);
const char *cupic_timing_names(void);

/* Plain streaming kernel of known traffic (kind 0: copy, bytes read and
 * bytes written; 1: read; 2: write) over a fresh device buffer of `bytes`,
 * timed over reps launches on the null stream: mean ms per launch.  The
 * achievable HBM bandwidth, and the reference that scales profiler byte
 * counters into bytes.  Allocates and frees its buffers. */
int cupic_bandwidth_probe(int64_t bytes, int32_t kind, int32_t reps, double *ms);

#pragma GCC visibility pop

#ifdef __cplusplus
}
#endif
#endif
