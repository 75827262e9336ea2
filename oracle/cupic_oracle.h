/*
 * cupic_oracle.h -- CPU restatement of cupic's electrostatic PIC cycle.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (cupic_amd/, include/)
 * links, loads or calls this code; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, as the checker.
 *
 * Two layers:
 *   1. FP64 reference semantics (orc_sim_*, orc_deposit_f64, orc_jacobi_*,
 *      orc_field, orc_push_f64, orc_cc_f64): a serial transcription of the
 *      reference kernels' intended behaviour, each function citing the
 *      reference file:line it follows.  Built with and without OpenMP; the
 *      OpenMP build is the CPU baseline.
 *   2. FP32 mirror (orc_*_f32, orc_mirror_*): the same cycle evaluated in IEEE
 *      binary32 with the exact operation order the device kernels use, so
 *      that integer outputs (cell keys, per-cell counts, offsets, the sort
 *      permutation, fixed-point charge) and the particle states themselves
 *      can be checked bit for bit.  The mirror is itself checked against
 *      layer 1 within FP32 tolerance.
 *
 * Grids are stored periodic in x: [nny][ncx] with node (i, j) at j*ncx + i.
 * The reference's column nnx-1 is the periodic image of column 0
 * (src/mesh.cu:236-240, 331-342, 414-423).
 */
#ifndef CUPIC_ORACLE_H
#define CUPIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int32_t ncx, ncy;      /* cells in x and y                                  */
    double ds, dt;         /* cell size [lambda_D], step [1/omega_pe]           */
    double n;              /* normalised density, particles per lambda_D^3      */
    double mi, kti;        /* ion mass [m_e], ion temperature [kTe]             */
    double me, kte;        /* 1, 1 (src/init.cu:487-494, 512-519)               */
    double phi_p;          /* probe bias [kTe/e]                                */
    double eps0;           /* normalised permittivity (== n, src/init.cu:612)   */
    uint64_t seed;         /* Philox key word 0                                 */
    uint64_t run_id;       /* Philox key word 1                                 */
    uint32_t quirks;       /* bit 0: reference injection half-kick sign (Q7)    */
    int32_t pad0;
    /* derived by orc_params_derive */
    double Lx, Ly;
    double dtin_e, dtin_i;
    int64_t per_row;       /* int(n*Lx*ds*ds) particles per row (src/init.cu:136) */
} orc_params;

#define ORC_QUIRK_Q7_REFERENCE_KICK 1u

/* ---- parameters ------------------------------------------------------ */
void orc_params_derive(orc_params *p);

/* ---- counter-based RNG (Philox-4x64-10, Salmon et al., SC'11) -------- */
void orc_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
double orc_u01_open(uint64_t r);
/* one particle's draws: purpose 0 = initial load, 1 = injection */
void orc_draw4(const orc_params *p, int species, int purpose, uint64_t index, uint64_t out[4]);

/* ---- FP64 layer ------------------------------------------------------ */
void orc_create_particles_f64(const orc_params *p, int species,
                              double *x, double *y, double *vx, double *vy);
void orc_deposit_f64(int64_t n, const double *x, const double *y, double ds,
                     int ncx, int ncy, double q, double *Q);
void orc_rho_from_Q(int ncx, int ncy, double ds, const double *Q, double *rho);
void orc_phi_linear(int ncx, int ncy, double phi_p, double *phi);
void orc_jacobi_sweeps(int ncx, int ncy, double ds, double eps0, const double *rho,
                       double *phi, double *work, int nsweeps, double *last_max_dphi);
double orc_residual_max(int ncx, int ncy, double ds, double eps0, const double *rho,
                        const double *phi);
int orc_poisson_solve(int ncx, int ncy, double ds, double eps0, const double *rho,
                      double *phi, int mode, double tol, int min_sweeps, int max_sweeps,
                      int check_every, double *resid_out);
void orc_field(int ncx, int ncy, double ds, const double *phi, double *Ex, double *Ey);
void orc_push_f64(int64_t n, double *x, double *y, double *vx, double *vy,
                  const double *Ex, const double *Ey, int ncx, int ncy, double ds,
                  double q, double m, double dt);
void orc_half_kick_f64(int64_t n, const double *x, const double *y, double *vx, double *vy,
                       const double *Ex, const double *Ey, int ncx, int ncy, double ds,
                       double q, double m, double dt);
/* injection of n_in particles of one species at the top boundary */
void orc_inject_f64(const orc_params *p, int species, uint64_t first_index, int n_in,
                    double t, const double *tin_seq, const double *Ex, const double *Ey,
                    double *x, double *y, double *vx, double *vy);
int orc_injection_count(double t, double dt, double *tin, double dtin, double *tin_seq,
                        int max_seq);

/* full FP64 simulation (the long-run oracle) */
typedef struct orc_sim orc_sim;
orc_sim *orc_sim_create(const orc_params *p, int64_t capacity);
void orc_sim_destroy(orc_sim *s);
void orc_sim_set_solver(orc_sim *s, int mode, double tol, double tol_init, int min_sweeps,
                        int max_sweeps, int check_every);
void orc_sim_init(orc_sim *s);                 /* t = 0 (src/init.cu:70-83)   */
void orc_sim_step(orc_sim *s, double t);       /* src/main.cu:61-73           */
int64_t orc_sim_count(const orc_sim *s, int species);
void orc_sim_get_particles(const orc_sim *s, int species, double *x, double *y,
                           double *vx, double *vy);
void orc_sim_set_particles(orc_sim *s, int species, int64_t n, const double *x,
                           const double *y, const double *vx, const double *vy);
void orc_sim_get_mesh(const orc_sim *s, double *rho, double *phi, double *Ex, double *Ey);
void orc_sim_set_phi(orc_sim *s, const double *phi);
void orc_sim_tallies(const orc_sim *s, int64_t probe[2], int64_t top[2], int64_t injected[2],
                     int *sweeps, double *resid);
void orc_sim_clock(const orc_sim *s, double tin[2], uint64_t ninj_total[2]);
void orc_sim_set_clock(orc_sim *s, const double tin[2], const uint64_t ninj_total[2],
                       int tin_started);

/* ---- FP32 mirror layer ---------------------------------------------- */
typedef struct orc_f32c {          /* FP32 constants, rounded from FP64 once */
    float inv_ds, Lx, Ly, dt;
    float qdtm[2];                 /* q*dt/m per species (0 = electrons)    */
    float hkick[2];                /* -0.5*q*dt/m per species              */
    int32_t ncx, ncy;
} orc_f32c;

void orc_f32c_make(const orc_params *p, orc_f32c *c);
int32_t orc_cell_of_f32(const orc_f32c *c, float x, float y, float *fx, float *fy,
                        int32_t *ic, int32_t *jc);
void orc_create_particles_f32(const orc_params *p, int species, float *x, float *y,
                              float *vx, float *vy);
void orc_deposit_fixed_f32(int64_t n, const float *x, const float *y, const orc_f32c *c,
                           int sign, int64_t *acc);
void orc_gather_f32(const orc_f32c *c, const float *Exy, float x, float y, float *ex,
                    float *ey);
/* stable counting sort by key*10+slot; returns count kept, writes perm */
int64_t orc_stable_sort_slots(int64_t n, const int32_t *key, const int8_t *slot,
                              int32_t nkeys, int64_t *perm, int64_t *offsets);
/* init mirror: generate, round to FP32 and sort by cell (stable) */
void orc_mirror_init(const orc_params *p, int species, float *x, float *y, float *vx,
                     float *vy, int64_t *offsets);
void orc_mirror_half_kick(const orc_f32c *c, int species, int64_t n, const float *x,
                          const float *y, float *vx, float *vy, const float *Exy);
/* one particle step of one species: push, wrap, absorb, classify, append
 * injected particles (already FP32) and stable-sort by (cell, source slot).
 * out arrays need n + n_in capacity.  Returns kept count. */
int64_t orc_mirror_step(const orc_f32c *c, int species, int64_t n, const float *x,
                        const float *y, const float *vx, const float *vy, const float *Exy,
                        int64_t n_in, const float *ix, const float *iy, const float *ivx,
                        const float *ivy, float *ox, float *oy, float *ovx, float *ovy,
                        int64_t *offsets, int64_t *probe, int64_t *top, int64_t *nfar);
void orc_inject_f32(const orc_params *p, int species, uint64_t first_index, int n_in,
                    double t, const double *tin_seq, const float *Exy, float *x, float *y,
                    float *vx, float *vy);
/* Right-hand side of the device solve from the int64 node charge:
 * b64 = acc * 2^-32 / (ds eps0) on rows 1..ncy-1, 0 on the Dirichlet rows;
 * b32 = (float)b64. */
void orc_mirror_rhs(const orc_params *p, const int64_t *acc, double *b64, float *b32);
/* The device's FP32 Jacobi with FP64 residual (mode 1): FP32 sweeps
 * 0.25f * ((((b + west) + east) + north) + south); FP64 residual
 * max |0.25 ((((b64 + w) + e) + n) + s - 4 phi)| checked before the first
 * sweep and after every check_every sweeps; stops at the first check with
 * residual <= tol, else after max_sweeps.  phi [ncy+1][ncx] in/out;
 * returns the sweeps run and the last residual. */
int orc_mirror_solve_f32(int ncx, int ncy, const double *b64, const float *b32, float *phi,
                         double tol, int max_sweeps, int check_every, double *resid_out);

#ifdef __cplusplus
}
#endif
#endif
