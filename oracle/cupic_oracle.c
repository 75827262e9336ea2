/*
 * cupic_oracle.c -- CPU restatement of cupic's PIC cycle (test infrastructure).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this file, as the checker.  See cupic_oracle.h for the two layers.
 *
 * Build (oracle/Makefile): gcc -O2 -ffp-contract=off, once serial and once
 * with -fopenmp (the CPU baseline).  -ffp-contract=off matters for the FP32
 * mirror: every float expression below is evaluated as written, one IEEE
 * operation at a time, which is what the device kernels are compiled to.
 *
 * Reference citations are file:line in AntonioTejero/cupic.
 */
#include "cupic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.1415926535897932 /* src/stdh.h:26 */

/* ===================================================================== */
/* parameters                                                            */
/* ===================================================================== */

/* Normalised units as derived in src/init.cu:451-789: lengths in lambda_D,
 * time in 1/omega_pe, mass in m_e, potential in kTe/e.  eps0 == n
 * (src/init.cu:612-630 reduces to the normalised density). */
void orc_params_derive(orc_params *p)
{
    if (p->me == 0.0) p->me = 1.0;     /* src/init.cu:487-494 */
    if (p->kte == 0.0) p->kte = 1.0;   /* src/init.cu:512-519 */
    if (p->eps0 == 0.0) p->eps0 = p->n;
    p->Lx = (double)(p->ds * (double)p->ncx); /* src/init.cu:560-568 */
    p->Ly = (double)(p->ds * (double)p->ncy); /* src/init.cu:572-580 */
    p->per_row = (int64_t)(p->n * p->Lx * p->ds * p->ds); /* src/init.cu:136 */
    /* src/init.cu:686-714 */
    p->dtin_i = sqrt(2.0 * ORC_PI * p->mi / p->kti) / (p->n * p->Lx * p->ds);
    p->dtin_e = sqrt(2.0 * ORC_PI) / (p->n * p->Lx * p->ds);
}

/* ===================================================================== */
/* Philox-4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:  */
/* as easy as 1, 2, 3", SC'11).  Replaces the reference's GSL streams     */
/* (src/init.cu:124,168-177; src/cc.cu:164,216-219).                      */
/* ===================================================================== */

static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t *hi, uint64_t *lo)
{
    unsigned __int128 p = (unsigned __int128)a * (unsigned __int128)b;
    *hi = (uint64_t)(p >> 64);
    *lo = (uint64_t)p;
}

void orc_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4])
{
    uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint64_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ULL, c0, &hi0, &lo0);
        mulhilo64(0xCA5A826395121157ULL, c2, &hi1, &lo1);
        uint64_t n0 = hi1 ^ c1 ^ k0;
        uint64_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B97F4A7C15ULL;
        k1 += 0xBB67AE8584CAA73BULL;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* uniform on the open interval (0,1): 52 random bits plus a half, exact */
double orc_u01_open(uint64_t r)
{
    return ((double)(r >> 12) + 0.5) * 0x1p-52;
}

void orc_draw4(const orc_params *p, int species, int purpose, uint64_t index, uint64_t out[4])
{
    uint64_t ctr[4] = {index, (uint64_t)species | ((uint64_t)purpose << 8), 0, 0};
    uint64_t key[2] = {p->seed, p->run_id};
    orc_philox4x64_10(ctr, key, out);
}

static double species_mass(const orc_params *p, int s) { return s == 0 ? p->me : p->mi; }
static double species_kt(const orc_params *p, int s) { return s == 0 ? p->kte : p->kti; }
static double species_q(int s) { return s == 0 ? -1.0 : 1.0; } /* src/init.cu:451-472 */

/* ===================================================================== */
/* FP64 layer                                                            */
/* ===================================================================== */

/* Initial load, src/init.cu:136, 159-179: per_row particles per y row,
 * x ~ U(0,1)*Lx, y = row*ds + U(0,1)*ds, v ~ N(0, sqrt(kT/m)).  Draws come
 * from Philox with counter (global index, species, purpose 0). */
static void draw_initial(const orc_params *p, int s, int64_t g, int64_t row, double *x,
                         double *y, double *vx, double *vy)
{
    uint64_t r[4];
    orc_draw4(p, s, 0, (uint64_t)g, r);
    double sigma = sqrt(species_kt(p, s) / species_mass(p, s));
    double u0 = orc_u01_open(r[0]), u1 = orc_u01_open(r[1]);
    double u2 = orc_u01_open(r[2]), u3 = orc_u01_open(r[3]);
    double rad = sqrt(-2.0 * log(u2));
    double th = 2.0 * ORC_PI * u3;
    *x = u0 * p->Lx;
    *y = (double)row * p->ds + u1 * p->ds;
    *vx = sigma * (rad * cos(th));
    *vy = sigma * (rad * sin(th));
}

void orc_create_particles_f64(const orc_params *p, int species, double *x, double *y,
                              double *vx, double *vy)
{
    int64_t N = p->per_row * p->ncy;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < N; g++) {
        int64_t row = g / p->per_row;
        draw_initial(p, species, g, row, &x[g], &y[g], &vx[g], &vy[g]);
    }
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* CIC weights as in src/mesh.cu:202-210 with the cell row taken from the
 * particle's own y (Q16). */
static inline void cic64(double x, double y, double ds, int ncx, int ncy, int *ic, int *jc,
                         double *dx, double *dy)
{
    int i = (int)floor(x / ds);
    int j = (int)floor(y / ds);
    i = clampi(i, 0, ncx - 1);
    j = clampi(j, 0, ncy - 1);
    *ic = i;
    *jc = j;
    *dx = fabs((double)i * ds - x) / ds;
    *dy = fabs((double)j * ds - y) / ds;
}

/* Charge deposition, src/mesh.cu:162-240: Q at the four vertices of the
 * particle's cell, periodic fold of column ncx onto column 0 (236-240). */
void orc_deposit_f64(int64_t n, const double *x, const double *y, double ds, int ncx,
                     int ncy, double q, double *Q)
{
    int64_t nodes = (int64_t)ncx * (ncy + 1);
#ifdef _OPENMP
#pragma omp parallel
    {
        double *loc = (double *)calloc((size_t)nodes, sizeof(double));
#pragma omp for schedule(static)
        for (int64_t k = 0; k < n; k++) {
            int ic, jc;
            double dx, dy;
            cic64(x[k], y[k], ds, ncx, ncy, &ic, &jc, &dx, &dy);
            int i1 = ic + 1 == ncx ? 0 : ic + 1;
            loc[(int64_t)jc * ncx + ic] += q * (1.0 - dx) * (1.0 - dy);
            loc[(int64_t)jc * ncx + i1] += q * dx * (1.0 - dy);
            loc[(int64_t)(jc + 1) * ncx + ic] += q * (1.0 - dx) * dy;
            loc[(int64_t)(jc + 1) * ncx + i1] += q * dx * dy;
        }
#pragma omp critical
        for (int64_t k = 0; k < nodes; k++) Q[k] += loc[k];
        free(loc);
    }
#else
    (void)nodes;
    for (int64_t k = 0; k < n; k++) {
        int ic, jc;
        double dx, dy;
        cic64(x[k], y[k], ds, ncx, ncy, &ic, &jc, &dx, &dy);
        int i1 = ic + 1 == ncx ? 0 : ic + 1;
        Q[(int64_t)jc * ncx + ic] += q * (1.0 - dx) * (1.0 - dy);
        Q[(int64_t)jc * ncx + i1] += q * dx * (1.0 - dy);
        Q[(int64_t)(jc + 1) * ncx + ic] += q * (1.0 - dx) * dy;
        Q[(int64_t)(jc + 1) * ncx + i1] += q * dx * dy;
    }
#endif
}

/* Volume normalisation: interior rows ds^3, the two boundary rows ds^3/2.
 * This is the intent of the commented block src/mesh.cu:242-268; the
 * committed src/mesh.cu:279-290 races with neighbouring blocks (Q2). */
void orc_rho_from_Q(int ncx, int ncy, double ds, const double *Q, double *rho)
{
    double v = ds * ds * ds;
    for (int j = 0; j <= ncy; j++) {
        double vol = (j == 0 || j == ncy) ? 0.5 * v : v;
        for (int i = 0; i < ncx; i++) rho[(int64_t)j * ncx + i] = Q[(int64_t)j * ncx + i] / vol;
    }
}

/* Initial potential, src/init.cu:235-241: phi(j) = (1 - j/ncy) * phi_p. */
void orc_phi_linear(int ncx, int ncy, double phi_p, double *phi)
{
    for (int j = 0; j <= ncy; j++)
        for (int i = 0; i < ncx; i++)
            phi[(int64_t)j * ncx + i] = (1.0 - (double)j / (double)ncy) * phi_p;
}

/* One Jacobi sweep over rows 1..ncy-1, src/mesh.cu:327-350, periodic in x
 * (331-342).  True Jacobi: reads phi, writes work (Q3 decision). */
static double jacobi_sweep(int ncx, int ncy, double ds, double eps0, const double *rho,
                           const double *phi, double *out)
{
    double mx = 0.0;
    double c = ds * ds;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int j = 1; j < ncy; j++) {
        const double *pc = phi + (int64_t)j * ncx;
        const double *pn = pc + ncx;
        const double *ps = pc - ncx;
        for (int i = 0; i < ncx; i++) {
            int iw = i == 0 ? ncx - 1 : i - 1;
            int ie = i == ncx - 1 ? 0 : i + 1;
            double rd = c * rho[(int64_t)j * ncx + i] / eps0;
            double v = 0.25 * (rd + pc[iw] + pc[ie] + pn[i] + ps[i]);
            double d = fabs(v - pc[i]);
            if (d > mx) mx = d;
            out[(int64_t)j * ncx + i] = v;
        }
    }
    return mx;
}

void orc_jacobi_sweeps(int ncx, int ncy, double ds, double eps0, const double *rho,
                       double *phi, double *work, int nsweeps, double *last_max_dphi)
{
    int64_t nodes = (int64_t)ncx * (ncy + 1);
    memcpy(work, phi, (size_t)nodes * sizeof(double));
    double mx = 0.0;
    for (int s = 0; s < nsweeps; s++) {
        mx = jacobi_sweep(ncx, ncy, ds, eps0, rho, phi, work);
        memcpy(phi + ncx, work + ncx, (size_t)ncx * (ncy - 1) * sizeof(double));
    }
    if (last_max_dphi) *last_max_dphi = mx;
}

/* max over interior nodes of |Jacobi update| = |residual|/4, in kTe/e --
 * the quantity the reference's stopping rule measures (src/mesh.cu:354). */
double orc_residual_max(int ncx, int ncy, double ds, double eps0, const double *rho,
                        const double *phi)
{
    double mx = 0.0;
    double c = ds * ds;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int j = 1; j < ncy; j++) {
        const double *pc = phi + (int64_t)j * ncx;
        for (int i = 0; i < ncx; i++) {
            int iw = i == 0 ? ncx - 1 : i - 1;
            int ie = i == ncx - 1 ? 0 : i + 1;
            double rd = c * rho[(int64_t)j * ncx + i] / eps0;
            double v = 0.25 * (rd + pc[iw] + pc[ie] + pc[i + ncx] + pc[i - ncx]);
            double d = fabs(v - pc[i]);
            if (d > mx) mx = d;
        }
    }
    return mx;
}

/* ---- mode 3: the fixed point of the sweeps, solved directly in FP64 ----
 * The sweeps above converge to the solution of 4 phi - (W + E + N + S) =
 * ds^2 rho / eps0 on rows 1..ncy-1, periodic in x, rows 0 and ncy fixed
 * (src/mesh.cu:297-381 iterates towards it).  Along x a DFT diagonalises
 * the periodic second difference (an FFT for power-of-two ncx, a direct
 * O(ncx^2) DFT otherwise); per wavenumber k the rows then solve the
 * tridiagonal system -x(j-1) + (2 + 4 sin^2(pi k / ncx)) x(j) - x(j+1) =
 * r(j), zero beyond the interior, by the Thomas algorithm (real matrix,
 * complex right-hand side); an inverse DFT gives the rows back. */

/* e^(sign 2 pi i t / n), t = 0 .. n-1 */
static void dft_table(int n, int sign, double *twr, double *twi)
{
    for (int t = 0; t < n; t++) {
        double a = 2.0 * ORC_PI * (double)t / (double)n;
        twr[t] = cos(a);
        twi[t] = sign * sin(a);
    }
}

/* in place, unnormalised: X_k = sum_t x_t tw[(t k) mod n] */
static void dft_row(int n, double *re, double *im, const double *twr, const double *twi,
                    double *sr, double *si)
{
    if ((n & (n - 1)) == 0) {
        /* iterative radix-2: bit reversal, then butterflies of growing span */
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double t = re[i]; re[i] = re[j]; re[j] = t;
                t = im[i]; im[i] = im[j]; im[j] = t;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int h = len >> 1, step = n / len;
            for (int j = 0; j < h; j++) {
                double wr = twr[j * step], wi = twi[j * step];
                for (int i = j; i < n; i += len) {
                    int k = i + h;
                    double tr = re[k] * wr - im[k] * wi, ti = re[k] * wi + im[k] * wr;
                    re[k] = re[i] - tr;
                    im[k] = im[i] - ti;
                    re[i] += tr;
                    im[i] += ti;
                }
            }
        }
        return;
    }
    for (int k = 0; k < n; k++) {
        double ar = 0.0, ai = 0.0;
        for (int t = 0; t < n; t++) {
            int q = (int)(((int64_t)t * k) % n);
            ar += re[t] * twr[q] - im[t] * twi[q];
            ai += re[t] * twi[q] + im[t] * twr[q];
        }
        sr[k] = ar;
        si[k] = ai;
    }
    memcpy(re, sr, (size_t)n * sizeof(double));
    memcpy(im, si, (size_t)n * sizeof(double));
}

static void direct_solve(int ncx, int ncy, double ds, double eps0, const double *rho,
                         double *phi)
{
    const int m = ncy - 1, n = ncx;
    const double c = ds * ds;
    double *xr = (double *)malloc((size_t)m * n * sizeof(double));
    double *xi = (double *)malloc((size_t)m * n * sizeof(double));
    double *fr = (double *)malloc((size_t)n * sizeof(double));
    double *fi = (double *)malloc((size_t)n * sizeof(double));
    double *br = (double *)malloc((size_t)n * sizeof(double));
    double *bi = (double *)malloc((size_t)n * sizeof(double));
    dft_table(n, -1, fr, fi);
    dft_table(n, +1, br, bi);
    /* rows 1..m of the right-hand side, the Dirichlet rows moved into the
     * first and last of them, forward DFT */
#pragma omp parallel
    {
        double *sr = (double *)malloc((size_t)n * sizeof(double));
        double *si = (double *)malloc((size_t)n * sizeof(double));
#pragma omp for schedule(static)
        for (int r = 0; r < m; r++) {
            double *re = xr + (int64_t)r * n, *im = xi + (int64_t)r * n;
            for (int i = 0; i < n; i++) {
                double v = c * rho[(int64_t)(r + 1) * n + i] / eps0;
                if (r == 0) v += phi[i];
                if (r == m - 1) v += phi[(int64_t)ncy * n + i];
                re[i] = v;
                im[i] = 0.0;
            }
            dft_row(n, re, im, fr, fi, sr, si);
        }
        free(sr);
        free(si);
    }
    /* per wavenumber, the tridiagonal system along y (Thomas) */
#pragma omp parallel
    {
        double *cp = (double *)malloc((size_t)m * sizeof(double));
#pragma omp for schedule(static)
        for (int k = 0; k < n; k++) {
            double sk = sin(ORC_PI * (double)k / (double)n);
            double d = 2.0 + 4.0 * sk * sk;
            double den = d;
            cp[0] = -1.0 / den;
            xr[k] /= den;
            xi[k] /= den;
            for (int j = 1; j < m; j++) {
                den = d + cp[j - 1];
                cp[j] = -1.0 / den;
                int64_t e = (int64_t)j * n + k, f = e - n;
                xr[e] = (xr[e] + xr[f]) / den;
                xi[e] = (xi[e] + xi[f]) / den;
            }
            for (int j = m - 2; j >= 0; j--) {
                int64_t e = (int64_t)j * n + k, f = e + n;
                xr[e] -= cp[j] * xr[f];
                xi[e] -= cp[j] * xi[f];
            }
        }
        free(cp);
    }
    /* inverse DFT into rows 1..m */
#pragma omp parallel
    {
        double *sr = (double *)malloc((size_t)n * sizeof(double));
        double *si = (double *)malloc((size_t)n * sizeof(double));
#pragma omp for schedule(static)
        for (int r = 0; r < m; r++) {
            double *re = xr + (int64_t)r * n, *im = xi + (int64_t)r * n;
            dft_row(n, re, im, br, bi, sr, si);
            for (int i = 0; i < n; i++) phi[(int64_t)(r + 1) * n + i] = re[i] / (double)n;
        }
        free(sr);
        free(si);
    }
    free(xr); free(xi); free(fr); free(fi); free(br); free(bi);
}

/* mode 0: the reference rule, src/mesh.cu:66-67, 94-114 -- sweep until
 * max|dphi| < tol and at least min_sweeps sweeps ran (the reference uses
 * max(nnx,nny)+1); capped at max_sweeps (the reference has no cap).
 * modes 1, 2: residual checked before every block of check_every sweeps.
 * mode 3: the direct solve above (no sweeps); the residual after it. */
int orc_poisson_solve(int ncx, int ncy, double ds, double eps0, const double *rho,
                      double *phi, int mode, double tol, int min_sweeps, int max_sweeps,
                      int check_every, double *resid_out)
{
    if (mode == 3) {
        direct_solve(ncx, ncy, ds, eps0, rho, phi);
        double res = orc_residual_max(ncx, ncy, ds, eps0, rho, phi);
        if (resid_out) *resid_out = res;
        return 0;
    }
    int64_t nodes = (int64_t)ncx * (ncy + 1);
    double *work = (double *)malloc((size_t)nodes * sizeof(double));
    memcpy(work, phi, (size_t)nodes * sizeof(double));
    int s = 0;
    double res = 0.0;
    if (mode == 0) {
        double err = tol * 10.0;
        while ((s < min_sweeps || err >= tol) && (max_sweeps <= 0 || s < max_sweeps)) {
            err = jacobi_sweep(ncx, ncy, ds, eps0, rho, phi, work);
            memcpy(phi + ncx, work + ncx, (size_t)ncx * (ncy - 1) * sizeof(double));
            s++;
        }
        res = err;
    } else {
        for (;;) {
            res = orc_residual_max(ncx, ncy, ds, eps0, rho, phi);
            if (res <= tol || s >= max_sweeps) break;
            int k = check_every;
            if (s + k > max_sweeps) k = max_sweeps - s;
            for (int q = 0; q < k; q++) {
                jacobi_sweep(ncx, ncy, ds, eps0, rho, phi, work);
                memcpy(phi + ncx, work + ncx, (size_t)ncx * (ncy - 1) * sizeof(double));
            }
            s += k;
        }
    }
    free(work);
    if (resid_out) *resid_out = res;
    return s;
}

/* E = -grad phi, src/mesh.cu:413-468: central differences on interior rows
 * with periodic x; one-sided Ey and Ex = 0 on rows 0 and ncy. */
void orc_field(int ncx, int ncy, double ds, const double *phi, double *Ex, double *Ey)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= ncy; j++) {
        for (int i = 0; i < ncx; i++) {
            int64_t k = (int64_t)j * ncx + i;
            if (j == 0) {
                Ex[k] = 0.0;
                Ey[k] = (phi[k] - phi[k + ncx]) / ds;
            } else if (j == ncy) {
                Ex[k] = 0.0;
                Ey[k] = (phi[k - ncx] - phi[k]) / ds;
            } else {
                int iw = i == 0 ? ncx - 1 : i - 1;
                int ie = i == ncx - 1 ? 0 : i + 1;
                Ex[k] = (phi[(int64_t)j * ncx + iw] - phi[(int64_t)j * ncx + ie]) / (2.0 * ds);
                Ey[k] = (phi[k - ncx] - phi[k + ncx]) / (2.0 * ds);
            }
        }
    }
}

/* CIC gather of q*E, src/particles.cu:145-177, with the cell row from the
 * particle's y (Q1: the committed code uses blockDim.x there). */
static inline void gather64(const double *Ex, const double *Ey, int ncx, int ncy, double ds,
                            double x, double y, double *ex, double *ey)
{
    int ic, jc;
    double dx, dy;
    cic64(x, y, ds, ncx, ncy, &ic, &jc, &dx, &dy);
    int i1 = ic + 1 == ncx ? 0 : ic + 1;
    int64_t k00 = (int64_t)jc * ncx + ic, k10 = (int64_t)jc * ncx + i1;
    int64_t k01 = k00 + ncx, k11 = k10 + ncx;
    *ex = Ex[k00] * (1.0 - dx) * (1.0 - dy) + Ex[k10] * dx * (1.0 - dy) +
          Ex[k01] * (1.0 - dx) * dy + Ex[k11] * dx * dy;
    *ey = Ey[k00] * (1.0 - dx) * (1.0 - dy) + Ey[k10] * dx * (1.0 - dy) +
          Ey[k01] * (1.0 - dx) * dy + Ey[k11] * dx * dy;
}

/* Leapfrog, src/particles.cu:220-223: v += dt*F/m; x += dt*v. */
void orc_push_f64(int64_t n, double *x, double *y, double *vx, double *vy, const double *Ex,
                  const double *Ey, int ncx, int ncy, double ds, double q, double m, double dt)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; k++) {
        double ex, ey;
        gather64(Ex, Ey, ncx, ncy, ds, x[k], y[k], &ex, &ey);
        double fx = q * ex, fy = q * ey;
        vx[k] += dt * fx / m;
        vy[k] += dt * fy / m;
        x[k] += dt * vx[k];
        y[k] += dt * vy[k];
    }
}

/* Initial velocity stagger, src/init.cu:825-826: v -= 0.5*dt*F/m. */
void orc_half_kick_f64(int64_t n, const double *x, const double *y, double *vx, double *vy,
                       const double *Ex, const double *Ey, int ncx, int ncy, double ds,
                       double q, double m, double dt)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; k++) {
        double ex, ey;
        gather64(Ex, Ey, ncx, ncy, ds, x[k], y[k], &ex, &ey);
        double fx = q * ex, fy = q * ey;
        vx[k] -= 0.5 * dt * fx / m;
        vy[k] -= 0.5 * dt * fy / m;
    }
}

/* Number of particles entering this step and their insertion times,
 * src/cc.cu:147, 179, 247: in = 1 + int((fpt - tin)/dtin) if tin < fpt,
 * with tin += dtin once per inserted particle. */
int orc_injection_count(double t, double dt, double *tin, double dtin, double *tin_seq,
                        int max_seq)
{
    double fpt = t + dt;
    int in = 0;
    if (*tin < fpt) in = 1 + (int)((fpt - *tin) / dtin);
    for (int k = 0; k < in; k++) {
        if (k < max_seq && tin_seq) tin_seq[k] = *tin;
        *tin += dtin;
    }
    return in;
}

/* One injected particle, src/cc.cu:214-247 (FP64).  E comes from the top
 * node row (disty = 1, src/cc.cu:222-238).  Q7: the reference subtracts
 * (fvt-tin)*E/m for both species; the default applies q*E/m. */
static void inject_one64(const orc_params *p, int s, uint64_t index, double t, double tin,
                         double ex_top_l, double ex_top_r, double ey_top_l, double ey_top_r,
                         int have_field, const double *Ex, const double *Ey, double *ox,
                         double *oy, double *ovx, double *ovy)
{
    uint64_t r[4];
    orc_draw4(p, s, 1, index, r);
    double m = species_mass(p, s);
    double sigma = sqrt(species_kt(p, s) / m);
    double fpt = t + p->dt, fvt = t + 0.5 * p->dt;
    double x = orc_u01_open(r[0]) * p->Lx;
    double y = p->Ly;
    double g = sqrt(-2.0 * log(orc_u01_open(r[1]))) * cos(2.0 * ORC_PI * orc_u01_open(r[2]));
    double vx = sigma * g;
    double vy = -(sigma * sqrt(-2.0 * log(orc_u01_open(r[3]))));
    int ic = (int)(x / p->ds);
    if (ic > p->ncx - 1) ic = p->ncx - 1;
    int jc = p->ncy - 1;
    int i1 = ic + 1 == p->ncx ? 0 : ic + 1;
    double distx = fabs((double)ic * p->ds - x) / p->ds;
    double disty = 1.0;
    double e00x, e10x, e01x, e11x, e00y, e10y, e01y, e11y;
    if (have_field) {
        int64_t k00 = (int64_t)jc * p->ncx + ic, k10 = (int64_t)jc * p->ncx + i1;
        e00x = Ex[k00]; e10x = Ex[k10]; e01x = Ex[k00 + p->ncx]; e11x = Ex[k10 + p->ncx];
        e00y = Ey[k00]; e10y = Ey[k10]; e01y = Ey[k00 + p->ncx]; e11y = Ey[k10 + p->ncx];
    } else {
        (void)ex_top_l; (void)ex_top_r; (void)ey_top_l; (void)ey_top_r;
        e00x = e10x = e01x = e11x = e00y = e10y = e01y = e11y = 0.0;
    }
    double epx = e00x * (1.0 - distx) * (1.0 - disty);
    epx += e10x * distx * (1.0 - disty);
    epx += e01x * (1.0 - distx) * disty;
    epx += e11x * distx * disty;
    double epy = e00y * (1.0 - distx) * (1.0 - disty);
    epy += e10y * distx * (1.0 - disty);
    epy += e01y * (1.0 - distx) * disty;
    epy += e11y * distx * disty;
    x += (fpt - tin) * vx;
    y += (fpt - tin) * vy;
    if ((p->quirks & ORC_QUIRK_Q7_REFERENCE_KICK) || species_q(s) < 0.0) {
        vx -= (fvt - tin) * epx / m;
        vy -= (fvt - tin) * epy / m;
    } else {
        vx += (fvt - tin) * epx / m;
        vy += (fvt - tin) * epy / m;
    }
    /* cyclic condition applied after injection (src/cc.cu:55, 63; Q6) */
    if (x < 0.0) x += p->Lx;
    if (x >= p->Lx) x -= p->Lx;
    *ox = x; *oy = y; *ovx = vx; *ovy = vy;
}

void orc_inject_f64(const orc_params *p, int species, uint64_t first_index, int n_in,
                    double t, const double *tin_seq, const double *Ex, const double *Ey,
                    double *x, double *y, double *vx, double *vy)
{
    for (int k = 0; k < n_in; k++)
        inject_one64(p, species, first_index + (uint64_t)k, t, tin_seq[k], 0, 0, 0, 0, 1, Ex,
                     Ey, &x[k], &y[k], &vx[k], &vy[k]);
}

/* ---- full FP64 simulation ------------------------------------------- */

struct orc_sim {
    orc_params p;
    int64_t cap;
    int64_t n[2];
    double *x[2], *y[2], *vx[2], *vy[2];
    double *tx, *ty, *tvx, *tvy; /* compaction scratch */
    double *Q, *rho, *phi, *Ex, *Ey;
    double tin[2];
    uint64_t ninj[2];
    int tin_started;
    int64_t probe[2], top[2], injected[2];
    int jmode, jmin, jmax, jcheck;
    double jtol, jtol_init;
    int sweeps;
    double resid;
    double *tin_seq;
    int tin_seq_cap;
};

orc_sim *orc_sim_create(const orc_params *pin, int64_t capacity)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    s->p = *pin;
    orc_params_derive(&s->p);
    int64_t N = s->p.per_row * s->p.ncy;
    if (capacity < N) capacity = N + N / 8 + 4096;
    s->cap = capacity;
    for (int sp = 0; sp < 2; sp++) {
        s->x[sp] = (double *)malloc((size_t)capacity * sizeof(double));
        s->y[sp] = (double *)malloc((size_t)capacity * sizeof(double));
        s->vx[sp] = (double *)malloc((size_t)capacity * sizeof(double));
        s->vy[sp] = (double *)malloc((size_t)capacity * sizeof(double));
    }
    s->tx = (double *)malloc((size_t)capacity * sizeof(double));
    s->ty = (double *)malloc((size_t)capacity * sizeof(double));
    s->tvx = (double *)malloc((size_t)capacity * sizeof(double));
    s->tvy = (double *)malloc((size_t)capacity * sizeof(double));
    int64_t nodes = (int64_t)s->p.ncx * (s->p.ncy + 1);
    s->Q = (double *)calloc((size_t)nodes, sizeof(double));
    s->rho = (double *)calloc((size_t)nodes, sizeof(double));
    s->phi = (double *)calloc((size_t)nodes, sizeof(double));
    s->Ex = (double *)calloc((size_t)nodes, sizeof(double));
    s->Ey = (double *)calloc((size_t)nodes, sizeof(double));
    s->jmode = 0;
    s->jtol = 1e-4;      /* src/main.cu:64 */
    s->jtol_init = 1e-3; /* src/init.cu:254 */
    s->jmin = (s->p.ncx + 1 > s->p.ncy + 1 ? s->p.ncx + 1 : s->p.ncy + 1) + 1;
    s->jmax = 0;
    s->jcheck = 50;
    s->tin_seq_cap = 1 << 20;
    s->tin_seq = (double *)malloc((size_t)s->tin_seq_cap * sizeof(double));
    return s;
}

void orc_sim_destroy(orc_sim *s)
{
    if (!s) return;
    for (int sp = 0; sp < 2; sp++) {
        free(s->x[sp]); free(s->y[sp]); free(s->vx[sp]); free(s->vy[sp]);
    }
    free(s->tx); free(s->ty); free(s->tvx); free(s->tvy);
    free(s->Q); free(s->rho); free(s->phi); free(s->Ex); free(s->Ey);
    free(s->tin_seq);
    free(s);
}

void orc_sim_set_solver(orc_sim *s, int mode, double tol, double tol_init, int min_sweeps,
                        int max_sweeps, int check_every)
{
    s->jmode = mode;
    s->jtol = tol;
    s->jtol_init = tol_init;
    s->jmin = min_sweeps;
    s->jmax = max_sweeps;
    s->jcheck = check_every > 0 ? check_every : 50;
}

static void sim_deposit_solve_field(orc_sim *s, double tol)
{
    const orc_params *p = &s->p;
    int64_t nodes = (int64_t)p->ncx * (p->ncy + 1);
    memset(s->Q, 0, (size_t)nodes * sizeof(double));
    orc_deposit_f64(s->n[0], s->x[0], s->y[0], p->ds, p->ncx, p->ncy, -1.0, s->Q);
    orc_deposit_f64(s->n[1], s->x[1], s->y[1], p->ds, p->ncx, p->ncy, +1.0, s->Q);
    orc_rho_from_Q(p->ncx, p->ncy, p->ds, s->Q, s->rho);
    s->sweeps = orc_poisson_solve(p->ncx, p->ncy, p->ds, p->eps0, s->rho, s->phi, s->jmode,
                                  tol, s->jmin, s->jmax, s->jcheck, &s->resid);
    orc_field(p->ncx, p->ncy, p->ds, s->phi, s->Ex, s->Ey);
}

/* init_sim with n_ini == 0, src/init.cu:70-83: create particles, linear phi,
 * deposit -> solve(1e-3) -> field, then the half-step velocity stagger. */
void orc_sim_init(orc_sim *s)
{
    const orc_params *p = &s->p;
    int64_t N = p->per_row * p->ncy;
    for (int sp = 0; sp < 2; sp++) {
        orc_create_particles_f64(p, sp, s->x[sp], s->y[sp], s->vx[sp], s->vy[sp]);
        s->n[sp] = N;
        s->ninj[sp] = 0;
        s->probe[sp] = s->top[sp] = s->injected[sp] = 0;
    }
    s->tin_started = 0;
    orc_phi_linear(p->ncx, p->ncy, p->phi_p, s->phi);
    sim_deposit_solve_field(s, s->jtol_init);
    for (int sp = 0; sp < 2; sp++)
        orc_half_kick_f64(s->n[sp], s->x[sp], s->y[sp], s->vx[sp], s->vy[sp], s->Ex, s->Ey,
                          p->ncx, p->ncy, p->ds, species_q(sp), species_mass(p, sp), p->dt);
}

/* cc for one species, src/cc.cu:51-63: remove particles whose y bin left
 * the domain (out_l at the probe, out_r at the top, src/cc.cu:182-185),
 * inject (src/cc.cu:179, 214-256) and wrap x (src/cc.cu:276-296). */
static void sim_cc_species(orc_sim *s, int sp, double t)
{
    const orc_params *p = &s->p;
    int64_t n = s->n[sp];
    int64_t probe = 0, top = 0, kept = 0;
    double *x = s->x[sp], *y = s->y[sp], *vx = s->vx[sp], *vy = s->vy[sp];
#ifdef _OPENMP
    int nt = omp_get_max_threads();
    int64_t *cnt = (int64_t *)calloc((size_t)nt * 3 + 3, sizeof(int64_t));
#pragma omp parallel num_threads(nt)
    {
        int tid = omp_get_thread_num();
        int64_t lo = n * tid / nt, hi = n * (tid + 1) / nt;
        int64_t c = 0, pr = 0, tp = 0;
        for (int64_t k = lo; k < hi; k++) {
            double b = floor(y[k] / p->ds);
            if (b < 0.0) pr++;
            else if (b >= (double)p->ncy) tp++;
            else c++;
        }
        cnt[3 * tid] = c; cnt[3 * tid + 1] = pr; cnt[3 * tid + 2] = tp;
#pragma omp barrier
        int64_t base = 0;
        for (int q = 0; q < tid; q++) base += cnt[3 * q];
        for (int64_t k = lo; k < hi; k++) {
            double b = floor(y[k] / p->ds);
            if (b < 0.0 || b >= (double)p->ncy) continue;
            s->tx[base] = x[k]; s->ty[base] = y[k]; s->tvx[base] = vx[k]; s->tvy[base] = vy[k];
            base++;
        }
    }
    for (int q = 0; q < nt; q++) { kept += cnt[3 * q]; probe += cnt[3 * q + 1]; top += cnt[3 * q + 2]; }
    free(cnt);
#else
    for (int64_t k = 0; k < n; k++) {
        double b = floor(y[k] / p->ds);
        if (b < 0.0) { probe++; continue; }
        if (b >= (double)p->ncy) { top++; continue; }
        s->tx[kept] = x[k]; s->ty[kept] = y[k]; s->tvx[kept] = vx[k]; s->tvy[kept] = vy[k];
        kept++;
    }
#endif
    /* swap compacted arrays in */
    double *t0 = s->x[sp]; s->x[sp] = s->tx; s->tx = t0;
    t0 = s->y[sp]; s->y[sp] = s->ty; s->ty = t0;
    t0 = s->vx[sp]; s->vx[sp] = s->tvx; s->tvx = t0;
    t0 = s->vy[sp]; s->vy[sp] = s->tvy; s->tvy = t0;
    x = s->x[sp]; y = s->y[sp]; vx = s->vx[sp]; vy = s->vy[sp];

    double dtin = sp == 0 ? p->dtin_e : p->dtin_i;
    int in = orc_injection_count(t, p->dt, &s->tin[sp], dtin, s->tin_seq, s->tin_seq_cap);
    if (in > s->tin_seq_cap) in = s->tin_seq_cap;
    if (kept + in > s->cap) in = (int)(s->cap - kept);
    orc_inject_f64(p, sp, s->ninj[sp], in, t, s->tin_seq, s->Ex, s->Ey, x + kept, y + kept,
                   vx + kept, vy + kept);
    s->ninj[sp] += (uint64_t)in;
    /* cyclic wrap of the surviving bulk (Q6: into [0, Lx)) */
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < kept; k++) {
        if (x[k] < 0.0) x[k] += p->Lx;
        if (x[k] >= p->Lx) x[k] -= p->Lx;
    }
    s->n[sp] = kept + in;
    s->probe[sp] = probe;
    s->top[sp] = top;
    s->injected[sp] = in;
}

void orc_sim_step(orc_sim *s, double t)
{
    const orc_params *p = &s->p;
    sim_deposit_solve_field(s, s->jtol);
    for (int sp = 0; sp < 2; sp++)
        orc_push_f64(s->n[sp], s->x[sp], s->y[sp], s->vx[sp], s->vy[sp], s->Ex, s->Ey,
                     p->ncx, p->ncy, p->ds, species_q(sp), species_mass(p, sp), p->dt);
    if (!s->tin_started) { /* src/cc.cu:35-36: first call sets tin = t + dtin */
        s->tin[0] = t + p->dtin_e;
        s->tin[1] = t + p->dtin_i;
        s->tin_started = 1;
    }
    sim_cc_species(s, 0, t);
    sim_cc_species(s, 1, t);
}

int64_t orc_sim_count(const orc_sim *s, int species) { return s->n[species]; }

void orc_sim_get_particles(const orc_sim *s, int sp, double *x, double *y, double *vx,
                           double *vy)
{
    size_t b = (size_t)s->n[sp] * sizeof(double);
    memcpy(x, s->x[sp], b); memcpy(y, s->y[sp], b);
    memcpy(vx, s->vx[sp], b); memcpy(vy, s->vy[sp], b);
}

void orc_sim_set_particles(orc_sim *s, int sp, int64_t n, const double *x, const double *y,
                           const double *vx, const double *vy)
{
    if (n > s->cap) n = s->cap;
    size_t b = (size_t)n * sizeof(double);
    memcpy(s->x[sp], x, b); memcpy(s->y[sp], y, b);
    memcpy(s->vx[sp], vx, b); memcpy(s->vy[sp], vy, b);
    s->n[sp] = n;
}

void orc_sim_get_mesh(const orc_sim *s, double *rho, double *phi, double *Ex, double *Ey)
{
    size_t b = (size_t)s->p.ncx * (s->p.ncy + 1) * sizeof(double);
    if (rho) memcpy(rho, s->rho, b);
    if (phi) memcpy(phi, s->phi, b);
    if (Ex) memcpy(Ex, s->Ex, b);
    if (Ey) memcpy(Ey, s->Ey, b);
}

void orc_sim_set_phi(orc_sim *s, const double *phi)
{
    memcpy(s->phi, phi, (size_t)s->p.ncx * (s->p.ncy + 1) * sizeof(double));
}

void orc_sim_tallies(const orc_sim *s, int64_t probe[2], int64_t top[2], int64_t injected[2],
                     int *sweeps, double *resid)
{
    for (int sp = 0; sp < 2; sp++) {
        probe[sp] = s->probe[sp];
        top[sp] = s->top[sp];
        injected[sp] = s->injected[sp];
    }
    *sweeps = s->sweeps;
    *resid = s->resid;
}

void orc_sim_clock(const orc_sim *s, double tin[2], uint64_t ninj_total[2])
{
    tin[0] = s->tin[0]; tin[1] = s->tin[1];
    ninj_total[0] = s->ninj[0]; ninj_total[1] = s->ninj[1];
}

void orc_sim_set_clock(orc_sim *s, const double tin[2], const uint64_t ninj_total[2],
                       int tin_started)
{
    s->tin[0] = tin[0]; s->tin[1] = tin[1];
    s->ninj[0] = ninj_total[0]; s->ninj[1] = ninj_total[1];
    s->tin_started = tin_started;
}

/* ===================================================================== */
/* FP32 mirror layer                                                     */
/* ===================================================================== */

void orc_f32c_make(const orc_params *p, orc_f32c *c)
{
    c->inv_ds = (float)(1.0 / p->ds);
    c->Lx = (float)p->Lx;
    c->Ly = (float)p->Ly;
    c->dt = (float)p->dt;
    for (int s = 0; s < 2; s++) {
        double q = species_q(s), m = species_mass(p, s);
        c->qdtm[s] = (float)(q * p->dt / m);
        c->hkick[s] = (float)(-0.5 * q * p->dt / m);
    }
    c->ncx = p->ncx;
    c->ncy = p->ncy;
}

/* Cell of an FP32 position and its in-cell fractions.  One product, one
 * floor, clamps -- the device computes exactly this. */
int32_t orc_cell_of_f32(const orc_f32c *c, float x, float y, float *fx, float *fy,
                        int32_t *ic, int32_t *jc)
{
    float tx = x * c->inv_ds;
    float ty = y * c->inv_ds;
    float bx = fminf(fmaxf(floorf(tx), 0.0f), (float)(c->ncx - 1));
    float by = fminf(fmaxf(floorf(ty), 0.0f), (float)(c->ncy - 1));
    float ax = tx - bx;
    float ay = ty - by;
    ax = fminf(fmaxf(ax, 0.0f), 1.0f);
    ay = fminf(fmaxf(ay, 0.0f), 1.0f);
    *fx = ax;
    *fy = ay;
    *ic = (int32_t)bx;
    *jc = (int32_t)by;
    return *jc * c->ncx + *ic;
}

/* Initial load rounded to FP32: x wrapped into [0, Lx), y nudged by whole
 * ulps until its FP32 cell row is the row it was drawn for. */
static void round_initial(const orc_f32c *c, int64_t row, double x, double y, double vx,
                          double vy, float *ox, float *oy, float *ovx, float *ovy)
{
    float xf = (float)x;
    if (xf < 0.0f) xf += c->Lx;
    if (xf >= c->Lx) xf -= c->Lx;
    float yf = (float)y;
    for (int q = 0; q < 64 && floorf(yf * c->inv_ds) > (float)row; q++) yf = nextafterf(yf, -INFINITY);
    for (int q = 0; q < 64 && floorf(yf * c->inv_ds) < (float)row; q++) yf = nextafterf(yf, INFINITY);
    *ox = xf; *oy = yf; *ovx = (float)vx; *ovy = (float)vy;
}

void orc_create_particles_f32(const orc_params *p, int species, float *x, float *y, float *vx,
                              float *vy)
{
    orc_f32c c;
    orc_f32c_make(p, &c);
    int64_t N = p->per_row * p->ncy;
    for (int64_t g = 0; g < N; g++) {
        int64_t row = g / p->per_row;
        double dx, dy, dvx, dvy;
        draw_initial(p, species, g, row, &dx, &dy, &dvx, &dvy);
        round_initial(&c, row, dx, dy, dvx, dvy, &x[g], &y[g], &vx[g], &vy[g]);
    }
}

/* int64 fixed-point CIC: 16-bit fractions per axis, so every particle puts
 * exactly 2^32 charge units on its four vertices (charge quantum 2^-32 e). */
void orc_deposit_fixed_f32(int64_t n, const float *x, const float *y, const orc_f32c *c,
                           int sign, int64_t *acc)
{
    int ncx = c->ncx;
    for (int64_t k = 0; k < n; k++) {
        float fx, fy;
        int32_t ic, jc;
        orc_cell_of_f32(c, x[k], y[k], &fx, &fy, &ic, &jc);
        int64_t wx = (int64_t)rintf(fx * 65536.0f);
        int64_t wy = (int64_t)rintf(fy * 65536.0f);
        int64_t gx = 65536 - wx, gy = 65536 - wy;
        int i1 = ic + 1 == ncx ? 0 : ic + 1;
        acc[(int64_t)jc * ncx + ic] += sign * (gx * gy);
        acc[(int64_t)jc * ncx + i1] += sign * (wx * gy);
        acc[(int64_t)(jc + 1) * ncx + ic] += sign * (gx * wy);
        acc[(int64_t)(jc + 1) * ncx + i1] += sign * (wx * wy);
    }
}

/* FP32 CIC gather from interleaved (Ex, Ey) nodes, fixed operation order:
 * the weighted corners accumulated with fused multiply-adds, as the device
 * does. */
void orc_gather_f32(const orc_f32c *c, const float *Exy, float x, float y, float *ex,
                    float *ey)
{
    float fx, fy;
    int32_t ic, jc;
    orc_cell_of_f32(c, x, y, &fx, &fy, &ic, &jc);
    int ncx = c->ncx;
    int i1 = ic + 1 == ncx ? 0 : ic + 1;
    int64_t k00 = (int64_t)jc * ncx + ic, k10 = (int64_t)jc * ncx + i1;
    int64_t k01 = k00 + ncx, k11 = k10 + ncx;
    float gx = 1.0f - fx, gy = 1.0f - fy;
    float w00 = gx * gy, w10 = fx * gy, w01 = gx * fy, w11 = fx * fy;
    *ex = fmaf(w11, Exy[2 * k11],
               fmaf(w01, Exy[2 * k01], fmaf(w10, Exy[2 * k10], w00 * Exy[2 * k00])));
    *ey = fmaf(w11, Exy[2 * k11 + 1],
               fmaf(w01, Exy[2 * k01 + 1], fmaf(w10, Exy[2 * k10 + 1], w00 * Exy[2 * k00 + 1])));
}

int64_t orc_stable_sort_slots(int64_t n, const int32_t *key, const int8_t *slot,
                              int32_t nkeys, int64_t *perm, int64_t *offsets)
{
    int64_t nb = (int64_t)nkeys * 10;
    int64_t *cnt = (int64_t *)calloc((size_t)nb + 1, sizeof(int64_t));
    int64_t kept = 0;
    for (int64_t k = 0; k < n; k++) {
        if (key[k] < 0) continue;
        cnt[(int64_t)key[k] * 10 + slot[k] + 1]++;
        kept++;
    }
    for (int64_t b = 0; b < nb; b++) cnt[b + 1] += cnt[b];
    if (offsets)
        for (int64_t k = 0; k <= nkeys; k++) offsets[k] = cnt[k * 10];
    for (int64_t k = 0; k < n; k++) {
        if (key[k] < 0) continue;
        perm[cnt[(int64_t)key[k] * 10 + slot[k]]++] = k;
    }
    free(cnt);
    return kept;
}

void orc_mirror_init(const orc_params *p, int species, float *x, float *y, float *vx,
                     float *vy, int64_t *offsets)
{
    orc_f32c c;
    orc_f32c_make(p, &c);
    int64_t N = p->per_row * p->ncy;
    float *tx = (float *)malloc((size_t)N * 4 * sizeof(float));
    float *ty = tx + N, *tvx = ty + N, *tvy = tvx + N;
    orc_create_particles_f32(p, species, tx, ty, tvx, tvy);
    int32_t *key = (int32_t *)malloc((size_t)N * sizeof(int32_t));
    int8_t *slot = (int8_t *)calloc((size_t)N, 1);
    int64_t *perm = (int64_t *)malloc((size_t)N * sizeof(int64_t));
    for (int64_t k = 0; k < N; k++) {
        float fx, fy;
        int32_t ic, jc;
        key[k] = orc_cell_of_f32(&c, tx[k], ty[k], &fx, &fy, &ic, &jc);
    }
    orc_stable_sort_slots(N, key, slot, p->ncx * p->ncy, perm, offsets);
    for (int64_t k = 0; k < N; k++) {
        int64_t s = perm[k];
        x[k] = tx[s]; y[k] = ty[s]; vx[k] = tvx[s]; vy[k] = tvy[s];
    }
    free(tx); free(key); free(slot); free(perm);
}

void orc_mirror_half_kick(const orc_f32c *c, int species, int64_t n, const float *x,
                          const float *y, float *vx, float *vy, const float *Exy)
{
    float hk = c->hkick[species];
    for (int64_t k = 0; k < n; k++) {
        float ex, ey;
        orc_gather_f32(c, Exy, x[k], y[k], &ex, &ey);
        vx[k] = fmaf(hk, ex, vx[k]);
        vy[k] = fmaf(hk, ey, vy[k]);
    }
}

int64_t orc_mirror_step(const orc_f32c *c, int species, int64_t n, const float *x,
                        const float *y, const float *vx, const float *vy, const float *Exy,
                        int64_t n_in, const float *ix, const float *iy, const float *ivx,
                        const float *ivy, float *ox, float *oy, float *ovx, float *ovy,
                        int64_t *offsets, int64_t *probe, int64_t *top, int64_t *nfar)
{
    int64_t tot = n + n_in;
    float *px = (float *)malloc((size_t)tot * 4 * sizeof(float));
    float *py = px + tot, *pvx = py + tot, *pvy = pvx + tot;
    int32_t *key = (int32_t *)malloc((size_t)tot * sizeof(int32_t));
    int8_t *slot = (int8_t *)malloc((size_t)tot);
    int64_t *perm = (int64_t *)malloc((size_t)tot * sizeof(int64_t));
    int64_t npr = 0, ntp = 0, nf = 0;
    float qdtm = c->qdtm[species];
    int ncx = c->ncx;
    for (int64_t k = 0; k < n; k++) {
        float fx, fy;
        int32_t oic, ojc;
        orc_cell_of_f32(c, x[k], y[k], &fx, &fy, &oic, &ojc);
        float ex, ey;
        orc_gather_f32(c, Exy, x[k], y[k], &ex, &ey);
        /* v += (q dt / m) E, x += dt v as fused multiply-adds (device order) */
        float nvx = fmaf(qdtm, ex, vx[k]);
        float nvy = fmaf(qdtm, ey, vy[k]);
        float nx = fmaf(c->dt, nvx, x[k]);
        float ny = fmaf(c->dt, nvy, y[k]);
        if (nx < 0.0f) nx = nx + c->Lx;
        if (nx >= c->Lx) nx = nx - c->Lx;
        px[k] = nx; py[k] = ny; pvx[k] = nvx; pvy[k] = nvy;
        float ty = ny * c->inv_ds;
        if (ty < 0.0f) { key[k] = -1; slot[k] = 0; npr++; continue; }
        if (ty >= (float)c->ncy) { key[k] = -1; slot[k] = 0; ntp++; continue; }
        int32_t nic, njc;
        key[k] = orc_cell_of_f32(c, nx, ny, &fx, &fy, &nic, &njc);
        int di = nic - oic, dj = njc - ojc;
        if (di == ncx - 1) di = -1;
        else if (di == -(ncx - 1)) di = 1;
        if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1) {
            slot[k] = (int8_t)((1 - di) + 3 * (1 - dj));
        } else {
            slot[k] = 9;
            nf++;
        }
    }
    for (int64_t k = 0; k < n_in; k++) {
        float fx, fy;
        int32_t ic, jc;
        px[n + k] = ix[k]; py[n + k] = iy[k]; pvx[n + k] = ivx[k]; pvy[n + k] = ivy[k];
        key[n + k] = orc_cell_of_f32(c, ix[k], iy[k], &fx, &fy, &ic, &jc);
        slot[n + k] = 9;
    }
    int64_t kept = orc_stable_sort_slots(tot, key, slot, c->ncx * c->ncy, perm, offsets);
    for (int64_t k = 0; k < kept; k++) {
        int64_t s = perm[k];
        ox[k] = px[s]; oy[k] = py[s]; ovx[k] = pvx[s]; ovy[k] = pvy[s];
    }
    *probe = npr;
    *top = ntp;
    *nfar = nf;
    free(px); free(key); free(slot); free(perm);
    return kept;
}

/* Injection evaluated in FP64 from an FP32 field, then rounded; x is wrapped
 * once more in FP32 in case rounding reached Lx, and a y that rounded up to
 * Ly (it is below Ly in FP64: vy < 0) becomes the largest float below Ly. */
void orc_inject_f32(const orc_params *p, int species, uint64_t first_index, int n_in,
                    double t, const double *tin_seq, const float *Exy, float *x, float *y,
                    float *vx, float *vy)
{
    orc_f32c c;
    orc_f32c_make(p, &c);
    int64_t nodes = (int64_t)p->ncx * (p->ncy + 1);
    double *Ex = (double *)malloc((size_t)nodes * sizeof(double));
    double *Ey = (double *)malloc((size_t)nodes * sizeof(double));
    for (int64_t k = 0; k < nodes; k++) { Ex[k] = Exy[2 * k]; Ey[k] = Exy[2 * k + 1]; }
    for (int k = 0; k < n_in; k++) {
        double dx, dy, dvx, dvy;
        inject_one64(p, species, first_index + (uint64_t)k, t, tin_seq[k], 0, 0, 0, 0, 1, Ex,
                     Ey, &dx, &dy, &dvx, &dvy);
        float xf = (float)dx;
        if (xf < 0.0f) xf += c.Lx;
        if (xf >= c.Lx) xf -= c.Lx;
        float yf = (float)dy;
        if (yf >= c.Ly) yf = nextafterf(c.Ly, 0.0f);
        x[k] = xf; y[k] = yf; vx[k] = (float)dvx; vy[k] = (float)dvy;
    }
    free(Ex); free(Ey);
}

/* ---- FP32 mirror of the device Poisson solve (mode 1) ------------------ */

void orc_mirror_rhs(const orc_params *p, const int64_t *acc, double *b64, float *b32)
{
    double bi = 0x1p-32 / (p->ds * p->eps0);
    int64_t nn = (int64_t)p->ncx * (p->ncy + 1);
    for (int64_t k = 0; k < nn; k++) {
        int64_t j = k / p->ncx;
        double bb = (j == 0 || j == p->ncy) ? 0.0 : (double)acc[k] * bi;
        b64[k] = bb;
        b32[k] = (float)bb;
    }
}

static double mirror_resid(int ncx, int ncy, const double *b64, const float *phi)
{
    double rmax = 0.0;
    for (int j = 1; j < ncy; j++)
        for (int i = 0; i < ncx; i++) {
            int64_t k = (int64_t)j * ncx + i;
            int iw = i == 0 ? ncx - 1 : i - 1, ie = i == ncx - 1 ? 0 : i + 1;
            double sum = (((b64[k] + (double)phi[(int64_t)j * ncx + iw]) +
                           (double)phi[(int64_t)j * ncx + ie]) + (double)phi[k + ncx]) +
                         (double)phi[k - ncx];
            double r = fabs(0.25 * (sum - 4.0 * (double)phi[k]));
            if (r > rmax) rmax = r;
        }
    return rmax;
}

int orc_mirror_solve_f32(int ncx, int ncy, const double *b64, const float *b32, float *phi,
                         double tol, int max_sweeps, int check_every, double *resid_out)
{
    int64_t nn = (int64_t)ncx * (ncy + 1);
    float *tmp = (float *)malloc((size_t)nn * sizeof(float));
    memcpy(tmp, phi, (size_t)nn * sizeof(float));
    if (check_every < 1) check_every = 1;
    int s = 0;
    double res = mirror_resid(ncx, ncy, b64, phi);
    while (!(res <= tol) && s < max_sweeps) {
        int k = check_every < max_sweeps - s ? check_every : max_sweeps - s;
        for (int q = 0; q < k; q++) {
            for (int j = 1; j < ncy; j++)
                for (int i = 0; i < ncx; i++) {
                    int64_t c = (int64_t)j * ncx + i;
                    int iw = i == 0 ? ncx - 1 : i - 1, ie = i == ncx - 1 ? 0 : i + 1;
                    float w = phi[(int64_t)j * ncx + iw], e = phi[(int64_t)j * ncx + ie];
                    tmp[c] = 0.25f * ((((b32[c] + w) + e) + phi[c + ncx]) + phi[c - ncx]);
                }
            for (int j = 1; j < ncy; j++)
                memcpy(phi + (int64_t)j * ncx, tmp + (int64_t)j * ncx, (size_t)ncx * sizeof(float));
        }
        s += k;
        res = mirror_resid(ncx, ncy, b64, phi);
    }
    free(tmp);
    if (resid_out) *resid_out = res;
    return s;
}
