// k_mesh.hip -- grid kernels: initial potential, charge conversion, Jacobi
// Poisson sweeps with a Dirichlet probe row, FP64 residual and iterative
// refinement, E = -grad phi.  All are HBM/L2-bandwidth kernels over the
// [ncy+1][ncx] node grid (periodic in x); none is a dense contraction.
#include "cupic_launch.h"

namespace cupic {

#define RUN_PROLOGUE()                                              \
    const int b_ = blockIdx.x;                                      \
    const int r_ = find_run(start, nruns, b_);                      \
    const RunDev &R = rds[r_];                                      \
    const int local = b_ - start[r_];                               \
    RunState *st = R.st;                                            \
    (void)local; (void)st;

// src/init.cu:235-241: phi(j) = (1 - j/ncy) phi_p in every phi buffer.
__global__ __launch_bounds__(kBlock) void k_phi_linear(const RunDev *rds, const int *start,
                                                       int nruns)
{
    RUN_PROLOGUE();
    int k = local * kBlock + threadIdx.x;
    if (k >= R.nnodes) return;
    int j = k / R.ncx;
    double v = (1.0 - (double)j / (double)R.ncy) * R.phi_p;
    R.phi32[0][k] = (float)v;
    R.phi32[1][k] = (float)v;
    R.phi64[k] = v;
    R.d32[0][k] = 0.0f;
    R.d32[1][k] = 0.0f;
    if (k == 0) st->phi_cur = 0;
}

// int64 node charge -> rho (volume ds^3, ds^3/2 on the two Dirichlet rows,
// the intent of src/mesh.cu:242-268) and the Jacobi right-hand side
// ds^2 rho / eps0 (src/mesh.cu:327).  Clears the accumulator.
__global__ __launch_bounds__(kBlock) void k_rho_convert(const RunDev *rds, const int *start,
                                                        int nruns)
{
    RUN_PROLOGUE();
    int k = local * kBlock + threadIdx.x;
    if (k >= R.nnodes) return;
    int j = k / R.ncx;
    long long a = (long long)R.acc[k];
    R.acc[k] = 0ull;
    bool edge = (j == 0 || j == R.ncy);
    double q = (double)a;
    R.rho[k] = (float)(q * (edge ? R.rho_edge : R.rho_interior));
    double bb = edge ? 0.0 : q * R.b_interior;
    R.b64[k] = bb;
    R.b32[k] = (float)bb;
}

__global__ __launch_bounds__(kBlock) void k_solver_begin(const RunDev *rds, const int *start,
                                                         int nruns, int use_args, int mode,
                                                         double tol, int min_sweeps,
                                                         int max_sweeps, int check_every)
{
    RUN_PROLOGUE();
    int mx = use_args ? max_sweeps : R.step_max;
    if (mx > R.max_sweeps_alloc) mx = R.max_sweeps_alloc;
    for (int q = threadIdx.x; q <= R.max_sweeps_alloc; q += kBlock) R.sweep_max[q] = 0u;
    for (int q = threadIdx.x; q <= R.max_sweeps_alloc + 1; q += kBlock) R.resid[q] = 0ull;
    if (threadIdx.x == 0) {
        st->solve_mode = use_args ? mode : R.step_mode;
        st->solve_tol = use_args ? tol : R.step_tol;
        st->solve_min = use_args ? min_sweeps : R.step_min;
        st->solve_max = mx;
        st->solve_check = use_args ? check_every : R.step_check;
        st->solve_done = 0;
        st->sweeps_done = 0;
        st->last_resid = 0.0;
    }
}

__device__ __forceinline__ void mark_done(const RunDev &R, RunState *st, int local, int s,
                                          double res)
{
    if (local == 0 && threadIdx.x == 0) {
        st->sweeps_done = s;
        st->last_resid = res;
        st->solve_done = 1;
    }
}

// One Jacobi sweep s on rows 1..ncy-1 (src/mesh.cu:297-381), double
// buffered (Q3), periodic in x.  Modes 0/1 relax phi itself with rhs b;
// mode 2 relaxes the correction d with rhs r (d starts at 0 each cycle).
// Convergence is decided on the device: mode 0 reads the previous sweep's
// max |dphi|, modes 1/2 the FP64 residual of the check that precedes this
// block of sweeps.
__global__ __launch_bounds__(kBlock) void k_jacobi_sweep(const RunDev *rds, const int *start,
                                                         int nruns, int s, int mode,
                                                         int check_slot, int first)
{
    RUN_PROLOGUE();
    if (st->solve_done) return;
    if (mode == 0) {
        if (s > 0 && s >= st->solve_min &&
            __uint_as_float(R.sweep_max[s - 1]) < (float)st->solve_tol) {
            mark_done(R, st, local, s, (double)__uint_as_float(R.sweep_max[s - 1]));
            return;
        }
        if (s >= st->solve_max) {
            mark_done(R, st, local, s, s > 0 ? (double)__uint_as_float(R.sweep_max[s - 1]) : 0.0);
            return;
        }
    } else if (check_slot >= 0) {
        double res = __longlong_as_double((long long)R.resid[check_slot]);
        if (res <= st->solve_tol || s >= st->solve_max) {
            mark_done(R, st, local, s, res);
            return;
        }
    }
    if (s >= st->solve_max) return;
    int k = local * kBlock + threadIdx.x;
    float dmax = 0.0f;
    if (k < R.nnodes) {
        int ncx = R.ncx;
        int j = k / ncx;
        int i = k - j * ncx;
        if (j >= 1 && j < R.ncy) {
            int kw = (i == 0) ? k + ncx - 1 : k - 1;
            int ke = (i == ncx - 1) ? k - ncx + 1 : k + 1;
            if (mode == 2) {
                const float *din = R.d32[s & 1];
                float *dout = R.d32[(s + 1) & 1];
                float rr = R.r32[k];
                float v;
                if (first) {
                    v = 0.25f * rr;
                } else {
                    v = 0.25f * ((((rr + din[kw]) + din[ke]) + din[k + ncx]) + din[k - ncx]);
                }
                dout[k] = v;
            } else {
                int pc = st->phi_cur;
                const float *in = R.phi32[(pc + s) & 1];
                float *out = R.phi32[(pc + s + 1) & 1];
                float c = in[k];
                float v = 0.25f * ((((R.b32[k] + in[kw]) + in[ke]) + in[k + ncx]) + in[k - ncx]);
                out[k] = v;
                dmax = fabsf(v - c);
            }
        }
    }
    if (mode == 0) {
        dmax = wave_max_f(dmax);
        if ((threadIdx.x & (kWave - 1)) == 0 && dmax > 0.0f)
            atomicMax(&R.sweep_max[s], __float_as_uint(dmax));
    }
}

// mode 2: phi64 += d, the correction solved by the sweeps since check s.
__global__ __launch_bounds__(kBlock) void k_ir_update(const RunDev *rds, const int *start,
                                                      int nruns, int s)
{
    RUN_PROLOGUE();
    if (st->solve_done) return;
    int k = local * kBlock + threadIdx.x;
    if (k >= R.nnodes) return;
    int j = k / R.ncx;
    if (j < 1 || j >= R.ncy) return;
    R.phi64[k] += (double)R.d32[s & 1][k];
}

// FP64 residual max over interior nodes of |0.25 (b + sum of neighbours) - phi|,
// the same measure as the reference's max |dphi| (src/mesh.cu:354).  Mode 1
// evaluates it on the FP32 phi; mode 2 on the FP64 master and also stores
// the correction right-hand side r = b + sum - 4 phi.
__global__ __launch_bounds__(kBlock) void k_residual(const RunDev *rds, const int *start,
                                                     int nruns, int c, int s, int mode)
{
    RUN_PROLOGUE();
    if (st->solve_done) return;
    int k = local * kBlock + threadIdx.x;
    double rmax = 0.0;
    if (k < R.nnodes) {
        int ncx = R.ncx;
        int j = k / ncx;
        int i = k - j * ncx;
        if (j >= 1 && j < R.ncy) {
            int kw = (i == 0) ? k + ncx - 1 : k - 1;
            int ke = (i == ncx - 1) ? k - ncx + 1 : k + 1;
            double pc_, pw, pe, pn, ps;
            if (mode == 2) {
                const double *p = R.phi64;
                pc_ = p[k]; pw = p[kw]; pe = p[ke]; pn = p[k + ncx]; ps = p[k - ncx];
            } else {
                const float *p = R.phi32[(st->phi_cur + s) & 1];
                pc_ = p[k]; pw = p[kw]; pe = p[ke]; pn = p[k + ncx]; ps = p[k - ncx];
            }
            double sum = (((R.b64[k] + pw) + pe) + pn) + ps;
            double rr = sum - 4.0 * pc_;
            if (mode == 2) R.r32[k] = (float)rr;
            rmax = fabs(0.25 * rr);
        }
    }
    rmax = wave_max_d(rmax);
    if ((threadIdx.x & (kWave - 1)) == 0 && rmax > 0.0)
        atomicMax(&R.resid[c], (unsigned long long)__double_as_longlong(rmax));
    (void)s;
}

__global__ __launch_bounds__(kBlock) void k_solver_end(const RunDev *rds, const int *start,
                                                       int nruns, int last_check)
{
    RUN_PROLOGUE();
    if (threadIdx.x != 0) return;
    if (!st->solve_done) {
        int s = st->solve_max;
        st->sweeps_done = s;
        if (st->solve_mode == 0)
            st->last_resid = s > 0 ? (double)__uint_as_float(R.sweep_max[s - 1]) : 0.0;
        else
            st->last_resid = __longlong_as_double((long long)R.resid[last_check]);
        st->solve_done = 1;
    }
    if (st->solve_mode != 2) st->phi_cur = (st->phi_cur + st->sweeps_done) & 1;
    st->sweeps_total += st->sweeps_done;
    st->solves += 1;
}

// E = -grad phi (src/mesh.cu:413-468): central differences with periodic x
// on interior rows; one-sided Ey and Ex = 0 on rows 0 and ncy.  Evaluated in
// FP64 from the FP32 (modes 0/1) or FP64 (mode 2) potential.
__global__ __launch_bounds__(kBlock) void k_field(const RunDev *rds, const int *start,
                                                  int nruns)
{
    RUN_PROLOGUE();
    int k = local * kBlock + threadIdx.x;
    if (k >= R.nnodes) return;
    int ncx = R.ncx;
    int j = k / ncx;
    int i = k - j * ncx;
    bool m2 = (st->solve_mode == 2);
    const float *p32 = R.phi32[st->phi_cur];
    const double *p64 = R.phi64;
    auto P = [&](int q) -> double { return m2 ? p64[q] : (double)p32[q]; };
    double ex, ey;
    if (j == 0) {
        ex = 0.0;
        ey = (P(k) - P(k + ncx)) / R.ds;
    } else if (j == R.ncy) {
        ex = 0.0;
        ey = (P(k - ncx) - P(k)) / R.ds;
    } else {
        int kw = (i == 0) ? k + ncx - 1 : k - 1;
        int ke = (i == ncx - 1) ? k - ncx + 1 : k + 1;
        ex = (P(kw) - P(ke)) / (2.0 * R.ds);
        ey = (P(k - ncx) - P(k + ncx)) / (2.0 * R.ds);
    }
    R.E[k] = make_float2((float)ex, (float)ey);
}

// ---------------------------------------------------------------- launchers

#define LAUNCH_K(kern, L, ...)                                                         \
    do {                                                                               \
        if ((L).nblocks > 0)                                                           \
            hipLaunchKernelGGL(kern, dim3((L).nblocks), dim3(kBlock), 0, (L).stream,   \
                               (L).rds, (L).start, (L).nruns, ##__VA_ARGS__);          \
    } while (0)

void launch_phi_linear(const Launch &L) { LAUNCH_K(k_phi_linear, L); }
void launch_rho_convert(const Launch &L) { LAUNCH_K(k_rho_convert, L); }
void launch_solver_begin(const Launch &L, int use_args, int mode, double tol, int min_sweeps,
                         int max_sweeps, int check_every)
{
    LAUNCH_K(k_solver_begin, L, use_args, mode, tol, min_sweeps, max_sweeps, check_every);
}
void launch_jacobi_sweep(const Launch &L, int s, int mode, int check_slot, int first)
{
    LAUNCH_K(k_jacobi_sweep, L, s, mode, check_slot, first);
}
void launch_ir_update(const Launch &L, int s) { LAUNCH_K(k_ir_update, L, s); }
void launch_residual(const Launch &L, int c, int s, int mode)
{
    LAUNCH_K(k_residual, L, c, s, mode);
}
void launch_solver_end(const Launch &L, int last_check) { LAUNCH_K(k_solver_end, L, last_check); }
void launch_field(const Launch &L) { LAUNCH_K(k_field, L); }

}  // namespace cupic
