// k_mesh.hip -- grid kernels: initial potential, charge conversion, Jacobi
// Poisson sweeps with a Dirichlet probe row, FP64 residual and iterative
// refinement, E = -grad phi.  All are HBM/L2-bandwidth kernels over the
// [ncy+1][ncx] node grid (periodic in x); none is a dense contraction.
#include "cupic_launch.h"

namespace cupic {


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
    R.phi32[2][k] = (float)v;
    R.phi64[k] = v;
    R.d32[0][k] = 0.0f;
    R.d32[1][k] = 0.0f;
    if (k == 0) st->phi_cur = 0;
}

// int64 node charge -> rho (volume ds^3, ds^3/2 on the two Dirichlet rows,
// the intent of src/mesh.cu:242-268) and the Jacobi right-hand side
// ds^2 rho / eps0 (src/mesh.cu:327).  Clears the accumulator.
__device__ void solver_begin_work(const RunDev &R, RunState *st, int use_args, int mode,
                                  double tol, int min_sweeps, int max_sweeps, int check_every);

// begin: block 0 also starts the step's solve with the run's own settings
// (the work of k_solver_begin), which saves a launch where a step runs the
// conversion and the solve back to back.
__global__ __launch_bounds__(kBlock) void k_rho_convert(const RunDev *rds, const int *start,
                                                        int nruns, int begin)
{
    RUN_PROLOGUE();
    if (begin && local == 0) solver_begin_work(R, st, 0, 0, 0.0, 0, 0, 0);
    int k = local * kBlock + threadIdx.x;
    if (k >= R.nnodes) return;
    int j = k / R.ncx;
    long long a = (long long)R.acc[k];
    R.acc[k] = 0ull;
    float rho;
    double bb;
    charge_to_rhs(R, a, j, &rho, &bb);
    R.rho[k] = rho;
    R.b64[k] = bb;
    R.b32[k] = (float)bb;
}

__device__ void solver_begin_work(const RunDev &R, RunState *st, int use_args, int mode,
                                  double tol, int min_sweeps, int max_sweeps, int check_every)
{
    int mx = use_args ? max_sweeps : R.step_max;
    if (mx > R.max_sweeps_alloc) mx = R.max_sweeps_alloc;
    const int md = use_args ? mode : R.step_mode;
    const int ck = max(1, use_args ? check_every : R.step_check);
    // clear the reduction slots this solve uses: one per sweep (mode 0), one
    // per residual check (modes 1, 2: sweeps 0, ck, 2 ck, ... and mx), two
    // for mode 3
    const int nsweep = md == 0 ? mx + 1 : 0;
    const int nres = md == 0 ? 0 : min(R.max_sweeps_alloc + 2, (mx + ck - 1) / ck + 2);
    for (int q = threadIdx.x; q < nsweep * kRedSlots; q += kBlock) R.sweep_max[q] = 0u;
    for (int q = threadIdx.x; q < nres * kRedSlots; q += kBlock) R.resid[q] = 0ull;
    if (threadIdx.x == 0) {
        st->solve_mode = use_args ? mode : R.step_mode;
        st->solve_tol = use_args ? tol : R.step_tol;
        st->solve_min = use_args ? min_sweeps : R.step_min;
        st->solve_max = mx;
        st->solve_check = use_args ? check_every : R.step_check;
        st->solve_done = 0;
        st->sweeps_done = 0;
        st->last_resid = 0.0;
        st->phi_final = st->phi_cur;
    }
}

__global__ __launch_bounds__(kBlock) void k_solver_begin(const RunDev *rds, const int *start,
                                                         int nruns, int use_args, int mode,
                                                         double tol, int min_sweeps,
                                                         int max_sweeps, int check_every)
{
    RUN_PROLOGUE();
    solver_begin_work(R, st, use_args, mode, tol, min_sweeps, max_sweeps, check_every);
}



// Ends a solve at sweep s with residual res; the solution is in phi buffer
// phi_at(phi_cur, par), or in the third buffer when par < 0 (a check inside
// a blocked launch).
__device__ __forceinline__ void mark_done(const RunDev &R, RunState *st, int local, int s,
                                          double res, int par)
{
    if (local == 0 && threadIdx.x == 0) {
        st->sweeps_done = s;
        st->last_resid = res;
        st->phi_final = par < 0 ? phi_third(st->phi_cur) : phi_at(st->phi_cur, par);
        st->solve_done = 1;
    }
}

// block-wide max of one value per thread, into a slotted atomic
template <int NW = kBlock / kWave, class T, class F>
__device__ __forceinline__ void block_max_atomic(T v, T *sh, F emit)
{
    int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        T u = __shfl_xor(v, o, kWave);
        v = v > u ? v : u;
    }
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T m = sh[0];
        for (int q = 1; q < NW; q++) m = m > sh[q] ? m : sh[q];
        emit(m);
    }
}

// One Jacobi sweep s on rows 1..ncy-1 (src/mesh.cu:297-381), double
// buffered (Q3), periodic in x, with the reference stopping rule (mode 0):
// the sweep first reads the previous sweep's max |dphi| and stops the solve
// once it is below tol after at least solve_min sweeps.
__global__ __launch_bounds__(kBlock) void k_jacobi_sweep(const RunDev *rds, const int *start,
                                                         int nruns, int s)
{
    RUN_PROLOGUE();
    __shared__ float sh[kBlock / kWave];
    if (st->solve_done) return;
    if (s > 0 && s >= st->solve_min && (double)read_sweep_max(R, s - 1) < st->solve_tol) {
        mark_done(R, st, local, s, (double)read_sweep_max(R, s - 1), s);
        return;
    }
    if (s >= st->solve_max) {
        mark_done(R, st, local, s, s > 0 ? (double)read_sweep_max(R, s - 1) : 0.0, s);
        return;
    }
    int k = local * kBlock + threadIdx.x;
    float dmax = 0.0f;
    if (k < R.nnodes) {
        int ncx = R.ncx;
        int j = k / ncx;
        int i = k - j * ncx;
        if (j >= 1 && j < R.ncy) {
            int kw = (i == 0) ? k + ncx - 1 : k - 1;
            int ke = (i == ncx - 1) ? k - ncx + 1 : k + 1;
            int pc = st->phi_cur;
            const CUPIC_GLOBAL float *in = gp(R.phi32[phi_at(pc, s)]);
            CUPIC_GLOBAL float *out = gp(R.phi32[phi_at(pc, s + 1)]);
            float c = in[k];
            float v = 0.25f * ((((gp(R.b32)[k] + in[kw]) + in[ke]) + in[k + ncx]) + in[k - ncx]);
            out[k] = v;
            dmax = fabsf(v - c);
        }
    }
    unsigned int *slot = &R.sweep_max[s * kRedSlots + (local & (kRedSlots - 1))];
    block_max_atomic(dmax, sh, [&](float m) {
        if (m > 0.0f) atomicMax(slot, __float_as_uint(m));
    });
}

// Neighbour values one lane away, as VALU data-parallel-primitive moves:
// wave_shr:1 gives lane l the value of lane l-1, wave_shl:1 that of lane
// l+1.  Lanes 0 and 63 read 0 (they are never updated nor stored); with
// zero fill the compiler folds each move into the add that consumes it.
__device__ __forceinline__ float lane_prev(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF,
                                                      true));
}

__device__ __forceinline__ float lane_next(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF,
                                                      true));
}

// nsw <= H - 1 Jacobi sweeps s0 .. s0+nsw-1 in one launch (modes 1, 2).
// A workgroup holds a 64 x (kJacWaves * kJacRows) node region -- a TX x TY
// tile with an H-node halo -- in registers: lane = region column, each
// wave kJacRows consecutive rows.  East/west neighbours come from the next
// lanes, north/south from the thread's own rows; only the rows at wave
// boundaries go through LDS.  The valid region shrinks by one node per
// sweep and still covers the tile and one ring after nsw sweeps, so the
// iterates equal those of nsw single sweeps (same operations per node).
// Mode 1 relaxes phi (rhs b), mode 2 the refinement correction d (rhs r),
// d starting at 0 when first.  par counts the launches so far (of the solve
// in mode 1, of the correction cycle in mode 2) and selects the input
// buffer; the output goes to the other one.
// Residual checks (mode 1): with cslot >= 0 the launch first reads the FP64
// residual of the check at sweep cat and ends the solve if it is met; that
// iterate is in buffer phi_at(phi_cur, cpar), or in the third buffer when
// cpar < 0 (the check fell inside the previous launch).  With rslot >= 0 the
// launch reduces the FP64 residual of its tile after rat of its sweeps (0:
// its input) into slot rslot; when 0 < rat < nsw it also stores the tile at
// that sweep into the third buffer.
// RK: 0, the launch reduces no residual; 1, at its start (rat = 0) or end
// (rat = nsw); 2, inside it (0 < rat < nsw, with the third-buffer copy and
// the FP64 rhs fetched a sweep ahead).  Separate instances keep the register
// count of the common ones low (53 VGPRs for RK 0, against 80 for RK 2),
// which matters to batches, whose launches fill the device many times.
template <int H, int RK>
__global__ __launch_bounds__(kJacThreads) void k_jacobi_block(const RunDev *rds, const int *start,
                                                         int nruns, int s0, int nsw, int mode,
                                                         int first, int par, int cslot, int cat,
                                                         int cpar, int rslot, int rat)
{
    RUN_PROLOGUE();
    constexpr int NR = kJacRows;
    constexpr int kJacTX = jac_tx(H), kJacTY = jac_ty(H), kJacH = H;
    // wave-boundary rows of waves -1 .. kJacWaves at index w + 1; the two
    // outer entries stay zero
    __shared__ float xch[2][kJacWaves + 2][2][kWave];
    __shared__ double shr[kJacWaves];
    // the run state and the residual slots in one round trip, ahead of the
    // exits that depend on them
    const int done = st->solve_done, phi_cur = st->phi_cur, smax = st->solve_max;
    const double tol = st->solve_tol;
    const double res = cslot >= 0 ? read_resid(R, cslot) : 0.0;
    if (done) return;
    if (cslot >= 0 && (res <= tol || cat >= smax)) {
        mark_done(R, st, local, cat, res, cpar);
        return;
    }
    if (s0 >= smax) return;
    const int ncx = R.ncx, ncy = R.ncy;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
    const int tiles_x = (ncx + kJacTX - 1) / kJacTX;
    const int by = local / tiles_x, bx = local - by * tiles_x;
    int gx = (bx * kJacTX - kJacH + lane) % ncx;
    if (gx < 0) gx += ncx;
    const int row0 = 1 + by * kJacTY - kJacH + NR * w;  // global row of v[0]
    const CUPIC_GLOBAL float *in, *rhs;
    CUPIC_GLOBAL float *out;
    if (mode == 2) {
        in = gp(R.d32[par & 1]);
        out = gp(R.d32[(par + 1) & 1]);
        rhs = gp(R.r32);
    } else {
        in = gp(R.phi32[phi_at(phi_cur, par)]);
        out = gp(R.phi32[phi_at(phi_cur, par + 1)]);
        rhs = gp(R.b32);
    }
    if (w == 0) {
        for (int q = 0; q < 2; q++) {
            xch[q][0][0][lane] = xch[q][0][1][lane] = 0.0f;
            xch[q][kJacWaves + 1][0][lane] = xch[q][kJacWaves + 1][1][lane] = 0.0f;
        }
    }
    float v[NR], b[NR];
    unsigned upd = 0;  // bit i: row i of this thread is updated
    const bool lane_in = lane >= 1 && lane <= kWave - 2;
    const bool zero_in = (mode == 2 && first);
#pragma unroll
    for (int i = 0; i < NR; i++) {
        int gy = row0 + i;
        int lr = NR * w + i;
        int gyc = min(max(gy, 0), ncy);  // rows outside the grid load a clamped row
        unsigned g = (unsigned)(gyc * ncx + gx);
        float vin = in[g], bin = rhs[g];
        bool ok = gy >= 0 && gy <= ncy;
        v[i] = (ok && !zero_in) ? vin : 0.0f;
        b[i] = ok ? bin : 0.0f;
        if (lane_in && lr >= 1 && lr <= NR * kJacWaves - 2 && gy >= 1 && gy < ncy)
            upd |= 1u << i;
    }
    const bool col_in = lane >= kJacH && lane < kJacH + kJacTX && bx * kJacTX + lane - kJacH < ncx;
    // the tile's nodes in a grid buffer
    auto store_tile = [&](CUPIC_GLOBAL float *dst) {
#pragma unroll
        for (int i = 0; i < NR; i++) {
            int lr = NR * w + i, gy = row0 + i;
            if (col_in && lr >= kJacH && lr < kJacH + kJacTY && gy >= 1 && gy < ncy)
                dst[gy * ncx + (bx * kJacTX + lane - kJacH)] = v[i];
        }
    };
    // the FP64 right-hand side of the tile's nodes (0 elsewhere), loaded a
    // sweep ahead of the residual that needs it
    double b64t[NR];
    auto load_b64 = [&]() {
        const CUPIC_GLOBAL double *b64 = gp(R.b64);
#pragma unroll
        for (int i = 0; i < NR; i++) {
            int lr = NR * w + i, gy = row0 + i;
            const bool in_tile = col_in && lr >= kJacH && lr < kJacH + kJacTY && gy >= 1 && gy < ncy;
            b64t[i] = in_tile ? b64[gy * ncx + (bx * kJacTX + lane - kJacH)] : 0.0;
        }
    };
    // FP64 residual of the tile's interior nodes at the current iterate,
    // given the neighbouring waves' boundary rows, into slot rslot
    auto tile_resid = [&](float below, float above) {
        double rmax = 0.0;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            float W = lane_prev(v[i]);
            float E = lane_next(v[i]);
            float S = i > 0 ? v[i - 1] : below;
            float N = i < NR - 1 ? v[i + 1] : above;
            int lr = NR * w + i, gy = row0 + i;
            if (col_in && lr >= kJacH && lr < kJacH + kJacTY && gy >= 1 && gy < ncy) {
                double sum = (((b64t[i] + (double)W) + (double)E) + (double)N) + (double)S;
                rmax = fmax(rmax, fabs(0.25 * (sum - 4.0 * (double)v[i])));
            }
        }
        unsigned long long *slot = &R.resid[rslot * kRedSlots + (local & (kRedSlots - 1))];
        block_max_atomic<kJacWaves>(rmax, shr, [&](double m) {
            if (m > 0.0) atomicMax(slot, (unsigned long long)__double_as_longlong(m));
        });
    };
    // Only the Dirichlet rows 0 and ncy, and the rows outside the grid beyond
    // them, must keep their values.  The region's edge lanes and edge rows
    // lack a neighbour, so they are wrong after one sweep whether they keep
    // their value or take the (finite) update, and the valid region shrinks
    // alike.  Waves without such rows -- all but those of the first and last
    // tile row -- therefore sweep without a mask.
    auto sweeps = [&](auto keep, int ja, int jb) {
#pragma unroll 1
        for (int j = ja; j < jb; j++) {
            float(*x)[2][kWave] = xch[j & 1];
            x[w + 1][0][lane] = v[0];
            x[w + 1][1][lane] = v[NR - 1];
            __syncthreads();
            float prev = x[w][1][lane];  // old value of the row below
            const float above = x[w + 2][0][lane];
#pragma unroll
            for (int i = 0; i < NR; i++) {
                const float cur = v[i];
                float W = lane_prev(cur);
                float E = lane_next(cur);
                float N = i < NR - 1 ? v[i + 1] : above;
                float nvv = 0.25f * ((((b[i] + W) + E) + N) + prev);
                prev = cur;
                v[i] = keep(i) ? cur : nvv;
            }
        }
    };
    // the FP64 residual at local sweep j, through the exchange buffer the
    // next sweep writes with the same rows
    auto resid_at = [&](int j) {
        float(*x)[2][kWave] = xch[j & 1];
        x[w + 1][0][lane] = v[0];
        x[w + 1][1][lane] = v[NR - 1];
        __syncthreads();
        tile_resid(x[w][1][lane], x[w + 2][0][lane]);
    };
    const int jres = RK > 0 && rslot >= 0 ? rat : -1;  // local sweep of the residual, if any
    auto run = [&](auto keep) {
        if (RK == 2) {
            // a check inside the launch: its residual, with the rhs fetched
            // during the sweep before, and the tile kept in the third buffer
            // in case the solve ends there
            sweeps(keep, 0, jres - 1);
            load_b64();
            sweeps(keep, jres - 1, jres);
            resid_at(jres);
            store_tile(gp(R.phi32[phi_third(phi_cur)]));
            sweeps(keep, jres, nsw);
        } else {
            if (RK == 1 && jres == 0) {
                // the residual of the input (the solve's first check)
                load_b64();
                resid_at(0);
            }
            sweeps(keep, 0, nsw);
        }
    };
    const bool dirichlet_wave =
        __builtin_amdgcn_readfirstlane((row0 <= 0 || row0 + NR - 1 >= ncy) ? 1 : 0) != 0;
    if (dirichlet_wave)
        run([&](int i) { return ((upd >> i) & 1u) == 0u; });
    else
        run([](int) { return false; });
    store_tile(out);
    if (RK == 1 && jres == nsw) {  // the residual after the last sweep
        load_b64();
        resid_at(nsw);
    }
}

// mode 2: phi64 += d, the correction solved by the last cycle of sweeps
// (d in buffer par).
__global__ __launch_bounds__(kBlock) void k_ir_update(const RunDev *rds, const int *start,
                                                      int nruns, int par)
{
    RUN_PROLOGUE();
    if (st->solve_done) return;
    int k = local * kBlock + threadIdx.x;
    if (k >= R.nnodes) return;
    int j = k / R.ncx;
    if (j < 1 || j >= R.ncy) return;
    R.phi64[k] += (double)R.d32[par & 1][k];
}

// FP64 residual max over interior nodes of |0.25 (b + sum of neighbours) - phi|,
// the same measure as the reference's max |dphi| (src/mesh.cu:354).  Mode 1
// evaluates it on the FP32 phi in buffer par; mode 2 on the FP64 master and
// also stores the correction right-hand side r = b + sum - 4 phi (mode 3
// reads the FP64 potential too).  Each thread takes kResNodes nodes, a
// block stride apart, so that few blocks meet at the atomic slots; the
// block's maximum goes to check slot c.
__device__ __forceinline__ void residual_block(const RunDev &R, RunState *st, double *sh,
                                               int local, int c, int mode, int par)
{
    const int ncx = R.ncx;
    const CUPIC_GLOBAL double *p64 = gp(R.phi64);
    const CUPIC_GLOBAL float *p32 = gp(R.phi32[phi_at(st->phi_cur, par)]);
    const CUPIC_GLOBAL double *b64 = gp(R.b64);
    double rmax = 0.0;
#pragma unroll
    for (int u = 0; u < kResNodes; u++) {
        const int k = (local * kResNodes + u) * kBlock + threadIdx.x;
        if (k >= R.nnodes) break;
        const int j = k / ncx;
        const int i = k - j * ncx;
        if (j >= 1 && j < R.ncy) {
            const int kw = (i == 0) ? k + ncx - 1 : k - 1;
            const int ke = (i == ncx - 1) ? k - ncx + 1 : k + 1;
            double pc_, pw, pe, pn, ps;
            if (mode >= 2) {
                pc_ = p64[k]; pw = p64[kw]; pe = p64[ke]; pn = p64[k + ncx]; ps = p64[k - ncx];
            } else {
                pc_ = p32[k]; pw = p32[kw]; pe = p32[ke]; pn = p32[k + ncx]; ps = p32[k - ncx];
            }
            const double sum = (((b64[k] + pw) + pe) + pn) + ps;
            const double rr = sum - 4.0 * pc_;
            if (mode == 2) gp(R.r32)[k] = (float)rr;
            rmax = fmax(rmax, fabs(0.25 * rr));
        }
    }
    unsigned long long *slot = &R.resid[c * kRedSlots + (local & (kRedSlots - 1))];
    block_max_atomic(rmax, sh, [&](double m) {
        if (m > 0.0) atomicMax(slot, (unsigned long long)__double_as_longlong(m));
    });
}

// end: the run's last block to finish also ends the solve with check c (the
// work of k_solver_end), which saves a launch after the direct solve.
__global__ __launch_bounds__(kBlock) void k_residual(const RunDev *rds, const int *start,
                                                     int nruns, int c, int mode, int par,
                                                     int end)
{
    RUN_PROLOGUE();
    __shared__ double sh[kBlock / kWave];
    const int done = st->solve_done;  // the same for the whole block
    if (done && !end) return;
    if (!done) residual_block(R, st, sh, local, c, mode, par);
    if (end) {
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            const unsigned nb = (unsigned)((R.nnodes + kBlock * kResNodes - 1) / (kBlock * kResNodes));
            if (atomicAdd(&st->solve_ticket, 1u) == nb - 1u) {
                __threadfence();
                st->solve_ticket = 0u;
                solver_end_work(R, st, c, par);
            }
        }
    }
}


__global__ __launch_bounds__(kBlock) void k_solver_end(const RunDev *rds, const int *start,
                                                       int nruns, int last_check, int par)
{
    RUN_PROLOGUE();
    if (threadIdx.x == 0) solver_end_work(R, st, last_check, par);
}

// E = -grad phi (src/mesh.cu:413-468): central differences with periodic x
// on interior rows; one-sided Ey and Ex = 0 on rows 0 and ncy.  Evaluated in
// FP64 from the FP32 (modes 0/1) or FP64 (modes 2/3) potential.
// kResid (a step's direct solve, launched instead of k_residual): the FP64
// residual of the solve, which reads the same neighbours, is reduced into
// check slot 0 here too; the next kernel ends the solve (solver_end_work).
template <bool kResid>
__global__ __launch_bounds__(kBlock) void k_field(const RunDev *rds, const int *start,
                                                  int nruns)
{
    RUN_PROLOGUE();
    const int k = local * kBlock + threadIdx.x;
    double rmax = 0.0;
    if (k < R.nnodes) {
        const int ncx = R.ncx;
        const int j = k / ncx;
        const int i = k - j * ncx;
        const bool m2 = kResid || (st->solve_mode >= 2);
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
            const int kw = (i == 0) ? k + ncx - 1 : k - 1;
            const int ke = (i == ncx - 1) ? k - ncx + 1 : k + 1;
            const double pw = P(kw), pe = P(ke), pn = P(k + ncx), ps = P(k - ncx);
            ex = (pw - pe) / (2.0 * R.ds);
            ey = (ps - pn) / (2.0 * R.ds);
            if (kResid) {
                const double sum = (((R.b64[k] + pw) + pe) + pn) + ps;
                rmax = fabs(0.25 * (sum - 4.0 * P(k)));
            }
        }
        R.E[k] = make_float2((float)ex, (float)ey);
    }
    if (kResid) {
        __shared__ double sh[kBlock / kWave];
        unsigned long long *slot = &R.resid[local & (kRedSlots - 1)];
        block_max_atomic(rmax, sh, [&](double m) {
            if (m > 0.0) atomicMax(slot, (unsigned long long)__double_as_longlong(m));
        });
    }
}

// ---------------------------------------------------------------- launchers

#define LAUNCH_K(kern, L, ...)                                                         \
    do {                                                                               \
        if ((L).nblocks > 0)                                                           \
            hipLaunchKernelGGL(kern, dim3((L).nblocks), dim3(kBlock), 0, (L).stream,   \
                               (L).rds, (L).start, (L).nruns, ##__VA_ARGS__);          \
    } while (0)

void launch_phi_linear(const Launch &L) { LAUNCH_K(k_phi_linear, L); }
void launch_rho_convert(const Launch &L, int begin) { LAUNCH_K(k_rho_convert, L, begin); }
void launch_solver_begin(const Launch &L, int use_args, int mode, double tol, int min_sweeps,
                         int max_sweeps, int check_every)
{
    LAUNCH_K(k_solver_begin, L, use_args, mode, tol, min_sweeps, max_sweeps, check_every);
}
void launch_jacobi_sweep(const Launch &L, int s) { LAUNCH_K(k_jacobi_sweep, L, s); }
void launch_jacobi_block(const Launch &L, int halo, int s0, int nsw, int mode, int first,
                         int par, int cslot, int cat, int cpar, int rslot, int rat)
{
    if (L.nblocks <= 0) return;
#define JAC_LAUNCH(HH, RR)                                                                    \
    hipLaunchKernelGGL((k_jacobi_block<HH, RR>), dim3(L.nblocks), dim3(kJacThreads), 0, L.stream, \
                       L.rds, L.start, L.nruns, s0, nsw, mode, first, par, cslot, cat, cpar,    \
                       rslot, rat)
#define JAC_HALO(HH)                                                                          \
    do {                                                                                      \
        if (rk == 2) JAC_LAUNCH(HH, 2);                                                       \
        else if (rk == 1) JAC_LAUNCH(HH, 1);                                                  \
        else JAC_LAUNCH(HH, 0);                                                               \
    } while (0)
    static_assert(kJacHSolo[0] == 28 && kJacHSolo[1] == 24 && kJacHSolo[2] == 20,
                  "one instance per halo");
    const int rk = rslot < 0 ? 0 : (rat > 0 && rat < nsw) ? 2 : 1;
    switch (halo) {
    case kJacHNarrow: JAC_HALO(kJacHNarrow); break;
    case kJacH: JAC_HALO(kJacH); break;
    case 20: JAC_HALO(20); break;
    case 24: JAC_HALO(24); break;
    case 28: JAC_HALO(28); break;
    default: break;  // no other halo is chosen
    }
#undef JAC_HALO
#undef JAC_LAUNCH
}
void launch_ir_update(const Launch &L, int par) { LAUNCH_K(k_ir_update, L, par); }
void launch_residual(const Launch &L, int c, int mode, int par, int end)
{
    LAUNCH_K(k_residual, L, c, mode, par, end);
}
void launch_solver_end(const Launch &L, int last_check, int par)
{
    LAUNCH_K(k_solver_end, L, last_check, par);
}
void launch_field(const Launch &L, int resid_end)
{
    if (resid_end)
        LAUNCH_K(k_field<true>, L);
    else
        LAUNCH_K(k_field<false>, L);
}

}  // namespace cupic
