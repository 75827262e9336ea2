// k_cc.hip -- boundary conditions and the cell-key sort bookkeeping:
// injection clock and flux-Maxwellian injection at y = Ly, ranks of the far
// list (particles that moved more than one cell, and injected ones), the
// per-cell scan that turns direction counts into new offsets and slot
// bases, the far-list scatter and the end-of-step bookkeeping (buffer flip,
// probe tally).
#include "cupic_clock.h"
#include "cupic_launch.h"

namespace cupic {


constexpr double kPi = 3.1415926535897932;  // src/stdh.h:26

// Injection clock (src/cc.cu:35-36, 147, 179, 247): tin starts at the first
// cc call as t + dtin; this step inserts in = 1 + int((t+dt-tin)/dtin)
// particles when tin < t+dt, advancing tin by dtin once per particle.  The
// counts and times must match the reference's FP64 host loop bit for bit;
// the chain of adds is evaluated as the exact arithmetic progressions of
// cupic_clock.h (one lane per species cuts it into segments, the block
// writes the times).  It reads only the clock, so a step can run it on a
// side stream while the field is solved.
constexpr int kClockSegs = 64;

__global__ __launch_bounds__(kBlock) void k_clock(const RunDev *rds, const int *start,
                                                  int nruns, int use_host_t, double t_host)
{
    RUN_PROLOGUE();
    __shared__ ClockSeg segs[2][kClockSegs];
    __shared__ int nseg[2];
    if (threadIdx.x == 0) {
        if (use_host_t) st->t = t_host;
        if (!st->tin_started) {
            st->tin[0] = st->t + R.sp[0].dtin;
            st->tin[1] = st->t + R.sp[1].dtin;
            st->tin_started = 1;
        }
    }
    __syncthreads();
    // lane 0 of waves 0 and 1: species 0 and 1
    if ((threadIdx.x & (kWave - 1)) == 0 && threadIdx.x < 2 * kWave) {
        const int s = threadIdx.x / kWave;
        const double t = st->t;
        const double fpt = t + R.dt;
        const double tin = st->tin[s];
        const double dtin = R.sp[s].dtin;
        int n = 0;
        if (tin < fpt) n = 1 + (int)((fpt - tin) / dtin);
        const int nst = min(n, R.max_inject);  // times kept for the injection
        double tend;
        int ns = clock_segments(tin, dtin, nst, segs[s], kClockSegs, &tend);
        if (ns < 0) {
            // more segments than the list holds (ties at many binades, or a
            // clock with a long way to go): plain adds for the rest
            CUPIC_GLOBAL double *seq = gp(R.inj_tin[s]);
            const ClockSeg &lastseg = segs[s][kClockSegs - 1];
            for (int k = lastseg.k0 + lastseg.len; k < nst; k++) {
                seq[k] = tend;
                tend += dtin;
            }
            ns = kClockSegs;
        }
        for (int k = nst; k < n; k++) tend += dtin;
        nseg[s] = ns;
        st->tin[s] = tend;
        int ne = n;
        unsigned int err = 0;
        if (ne > R.max_inject) { ne = R.max_inject; err |= kErrInjectCap; }
        if (st->N[s] + ne > R.cap) { ne = max(0, R.cap - st->N[s]); err |= kErrCapacity; }
        if (err) atomicOr(&st->err, err);
        st->n_in[s] = ne;
        st->ninj_first[s] = st->ninj[s];
        st->ninj[s] += (unsigned long long)ne;
    }
    __syncthreads();
    for (int s = 0; s < 2; s++) {
        CUPIC_GLOBAL double *seq = gp(R.inj_tin[s]);
        for (int q = 0; q < nseg[s]; q++) {
            const ClockSeg g = segs[s][q];
            for (int j = threadIdx.x; j < g.len; j += kBlock) seq[g.k0 + j] = clock_value(g, j);
        }
    }
}

// One injected particle (src/cc.cu:214-247), evaluated in FP64: x = U Lx,
// y = Ly, vx ~ N(0, sigma), vy = -Rayleigh(sigma); E interpolated from the
// top node row (disty = 1); partial push x += (t+dt-tin) v and
// v += q (t+dt/2-tin) E/m (Q7; the committed sign for both species is kept
// behind a quirk flag); wrapped in x, rounded to FP32, wrapped again; a y
// rounded up to Ly becomes the largest float below Ly.
__global__ __launch_bounds__(kBlock) void k_inject(const RunDev *rds, const int *start,
                                                   int nruns)
{
    RUN_PROLOGUE();
    int bps = (R.max_inject + kBlock - 1) / kBlock;
    int s = local / bps;
    int k = (local - s * bps) * kBlock + threadIdx.x;
    if (s > 1 || k >= st->n_in[s]) return;
    double tin = R.inj_tin[s][k];
    unsigned long long r[4];
    draw4(R, s, 1, st->ninj_first[s] + (unsigned long long)k, r);
    const SpecConst &S = R.sp[s];
    double t = st->t;
    double fpt = t + R.dt, fvt = t + 0.5 * R.dt;
    double x = u01_open(r[0]) * R.Lx;
    double y = R.Ly;
    double g = sqrt(-2.0 * log(u01_open(r[1]))) * cos(2.0 * kPi * u01_open(r[2]));
    double vx = S.sigma * g;
    double vy = -(S.sigma * sqrt(-2.0 * log(u01_open(r[3]))));
    int ic = (int)(x / R.ds);
    if (ic > R.ncx - 1) ic = R.ncx - 1;
    int jc = R.ncy - 1;
    int i1 = ic + 1 == R.ncx ? 0 : ic + 1;
    double distx = fabs((double)ic * R.ds - x) / R.ds;
    double disty = 1.0;
    float2 e00 = R.E[jc * R.ncx + ic], e10 = R.E[jc * R.ncx + i1];
    float2 e01 = R.E[(jc + 1) * R.ncx + ic], e11 = R.E[(jc + 1) * R.ncx + i1];
    double epx = (double)e00.x * (1.0 - distx) * (1.0 - disty);
    epx += (double)e10.x * distx * (1.0 - disty);
    epx += (double)e01.x * (1.0 - distx) * disty;
    epx += (double)e11.x * distx * disty;
    double epy = (double)e00.y * (1.0 - distx) * (1.0 - disty);
    epy += (double)e10.y * distx * (1.0 - disty);
    epy += (double)e01.y * (1.0 - distx) * disty;
    epy += (double)e11.y * distx * disty;
    x += (fpt - tin) * vx;
    y += (fpt - tin) * vy;
    if ((R.quirks & 1u) || S.q < 0.0) {
        vx -= (fvt - tin) * epx / S.m;
        vy -= (fvt - tin) * epy / S.m;
    } else {
        vx += (fvt - tin) * epx / S.m;
        vy += (fvt - tin) * epy / S.m;
    }
    if (x < 0.0) x += R.Lx;
    if (x >= R.Lx) x -= R.Lx;
    float xf = (float)x;
    if (xf < 0.0f) xf = xf + R.Lx_f;
    if (xf >= R.Lx_f) xf = xf - R.Lx_f;
    FarRec rec;
    rec.x = xf;
    rec.y = (float)y;
    // y < Ly in FP64 (vy < 0), but rounding can reach Ly: keep it inside
    if (rec.y >= R.Ly_f) rec.y = __int_as_float(__float_as_int(R.Ly_f) - 1);
    rec.vx = (float)vx;
    rec.vy = (float)vy;
    float fx, fy;
    int nic, njc;
    cell_of(R, rec.x, rec.y, fx, fy, nic, njc);
    rec.key = njc * R.ncx + nic;
    rec.idx = st->N[s] + k;
    rec.rank = 0;
    rec.pad = 0;
    R.far[s][far_mov_cap(R) + k] = rec;  // k < n_in <= max_inject
    atomicAdd(&R.far_cnt[s][rec.key], 1);
}

// Count of particles arriving in new cell (i, j) from neighbour slot sl
// (source offset (sl%3-1, sl/3-1), periodic in x): the source cell's count
// for the direction that leads here.  cnt is direction-major.
__device__ __forceinline__ int src_count(const CUPIC_GLOBAL int *cnt, int ncx, int ncy,
                                         int ncells, int i, int j, int sl)
{
    int dj = sl / 3 - 1, di = sl - 3 * (sl / 3) - 1;
    int jj = j + dj;
    if (jj < 0 || jj >= ncy) return 0;
    int ii = i + di;
    if (ii < 0) ii += ncx;
    if (ii >= ncx) ii -= ncx;
    return cnt[(size_t)(8 - sl) * ncells + jj * ncx + ii];
}

// Block-wide sum of one value per thread (thread 0 gets it).
__device__ __forceinline__ int block_sum(int v, int *sh)
{
    int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int q = 0; q < kBlock / kWave; q++) t += sh[q];
    return t;
}

// Per-block sums of the new state's cell totals (near arrivals plus the far
// list).  One thread per cell (kScanCells per block, interleaved by kBlock
// so that every access is coalesced).
constexpr int kScanPer = kScanCells / kBlock;  // cells per thread

__global__ __launch_bounds__(kBlock) void k_scan_a(const RunDev *rds, const int *start,
                                                   int nruns)
{
    RUN_PROLOGUE();
    __shared__ int sh[kBlock / kWave];
    const int ncells = R.ncells, ncx = R.ncx, ncy = R.ncy;
    int nblk = (ncells + kScanCells - 1) / kScanCells;
    int s = local / nblk;
    int blk = local - s * nblk;
    const CUPIC_GLOBAL int *cnt = gp(R.cnt[s]);
    const CUPIC_GLOBAL int *farc = gp(R.far_cnt[s]);
    // every load first, then the sums
    int a[kScanPer][10];
#pragma unroll
    for (int u = 0; u < kScanPer; u++) {
        const int k = blk * kScanCells + u * kBlock + threadIdx.x;
#pragma unroll
        for (int sl = 0; sl < 10; sl++) a[u][sl] = 0;
        if (k < ncells) {
            const int j = k / ncx, i = k - j * ncx;
            a[u][9] = farc[k];
#pragma unroll
            for (int sl = 0; sl < 9; sl++) a[u][sl] = src_count(cnt, ncx, ncy, ncells, i, j, sl);
        }
    }
    int sum = 0;
#pragma unroll
    for (int u = 0; u < kScanPer; u++)
#pragma unroll
        for (int sl = 0; sl < 10; sl++) sum += a[u][sl];
    int total = block_sum(sum, sh);
    if (threadIdx.x == 0) R.blk[s][blk] = total;
}

// New offsets and slot bases (slot-major): base[sl][k] = off[k] + arrivals
// from slots before sl; slot 9 (far list) last.  Clears the far counters.
// Each block starts from the sum of the block totals before it (at most a
// few hundred, summed by the block itself); the last one also sets N_new.
// Every load is issued before the first barrier: the block totals, and for
// each of the thread's cells its far count and nine arrival counts, from
// which the cell's total is summed again here.  The thread's kScanPer cells
// are scanned together: one wave scan each, then one exchange of the wave
// totals.
__global__ __launch_bounds__(kBlock) void k_scan_c(const RunDev *rds, const int *start,
                                                   int nruns)
{
    RUN_PROLOGUE();
    __shared__ int sh[kBlock / kWave];
    __shared__ int wt[kScanPer][kBlock / kWave];
    const int ncells = R.ncells, ncx = R.ncx, ncy = R.ncy;
    int nblk = (ncells + kScanCells - 1) / kScanCells;
    int s = local / nblk;
    int blk = local - s * nblk;
    const CUPIC_GLOBAL int *cnt = gp(R.cnt[s]);
    CUPIC_GLOBAL int *farc = gp(R.far_cnt[s]);
    // the parity as a scalar, so that the output pointer is a scalar load
    // and no store waits for it
    const int cur = __builtin_amdgcn_readfirstlane(gp(st)->cur);
    CUPIC_GLOBAL int *noff = gp(cur ? R.off[s][0] : R.off[s][1]);
    // block totals before this one (the first 4 kBlock in flight with the
    // counts; more only past 4 M cells per species)
    const CUPIC_GLOBAL int *bt = gp(R.blk[s]);
    int v[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int q = m * kBlock + threadIdx.x;
        v[m] = q < blk ? bt[q] : 0;
    }
    int a[kScanPer][9], t[kScanPer];
#pragma unroll
    for (int u = 0; u < kScanPer; u++) {
        const int k = blk * kScanCells + u * kBlock + threadIdx.x;
        t[u] = 0;
#pragma unroll
        for (int sl = 0; sl < 9; sl++) a[u][sl] = 0;
        if (k < ncells) {
            const int j = k / ncx, i = k - j * ncx;
            t[u] = farc[k];
#pragma unroll
            for (int sl = 0; sl < 9; sl++) a[u][sl] = src_count(cnt, ncx, ncy, ncells, i, j, sl);
        }
    }
    int pre = (v[0] + v[1]) + (v[2] + v[3]);
    for (int q = 4 * kBlock + threadIdx.x; q < blk; q += kBlock) pre += bt[q];
    pre = block_sum(pre, sh);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
    int x[kScanPer];
#pragma unroll
    for (int u = 0; u < kScanPer; u++) {
#pragma unroll
        for (int sl = 0; sl < 9; sl++) t[u] += a[u][sl];
        x[u] = t[u];
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(x[u], o, kWave);
            if (lane >= o) x[u] += y;
        }
        if (lane == kWave - 1) wt[u][w] = x[u];
    }
    __syncthreads();
    // exclusive prefix of cell (u, thread): the cells of earlier u, then
    // the earlier waves of this u, then this wave's lanes before it
    int base_u = pre;
    CUPIC_GLOBAL int *base = gp(R.base[s]);
    CUPIC_GLOBAL int *farcur = gp(R.far_cur[s]);
#pragma unroll
    for (int u = 0; u < kScanPer; u++) {
        int before = 0, all = 0;
#pragma unroll
        for (int q = 0; q < kBlock / kWave; q++) {
            const int v = wt[u][q];
            before += q < w ? v : 0;
            all += v;
        }
        const int k = blk * kScanCells + u * kBlock + threadIdx.x;
        if (k < ncells) {
            int b = base_u + before + x[u] - t[u];
            noff[k] = b;
#pragma unroll
            for (int sl = 0; sl < 9; sl++) {
                base[(size_t)sl * ncells + k] = b;
                b += a[u][sl];
            }
            base[(size_t)9 * ncells + k] = b;
            farc[k] = 0;
            farcur[k] = 0;
        }
        base_u += all;
    }
    if (blk == nblk - 1 && threadIdx.x == 0) {
        st->N_new[s] = base_u;
        noff[ncells] = base_u;
    }
}

// Stable rank of a far entry within its cell: entries of the same cell with
// a smaller input index (a handful per cell).
__global__ __launch_bounds__(kBlock) void k_far_rank(const RunDev *rds, const int *start,
                                                     int nruns)
{
    RUN_PROLOGUE();
    const int bps = far_blocks(R);
    const int s = local / bps;
    if (s > 1) return;
    const int F = far_len(R, st, s);
    const int cur = __builtin_amdgcn_readfirstlane(gp(st)->cur);
    const int *const noff = cur ? R.off[s][0] : R.off[s][1];
    const float4 *const slots = cur ? R.part[s][0] : R.part[s][1];
    for (int v = (local - s * bps) * kBlock + threadIdx.x; v < F; v += bps * kBlock) {
        const int e = far_entry(R, st, s, v);
        const FarRec *far = R.far[s];
        int key = far[e].key, idx = far[e].idx;
        int b = R.base[s][9 * R.ncells + key];
        int n = noff[key + 1] - b;
        // four slots per round trip: their entry indices, then the entries'
        // input indices (an index outside the list reads entry 0 and counts
        // for nothing)
        int rank = 0;
        for (int u = 0; u < n; u += 4) {
            int q[4];
#pragma unroll
            for (int k = 0; k < 4; k++)
                q[k] = u + k < n ? reinterpret_cast<const int *>(slots + (b + u + k))[0] : -1;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const bool ok = (unsigned)q[k] < (unsigned)R.far_cap;
                const int other = far[ok ? q[k] : 0].idx;
                rank += (ok && other < idx) ? 1 : 0;
            }
        }
        R.far[s][e].rank = rank;
    }
}

// End of a step: flip the particle buffers, commit the counts and the probe
// tally (probe charge accumulated in FP64: ions +1, electrons -1).
__device__ void finish_step(const RunDev &R, RunState *st, int advance_t)
{
    double dq = 0.0;
    for (int s = 0; s < 2; s++) {
        st->last_probe[s] = st->probe[s];
        st->last_top[s] = st->top[s];
        st->last_injected[s] = (unsigned long long)st->n_in[s];
        st->total_probe[s] += st->probe[s];
        st->total_top[s] += st->top[s];
        st->total_injected[s] += (unsigned long long)st->n_in[s];
        st->total_undecided[s] += st->undecided[s];
        st->total_far[s] += (unsigned long long)st->far_n[s];
        st->undecided[s] = 0;
        st->particle_steps += (unsigned long long)st->N[s];
        dq += R.sp[s].q * (double)st->probe[s];
        st->probe[s] = 0;
        st->top[s] = 0;
        st->far_n[s] = 0;
        st->n_in[s] = 0;
        st->N[s] = st->N_new[s];
    }
    st->probe_charge_total += dq;
    st->cur ^= 1;
    st->step += 1;
    if (advance_t) st->t += R.dt;
}

// Far-list entries to their sorted positions, with their charge (global
// int64 atomics; the list is short).  The run's last block to take a ticket
// then ends the step (finish_step).  No fence is needed: the step's end
// reads only state that earlier kernels wrote, and every block has used the
// state it changes (the list length, the parity) before taking its ticket.
__global__ __launch_bounds__(kBlock) void k_far_scatter(const RunDev *rds, const int *start,
                                                        int nruns, int advance_t)
{
    RUN_PROLOGUE();
    const int bps = far_blocks(R);
    const int s = local / bps;  // launched with 2 bps blocks per run: s is 0 or 1
    const int F = far_len(R, st, s);
    const int o = __builtin_amdgcn_readfirstlane(gp(st)->cur) ^ 1;
    for (int v = (local - s * bps) * kBlock + threadIdx.x; v < F; v += bps * kBlock) {
        FarRec rec = R.far[s][far_entry(R, st, s, v)];
        int dest = R.base[s][9 * R.ncells + rec.key] + rec.rank;
        if ((unsigned)dest >= (unsigned)R.cap) {
            atomicOr(&st->err, kErrSortInvariant);
            continue;
        }
        st4(R.part[s][o], dest, v4f{rec.x, rec.y, rec.vx, rec.vy});
        R.hint[s][o][dest] = 0;  // placed without a hint: pushed in full next step
        float fx, fy;
        int ic, jc;
        cell_of(R, rec.x, rec.y, fx, fy, ic, jc);
        unsigned wx, wy;
        fixed_weights(fx, fy, wx, wy);
        long long sgn = s == 0 ? -1 : 1;
        long long gx = 65536 - (long long)wx, gy = 65536 - (long long)wy;
        int i1 = ic + 1 == R.ncx ? 0 : ic + 1;
        atomicAdd(&R.acc[jc * R.ncx + ic], (unsigned long long)(sgn * gx * gy));
        atomicAdd(&R.acc[jc * R.ncx + i1], (unsigned long long)(sgn * (long long)wx * gy));
        atomicAdd(&R.acc[(jc + 1) * R.ncx + ic], (unsigned long long)(sgn * gx * (long long)wy));
        atomicAdd(&R.acc[(jc + 1) * R.ncx + i1],
                  (unsigned long long)(sgn * (long long)wx * (long long)wy));
    }
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(&st->fin_ticket, 1u) == (unsigned)(2 * bps - 1)) {
        st->fin_ticket = 0u;
        finish_step(R, st, advance_t);
    }
}

__global__ __launch_bounds__(kBlock) void k_init_finish(const RunDev *rds, const int *start,
                                                        int nruns)
{
    RUN_PROLOGUE();
    if (threadIdx.x != 0) return;
    st->cur = 0;
    for (int s = 0; s < 2; s++) {
        st->N[s] = R.n_init;
        st->N_new[s] = R.n_init;
        st->n_in[s] = 0;
        st->far_n[s] = 0;
        st->probe[s] = st->top[s] = 0;
        st->last_probe[s] = st->last_top[s] = st->last_injected[s] = 0;
        st->total_probe[s] = st->total_top[s] = st->total_injected[s] = 0;
        st->undecided[s] = st->total_undecided[s] = st->total_far[s] = 0;
        st->ninj[s] = 0;
        st->ninj_first[s] = 0;
        st->tin[s] = 0.0;
    }
    st->tin_started = 0;
    st->probe_charge_total = 0.0;
    st->particle_steps = 0;
    st->sweeps_total = 0;
    st->solves = 0;
    st->step = 0;
    st->t = 0.0;
    st->err = 0;
    st->fin_ticket = 0u;
    st->solve_ticket = 0u;
}

// ---------------------------------------------------------------- launchers

#define LAUNCH_K(kern, L, ...)                                                         \
    do {                                                                               \
        if ((L).nblocks > 0)                                                           \
            hipLaunchKernelGGL(kern, dim3((L).nblocks), dim3(kBlock), 0, (L).stream,   \
                               (L).rds, (L).start, (L).nruns, ##__VA_ARGS__);          \
    } while (0)

void launch_clock(const Launch &L, int use_host_t, double t_host)
{
    LAUNCH_K(k_clock, L, use_host_t, t_host);
}
void launch_inject(const Launch &L) { LAUNCH_K(k_inject, L); }
void launch_far_rank(const Launch &L) { LAUNCH_K(k_far_rank, L); }
void launch_scan_a(const Launch &L) { LAUNCH_K(k_scan_a, L); }
void launch_scan_c(const Launch &L) { LAUNCH_K(k_scan_c, L); }
void launch_far_scatter(const Launch &L, int advance_t) { LAUNCH_K(k_far_scatter, L, advance_t); }
void launch_init_finish(const Launch &L) { LAUNCH_K(k_init_finish, L); }

}  // namespace cupic
