// cupic_api.cpp -- host runtime behind the C-ABI (include/cupic.h).
//
// Owns each run's device buffers, sequences the kernels of the PIC cycle on
// the run's stream, batches runs behind one set of launches (optionally one
// HIP graph per step) and keeps the error convention of returning status
// codes instead of exiting (the reference prints and exits,
// src/cuda.cu:15-54).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cupic.h"
#include "cupic_launch.h"

using namespace cupic;

namespace {

thread_local std::string g_err;

enum TimeClass : int {
    kTDeposit = 0,   // charge conversion (+ standalone deposit)
    kTSolve,         // whole Poisson solve
    kTField,
    kTPushRank,      // pass 1
    kTInject,        // clock + injection + far ranks
    kTScan,
    kTScatter,       // pass 2 (+ fused deposition)
    kTFinish,        // far scatter + finish
    kTClasses
};
const char *kTimeNames = "deposit,solve,field,push_rank,inject,scan,scatter,finish";

int set_err(std::string *dst, int code, const std::string &msg)
{
    if (dst) *dst = msg;
    g_err = msg;
    return code;
}

}  // namespace

#define HIPTRY(dst, expr)                                                               \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_err(dst, CUPIC_ERR_HIP,                                          \
                           std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

struct cupic_run {
    cupic_params p{};
    RunDev h{};
    RunDev *d_rd = nullptr;
    int *d_plan = nullptr;
    int nblk[kShapeCount] = {};
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_field = nullptr;
    std::vector<void *> allocs;
    std::string err;
    int moved = 0;          // particle_mover done, cc pending
    bool acc_valid = false; // acc holds the charge of the current particles
    bool clock_forked = false;  // the coming cc's clock already runs on the side stream
    bool initialised = false;
    double t_next = 0.0;
    int64_t footprint = 0;
    bool timing = false;
    unsigned t_mask = (1u << kTClasses) - 1;  // classes timed while timing is on
    struct Ev { int cls; hipEvent_t a, b; };
    std::vector<Ev> evs;
    std::vector<hipEvent_t> pool;
    double t_ms[kTClasses] = {};
    int64_t t_n[kTClasses] = {};
};

struct cupic_batch {
    std::vector<cupic_run *> runs;
    RunDev *d_rds = nullptr;
    int *d_plan = nullptr;   // kShapeCount rows: block prefixes, then the block map
    int64_t plan_off[kShapeCount] = {};   // row s of d_plan
    int nblk[kShapeCount] = {};
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_field = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = false;
    int mode = 0, check = 50, max_sweeps = 0;
    std::string err;
};

namespace cupic {

// The direct solve's method for a run: its FFT instance (fft_class), or 0
// for the dense products.
int direct_class(const cupic_params &p)
{
    return p.direct_method == CUPIC_DIRECT_DENSE
               ? 0
               : fft_class(p.ncx, p.ncy, p.direct_method == CUPIC_DIRECT_FFT);
}

// Direct solve: output tiles of one product over the m = ncy - 1 interior
// rows and ncx columns, and the k slices that bring a product near
// kGSlices workgroups (never more slices than k tiles).
int dft_tiles(int ncx, int m) { return ((m + kGM - 1) / kGM) * ((ncx + kGN - 1) / kGN); }
int dft_slices(int ncx, int m, int Kd)
{
    const int want = (kGSlices + dft_tiles(ncx, m) - 1) / dft_tiles(ncx, m);
    return std::max(1, std::min({want, kGMaxSlices, (Kd + kGK - 1) / kGK}));
}

int shape_blocks(const RunDev &h, Shape s)
{
    switch (s) {
    case kShapeNodes: return (h.nnodes + kBlock - 1) / kBlock;
    case kShapeResid: return (h.nnodes + kBlock * kResNodes - 1) / (kBlock * kResNodes);
    case kShapeTasks: return (2 * h.ncy * h.tpr + 3) / 4;
    case kShapeScatter: return 2 * far_blocks(h) + (2 * h.ncy * h.tpr + 3) / 4;
    case kShapeCountTasks: return (2 * h.ncy * h.tpr1 + 3) / 4;
    case kShapeCellBlk: return 2 * ((h.ncells + kScanCells - 1) / kScanCells);
    case kShapeRunSp: return 2;
    case kShapeRun: return 1;
    case kShapeFar: return 2 * far_blocks(h);
    case kShapeInject: return 2 * ((h.max_inject + kBlock - 1) / kBlock);
    case kShapeRows: return 2 * h.ncy;
    case kShapeInit: return 2 * ((h.n_init + kBlock - 1) / kBlock);
    case kShapeJacBlk: return jac_blocks(h.ncx, h.ncy, kJacH);
    case kShapeJacBlkN: return jac_blocks(h.ncx, h.ncy, kJacHNarrow);
    case kShapeDftInit: {
        if (!h.dft_d) return 0;
        const int64_t m = h.ncy - 1;
        if (h.dft_fft) {  // twiddles and chirps of both lengths, then the eigenvalue parts
            const FftTabs t = fft_tabs(h.ncx, h.ncy);
            return (int)(((int64_t)t.bh_x + h.ncx + m + kBlock - 1) / kBlock);
        }
        return (int)(((int64_t)h.ncx * h.ncx + kBlock - 1) / kBlock + (m * m + kBlock - 1) / kBlock +
                     (h.ncx + m + kBlock - 1) / kBlock);
    }
    case kShapeDftGemmX:
        if (!h.dft_f) return 0;
        return dft_tiles(h.ncx, h.ncy - 1) * h.nsx;
    case kShapeDftGemmY:
        if (!h.dft_f) return 0;
        return dft_tiles(h.ncx, h.ncy - 1) * h.nsy;
    case kShapeDftReduce:
        if (!h.dft_f) return 0;
        return (int)(((int64_t)(h.ncy - 1) * h.ncx + kBlock - 1) / kBlock);
    case kShapeFftRowsS: return h.dft_fft == 1 ? h.ncy / 2 : 0;  // ceil(m / 2), m = ncy - 1
    case kShapeFftColsS: return h.dft_fft == 1 ? (h.ncx + 1) / 2 : 0;
    case kShapeFftRowsM: return h.dft_fft == 2 ? h.ncy / 2 : 0;
    case kShapeFftColsM: return h.dft_fft == 2 ? (h.ncx + 1) / 2 : 0;
    case kShapeFftRowsL: return h.dft_fft == 3 ? h.ncy / 2 : 0;
    case kShapeFftColsL: return h.dft_fft == 3 ? (h.ncx + 1) / 2 : 0;
    case kShapeFftTabS: return h.dft_fft == 1 ? 2 : 0;
    case kShapeFftTabM: return h.dft_fft == 2 ? 2 : 0;
    case kShapeFftTabL: return h.dft_fft == 3 ? 2 : 0;
    default: return 0;
    }
}

}  // namespace cupic

// ---------------------------------------------------------------- kernel sequences

namespace {

// A sequence target: a single run or a batch.
struct Target {
    const RunDev *rds;
    const int *plan;      // kShapeCount rows: row s at off[s], or at 2 s for a single run
                          // (block prefixes 0, nblk[s])
    const int *nblk;
    int nruns;
    hipStream_t stream;
    cupic_run *timed;     // non-null: record timing events for this run
    hipStream_t side = nullptr;   // injection clock runs here when set
    hipEvent_t fork = nullptr, join = nullptr, field = nullptr;
    bool seg_stayers = false;     // scatter instance for low particles per cell
    const int64_t *off = nullptr;  // batch plan rows (null: a single run's)
    int jac_h = 0;                // a single run's blocked-Jacobi halo (0: by block count)
    int jac_nb = 0;               // and its workgroups
    Launch L(Shape s) const
    {
        return Launch{rds, plan + (off ? off[s] : 2 * (int64_t)s), nruns, nblk[s], stream};
    }
    Launch Lside(Shape s) const
    {
        return Launch{rds, plan + (off ? off[s] : 2 * (int64_t)s), nruns, nblk[s], side};
    }
};

Target target_of(cupic_run *r)
{
    Target T{r->d_rd, r->d_plan, r->nblk, 1, r->stream, r->timing ? r : nullptr};
    T.side = r->side;
    T.fork = r->ev_fork;
    T.join = r->ev_join;
    T.field = r->ev_field;
    T.seg_stayers = (int64_t)r->p.per_row < (int64_t)kSegStayersPpc * r->p.ncx;
    T.jac_h = jac_solo_halo(r->h.ncx, r->h.ncy);
    T.jac_nb = jac_blocks(r->h.ncx, r->h.ncy, T.jac_h);
    return T;
}

// Starts the injection clock of the coming cc on the side stream.
// A step's injection on the side stream: the clock at the step's start
// (seq_clock_fork), the injection once the field is ready (seq_inject_fork,
// it reads E), beside the counting pass, whose part of the far list and
// whose counts it does not touch; seq_cc with clock_forked set waits for it.
void seq_clock_fork(const Target &T, int use_host_t, double t)
{
    (void)hipEventRecord(T.fork, T.stream);
    (void)hipStreamWaitEvent(T.side, T.fork, 0);
    launch_clock(T.Lside(kShapeRun), use_host_t, t);
}

void seq_inject_fork(const Target &T)
{
    (void)hipEventRecord(T.field, T.stream);
    (void)hipStreamWaitEvent(T.side, T.field, 0);
    launch_inject(T.Lside(kShapeInject));
    (void)hipEventRecord(T.join, T.side);
}

struct TimeScope {
    cupic_run *r;
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    TimeScope(const Target &T, int c)
        : r(T.timed && ((T.timed->t_mask >> c) & 1u) ? T.timed : nullptr), cls(c)
    {
        if (!r) return;
        a = take(); b = take();
        (void)hipEventRecord(a, r->stream);
    }
    ~TimeScope()
    {
        if (!r) return;
        (void)hipEventRecord(b, r->stream);
        r->evs.push_back({cls, a, b});
    }
    hipEvent_t take()
    {
        if (!r->pool.empty()) { hipEvent_t e = r->pool.back(); r->pool.pop_back(); return e; }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
};

// begun: the solve was started by the charge conversion (launch_rho_convert
// with begin set, use_args 0), so no k_solver_begin launch.  field_ends: a
// direct solve leaves its residual and its end to the field kernel that
// follows (seq_field with resid_end set).
void seq_solve(const Target &T, int use_args, int mode, double tol, int min_sweeps,
               int max_sweeps, int check, bool begun = false, bool field_ends = false)
{
    TimeScope ts(T, kTSolve);
    if (!begun)
        launch_solver_begin(T.L(kShapeRun), use_args, mode, tol, min_sweeps, max_sweeps, check);
    Launch N = T.L(kShapeNodes);
    if (mode == 0) {
        for (int s = 0; s < max_sweeps; s++) launch_jacobi_sweep(N, s);
        launch_solver_end(T.L(kShapeRun), 0, max_sweeps);
        return;
    }
    if (mode == 3) {
        // direct FP64 solve of the same system, then its FP64 residual: by
        // FFT where the lengths allow it, else by dense products (in a
        // batch each run takes its own; launches with no blocks are skipped)
        static const Shape rows[3] = {kShapeFftRowsS, kShapeFftRowsM, kShapeFftRowsL};
        static const Shape cols[3] = {kShapeFftColsS, kShapeFftColsM, kShapeFftColsL};
        for (int step = 0; step < 3; step++)
            for (int c = 0; c < 3; c++) launch_fft(T.L(step == 1 ? cols[c] : rows[c]), step, c + 1);
        for (int step = 0; step < 4; step++) {
            launch_dft_gemm(T.L(step == 0 || step == 3 ? kShapeDftGemmX : kShapeDftGemmY), step);
            launch_dft_reduce(T.L(kShapeDftReduce), step);
        }
        if (!field_ends)
            launch_residual(T.L(kShapeResid), 0, 3, 0, 1);  // its last block ends the solve
        return;
    }
    const bool narrow = T.jac_h > 0 ? T.jac_h == kJacHNarrow
                                    : T.nblk[kShapeJacBlk] >= kJacNarrowMinBlocks;
    const int halo = T.jac_h > 0 ? T.jac_h : narrow ? kJacHNarrow : kJacH;
    Launch B = T.L(narrow ? kShapeJacBlkN : kShapeJacBlk);
    if (T.jac_h > 0) B.nblocks = T.jac_nb;  // a single run: blocks map one to one
    if (check < 1) check = 1;
    const int per = halo - 1;  // sweeps per blocked launch
    int c = 0, s = 0, nlaunch = 0, dpar = 0;
    if (mode == 1) {
        // FP64 residual checks at sweeps 0, check, 2 check, ... and
        // max_sweeps, check k in slot k, each reduced by the launch whose
        // sweeps reach it.  With the wide halo (one run, short launches
        // whose fixed cost dominates) launches of `per` sweeps run straight
        // through the check points: a check inside a launch keeps its iterate
        // in the third phi buffer, the next launch reads the check and ends
        // the solve there if it is met, and the first check is reduced by the
        // first launch from its input.  With the narrow halo (a batch whose
        // launches fill the device many times over) launches end at every
        // check, which keeps the instance with the inside check, and its
        // registers, out of the way.
        const bool through = !narrow;
        const int step = std::min(per, check);  // so at most one check per launch
        const bool fuse0 = through && check > per && max_sweeps > per;
        if (!fuse0) launch_residual(T.L(kShapeResid), 0, mode, 0, 0);
        int cslot = fuse0 ? -1 : 0, cat = 0, cpar = 0;  // the check the next launch reads
        while (s < max_sweeps) {
            int nsw = std::min(step, max_sweeps - s);
            int rslot = -1, rat = 0, at = 0;
            if (fuse0 && nlaunch == 0) {
                rslot = 0;
            } else {
                at = std::min((c + 1) * check, max_sweeps);
                if (!through) nsw = std::min(nsw, at - s);
                if (at <= s + nsw) {
                    rslot = c + 1;
                    rat = at - s;
                }
            }
            launch_jacobi_block(B, halo, s, nsw, 1, 0, nlaunch & 1, cslot, cat, cpar, rslot, rat);
            cslot = -1;
            if (rslot >= 0) {
                c = rslot;
                cslot = rslot;
                cat = at;
                // where the checked iterate is: this launch's input, its
                // output (the next launch's input) or the third buffer
                cpar = rat == 0 ? nlaunch : rat == nsw ? nlaunch + 1 : -1;
            }
            nlaunch++;
            s += nsw;
        }
        launch_solver_end(T.L(kShapeRun), c, nlaunch & 1);
        return;
    }
    for (;;) {
        if (c > 0) launch_ir_update(N, dpar);
        launch_residual(T.L(kShapeResid), c, mode, 0, 0);
        if (s >= max_sweeps) break;
        int k = std::min(check, max_sweeps - s);
        int dl = 0;
        for (int q = 0; q < k;) {
            int nsw = std::min(per, k - q);
            launch_jacobi_block(B, halo, s + q, nsw, 2, q == 0 ? 1 : 0, dl & 1, q == 0 ? c : -1,
                                s + q, dl & 1, -1, 0);
            dl++;
            q += nsw;
        }
        dpar = dl & 1;
        s += k;
        c += 1;
    }
    launch_solver_end(T.L(kShapeRun), c, nlaunch & 1);
}

}  // namespace

namespace cupic {

void seq_solve_one(const RunDev *d_rd, const int *d_plan, const int *nblk, int ncx, int ncy,
                   hipStream_t stream, int mode, double tol, int min_sweeps, int max_sweeps,
                   int check_every)
{
    Target T{d_rd, d_plan, nblk, 1, stream, nullptr};
    T.jac_h = jac_solo_halo(ncx, ncy);
    T.jac_nb = jac_blocks(ncx, ncy, T.jac_h);
    seq_solve(T, 1, mode, tol, min_sweeps, max_sweeps, check_every);
}

int set_global_error(int code, const char *msg) { return set_err(nullptr, code, msg); }

}  // namespace cupic

namespace {

void seq_rho(const Target &T, int begin_solve = 0)
{
    TimeScope ts(T, kTDeposit);
    launch_rho_convert(T.L(kShapeNodes), begin_solve);
}

void seq_field(const Target &T, int resid_end = 0)
{
    TimeScope ts(T, kTField);
    launch_field(T.L(kShapeNodes), resid_end);
}

void seq_mover(const Target &T, int end_solve = 0)
{
    TimeScope ts(T, kTPushRank);
    launch_push_rank(T.L(kShapeCountTasks), end_solve);
}

void seq_cc(const Target &T, int use_host_t, double t, int advance_t, bool clock_forked)
{
    {
        TimeScope ts(T, kTInject);
        if (clock_forked) {  // clock and injection ran on the side stream
            (void)hipStreamWaitEvent(T.stream, T.join, 0);
        } else {
            launch_clock(T.L(kShapeRun), use_host_t, t);
            launch_inject(T.L(kShapeInject));
        }
    }
    {
        TimeScope ts(T, kTScan);
        launch_scan_a(T.L(kShapeCellBlk));
        launch_scan_c(T.L(kShapeCellBlk));
    }
    {
        TimeScope ts(T, kTScatter);
        launch_scatter(T.L(kShapeScatter), T.seg_stayers);  // also places the far list
    }
    {
        TimeScope ts(T, kTFinish);
        launch_far_rank(T.L(kShapeFar));
        launch_far_scatter(T.L(kShapeFar), advance_t);
    }
}

template <class T>
T *dev_alloc(cupic_run *r, size_t count, int *status)
{
    void *p = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) {
        *status = CUPIC_ERR_ALLOC;
        return nullptr;
    }
    (void)hipMemset(p, 0, bytes);
    r->allocs.push_back(p);
    r->footprint += (int64_t)bytes;
    return (T *)p;
}

int check_launch(cupic_run *r, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_err(&r->err, CUPIC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return CUPIC_OK;
}

int read_state(cupic_run *r, RunState *out)
{
    HIPTRY(&r->err, hipMemcpyAsync(out, r->h.st, sizeof(RunState), hipMemcpyDeviceToHost,
                                   r->stream));
    HIPTRY(&r->err, hipStreamSynchronize(r->stream));
    if (out->err) {
        std::string m = "device error flags:";
        if (out->err & kErrFarOverflow) m += " far-list overflow;";
        if (out->err & kErrCapacity) m += " particle capacity exceeded;";
        if (out->err & kErrInjectCap) m += " injection count above max_inject;";
        if (out->err & kErrSortInvariant) m += " sort invariant broken;";
        if (out->err & kErrHint) m += " push hint contradicted by the push;";
        return set_err(&r->err, CUPIC_ERR_CAPACITY, m);
    }
    return CUPIC_OK;
}

int write_state(cupic_run *r, const RunState &s)
{
    HIPTRY(&r->err, hipMemcpyAsync(r->h.st, &s, sizeof(RunState), hipMemcpyHostToDevice,
                                   r->stream));
    HIPTRY(&r->err, hipStreamSynchronize(r->stream));
    return CUPIC_OK;
}

// host copy of the device cell_of (same float operations, no contraction)
inline int host_cell(const RunDev &R, float x, float y)
{
    float tx = x * R.inv_ds;
    float ty = y * R.inv_ds;
    float bx = std::fmin(std::fmax(std::floor(tx), 0.0f), (float)(R.ncx - 1));
    float by = std::fmin(std::fmax(std::floor(ty), 0.0f), (float)(R.ncy - 1));
    return (int)by * R.ncx + (int)bx;
}

int validate(const cupic_params *p, std::string *err)
{
    if (p->ncx < 4 || p->ncy < 2) return set_err(err, CUPIC_ERR_ARG, "need ncx >= 4 and ncy >= 2");
    if (p->ncx > 4096) return set_err(err, CUPIC_ERR_ARG, "ncx above 4096 is not supported");
    if (!(p->ds > 0) || !(p->dt > 0) || !(p->n > 0))
        return set_err(err, CUPIC_ERR_ARG, "ds, dt and n must be positive");
    if (!(p->mi > 0) || !(p->kti > 0))
        return set_err(err, CUPIC_ERR_ARG, "mi and kti must be positive");
    if (p->jacobi_mode < 0 || p->jacobi_mode > 3)
        return set_err(err, CUPIC_ERR_ARG, "jacobi_mode must be 0, 1, 2 or 3");
    if (p->tile_cells < 0 || p->tile_cells > kCmax)
        return set_err(err, CUPIC_ERR_ARG, "tile_cells must be 0 (automatic) or 1..16");
    if (p->direct_method < CUPIC_DIRECT_AUTO || p->direct_method > CUPIC_DIRECT_FFT)
        return set_err(err, CUPIC_ERR_ARG,
                       "direct_method must be 0 (automatic), 1 (dense) or 2 (FFT)");
    if (p->jacobi_max_sweeps < 1 || p->jacobi_max_sweeps > 1000000)
        return set_err(err, CUPIC_ERR_ARG, "jacobi_max_sweeps out of range");
    double ncells = (double)p->ncx * p->ncy;
    if (ncells * 10.0 > 2.0e9) return set_err(err, CUPIC_ERR_ARG, "grid too large");
    if ((double)p->per_row * p->ncy > 1.9e9)
        return set_err(err, CUPIC_ERR_ARG, "more than 1.9e9 particles per species");
    return CUPIC_OK;
}

struct Sizes {
    int64_t cap, n_init, far_cap, max_inject, nnodes, ncells, nblk_scan;
    int C;
};

Sizes sizes_of(const cupic_params *p, int64_t capacity_hint)
{
    Sizes z;
    z.n_init = p->per_row * p->ncy;
    double per_step = p->dt / std::min(p->dtin_e, p->dtin_i);
    z.max_inject = (int64_t)std::ceil(per_step * 2.0) + 1024;
    z.far_cap = std::max<int64_t>(65536, 4 * z.max_inject);
    z.cap = capacity_hint > 0 ? capacity_hint : z.n_init + z.n_init / 8 + 4 * z.max_inject + 4096;
    z.nnodes = (int64_t)p->ncx * (p->ncy + 1);
    z.ncells = (int64_t)p->ncx * p->ncy;
    z.nblk_scan = (z.ncells + kScanCells - 1) / kScanCells;
    double ppc = (double)p->per_row / p->ncx;
    int C = p->tile_cells > 0 ? p->tile_cells
                              : (int)std::lround(2048.0 / std::max(ppc, 1.0));
    z.C = std::max(1, std::min(kCmax, C));
    return z;
}

// Slot indices, offsets and counters are int32, so a species buffer holds
// at most kMaxCap slots; kernels address slots through float4 pointers
// (64-bit byte offsets), never as 4 * slot int words.
constexpr int64_t kMaxCap = 2000000000LL;

int check_capacity(const Sizes &z, std::string *err)
{
    if (z.cap < z.n_init) return set_err(err, CUPIC_ERR_ARG, "capacity below initial load");
    if (z.cap > kMaxCap) return set_err(err, CUPIC_ERR_ARG, "capacity too large (above 2e9 slots)");
    return CUPIC_OK;
}

// [ncy+1][ncx] periodic -> reference layout [nny][nnx] with column nnx-1 = column 0
template <class T>
void to_reference(const std::vector<T> &in, int ncx, int ncy, T *out)
{
    int nnx = ncx + 1;
    for (int j = 0; j <= ncy; j++) {
        for (int i = 0; i < ncx; i++) out[(size_t)j * nnx + i] = in[(size_t)j * ncx + i];
        out[(size_t)j * nnx + ncx] = in[(size_t)j * ncx];
    }
}

// charge_deposition; begin_solve: the conversion also starts the step's
// solve (a seq_solve with begun set must follow)
int charge_deposition(cupic_run *r, int begin_solve)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    if (r->moved) return set_err(&r->err, CUPIC_ERR_STATE, "charge_deposition between mover and cc");
    Target T = target_of(r);
    if (!r->acc_valid) {
        HIPTRY(&r->err, hipMemsetAsync(r->h.acc, 0, (size_t)r->h.nnodes * 8, r->stream));
        TimeScope ts(T, kTDeposit);
        launch_deposit(T.L(kShapeTasks));
    }
    seq_rho(T, begin_solve);
    r->acc_valid = false;
    return check_launch(r, "charge_deposition");
}

// particle_mover; end_solve: the counting pass also ends the step's direct
// solve (after seq_field with resid_end set)
int particle_mover(cupic_run *r, int end_solve)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    if (!r->initialised) return set_err(&r->err, CUPIC_ERR_STATE, "run not initialised");
    if (r->moved) return set_err(&r->err, CUPIC_ERR_STATE, "particle_mover called twice");
    if (r->acc_valid) {  // an unconsumed deposit would be added to
        HIPTRY(&r->err, hipMemsetAsync(r->h.acc, 0, (size_t)r->h.nnodes * 8, r->stream));
        r->acc_valid = false;
    }
    seq_mover(target_of(r), end_solve);
    r->moved = 1;
    return check_launch(r, "particle_mover");
}

}  // namespace

// ---------------------------------------------------------------- C-ABI

extern "C" {

int cupic_init_dev(int device, char *name, int name_len, int *compute_units)
{
    int n = 0;
    HIPTRY(nullptr, hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_err(nullptr, CUPIC_ERR_ARG, "no such device");
    HIPTRY(nullptr, hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPTRY(nullptr, hipGetDeviceProperties(&prop, device));
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    return noalloc_setup(device);
}

int cupic_params_default(cupic_params *p)
{
    if (!p) return set_err(nullptr, CUPIC_ERR_ARG, "null params");
    std::memset(p, 0, sizeof(*p));
    p->ncx = 128;
    p->ncy = 64;
    p->ds = 0.5;
    p->dt = 0.05;
    p->n = 20.0 / (0.5 * 0.5 * 0.5);
    p->mi = 100.0;
    p->kti = 1.0;
    p->me = 1.0;
    p->kte = 1.0;
    p->phi_p = -5.0;
    p->eps0 = 0.0;
    p->seed = 20261203ull;
    p->run_id = 0;
    p->quirks = 0;
    p->jacobi_mode = CUPIC_JACOBI_RESIDUAL;
    p->jacobi_min_sweeps = 0;
    p->jacobi_max_sweeps = 500;
    p->jacobi_check_every = 50;
    p->jacobi_tol = 1e-6;
    p->jacobi_tol_init = 1e-6;
    return cupic_params_derive(p);
}

// Normalised derived values, src/init.cu:451-789 (eps0 reduces to n).
int cupic_params_derive(cupic_params *p)
{
    if (!p) return set_err(nullptr, CUPIC_ERR_ARG, "null params");
    const double pi = 3.1415926535897932;  // src/stdh.h:26
    if (p->me == 0.0) p->me = 1.0;
    if (p->kte == 0.0) p->kte = 1.0;
    if (p->eps0 == 0.0) p->eps0 = p->n;
    if (p->jacobi_check_every <= 0) p->jacobi_check_every = 50;
    p->Lx = p->ds * (double)p->ncx;
    p->Ly = p->ds * (double)p->ncy;
    p->per_row = (int64_t)(p->n * p->Lx * p->ds * p->ds);
    p->dtin_i = std::sqrt(2.0 * pi * p->mi / p->kti) / (p->n * p->Lx * p->ds);
    p->dtin_e = std::sqrt(2.0 * pi) / (p->n * p->Lx * p->ds);
    return validate(p, nullptr);
}

// Positional reader of input/input_data (src/init.cu:424-447): a header
// line, then 13 "key = value;" lines read by position (keys ignored).
int cupic_params_from_input_file(const char *path, cupic_params *p, int32_t *n_ini,
                                 int32_t *n_prev, int32_t *n_save, int32_t *n_fin)
{
    if (!path || !p) return set_err(nullptr, CUPIC_ERR_ARG, "null argument");
    FILE *f = std::fopen(path, "r");
    if (!f) return set_err(nullptr, CUPIC_ERR_IO, std::string("cannot open ") + path);
    char line[256];
    double v[14] = {};
    int got = 0;
    if (!std::fgets(line, sizeof line, f)) { std::fclose(f); return set_err(nullptr, CUPIC_ERR_IO, "empty input file"); }
    for (int k = 1; k <= 13; k++) {
        if (!std::fgets(line, sizeof line, f)) break;
        double d;
        if (std::sscanf(line, "%*s = %lf;", &d) != 1) break;
        v[k] = d;
        got++;
    }
    std::fclose(f);
    if (got != 13) return set_err(nullptr, CUPIC_ERR_IO, "input file: expected 13 value lines");
    const double CST_ME = 9.109e-31, CST_E = 1.602e-19, CST_KB = 1.381e-23,
                 CST_EPSILON = 8.854e-12;  // src/init.h:24-27
    double ne = v[5], Te = v[6], beta = v[7], gamma = v[8], phi_v = v[9];
    cupic_params_default(p);
    p->ncx = (int32_t)v[10];
    p->ncy = (int32_t)v[11];
    p->ds = v[12];
    p->dt = v[13];
    double Dl = std::sqrt(CST_EPSILON * CST_KB * Te / (ne * CST_E * CST_E));  // src/init.cu:729
    p->n = ne * Dl * Dl * Dl;                                                // src/init.cu:551
    p->mi = gamma;
    p->kti = beta;
    p->phi_p = phi_v * (CST_E / (CST_KB * Te));                              // src/init.cu:534
    double eps = CST_EPSILON;                                                // src/init.cu:622-626
    eps /= std::pow(Dl * std::sqrt(CST_ME / (CST_KB * Te)), 2);
    eps /= CST_E * CST_E;
    eps *= Dl * Dl * Dl;
    eps *= CST_ME;
    p->eps0 = eps;
    p->jacobi_mode = CUPIC_JACOBI_REFERENCE;
    p->jacobi_tol = 1.0e-4;       // src/main.cu:64
    p->jacobi_tol_init = 1.0e-3;  // src/init.cu:254
    p->jacobi_min_sweeps = std::max(p->ncx + 1, p->ncy + 1) + 1;
    p->jacobi_max_sweeps = std::max(500, p->jacobi_min_sweeps + 1);
    if (n_ini) *n_ini = (int32_t)v[1];
    if (n_prev) *n_prev = (int32_t)v[2];
    if (n_save) *n_save = (int32_t)v[3];
    if (n_fin) *n_fin = (int32_t)v[4];
    return cupic_params_derive(p);
}

int cupic_run_params(const cupic_run *r, cupic_params *out)
{
    if (!r || !out) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    *out = r->p;
    return CUPIC_OK;
}

int cupic_run_set_tile(cupic_run *r, int32_t tile)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    if (tile < 0 || tile > kCmax) return set_err(&r->err, CUPIC_ERR_ARG, "tile must be 0..16");
    if (r->moved) return set_err(&r->err, CUPIC_ERR_STATE, "set_tile between mover and cc");
    cupic_params q = r->p;
    q.tile_cells = tile;
    Sizes z = sizes_of(&q, r->h.cap);
    HIPTRY(&r->err, hipStreamSynchronize(r->stream));
    r->p.tile_cells = tile;
    r->h.C = z.C;
    r->h.tpr = (r->h.ncx + z.C - 1) / z.C;
    std::vector<int> plan(kShapeCount * 2);
    for (int s = 0; s < kShapeCount; s++) {
        r->nblk[s] = shape_blocks(r->h, (Shape)s);
        plan[2 * s] = 0;
        plan[2 * s + 1] = r->nblk[s];
    }
    HIPTRY(&r->err, hipMemcpy(r->d_rd, &r->h, sizeof(RunDev), hipMemcpyHostToDevice));
    HIPTRY(&r->err, hipMemcpy(r->d_plan, plan.data(), plan.size() * sizeof(int),
                              hipMemcpyHostToDevice));
    return CUPIC_OK;
}

int cupic_run_set_hint_mode(cupic_run *r, int32_t mode)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    if (mode < kHintOff || mode > kHintChecked)
        return set_err(&r->err, CUPIC_ERR_ARG, "hint mode must be 0, 1 or 2");
    if (r->moved) return set_err(&r->err, CUPIC_ERR_STATE, "set_hint_mode between mover and cc");
    HIPTRY(&r->err, hipStreamSynchronize(r->stream));
    r->h.hint_mode = mode;
    HIPTRY(&r->err, hipMemcpy(r->d_rd, &r->h, sizeof(RunDev), hipMemcpyHostToDevice));
    return CUPIC_OK;
}

int cupic_run_set_statistics(cupic_run *r, int32_t on)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    if (r->moved) return set_err(&r->err, CUPIC_ERR_STATE, "set_statistics between mover and cc");
    HIPTRY(&r->err, hipStreamSynchronize(r->stream));
    r->h.count_stats = on ? 1 : 0;
    HIPTRY(&r->err, hipMemcpy(r->d_rd, &r->h, sizeof(RunDev), hipMemcpyHostToDevice));
    return CUPIC_OK;
}

int cupic_run_footprint(const cupic_params *pin, int64_t capacity_hint, int64_t *bytes)
{
    if (!pin || !bytes) return set_err(nullptr, CUPIC_ERR_ARG, "null argument");
    cupic_params p = *pin;
    int st = cupic_params_derive(&p);
    if (st) return st;
    Sizes z = sizes_of(&p, capacity_hint);
    if ((st = check_capacity(z, nullptr))) return st;
    int64_t b = 0;
    b += 2 * 2 * 4 * z.cap * 4;                  // particles
    b += 2 * 2 * ((z.cap + 255) & ~(int64_t)255);  // push hints
    b += 2 * 2 * (z.ncells + 1) * 4;             // offsets
    b += 2 * z.ncells * (9 + 10 + 1 + 1) * 4;  // cnt, base, far_cnt, far_cur
    b += 2 * z.far_cap * (int64_t)sizeof(FarRec) + 2 * z.max_inject * 8;
    b += z.nnodes * 64;                          // grids
    b += (int64_t)(p.jacobi_max_sweeps + 2) * 12 * kRedSlots;
    if (p.jacobi_mode == 3) {
        const int64_t nx = p.ncx, m = p.ncy - 1;
        if (direct_class(p)) {
            b += 16 * (int64_t)fft_tabs(p.ncx, p.ncy).total + 8 * (nx + m + 2 * m * nx);
        } else {
            const int64_t ns = std::max(dft_slices(p.ncx, p.ncy - 1, p.ncx),
                                        dft_slices(p.ncx, p.ncy - 1, p.ncy - 1));
            b += 8 * (2 * nx * nx + m * m + nx + m + (2 + ns) * m * nx);
        }
    }
    *bytes = b;
    return CUPIC_OK;
}

int cupic_run_create(const cupic_params *pin, int64_t capacity_hint, cupic_run **out)
{
    if (!pin || !out) return set_err(nullptr, CUPIC_ERR_ARG, "null argument");
    *out = nullptr;
    cupic_params p = *pin;
    int status = cupic_params_derive(&p);
    if (status) return status;
    Sizes z = sizes_of(&p, capacity_hint);
    if ((status = check_capacity(z, nullptr))) return status;
    cupic_run *r = new cupic_run();
    r->p = p;
    RunDev &h = r->h;
    h.ncx = p.ncx;
    h.ncy = p.ncy;
    h.ncells = (int)z.ncells;
    h.nnodes = (int)z.nnodes;
    h.C = z.C;
    h.tpr = (p.ncx + z.C - 1) / z.C;
    h.C1 = std::min(kCountCells, p.ncx);
    h.tpr1 = (p.ncx + h.C1 - 1) / h.C1;
    h.cap = (int)z.cap;
    h.far_cap = (int)z.far_cap;
    h.max_inject = (int)z.max_inject;
    h.n_init = (int)z.n_init;
    h.per_row = (int)p.per_row;
    h.max_sweeps_alloc = p.jacobi_max_sweeps;
    h.step_mode = p.jacobi_mode;
    h.step_min = p.jacobi_min_sweeps;
    h.step_max = p.jacobi_max_sweeps;
    h.step_check = p.jacobi_check_every;
    h.step_tol = p.jacobi_tol;
    h.init_tol = p.jacobi_tol_init;
    h.inv_ds = (float)(1.0 / p.ds);
    h.Lx_f = (float)p.Lx;
    h.Ly_f = (float)p.Ly;
    h.dt_f = (float)p.dt;
    h.ds = p.ds;
    h.dt = p.dt;
    h.Lx = p.Lx;
    h.Ly = p.Ly;
    h.eps0 = p.eps0;
    h.phi_p = p.phi_p;
    h.b_interior = 0x1p-32 / (p.ds * p.eps0);
    h.rho_interior = 0x1p-32 / (p.ds * p.ds * p.ds);
    h.rho_edge = 0x1p-32 / (0.5 * p.ds * p.ds * p.ds);
    for (int s = 0; s < 2; s++) {
        double q = s == 0 ? -1.0 : 1.0;
        double m = s == 0 ? p.me : p.mi;
        double kt = s == 0 ? p.kte : p.kti;
        h.sp[s].q = q;
        h.sp[s].m = m;
        h.sp[s].sigma = std::sqrt(kt / m);
        h.sp[s].dtin = s == 0 ? p.dtin_e : p.dtin_i;
        h.sp[s].qdtm = (float)(q * p.dt / m);
        h.sp[s].hkick = (float)(-0.5 * q * p.dt / m);
    }
    h.seed = p.seed;
    h.run_id = p.run_id;
    h.quirks = p.quirks;
    h.hint_mode = kHintOn;
    int st = CUPIC_OK;
    for (int s = 0; s < 2 && !st; s++) {
        for (int b = 0; b < 2; b++) {
            h.part[s][b] = dev_alloc<float4>(r, z.cap, &st);
            // padded to whole 256-byte lines: the counting pass reads 4-byte words
            h.hint[s][b] = dev_alloc<unsigned char>(r, ((size_t)z.cap + 255) & ~(size_t)255, &st);
            h.off[s][b] = dev_alloc<int>(r, z.ncells + 1, &st);
        }
        h.cnt[s] = dev_alloc<int>(r, z.ncells * 9, &st);
        h.base[s] = dev_alloc<int>(r, z.ncells * 10, &st);
        h.far_cnt[s] = dev_alloc<int>(r, z.ncells, &st);
        h.far_cur[s] = dev_alloc<int>(r, z.ncells, &st);
        h.far[s] = dev_alloc<FarRec>(r, z.far_cap, &st);
        h.inj_tin[s] = dev_alloc<double>(r, z.max_inject, &st);
        h.blk[s] = dev_alloc<int>(r, z.nblk_scan, &st);
    }
    h.acc = dev_alloc<unsigned long long>(r, z.nnodes, &st);
    h.rho = dev_alloc<float>(r, z.nnodes, &st);
    h.b32 = dev_alloc<float>(r, z.nnodes, &st);
    h.b64 = dev_alloc<double>(r, z.nnodes, &st);
    h.phi32[0] = dev_alloc<float>(r, z.nnodes, &st);
    h.phi32[1] = dev_alloc<float>(r, z.nnodes, &st);
    h.phi32[2] = dev_alloc<float>(r, z.nnodes, &st);
    h.phi64 = dev_alloc<double>(r, z.nnodes, &st);
    h.d32[0] = dev_alloc<float>(r, z.nnodes, &st);
    h.d32[1] = dev_alloc<float>(r, z.nnodes, &st);
    h.r32 = dev_alloc<float>(r, z.nnodes, &st);
    h.E = dev_alloc<float2>(r, z.nnodes, &st);
    h.sweep_max = dev_alloc<unsigned int>(r, ((size_t)p.jacobi_max_sweeps + 1) * kRedSlots, &st);
    h.resid = dev_alloc<unsigned long long>(r, ((size_t)p.jacobi_max_sweeps + 2) * kRedSlots, &st);
    if (p.jacobi_mode == 3) {
        const size_t nx = (size_t)p.ncx, m = (size_t)p.ncy - 1;
        h.dft_fft = direct_class(p);
        if (h.dft_fft) {
            h.dft_tw = dev_alloc<double>(r, 2 * (size_t)fft_tabs(p.ncx, p.ncy).total, &st);
        } else {
            h.nsx = dft_slices(p.ncx, p.ncy - 1, p.ncx);
            h.nsy = dft_slices(p.ncx, p.ncy - 1, p.ncy - 1);
            h.dft_f = dev_alloc<double>(r, nx * nx, &st);
            h.dft_g = dev_alloc<double>(r, nx * nx, &st);
            h.dft_s = dev_alloc<double>(r, m * m, &st);
            h.dft_p = dev_alloc<double>(r, (size_t)std::max(h.nsx, h.nsy) * m * nx, &st);
        }
        h.dft_d = dev_alloc<double>(r, nx + m, &st);
        h.dft_z = dev_alloc<double>(r, m * nx, &st);
        h.dft_y = dev_alloc<double>(r, m * nx, &st);
    }
    h.st = dev_alloc<RunState>(r, 1, &st);
    r->d_rd = dev_alloc<RunDev>(r, 1, &st);
    r->d_plan = dev_alloc<int>(r, (size_t)kShapeCount * 2, &st);
    if (st) {
        cupic_run_destroy(r);
        return set_err(nullptr, CUPIC_ERR_ALLOC, "device allocation failed");
    }
    std::vector<int> plan(kShapeCount * 2);
    for (int s = 0; s < kShapeCount; s++) {
        r->nblk[s] = shape_blocks(h, (Shape)s);
        plan[2 * s] = 0;
        plan[2 * s + 1] = r->nblk[s];
    }
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&r->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&r->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&r->ev_field, hipEventDisableTiming) != hipSuccess ||
        hipMemcpy(r->d_rd, &h, sizeof(RunDev), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(r->d_plan, plan.data(), plan.size() * sizeof(int), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        cupic_run_destroy(r);
        return set_err(nullptr, CUPIC_ERR_HIP, "run setup failed");
    }
    if (h.dft_d) {
        launch_dft_init(target_of(r).L(kShapeDftInit));
        static const Shape tabs[3] = {kShapeFftTabS, kShapeFftTabM, kShapeFftTabL};
        for (int c = 0; c < 3; c++) launch_fft(target_of(r).L(tabs[c]), 3, c + 1);
        if (hipStreamSynchronize(r->stream) != hipSuccess) {
            cupic_run_destroy(r);
            return set_err(nullptr, CUPIC_ERR_HIP, "direct-solve tables failed");
        }
    }
    *out = r;
    return CUPIC_OK;
}

int cupic_run_destroy(cupic_run *r)
{
    if (!r) return CUPIC_OK;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    for (auto &e : r->evs) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    for (auto e : r->pool) (void)hipEventDestroy(e);
    if (r->side) (void)hipStreamSynchronize(r->side);
    for (void *p : r->allocs) (void)hipFree(p);
    if (r->ev_fork) (void)hipEventDestroy(r->ev_fork);
    if (r->ev_join) (void)hipEventDestroy(r->ev_join);
    if (r->ev_field) (void)hipEventDestroy(r->ev_field);
    if (r->side) (void)hipStreamDestroy(r->side);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return CUPIC_OK;
}

const char *cupic_last_error(const cupic_run *r)
{
    if (r && !r->err.empty()) return r->err.c_str();
    return g_err.c_str();
}

int cupic_synchronize(cupic_run *r)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    HIPTRY(&r->err, hipStreamSynchronize(r->stream));
    RunState s;
    return read_state(r, &s);
}

// ---- init

int cupic_create_particles(cupic_run *r)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    Target T = target_of(r);
    launch_init_finish(T.L(kShapeRun));
    launch_create(T.L(kShapeInit));
    launch_init_sort(T.L(kShapeRows));
    r->acc_valid = false;
    r->moved = 0;
    r->initialised = true;
    r->t_next = r->p.dt;
    return check_launch(r, "create_particles");
}

int cupic_charge_deposition(cupic_run *r) { return charge_deposition(r, 0); }

int cupic_initialize_mesh(cupic_run *r)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    Target T = target_of(r);
    launch_phi_linear(T.L(kShapeNodes));
    r->acc_valid = false;
    int st = cupic_charge_deposition(r);
    if (st) return st;
    seq_solve(T, 1, r->p.jacobi_mode, r->p.jacobi_tol_init, r->p.jacobi_min_sweeps,
              r->p.jacobi_max_sweeps, r->p.jacobi_check_every);
    seq_field(T);
    return check_launch(r, "initialize_mesh");
}

int cupic_adjust_leap_frog(cupic_run *r)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    launch_half_kick(target_of(r).L(kShapeInit));
    return check_launch(r, "adjust_leap_frog");
}

int cupic_init_sim(cupic_run *r, double *t)
{
    int st = cupic_create_particles(r);
    if (!st) st = cupic_initialize_mesh(r);
    if (!st) st = cupic_adjust_leap_frog(r);
    if (!st && t) *t = 0.0;
    return st;
}

// ---- mesh

int cupic_poisson_solver(cupic_run *r, double tol, int32_t *sweeps, double *resid)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    if (r->moved) return set_err(&r->err, CUPIC_ERR_STATE, "poisson_solver between mover and cc");
    Target T = target_of(r);
    seq_solve(T, 1, r->p.jacobi_mode, tol < 0 ? r->p.jacobi_tol : tol, r->p.jacobi_min_sweeps,
              r->p.jacobi_max_sweeps, r->p.jacobi_check_every);
    int st = check_launch(r, "poisson_solver");
    if (st) return st;
    if (sweeps || resid) {
        RunState s;
        st = read_state(r, &s);
        if (sweeps) *sweeps = s.sweeps_done;
        if (resid) *resid = s.last_resid;
    }
    return st;
}

int cupic_field_solver(cupic_run *r)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    if (r->moved) return set_err(&r->err, CUPIC_ERR_STATE, "field_solver between mover and cc");
    seq_field(target_of(r));
    return check_launch(r, "field_solver");
}

// ---- particles and boundary conditions

int cupic_particle_mover(cupic_run *r) { return particle_mover(r, 0); }

int cupic_cc(cupic_run *r, double t)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    if (!r->moved) return set_err(&r->err, CUPIC_ERR_STATE, "cc without particle_mover");
    seq_cc(target_of(r), 1, t, 0, r->clock_forked);
    r->clock_forked = false;
    r->moved = 0;
    r->acc_valid = true;
    r->t_next = t + r->p.dt;
    return check_launch(r, "cc");
}

int cupic_step(cupic_run *r, double t)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    if (r->moved) return set_err(&r->err, CUPIC_ERR_STATE, "step between mover and cc");
    int st = charge_deposition(r, 1);
    if (!st) {
        seq_clock_fork(target_of(r), 1, t);
        r->clock_forked = true;
    }
    if (!st) {
        const bool direct = r->p.jacobi_mode == 3;
        seq_solve(target_of(r), 0, r->p.jacobi_mode, 0, 0, r->p.jacobi_max_sweeps,
                  r->p.jacobi_check_every, true, direct);
        seq_field(target_of(r), direct);
        seq_inject_fork(target_of(r));
        st = check_launch(r, "solve/field");
        if (!st) st = particle_mover(r, direct);
    }
    if (!st) st = cupic_cc(r, t);
    return st;
}

// ---- diagnostics

int cupic_number_of_particles(cupic_run *r, int species, int64_t *n)
{
    if (!r || !n || species < 0 || species > 1) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    RunState s;
    int st = read_state(r, &s);
    *n = s.N[species];
    return st;
}

int cupic_probe_tally(cupic_run *r, int64_t last_probe[2], int64_t last_top[2],
                      int64_t last_injected[2], int64_t total_probe[2], double *charge)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    RunState s;
    int st = read_state(r, &s);
    for (int k = 0; k < 2; k++) {
        if (last_probe) last_probe[k] = (int64_t)s.last_probe[k];
        if (last_top) last_top[k] = (int64_t)s.last_top[k];
        if (last_injected) last_injected[k] = (int64_t)s.last_injected[k];
        if (total_probe) total_probe[k] = (int64_t)s.total_probe[k];
    }
    if (charge) *charge = s.probe_charge_total;
    return st;
}

int cupic_last_solve(cupic_run *r, int32_t *sweeps, double *resid)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    RunState s;
    int st = read_state(r, &s);
    if (sweeps) *sweeps = s.sweeps_done;
    if (resid) *resid = s.last_resid;
    return st;
}

int cupic_solver_totals(cupic_run *r, int64_t *sweeps_total, int64_t *solves)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    RunState s;
    int st = read_state(r, &s);
    if (sweeps_total) *sweeps_total = s.sweeps_total;
    if (solves) *solves = s.solves;
    return st;
}

int cupic_run_counters(cupic_run *r, int64_t total_far[2], int64_t total_undecided[2],
                       int64_t *steps, int64_t *particle_steps)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    RunState s;
    int st = read_state(r, &s);
    for (int k = 0; k < 2; k++) {
        if (total_far) total_far[k] = (int64_t)s.total_far[k];
        if (total_undecided) total_undecided[k] = (int64_t)s.total_undecided[k];
    }
    if (steps) *steps = s.step;
    if (particle_steps) *particle_steps = (int64_t)s.particle_steps;
    return st;
}

int cupic_get_particles(cupic_run *r, int species, float *x, float *y, float *vx, float *vy,
                        int64_t capacity, int64_t *n)
{
    if (!r || species < 0 || species > 1) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    if (r->moved) return set_err(&r->err, CUPIC_ERR_STATE, "particles are pending cc");
    RunState s;
    int st = read_state(r, &s);
    if (st && st != CUPIC_ERR_CAPACITY) return st;
    int64_t N = s.N[species];
    if (n) *n = N;
    if (capacity < N) return set_err(&r->err, CUPIC_ERR_ARG, "host buffer too small");
    std::vector<float4> buf((size_t)N);
    HIPTRY(&r->err, hipMemcpy(buf.data(), r->h.part[species][s.cur], (size_t)N * sizeof(float4),
                              hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < N; k++) {
        if (x) x[k] = buf[k].x;
        if (y) y[k] = buf[k].y;
        if (vx) vx[k] = buf[k].z;
        if (vy) vy[k] = buf[k].w;
    }
    return st;
}

int cupic_set_particles(cupic_run *r, int species, const float *x, const float *y,
                        const float *vx, const float *vy, int64_t n)
{
    if (!r || species < 0 || species > 1 || n < 0 || !x || !y || !vx || !vy)
        return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    if (r->moved) return set_err(&r->err, CUPIC_ERR_STATE, "particles are pending cc");
    if (n > r->h.cap) return set_err(&r->err, CUPIC_ERR_CAPACITY, "more particles than capacity");
    const RunDev &h = r->h;
    RunState s;
    int st = read_state(r, &s);
    if (st && st != CUPIC_ERR_CAPACITY) return st;
    // stable counting sort by cell on the host
    std::vector<int> key((size_t)n), cnt((size_t)h.ncells + 1, 0);
    for (int64_t k = 0; k < n; k++) {
        key[k] = host_cell(h, x[k], y[k]);
        cnt[key[k] + 1]++;
    }
    for (int k = 0; k < h.ncells; k++) cnt[k + 1] += cnt[k];
    std::vector<int> off(cnt);
    std::vector<float4> buf((size_t)n);
    for (int64_t k = 0; k < n; k++) {
        int d = cnt[key[k]]++;
        buf[d] = make_float4(x[k], y[k], vx[k], vy[k]);
    }
    HIPTRY(&r->err, hipMemcpy(h.part[species][s.cur], buf.data(), (size_t)n * sizeof(float4),
                              hipMemcpyHostToDevice));
    HIPTRY(&r->err, hipMemcpy(h.off[species][s.cur], off.data(), off.size() * sizeof(int),
                              hipMemcpyHostToDevice));
    HIPTRY(&r->err, hipMemset(h.hint[species][s.cur], 0, (size_t)std::max<int64_t>(n, 1)));
    s.N[species] = (int)n;
    s.N_new[species] = (int)n;
    r->acc_valid = false;
    r->initialised = true;
    return write_state(r, s);
}

int cupic_read_particle_file(cupic_run *r, int species, const char *path)
{
    if (!r || !path || species < 0 || species > 1)
        return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    FILE *f = std::fopen(path, "r");
    if (!f) return set_err(&r->err, CUPIC_ERR_IO, std::string("cannot open ") + path);
    const bool skip_first = (r->p.quirks & CUPIC_QUIRK_Q10_SKIP_FIRST_LINE) != 0;
    std::vector<float> x, y, vx, vy;
    std::string line;
    long lineno = 0;
    int st = CUPIC_OK;
    for (int c = std::fgetc(f); c != EOF || !line.empty(); c = std::fgetc(f)) {
        if (c != EOF && c != '\n') {
            line.push_back((char)c);
            continue;
        }
        lineno++;
        const bool blank = line.find_first_not_of(" \t\r\v\f") == std::string::npos;
        if (!(skip_first && lineno == 1) && !blank) {
            double a, b, v, w;
            char extra;
            if (std::sscanf(line.c_str(), " %le %le %le %le %c", &a, &b, &v, &w, &extra) != 4) {
                st = set_err(&r->err, CUPIC_ERR_IO,
                             std::string(path) + ": line " + std::to_string(lineno) +
                                 " is not four numbers");
                break;
            }
            x.push_back((float)a);
            y.push_back((float)b);
            vx.push_back((float)v);
            vy.push_back((float)w);
        }
        line.clear();
        if (c == EOF) break;
    }
    std::fclose(f);
    if (st) return st;
    const float none = 0.0f;
    auto ptr = [&](std::vector<float> &v) { return v.empty() ? &none : v.data(); };
    return cupic_set_particles(r, species, ptr(x), ptr(y), ptr(vx), ptr(vy), (int64_t)x.size());
}

int cupic_init_sim_restart(cupic_run *r, int32_t n_ini, const char *ions_path,
                           const char *electrons_path, double *t)
{
    if (!r || !ions_path || !electrons_path) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    if (n_ini <= 0) return set_err(&r->err, CUPIC_ERR_ARG, "n_ini must be positive for a restart");
    // a fresh run state (counts, tallies, clocks), then the two files
    launch_init_finish(target_of(r).L(kShapeRun));
    int st = check_launch(r, "init_sim_restart");
    r->moved = 0;
    r->acc_valid = false;
    if (!st) st = cupic_read_particle_file(r, 1, ions_path);   // src/init.cu:338-339
    if (!st) st = cupic_read_particle_file(r, 0, electrons_path);
    if (!st) st = cupic_initialize_mesh(r);
    if (st) return st;
    const double t0 = (double)n_ini * r->p.dt;  // src/init.cu:86
    r->t_next = t0 + r->p.dt;
    if (t) *t = t0;
    return CUPIC_OK;
}

int cupic_input_file_set_n_ini(const char *path, int32_t n_ini)
{
    if (!path || n_ini < 0) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    FILE *f = std::fopen(path, "r");
    if (!f) return set_err(nullptr, CUPIC_ERR_IO, std::string("cannot open ") + path);
    std::vector<std::string> lines;
    std::string line;
    for (int c = std::fgetc(f); c != EOF; c = std::fgetc(f)) {
        if (c == '\n') {
            lines.push_back(line);
            line.clear();
        } else {
            line.push_back((char)c);
        }
    }
    if (!line.empty()) lines.push_back(line);
    std::fclose(f);
    if (lines.size() < 2) return set_err(nullptr, CUPIC_ERR_IO, std::string(path) + ": no n_ini line");
    lines[1] = "n_ini = " + std::to_string(n_ini) + ";";  // src/main.cu:104-105
    const std::string tmp = std::string(path) + ".new";
    FILE *o = std::fopen(tmp.c_str(), "w");
    if (!o) return set_err(nullptr, CUPIC_ERR_IO, "cannot write " + tmp);
    bool ok = true;
    for (const std::string &l : lines) ok = ok && std::fprintf(o, "%s\n", l.c_str()) >= 0;
    ok = (std::fclose(o) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0)
        return set_err(nullptr, CUPIC_ERR_IO, std::string("cannot replace ") + path);
    return CUPIC_OK;
}

int cupic_density_profile(cupic_run *r, int species, int64_t *per_row)
{
    if (!r || !per_row || species < 0 || species > 1)
        return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    std::vector<int32_t> off((size_t)r->h.ncells + 1);
    int st = cupic_get_offsets(r, species, off.data());
    if (st) return st;
    for (int j = 0; j < r->h.ncy; j++)
        per_row[j] = (int64_t)off[(size_t)(j + 1) * r->h.ncx] - off[(size_t)j * r->h.ncx];
    return CUPIC_OK;
}

// Row bookmarks as the reference keeps them (src/cc.cu:760-763,
// src/init.cu:401-407): bm[2j] and bm[2j+1] are the first and last index of
// the particles in cell row j, both -1 for an empty row.  Rows are contiguous
// in the cell-sorted order, so they follow from the cell offsets.
int cupic_get_bookmarks(cupic_run *r, int species, int32_t *bm)
{
    if (!r || !bm || species < 0 || species > 1) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    std::vector<int32_t> off((size_t)r->h.ncells + 1);
    int st = cupic_get_offsets(r, species, off.data());
    if (st) return st;
    for (int j = 0; j < r->h.ncy; j++) {
        int32_t a = off[(size_t)j * r->h.ncx], b = off[(size_t)(j + 1) * r->h.ncx];
        bm[2 * j] = b > a ? a : -1;
        bm[2 * j + 1] = b > a ? b - 1 : -1;
    }
    return CUPIC_OK;
}

int cupic_get_offsets(cupic_run *r, int species, int32_t *offsets)
{
    if (!r || !offsets || species < 0 || species > 1) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    RunState s;
    int st = read_state(r, &s);
    if (st && st != CUPIC_ERR_CAPACITY) return st;
    HIPTRY(&r->err, hipMemcpy(offsets, r->h.off[species][s.cur],
                              ((size_t)r->h.ncells + 1) * sizeof(int), hipMemcpyDeviceToHost));
    return st;
}

int cupic_get_mesh(cupic_run *r, int which, void *host, int64_t count)
{
    if (!r || !host) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    const RunDev &h = r->h;
    int64_t need = (int64_t)(h.ncx + 1) * (h.ncy + 1);
    if (count < need) return set_err(&r->err, CUPIC_ERR_ARG, "host buffer too small");
    RunState s;
    int st = read_state(r, &s);
    if (st && st != CUPIC_ERR_CAPACITY) return st;
    size_t nn = (size_t)h.nnodes;
    switch (which) {
    case CUPIC_MESH_RHO: {
        std::vector<float> v(nn);
        HIPTRY(&r->err, hipMemcpy(v.data(), h.rho, nn * 4, hipMemcpyDeviceToHost));
        to_reference(v, h.ncx, h.ncy, (float *)host);
        break;
    }
    case CUPIC_MESH_PHI: {
        std::vector<double> v(nn);
        if (r->p.jacobi_mode >= 2) {
            HIPTRY(&r->err, hipMemcpy(v.data(), h.phi64, nn * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> f(nn);
            HIPTRY(&r->err, hipMemcpy(f.data(), h.phi32[s.phi_cur], nn * 4, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < nn; k++) v[k] = f[k];
        }
        to_reference(v, h.ncx, h.ncy, (double *)host);
        break;
    }
    case CUPIC_MESH_EX:
    case CUPIC_MESH_EY: {
        std::vector<float2> e(nn);
        HIPTRY(&r->err, hipMemcpy(e.data(), h.E, nn * 8, hipMemcpyDeviceToHost));
        std::vector<float> v(nn);
        for (size_t k = 0; k < nn; k++) v[k] = which == CUPIC_MESH_EX ? e[k].x : e[k].y;
        to_reference(v, h.ncx, h.ncy, (float *)host);
        break;
    }
    case CUPIC_MESH_CHARGE: {
        if (!r->acc_valid)
            return set_err(&r->err, CUPIC_ERR_STATE,
                           "no pending charge: it is consumed by charge_deposition");
        std::vector<long long> v(nn);
        HIPTRY(&r->err, hipMemcpy(v.data(), h.acc, nn * 8, hipMemcpyDeviceToHost));
        to_reference(v, h.ncx, h.ncy, (long long *)host);
        break;
    }
    default:
        return set_err(&r->err, CUPIC_ERR_ARG, "unknown mesh quantity");
    }
    return CUPIC_OK;
}

int cupic_set_phi(cupic_run *r, const double *phi, int64_t count)
{
    if (!r || !phi) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    const RunDev &h = r->h;
    int nnx = h.ncx + 1;
    if (count < (int64_t)nnx * (h.ncy + 1)) return set_err(&r->err, CUPIC_ERR_ARG, "array too small");
    size_t nn = (size_t)h.nnodes;
    std::vector<double> v(nn);
    std::vector<float> f(nn);
    for (int j = 0; j <= h.ncy; j++)
        for (int i = 0; i < h.ncx; i++) {
            v[(size_t)j * h.ncx + i] = phi[(size_t)j * nnx + i];
            f[(size_t)j * h.ncx + i] = (float)phi[(size_t)j * nnx + i];
        }
    HIPTRY(&r->err, hipStreamSynchronize(r->stream));
    HIPTRY(&r->err, hipMemcpy(h.phi64, v.data(), nn * 8, hipMemcpyHostToDevice));
    for (int q = 0; q < 3; q++)
        HIPTRY(&r->err, hipMemcpy(h.phi32[q], f.data(), nn * 4, hipMemcpyHostToDevice));
    return CUPIC_OK;
}

int cupic_get_clock(cupic_run *r, double tin[2], uint64_t injected_total[2], int32_t *started)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    RunState s;
    int st = read_state(r, &s);
    for (int k = 0; k < 2; k++) {
        if (tin) tin[k] = s.tin[k];
        if (injected_total) injected_total[k] = s.ninj[k];
    }
    if (started) *started = s.tin_started;
    return st;
}

int cupic_set_clock(cupic_run *r, const double tin[2], const uint64_t injected_total[2],
                    int32_t started)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    RunState s;
    int st = read_state(r, &s);
    if (st && st != CUPIC_ERR_CAPACITY) return st;
    for (int k = 0; k < 2; k++) {
        if (tin) s.tin[k] = tin[k];
        if (injected_total) s.ninj[k] = injected_total[k];
    }
    s.tin_started = started;
    return write_state(r, s);
}

// Text snapshots in the reference formats: " %.17e %.17e %.17e %.17e \n" per
// particle (src/diagnostic.cu:76-79) and " %d %d %.17e \n" per node with a
// blank line after each column (src/diagnostic.cu:113-120); ".dat" is
// appended to the path as the reference does.
int cupic_particles_snapshot(cupic_run *r, int species, const char *path, double t)
{
    (void)t;
    if (!r || !path) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    int64_t n = 0;
    int st = cupic_number_of_particles(r, species, &n);
    if (st) return st;
    std::vector<float> x(n), y(n), vx(n), vy(n);
    st = cupic_get_particles(r, species, x.data(), y.data(), vx.data(), vy.data(), n, &n);
    if (st) return st;
    std::string fn = std::string(path) + ".dat";
    FILE *f = std::fopen(fn.c_str(), "w");
    if (!f) return set_err(&r->err, CUPIC_ERR_IO, "cannot write " + fn);
    for (int64_t k = 0; k < n; k++)
        std::fprintf(f, " %.17e %.17e %.17e %.17e \n", (double)x[k], (double)y[k], (double)vx[k],
                     (double)vy[k]);
    std::fclose(f);
    return CUPIC_OK;
}

// save_bm (src/diagnostic.cu:137-167): " %d %d %d \n" per row with the even
// index 2j first, as the reference's loop prints it.
int cupic_bookmarks_snapshot(cupic_run *r, int species, const char *path)
{
    if (!r || !path) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    std::vector<int32_t> bm(2 * (size_t)r->h.ncy);
    int st = cupic_get_bookmarks(r, species, bm.data());
    if (st) return st;
    std::string fn = std::string(path) + ".dat";
    FILE *f = std::fopen(fn.c_str(), "w");
    if (!f) return set_err(&r->err, CUPIC_ERR_IO, "cannot write " + fn);
    for (int i = 0; i < 2 * r->h.ncy; i += 2) std::fprintf(f, " %d %d %d \n", i, bm[i], bm[i + 1]);
    std::fclose(f);
    return CUPIC_OK;
}

// save_bins (src/diagnostic.cu:171-215): " %d %d \n" per particle, its index
// in storage order and int(y/ds) with y widened to double as the reference's
// float/double division does.
int cupic_bins_snapshot(cupic_run *r, int species, const char *path)
{
    if (!r || !path) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    int64_t n = 0;
    int st = cupic_number_of_particles(r, species, &n);
    if (st) return st;
    std::vector<float> x(n), y(n), vx(n), vy(n);
    st = cupic_get_particles(r, species, x.data(), y.data(), vx.data(), vy.data(), n, &n);
    if (st) return st;
    std::string fn = std::string(path) + ".dat";
    FILE *f = std::fopen(fn.c_str(), "w");
    if (!f) return set_err(&r->err, CUPIC_ERR_IO, "cannot write " + fn);
    for (int64_t k = 0; k < n; k++) std::fprintf(f, " %d %d \n", (int)k, (int)((double)y[k] / r->h.ds));
    std::fclose(f);
    return CUPIC_OK;
}

int cupic_mesh_snapshot(cupic_run *r, int which, const char *path)
{
    if (!r || !path) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    int nnx = r->h.ncx + 1, nny = r->h.ncy + 1;
    size_t nn = (size_t)nnx * nny;
    std::vector<double> v(nn);
    int st;
    if (which == CUPIC_MESH_PHI) {
        st = cupic_get_mesh(r, which, v.data(), (int64_t)nn);
    } else if (which == CUPIC_MESH_RHO || which == CUPIC_MESH_EX || which == CUPIC_MESH_EY) {
        std::vector<float> f(nn);
        st = cupic_get_mesh(r, which, f.data(), (int64_t)nn);
        for (size_t k = 0; k < nn; k++) v[k] = f[k];
    } else {
        return set_err(&r->err, CUPIC_ERR_ARG, "unknown mesh quantity");
    }
    if (st) return st;
    std::string fn = std::string(path) + ".dat";
    FILE *f = std::fopen(fn.c_str(), "w");
    if (!f) return set_err(&r->err, CUPIC_ERR_IO, "cannot write " + fn);
    for (int i = 0; i < nnx; i++) {
        for (int j = 0; j < nny; j++) std::fprintf(f, " %d %d %.17e \n", i, j, v[(size_t)i + (size_t)j * nnx]);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    return CUPIC_OK;
}

// ---- timing

int cupic_timing_enable(cupic_run *r, int32_t on)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    r->timing = on != 0;
    r->t_mask = (1u << kTClasses) - 1;
    return CUPIC_OK;
}

int cupic_timing_select(cupic_run *r, uint32_t class_mask)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    if (class_mask == 0 || (class_mask >> kTClasses) != 0)
        return set_err(&r->err, CUPIC_ERR_ARG, "class mask must select classes 0..7");
    r->timing = true;
    r->t_mask = class_mask;
    return CUPIC_OK;
}

int cupic_timing_read(cupic_run *r, double *ms, int64_t *launches, int32_t n)
{
    if (!r) return set_err(nullptr, CUPIC_ERR_ARG, "null run");
    HIPTRY(&r->err, hipStreamSynchronize(r->stream));
    for (auto &e : r->evs) {
        float v = 0.0f;
        if (hipEventElapsedTime(&v, e.a, e.b) == hipSuccess) {
            r->t_ms[e.cls] += v;
            r->t_n[e.cls] += 1;
        }
        r->pool.push_back(e.a);
        r->pool.push_back(e.b);
    }
    r->evs.clear();
    for (int k = 0; k < n && k < kTClasses; k++) {
        if (ms) ms[k] = r->t_ms[k];
        if (launches) launches[k] = r->t_n[k];
        r->t_ms[k] = 0.0;
        r->t_n[k] = 0;
    }
    return CUPIC_OK;
}

const char *cupic_timing_names(void) { return kTimeNames; }

// ---- batches

}  // extern "C"

namespace {

Target batch_target(cupic_batch *b)
{
    Target T{b->d_rds, b->d_plan, b->nblk, (int)b->runs.size(), b->stream, nullptr};
    T.off = b->plan_off;
    T.side = b->side;
    T.fork = b->ev_fork;
    T.join = b->ev_join;
    T.field = b->ev_field;
    // one scatter launch serves every run: the instance follows the
    // particles per cell over the whole batch
    int64_t part = 0, cells = 0;
    for (cupic_run *r : b->runs) {
        part += (int64_t)r->p.per_row * r->p.ncy;
        cells += (int64_t)r->p.ncx * r->p.ncy;
    }
    T.seg_stayers = part < (int64_t)kSegStayersPpc * cells;
    return T;
}

// One step of every run in the batch; t lives on the device and advances.
void batch_step_sequence(cupic_batch *b, const Target &T)
{
    launch_rho_convert(T.L(kShapeNodes), 1);
    seq_clock_fork(T, 0, 0.0);
    seq_solve(T, 0, b->mode, 0, 0, b->max_sweeps, b->check, true, b->mode == 3);
    launch_field(T.L(kShapeNodes), b->mode == 3);
    seq_inject_fork(T);
    launch_push_rank(T.L(kShapeCountTasks), b->mode == 3);
    seq_cc(T, 0, 0.0, 1, true);
}

}  // namespace

extern "C" {

// HIP failures while a batch is being built release what was built so far
#define BATCH_TRY(expr)                                                                 \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            if (b->stream) (void)hipStreamEndCapture(b->stream, &b->graph);             \
            cupic_batch_destroy(b);                                                     \
            return set_err(nullptr, CUPIC_ERR_HIP,                                      \
                           std::string(#expr) + ": " + hipGetErrorString(e_));          \
        }                                                                               \
    } while (0)

int cupic_batch_create(cupic_run *const *runs, int32_t nruns, int32_t use_graph,
                       cupic_batch **out)
{
    if (!runs || nruns < 1 || !out) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    *out = nullptr;
    cupic_batch *b = new cupic_batch();
    b->use_graph = use_graph != 0;
    b->mode = runs[0]->p.jacobi_mode;
    b->check = runs[0]->p.jacobi_check_every;
    b->max_sweeps = runs[0]->p.jacobi_max_sweeps;
    std::vector<RunDev> hs;
    for (int k = 0; k < nruns; k++) {
        cupic_run *r = runs[k];
        if (!r || !r->initialised || r->moved) {
            cupic_batch_destroy(b);
            return set_err(nullptr, CUPIC_ERR_STATE, "batch runs must be initialised and idle");
        }
        if (r->p.jacobi_mode != b->mode || r->p.jacobi_check_every != b->check ||
            r->p.jacobi_max_sweeps != b->max_sweeps) {
            cupic_batch_destroy(b);
            return set_err(nullptr, CUPIC_ERR_ARG, "batch runs must share the Jacobi schedule");
        }
        b->runs.push_back(r);
        hs.push_back(r->h);
    }
    // every run enters the batch with its charge accumulated and its clock at
    // the coming step's t
    for (cupic_run *r : b->runs) {
        if (!r->acc_valid) {
            BATCH_TRY(hipMemsetAsync(r->h.acc, 0, (size_t)r->h.nnodes * 8, r->stream));
            launch_deposit(target_of(r).L(kShapeTasks));
            r->acc_valid = true;
        }
        RunState s;
        int st = read_state(r, &s);
        if (st) { cupic_batch_destroy(b); return st; }
        s.t = r->t_next;
        st = write_state(r, s);
        if (st) { cupic_batch_destroy(b); return st; }
    }
    // row s: nruns + 1 block prefixes, then from plan_map_at(nruns) the
    // (run, local block) pair of each of the row's blocks
    std::vector<int> nbs((size_t)kShapeCount * nruns);
    size_t total = 0;
    for (int s = 0; s < kShapeCount; s++) {
        b->plan_off[s] = (int64_t)total;
        int acc = 0;
        for (int k = 0; k < nruns; k++) acc += (nbs[(size_t)s * nruns + k] = shape_blocks(hs[k], (Shape)s));
        b->nblk[s] = acc;
        total += plan_map_at(nruns) + 2 * (size_t)acc;
    }
    std::vector<int> plan(total, 0);
    for (int s = 0; s < kShapeCount; s++) {
        int *row = plan.data() + b->plan_off[s];
        int *map = row + plan_map_at(nruns);
        int acc = 0;
        for (int k = 0; k < nruns; k++) {
            row[k] = acc;
            const int nb = nbs[(size_t)s * nruns + k];
            for (int q = 0; q < nb; q++) {
                map[2 * (acc + q)] = k;
                map[2 * (acc + q) + 1] = q;
            }
            acc += nb;
        }
        row[nruns] = acc;
    }
    if (hipMalloc(&b->d_rds, hs.size() * sizeof(RunDev)) != hipSuccess ||
        hipMalloc(&b->d_plan, plan.size() * sizeof(int)) != hipSuccess) {
        cupic_batch_destroy(b);
        return set_err(nullptr, CUPIC_ERR_ALLOC, "batch allocation failed");
    }
    BATCH_TRY(hipMemcpy(b->d_rds, hs.data(), hs.size() * sizeof(RunDev), hipMemcpyHostToDevice));
    BATCH_TRY(hipMemcpy(b->d_plan, plan.data(), plan.size() * sizeof(int), hipMemcpyHostToDevice));
    BATCH_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BATCH_TRY(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
    BATCH_TRY(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    BATCH_TRY(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
    BATCH_TRY(hipEventCreateWithFlags(&b->ev_field, hipEventDisableTiming));
    BATCH_TRY(hipDeviceSynchronize());
    if (b->use_graph) {
        Target T = batch_target(b);
        BATCH_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        batch_step_sequence(b, T);
        BATCH_TRY(hipStreamEndCapture(b->stream, &b->graph));
        BATCH_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    *out = b;
    return CUPIC_OK;
}

int cupic_batch_launch(cupic_batch *b, int32_t nsteps)
{
    if (!b || nsteps < 0) return set_err(nullptr, CUPIC_ERR_ARG, "bad argument");
    Target T = batch_target(b);
    for (int k = 0; k < nsteps; k++) {
        if (b->use_graph) {
            HIPTRY(&b->err, hipGraphLaunch(b->exec, b->stream));
        } else {
            batch_step_sequence(b, T);
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_err(&b->err, CUPIC_ERR_HIP, std::string("batch step: ") + hipGetErrorString(e));
    for (cupic_run *r : b->runs) {
        for (int k = 0; k < nsteps; k++) r->t_next += r->p.dt;   // as the device clock
        r->acc_valid = true;
    }
    return CUPIC_OK;
}

int cupic_batch_wait(cupic_batch *b)
{
    if (!b) return set_err(nullptr, CUPIC_ERR_ARG, "null batch");
    HIPTRY(&b->err, hipStreamSynchronize(b->stream));
    return CUPIC_OK;
}

int cupic_batch_step(cupic_batch *b, int32_t nsteps)
{
    int st = cupic_batch_launch(b, nsteps);
    return st ? st : cupic_batch_wait(b);
}

int cupic_batch_destroy(cupic_batch *b)
{
    if (!b) return CUPIC_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->side) (void)hipStreamSynchronize(b->side);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->ev_field) (void)hipEventDestroy(b->ev_field);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->d_rds) (void)hipFree(b->d_rds);
    if (b->d_plan) (void)hipFree(b->d_plan);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CUPIC_OK;
}

const char *cupic_batch_last_error(const cupic_batch *b)
{
    if (b && !b->err.empty()) return b->err.c_str();
    return g_err.c_str();
}

}  // extern "C"
