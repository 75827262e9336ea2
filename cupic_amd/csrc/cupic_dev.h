// cupic_dev.h -- device-side data layout and shared device functions.
//
// Layout in HBM (per run):
//   particles   FP32 (x, y, vx, vy) as one 16-byte record per particle (AoS),
//               per species, two buffers (ping-pong), sorted by cell key
//               jc*ncx + ic with per-cell offsets; beside each buffer one
//               push-hint byte per particle slot.
//   grids       periodic in x, [ncy+1][ncx]: node (i, j) at j*ncx + i.
//               charge: int64 fixed point (2^-32 e); rhs: FP32 and FP64;
//               phi: FP32, three buffers (modes 0, 1) / FP64 master + FP32
//               correction ping-pong (mode 2) / FP64 (mode 3, direct);
//               E: float2 (Ex, Ey).
//   direct      mode 3 only: DFT and sine bases, eigenvalue tables, transformed
//               rows, per-slice partial sums.
//   sort meta   cnt[old cell][9 directions], base[new cell][10 slots],
//               far list of particles that moved more than one cell or were
//               injected.
//
// Every kernel is written for a batch of runs: blockIdx.x is mapped to
// (run, local block) through a prefix array, so the plain path is a batch
// of one.  All run-dependent control state lives in device memory
// (RunState) so a whole step can be captured in a HIP graph.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cupic {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kCmax = 16;           // max cells per wave task (tile_cells)
// Below this many particles per cell the scatter sums the stayers of a chunk
// that spans cells with one segmented scan (most chunks span cells there);
// above it, per cell as the wave leaves the cell.
constexpr int kSegStayersPpc = 128;
constexpr int kTileW = kCmax + 3;   // deposit tile width (node columns)
constexpr int kCountCells = 32;     // cells per wave strip of the counting pass (any tile)
constexpr int kScanCells = 1024;    // cells per scan block
constexpr int kRedSlots = 16;       // atomic slots per reduced quantity
constexpr int kResNodes = 8;        // FP64 residual: nodes per thread
constexpr int kFarBlocks = 64;      // far-list kernels: at most this many blocks per species
constexpr int kGM = 64;             // direct solve: GEMM output tile rows (4 waves x 16)
constexpr int kGN = 64;             // direct solve: GEMM output tile columns (4 chains per wave)
constexpr int kGK = 32;             // direct solve: GEMM k tile
constexpr int kGSlices = 512;       // direct solve: workgroups a product aims for (k slices)
constexpr int kGMaxSlices = 2;   // at most: each slice adds a partial grid to write and sum
// Direct solve by FFT: a transform of length n (ncx along x, 2 ncy along y)
// runs as one FFT of length n when n is a power of two, else (where the
// run's direct_method allows it) by Bluestein's chirp z-transform as two
// FFTs of the power of two P >= 2 n - 1.  Three kernel instances, for FFT
// lengths up to kFftS, kFftM and kFftL (LDS: two arrays of the length in
// complex doubles).
constexpr int kFftS = 1024;
constexpr int kFftM = 2048;
constexpr int kFftL = 4096;
__host__ __device__ constexpr bool pow2_len(int n) { return n >= 1 && (n & (n - 1)) == 0; }
// the FFT length that transforms n points (n itself, or Bluestein's P)
__host__ __device__ constexpr int fft_len(int n)
{
    if (pow2_len(n)) return n;
    int p = 1;
    while (p < 2 * n - 1) p *= 2;
    return p;
}
// 0: the dense products; 1, 2, 3: FFT, lengths up to kFftS, kFftM, kFftL.
// any_len: Bluestein's transform for lengths that are not powers of two
// (else those grids take the dense products).
__host__ __device__ constexpr int fft_class(int ncx, int ncy, bool any_len)
{
    const bool fits = ncx >= 2 && ncx <= kFftL && ncy >= 2 && ncy <= kFftL &&
                      (any_len || (pow2_len(ncx) && pow2_len(2 * ncy)));
    const int lx = fits ? fft_len(ncx) : 2 * kFftL;
    const int ly = fits ? fft_len(2 * ncy) : 2 * kFftL;
    const int l = lx > ly ? lx : ly;
    return l > kFftL ? 0 : (l <= kFftS ? 1 : (l <= kFftM ? 2 : 3));
}
// The FFT tables of a run, in complex doubles (re, im) from dft_tw:
// twiddles e^(-2 pi i t / Lx) (Lx = fft_len(ncx)) and e^(-2 pi i t / Ly)
// (Ly = fft_len(2 ncy)), the chirps w_n = e^(-i pi n^2 / N) for N = ncx and
// N = 2 ncy, and Bluestein's filter spectra FFT(conj w, wrapped) / Lx and
// / Ly (used only where the length is not a power of two).
struct FftTabs {
    int lx, ly;
    int tw_x, tw_y, ch_x, ch_y, bh_x, bh_y;  // offsets in complex doubles
    int total;
};
__host__ __device__ inline FftTabs fft_tabs(int ncx, int ncy)
{
    FftTabs t;
    t.lx = fft_len(ncx);
    t.ly = fft_len(2 * ncy);
    t.tw_x = 0;
    t.tw_y = t.lx;
    t.ch_x = t.tw_y + t.ly;
    t.ch_y = t.ch_x + ncx;
    t.bh_x = t.ch_y + 2 * ncy;
    t.bh_y = t.bh_x + t.lx;
    t.total = t.bh_y + t.ly;
    return t;
}
constexpr int kJacRows = 8;         // blocked Jacobi: rows per wave
constexpr int kJacWaves = 8;        // blocked Jacobi: waves per workgroup
constexpr int kJacThreads = kJacWaves * 64;
// Halo of the blocked Jacobi, at most halo - 1 sweeps per launch.  A wide
// halo means fewer launches but more redundant work per node: it pays for a
// single run, whose launches are short; a batch of many runs, whose
// launches fill the device many times over, takes the narrow one.
constexpr int kJacH = 16;
constexpr int kJacHNarrow = 10;
constexpr int kJacNarrowMinBlocks = 1024;   // wide-halo workgroups from which to go narrow
__host__ __device__ constexpr int jac_tx(int h) { return 64 - 2 * h; }
__host__ __device__ constexpr int jac_ty(int h) { return kJacRows * kJacWaves - 2 * h; }
// workgroups of a blocked-Jacobi launch with halo h over an ncx x ncy grid
__host__ __device__ constexpr int jac_blocks(int ncx, int ncy, int h)
{
    return ((ncx + jac_tx(h) - 1) / jac_tx(h)) * ((ncy - 1 + jac_ty(h) - 1) / jac_ty(h));
}
// A single run whose grid leaves compute units idle takes a wider halo
// still: the widest of these whose launch has at most one workgroup per
// compute unit (MI355X: 256), as the redundant work then runs on units that
// would have none, while each launch saved is about 7 us of fixed cost.
constexpr int kJacHSolo[3] = {28, 24, 20};
constexpr int kJacSoloMaxBlocks = 256;
// The halo of a single run's blocked Jacobi.
inline int jac_solo_halo(int ncx, int ncy)
{
    for (int h : kJacHSolo)
        if (jac_blocks(ncx, ncy, h) <= kJacSoloMaxBlocks) return h;
    return jac_blocks(ncx, ncy, kJacH) >= kJacNarrowMinBlocks ? kJacHNarrow : kJacH;
}

enum Dir : int { kDirFar = 9, kDirProbe = 10, kDirTop = 11, kDirNone = 15 };

struct FarRec {             // 32 B
    int32_t key;            // new cell
    int32_t idx;            // input index (injected: N + k)
    int32_t rank;           // stable rank within (key, far slot)
    int32_t pad;
    float x, y, vx, vy;     // final state
};

struct SpecConst {
    float qdtm;             // q dt / m
    float hkick;            // -q dt / (2 m)
    double q, m, sigma, dtin;
};

// Mutable per-run control state (device memory).
struct RunState {
    int cur;                        // current particle buffer
    int phi_cur;                    // current FP32 phi buffer (modes 0, 1)
    int N[2];                       // particles per species (current buffer)
    int N_new[2];
    int n_in[2];                    // injected this step
    int far_n[2];                   // fast movers in the far list (see far_entry)
    int tin_started;
    int step;
    double t;
    double tin[2];
    unsigned long long ninj[2];     // injected so far (RNG sequence)
    unsigned long long ninj_first[2];
    // solver
    int solve_mode, solve_min, solve_max, solve_check;
    double solve_tol;
    int solve_done;
    int phi_final;                  // phi buffer holding the solution
    int sweeps_done;
    double last_resid;
    long long sweeps_total;         // sweeps over all solves
    long long solves;
    // tallies (this step and running)
    unsigned long long probe[2], top[2];
    unsigned long long last_probe[2], last_top[2], last_injected[2];
    unsigned long long total_probe[2], total_top[2], total_injected[2];
    // counting pass: particles pushed in full (hint missing or overturnable)
    // this step, and running totals of those and of far movers
    unsigned long long undecided[2];
    unsigned long long total_undecided[2], total_far[2];
    unsigned long long particle_steps;  // sum over steps of the particles moved
    double probe_charge_total;
    unsigned int err;               // error bits, see kErr*
    unsigned int fin_ticket;        // k_far_scatter blocks done this step (the last ends it)
    unsigned int solve_ticket;      // k_residual blocks done (end: the last ends the solve)
};

enum : unsigned int {
    kErrFarOverflow = 1u,
    kErrCapacity = 2u,
    kErrInjectCap = 4u,
    kErrSortInvariant = 8u,
    kErrHint = 16u,                 // checked hint mode: a hint disagreed with the push
};

// Push hints (one byte per particle slot, written with the particle by the
// scatter): low 4 bits the direction 0..8 the particle will move if its
// velocity does not change, high 4 bits a code c >= 1 stating that the
// position it would reach lies at least 2^(c-16) cells from every wall of
// that cell; 0 means no hint.  The counting pass takes the direction from
// the hint whenever the strip's field cannot move the particle by that
// margin.
enum : int { kHintOff = 0, kHintOn = 1, kHintChecked = 2 };

// Immutable per-run description (device memory).
struct RunDev {
    int ncx, ncy, ncells, nnodes;
    int C, tpr;                     // cells per wave task, tasks per row
    int C1, tpr1;                   // the same for the counting pass
    int cap;                        // particles per species per buffer
    int far_cap, max_inject;
    int n_init;                     // initial particles per species
    int per_row;
    int max_sweeps_alloc;           // size of the per-sweep reduction slots
    int step_mode, step_min, step_max, step_check;
    double step_tol, init_tol;
    float inv_ds, Lx_f, Ly_f, dt_f;
    double ds, dt, Lx, Ly, eps0, phi_p;
    double b_interior;              // 2^-32 / (ds eps0): charge units -> rhs
    double rho_interior, rho_edge;  // 2^-32 / ds^3, 2^-32 / (ds^3 / 2)
    SpecConst sp[2];
    unsigned long long seed, run_id;
    unsigned int quirks;
    int hint_mode;                  // kHintOff / kHintOn / kHintChecked
    int count_stats;                // count fully pushed particles (one same-address
                                    // atomic per wave of the counting pass: off by default)
    float4 *part[2][2];             // [species][buffer], (x, y, vx, vy) per particle
    unsigned char *hint[2][2];      // [species][buffer], push hint per particle slot
    int *off[2][2];                 // [species][buffer], ncells + 1
    int *cnt[2];                    // [9][ncells] arrivals per (direction, old cell)
    int *base[2];                   // [10][ncells] slot bases per new cell
    int *far_cnt[2];                // ncells, far entries per new cell
    int *far_cur[2];                // ncells, placement cursors
    FarRec *far[2];                 // far_cap
    double *inj_tin[2];             // max_inject
    int *blk[2];                    // scan block sums / offsets
    unsigned long long *acc;        // nnodes, int64 charge
    float *rho;                     // nnodes
    float *b32;                     // nnodes, Jacobi rhs
    double *b64;                    // nnodes
    float *phi32[3];                // nnodes, modes 0/1 (see phi_at)
    double *phi64;                  // nnodes, mode 2 master
    float *d32[2];                  // nnodes, mode 2 correction
    float *r32;                     // nnodes, mode 2 correction rhs
    float2 *E;                      // nnodes
    unsigned int *sweep_max;        // (max_sweeps_alloc + 1) * kRedSlots (float bits)
    unsigned long long *resid;      // (max_sweeps_alloc + 2) * kRedSlots (double bits)
    int nsx, nsy;                   // mode 3: k slices of the x and y products
    double *dft_f;                  // mode 3: [ncx][ncx] forward real DFT basis
    double *dft_g;                  // mode 3: [ncx][ncx] inverse basis
    double *dft_s;                  // mode 3: [m][m] orthonormal sine basis (m = ncy - 1)
    double *dft_d;                  // mode 3: [ncx + m] eigenvalue parts per column, per row
    double *dft_z, *dft_y;          // mode 3: [m][ncx] transformed rows (two stages)
    double *dft_p;                  // mode 3: [max(nsx, nsy)][m][ncx] partial sums
    int dft_fft;                    // mode 3: the FFT instance, fft_class (0: the dense products)
    double *dft_tw;                 // mode 3, FFT: the tables of fft_tabs (re, im pairs)
    RunState *st;
};

// Blocks per species of the far-list kernels: grid-stride loops over the
// entries, one thread per entry of a usual list (max_inject is twice a
// step's injection plus a margin; fast movers are rare), at most
// kFarBlocks, so that a batch of many small runs launches few idle blocks.
__host__ __device__ inline int far_blocks(const RunDev &h)
{
    const int b = (h.max_inject + kBlock - 1) / kBlock;
    return b < 1 ? 1 : (b < kFarBlocks ? b : kFarBlocks);
}

// A species' far list holds the fast movers of the counting pass from slot
// 0 (at most far_mov_cap of them) and the injected particles from slot
// far_mov_cap (at most max_inject), so that the two can be written
// concurrently.  Its entries in order, movers then injected, are numbered
// v = 0 .. far_len - 1.
__host__ __device__ inline int far_mov_cap(const RunDev &h) { return h.far_cap - h.max_inject; }

__device__ inline int far_len(const RunDev &R, const RunState *st, int s)
{
    return min(st->far_n[s], far_mov_cap(R)) + st->n_in[s];
}

__device__ inline int far_entry(const RunDev &R, const RunState *st, int s, int v)
{
    const int fm = min(st->far_n[s], far_mov_cap(R));
    return v < fm ? v : far_mov_cap(R) + (v - fm);
}

// ---------------------------------------------------------------- helpers

// Pointers held in RunDev are generic; casting them to the global address
// space lets the compiler emit global (not flat) memory instructions, whose
// completion is tracked apart from LDS traffic.
#define CUPIC_GLOBAL __attribute__((address_space(1)))

template <class T>
__device__ __forceinline__ const CUPIC_GLOBAL T *gp(const T *p)
{
    return (const CUPIC_GLOBAL T *)p;
}

template <class T>
__device__ __forceinline__ CUPIC_GLOBAL T *gp(T *p)
{
    return (CUPIC_GLOBAL T *)p;
}

typedef float v4f __attribute__((ext_vector_type(4)));

// one particle record (x, y, vx, vy): a single 16-byte access per lane
__device__ __forceinline__ v4f ld4(const float4 *p, int i)
{
    return ((const CUPIC_GLOBAL v4f *)p)[i];
}

__device__ __forceinline__ void st4(float4 *p, int i, v4f v)
{
    ((CUPIC_GLOBAL v4f *)p)[i] = v;
}


// float2 is a class type that cannot be copied through an address-space
// reference; load it as a native two-element vector.
__device__ __forceinline__ float2 ld_f2(const float2 *p, int k)
{
    typedef float v2f __attribute__((ext_vector_type(2)));
    v2f v = ((const CUPIC_GLOBAL v2f *)p)[k];
    return make_float2(v.x, v.y);
}

// A batch launch's plan row: nruns + 1 block prefixes, then, from this
// (even) offset, the (run, local block) pair of every block, so that a block
// finds its run with one load instead of a search over the prefixes.
__host__ __device__ constexpr int plan_map_at(int nruns) { return (nruns + 2) & ~1; }

// FP32 potential buffers of modes 0 and 1.  A solve starts from buffer
// a = phi_cur and alternates between a and phi_alt(a): after `par` sweeps
// (mode 0) or blocked launches (mode 1) the iterate is in phi_at(a, par).
// The third buffer, phi_third(a), receives the iterate of a residual check
// that falls inside a blocked launch, in case the solve stops there.
__host__ __device__ inline int phi_alt(int a) { return a == 0 ? 1 : 0; }
__host__ __device__ inline int phi_at(int a, int par) { return (par & 1) ? phi_alt(a) : a; }
__host__ __device__ inline int phi_third(int a) { return 3 - a - phi_alt(a); }

// Reductions go to kRedSlots addresses per quantity (block b uses slot
// b % kRedSlots) so that thousands of blocks do not serialise on one
// atomic; readers take the max over the slots.
__device__ __forceinline__ double read_resid(const RunDev &R, int c)
{
    double m = 0.0;
#pragma unroll
    for (int q = 0; q < kRedSlots; q++)
        m = fmax(m, __longlong_as_double((long long)__hip_atomic_load(
                        &R.resid[c * kRedSlots + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
    return m;
}

__device__ __forceinline__ float read_sweep_max(const RunDev &R, int s)
{
    float m = 0.0f;
#pragma unroll
    for (int q = 0; q < kRedSlots; q++) m = fmaxf(m, __uint_as_float(R.sweep_max[s * kRedSlots + q]));
    return m;
}

// End of a solve (one thread): sweep count and last residual, where the
// solution is, and the running totals.
__device__ inline void solver_end_work(const RunDev &R, RunState *st, int last_check, int par)
{
    if (!st->solve_done) {
        int s = st->solve_mode == 3 ? 0 : st->solve_max;
        st->sweeps_done = s;
        if (st->solve_mode == 0)
            st->last_resid = s > 0 ? (double)read_sweep_max(R, s - 1) : 0.0;
        else
            st->last_resid = read_resid(R, last_check);
        st->phi_final = phi_at(st->phi_cur, par);
        st->solve_done = 1;
    }
    if (st->solve_mode != 2) st->phi_cur = st->phi_final;
    st->sweeps_total += st->sweeps_done;
    st->solves += 1;
}

// Node charge a (int64, 2^-32 of a macro-particle charge) of row j -> rho
// (volume ds^3, ds^3/2 on the two Dirichlet rows, the intent of
// src/mesh.cu:242-268) and the Jacobi right-hand side ds^2 rho / eps0
// (src/mesh.cu:327), 0 on the Dirichlet rows.
__device__ __forceinline__ void charge_to_rhs(const RunDev &R, long long a, int j, float *rho,
                                              double *b)
{
    const bool edge = (j == 0 || j == R.ncy);
    const double q = (double)a;
    *rho = (float)(q * (edge ? R.rho_edge : R.rho_interior));
    *b = edge ? 0.0 : q * R.b_interior;
}

// Prologue of every kernel that takes (rds, start, nruns): the run R of this
// workgroup, its block index `local` within the run, and the run state st.
// start[0] is 0, so a launch for one run needs no load from start.
#define RUN_PROLOGUE()                                                  \
    const int b_ = blockIdx.x;                                          \
    const int2 rl_ = nruns == 1 ? make_int2(0, b_)                      \
                                : reinterpret_cast<const int2 *>(       \
                                      start + plan_map_at(nruns))[b_];  \
    const int r_ = __builtin_amdgcn_readfirstlane(rl_.x); /* uniform */ \
    const RunDev &R = rds[r_];                                          \
    const int local = __builtin_amdgcn_readfirstlane(rl_.y);            \
    RunState *st = R.st;                                                \
    (void)local; (void)st;

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float wave_max_f(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double wave_max_d(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// ---------------------------------------------------------------- Philox-4x64-10

__device__ __forceinline__ void philox4x64_10(unsigned long long c[4], unsigned long long k0,
                                              unsigned long long k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        unsigned long long hi0 = __umul64hi(0xD2E7470EE14C6C93ULL, c[0]);
        unsigned long long lo0 = 0xD2E7470EE14C6C93ULL * c[0];
        unsigned long long hi1 = __umul64hi(0xCA5A826395121157ULL, c[2]);
        unsigned long long lo1 = 0xCA5A826395121157ULL * c[2];
        unsigned long long n0 = hi1 ^ c[1] ^ k0;
        unsigned long long n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B97F4A7C15ULL;
        k1 += 0xBB67AE8584CAA73BULL;
    }
}

// counter (index, species | purpose << 8, 0, 0), key (seed, run id)
__device__ __forceinline__ void draw4(const RunDev &R, int species, int purpose,
                                      unsigned long long index, unsigned long long out[4])
{
    out[0] = index;
    out[1] = (unsigned long long)species | ((unsigned long long)purpose << 8);
    out[2] = 0;
    out[3] = 0;
    philox4x64_10(out, R.seed, R.run_id);
}

__device__ __forceinline__ double u01_open(unsigned long long r)
{
    return ((double)(r >> 12) + 0.5) * 0x1p-52;
}

// ---------------------------------------------------------------- cell keys

// Geometry a particle kernel needs, copied once into registers: reading it
// through the RunDev reference inside a loop that stores particles would
// reload it after every store (the compiler cannot rule out aliasing).
struct Geo {
    int ncx, ncy;
    float inv_ds, Lx_f, Ly_f, dt_f;
    float dtinv;                    // dt / ds in FP32 (push hints)
};

__device__ __forceinline__ Geo geo_of(const RunDev &R)
{
    Geo g;
    g.ncx = R.ncx;
    g.ncy = R.ncy;
    g.inv_ds = R.inv_ds;
    g.Lx_f = R.Lx_f;
    g.Ly_f = R.Ly_f;
    g.dt_f = R.dt_f;
    g.dtinv = R.dt_f * R.inv_ds;
    return g;
}

// Cell of an FP32 position and its in-cell fractions: one product, one
// floor and clamps, evaluated without contraction.
template <class G>
__device__ __forceinline__ void cell_of(const G &R, float x, float y, float &fx, float &fy,
                                        int &ic, int &jc)
{
    float tx = x * R.inv_ds;
    float ty = y * R.inv_ds;
    float bx = floorf(tx), by = floorf(ty);
    fx = tx - bx;
    fy = ty - by;
    // inside the grid the clamps below change nothing; outside (or NaN) they
    // pin the cell to the edge and the fractions into [0, 1].  The wave
    // branches around them (a ballot) so that they cost nothing when no lane
    // needs them, which is almost always.
    const bool out = !(tx >= 0.0f && tx < (float)R.ncx && ty >= 0.0f && ty < (float)R.ncy);
    if (__builtin_expect(__ballot(out) != 0ull, 0) && out) {
        bx = fminf(fmaxf(bx, 0.0f), (float)(R.ncx - 1));
        by = fminf(fmaxf(by, 0.0f), (float)(R.ncy - 1));
        fx = fminf(fmaxf(tx - bx, 0.0f), 1.0f);
        fy = fminf(fmaxf(ty - by, 0.0f), 1.0f);
    }
    ic = (int)bx;
    jc = (int)by;
}

// CIC gather from interleaved (Ex, Ey), fixed operation order.
// CIC interpolation of (Ex, Ey) from the four corners of a cell, weights
// (1-fx)(1-fy), fx(1-fy), (1-fx)fy, fx fy accumulated in that order with
// fused multiply-adds.
__device__ __forceinline__ void cic_interp(float fx, float fy, float2 e00, float2 e10,
                                           float2 e01, float2 e11, float &ex, float &ey)
{
    float gx = 1.0f - fx, gy = 1.0f - fy;
    float w00 = gx * gy, w10 = fx * gy, w01 = gx * fy, w11 = fx * fy;
    ex = fmaf(w11, e11.x, fmaf(w01, e01.x, fmaf(w10, e10.x, w00 * e00.x)));
    ey = fmaf(w11, e11.y, fmaf(w01, e01.y, fmaf(w10, e10.y, w00 * e00.y)));
}

template <class G>
__device__ __forceinline__ void gather(const G &R, const float2 *__restrict__ Eg, float fx,
                                       float fy, int ic, int jc, float &ex, float &ey)
{
    int i1 = (ic + 1 == R.ncx) ? 0 : ic + 1;
    int k00 = jc * R.ncx + ic, k10 = jc * R.ncx + i1;
    int k01 = k00 + R.ncx, k11 = k10 + R.ncx;
    float2 e00 = ld_f2(Eg, k00), e10 = ld_f2(Eg, k10), e01 = ld_f2(Eg, k01), e11 = ld_f2(Eg, k11);
    cic_interp(fx, fy, e00, e10, e01, e11, ex, ey);
}

// The leapfrog update v += (q dt / m) E, x += dt v, each as one fused
// multiply-add.
__device__ __forceinline__ void leapfrog(float qdtm, float dt, float ex, float ey, float &x,
                                         float &y, float &vx, float &vy)
{
    vx = fmaf(qdtm, ex, vx);
    vy = fmaf(qdtm, ey, vy);
    x = fmaf(dt, vx, x);
    y = fmaf(dt, vy, y);
}

struct Pushed {
    float x, y, vx, vy;
    float fx, fy;      // fractions in the new cell
    int oic, ojc;      // old cell
    int nic, njc;      // new cell
    int di;            // x displacement in cells (-1, 0, 1 for near movers)
    int dir;           // 0..8 near, kDirFar, kDirProbe, kDirTop
};

// New cell, in-cell fractions and direction of a particle pushed to
// (o.x, o.y) from cell (oic, ojc), without branches: the direction of an
// absorbed particle (y < 0: probe, y >= Ly: top) overrides the cell's, and
// its cell fields, clamped by cell_of, are not read.
template <class G>
__device__ __forceinline__ void classify(const G &R, Pushed &o, int oic, int ojc)
{
    float fx, fy;
    int nic, njc;
    cell_of(R, o.x, o.y, fx, fy, nic, njc);
    o.fx = fx; o.fy = fy; o.nic = nic; o.njc = njc;
    int di = nic - oic;
    const int dj = njc - ojc;
    di = di == R.ncx - 1 ? -1 : (di == -(R.ncx - 1) ? 1 : di);
    o.di = di;
    const float ty = o.y * R.inv_ds;
    const bool near = di >= -1 && di <= 1 && dj >= -1 && dj <= 1;
    const int dcell = near ? (di + 1) + 3 * (dj + 1) : (int)kDirFar;
    o.dir = ty < 0.0f ? (int)kDirProbe : (ty >= (float)R.ncy ? (int)kDirTop : dcell);
}

// Leapfrog push of one particle (src/particles.cu:110-231 semantics),
// periodic wrap into [0, Lx) (src/cc.cu:775-810, Q6), absorption at y < 0
// (probe) and y >= Ly (top) (src/cc.cu:182-185), new cell and direction.
template <class G>
__device__ __forceinline__ Pushed push_one(const G &R, const float2 *__restrict__ E,
                                           float qdtm, float x, float y, float vx, float vy)
{
    Pushed o;
    float fx, fy;
    int oic, ojc;
    cell_of(R, x, y, fx, fy, oic, ojc);
    float ex, ey;
    gather(R, E, fx, fy, oic, ojc, ex, ey);
    float nx = x, ny = y, nvx = vx, nvy = vy;
    leapfrog(qdtm, R.dt_f, ex, ey, nx, ny, nvx, nvy);
    if (nx < 0.0f) nx = nx + R.Lx_f;
    if (nx >= R.Lx_f) nx = nx - R.Lx_f;
    o.x = nx; o.y = ny; o.vx = nvx; o.vy = nvy;
    o.oic = oic; o.ojc = ojc;
    classify(R, o, oic, ojc);
    return o;
}

// 16-bit-per-axis fixed-point weights; every particle carries 2^32 units.
// Fractions lie in [0, 1], so the weights are at most 2^16; the mask (a
// no-op on them) tells the compiler so, and products of weights become
// full-rate 24-bit multiplies.
__device__ __forceinline__ void fixed_weights(float fx, float fy, unsigned &wx, unsigned &wy)
{
    wx = (unsigned)rintf(fx * 65536.0f) & 0x1FFFFu;
    wy = (unsigned)rintf(fy * 65536.0f) & 0x1FFFFu;
}

// ---------------------------------------------------------------- push hints
//
// Bound behind the hints.  A particle stored in cell (ic, jc) with in-cell
// fraction fx = tx - ic (exact: tx = fl(x inv_ds), Sterbenz) and velocity vx
// gets the predicted fraction F = fl(fma(vx, fl(dt inv_ds), fx)).  Its next
// push computes vx' = fl(fma(qdtm, ex, vx)), x' = fl(fma(dt, vx', x)),
// tx' = fl(x' inv_ds) and the cell floor(tx').  With u = 2^-24, positions
// inside a grid of at most M = max(ncx, ncy) cells and |F| < 2,
//   |(tx' - ic) - F| <= |dt qdtm inv_ds| |ex| + (3.2 M + 7) u,
// and |ex| <= (1 + 7u) max |E| over the four gather nodes.  The hint's
// distance m of F to the walls of its cell floor(F) is exact up to u
// (F - floor(F) is exact but for -1/2 < F < 0, 1 - that within u/2), so
// m >= 2^(c-16) - u.  Hence, with
//   thr >= |dt qdtm inv_ds| Emax (1 + 2^-14) + 16 (M + 1) u,
// a code with 2^(c-16) > thr puts tx' strictly inside the hinted cell: the
// push lands there, with no wrap in x (hinted columns 1 .. ncx-2) and no
// absorption in y (rows 0 .. ncy-1).  A non-finite F gives no hint; a
// non-finite field gives no threshold.

// Hint of a particle stored in cell (ic, jc) with fractions (fx, fy) and
// velocity (vx, vy).
template <class G>
__device__ __forceinline__ unsigned push_hint(const G &R, float fx, float fy, float vx, float vy,
                                              int ic, int jc)
{
    const float fxp = fmaf(vx, R.dtinv, fx), fyp = fmaf(vy, R.dtinv, fy);
    if (!(fxp >= -1.0f && fxp < 2.0f && fyp >= -1.0f && fyp < 2.0f)) return 0u;  // NaN: no hint
    const float kx = floorf(fxp), ky = floorf(fyp);
    const int di = (int)kx, dj = (int)ky;
    const int gi = ic + di, gj = jc + dj;
    if (gi < 1 || gi > R.ncx - 2 || gj < 0 || gj > R.ncy - 1) return 0u;
    const float gx = fxp - kx, gy = fyp - ky;
    const float m = fminf(fminf(gx, 1.0f - gx), fminf(gy, 1.0f - gy));
    // floor(log2 m) + 16 from the exponent field (0 <= m <= 1/2; zero and
    // subnormal m give a negative code)
    const int c = (int)((__float_as_uint(m) >> 23) & 0xFFu) - 127 + 16;
    if (c < 1) return 0u;
    return ((unsigned)min(c, 15) << 4) | (unsigned)((di + 1) + 3 * (dj + 1));
}

// Smallest hint code the counting pass may trust for a strip whose field
// nodes have max(|Ex|, |Ey|) = emax (16: none).
__device__ __forceinline__ int hint_code_needed(float qdtm, float dt, float inv_ds, int ncx,
                                                int ncy, float emax, bool finite)
{
    const float a = fabsf(qdtm) * dt * inv_ds * emax * (1.0f + 0x1p-14f);
    const float thr = a + 16.0f * (float)(max(ncx, ncy) + 1) * 0x1p-24f;
    if (!finite || !(thr < 0.25f)) return 16;
    const int e = (int)((__float_as_uint(thr) >> 23) & 0xFFu) - 127;  // thr is normal here
    return max(1, e + 17);
}

}  // namespace cupic
