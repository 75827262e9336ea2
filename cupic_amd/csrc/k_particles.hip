// k_particles.hip -- particle kernels.
//
// The stable cell-key counting sort of each step needs the number of
// particles per (old cell, direction) before any particle can be stored:
//
//   pass 1 (k_push_rank): classify each particle by the direction it moves
//     (9 neighbour cells, far, probe, top) and count per (old cell,
//     direction); write the counts, the far list and the absorption
//     tallies.  The direction comes from the particle's one-byte push hint
//     whenever the strip's field is too weak to overturn it, so most
//     particles cost one byte here; the rest are loaded and pushed.
//   scan (k_cc.hip): new-cell offsets and per-slot bases.
//   pass 2 (k_scatter): the push of every particle and its rank within
//     (old cell, direction), then a store of the new state and its hint at
//     the sorted position and the int64 fixed-point CIC charge of the new
//     positions (the next step's deposition), reduced per wave by cell in
//     registers and per wave-strip in LDS before HBM atomics.  Each wave
//     owns a strip of whole old cells, so ranks are deterministic.
//
// Order within a new cell: by source slot (the 3x3 neighbourhood of old
// cells in row-major order, then far movers and injected particles), and by
// input order within a slot -- a stable counting sort by (cell, slot).
#include "cupic_launch.h"

namespace cupic {


constexpr double kPi = 3.1415926535897932;  // src/stdh.h:26

// ---------------------------------------------------------------- init

// Initial load (src/init.cu:136, 159-179): per_row particles per y row,
// x ~ U(0,1) Lx, y = row ds + U(0,1) ds, v ~ N(0, sqrt(kT/m)).  FP64 draws
// from Philox (counter = particle index, species, purpose 0), rounded to
// FP32: x wrapped into [0, Lx), y moved by whole ulps into its own row.
__global__ __launch_bounds__(kBlock) void k_create(const RunDev *rds, const int *start,
                                                   int nruns)
{
    RUN_PROLOGUE();
    int bps = (R.n_init + kBlock - 1) / kBlock;
    int s = local / bps;
    int g = (local - s * bps) * kBlock + threadIdx.x;
    if (s > 1 || g >= R.n_init) return;
    int row = g / R.per_row;
    unsigned long long r[4];
    draw4(R, s, 0, (unsigned long long)g, r);
    double sigma = R.sp[s].sigma;
    double u0 = u01_open(r[0]), u1 = u01_open(r[1]);
    double u2 = u01_open(r[2]), u3 = u01_open(r[3]);
    double rad = sqrt(-2.0 * log(u2));
    double th = 2.0 * kPi * u3;
    double x = u0 * R.Lx;
    double y = (double)row * R.ds + u1 * R.ds;
    double vx = sigma * (rad * cos(th));
    double vy = sigma * (rad * sin(th));
    float xf = (float)x;
    if (xf < 0.0f) xf = xf + R.Lx_f;
    if (xf >= R.Lx_f) xf = xf - R.Lx_f;
    float yf = (float)y;
    for (int q = 0; q < 64 && floorf(yf * R.inv_ds) > (float)row; q++) yf = nextafterf(yf, -INFINITY);
    for (int q = 0; q < 64 && floorf(yf * R.inv_ds) < (float)row; q++) yf = nextafterf(yf, INFINITY);
    st4(R.part[s][1], g, v4f{xf, yf, (float)vx, (float)vy});
}

constexpr int kInitMaxNcx = 4096;

// Stable sort of each generated row by cell column (buffer 1 -> buffer 0)
// and the row's cell offsets.  One workgroup per (species, row); ranks
// within a 256-particle chunk are counted directly in LDS.  Init only.
__global__ __launch_bounds__(kBlock) void k_init_sort(const RunDev *rds, const int *start,
                                                      int nruns)
{
    RUN_PROLOGUE();
    __shared__ int hist[kInitMaxNcx];
    __shared__ int runc[kInitMaxNcx];
    __shared__ int chunk[kBlock];
    __shared__ int part[kBlock];
    int s = local / R.ncy;
    int row = local - s * R.ncy;
    int ncx = R.ncx;
    int pr = R.per_row;
    int first = row * pr;
    const float4 *IN = R.part[s][1];
    float4 *OUT = R.part[s][0];
    for (int c = threadIdx.x; c < ncx; c += kBlock) { hist[c] = 0; runc[c] = 0; }
    __syncthreads();
    for (int q = threadIdx.x; q < pr; q += kBlock) {
        float fx, fy;
        int ic, jc;
        v4f p = ld4(IN, first + q);
        cell_of(R, p.x, p.y, fx, fy, ic, jc);
        atomicAdd(&hist[ic], 1);
    }
    __syncthreads();
    // exclusive scan of hist (ncx entries) -> column starts
    int per = (ncx + kBlock - 1) / kBlock;
    int lo = threadIdx.x * per, hi = min(lo + per, ncx);
    int sum = 0;
    for (int c = lo; c < hi; c++) sum += hist[c];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int t = 0; t < kBlock; t++) { int v = part[t]; part[t] = acc; acc += v; }
    }
    __syncthreads();
    int acc = part[threadIdx.x];
    for (int c = lo; c < hi; c++) { int v = hist[c]; hist[c] = acc; acc += v; }
    __syncthreads();
    int *off = R.off[s][0];
    for (int c = threadIdx.x; c < ncx; c += kBlock) off[row * ncx + c] = first + hist[c];
    if (row == R.ncy - 1 && threadIdx.x == 0) off[R.ncells] = R.n_init;
    for (int base = 0; base < pr; base += kBlock) {
        int q = base + threadIdx.x;
        bool valid = q < pr;
        int ic = -1;
        float fx, fy;
        int jc;
        v4f p = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        if (valid) {
            p = ld4(IN, first + q);
            cell_of(R, p.x, p.y, fx, fy, ic, jc);
        }
        chunk[threadIdx.x] = ic;
        __syncthreads();
        int rank = 0, after = 0;
        if (valid) {
            for (int t = 0; t < (int)threadIdx.x; t++) rank += (chunk[t] == ic);
            for (int t = threadIdx.x + 1; t < kBlock; t++) after += (chunk[t] == ic);
            int dest = first + hist[ic] + runc[ic] + rank;
            st4(OUT, dest, p);
            R.hint[s][0][dest] = 0;
        }
        __syncthreads();
        if (valid && after == 0) runc[ic] += rank + 1;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- deposit tile

// Adds the fixed-point CIC charge of the wave's depositing lanes to a
// wave-private LDS node tile.  Lanes are grouped by tile cell; each group
// reduces (sum wx | sum wy << 32) and (sum wx*wy | count << 48), from which
// the four corner sums follow exactly:
//   c00 = 2^32 n - 2^16 (Swx + Swy) + Swxwy, c10 = 2^16 Swx - Swxwy,
//   c01 = 2^16 Swy - Swxwy,                 c11 = Swxwy.
__device__ __forceinline__ void wave_deposit(unsigned long long *tile, int TW, int dkey,
                                             unsigned wx, unsigned wy, int lane)
{
    unsigned long long v1 = 0, v2 = 0;
    if (dkey >= 0) {
        v1 = (unsigned long long)wx | ((unsigned long long)wy << 32);
        v2 = (unsigned long long)wx * (unsigned long long)wy + (1ull << 48);
    }
    unsigned long long pending = __ballot(dkey >= 0);
    while (pending) {
        int leader = __ffsll((long long)pending) - 1;
        int lk = __shfl(dkey, leader, kWave);
        bool mine = (dkey == lk);
        unsigned long long m = __ballot(mine);
        unsigned long long a1 = wave_sum_u64(mine ? v1 : 0ull);
        unsigned long long a2 = wave_sum_u64(mine ? v2 : 0ull);
        if (lane == leader) {
            long long sx = (long long)(a1 & 0xFFFFFFFFull);
            long long sy = (long long)(a1 >> 32);
            long long n = (long long)(a2 >> 48);
            long long sxy = (long long)(a2 & ((1ull << 48) - 1));
            long long c11 = sxy;
            long long c10 = (sx << 16) - sxy;
            long long c01 = (sy << 16) - sxy;
            long long c00 = (n << 32) - ((sx + sy) << 16) + sxy;
            atomicAdd(&tile[lk], (unsigned long long)c00);
            atomicAdd(&tile[lk + 1], (unsigned long long)c10);
            atomicAdd(&tile[lk + TW], (unsigned long long)c01);
            atomicAdd(&tile[lk + TW + 1], (unsigned long long)c11);
        }
        pending &= ~m;
    }
}

// Tile rows j0-1 .. j0+2, columns col0-1 .. col0+nc+1 (wrapped) -> HBM.  The
// tile holds charge counts of one species; electrons (neg) enter negated.
__device__ __forceinline__ void flush_tile(const RunDev &R, const unsigned long long *tile,
                                           int TW, int j0, int col0, int lane, bool neg)
{
    for (int q = lane; q < 4 * TW; q += kWave) {
        unsigned long long v = tile[q];
        if (v == 0ull) continue;
        if (neg) v = 0ull - v;
        int row = j0 - 1 + q / TW;
        int col = col0 - 1 + (q % TW);
        if (col < 0) col += R.ncx;
        if (col >= R.ncx) col -= R.ncx;
        if (row < 0 || row > R.ncy) continue;
        atomicAdd(&R.acc[row * R.ncx + col], v);  // global: acc lives in HBM
    }
}

struct TaskInfo {
    int s, j0, col0, nc;
    bool ok;
};

// Wave task of workgroup `local` for strips of C cells, tpr strips per row.
__device__ __forceinline__ TaskInfo task_of(const RunDev &R, int local, int C, int tpr)
{
    TaskInfo t;
    // the wave index is uniform: keep the task and everything derived from
    // it in scalar registers
    int task = local * (kBlock / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int per_species = R.ncy * tpr;
    t.ok = task < 2 * per_species;
    t.s = task / per_species;
    int t2 = task - t.s * per_species;
    t.j0 = t2 / tpr;
    t.col0 = (t2 - t.j0 * tpr) * C;
    t.nc = min(C, R.ncx - t.col0);
    return t;
}

// Standalone deposition of the current (sorted) state: charge_deposition
// (src/mesh.cu:15-51) in int64 fixed point.
__global__ __launch_bounds__(kBlock) void k_deposit(const RunDev *rds, const int *start,
                                                    int nruns)
{
    RUN_PROLOGUE();
    __shared__ unsigned long long tiles[kBlock / kWave][4 * kTileW];
    TaskInfo T = task_of(R, local, R.C, R.tpr);
    if (!T.ok) return;
    int lane = threadIdx.x & (kWave - 1);
    unsigned long long *tile = tiles[threadIdx.x >> 6];
    int TW = T.nc + 3;
    for (int q = lane; q < 4 * TW; q += kWave) tile[q] = 0ull;
    wave_lds_sync();
    int cur = st->cur;
    const int *off = R.off[T.s][cur];
    int c0 = T.j0 * R.ncx + T.col0;
    int p0 = off[c0], p1 = off[c0 + T.nc];
    const float4 *P = R.part[T.s][cur];
    long long sgn = T.s == 0 ? -1 : 1;
    for (int i0 = p0; i0 < p1; i0 += kWave) {
        int i = i0 + lane;
        int dkey = -1;
        unsigned wx = 0, wy = 0;
        if (i < p1) {
            float fx, fy;
            int ic, jc;
            v4f p = ld4(P, i);
            cell_of(R, p.x, p.y, fx, fy, ic, jc);
            int lc = ic - T.col0 + 1, lr = jc - (T.j0 - 1);
            fixed_weights(fx, fy, wx, wy);
            if (lr >= 0 && lr <= 2 && lc >= 0 && lc <= T.nc + 1) {
                dkey = lr * TW + lc;
            } else {  // outside the strip: straight to HBM
                int i1 = ic + 1 == R.ncx ? 0 : ic + 1;
                long long gx = 65536 - (long long)wx, gy = 65536 - (long long)wy;
                atomicAdd(&R.acc[jc * R.ncx + ic], (unsigned long long)(sgn * gx * gy));
                atomicAdd(&R.acc[jc * R.ncx + i1], (unsigned long long)(sgn * (long long)wx * gy));
                atomicAdd(&R.acc[(jc + 1) * R.ncx + ic], (unsigned long long)(sgn * gx * (long long)wy));
                atomicAdd(&R.acc[(jc + 1) * R.ncx + i1],
                          (unsigned long long)(sgn * (long long)wx * (long long)wy));
            }
        }
        wave_deposit(tile, TW, dkey, wx, wy, lane);
    }
    wave_lds_sync();
    flush_tile(R, tile, TW, T.j0, T.col0, lane, T.s == 0);
}

// Initial velocity stagger, src/init.cu:793-833: v -= dt/2 * qE/m.
__global__ __launch_bounds__(kBlock) void k_half_kick(const RunDev *rds, const int *start,
                                                      int nruns)
{
    RUN_PROLOGUE();
    int bps = (R.n_init + kBlock - 1) / kBlock;
    int s = local / bps;
    if (s > 1) return;
    int cur = st->cur;
    int n = st->N[s];
    float hk = R.sp[s].hkick;
    float4 *P = R.part[s][cur];
    for (int i = (local - s * bps) * kBlock + threadIdx.x; i < n; i += bps * kBlock) {
        float fx, fy, ex, ey;
        int ic, jc;
        v4f p = ld4(P, i);
        cell_of(R, p.x, p.y, fx, fy, ic, jc);
        gather(R, R.E, fx, fy, ic, jc, ex, ey);
        p.z = fmaf(hk, ex, p.z);
        p.w = fmaf(hk, ey, p.w);
        st4(P, i, p);
        R.hint[s][cur][i] = 0;
    }
}

// ---------------------------------------------------------------- push passes

// popcount of m over the lanes below this one
__device__ __forceinline__ int mbcnt64(unsigned long long m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// value of lane l-1 (lane 0 keeps its own), a VALU wave_shr:1 move
__device__ __forceinline__ int dpp_prev_int(int v)
{
    return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xF, 0xF, false);
}

// Adds one particle group's fixed-point charge (reduced sums, see
// wave_deposit) to the four corners of tile cell key.
// Packing: a1 = sum wx | sum wy << 32, a2 = sum wx*wy | count << 48; exact
// while a group holds fewer than 65536 particles.
__device__ __forceinline__ void tile_add(unsigned long long *tile, int TW, int key,
                                         unsigned long long a1, unsigned long long a2)
{
    unsigned long long sx = a1 & 0xFFFFFFFFull;
    unsigned long long sy = a1 >> 32;
    unsigned long long n = a2 >> 48;
    unsigned long long sxy = a2 & ((1ull << 48) - 1);
    atomicAdd(&tile[key], (n << 32) - ((sx + sy) << 16) + sxy);
    atomicAdd(&tile[key + 1], (sx << 16) - sxy);
    atomicAdd(&tile[key + TW], (sy << 16) - sxy);
    atomicAdd(&tile[key + TW + 1], sxy);
}

// One particle's charge: the four corner products of its 16-bit weights
// (at most 2^16 each, so full-rate 24-bit multiplies) into tile cell key.
__device__ __forceinline__ void tile_add_one(unsigned long long *tile, int TW, int key,
                                             unsigned wx, unsigned wy)
{
    // the masks are no-ops (weights lie in [0, 2^16]) that keep every factor
    // visibly below 2^24
    const unsigned long long gx = (65536u - wx) & 0x1FFFFu, gy = (65536u - wy) & 0x1FFFFu;
    atomicAdd(&tile[key], gx * gy);
    atomicAdd(&tile[key + 1], (unsigned long long)wx * gy);
    atomicAdd(&tile[key + TW], gx * (unsigned long long)wy);
    atomicAdd(&tile[key + TW + 1], (unsigned long long)wx * (unsigned long long)wy);
}

// value + the DPP move kCtrl of it (0 where the source lane is out of its
// row or the row is masked off): one step of a whole-wave scan, which needs
// no lane conditions
template <int kCtrl, int kRowMask>
__device__ __forceinline__ unsigned dpp_add_u32(unsigned v)
{
    return v + (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, kRowMask, 0xF, true);
}

// The same step for a 64-bit value held as two 32-bit words: the low words
// add with a carry into the high ones.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ void dpp_add_u64w(unsigned &lo, unsigned &hi)
{
    const unsigned dl = (unsigned)__builtin_amdgcn_update_dpp(0, (int)lo, kCtrl, kRowMask, 0xF, true);
    const unsigned dh = (unsigned)__builtin_amdgcn_update_dpp(0, (int)hi, kCtrl, kRowMask, 0xF, true);
    const unsigned nl = lo + dl;
    hi = hi + dh + (nl < lo ? 1u : 0u);
    lo = nl;
}

// Inclusive prefix sum over the whole wave (the total lands in lane 63):
// Hillis-Steele within rows of 16 lanes with row_shr, then across rows with
// row_bcast:15 and row_bcast:31.  The zero fill of the DPP moves gives each
// lane exactly the partial sums it must add, so no lane conditions are
// needed.  a1's halves (sums of at most 60000 weights of 2^16 or less) are
// added as 32-bit values.
__device__ __forceinline__ void wave_scan_full(unsigned &sx, unsigned &sy,
                                               unsigned long long &a2)
{
    unsigned lo = (unsigned)a2, hi = (unsigned)(a2 >> 32);
    sx = dpp_add_u32<0x111, 0xF>(sx); sy = dpp_add_u32<0x111, 0xF>(sy);
    dpp_add_u64w<0x111, 0xF>(lo, hi);
    sx = dpp_add_u32<0x112, 0xF>(sx); sy = dpp_add_u32<0x112, 0xF>(sy);
    dpp_add_u64w<0x112, 0xF>(lo, hi);
    sx = dpp_add_u32<0x114, 0xF>(sx); sy = dpp_add_u32<0x114, 0xF>(sy);
    dpp_add_u64w<0x114, 0xF>(lo, hi);
    sx = dpp_add_u32<0x118, 0xF>(sx); sy = dpp_add_u32<0x118, 0xF>(sy);
    dpp_add_u64w<0x118, 0xF>(lo, hi);
    sx = dpp_add_u32<0x142, 0xA>(sx); sy = dpp_add_u32<0x142, 0xA>(sy);
    dpp_add_u64w<0x142, 0xA>(lo, hi);
    sx = dpp_add_u32<0x143, 0xC>(sx); sy = dpp_add_u32<0x143, 0xC>(sy);
    dpp_add_u64w<0x143, 0xC>(lo, hi);
    a2 = (unsigned long long)lo | ((unsigned long long)hi << 32);
}

// Reduces the per-lane stayer sums of old cell cl over the wave and adds
// them to the tile cell of cl (stayers keep their cell).
__device__ __forceinline__ void flush_stayers(unsigned long long *tile, int TW, int cl,
                                              unsigned long long a1, unsigned long long a2,
                                              int lane)
{
    unsigned sx = (unsigned)a1, sy = (unsigned)(a1 >> 32);
    wave_scan_full(sx, sy, a2);
    a1 = (unsigned long long)sx | ((unsigned long long)sy << 32);
    if (lane == kWave - 1 && (a2 >> 48) != 0ull) tile_add(tile, TW, 1 * TW + (cl + 1), a1, a2);
}

// LDS tile of E on the two node rows of a strip: columns col0 .. col0+nc.
__device__ __forceinline__ void load_E_tile(const RunDev &R, float2 *Et, int j0, int col0,
                                            int nc, int lane)
{
    int w = nc + 1;
    for (int q = lane; q < 2 * w; q += kWave) {
        int r = q >= w ? 1 : 0, c = col0 + (q - r * w);  // q < 2 w: two node rows
        if (c >= R.ncx) c -= R.ncx;
        Et[q] = ld_f2(R.E, (j0 + r) * R.ncx + c);
    }
}

// The gather of push_one with the four nodes taken from the strip tile.
__device__ __forceinline__ Pushed push_tile(const Geo &R, const float2 *Et, int nc, int col0,
                                            int j0, const float2 *__restrict__ E, float qdtm,
                                            float x, float y, float vx, float vy, bool &in_strip)
{
    float fx, fy;
    int oic, ojc;
    cell_of(R, x, y, fx, fy, oic, ojc);
    int cl = oic - col0;
    in_strip = (ojc == j0) && cl >= 0 && cl < nc;
    // A particle outside the strip means the sort invariant is broken: it is
    // reported (kErrSortInvariant) and reads a clamped tile entry instead of
    // a global load, which would stall the prefetch of every chunk.
    cl = min(max(cl, 0), nc - 1);
    int w = nc + 1;
    float2 e00 = Et[cl], e10 = Et[cl + 1], e01 = Et[w + cl], e11 = Et[w + cl + 1];
    float ex, ey;
    cic_interp(fx, fy, e00, e10, e01, e11, ex, ey);
    (void)E;
    Pushed o;
    float nx = x, ny = y, nvx = vx, nvy = vy;
    leapfrog(qdtm, R.dt_f, ex, ey, nx, ny, nvx, nvy);
    if (nx < 0.0f) nx = nx + R.Lx_f;
    if (nx >= R.Lx_f) nx = nx - R.Lx_f;
    o.x = nx; o.y = ny; o.vx = nvx; o.vy = nvy;
    o.oic = oic; o.ojc = ojc;
    classify(R, o, oic, ojc);
    return o;
}

constexpr int kQueue = 512;   // pass 1: ring of slots awaiting a full push, per wave
constexpr int kFarList = 64;  // pass 1: far movers listed per wave strip

// Pass 1: one wave per strip of whole old cells counts arrivals per (old
// cell, direction), appends far movers to the far list and tallies
// absorptions.  A particle whose hint the strip's field cannot overturn is
// counted from its hint byte alone (movers need their old cell, found from
// the strip's cell offsets).  The slots of the others are queued in LDS and
// pushed 64 at a time, so the full push runs on full waves.  With
// kHintChecked every particle is pushed and each trusted hint is compared
// with its push.
__device__ __forceinline__ void count_task(const RunDev &R, RunState *st, const TaskInfo &T,
                                           int *cntw, int *gonew, int *bnd, int *queue,
                                           int *farq, float2 *Et, int lane)
{
    const int s = T.s, j0 = T.j0, col0 = T.col0, nc = T.nc;
    const Geo G = geo_of(R);
    const int far_cap = far_mov_cap(R);  // fast movers' part of the far list
    int *const far_cnt = R.far_cnt[s];
    FarRec *const farl = R.far[s];
    const float qdtm = R.sp[s].qdtm;
    const int mode = R.hint_mode;
    const int c0 = j0 * R.ncx + col0;
    // The strip's E tile (2 (nc + 1) nodes, two per lane) does not depend on
    // the buffer parity: its loads go out first, beside the parity's, and
    // the first hint words follow the strip's offsets (scalar loads, so that
    // they do not wait for the E tile).
    static_assert(kCountCells + 1 <= kWave, "prologue staging");
    const int w = nc + 1;
    float2 e2[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int q = k * kWave + lane;
        e2[k] = make_float2(0.0f, 0.0f);
        if (q < 2 * w) {
            const int r = q >= w ? 1 : 0;  // q < 2 w: two node rows
            int c = col0 + (q - r * w);
            if (c >= G.ncx) c -= G.ncx;
            e2[k] = ld_f2(R.E, (j0 + r) * G.ncx + c);
        }
    }
    const int cur = __builtin_amdgcn_readfirstlane(gp(st)->cur);
    const CUPIC_GLOBAL int *off = gp(cur ? R.off[s][1] : R.off[s][0]);
    const int p0u = __builtin_amdgcn_readfirstlane(off[c0]);
    const int p1u = __builtin_amdgcn_readfirstlane(off[c0 + nc]);
    const float4 *PIN = cur ? R.part[s][1] : R.part[s][0];
    const CUPIC_GLOBAL unsigned char *HIN = gp(cur ? R.hint[s][1] : R.hint[s][0]);
    const CUPIC_GLOBAL unsigned *HW = (const CUPIC_GLOBAL unsigned *)HIN;
    // Hints are read as words of four slots, lane l covering slots
    // w + 4l .. w + 4l + 3 of the 256-slot window starting at w; the words of
    // the next window are in flight while one is processed.  Every lane
    // loads a clamped, in-bounds word whatever the mode, so that the count
    // of loads in flight is the same on every path and the wait for one
    // window does not wait for the next; bytes of slots outside the strip,
    // and all bytes unless the mode trusts hints, are ignored.
    auto load_word = [&](int w0) -> unsigned {
        return HW[max(0, min(w0 + 4 * lane, p1u - 1)) >> 2];
    };
    const int a0 = p0u & ~3;
    unsigned wa = load_word(a0);
    // lane k <= nc: first slot of strip cell k
    const int bv = lane <= nc ? off[c0 + lane] : 0x7FFFFFFF;
    // E tile of the strip and its largest component
    float em = 0.0f;
    bool fin = true;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        if (k * kWave + lane < 2 * w) {
            Et[k * kWave + lane] = e2[k];
            const float ax = fabsf(e2[k].x), ay = fabsf(e2[k].y);
            fin = fin && ax < INFINITY && ay < INFINITY;  // false for NaN
            em = fmaxf(em, fmaxf(ax, ay));
        }
    }
    em = wave_max_f(em);
    const int c_need =
        hint_code_needed(qdtm, G.dt_f, G.inv_ds, G.ncx, G.ncy, em, __ballot(!fin) == 0ull);
    // a hint byte is trusted when its code (high nibble) is at least c_need:
    // code + code_add carries into bit 4 of the byte (c_need 16: never)
    const unsigned code_add = mode == kHintOn ? (unsigned)(16 - c_need) * 0x01010101u : 0u;
    for (int q = lane; q < nc * 9; q += kWave) cntw[q] = 0;
    for (int q = lane; q < nc; q += kWave) gonew[q] = 0;
    if (lane <= nc) bnd[lane] = bv;
    wave_lds_sync();
    const unsigned long long lt = (1ull << lane) - 1ull;
    unsigned long long probe_n = 0, top_n = 0, full_n = 0;
    bool bad = false, wrong = false;

    // full push of queued slot i (lanes with ok false take no part)
    // Full push of slot i (lanes with ok false take no part): counts, checks
    // and the absorption tallies.  Far movers are only listed here (slot
    // numbers in LDS); their records are written after the loop, which keeps
    // the loop free of HBM stores and atomics.
    int nfar = 0;
    auto push_full = [&](int i, bool ok) {
        int dir = kDirNone;
        const v4f p = ld4(PIN, i);
        const unsigned hh = HIN[i];
        if (ok) {
            bool in_strip;
            const Pushed o = push_tile(G, Et, nc, col0, j0, R.E, qdtm, p.x, p.y, p.z, p.w, in_strip);
            dir = o.dir;
            if (dir < 9 && !in_strip) {  // not in this strip: the sort invariant is broken
                dir = kDirFar;
                bad = true;
            }
            const int cl = o.oic - col0;
            if (dir != 4 && cl >= 0 && cl < nc) {
                if (dir < 9) atomicAdd(&cntw[__mul24(cl, 9) + dir], 1);
                else atomicAdd(&gonew[cl], 1);
            }
            if (mode == kHintChecked && (int)(hh >> 4) >= c_need && dir != (int)(hh & 15u))
                wrong = true;
        }
        const unsigned long long fm = __ballot(dir == kDirFar);
        if (dir == kDirFar) {
            const int k = nfar + mbcnt64(fm);
            if (k < kFarList) farq[k] = i;
        }
        nfar += __popcll(fm);
        probe_n += __popcll(__ballot(dir == kDirProbe));
        top_n += __popcll(__ballot(dir == kDirTop));
    };
    // Far record of slot i (the same push again), to the run's far list.
    auto emit_far = [&](int i, bool ok) {
        int dir = kDirNone;
        Pushed o = {};
        if (ok) {
            const v4f p = ld4(PIN, i);
            bool in_strip;
            o = push_tile(G, Et, nc, col0, j0, R.E, qdtm, p.x, p.y, p.z, p.w, in_strip);
            dir = o.dir;
            if (dir < 9 && !in_strip) dir = kDirFar;
        }
        const unsigned long long fm = __ballot(dir == kDirFar);
        if (!fm) return;
        const int leader = __ffsll((long long)fm) - 1;
        int fb = 0;
        if (lane == leader) fb = atomicAdd(&st->far_n[s], __popcll(fm));
        fb = __shfl(fb, leader, kWave);
        if (dir == kDirFar) {
            const int slot = fb + __popcll(fm & lt);
            if (slot < far_cap) {
                CUPIC_GLOBAL FarRec *f = gp(farl) + slot;
                const int key = o.njc * G.ncx + o.nic;
                f->key = key;
                f->idx = i;
                f->rank = 0;
                f->pad = 0;
                f->x = o.x; f->y = o.y; f->vx = o.vx; f->vy = o.vy;
                atomicAdd(&far_cnt[key], 1);
            } else {
                atomicOr(&st->err, kErrFarOverflow);
            }
        }
    };

    int cb = 0;  // strip cell of the window's first slot
    int qh = 0;  // queue: ring of kQueue slots, head qh, qn entries
    int qn = 0;
    auto window = [&](int w0, unsigned word) {
        while (cb + 1 < nc && __builtin_amdgcn_readlane(bv, cb + 1) <= w0) cb++;
        // The four hint bytes are classified together, one flag per byte
        // (bit 0): valid (slot in the strip), use (code >= c_need, as
        // code + 16 - c_need >= 16), mov (trusted and direction != 4).
        unsigned valid = 0x01010101u;
        if (w0 < p0u || w0 + 4 * kWave > p1u) {  // the strip's first or last window
            valid = 0u;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int i = w0 + 4 * lane + q;
                valid |= (i >= p0u && i < p1u ? 1u : 0u) << (8 * q);
            }
        }
        const unsigned use = (((word >> 4) & 0x0F0F0F0Fu) + code_add) >> 4 & valid;
        const unsigned mov = (((word & 0x0F0F0F0Fu) ^ 0x04040404u) + 0x0F0F0F0Fu) >> 4 & use;
        const unsigned full = valid & ~use;
        // movers are counted by (strip cell, direction); a window without
        // any (most of the slow species') skips the cells and the atomics
        if (__ballot(mov != 0u)) {
            int cl[4];
#pragma unroll
            for (int q = 0; q < 4; q++) cl[q] = cb;
            for (int k = cb + 1; k < nc; k++) {
                const int bk = __builtin_amdgcn_readlane(bv, k);
                if (bk >= w0 + 4 * kWave) break;
#pragma unroll
                for (int q = 0; q < 4; q++) cl[q] += (w0 + 4 * lane + q >= bk);
            }
#pragma unroll
            for (int q = 0; q < 4; q++)
                if ((mov >> (8 * q)) & 1u)
                    atomicAdd(&cntw[__mul24(cl[q], 9) + (int)((word >> (8 * q)) & 15u)], 1);
        }
        if (__ballot(full != 0u)) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const bool f = ((full >> (8 * q)) & 1u) != 0u;
                const unsigned long long fq = __ballot(f);
                if (fq == 0ull) continue;
                if (f) queue[(qh + qn + mbcnt64(fq)) & (kQueue - 1)] = w0 + 4 * lane + q;
                qn += __popcll(fq);
                full_n += __popcll(fq);
            }
        }
        // at most kWave - 1 + 4 kWave entries: the ring never overfills
        while (qn >= kWave) {
            wave_lds_sync();
            push_full(queue[(qh + lane) & (kQueue - 1)], true);
            qh = (qh + kWave) & (kQueue - 1);
            qn -= kWave;
        }
    };
    for (int w0 = a0; w0 < p1u; w0 += 8 * kWave) {
        const unsigned wb = load_word(w0 + 4 * kWave);
        window(w0, wa);
        wa = load_word(w0 + 8 * kWave);
        if (w0 + 4 * kWave < p1u) window(w0 + 4 * kWave, wb);
    }
    if (qn > 0) {
        wave_lds_sync();
        push_full(lane < qn ? queue[(qh + lane) & (kQueue - 1)] : p0u, lane < qn);
    }
    // far records: from the list, or, if it overflowed, by pushing the whole
    // strip again (a strip with more than kFarList far movers is rare)
    wave_lds_sync();
    if (nfar <= kFarList) {
        for (int b0 = 0; b0 < nfar; b0 += kWave)
            emit_far(b0 + lane < nfar ? farq[b0 + lane] : p0u, b0 + lane < nfar);
    } else {
        for (int i0 = p0u; i0 < p1u; i0 += kWave) emit_far(i0 + lane, i0 + lane < p1u);
    }
    wave_lds_sync();
    // stayers = population - movers - departures (probe, top, far)
    for (int cl = lane; cl < nc; cl += kWave) {
        int moved = gonew[cl];
#pragma unroll
        for (int d = 0; d < 9; d++)
            if (d != 4) moved += cntw[cl * 9 + d];
        cntw[cl * 9 + 4] = bnd[cl + 1] - bnd[cl] - moved;
    }
    wave_lds_sync();
    CUPIC_GLOBAL int *cnt = gp(R.cnt[s]);
    const int ncells = R.ncells;
    // direction by direction: consecutive cells, coalesced (nc < kWave)
    if (lane < nc) {
#pragma unroll
        for (int d = 0; d < 9; d++) cnt[d * ncells + c0 + lane] = cntw[__mul24(lane, 9) + d];
    }
    if (lane == 0) {
        if (probe_n) atomicAdd(&st->probe[s], probe_n);
        if (top_n) atomicAdd(&st->top[s], top_n);
        if (R.count_stats && full_n) atomicAdd(&st->undecided[s], full_n);
    }
    if (__ballot(bad) && lane == 0) atomicOr(&st->err, kErrSortInvariant);
    if (__ballot(wrong) && lane == 0) atomicOr(&st->err, kErrHint);
}

// Pass 2: one wave per strip replays the identical push, ranks every near
// particle within (old cell, direction) with each counter started at its
// slot base, stores the particle and its hint at the sorted position and
// deposits its charge.
template <bool kSegStayers>
__device__ __forceinline__ void scatter_task(const RunDev &R, RunState *st, const TaskInfo &T,
                                             int *cntw, float2 *Et, unsigned long long *tile,
                                             int lane)
{
    const int s = T.s, j0 = T.j0, col0 = T.col0, nc = T.nc;
    const Geo G = geo_of(R);
    const int cap = R.cap;
    const int ncells = R.ncells;
    const float2 *E = R.E;
    const float qdtm = R.sp[s].qdtm;
    const int TW = nc + 3;
    const int c0 = j0 * R.ncx + col0;
    // The strip's E tile (2 (nc + 1) nodes) and slot bases (9 nc) do not
    // depend on the buffer parity: their loads go out first, beside the
    // parity's, and the first chunk's records follow the strip's offsets,
    // so that a task waits for three round trips, not six, before its first
    // particle.  Counters start at the base of the slot each (cell,
    // direction) fills.
    static_assert(2 * (kCmax + 1) <= kWave && 9 * kCmax <= 3 * kWave, "prologue staging");
    float2 ev = make_float2(0.0f, 0.0f);
    if (lane < 2 * (nc + 1)) {
        const int r = lane > nc ? 1 : 0;
        int c = col0 + (lane - r * (nc + 1));
        if (c >= R.ncx) c -= R.ncx;
        ev = ld_f2(E, (j0 + r) * R.ncx + c);
    }
    int bv[3];
    {
        const CUPIC_GLOBAL int *base = gp(R.base[s]);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int q = k * kWave + lane;
            const int cl = q / 9, d = q - 9 * cl;
            const int di = d % 3 - 1, jj = j0 + d / 3 - 1;
            int ii = col0 + cl + di;
            if (ii < 0) ii += R.ncx;
            if (ii >= R.ncx) ii -= R.ncx;
            bv[k] = 0;
            if (q < 9 * nc && jj >= 0 && jj < R.ncy)
                bv[k] = base[(8 - d) * ncells + jj * R.ncx + ii];
        }
    }
    const int cur = __builtin_amdgcn_readfirstlane(gp(st)->cur);
    const CUPIC_GLOBAL int *off = gp(cur ? R.off[s][1] : R.off[s][0]);
    const int p0 = __builtin_amdgcn_readfirstlane(off[c0]);
    const int p1 = __builtin_amdgcn_readfirstlane(off[c0 + nc]);
    const float4 *PIN = cur ? R.part[s][1] : R.part[s][0];
    float4 *POUT = cur ? R.part[s][0] : R.part[s][1];
    CUPIC_GLOBAL unsigned char *HOUT = gp(cur ? R.hint[s][0] : R.hint[s][1]);
    v4f nxt = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    if (p0 + lane < p1) nxt = ld4(PIN, p0 + lane);
    if (lane < 2 * (nc + 1)) Et[lane] = ev;
#pragma unroll
    for (int k = 0; k < 3; k++)
        if (k * kWave + lane < 9 * nc) cntw[k * kWave + lane] = bv[k];
    for (int q = lane; q < 4 * TW; q += kWave) tile[q] = 0ull;
    wave_lds_sync();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const unsigned long long upto = lt | (1ull << lane);
    bool bad = false;
    unsigned long long acc1 = 0ull, acc2 = 0ull;  // stayer sums of old cell acc_cl
    int acc_cl = -1, acc_n = 0;
    // The next chunk is in flight while one is processed.  A chunk's stores
    // are issued during the next chunk, after its prefetch: loads and stores
    // share one wait counter here, so the wait for a prefetch also waits for
    // every store issued before it, and stores one chunk old have completed.
    v4f prec = v4f{0.0f, 0.0f, 0.0f, 0.0f};  // the previous chunk's store
    int pdest = -1;
    unsigned phint = 0u;
    for (int i0 = p0; i0 < p1; i0 += kWave) {
        const int i = i0 + lane;
        const bool valid = i < p1;
        const v4f cur4 = nxt;
        const float x = cur4.x, y = cur4.y, vx = cur4.z, vy = cur4.w;
        if (i + kWave < p1) nxt = ld4(PIN, i + kWave);
        if (pdest >= 0) {
            st4(POUT, pdest, prec);
            HOUT[pdest] = (unsigned char)phint;
        }
        pdest = -1;
        bool in_strip;
        Pushed o = push_tile(G, Et, nc, col0, j0, E, qdtm, x, y, vx, vy, in_strip);
        int dir = valid ? o.dir : (int)kDirNone;
        int cl = o.oic - col0;
        if (dir < 9 && !in_strip) {  // not in this strip: the sort invariant is broken
            dir = kDirFar;
            bad = true;
        }
        // rank within (old cell, direction); lanes of one old cell are contiguous
        int rank = 0;
        int cl0 = 0;
        const int clp = dpp_prev_int(cl);
        const bool head = valid && (lane == 0 || cl != clp);
        const unsigned long long heads = __ballot(head);
        const unsigned long long near_m = __ballot(dir < 9);
        const bool one_cell = (heads == 1ull);
        if (one_cell) {
            // the whole chunk lies in one old cell: per direction present,
            // rank = counter + earlier lanes with that direction
            cl0 = __builtin_amdgcn_readfirstlane(cl);
            unsigned long long pend = near_m;
            while (pend) {
                int d = __builtin_amdgcn_readlane(dir, __ffsll((long long)pend) - 1);
                unsigned long long m = __ballot(dir == d);
                int base_c = cntw[cl0 * 9 + d];
                if (dir == d) rank = base_c + mbcnt64(m);
                if (lane == 63 - __clzll((long long)m)) cntw[cl0 * 9 + d] = base_c + __popcll(m);
                pend &= ~m;
            }
        } else {
            unsigned long long mymask = 0;
#pragma unroll
            for (int dd = 0; dd < 9; dd++) {
                unsigned long long m = __ballot(dir == dd);
                if (dir == dd) mymask = m;
            }
            const int seg_start = valid ? 63 - __clzll((long long)(heads & upto)) : lane;
            bool last = false;
            if (dir < 9) {
                unsigned long long after = heads & ~upto;
                unsigned long long hi = after ? ((1ull << (__ffsll((long long)after) - 1)) - 1ull) : ~0ull;
                unsigned long long seg = hi & ~((1ull << seg_start) - 1ull);
                unsigned long long grp = mymask & seg;
                rank = cntw[__mul24(cl, 9) + dir] + __popcll(grp & lt);
                last = (grp & ~upto) == 0ull;
            }
            wave_lds_sync();
            if (dir < 9 && last) cntw[__mul24(cl, 9) + dir] = rank + 1;
        }
        wave_lds_sync();
        unsigned wx = 0, wy = 0;
        if (dir < 9) {
            int dest = rank;  // counters hold slot base + rank
            if ((unsigned)dest < (unsigned)cap) {
                prec = v4f{o.x, o.y, o.vx, o.vy};
                pdest = dest;
                phint = push_hint(G, o.fx, o.fy, o.vx, o.vy, o.nic, o.njc);
            } else {
                bad = true;
            }
            fixed_weights(o.fx, o.fy, wx, wy);
        }
        const bool stay = (dir == 4);
        unsigned long long v1 = 0, v2 = 0;
        if (dir < 9) {
            v1 = (unsigned long long)wx | ((unsigned long long)wy << 32);
            v2 = (unsigned long long)wx * (unsigned long long)wy + (1ull << 48);
        }
        // stayers: per-lane sums while the wave stays in one old cell,
        // reduced once per cell (or every 60000 particles); a chunk that
        // spans cells feeds each cell segment in turn; movers add their own
        const unsigned long long stay_m = __ballot(stay);
        if (one_cell) {
            const int ns = __popcll(stay_m);
            if (cl0 != acc_cl || acc_n + ns >= 60000) {
                if (acc_cl >= 0) flush_stayers(tile, TW, acc_cl, acc1, acc2, lane);
                acc1 = 0ull;
                acc2 = 0ull;
                acc_cl = cl0;
                acc_n = 0;
            }
            if (stay) { acc1 += v1; acc2 += v2; }
            acc_n += ns;
        } else if (kSegStayers) {
            // a chunk that spans cells: the running sums are flushed; the
            // stayer sums of each cell segment are its last lane's prefix
            // over the whole wave minus the prefix just before the segment
            // (read from that lane), which the last lane adds to its cell
            if (acc_cl >= 0) flush_stayers(tile, TW, acc_cl, acc1, acc2, lane);
            acc1 = 0ull;
            acc2 = 0ull;
            acc_cl = -1;
            acc_n = 0;
            const int sstart = valid ? 63 - __clzll((long long)(heads & upto)) : lane;
            unsigned px = stay ? (unsigned)v1 : 0u, py = stay ? (unsigned)(v1 >> 32) : 0u;
            unsigned long long p2 = stay ? v2 : 0ull;
            wave_scan_full(px, py, p2);
            const int src = max(sstart - 1, 0);
            unsigned bx = (unsigned)__shfl((int)px, src, kWave);
            unsigned by = (unsigned)__shfl((int)py, src, kWave);
            unsigned bl = (unsigned)__shfl((int)(unsigned)p2, src, kWave);
            unsigned bh = (unsigned)__shfl((int)(unsigned)(p2 >> 32), src, kWave);
            if (sstart == 0) bx = by = bl = bh = 0u;
            const unsigned long long s1 =
                (unsigned long long)(px - bx) | ((unsigned long long)(py - by) << 32);
            const unsigned long long s2 = p2 - (((unsigned long long)bh << 32) | bl);
            const bool tail =
                valid && (lane == kWave - 1 || i + 1 >= p1 || ((heads >> (lane + 1)) & 1ull));
            if (tail && cl >= 0 && cl < nc && (s2 >> 48) != 0ull)
                tile_add(tile, TW, 1 * TW + (cl + 1), s1, s2);
        } else {
            unsigned long long rem = heads;
            while (rem) {
                const int h = __ffsll((long long)rem) - 1;
                rem &= rem - 1ull;
                const unsigned long long end =
                    rem ? ((1ull << (__ffsll((long long)rem) - 1)) - 1ull) : ~0ull;
                const unsigned long long M = end & ~((1ull << h) - 1ull);
                const int c = __builtin_amdgcn_readlane(cl, h);
                if (c < 0 || c >= nc) continue;  // outside the strip: no stayers
                const int ns = __popcll(stay_m & M);
                if (c != acc_cl || acc_n + ns >= 60000) {
                    if (acc_cl >= 0) flush_stayers(tile, TW, acc_cl, acc1, acc2, lane);
                    acc1 = 0ull;
                    acc2 = 0ull;
                    acc_cl = c;
                    acc_n = 0;
                }
                if (stay && ((M >> lane) & 1ull)) { acc1 += v1; acc2 += v2; }
                acc_n += ns;
            }
        }
        if (dir < 9 && !stay) {
            int key = __mul24(o.njc - (j0 - 1), TW) + (cl + o.di + 1);
            tile_add_one(tile, TW, key, wx, wy);
        }
    }
    if (pdest >= 0) {
        st4(POUT, pdest, prec);
        HOUT[pdest] = (unsigned char)phint;
    }
    if (acc_cl >= 0) flush_stayers(tile, TW, acc_cl, acc1, acc2, lane);
    wave_lds_sync();
    flush_tile(R, tile, TW, j0, col0, lane, s == 0);
    if (__ballot(bad) && lane == 0) atomicOr(&st->err, kErrSortInvariant);
}

// end_solve: block 0 first ends the step's direct solve, whose residual the
// field kernel reduced (solver_end_work), which saves a launch.
__global__ __launch_bounds__(kBlock) void k_push_rank(const RunDev *rds, const int *start,
                                                      int nruns, int end_solve)
{
    RUN_PROLOGUE();
    if (end_solve && local == 0 && threadIdx.x == 0) solver_end_work(R, st, 0, 0);
    __shared__ int cnts[kBlock / kWave][kCountCells * 9];
    __shared__ int gone[kBlock / kWave][kCountCells];
    __shared__ int bnds[kBlock / kWave][kCountCells + 1];
    __shared__ int queues[kBlock / kWave][kQueue];
    __shared__ int farqs[kBlock / kWave][kFarList];
    __shared__ float2 etiles[kBlock / kWave][2 * (kCountCells + 1)];
    TaskInfo T = task_of(R, local, R.C1, R.tpr1);
    if (!T.ok) return;
    int w = threadIdx.x >> 6;
    count_task(R, st, T, cnts[w], gone[w], bnds[w], queues[w], farqs[w], etiles[w],
               threadIdx.x & (kWave - 1));
}

// Far list, after the scan: each entry takes a place in its cell's far
// region (any order) and writes its list index into the first word of that
// slot of the output buffer, which the far scatter overwrites last.  Blocks
// 0 .. 2 far_blocks - 1 of the scatter launch (far_blocks per species): the
// work depends on the scan alone, as the scatter's does, and the scatter
// does not touch far slots.
__device__ __forceinline__ void far_place(const RunDev &R, RunState *st, int local)
{
    const int bps = far_blocks(R);
    const int s = local / bps;
    const int F = far_len(R, st, s);
    const int cur = __builtin_amdgcn_readfirstlane(gp(st)->cur);
    float4 *const out = cur ? R.part[s][0] : R.part[s][1];
    for (int v = (local - s * bps) * kBlock + threadIdx.x; v < F; v += bps * kBlock) {
        const int e = far_entry(R, st, s, v);
        int key = R.far[s][e].key;
        int slot = atomicAdd(&R.far_cur[s][key], 1);
        int pos = R.base[s][9 * R.ncells + key] + slot;
        // the slot is addressed as a float4 (64-bit offset): 4 * pos as an
        // int word index would overflow past 2^29 slots
        if ((unsigned)pos < (unsigned)R.cap)
            reinterpret_cast<int *>(out + pos)[0] = e;  // first word
    }
}

template <bool kSegStayers>
__global__ __launch_bounds__(kBlock) void k_scatter(const RunDev *rds, const int *start,
                                                    int nruns)
{
    RUN_PROLOGUE();
    __shared__ int cnts[kBlock / kWave][kCmax * 9];
    __shared__ float2 etiles[kBlock / kWave][2 * (kCmax + 1)];
    __shared__ unsigned long long tiles[kBlock / kWave][4 * kTileW];
    const int nfar = 2 * far_blocks(R);
    if (local < nfar) {
        far_place(R, st, local);
        return;
    }
    TaskInfo T = task_of(R, local - nfar, R.C, R.tpr);
    if (!T.ok) return;
    int w = threadIdx.x >> 6;
    scatter_task<kSegStayers>(R, st, T, cnts[w], etiles[w], tiles[w], threadIdx.x & (kWave - 1));
}

// ---------------------------------------------------------------- launchers

#define LAUNCH_K(kern, L, ...)                                                         \
    do {                                                                               \
        if ((L).nblocks > 0)                                                           \
            hipLaunchKernelGGL(kern, dim3((L).nblocks), dim3(kBlock), 0, (L).stream,   \
                               (L).rds, (L).start, (L).nruns, ##__VA_ARGS__);          \
    } while (0)

void launch_create(const Launch &L) { LAUNCH_K(k_create, L); }
void launch_init_sort(const Launch &L) { LAUNCH_K(k_init_sort, L); }
void launch_deposit(const Launch &L) { LAUNCH_K(k_deposit, L); }
void launch_half_kick(const Launch &L) { LAUNCH_K(k_half_kick, L); }
void launch_push_rank(const Launch &L, int end_solve) { LAUNCH_K(k_push_rank, L, end_solve); }
void launch_scatter(const Launch &L, bool seg_stayers)
{
    if (seg_stayers)
        LAUNCH_K(k_scatter<true>, L);
    else
        LAUNCH_K(k_scatter<false>, L);
}

}  // namespace cupic
