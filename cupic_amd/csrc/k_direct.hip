// k_direct.hip -- direct FP64 solve of the Jacobi system (Poisson mode 3).
//
// The Jacobi iteration of src/mesh.cu:297-381 converges to the solution of
// the 5-point system 4 phi - (sum of the four neighbours) = b on rows
// 1..ncy-1, periodic in x, with the Dirichlet rows 0 (probe) and ncy.  This
// mode computes that fixed point directly, as four FP64 matrix products on
// the matrix cores (m = ncy - 1 interior rows):
//
//   1. R^ = R F: a real DFT along x of the right-hand side rows, with the
//      Dirichlet rows moved into the first and last of them (the known
//      neighbours of rows 1 and m).  F [ncx][ncx] holds cos(2 pi i k / ncx)
//      for k = 0 .. ncx/2 and -sin(2 pi i k / ncx) for k = 1 .. (ncx-1)/2:
//      the ncx real coefficients of a real row.
//   2. Y = S R^, scaled by 1 / mu(l, k): per wavenumber k the rows satisfy
//      the tridiagonal system -x(j-1) + (4 - 2 cos(2 pi k / ncx)) x(j)
//      - x(j+1) = r(j) with x(0) = x(m+1) = 0, which the orthonormal sine
//      basis S(l, j) = sqrt(2 / (m+1)) sin(pi (l+1)(j+1) / (m+1))
//      diagonalises, with eigenvalues
//        mu(l, k) = 4 sin^2(pi k / ncx) + 4 sin^2(pi (l+1) / (2 (m+1)))
//      (written as sums of squares, so the smallest one, of order 1e-5 at
//      512 rows, is not a difference of numbers near 2).
//   3. X^ = S Y.
//   4. phi = X^ G, G [ncx][ncx] the inverse basis with the weights 1 or 2 of
//      the half spectrum, into the FP64 potential and its FP32 copy.
//
// Each product is split along k into ns slices so that a launch has enough
// workgroups for the device; the slices' partial sums are added by a second
// kernel in a fixed order, so the result does not depend on scheduling.
// At 1024 x 512 nodes: 2 x 0.54 G + 2 x 0.27 G FMA.  Nothing here depends on
// other runs of a batch.
#include "cupic_launch.h"

namespace cupic {


typedef double v4d __attribute__((ext_vector_type(4)));

// Columns of the real spectrum: cos columns c < kc(ncx) for k = c, then sin
// columns for k = c - kc + 1.
__host__ __device__ inline int dft_cos_cols(int ncx) { return ncx / 2 + 1; }
__device__ __forceinline__ int dft_k_of(int c, int ncx)
{
    const int kc = dft_cos_cols(ncx);
    return c < kc ? c : c - kc + 1;
}

// One k-slice of an FP64 matrix product on the matrix cores: a 256-thread
// workgroup computes a kGM x kGN block of A B over k in [k_begin, k_end)
// (A: M x Kd through load_a, B: Kd x N row-major) with
// v_mfma_f64_16x16x4f64.  Wave w owns rows 16w..16w+15 of the block and
// kGN / 16 output blocks of 16 x 16 (independent accumulation chains); the k
// dimension is staged through LDS kGK at a time.  Operand layout of the
// instruction: lane l supplies A[l % 16][k + l / 16] and B[k + l / 16][l % 16];
// it receives D[l / 16 + 4 r][l % 16] in result register r.  Every element
// of the block is stored, zero when the slice is empty.  As and Bs are the
// kernel's LDS staging tiles (one pair, whatever the operands).
template <class LoadA, class StoreC>
__device__ __forceinline__ void gemm_slice(int M, int N, int k_begin, int k_end, int tile,
                                           LoadA load_a, const double *__restrict__ B,
                                           StoreC store_c, double (*As)[kGM + 1],
                                           double (*Bs)[kGN + 1])
{
    constexpr int NB = kGN / 16;
    const int ntn = (N + kGN - 1) / kGN;
    const int tm = tile / ntn, tn = tile - tm * ntn;
    const int row0 = tm * kGM, col0 = tn * kGN;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    constexpr int QA = kGM * kGK / kBlock, QB = kGK * kGN / kBlock;
    v4d acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = v4d{0.0, 0.0, 0.0, 0.0};
    // the next k tile is loaded into registers while the matrix cores work
    // on the current one
    double ra[QA], rb[QB];
    auto load_tile = [&](int k0) {
#pragma unroll
        for (int q = 0; q < QA; q++) {
            const int e = threadIdx.x + q * kBlock;  // kGM rows x kGK k, k fastest
            const int r = e / kGK, kk = e - r * kGK;
            const int gr = row0 + r, gk = k0 + kk;
            ra[q] = (gr < M && gk < k_end) ? load_a(gr, gk) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < QB; q++) {
            const int e = threadIdx.x + q * kBlock;  // kGK k x kGN columns, columns fastest
            const int kk = e / kGN, c = e - kk * kGN;
            const int gk = k0 + kk, gc = col0 + c;
            rb[q] = (gk < k_end && gc < N) ? B[(size_t)gk * N + gc] : 0.0;
        }
    };
    if (k_begin < k_end) load_tile(k_begin);
    for (int k0 = k_begin; k0 < k_end; k0 += kGK) {
#pragma unroll
        for (int q = 0; q < QA; q++) {
            const int e = threadIdx.x + q * kBlock;
            As[e - (e / kGK) * kGK][e / kGK] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < QB; q++) {
            const int e = threadIdx.x + q * kBlock;
            Bs[e / kGN][e - (e / kGN) * kGN] = rb[q];
        }
        __syncthreads();
        if (k0 + kGK < k_end) load_tile(k0 + kGK);
#pragma unroll
        for (int kk = 0; kk < kGK; kk += 4) {
            const double a = As[kk + lk][16 * w + li];
#pragma unroll
            for (int nb = 0; nb < NB; nb++) {
                const double bv = Bs[kk + lk][16 * nb + li];
                acc[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[nb], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int gr = row0 + 16 * w + lk + 4 * r, gc = col0 + 16 * nb + li;
            if (gr < M && gc < N) store_c(gr, gc, acc[nb][r]);
        }
}

// The k range of slice s of ns over Kd, in whole kGK tiles.
__device__ __forceinline__ void slice_range(int Kd, int ns, int s, int &k_begin, int &k_end)
{
    const int nkt = (Kd + kGK - 1) / kGK;
    const int per = (nkt + ns - 1) / ns;
    k_begin = min(Kd, s * per * kGK);
    k_end = min(Kd, (s + 1) * per * kGK);
}

// Set-up tables: F and G (angles reduced exactly: (i k) mod ncx), the sine
// basis S (reduced mod 2(m+1)), and the eigenvalue parts 4 sin^2(pi k/ncx)
// per column and 4 sin^2(pi (l+1) / (2(m+1))) per row.
__global__ __launch_bounds__(kBlock) void k_dft_init(const RunDev *rds, const int *start,
                                                     int nruns)
{
    RUN_PROLOGUE();
    const int ncx = R.ncx, m = R.ncy - 1;
    const long long nfg = (long long)ncx * ncx, nss = (long long)m * m;
    const int bfg = (int)((nfg + kBlock - 1) / kBlock), bss = (int)((nss + kBlock - 1) / kBlock);
    const long long e0 = (long long)local * kBlock + threadIdx.x;
    if (local < bfg) {
        if (e0 >= nfg) return;
        const int i = (int)(e0 / ncx), c = (int)(e0 - (long long)i * ncx);
        const int kc = dft_cos_cols(ncx), k = dft_k_of(c, ncx);
        const long long ik = ((long long)i * k) % ncx;
        const double th = 2.0 * (double)ik / (double)ncx;
        const double cs = cospi(th), sn = sinpi(th);
        const double wgt = (k == 0 || 2 * k == ncx) ? 1.0 : 2.0;
        R.dft_f[e0] = c < kc ? cs : -sn;                                      // F[i][c]
        R.dft_g[(size_t)c * ncx + i] = (c < kc ? wgt * cs : -wgt * sn) / (double)ncx;  // G[c][i]
        return;
    }
    const long long e1 = e0 - (long long)bfg * kBlock;
    if (local < bfg + bss) {
        if (e1 >= nss) return;
        const int l = (int)(e1 / m), j = (int)(e1 - (long long)l * m);
        const long long q = ((long long)(l + 1) * (j + 1)) % (2LL * (m + 1));
        R.dft_s[e1] = sqrt(2.0 / (double)(m + 1)) * sinpi((double)q / (double)(m + 1));
        return;
    }
    const long long e2 = e1 - (long long)bss * kBlock;
    if (e2 < ncx) {
        const double s = sinpi((double)dft_k_of((int)e2, ncx) / (double)ncx);
        R.dft_d[e2] = 4.0 * s * s;
    } else if (e2 < (long long)ncx + m) {
        const int l = (int)(e2 - ncx);
        const double s = sinpi((double)(l + 1) / (2.0 * (double)(m + 1)));
        R.dft_d[e2] = 4.0 * s * s;
    }
}

// The four products, one kernel each: workgroup local = (slice, tile).
// step 0: R F (R = b rows 1..m, the Dirichlet rows added to rows 1 and m);
// step 1: S R^; step 2: S Y; step 3: X^ G.
__global__ __launch_bounds__(kBlock) void k_dft_gemm(const RunDev *rds, const int *start,
                                                     int nruns, int step)
{
    RUN_PROLOGUE();
    __shared__ double As[kGK][kGM + 1];
    __shared__ double Bs[kGK][kGN + 1];
    const int ncx = R.ncx, ncy = R.ncy, m = ncy - 1;
    const bool xdir = (step == 0 || step == 3);
    const int ns = xdir ? R.nsx : R.nsy;
    const int tiles = ((m + kGM - 1) / kGM) * ((ncx + kGN - 1) / kGN);
    const int s = local / tiles, tile = local - s * tiles;
    if (s >= ns) return;
    int k_begin, k_end;
    slice_range(xdir ? ncx : m, ns, s, k_begin, k_end);
    CUPIC_GLOBAL double *P = gp(R.dft_p) + (size_t)s * m * ncx;
    auto store = [&](int r, int c, double v) { P[(size_t)r * ncx + c] = v; };
    if (step == 0) {
        const CUPIC_GLOBAL double *phi = gp(R.phi64);
        const CUPIC_GLOBAL double *b = gp(R.b64);
        gemm_slice(
            m, ncx, k_begin, k_end, tile,
            [&](int r, int k) -> double {
                double v = b[(size_t)(r + 1) * ncx + k];
                if (r == 0) v += phi[k];                            // phi row 0 (probe)
                if (r == m - 1) v += phi[(size_t)ncy * ncx + k];    // phi row ncy
                return v;
            },
            R.dft_f, store, As, Bs);
    } else if (step == 3) {
        const CUPIC_GLOBAL double *X = gp(R.dft_z);
        gemm_slice(
            m, ncx, k_begin, k_end, tile,
            [&](int r, int k) -> double { return X[(size_t)r * ncx + k]; }, R.dft_g, store, As,
            Bs);
    } else {
        const CUPIC_GLOBAL double *S = gp(R.dft_s);
        gemm_slice(
            m, ncx, k_begin, k_end, tile,
            [&](int l, int j) -> double { return S[(size_t)l * m + j]; },
            step == 1 ? R.dft_z : R.dft_y, store, As, Bs);
    }
}

// Sum of the slices in slice order, then: step 0 -> R^ (dft_z); step 1 ->
// Y (dft_y) divided by mu; step 2 -> X^ (dft_z); step 3 -> rows 1..m of the
// FP64 potential and its FP32 copy.
__global__ __launch_bounds__(kBlock) void k_dft_reduce(const RunDev *rds, const int *start,
                                                       int nruns, int step)
{
    RUN_PROLOGUE();
    const int ncx = R.ncx, m = R.ncy - 1;
    const long long n = (long long)m * ncx;
    const long long q = (long long)local * kBlock + threadIdx.x;
    if (q >= n) return;
    const int ns = (step == 0 || step == 3) ? R.nsx : R.nsy;
    const CUPIC_GLOBAL double *P = gp(R.dft_p);
    double v = P[q];
    for (int s = 1; s < ns; s++) v += P[(size_t)s * n + q];
    const int r = (int)(q / ncx), c = (int)(q - (long long)r * ncx);
    if (step == 0 || step == 2) {
        gp(R.dft_z)[q] = v;
    } else if (step == 1) {
        gp(R.dft_y)[q] = v / (R.dft_d[c] + R.dft_d[ncx + r]);
    } else {
        const size_t k = (size_t)(r + 1) * ncx + c;
        gp(R.phi64)[k] = v;
        gp(R.phi32[st->phi_cur])[k] = (float)v;
    }
}

#define LAUNCH_K(kern, L, ...)                                                         \
    do {                                                                               \
        if ((L).nblocks > 0)                                                           \
            hipLaunchKernelGGL(kern, dim3((L).nblocks), dim3(kBlock), 0, (L).stream,   \
                               (L).rds, (L).start, (L).nruns, ##__VA_ARGS__);          \
    } while (0)

void launch_dft_init(const Launch &L) { LAUNCH_K(k_dft_init, L); }
void launch_dft_gemm(const Launch &L, int step) { LAUNCH_K(k_dft_gemm, L, step); }
void launch_dft_reduce(const Launch &L, int step) { LAUNCH_K(k_dft_reduce, L, step); }

}  // namespace cupic
