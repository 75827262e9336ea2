// k_direct.hip -- direct FP64 solve of the Jacobi system (Poisson mode 3).
//
// The Jacobi iteration of src/mesh.cu:297-381 converges to the solution of
// the 5-point system 4 phi - (sum of the four neighbours) = b on rows
// 1..ncy-1, periodic in x, with the Dirichlet rows 0 (probe) and ncy.  This
// mode computes that fixed point directly, by fast transforms where their
// lengths fit the FFT kernels (the FFT path below, any ncx and ncy whose
// FFT lengths are at most kFftL), else as four FP64 matrix products on the
// matrix cores (m = ncy - 1 interior rows):
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
    if (R.dft_fft) {
        // twiddles e^(-2 pi i t / L) for L = Lx, then L = Ly (t / L exact),
        // the chirps w_n = e^(-i pi n^2 / N) for N = ncx, then N = 2 ncy
        // (n^2 reduced exactly mod 2N), then the eigenvalue parts as below;
        // Bluestein's filter spectra follow from the chirps (k_fft step 3)
        const FftTabs T = fft_tabs(ncx, R.ncy);
        int e = local * kBlock + threadIdx.x;
        double2 *tw = reinterpret_cast<double2 *>(R.dft_tw);
        if (e < T.ch_x) {
            const int L = e < T.tw_y ? T.lx : T.ly, t = e < T.tw_y ? e : e - T.tw_y;
            const double a = 2.0 * (double)t / (double)L;
            tw[e] = make_double2(cospi(a), -sinpi(a));
            return;
        }
        if (e < T.bh_x) {
            const int N = e < T.ch_y ? ncx : 2 * R.ncy;
            const long long n = e < T.ch_y ? e - T.ch_x : e - T.ch_y;
            const double a = (double)((n * n) % (2LL * N)) / (double)N;
            tw[e] = make_double2(cospi(a), -sinpi(a));
            return;
        }
        e -= T.bh_x;
        if (e < ncx) {
            const double sk = sinpi((double)dft_k_of(e, ncx) / (double)ncx);
            R.dft_d[e] = 4.0 * sk * sk;
        } else if (e < ncx + m) {
            const int l = e - ncx;
            const double sl = sinpi((double)(l + 1) / (2.0 * (double)(m + 1)));
            R.dft_d[e] = 4.0 * sl * sl;
        }
        return;
    }
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

// ---------------------------------------------------------------- FFT path
//
// Where both transforms fit (fft_class), the same solve runs as fast
// transforms, three kernels of one workgroup per row or column pair:
//   0. rows r, r+1 of R (the right-hand side with the Dirichlet rows added,
//      as for the dense product) as one complex row z = R_r + i R_{r+1}; a
//      forward DFT of length ncx split into the two real spectra
//      X1 = (Z_k + conj Z_{-k}) / 2, X2 = (Z_k - conj Z_{-k}) / 2i, stored as
//      the dense path's ncx real coefficients per row, column-major
//      (dft_z[c][r]): Re X_k in the cos columns, Im X_k in the sin columns;
//   1. columns c, c+1 as one complex column x + i y (c+1 absent for the
//      last column of an odd ncx): the sine transform of length m as a DFT
//      of length 2(m+1) of the odd extension
//      (0, v_0 .. v_{m-1}, 0, -v_{m-1} .. -v_0), whose coefficient l+1 is
//      -2i X_l + 2 Y_l; scaled by sqrt(2/(m+1)) and divided by mu(l, k),
//      then transformed again the same way (S is its own inverse), into
//      dft_y[c][r];
//   2. rows r, r+1: both real spectra as one complex spectrum extended by
//      Hermitian symmetry, transformed back (conjugate, forward DFT,
//      conjugate, / ncx) into the two rows of the potential.
// A DFT of power-of-two length is one FFT.  Any other length N runs as
// Bluestein's chirp z-transform: with w_n = e^(-i pi n^2 / N),
// nk = (n^2 + k^2 - (k-n)^2) / 2 gives X_k = w_k sum_n (x_n w_n) conj w_(k-n),
// a cyclic convolution of length P >= 2N - 1 (P a power of two): the
// chirped input, zero-padded, is transformed, multiplied by the filter
// spectrum FFT(b) / P (b_j = conj w_j and b_(P-j) = conj w_j, set up once per
// run), and transformed back (conjugate, FFT, conjugate).
// The FFT: iterative Stockham in LDS (natural order in and out), radix-4
// stages and, when log2 L is odd, a last radix-2 stage; each stage reads
// its butterflies into registers, synchronises and writes in place.  The
// twiddles e^(-2 pi i t / L) come from the run's table.

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Per-thread unroll bounds of an instance (one transform per workgroup):
// radix-4 butterflies per thread and stage, points per thread, and sine
// coefficients per thread (m < P / 2).
template <int LMAX>
struct FftDims {
    static constexpr int kB4 = LMAX / (4 * kBlock) > 1 ? LMAX / (4 * kBlock) : 1;
    static constexpr int kE = 4 * kB4;
    static constexpr int kU = 2 * kB4;
};

// FFT of length L (a power of two) of buf[0 .. L) in place, by the
// workgroup.
template <int LMAX>
__device__ void fft_lds(double2 *buf, const double2 *tw, int L)
{
    const int T = kBlock, lt = threadIdx.x;
    constexpr int kB4 = FftDims<LMAX>::kB4, kB2 = 2 * kB4;
    const int lg = 31 - __clz(L);
    int Ns = 1;
    for (int stage = 0; stage < lg / 2; stage++) {
        const int nb = L >> 2, ts = L / (4 * Ns);
        double2 o[kB4][4];
#pragma unroll
        for (int u = 0; u < kB4; u++) {
            const int j = lt + u * T;
            if (j < nb) {
                const int k = j & (Ns - 1);
                const double2 v0 = buf[j];
                const double2 v1 = cmul(buf[j + nb], tw[k * ts]);
                const double2 v2 = cmul(buf[j + 2 * nb], tw[2 * k * ts]);
                const double2 v3 = cmul(buf[j + 3 * nb], tw[3 * k * ts]);
                const double2 a = make_double2(v0.x + v2.x, v0.y + v2.y);
                const double2 b = make_double2(v0.x - v2.x, v0.y - v2.y);
                const double2 c = make_double2(v1.x + v3.x, v1.y + v3.y);
                const double2 d = make_double2(v1.y - v3.y, v3.x - v1.x);  // -i (v1 - v3)
                o[u][0] = make_double2(a.x + c.x, a.y + c.y);
                o[u][1] = make_double2(b.x + d.x, b.y + d.y);
                o[u][2] = make_double2(a.x - c.x, a.y - c.y);
                o[u][3] = make_double2(b.x - d.x, b.y - d.y);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kB4; u++) {
            const int j = lt + u * T;
            if (j < nb) {
                const int k = j & (Ns - 1);
                const int d = (j - k) * 4 + k;
#pragma unroll
                for (int r = 0; r < 4; r++) buf[d + r * Ns] = o[u][r];
            }
        }
        __syncthreads();
        Ns *= 4;
    }
    if (lg & 1) {
        const int nb = L >> 1;  // Ns = L / 2: twiddle stride 1
        double2 o[kB2][2];
#pragma unroll
        for (int u = 0; u < kB2; u++) {
            const int j = lt + u * T;
            if (j < nb) {
                const int k = j & (Ns - 1);
                const double2 v0 = buf[j];
                const double2 v1 = cmul(buf[j + nb], tw[k]);
                o[u][0] = make_double2(v0.x + v1.x, v0.y + v1.y);
                o[u][1] = make_double2(v0.x - v1.x, v0.y - v1.y);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kB2; u++) {
            const int j = lt + u * T;
            if (j < nb) {
                const int k = j & (Ns - 1);
                const int d = (j - k) * 2 + k;
                buf[d] = o[u][0];
                buf[d + Ns] = o[u][1];
            }
        }
        __syncthreads();
    }
}

typedef double v2d __attribute__((ext_vector_type(2)));

// Complex entry i of a run's FFT table section at offset off (fft_tabs),
// as a global load.
struct CTab {
    const double *p;
    __device__ __forceinline__ double2 operator[](int i) const
    {
        const v2d v = ((const CUPIC_GLOBAL v2d *)p)[i];
        return make_double2(v.x, v.y);
    }
};
__device__ __forceinline__ CTab fft_tab(const RunDev &R, int off)
{
    return CTab{R.dft_tw + 2 * (size_t)off};
}

// Forward DFT of length N of buf[0 .. N) in place (the caller has written
// buf and synchronised; tw holds the twiddles of the FFT length P, P = N
// for a power of two, else Bluestein's: ch the chirps w_n, bh the filter
// spectrum).  Point n of every pass is handled by thread n mod kBlock.
template <int LMAX>
__device__ void dft_lds(double2 *buf, const double2 *tw, int N, int P, CTab ch, CTab bh)
{
    const int T = kBlock, lt = threadIdx.x;
    if (N == P) {
        fft_lds<LMAX>(buf, tw, N);
        return;
    }
    // the chirps and the filter spectrum are read where they are used (L2
    // hits), so that nothing stays live across the FFTs: registers are the
    // kernel's, and power-of-two runs take the same kernel
    constexpr int kE = FftDims<LMAX>::kE;
    double2 t[kE];
#pragma unroll
    for (int u = 0; u < kE; u++) t[u] = ch[min(lt + u * T, N - 1)];
#pragma unroll
    for (int u = 0; u < kE; u++) {
        const int n = lt + u * T;
        if (n < P) buf[n] = n < N ? cmul(buf[n], t[u]) : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<LMAX>(buf, tw, P);
#pragma unroll
    for (int u = 0; u < kE; u++) t[u] = bh[min(lt + u * T, P - 1)];
#pragma unroll
    for (int u = 0; u < kE; u++) {
        const int k = lt + u * T;
        if (k < P) {
            const double2 z = cmul(buf[k], t[u]);
            buf[k] = make_double2(z.x, -z.y);
        }
    }
    __syncthreads();
    fft_lds<LMAX>(buf, tw, P);
#pragma unroll
    for (int u = 0; u < kE; u++) t[u] = ch[min(lt + u * T, N - 1)];
#pragma unroll
    for (int u = 0; u < kE; u++) {
        const int n = lt + u * T;
        if (n < N) {
            const double2 z = buf[n];
            buf[n] = cmul(make_double2(z.x, -z.y), t[u]);
        }
    }
    __syncthreads();
}

// The twiddles of FFT length P into LDS, loaded by the whole workgroup
// (kT per thread) before any LDS store of the caller.
template <int LMAX>
struct TwLoad {
    static constexpr int kT = LMAX / kBlock;
    double2 t[kT];
    __device__ __forceinline__ void load(CTab twg, int P)
    {
#pragma unroll
        for (int u = 0; u < kT; u++) t[u] = twg[min((int)threadIdx.x + u * kBlock, P - 1)];
    }
    __device__ __forceinline__ void store(double2 *tw, int P) const
    {
#pragma unroll
        for (int u = 0; u < kT; u++)
            if ((int)threadIdx.x + u * kBlock < P) tw[threadIdx.x + u * kBlock] = t[u];
    }
};

// The global loads of a kernel are all issued before the first LDS store
// (kE points per thread in registers), so that they cost one round trip.

// step 0: row pair -> the two rows' real spectra, column-major in dft_z
template <int LMAX>
__device__ void fft_rows_fwd(const RunDev &R, int g, double2 *buf, double2 *tw)
{
    constexpr int kE = FftDims<LMAX>::kE;
    const int N = R.ncx, ncy = R.ncy, m = ncy - 1;
    const FftTabs T = fft_tabs(N, ncy);
    const int P = T.lx, lt = threadIdx.x;
    const int r1 = 2 * g, r2 = r1 + 1;
    const CUPIC_GLOBAL double *b = gp(R.b64);
    const CUPIC_GLOBAL double *phi = gp(R.phi64);
    // rows r1 and r2 (r2 clamped: row m does not exist and reads as 0) and
    // the Dirichlet rows, loaded without conditions so that no load waits
    // for another; the Dirichlet values are added to rows 0 and m - 1
    const int ra = r1 + 1, rb = min(r2, m - 1) + 1;
    TwLoad<LMAX> tl;
    tl.load(fft_tab(R, T.tw_x), P);
    double2 v[kE], pd[kE];
#pragma unroll
    for (int u = 0; u < kE; u++) {
        const int i = min(lt + u * kBlock, N - 1);
        v[u] = make_double2(b[(size_t)ra * N + i], b[(size_t)rb * N + i]);
        pd[u] = make_double2(phi[i], phi[(size_t)ncy * N + i]);
    }
#pragma unroll
    for (int u = 0; u < kE; u++) {
        double x = v[u].x, y = r2 < m ? v[u].y : 0.0;
        if (r1 == 0) x += pd[u].x;                   // phi row 0 (probe)
        if (r1 == m - 1) x += pd[u].y;               // phi row ncy
        if (r2 == m - 1) y += pd[u].y;
        v[u] = make_double2(x, y);
    }
    tl.store(tw, P);
#pragma unroll
    for (int u = 0; u < kE; u++) {
        const int i = lt + u * kBlock;
        if (i < N) buf[i] = v[u];
    }
    __syncthreads();
    dft_lds<LMAX>(buf, tw, N, P, fft_tab(R, T.ch_x), fft_tab(R, T.bh_x));
    const int kc = dft_cos_cols(N), ks = N - kc;
    CUPIC_GLOBAL double *Z = gp(R.dft_z);
    for (int k = lt; k <= N / 2; k += kBlock) {
        const double2 a = buf[k], c = buf[k == 0 ? 0 : N - k];
        const double x1r = 0.5 * (a.x + c.x), x1i = 0.5 * (a.y - c.y);
        const double x2r = 0.5 * (a.y + c.y), x2i = -0.5 * (a.x - c.x);
        Z[(size_t)k * m + r1] = x1r;
        if (r2 < m) Z[(size_t)k * m + r2] = x2r;
        if (k >= 1 && k <= ks) {
            Z[(size_t)(kc + k - 1) * m + r1] = x1i;
            if (r2 < m) Z[(size_t)(kc + k - 1) * m + r2] = x2i;
        }
    }
}

// step 1: column pair: Y = S R^ / mu, then X^ = S Y, column-major in dft_y
template <int LMAX>
__device__ void fft_cols(const RunDev &R, int g, double2 *buf, double2 *tw)
{
    constexpr int kE = FftDims<LMAX>::kE, kU = FftDims<LMAX>::kU;
    const int N = R.ncx, M = R.ncy, m = M - 1, L = 2 * M;
    const FftTabs T = fft_tabs(N, M);
    const int P = T.ly, lt = threadIdx.x;
    const int c1 = 2 * g, c2 = c1 + 1;
    const bool two = c2 < N;        // an odd ncx leaves the last column alone
    const int c2r = two ? c2 : c1;  // its partner reads column c1 and is not stored
    const CUPIC_GLOBAL double *Z = gp(R.dft_z);
    const CUPIC_GLOBAL double *D = gp(R.dft_d);
    const double sc = sqrt(2.0 / (double)M);
    // the odd extension (0, v_0 .. v_{m-1}, 0, -v_{m-1} .. -v_0) of the
    // column values v_j = x_j + i y_j: element n < M is v_{n-1}, element
    // n > M is -v_{L-n-1}
    TwLoad<LMAX> tl;
    tl.load(fft_tab(R, T.tw_y), P);
    double2 v[kE];
    double dl[kU];
    const double d1 = D[c1], d2 = D[c2r];
#pragma unroll
    for (int u = 0; u < kE; u++) {
        const int n = min(lt + u * kBlock, L - 1);
        const int j = min(max(n < M ? n - 1 : L - n - 1, 0), m - 1);  // clamped for n = 0, M
        v[u] = make_double2(Z[(size_t)c1 * m + j], Z[(size_t)c2r * m + j]);
    }
#pragma unroll
    for (int u = 0; u < kU; u++) dl[u] = D[N + min(lt + u * kBlock, m - 1)];
#pragma unroll
    for (int u = 0; u < kE; u++) {
        const int n = lt + u * kBlock;
        v[u] = (n == 0 || n == M) ? make_double2(0.0, 0.0)
                                  : (n < M ? v[u] : make_double2(-v[u].x, -v[u].y));
    }
    tl.store(tw, P);
#pragma unroll
    for (int u = 0; u < kE; u++) {
        const int n = lt + u * kBlock;
        if (n < L) buf[n] = v[u];
    }
    __syncthreads();
    const CTab ch = fft_tab(R, T.ch_y), bh = fft_tab(R, T.bh_y);
    dft_lds<LMAX>(buf, tw, L, P, ch, bh);
    // coefficient l + 1 is -2i X_l + 2 Y_l: X_l = -Im / 2, Y_l = Re / 2
    double2 s[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
        const int l = lt + u * kBlock;
        if (l < m) {
            const double2 V = buf[l + 1];
            s[u] = make_double2((sc * (-0.5 * V.y)) / (d1 + dl[u]), (sc * (0.5 * V.x)) / (d2 + dl[u]));
        }
    }
    __syncthreads();
    if (lt == 0) buf[0] = buf[M] = make_double2(0.0, 0.0);
#pragma unroll
    for (int u = 0; u < kU; u++) {
        const int l = lt + u * kBlock;
        if (l < m) {
            buf[l + 1] = s[u];
            buf[L - l - 1] = make_double2(-s[u].x, -s[u].y);
        }
    }
    __syncthreads();
    dft_lds<LMAX>(buf, tw, L, P, ch, bh);
    CUPIC_GLOBAL double *Y = gp(R.dft_y);
    for (int j = lt; j < m; j += kBlock) {
        const double2 V = buf[j + 1];
        Y[(size_t)c1 * m + j] = sc * (-0.5 * V.y);
        if (two) Y[(size_t)c2 * m + j] = sc * (0.5 * V.x);
    }
}

// step 2: row pair of X^ -> rows r1 + 1, r2 + 1 of the FP64 potential and
// its FP32 copy
template <int LMAX>
__device__ void fft_rows_inv(const RunDev &R, const RunState *st, int g, double2 *buf,
                             double2 *tw)
{
    constexpr int kE = FftDims<LMAX>::kE;
    const int N = R.ncx, m = R.ncy - 1;
    const FftTabs T = fft_tabs(N, R.ncy);
    const int P = T.lx, lt = threadIdx.x;
    const int r1 = 2 * g, r2 = r1 + 1;
    const CUPIC_GLOBAL double *X = gp(R.dft_y);
    const int kc = dft_cos_cols(N), ks = N - kc;
    // coefficient k of row r as a complex number (Hermitian beyond N / 2)
    // rows r1 and r2 (r2 clamped, then zeroed when it is row m) of the cos
    // coefficient kk and the sin coefficient kk (index clamped, zeroed
    // outside 1 .. ks), loaded without conditions
    const int rb = min(r2, m - 1);
    TwLoad<LMAX> tl;
    tl.load(fft_tab(R, T.tw_x), P);
    double2 v[kE];
    double4 q[kE];
#pragma unroll
    for (int u = 0; u < kE; u++) {
        const int k = min(lt + u * kBlock, N - 1);
        const int kk = k <= N / 2 ? k : N - k;
        const int cs = kc + min(max(kk, 1), max(ks, 1)) - 1;
        q[u] = make_double4(X[(size_t)kk * m + r1], X[(size_t)cs * m + r1], X[(size_t)kk * m + rb],
                            X[(size_t)cs * m + rb]);
    }
#pragma unroll
    for (int u = 0; u < kE; u++) {
        const int k = min(lt + u * kBlock, N - 1);
        const int kk = k <= N / 2 ? k : N - k;
        const bool sin_ok = kk >= 1 && kk <= ks;
        const double sg = k <= N / 2 ? 1.0 : -1.0;  // Hermitian: conj beyond N / 2
        const double2 a = make_double2(q[u].x, sin_ok ? sg * q[u].y : 0.0);
        double2 b = make_double2(q[u].z, sin_ok ? sg * q[u].w : 0.0);
        if (r2 >= m) b = make_double2(0.0, 0.0);
        v[u] = make_double2(a.x - b.y, -(a.y + b.x));  // conj(a + i b)
    }
    tl.store(tw, P);
#pragma unroll
    for (int u = 0; u < kE; u++) {
        const int k = lt + u * kBlock;
        if (k < N) buf[k] = v[u];
    }
    __syncthreads();
    dft_lds<LMAX>(buf, tw, N, P, fft_tab(R, T.ch_x), fft_tab(R, T.bh_x));
    // / N: for a power of two the product with the exact inverse (the same
    // value, without the division's instruction sequence)
    const double dn = (double)N, inv = 1.0 / dn;
    const bool p2 = pow2_len(N);
    CUPIC_GLOBAL double *phi = gp(R.phi64);
    CUPIC_GLOBAL float *phi32 = gp(R.phi32[st->phi_cur]);
    for (int i = lt; i < N; i += kBlock) {
        const double2 z = buf[i];
        const double v1 = p2 ? z.x * inv : z.x / dn, v2 = p2 ? -z.y * inv : -z.y / dn;
        phi[(size_t)(r1 + 1) * N + i] = v1;
        phi32[(size_t)(r1 + 1) * N + i] = (float)v1;
        if (r2 < m) {
            phi[(size_t)(r2 + 1) * N + i] = v2;
            phi32[(size_t)(r2 + 1) * N + i] = (float)v2;
        }
    }
}

// step 3 (set-up, after k_dft_init): block 0 the x filter spectrum, block 1
// the y one, FFT(b) / P with b_j = conj w_j for j < N, b_(P-j) = conj w_j
// for 0 < j < N and zero between (P >= 2N - 1: the two ranges are
// disjoint); nothing for a power-of-two length
template <int LMAX>
__device__ void fft_filter(const RunDev &R, int dirn, double2 *buf, double2 *tw)
{
    const FftTabs T = fft_tabs(R.ncx, R.ncy);
    const int N = dirn == 0 ? R.ncx : 2 * R.ncy;
    const int P = dirn == 0 ? T.lx : T.ly;
    if (N == P) return;
    const CTab twg = fft_tab(R, dirn == 0 ? T.tw_x : T.tw_y);
    const CTab ch = fft_tab(R, dirn == 0 ? T.ch_x : T.ch_y);
    for (int j = threadIdx.x; j < P; j += kBlock) {
        tw[j] = twg[j];
        const int n = j < N ? j : (j > P - N ? P - j : -1);
        double2 w = make_double2(0.0, 0.0);
        if (n >= 0) w = ch[n];
        buf[j] = make_double2(w.x, -w.y);
    }
    __syncthreads();
    fft_lds<LMAX>(buf, tw, P);
    double2 *bh = reinterpret_cast<double2 *>(R.dft_tw) + (dirn == 0 ? T.bh_x : T.bh_y);
    const double ip = 1.0 / (double)P;  // a power of two: exact
    for (int k = threadIdx.x; k < P; k += kBlock) {
        const double2 z = buf[k];
        bh[k] = make_double2(z.x * ip, z.y * ip);
    }
}

template <int LMAX>
__global__ __launch_bounds__(kBlock) void k_fft(const RunDev *rds, const int *start, int nruns,
                                                int step)
{
    RUN_PROLOGUE();
    __shared__ double2 buf[LMAX];
    __shared__ double2 tw[LMAX];
    if (step == 0)
        fft_rows_fwd<LMAX>(R, local, buf, tw);
    else if (step == 1)
        fft_cols<LMAX>(R, local, buf, tw);
    else if (step == 2)
        fft_rows_inv<LMAX>(R, st, local, buf, tw);
    else
        fft_filter<LMAX>(R, local, buf, tw);
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
void launch_fft(const Launch &L, int step, int cls)
{
    if (cls == 1)
        LAUNCH_K(k_fft<kFftS>, L, step);
    else if (cls == 2)
        LAUNCH_K(k_fft<kFftM>, L, step);
    else
        LAUNCH_K(k_fft<kFftL>, L, step);
}

}  // namespace cupic
