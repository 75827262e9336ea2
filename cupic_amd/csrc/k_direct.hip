// k_direct.hip -- direct FP64 solve of the Jacobi system (Poisson mode 3).
//
// The Jacobi iteration of src/mesh.cu:297-381 converges to the solution of
// the 5-point system 4 phi - (sum of the four neighbours) = b on rows
// 1..ncy-1, periodic in x, with the Dirichlet rows 0 (probe) and ncy.  This
// mode computes that fixed point directly:
//
//   1. a real DFT along x of the right-hand side and of the two Dirichlet
//      rows, as an FP64 matrix product Z_hat = Z F, F = [cos | -sin] for the
//      wavenumbers k = 0 .. ncx/2 (the spectrum of real data is symmetric);
//   2. per wavenumber, the tridiagonal system in y
//        -x(j-1) + (4 - 2 cos(2 pi k / ncx)) x(j) - x(j+1) = r(j),
//      by the Thomas algorithm with its multipliers tabulated at set-up;
//   3. the inverse DFT, another FP64 product phi = X_hat G, G = the
//      inverse basis with the weights 1 or 2 of the half spectrum.
//
// Cost at 1024 x 512 nodes: 2 x 0.54 G FMA plus 1,026 tridiagonal sweeps of
// 511 rows, against 500 Jacobi sweeps that reduce the longest wavelengths
// by about 1.4 % per step on such grids.  Nothing here depends on other runs
// of a batch.
#include "cupic_launch.h"

namespace cupic {

#define RUN_PROLOGUE()                                              \
    const int b_ = blockIdx.x;                                      \
    const int r_ = find_run(start, nruns, b_);                      \
    const RunDev &R = rds[r_];                                      \
    const int local = b_ - start[r_];                               \
    RunState *st = R.st;                                            \
    (void)local; (void)st;

// FP64 GEMM tile on the matrix cores: a 256-thread workgroup computes a
// kGM x kGN block of C = A B (A: M x Kd through load_a, B: Kd x N row-major)
// with v_mfma_f64_16x16x4f64.  Wave w owns rows 16w..16w+15 of the block
// and kGN / 16 output blocks of 16 x 16; the k dimension is staged through
// LDS kGK at a time.  Operand layout of the instruction: lane l supplies
// A[l % 16][k + l / 16] and B[k + l / 16][l % 16]; it receives
// D[l / 16 + 4 r][l % 16] in result register r.
typedef double v4d __attribute__((ext_vector_type(4)));

template <class LoadA, class StoreC>
__device__ __forceinline__ void gemm_tile(int M, int N, int Kd, int tile, LoadA load_a,
                                          const double *__restrict__ B, StoreC store_c)
{
    constexpr int NB = kGN / 16;
    __shared__ double As[kGK][kGM + 1];
    __shared__ double Bs[kGK][kGN + 1];
    const int ntn = (N + kGN - 1) / kGN;
    const int tm = tile / ntn, tn = tile - tm * ntn;
    const int row0 = tm * kGM, col0 = tn * kGN;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    v4d acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < Kd; k0 += kGK) {
#pragma unroll
        for (int q = 0; q < kGM * kGK / kBlock; q++) {
            int e = threadIdx.x + q * kBlock;   // kGM rows x kGK k, k fastest
            int r = e / kGK, kk = e - r * kGK;
            int gr = row0 + r, gk = k0 + kk;
            As[kk][r] = (gr < M && gk < Kd) ? load_a(gr, gk) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kGK * kGN / kBlock; q++) {
            int e = threadIdx.x + q * kBlock;   // kGK k x kGN columns, columns fastest
            int kk = e / kGN, c = e - kk * kGN;
            int gk = k0 + kk, gc = col0 + c;
            Bs[kk][c] = (gk < Kd && gc < N) ? B[(size_t)gk * N + gc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kGK; kk += 4) {
            double a = As[kk + lk][16 * w + li];
#pragma unroll
            for (int nb = 0; nb < NB; nb++) {
                double b = Bs[kk + lk][16 * nb + li];
                acc[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[nb], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int gr = row0 + 16 * w + lk + 4 * r, gc = col0 + 16 * nb + li;
            if (gr < M && gc < N) store_c(gr, gc, acc[nb][r]);
        }
}

// Set-up: the forward basis F [ncx][2K] and inverse basis G [2K][ncx]
// (angles reduced exactly: (i k) mod ncx), then the Thomas multipliers
// g(j, k) = 1 / (d_k - g(j-1, k)), g(1, k) = 1 / d_k, stored [ncy-1][K].
__global__ __launch_bounds__(kBlock) void k_dft_init(const RunDev *rds, const int *start,
                                                     int nruns)
{
    RUN_PROLOGUE();
    const int ncx = R.ncx, K = R.nk, W = 2 * K, m = R.ncy - 1;
    const long long nbasis = (long long)ncx * W;
    const int nb_blocks = (int)((nbasis + kBlock - 1) / kBlock);
    if (local < nb_blocks) {
        long long e = (long long)local * kBlock + threadIdx.x;
        if (e >= nbasis) return;
        int i = (int)(e / W), c = (int)(e - (long long)i * W);
        int k = c < K ? c : c - K;
        long long ik = ((long long)i * k) % ncx;
        double th = 2.0 * (double)ik / (double)ncx;
        double cs = cospi(th), sn = sinpi(th);
        double w = (k == 0 || 2 * k == ncx) ? 1.0 : 2.0;
        R.dft_f[e] = c < K ? cs : -sn;
        R.dft_g[(size_t)c * ncx + i] = (c < K ? w * cs : -w * sn) / (double)ncx;
        return;
    }
    int k = (local - nb_blocks) * kBlock + threadIdx.x;
    if (k >= K) return;
    double d = 4.0 - 2.0 * cospi(2.0 * (double)k / (double)ncx);
    double g = 1.0 / d;
    R.dft_t[k] = g;
    for (int j = 2; j <= m; j++) {
        g = 1.0 / (d - g);
        R.dft_t[(size_t)(j - 1) * K + k] = g;
    }
}

// Z_hat = Z F: rows 0 and ncy of Z are the Dirichlet rows of phi, rows
// 1..ncy-1 the right-hand side b.
__global__ __launch_bounds__(kBlock) void k_dft_fwd(const RunDev *rds, const int *start,
                                                    int nruns)
{
    RUN_PROLOGUE();
    const int ncx = R.ncx, ncy = R.ncy, W = 2 * R.nk;
    const CUPIC_GLOBAL double *phi = gp(R.phi64);
    const CUPIC_GLOBAL double *b = gp(R.b64);
    CUPIC_GLOBAL double *Z = gp(R.dft_z);
    gemm_tile(
        ncy + 1, W, ncx, local,
        [&](int r, int k) -> double {
            size_t q = (size_t)r * ncx + k;
            return (r == 0 || r == ncy) ? phi[q] : b[q];
        },
        R.dft_f, [&](int r, int c, double v) { Z[(size_t)r * W + c] = v; });
}

// One tridiagonal system per column of Z_hat (real and imaginary parts of
// each wavenumber), solved in place on rows 1..ncy-1; the Dirichlet rows
// enter the first and last equations.  Rows go through registers kTri at a
// time, so the loads of a block are in flight together and only the
// recurrence itself is sequential.
__global__ __launch_bounds__(kBlock) void k_dft_tri(const RunDev *rds, const int *start,
                                                    int nruns)
{
    RUN_PROLOGUE();
    constexpr int kTri = 16;
    const int K = R.nk, W = 2 * K, ncy = R.ncy, m = ncy - 1;
    const int c = local * kBlock + threadIdx.x;
    if (c >= W) return;
    const int k = c < K ? c : c - K;
    CUPIC_GLOBAL double *Z = gp(R.dft_z);
    const CUPIC_GLOBAL double *T = gp(R.dft_t);
    const double z0 = Z[c], zn = Z[(size_t)ncy * W + c];
    double y = 0.0;
    for (int j0 = 1; j0 <= m; j0 += kTri) {
        double r[kTri], g[kTri];
#pragma unroll
        for (int q = 0; q < kTri; q++) {
            int j = j0 + q;
            r[q] = j <= m ? Z[(size_t)j * W + c] : 0.0;
            g[q] = j <= m ? T[(size_t)(j - 1) * K + k] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kTri; q++) {
            int j = j0 + q;
            if (j <= m) {
                double rr = r[q];
                if (j == 1) rr += z0;
                if (j == m) rr += zn;
                y = (rr + y) * g[q];
                Z[(size_t)j * W + c] = y;
            }
        }
    }
    // back substitution x(j) = y(j) + g(j) x(j+1), x(m) = y(m), from row m-1 down
    double x = y;
    for (int j1 = m - 1; j1 >= 1; j1 -= kTri) {
        double r[kTri], g[kTri];
#pragma unroll
        for (int q = 0; q < kTri; q++) {
            int j = j1 - q;
            r[q] = j >= 1 ? Z[(size_t)j * W + c] : 0.0;
            g[q] = j >= 1 ? T[(size_t)(j - 1) * K + k] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kTri; q++) {
            int j = j1 - q;
            if (j >= 1) {
                x = r[q] + g[q] * x;
                Z[(size_t)j * W + c] = x;
            }
        }
    }
}

// phi (rows 1..ncy-1) = X_hat G, into the FP64 potential and its FP32 copy.
__global__ __launch_bounds__(kBlock) void k_dft_inv(const RunDev *rds, const int *start,
                                                    int nruns)
{
    RUN_PROLOGUE();
    const int ncx = R.ncx, W = 2 * R.nk, m = R.ncy - 1;
    const CUPIC_GLOBAL double *Z = gp(R.dft_z);
    CUPIC_GLOBAL double *phi = gp(R.phi64);
    CUPIC_GLOBAL float *p32 = gp(R.phi32[st->phi_cur]);
    gemm_tile(
        m, ncx, W, local, [&](int r, int k) -> double { return Z[(size_t)(r + 1) * W + k]; },
        R.dft_g, [&](int r, int c, double v) {
            size_t q = (size_t)(r + 1) * ncx + c;
            phi[q] = v;
            p32[q] = (float)v;
        });
}

#define LAUNCH_K(kern, L, ...)                                                         \
    do {                                                                               \
        if ((L).nblocks > 0)                                                           \
            hipLaunchKernelGGL(kern, dim3((L).nblocks), dim3(kBlock), 0, (L).stream,   \
                               (L).rds, (L).start, (L).nruns, ##__VA_ARGS__);          \
    } while (0)

void launch_dft_init(const Launch &L) { LAUNCH_K(k_dft_init, L); }
void launch_dft_fwd(const Launch &L) { LAUNCH_K(k_dft_fwd, L); }
void launch_dft_tri(const Launch &L) { LAUNCH_K(k_dft_tri, L); }
void launch_dft_inv(const Launch &L) { LAUNCH_K(k_dft_inv, L); }

}  // namespace cupic
