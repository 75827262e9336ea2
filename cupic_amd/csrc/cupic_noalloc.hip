// cupic_noalloc.hip -- the Jacobi Poisson solve of cupic_poisson_solver on
// caller-owned device buffers, with every temporary (run description and
// state, right-hand side, the two other iterate buffers, reduction slots) in
// a static device array instead of device allocations.  It runs the same
// kernels in the same sequence as a run's solve, so its results are the
// run's bit for bit.
#include <cstddef>
#include <cstring>
#include <mutex>
#include <string>

#include "cupic.h"
#include "cupic_launch.h"

namespace cupic {
namespace {

constexpr size_t kScratchBytes = CUPIC_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// One solve's temporaries.  In host code `scratch` is a host-side stand-in
// for the array, not a device address: host code uses the address that
// noalloc_setup looks up for each device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};  // device address of scratch, per device

// head of the scratch: the run description the kernels read, its state and
// the block prefixes of each work shape (a batch of one run)
struct Header {
    RunDev rd;
    RunState st;
    int plan[kShapeCount * 2];
};

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// The right-hand side and rho from the caller's charge (left unchanged), and
// the caller's potential, Dirichlet rows included, into the two other iterate
// buffers (a solve may end in either of them).
__global__ __launch_bounds__(kBlock) void k_noalloc_stage(const RunDev *rds, const int *start,
                                                          int nruns, const long long *charge)
{
    RUN_PROLOGUE();
    const int k = local * kBlock + threadIdx.x;
    if (k >= R.nnodes) return;
    float rho;
    double bb;
    charge_to_rhs(R, charge[k], k / R.ncx, &rho, &bb);
    if (R.rho) R.rho[k] = rho;
    R.b64[k] = bb;
    R.b32[k] = (float)bb;
    const float v = R.phi32[0][k];
    R.phi32[1][k] = v;
    R.phi32[2][k] = v;
}

// The solution into the caller's buffer (buffer 0) when the solve ended in
// another one.
__global__ __launch_bounds__(kBlock) void k_noalloc_publish(const RunDev *rds, const int *start,
                                                            int nruns)
{
    RUN_PROLOGUE();
    const int pc = st->phi_cur;
    const int k = local * kBlock + threadIdx.x;
    if (pc == 0 || k >= R.nnodes) return;
    R.phi32[0][k] = R.phi32[pc][k];
}

}  // namespace

int noalloc_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= kMaxDevices)
        return set_global_error(CUPIC_ERR_ARG, "no-allocation scratch: device index out of range");
    if (g_ptr[device]) return CUPIC_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_global_error(CUPIC_ERR_HIP, (std::string("hipGetSymbolAddress(scratch): ") +
                                                hipGetErrorString(e)).c_str());
    g_ptr[device] = (unsigned char *)p;
    return CUPIC_OK;
}

}  // namespace cupic

extern "C" int cupic_poisson_solver_noalloc(const cupic_params *pin, const int64_t *charge,
                                            float *phi, float *rho, double tol, void *stream_arg,
                                            int32_t *sweeps, double *resid)
{
    using namespace cupic;
    if (!pin || !charge || !phi) return set_global_error(CUPIC_ERR_ARG, "null argument");
    cupic_params p = *pin;
    int st = cupic_params_derive(&p);
    if (st) return st;
    if (p.jacobi_mode != CUPIC_JACOBI_REFERENCE && p.jacobi_mode != CUPIC_JACOBI_RESIDUAL)
        return set_global_error(CUPIC_ERR_ARG,
                                "no-allocation solve: jacobi_mode must be 0 (reference rule) "
                                "or 1 (FP64 residual)");
    // scratch layout after the header: b64, b32, the two other phi buffers,
    // the per-sweep (mode 0) and per-check (mode 1) reduction slots
    const size_t nn = (size_t)p.ncx * (size_t)(p.ncy + 1);
    const size_t msw = (size_t)p.jacobi_max_sweeps;
    size_t off = up256(sizeof(Header));
    const size_t o_b64 = off; off += up256(8 * nn);
    const size_t o_b32 = off; off += up256(4 * nn);
    const size_t o_p1 = off; off += up256(4 * nn);
    const size_t o_p2 = off; off += up256(4 * nn);
    const size_t o_sm = off; off += up256((msw + 1) * kRedSlots * sizeof(unsigned int));
    const size_t o_rs = off; off += up256((msw + 2) * kRedSlots * sizeof(unsigned long long));
    if (off > kScratchBytes)
        return set_global_error(
            CUPIC_ERR_ARG,
            (std::string("no-allocation solve: this grid and sweep cap need ") +
             std::to_string(off) + " bytes of scratch, above the limit of " +
             std::to_string(kScratchBytes) + " (CUPIC_NOALLOC_SCRATCH_BYTES)").c_str());
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return set_global_error(CUPIC_ERR_HIP, (std::string("hipGetDevice: ") +
                                                hipGetErrorString(e)).c_str());
    if ((st = noalloc_setup(dev))) return st;
    unsigned char *base;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = g_ptr[dev];
    }
    Header h;
    std::memset(&h, 0, sizeof h);
    RunDev &R = h.rd;
    R.ncx = p.ncx;
    R.ncy = p.ncy;
    R.ncells = p.ncx * p.ncy;
    R.nnodes = (int)nn;
    R.C = R.C1 = 1;
    R.tpr = R.tpr1 = p.ncx;
    R.max_sweeps_alloc = p.jacobi_max_sweeps;
    R.step_mode = p.jacobi_mode;
    R.step_min = p.jacobi_min_sweeps;
    R.step_max = p.jacobi_max_sweeps;
    R.step_check = p.jacobi_check_every;
    R.step_tol = p.jacobi_tol;
    R.init_tol = p.jacobi_tol_init;
    R.ds = p.ds;
    R.dt = p.dt;
    R.eps0 = p.eps0;
    R.phi_p = p.phi_p;
    // the constants of a run (cupic_run_create)
    R.b_interior = 0x1p-32 / (p.ds * p.eps0);
    R.rho_interior = 0x1p-32 / (p.ds * p.ds * p.ds);
    R.rho_edge = 0x1p-32 / (0.5 * p.ds * p.ds * p.ds);
    R.rho = rho;
    R.b64 = (double *)(base + o_b64);
    R.b32 = (float *)(base + o_b32);
    R.phi32[0] = phi;
    R.phi32[1] = (float *)(base + o_p1);
    R.phi32[2] = (float *)(base + o_p2);
    R.sweep_max = (unsigned int *)(base + o_sm);
    R.resid = (unsigned long long *)(base + o_rs);
    R.st = (RunState *)(base + offsetof(Header, st));
    h.st.phi_cur = 0;  // the caller's buffer holds the starting iterate
    int nblk[kShapeCount];
    for (int s = 0; s < kShapeCount; s++) {
        nblk[s] = shape_blocks(R, (Shape)s);
        h.plan[2 * s] = 0;
        h.plan[2 * s + 1] = nblk[s];
    }
    const RunDev *d_rd = (const RunDev *)(base + offsetof(Header, rd));
    const int *d_plan = (const int *)(base + offsetof(Header, plan));
    hipStream_t stream = (hipStream_t)stream_arg;
    // h is staged from this call's stack; the call synchronizes on the stream
    // before it returns, so the copy has completed by then
    e = hipMemcpyAsync(base, &h, sizeof h, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess)
        return set_global_error(CUPIC_ERR_HIP, (std::string("hipMemcpyAsync: ") +
                                                hipGetErrorString(e)).c_str());
    const int nb = nblk[kShapeNodes];
    hipLaunchKernelGGL(k_noalloc_stage, dim3(nb), dim3(kBlock), 0, stream, d_rd, d_plan, 1,
                       (const long long *)charge);
    seq_solve_one(d_rd, d_plan, nblk, R.ncx, R.ncy, stream, p.jacobi_mode,
                  tol < 0 ? p.jacobi_tol : tol, p.jacobi_min_sweeps, p.jacobi_max_sweeps,
                  p.jacobi_check_every);
    hipLaunchKernelGGL(k_noalloc_publish, dim3(nb), dim3(kBlock), 0, stream, d_rd, d_plan, 1);
    e = hipGetLastError();
    RunState out;
    if (e == hipSuccess)
        e = hipMemcpyAsync(&out, base + offsetof(Header, st), sizeof out, hipMemcpyDeviceToHost,
                           stream);
    hipError_t es = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = es;
    if (e != hipSuccess)
        return set_global_error(CUPIC_ERR_HIP, (std::string("no-allocation solve: ") +
                                                hipGetErrorString(e)).c_str());
    if (sweeps) *sweeps = out.sweeps_done;
    if (resid) *resid = out.last_resid;
    return CUPIC_OK;
}
