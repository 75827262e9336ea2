// cupic_launch.h -- host-side launchers for the device kernels.
//
// A Launch names a batch of runs (device array of RunDev), the prefix array
// that maps blockIdx.x to (run, local block) for the kernel's work shape, and
// the stream.  The plain path passes a batch of one.
#pragma once

#include <hip/hip_runtime.h>

#include "cupic_dev.h"

namespace cupic {

struct Launch {
    const RunDev *rds;     // device array, nruns entries
    const int *start;      // device array, nruns + 1 block prefixes; with
                           // nruns > 1 followed, from plan_map_at(nruns), by
                           // the (run, local block) pair of every block
    int nruns;
    int nblocks;
    hipStream_t stream;
};

// Work shapes: blocks per run for each kernel family.
enum Shape : int {
    kShapeNodes = 0,   // ceil(nnodes / 256)
    kShapeResid,       // ceil(nnodes / (256 kResNodes))
    kShapeTasks,       // ceil(2 * ncy * tpr / 4): one wave per (species, row, column block)
    kShapeScatter,     // 2 far_blocks() + the kShapeTasks blocks: far-list places, then tasks
    kShapeCountTasks,  // ceil(2 * ncy * tpr1 / 4): the same for the counting pass
    kShapeCellBlk,     // 2 * ceil(ncells / kScanCells)
    kShapeRunSp,       // 2: one block per species
    kShapeRun,         // 1
    kShapeFar,         // 2 * far_blocks(): grid-stride over the far list
    kShapeInject,      // 2 * ceil(max_inject / 256)
    kShapeRows,        // 2 * ncy
    kShapeInit,        // 2 * ceil(n_init / 256)
    kShapeJacBlk,      // ceil(ncx / jac_tx(kJacH)) * ceil((ncy - 1) / jac_ty(kJacH))
    kShapeJacBlkN,     // the same for the narrow halo kJacHNarrow
    kShapeDftInit,     // dense: ceil(ncx^2 / 256) + ceil(m^2 / 256) + ceil((ncx + m) / 256),
                       // m = ncy - 1; FFT: ceil((twiddles + chirps + ncx + m) / 256)
    kShapeDftGemmX,    // ceil(m / kGM) * ceil(ncx / kGN) * nsx
    kShapeDftGemmY,    // ceil(m / kGM) * ceil(ncx / kGN) * nsy
    kShapeDftReduce,   // ceil(m * ncx / 256)
    kShapeFftRowsS,    // ceil(m / 2) when fft_class is 1: row pairs
    kShapeFftColsS,    // ceil(ncx / 2) when fft_class is 1: column pairs
    kShapeFftTabS,     // 2 when fft_class is 1: Bluestein filter spectra (x, y)
    kShapeFftRowsM,    // the same three for fft_class 2
    kShapeFftColsM,
    kShapeFftTabM,
    kShapeFftRowsL,    // and for fft_class 3
    kShapeFftColsL,
    kShapeFftTabL,
    kShapeCount
};

int shape_blocks(const RunDev &h, Shape s);

// mesh
void launch_phi_linear(const Launch &L);
// begin: block 0 also starts the solve with the run's step settings
void launch_rho_convert(const Launch &L, int begin);
void launch_solver_begin(const Launch &L, int use_args, int mode, double tol, int min_sweeps,
                         int max_sweeps, int check_every);
void launch_jacobi_sweep(const Launch &L, int s);
void launch_jacobi_block(const Launch &L, int halo, int s0, int nsw, int mode, int first,
                         int par, int cslot, int cat, int cpar, int rslot, int rat);
void launch_ir_update(const Launch &L, int par);
// end: the last block also ends the solve at check c (as launch_solver_end)
void launch_residual(const Launch &L, int c, int mode, int par, int end);
void launch_solver_end(const Launch &L, int last_check, int par);
// resid_end (direct solve only, instead of its launch_residual): the field
// kernel also reduces the solve's residual; launch_push_rank with end_solve
// set must follow
void launch_field(const Launch &L, int resid_end);

// direct solve (mode 3)
void launch_dft_init(const Launch &L);
void launch_dft_gemm(const Launch &L, int step);
void launch_dft_reduce(const Launch &L, int step);
// step 0: forward x transform of row pairs; 1: the y solve of column pairs;
// 2: inverse x transform into the potential; 3: Bluestein's filter spectra
// (set-up, after launch_dft_init).  cls: the instance, fft_class 1, 2 or 3.
void launch_fft(const Launch &L, int step, int cls);

// particles
void launch_create(const Launch &L);
void launch_init_sort(const Launch &L);
void launch_deposit(const Launch &L);
void launch_half_kick(const Launch &L);
// end_solve: block 0 also ends a direct solve whose residual the field
// kernel reduced (launch_field with resid_end set)
void launch_push_rank(const Launch &L, int end_solve);
// seg_stayers: the instance for fewer than kSegStayersPpc particles per cell
void launch_scatter(const Launch &L, bool seg_stayers);  // kShapeScatter

// boundary conditions and sort
void launch_clock(const Launch &L, int use_host_t, double t_host);
void launch_inject(const Launch &L);
void launch_far_rank(const Launch &L);
void launch_scan_a(const Launch &L);
void launch_scan_c(const Launch &L);
// the far list to its slots; the run's last block ends the step
void launch_far_scatter(const Launch &L, int advance_t);
void launch_init_finish(const Launch &L);

// host sequences and state shared between translation units
// The Poisson solve (any mode) of the one run described by d_rd, an ncx x
// ncy grid, on stream, with the blocks per work shape in nblk and their
// prefixes in d_plan (cupic_api.cpp).
void seq_solve_one(const RunDev *d_rd, const int *d_plan, const int *nblk, int ncx, int ncy,
                   hipStream_t stream, int mode, double tol, int min_sweeps, int max_sweeps,
                   int check_every);
// Records msg as the last error of calls without a run (cupic_last_error(NULL));
// returns code (cupic_api.cpp).
int set_global_error(int code, const char *msg);
// Looks up the address of the no-allocation scratch array on the current
// device, once per device (cupic_noalloc.hip).
int noalloc_setup(int device);

}  // namespace cupic
