// Box-grouped launches of the scheduled step for batches (gfx950).
//
// A batch replay runs one step of every member.  Launched member by member, each of the eight
// kernels of a step fills only a few CUs for the smaller boxes (a 32^3 deposit is 16 workgroups)
// and costs its fixed start-up time once per box, and a captured graph runs its per-box branches one
// kernel at a time.  Here the members of one mesh size share every launch: the grid is the one-box
// grid times the number of boxes, and a workgroup reads its box's pointers from `boxes[blockIdx.y]`
// (blockIdx.z for the far-particle pass, whose y dimension is in use) and runs the same kernel body
// as the one-box launch, so a member's results do not depend on the grouping.
#pragma once
#include "common.hpp"
#include "deposit.hpp"
#include "fft.hpp"
#include "fft_plane.hpp"
#include "particles_mixed.hpp"

namespace pm3d {

// Per-box pointers and constants of one mixed-precision member (device copy, one per member).
struct MxBox {
    int* cell;
    float* off;
    float* p;
    float* acc;
    long long n;
    int N, Np;
    unsigned long long* dscal;  // [1] sticky bad count
    DevSched* sched;
    const DevStep* tab;
    double* jhist;
    unsigned long long* mesh;
    int* far;
    unsigned int* farcnt;
    cplx<float>* spec;   // half spectrum; its Nyquist plane follows at + N*N*(N/2)
    cplx<float>* gspec;  // three gradient spectra of N*N*(N/2) + N*N points each
    float* grad;         // component-major gradient mesh
    const cplx<float>* tw;
    double* bmax;                // per-block maxima / minima
    unsigned long long* qslot;   // deposit maximum slot
    FarRows fr;                  // far-particle row lists (cnt == nullptr: none)
    double kval;                 // 1/(N dx)
    double inv_c3;               // (Np / N)^3: 1 / (cells per lattice site)^3
    double qscale, nmean;
    SiteDiv sdv;                 // lattice site of a particle index (row-list test)
};

// Pointers read from a descriptor in memory are generic to the compiler, which then reaches memory
// through them with flat instructions: a 64-bit address per lane, no scalar base, and completion
// counted against the LDS counter as well, so every LDS wait of a kernel (the deposit's atomics and
// barriers) also waits for its outstanding flat loads.  They all point to global memory; a round
// trip through the global address space lets the compiler use global instructions.
template <class T>
__device__ __forceinline__ T* as_global(T* p) {
    return (T*)(__attribute__((address_space(1))) T*)p;
}

// Box k's descriptor with every pointer marked global (as_global).
__device__ __forceinline__ MxBox load_box(const MxBox* __restrict__ boxes, unsigned k) {
    MxBox b = boxes[k];
    b.cell = as_global(b.cell);
    b.off = as_global(b.off);
    b.p = as_global(b.p);
    b.acc = as_global(b.acc);
    b.dscal = as_global(b.dscal);
    b.sched = as_global(b.sched);
    b.tab = as_global(b.tab);
    b.jhist = as_global(b.jhist);
    b.mesh = as_global(b.mesh);
    b.far = as_global(b.far);
    b.farcnt = as_global(b.farcnt);
    b.spec = as_global(b.spec);
    b.gspec = as_global(b.gspec);
    b.grad = as_global(b.grad);
    b.tw = as_global(b.tw);
    b.bmax = as_global(b.bmax);
    b.qslot = as_global(b.qslot);
    b.fr.cnt = as_global(b.fr.cnt);
    b.fr.list = as_global(b.fr.list);
    b.fr.ovl = as_global(b.fr.ovl);
    b.fr.ovcnt = as_global(b.fr.ovcnt);
    return b;
}

// A member whose schedule has ended (table done, or ended at its first shell crossing) skips the
// step: every kernel of a replay sees the same counter (it changes only in the step's last kernel).
__device__ __forceinline__ bool box_done(const MxBox& b) { return b.sched->counter >= b.sched->nsteps; }

__device__ __forceinline__ SrcMx box_src(const MxBox& b) {
    return SrcMx{b.cell, b.cell + b.n, b.cell + 2 * b.n, b.off, b.off + b.n, b.off + 2 * b.n, b.N};
}

// The previous step's gather fused with this step's kicks and drift (gather_kick_drift_mx_body).
__global__ __launch_bounds__(kThreads) void k_gather_kick_drift_mx_boxes(const MxBox* __restrict__ boxes, XMap xm) {
    const MxBox b = load_box(boxes, blockIdx.y);
    if (box_done(b)) return;
    gather_kick_drift_mx_body(b.cell, b.off, b.p, b.n, b.N, b.grad, 1, 0.f, 0.f, 0.f, b.dscal + 1, b.sched, b.tab, xm,
                              b.fr, b.sdv);
}

template <int BX, int BY, int N>
__global__ __launch_bounds__(SrcMx::kColThreads) void k_deposit_col_boxes(const MxBox* __restrict__ boxes, XMap xm) {
    const MxBox b = load_box(boxes, blockIdx.y);
    if (box_done(b)) return;
    deposit_col_body<SrcMx, BX, BY, N, false, SrcMx::kColThreads>(box_src(b), b.mesh, b.far, b.farcnt, b.fr, xm);
}

__global__ __launch_bounds__(kThreads) void k_deposit_far_seg_boxes(const MxBox* __restrict__ boxes, int seg) {
    const MxBox b = load_box(boxes, blockIdx.z);
    if (box_done(b)) return;
    deposit_far_seg_body<SrcMx, false>(box_src(b), b.mesh, b.far, b.farcnt, seg, b.N, nullptr);
}

__global__ __launch_bounds__(kThreads) void k_deposit_far_bal_boxes(const MxBox* __restrict__ boxes, int ncol, int seg) {
    const MxBox b = load_box(boxes, blockIdx.y);
    if (box_done(b)) return;
    deposit_far_bal_body<SrcMx, false>(box_src(b), b.mesh, b.far, b.farcnt, ncol, seg, b.N, blockIdx.x, gridDim.x, b.fr,
                                       1);
}

template <int N>
__global__ __launch_bounds__((PlaneTile<float, N>::NT)) void k_fwd_plane_boxes(const MxBox* __restrict__ boxes, XMap xm) {
    const MxBox b = load_box(boxes, blockIdx.y);
    if (box_done(b)) return;
    constexpr size_t specN = (size_t)N * N * (N / 2);
    fwd_plane_body<float, N, true, false>(b.mesh, b.spec, b.spec + specN, b.tw, b.bmax, b.qscale, b.nmean, xm,
                                          b.fr.ovcnt);
}

// The Poisson x pass with the scheduled Green's function factor; its last block reduces the
// plane forward pass's 2N per-block deposit maxima into the box's slot.
template <int N, int V>
__global__ __launch_bounds__(kThreads) void k_x_poisson_boxes(const MxBox* __restrict__ boxes, XMap xm) {
    const MxBox b = load_box(boxes, blockIdx.y);
    if (box_done(b)) return;
    constexpr size_t specN = (size_t)N * N * (N / 2), tot = specN + (size_t)N * N;
    x_pass_body<float, N, 2, V>(b.spec, b.spec + specN, b.gspec, b.gspec + specN, tot, tot, b.tw, b.kval, 0.0, 0.0, 0,
                                b.sched, b.tab, b.bmax, 2 * N, b.qslot, xm);
}

template <int N, int Q>
__global__ __launch_bounds__((InvTile<float, N, Q>::NT)) void k_inv_plane_boxes(const MxBox* __restrict__ boxes, XMap xm) {
    const MxBox b = load_box(boxes, blockIdx.y);
    if (box_done(b)) return;
    constexpr size_t specN = (size_t)N * N * (N / 2), tot = specN + (size_t)N * N;
    inv_plane_body<float, N, Q>(b.gspec, b.gspec + specN, tot, tot, b.grad, (size_t)N * N * N, b.tw,
                                (float)(1.0 / ((double)N * N * N)), xm);
}

__global__ __launch_bounds__(kThreads) void k_min_jacobian_mx_boxes(const MxBox* __restrict__ boxes) {
    const MxBox b = load_box(boxes, blockIdx.y);
    if (box_done(b)) return;
    min_jacobian_mx4_body(MxPos{b.cell, b.off, b.n}, b.N, b.inv_c3, b.sdv, b.bmax);  // grouped members: Np = N
}

__global__ __launch_bounds__(kThreads) void k_sched_advance_boxes(const MxBox* __restrict__ boxes, int nb) {
    const MxBox b = load_box(boxes, blockIdx.y);
    sched_advance_body(b.sched, b.bmax, nb, b.jhist, b.dscal + 1);
}

}  // namespace pm3d
