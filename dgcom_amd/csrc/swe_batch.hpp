// Ensembles: independent runs of different meshes, orders and forcings
// advanced together. Each member keeps its own statics (a dg_ctx form) and its
// own adaptive time step; the members of one order share one launch per kernel
// over a work list of element blocks (the blocks of every member of that order,
// member after member), so that a stage of the whole ensemble is a few large
// launches instead of a few per member. The arithmetic per element is the
// single-run kernels' (stage_block, dt_block), so every member's run is bit for
// bit the run it makes alone. Included by dgswe.hip after swe_kernels.hpp.
#pragma once

#include "swe_kernels.hpp"

namespace dgswe {

struct BatchMember {
    StageParams sp;           // statics and physics; the state pointers, stage coefficients and queues are set per launch
    DtParams dp;              // statics of the time step; the state pointer is set per launch
    double* x[2];             // the state, ping-pong between steps
    double* y[4];             // stage values
    uint32_t block0;          // first block of the member in its order's work list
    uint32_t n_blocks;
    const double* block_min;  // its order's per-block time steps
};

struct BatchCoef { double a, b, c, eta; };  // one stage of one member

struct BatchArgs {
    const BatchMember* mem;
    const uint32_t* block_member;  // [n_blocks]: the member of each block of this order
    const BatchCoef* coef;         // [members][4]: this step's stage coefficients
    uint32_t* queue;               // [n_blocks][EPB]: subcell elements per block
    uint8_t* queue_count;          // [n_blocks]
    double* block_min;             // [n_blocks]
    unsigned int n_blocks, xcd_stride;
    int q, ns, parity;
};

// The member's stage parameters for stage A.q: base = the state, input = the
// previous stage value, out = the next stage value (the new state at the last)
template <int N>
__device__ __forceinline__ void batch_stage_params(const BatchArgs& A, uint32_t m, StageParams& p) {
    constexpr int NN = (N + 1) * (N + 1);
    const BatchMember& M = A.mem[m];
    p = M.sp;
    const size_t nt = (size_t)p.n_elements * NN;
    double* const x0 = A.parity ? M.x[1] : M.x[0];
    double* const x1 = A.parity ? M.x[0] : M.x[1];
    double* const yq = A.q == 0 ? M.y[0] : A.q == 1 ? M.y[1] : A.q == 2 ? M.y[2] : M.y[3];
    double* const yp = A.q == 1 ? M.y[0] : A.q == 2 ? M.y[1] : M.y[2];
    const double* base = x0;
    const double* input = A.q == 0 ? x0 : yp;
    double* out = A.q == A.ns - 1 ? x1 : yq;
    p.bh = base; p.bhu = base + nt; p.bhv = base + 2 * nt;
    p.ih = input; p.ihu = input + nt; p.ihv = input + 2 * nt;
    p.oh = out; p.ohu = out + nt; p.ohv = out + 2 * nt;
    const BatchCoef cf = A.coef[(size_t)m * 4 + A.q];
    p.a = cf.a; p.b = cf.b; p.c = cf.c; p.eta_e = cf.eta;
    p.rhs_only = 0;
    p.fstar_out = nullptr;
    p.rh = p.rhu = p.rhv = nullptr;
}

// The wet kernel of one order over the ensemble's blocks of that order
template <int N>
__global__ __launch_bounds__(64) DG_WET_OCCUPANCY void batch_wet_kernel(BatchArgs A) {
    constexpr int NN = (N + 1) * (N + 1), EPB = 64 / NN;
    const unsigned int blk = xcd_block(blockIdx.x, A.xcd_stride);
    if (blk >= A.n_blocks) return;
    const uint32_t m = A.block_member[blk];
    StageParams p;
    batch_stage_params<N>(A, m, p);
    const uint32_t block0 = A.mem[m].block0;
    p.queue = A.queue + (size_t)block0 * EPB;
    p.queue_count = A.queue_count + block0;
    const int64_t lb = (int64_t)(blk - block0);
    const int e = threadIdx.x / NN;
    const int64_t k = lb * EPB + (e < EPB ? e : 0);
    stage_block<N, false>(p, k < p.n_elements ? k : p.n_elements - 1, e < EPB && lb * EPB + e < p.n_elements);
}

// Compaction of the per-block subcell lists of one order, with the member of
// each entry (compact_kernel's scan, after chunk_count_kernel)
__global__ __launch_bounds__(256) void batch_compact_kernel(const uint8_t* counts, const uint32_t* queue,
                                                            const uint32_t* block_member, unsigned int n_blocks,
                                                            const unsigned int* chunk_total, int epb, uint32_t* list,
                                                            uint32_t* list_member, unsigned int* list_count) {
    const unsigned int b0 = blockIdx.x * 1024u + threadIdx.x * 4u;
    unsigned int cnt[4], s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) { cnt[q] = b0 + q < n_blocks ? counts[b0 + q] : 0u; s += cnt[q]; }
    unsigned int base, excl, total;
    chunk_offsets(chunk_total, s, base, excl, total);
    unsigned int off = base + excl;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        for (unsigned int r = 0; r < cnt[q]; ++r) {
            list[off] = queue[(size_t)(b0 + q) * epb + r];
            list_member[off] = block_member[b0 + q];
            ++off;
        }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *list_count = base + total;
}

// The listed shallow elements of one order, EPB per workgroup pass: a pass takes
// the elements of one member at a time (the first unprocessed entry's), so that
// the member's parameters stay uniform over the wavefront.
template <int N>
__global__ __launch_bounds__(64) DG_SUB_OCCUPANCY void batch_subcell_kernel(BatchArgs A, const uint32_t* list,
                                                           const uint32_t* list_member, const unsigned int* count_ptr) {
    constexpr int NN = (N + 1) * (N + 1), EPB = 64 / NN;
    const unsigned int count = *count_ptr;
    const int e = threadIdx.x / NN;
    for (size_t first = (size_t)blockIdx.x * EPB; first < count; first += (size_t)gridDim.x * EPB) {
        const size_t last = first + EPB < count ? first + EPB : count;
        const bool mine = e < EPB && first + e < count;
        const uint32_t my_member = mine ? list_member[first + e] : 0xFFFFFFFFu;
        size_t next = first;
        while (next < last) {
            const uint32_t m = list_member[next];  // uniform
            StageParams p;
            batch_stage_params<N>(A, m, p);
            const bool active = mine && my_member == m;
            stage_block<N, true>(p, active ? (int64_t)list[first + e] : 0, active);
            lds_sync();
            // the entries are in member order: skip this member's
            while (next < last && list_member[next] == m) ++next;
        }
    }
}

// Per-block time steps of one order's blocks (one workgroup per block)
template <int N>
__global__ __launch_bounds__(64) void batch_dt_kernel(BatchArgs A) {
    const unsigned int blk = xcd_block(blockIdx.x, A.xcd_stride);
    if (blk >= A.n_blocks) return;
    const uint32_t m = A.block_member[blk];
    const BatchMember& M = A.mem[m];
    DtParams p;  // field by field (an aggregate copy of the record would go through scratch)
    p.corners = M.dp.corners; p.emass = M.dp.emass; p.links = M.dp.links; p.affine = M.dp.affine; p.n_elements = M.dp.n_elements;
    p.cfl = M.dp.cfl; p.g = M.dp.g; p.h_min = M.dp.h_min; p.h_dry = M.dp.h_dry;
    p.pcfl = M.dp.pcfl; p.max_cfl = M.dp.max_cfl; p.max_cfl_sub = M.dp.max_cfl_sub; p.positivity = M.dp.positivity;
    p.element_dt = nullptr; p.slots = nullptr; p.xcd_stride = 0;
    const size_t nt = (size_t)p.n_elements * (N + 1) * (N + 1);
    p.h = A.parity ? M.x[1] : M.x[0];  // a select, not an indexed load (that would copy the record to scratch)
    p.hu = p.h + nt; p.hv = p.h + 2 * nt;
    const double v = dt_block<N>(p, (int64_t)(blk - M.block0));
    if (threadIdx.x == 0) A.block_min[blk] = v;
}

// Each member's time step: the minimum over its blocks (one workgroup per member)
__global__ __launch_bounds__(256) void batch_member_min_kernel(const BatchMember* mem, int n_members, double* member_dt) {
    const int m = blockIdx.x;
    if (m >= n_members) return;
    __shared__ double red[256];
    const BatchMember& M = mem[m];
    double v = INFINITY;
    for (uint32_t b = threadIdx.x; b < M.n_blocks; b += blockDim.x) v = fmin_(v, M.block_min[M.block0 + b]);
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmin_(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) member_dt[m] = red[0];
}

}  // namespace dgswe
