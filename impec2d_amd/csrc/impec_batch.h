// impec_batch.h -- batched launches over the members of an ensemble.
//
// A kernel body that would run over one context's grid is written once as a
// device function of the launch geometry it sees (G: bx(), by(), gx(), gy()).
// The single-context kernel passes the hardware grid; the batched kernel runs
// one flat grid over every member: member m owns the workgroups
// [start[m], start[m+1]) and sees them as its own gx x gy grid, with its own
// argument record tab[m].  Each member's workgroups therefore compute exactly
// what that member's own launch computes, partial sums included.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>

namespace impec {

// Build-time option, off by default.  The dispatcher hands the workgroups of a
// launch to the eight XCDs in turn (workgroup k to XCD k mod 8), and each XCD
// has its own L2.  With IMPEC_XCD_SWIZZLE=1, xcd_block renumbers the first
// 8 * floor(nb / 8) workgroups so that each XCD receives one consecutive range
// of logical workgroups, and tiles that share cache lines (edge columns of
// neighbouring strips, halo rows of neighbouring tiles) meet in one L2.
// Measured at 4096^2 on MI355X: FETCH_SIZE drops by up to 46% per kernel
// (the sweeps, velocities, CN tensors, CG update), yet every kernel runs 2-15%
// slower and the step 5% slower.  The duplicate fetches of the round-robin
// order are served by the Infinity Cache, while eight XCDs streaming eight
// distant regions lose DRAM locality.  A bijection on [0, nb); no kernel of
// this library waits on another workgroup, so the order carries no progress
// guarantee either way.
#ifndef IMPEC_XCD_SWIZZLE
#define IMPEC_XCD_SWIZZLE 0
#endif
__device__ __forceinline__ int xcd_block(int b, int nb) {
#if IMPEC_XCD_SWIZZLE
    const int per = nb >> 3;
    return b < (per << 3) ? (b & 7) * per + (b >> 3) : b;
#else
    (void)nb;
    return b;
#endif
}

// the hardware grid (renumbered by xcd_block when that is enabled, x fastest
// as the dispatcher walks them); partial sums indexed by by() * gx() + bx()
// land where the plain launch puts them
struct HwGrid {
#if IMPEC_XCD_SWIZZLE
    __device__ __forceinline__ int lin() const {
        return xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    }
    __device__ __forceinline__ int bx() const { return gridDim.y == 1 ? lin() : lin() % gridDim.x; }
    __device__ __forceinline__ int by() const { return gridDim.y == 1 ? 0 : lin() / gridDim.x; }
#else
    __device__ __forceinline__ int bx() const { return blockIdx.x; }
    __device__ __forceinline__ int by() const { return blockIdx.y; }
#endif
    __device__ __forceinline__ int gx() const { return gridDim.x; }
    __device__ __forceinline__ int gy() const { return gridDim.y; }
};

struct VGrid {
    int x, y, nx, ny;
    __device__ __forceinline__ int bx() const { return x; }
    __device__ __forceinline__ int by() const { return y; }
    __device__ __forceinline__ int gx() const { return nx; }
    __device__ __forceinline__ int gy() const { return ny; }
};

// member owning flat index b: the last m with start[m] <= b (start ascending,
// start[0] = 0); b is uniform, so the search runs on the scalar unit
__device__ __forceinline__ int member_of(const int *__restrict__ start, int nmem, int b) {
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// a batched launch: nmem argument records and the first workgroup (or wave)
// of each member, in device memory
template <class E>
struct Batch {
    E *tab = nullptr;
    int *start = nullptr;    // nmem + 1 entries
    int nmem = 0;
    int total = 0;           // workgroups (or waves) of the whole launch
};

template <class E>
inline void batch_free(Batch<E> &b) {
    if (b.tab) (void)hipFree(b.tab);
    if (b.start) (void)hipFree(b.start);
    b = Batch<E>();
}

// uploads the records and the prefix sums of `units` (workgroups or waves
// per member); members with zero units are dropped
template <class E>
inline hipError_t batch_upload(Batch<E> &b, const std::vector<E> &recs, const std::vector<int> &units) {
    batch_free(b);
    std::vector<E> keep;
    std::vector<int> start;
    int acc = 0;
    for (size_t k = 0; k < recs.size(); k++) {
        if (units[k] <= 0) continue;
        keep.push_back(recs[k]);
        start.push_back(acc);
        acc += units[k];
    }
    start.push_back(acc);
    b.nmem = (int)keep.size();
    b.total = acc;
    if (b.nmem == 0) return hipSuccess;
    hipError_t e = hipMalloc(&b.tab, sizeof(E) * keep.size());
    if (e == hipSuccess) e = hipMalloc(&b.start, sizeof(int) * start.size());
    if (e == hipSuccess) e = hipMemcpy(b.tab, keep.data(), sizeof(E) * keep.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.start, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice);
    return e;
}

}  // namespace impec

namespace impec {

// Workgroup loop of a batched kernel launched with at most as many
// workgroups as units: workgroup g takes the contiguous units
// [g*per, (g+1)*per), finds its first member once and steps to the next member
// when it crosses one (every listed member has at least one unit), so the
// member search is not repeated per unit.  f(R, G) runs the body for record R
// at position G of that member's own grid; a barrier separates the units, so
// bodies may use LDS.
template <class E, class F>
__device__ __forceinline__ void batch_for(const E *__restrict__ tab, const int *__restrict__ start, int nmem, F &&f) {
    const int total = start[nmem];
    const int per = (total + (int)gridDim.x - 1) / (int)gridDim.x;
    int vb = (int)blockIdx.x * per;
    const int end = vb + per < total ? vb + per : total;
    if (vb >= end) return;
    int k = member_of(start, nmem, vb);
    int kend = start[k + 1];
    for (; vb < end; vb++) {
        if (vb >= kend) {
            k++;
            kend = start[k + 1];
        }
        const E R = tab[k];
        const int loc = vb - start[k];
        f(R, VGrid{loc % R.gx, loc / R.gx, R.gx, R.gy});
        __syncthreads();
    }
}

// The members of a batch that are in a given loop state, compacted on the
// device: the same loop over the listed members only.
struct ActList {
    int *member;     // batch-local member ids, ascending
    int *start;      // first unit of each listed member; start[n] = all listed units
    int *n;          // number of listed members
};

template <class E, class F>
__device__ __forceinline__ void act_for(const E *__restrict__ tab, const ActList &A, F &&f) {
    const int nact = A.n[0];
    if (nact == 0) return;
    const int total = A.start[nact];
    const int per = (total + (int)gridDim.x - 1) / (int)gridDim.x;
    int vb = (int)blockIdx.x * per;
    const int end = vb + per < total ? vb + per : total;
    if (vb >= end) return;
    int k = member_of(A.start, nact, vb);
    int kend = A.start[k + 1];
    for (; vb < end; vb++) {
        if (vb >= kend) {
            k++;
            kend = A.start[k + 1];
        }
        const E R = tab[A.member[k]];
        const int loc = vb - A.start[k];
        f(R, VGrid{loc % R.gx, loc / R.gx, R.gx, R.gy});
        __syncthreads();
    }
}

// workgroups of a batched launch over `units` units
constexpr int BATCH_WG = 4096;
inline int batch_grid(int units) { return units < BATCH_WG ? units : BATCH_WG; }

}  // namespace impec
