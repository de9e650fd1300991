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

struct HwGrid {
    __device__ __forceinline__ int bx() const { return blockIdx.x; }
    __device__ __forceinline__ int by() const { return blockIdx.y; }
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

// Prologue of a batched kernel over workgroups: member m_ of flat workgroup
// blockIdx.x, its record R (which carries the member's own grid gx x gy) and
// the workgroup's position G in that grid.
#define IMPEC_BATCH_BLOCK(R, G, tab, start, nmem)                                  \
    const int m_ = ::impec::member_of((start), (nmem), (int)blockIdx.x);           \
    const auto R = (tab)[m_];                                                      \
    const int loc_ = (int)blockIdx.x - (start)[m_];                                \
    const ::impec::VGrid G{loc_ % R.gx, loc_ / R.gx, R.gx, R.gy};
