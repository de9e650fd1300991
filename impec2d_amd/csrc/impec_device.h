// impec_device.h -- device-side helpers shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include "impec_internal.h"
#include "impec_batch.h"

namespace impec {

constexpr int WAVE = 64;

__device__ __forceinline__ int tid_linear() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}
__device__ __forceinline__ int block_threads() { return blockDim.x * blockDim.y * blockDim.z; }

enum RedOp { RED_SUM = 0, RED_MAX = 1, RED_MIN = 2 };

template <int OP>
__device__ __forceinline__ double red_combine(double a, double b) {
    if (OP == RED_SUM) return a + b;
    if (OP == RED_MAX) return b > a ? b : a;
    return b < a ? b : a;
}

template <int OP>
__device__ __forceinline__ double red_identity() {
    if (OP == RED_SUM) return 0.0;
    if (OP == RED_MAX) return -DBL_MAX;
    return DBL_MAX;
}

// Deterministic block reduction of NV values: wavefront shuffle tree, then
// wave 0 combines the per-wave results in index order.  Thread 0 returns the
// block result in v[].  Every thread of the block must call it.
template <int OP, int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV]) {
    __shared__ double sh[NV][16];
    const int t = tid_linear(), lane = t & (WAVE - 1), wv = t / WAVE;
    const int nw = (block_threads() + WAVE - 1) / WAVE;
#pragma unroll
    for (int k = 0; k < NV; k++)
        for (int off = WAVE / 2; off > 0; off >>= 1) v[k] = red_combine<OP>(v[k], __shfl_down(v[k], off, WAVE));
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[k][wv] = v[k];
    __syncthreads();
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = sh[k][0];
            for (int w = 1; w < nw; w++) s = red_combine<OP>(s, sh[k][w]);
            v[k] = s;
        }
    }
}

// the share of thread t of T threads in a combine over n partials p[0..n)
// read with stride T, eight independent loads in flight (a fixed order: the
// result does not depend on timing)
// DEPTH: strides of eight loaded before any is combined (more loads in flight,
// the same combination order)
template <int OP, int DEPTH = 1>
__device__ __forceinline__ double strided_combine_at(const double *__restrict__ p, int n, int t, int T) {
    double acc[8];
#pragma unroll
    for (int u = 0; u < 8; u++) acc[u] = red_identity<OP>();
    int b = t;
    if (DEPTH > 1)
        for (; b + (8 * DEPTH - 1) * T < n; b += 8 * DEPTH * T) {
            double v[8 * DEPTH];
#pragma unroll
            for (int u = 0; u < 8 * DEPTH; u++) v[u] = p[b + u * T];
#pragma unroll
            for (int q = 0; q < DEPTH; q++)
#pragma unroll
                for (int u = 0; u < 8; u++) acc[u] = red_combine<OP>(acc[u], v[8 * q + u]);
        }
    for (; b + 7 * T < n; b += 8 * T) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = p[b + u * T];
#pragma unroll
        for (int u = 0; u < 8; u++) acc[u] = red_combine<OP>(acc[u], v[u]);
    }
    // the rest (fewer than eight strides): loads issued together, combined in
    // order into acc[0]; a missing one is the identity, which leaves the
    // combination unchanged
    {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = b + u * T < n ? p[b + u * T] : red_identity<OP>();
#pragma unroll
        for (int u = 0; u < 8; u++) acc[0] = red_combine<OP>(acc[0], v[u]);
    }
#pragma unroll
    for (int u = 1; u < 8; u++) acc[0] = red_combine<OP>(acc[0], acc[u]);
    return acc[0];
}
// the share of thread threadIdx.x of blockDim.x threads
template <int OP>
__device__ __forceinline__ double strided_combine(const double *__restrict__ p, int n) {
    return strided_combine_at<OP>(p, n, threadIdx.x, blockDim.x);
}

// strided_combine of NV vectors p + q * vstride (q < NV) at once: each vector
// combined in strided_combine's order, the loads of all NV issued together
template <int OP, int NV>
__device__ __forceinline__ void strided_combine_nv(const double *__restrict__ p, long vstride, int n, double (&out)[NV]) {
    const int T = blockDim.x;
    double acc[NV][8];
#pragma unroll
    for (int q = 0; q < NV; q++)
#pragma unroll
        for (int u = 0; u < 8; u++) acc[q][u] = red_identity<OP>();
    int b = threadIdx.x;
    for (; b + 7 * T < n; b += 8 * T) {
        double v[NV][8];
#pragma unroll
        for (int q = 0; q < NV; q++)
#pragma unroll
            for (int u = 0; u < 8; u++) v[q][u] = p[q * vstride + b + u * T];
#pragma unroll
        for (int q = 0; q < NV; q++)
#pragma unroll
            for (int u = 0; u < 8; u++) acc[q][u] = red_combine<OP>(acc[q][u], v[q][u]);
    }
    {
        double v[NV][8];
#pragma unroll
        for (int q = 0; q < NV; q++)
#pragma unroll
            for (int u = 0; u < 8; u++) v[q][u] = b + u * T < n ? p[q * vstride + b + u * T] : red_identity<OP>();
#pragma unroll
        for (int q = 0; q < NV; q++)
#pragma unroll
            for (int u = 0; u < 8; u++) acc[q][0] = red_combine<OP>(acc[q][0], v[q][u]);
    }
#pragma unroll
    for (int q = 0; q < NV; q++) {
#pragma unroll
        for (int u = 1; u < 8; u++) acc[q][0] = red_combine<OP>(acc[q][0], acc[q][u]);
        out[q] = acc[q][0];
    }
}

// mu(c) of EllipticSolver.cpp:692-695
__device__ __forceinline__ double mu_of(double c, double mob1over4, double mu_res) {
    double x = (1. - c + mob1over4 * c);
    return 1. / (x * x * x * x) * mu_res;
}

// face mobility between two cells from the stored per-cell quantity w
// (EllipticSolver.cpp:478-524): harmonic 2/(w_l + w_r), mean 1/2 (w_l + w_r)
__device__ __forceinline__ double face_mob(double wl, double wr, int harmonic) {
    return harmonic ? 2. / (wl + wr) : 1. / 2 * (wl + wr);
}

// vertex mobility from the four cells around it, in the reference's
// summation order (i+1,j), (i,j), (i,j+1), (i+1,j+1) (EllipticSolver.cpp:697-754)
__device__ __forceinline__ double vertex_mob(double w_ip1j, double w_ij, double w_ijp1, double w_ip1jp1,
                                             int harmonic) {
    double s = w_ip1j + w_ij + w_ijp1 + w_ip1jp1;
    return harmonic ? 4. / s : 1. / 4 * s;
}

// lane l receives v of lane l-1 (lane 0 receives 0): DPP wave_shr:1 (a register move,
// no memory traffic).  bound_ctrl set: the lane with no source reads 0, so the
// move needs no prior value in its destination.
__device__ __forceinline__ double lane_prev(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
// lane l receives v of lane l+1 (lane 63 receives 0): DPP wave_shl:1
__device__ __forceinline__ double lane_next(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_prev(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_next(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

}  // namespace impec
