// noalloc.hip -- the allocation-free variant of impec_max_stable_timestep.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// serving loops) pass the face velocities they hold on the device, in the
// reference's host layout (Array3d v, component last, base -b), and a device
// word for the result.  The reduction's partials live in `scratch`, a
// file-scope device array of NOALLOC_PARTS x 3 doubles that every call on a
// GPU shares: each call writes all the partials it reads (one per workgroup
// of its launch), so nothing has to start at zero, and calls that overlap on
// one GPU race on it.  The per-cell arithmetic and the combine are those of
// impec_max_stable_timestep (HyperbolicSolver::MaxStableTimestep,
// HyperbolicSolver.cpp:93-198): max and min are exact, so both entry points
// return the same time step bit for bit.
#include "impec_device.h"
#include <float.h>
#include <math.h>
#include <mutex>
#include <stdarg.h>
#include <stdio.h>

namespace impec {

// workgroups of the reduction (and partials per quantity); 256 threads each
constexpr int NOALLOC_PARTS = 2048;
constexpr int NOALLOC_MAX_DEVICES = 64;
// velocity arrays up to 2^31 - 1 doubles (int offsets inside the kernel)
constexpr long long NOALLOC_MAX_ELEMS = 2147483647LL;

static __device__ __align__(256) double scratch[3 * NOALLOC_PARTS];

// the device address of `scratch` on each GPU, taken once per device with
// that device current (host code never passes the array's name itself)
static void *scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex scratch_mu;

static thread_local char noalloc_err[256];

int noalloc_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(noalloc_err, sizeof(noalloc_err), fmt, ap);
    va_end(ap);
    return code;
}

const char *noalloc_last_error() { return noalloc_err; }

struct RefCfl {
    int N1, N2, B, ldv;   // ldv = N2 + 1 + 2B interleaved pairs per row
    int mode, diff;       // advection-only (test mode); explicit dispersion limit
    double dx, dy, al, at;
};

// the per-cell terms of cfl_body, on the interleaved layout
__global__ void __launch_bounds__(256) k_cfl_ref(RefCfl k, const double *__restrict__ V, double *__restrict__ part) {
    const long long n = (long long)k.N1 * k.N2;
    double mx[1] = {0.0}, mn[2] = {DBL_MAX, 1.e10};
    for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n;
         id += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(id / k.N2), j = (int)(id - (long long)i * k.N2);
        const long long o = ((long long)(i + k.B) * k.ldv + (j + k.B)) * 2;
        const double u0 = V[o], u1 = V[o + 2LL * k.ldv], v0 = V[o + 1], v1 = V[o + 3];
        double uc = fabs(0.5 * (u0 + u1));
        double vc = fabs(0.5 * (v0 + v1));
        if (uc < DBL_MIN) uc = DBL_MIN;
        if (vc < DBL_MIN) vc = DBL_MIN;
        if (mx[0] < uc) mx[0] = uc;
        if (mx[0] < vc) mx[0] = vc;
        if (k.mode) {
            const double a = k.dx / uc, b = k.dy / vc;
            const double dt = b < a ? b : a;
            if (dt < mn[0]) mn[0] = dt;
        }
        if (k.diff) {
            const double vij = 0.5 * (v0 + v1), uij = 0.5 * (u0 + u1);
            const double vij2 = vij * vij, uij2 = uij * uij;
            const double absu = sqrt(vij2 + uij2);
            const double d11 = k.dx / absu * (k.al * uij2 + k.at * vij2);
            const double d22 = k.dx / absu * (k.al * vij2 + k.at * uij2);
            const double dt = 1. / 2 * k.dx * k.dx / (d11 + d22);
            if (dt < mn[1]) mn[1] = dt;
        }
    }
    block_reduce<RED_MAX, 1>(mx);
    block_reduce<RED_MIN, 2>(mn);
    if (tid_linear() == 0) {
        part[0 * NOALLOC_PARTS + blockIdx.x] = mx[0];
        part[1 * NOALLOC_PARTS + blockIdx.x] = mn[0];
        part[2 * NOALLOC_PARTS + blockIdx.x] = mn[1];
    }
}

struct RefDt {
    int nb, is_test, diffusion, cn;
    double dx, dy, cfl, dt_diff_limiter;
};

// combine the partials and form the step as launch_cfl does on the host
__global__ void __launch_bounds__(256) k_cfl_ref_dt(RefDt d, const double *__restrict__ part, double *__restrict__ dt_out) {
    double mx[1] = {0.0}, mn[2] = {DBL_MAX, 1.e10};
    for (int b = threadIdx.x; b < d.nb; b += blockDim.x) {
        const double a = part[b], m0 = part[NOALLOC_PARTS + b], m1 = part[2 * NOALLOC_PARTS + b];
        if (a > mx[0]) mx[0] = a;
        if (m0 < mn[0]) mn[0] = m0;
        if (m1 < mn[1]) mn[1] = m1;
    }
    block_reduce<RED_MAX, 1>(mx);
    block_reduce<RED_MIN, 2>(mn);
    if (threadIdx.x != 0) return;
    double dt;
    if (d.is_test) {
        dt = mn[0] * d.cfl;   // MaxStableAdvectionTimestep (HyperbolicSolver.cpp:100-121)
    } else {                  // MaxStableTimestepCombined (HyperbolicSolver.cpp:123-198)
        const double maxv = mx[0];
        const double a = d.dx / maxv, b = d.dy / maxv;
        double dt_adv = b < a ? b : a;
        dt_adv *= d.cfl;
        double dt_diff = DBL_MAX;
        if (d.diffusion) dt_diff = mn[1] * d.dt_diff_limiter;
        if (d.cn) dt = dt_adv;
        else dt = dt_adv < dt_diff ? dt_adv : dt_diff;
    }
    dt_out[0] = dt;
}

static int scratch_address(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return noalloc_error(IMPEC_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return noalloc_error(IMPEC_EINVAL, "device %d: at most %d devices", dev, NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return noalloc_error(IMPEC_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                                 hipGetErrorString(e));
        scratch_ptr[dev] = p;
    }
    *out = (double *)scratch_ptr[dev];
    return IMPEC_OK;
}

}  // namespace impec

using namespace impec;

extern "C" int impec_max_stable_timestep_noalloc(const impec_params *P, const double *v_dev, double *dt_dev,
                                                 void *stream) {
    if (!P || !v_dev || !dt_dev) return noalloc_error(IMPEC_EINVAL, "null argument");
    if (P->N1 < 1 || P->N2 < 1 || P->num_ghostcells != 2)
        return noalloc_error(IMPEC_EINVAL, "grid %d x %d with %d ghost cells: need N1, N2 >= 1 and 2 ghost cells",
                             P->N1, P->N2, P->num_ghostcells);
    const long long b = P->num_ghostcells;
    const long long elems = ((long long)P->N1 + 1 + 2 * b) * ((long long)P->N2 + 1 + 2 * b) * 2;
    if (elems > NOALLOC_MAX_ELEMS)
        return noalloc_error(IMPEC_EINVAL, "velocity array of %lld doubles exceeds the maximum of %lld", elems,
                             NOALLOC_MAX_ELEMS);
    if (!(P->x_ub > P->x_lb) || !(P->y_ub > P->y_lb) || !(P->cfl > 0))
        return noalloc_error(IMPEC_EINVAL, "domain or cfl not positive");
    double *part = nullptr;
    int st = scratch_address(&part);
    if (st) return st;
    RefCfl k;
    k.N1 = P->N1; k.N2 = P->N2; k.B = (int)b; k.ldv = P->N2 + 1 + 2 * (int)b;
    k.mode = P->is_test;
    k.diff = P->diffusion && P->para_method != IMPEC_PARA_CN && !P->is_test;
    k.dx = (P->x_ub - P->x_lb) / P->N1;
    k.dy = (P->y_ub - P->y_lb) / P->N2;
    k.al = P->alpha_l; k.at = P->alpha_t;
    const long long n = (long long)P->N1 * P->N2;
    long long nb = (n + 255) / 256;
    if (nb > NOALLOC_PARTS) nb = NOALLOC_PARTS;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_cfl_ref, dim3((unsigned)nb), dim3(256), 0, s, k, v_dev, part);
    RefDt d;
    d.nb = (int)nb; d.is_test = P->is_test; d.diffusion = P->diffusion;
    d.cn = P->para_method == IMPEC_PARA_CN;
    d.dx = k.dx; d.dy = k.dy; d.cfl = P->cfl; d.dt_diff_limiter = P->dt_diff_limiter;
    hipLaunchKernelGGL(k_cfl_ref_dt, dim3(1), dim3(256), 0, s, d, part, dt_dev);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return noalloc_error(IMPEC_EHIP, "launch: %s", hipGetErrorString(e));
    return IMPEC_OK;
}
