// k_cnn.hip -- the learned initial guess of the pressure solve: a small CNN
// (3x3 convolutions, 32 channels, ReLU) that maps (log mobility, previous
// pressure, source) to the pressure increment, with its own forward, backward
// and Adam kernels.  The 32->32 layers run on MFMA (v_mfma_f32_32x32x16_bf16:
// BF16 weights and activations, FP32 accumulation); the 3->32 input layer and
// the 32->1 output layer are FP32 dot products.  Weight gradients are summed
// in fixed order (per-block partials, then one pass per weight), so training
// is deterministic.  Multigrid-CG decides convergence: the network only supplies
// the initial guess and every solve still runs to the same tolerance.
//
// Tensors are NHWC: [image][row][column][channel].  Convolution:
//   y[p][co] = b[co] + sum_{tap, ci} W[tap][ci][co] x[p + off(tap)][ci],
//   off(tap) = (tap / 3 - 1, tap % 3 - 1), zero outside the image.
#include "impec_device.h"
#include <math.h>
#include <algorithm>
#include <random>
#include <vector>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int CH = 32;          // hidden channels (one 32x32 MFMA tile)
constexpr int TAPS = 9;
constexpr int MID_TR = 8;       // row tiles (of 4 rows) per workgroup of the MFMA convolution
constexpr int LDS_PIX = 40;     // LDS pitch of one pixel's channels (bf16), padded against bank conflicts

struct impec_cnn {
    int L;                      // conv layers: 3->32, (L-2) x 32->32, 32->1
    int crop, capacity;
    unsigned long long seed;
    long long nparam;
    std::vector<long long> woff, boff;   // offsets of W_l, b_l in the parameter vector
    float *param = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    long long adam_t = 0;
    __bf16 *wf = nullptr;       // MFMA forward weights of the mid layers, [l][tap][co][ci]
    __bf16 *wb = nullptr;       // MFMA backward-data weights, [l][tap'][ci][co] = W[8-tap'][ci][co]
    float *pairs = nullptr;     // training pairs: capacity x crop x crop x 4 (3 features + target)
    int count = 0, next = 0;
    float tscale = 0.f;         // target scale (set by the first collected field)
    int loss_kind = 1;          // 0: mean squared error, 1: H1 seminorm + small L2 (see k_loss)
    // work buffers, grown on demand
    long long act_pix = 0;      // pixels per activation buffer
    int act_nbuf = 0;           // activation buffers: L-1 when training, 2 for inference
    __bf16 *act = nullptr;
    long long train_pix = 0;    // pixels of the training-only buffers
    __bf16 *dz = nullptr;       // two backward buffers
    float *x0 = nullptr, *y = nullptr, *dy = nullptr, *t = nullptr;
    long long part_len = 0;
    float *part = nullptr;      // per-block partial gradients
    double *stat = nullptr;     // field statistics (device)
    double *stat_part = nullptr;
    int *idx = nullptr;         // batch pair indices
    int idx_len = 0;
    hipStream_t stream = nullptr;
    char err[256];
};

namespace impec {

static int cnn_err(impec_cnn *n, int code, const char *msg) {
    snprintf(n->err, sizeof(n->err), "%s", msg);
    return code;
}
#define CNN_CHECK(net, call)                                                                    \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            snprintf((net)->err, sizeof((net)->err), "%s: %s", #call, hipGetErrorString(e_));   \
            return IMPEC_EHIP;                                                                  \
        }                                                                                       \
    } while (0)

__device__ __forceinline__ int tap_dy(int tap) { return tap / 3 - 1; }
__device__ __forceinline__ int tap_dx(int tap) { return tap % 3 - 1; }

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
// 3 -> 32, FP32 inputs, bias + ReLU, BF16 output; one thread per pixel
__global__ void __launch_bounds__(256) k_conv_first(const float *__restrict__ x, const float *__restrict__ W,
                                                    const float *__restrict__ b, __bf16 *__restrict__ y, int N, int H,
                                                    int Wd) {
    __shared__ float sw[TAPS * 3 * CH + CH];
    for (int k = threadIdx.x; k < TAPS * 3 * CH + CH; k += blockDim.x) sw[k] = k < TAPS * 3 * CH ? W[k] : b[k - TAPS * 3 * CH];
    __syncthreads();
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)N * H * Wd) return;
    const int n = (int)(p / ((long long)H * Wd)), rem = (int)(p - (long long)n * H * Wd), r = rem / Wd, c = rem % Wd;
    float acc[CH];
#pragma unroll
    for (int co = 0; co < CH; co++) acc[co] = sw[TAPS * 3 * CH + co];
    for (int tap = 0; tap < TAPS; tap++) {
        const int rr = r + tap_dy(tap), cc = c + tap_dx(tap);
        if (rr < 0 || rr >= H || cc < 0 || cc >= Wd) continue;
        const float *xp = x + (((long long)n * H + rr) * Wd + cc) * 3;
        for (int ci = 0; ci < 3; ci++) {
            const float xv = xp[ci];
            const float *wr = sw + (tap * 3 + ci) * CH;
#pragma unroll
            for (int co = 0; co < CH; co++) acc[co] = fmaf(wr[co], xv, acc[co]);
        }
    }
    __bf16 *yp = y + p * CH;
#pragma unroll
    for (int co = 0; co < CH; co++) yp[co] = (__bf16)fmaxf(acc[co], 0.f);
}

// 32 -> 32 on MFMA.  FWD: y = relu(conv(x) + b).  !FWD (backward data):
// y = conv(x) * (mask > 0), with the flipped, transposed weights.  A workgroup
// (4 waves) covers 32 columns and MID_TR tiles of 4 rows, one row per wave;
// the weights stay in registers (18 K-steps x 8 BF16 per lane), the input
// rows (+1 halo) are staged in LDS.
template <bool FWD>
__global__ void __launch_bounds__(256) k_conv_mid(const __bf16 *__restrict__ x, const __bf16 *__restrict__ wt,
                                                  const float *__restrict__ bias, const __bf16 *__restrict__ mask,
                                                  __bf16 *__restrict__ y, int N, int H, int Wd) {
    __shared__ __bf16 tile[6 * 34 * LDS_PIX];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = blockIdx.x * 32, n = blockIdx.z;
    const int i = lane & 31, half = lane >> 5;
    // B operand: lane holds W[tap][co = i][ci = 16 ks + 8 half + 0..7]
    bf16x8 wreg[TAPS][2];
#pragma unroll
    for (int tap = 0; tap < TAPS; tap++)
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
            wreg[tap][ks] = *(const bf16x8 *)(wt + ((long long)(tap * CH + i) * CH + 16 * ks + 8 * half));
    float bco = 0.f;
    if (FWD) bco = bias[i];
    for (int tr = 0; tr < MID_TR; tr++) {
        const int r0 = (blockIdx.y * MID_TR + tr) * 4;
        if (r0 >= H) break;   // uniform over the workgroup
        __syncthreads();
        // stage rows r0-1 .. r0+4, columns c0-1 .. c0+32 (zero outside)
        for (int k = threadIdx.x; k < 6 * 34 * 4; k += 256) {
            const int q = k & 3, pix = k >> 2, rr = pix / 34, cc = pix % 34;
            const int gr = r0 - 1 + rr, gc = c0 - 1 + cc;
            bf16x8 v;
            if (gr >= 0 && gr < H && gc >= 0 && gc < Wd)
                v = *(const bf16x8 *)(x + (((long long)n * H + gr) * Wd + gc) * CH + 8 * q);
            else
                for (int e = 0; e < 8; e++) v[e] = (__bf16)0.f;
            *(bf16x8 *)(tile + (rr * 34 + cc) * LDS_PIX + 8 * q) = v;
        }
        __syncthreads();
        const int row = r0 + wave;
        f32x16 acc = {};
#pragma unroll
        for (int tap = 0; tap < TAPS; tap++) {
            const int rr = wave + 1 + tap_dy(tap), cc = i + 1 + tap_dx(tap);
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const bf16x8 a = *(const bf16x8 *)(tile + (rr * 34 + cc) * LDS_PIX + 16 * ks + 8 * half);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, wreg[tap][ks], acc, 0, 0, 0);
            }
        }
        if (row < H) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int pi = 8 * (r >> 2) + 4 * half + (r & 3);   // pixel (row of D); channel = i
                const int col = c0 + pi;
                if (col >= Wd) continue;
                const long long o = (((long long)n * H + row) * Wd + col) * CH + i;
                float v = acc[r];
                if (FWD) v = fmaxf(v + bco, 0.f);
                else v = (float)mask[o] > 0.f ? v : 0.f;
                y[o] = (__bf16)v;
            }
        }
    }
}

// 32 -> 1, FP32 output (no activation); one thread per pixel
__global__ void __launch_bounds__(256) k_conv_last(const __bf16 *__restrict__ x, const float *__restrict__ W,
                                                   const float *__restrict__ b, float *__restrict__ y, int N, int H,
                                                   int Wd) {
    __shared__ float sw[TAPS * CH + 1];
    for (int k = threadIdx.x; k < TAPS * CH + 1; k += blockDim.x) sw[k] = k < TAPS * CH ? W[k] : b[0];
    __syncthreads();
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)N * H * Wd) return;
    const int n = (int)(p / ((long long)H * Wd)), rem = (int)(p - (long long)n * H * Wd), r = rem / Wd, c = rem % Wd;
    float acc = sw[TAPS * CH];
    for (int tap = 0; tap < TAPS; tap++) {
        const int rr = r + tap_dy(tap), cc = c + tap_dx(tap);
        if (rr < 0 || rr >= H || cc < 0 || cc >= Wd) continue;
        const __bf16 *xp = x + (((long long)n * H + rr) * Wd + cc) * CH;
        for (int ci = 0; ci < CH; ci++) acc = fmaf(sw[tap * CH + ci], (float)xp[ci], acc);
    }
    y[p] = acc;
}

// ---------------------------------------------------------------------------
// loss and backward
// ---------------------------------------------------------------------------
// Losses over e = y - t, normalised by the pixel count:
//   MSE (kind 0):  sum e^2
//   H1  (kind 1):  sum over the image's inner faces (e_p - e_q)^2 + H1_L2 sum e^2
//                  (the gradient mismatch, which is what an initial guess of an
//                  elliptic solve is judged by, plus a small L2 term that
//                  fixes the constant)
// dy = dLoss/dy; per-block partials of the loss
constexpr float H1_L2 = 0.01f;
__global__ void k_loss(const float *__restrict__ y, const float *__restrict__ t, float *__restrict__ dy, int N, int H,
                       int Wd, int kind, double *__restrict__ part) {
    const long long npix = (long long)N * H * Wd;
    double s[1] = {0.};
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += (long long)gridDim.x * blockDim.x) {
        const float e = y[p] - t[p];
        if (kind == 0) {
            dy[p] = 2.f * e / (float)npix;
            s[0] += (double)e * e;
            continue;
        }
        const int rem = (int)(p % ((long long)H * Wd)), r = rem / Wd, c = rem % Wd;
        float g = 2.f * H1_L2 * e;
        double en = (double)H1_L2 * e * e;
        if (c + 1 < Wd) { const float d = e - (y[p + 1] - t[p + 1]); g += 2.f * d; en += (double)d * d; }
        if (c > 0) { const float d = e - (y[p - 1] - t[p - 1]); g += 2.f * d; }
        if (r + 1 < H) { const float d = e - (y[p + Wd] - t[p + Wd]); g += 2.f * d; en += (double)d * d; }
        if (r > 0) { const float d = e - (y[p - Wd] - t[p - Wd]); g += 2.f * d; }
        dy[p] = g / (float)npix;
        s[0] += en;
    }
    block_reduce<RED_SUM, 1>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// data gradient of the output layer, masked by the ReLU of its input:
// dz[p][ci] = (sum_tap W[tap][ci] dy[p - off(tap)]) * (a[p][ci] > 0)
__global__ void __launch_bounds__(256) k_last_back(const float *__restrict__ dy, const float *__restrict__ W,
                                                   const __bf16 *__restrict__ a, __bf16 *__restrict__ dz, int N, int H,
                                                   int Wd) {
    __shared__ float sw[TAPS * CH];
    for (int k = threadIdx.x; k < TAPS * CH; k += blockDim.x) sw[k] = W[k];
    __syncthreads();
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)N * H * Wd) return;
    const int n = (int)(p / ((long long)H * Wd)), rem = (int)(p - (long long)n * H * Wd), r = rem / Wd, c = rem % Wd;
    float acc[CH];
#pragma unroll
    for (int ci = 0; ci < CH; ci++) acc[ci] = 0.f;
    for (int tap = 0; tap < TAPS; tap++) {
        const int rr = r - tap_dy(tap), cc = c - tap_dx(tap);
        if (rr < 0 || rr >= H || cc < 0 || cc >= Wd) continue;
        const float d = dy[((long long)n * H + rr) * Wd + cc];
#pragma unroll
        for (int ci = 0; ci < CH; ci++) acc[ci] = fmaf(sw[tap * CH + ci], d, acc[ci]);
    }
#pragma unroll
    for (int ci = 0; ci < CH; ci++) dz[p * CH + ci] = (__bf16)((float)a[p * CH + ci] > 0.f ? acc[ci] : 0.f);
}

// weight and bias gradient partials of one layer over a tile of 4 x 32 pixels:
//   dW[tap][ci][co] += x[p + off(tap)][ci] dz[p][co],  db[co] += dz[p][co]
// XT / ZT: element types of x (CIN channels) and dz (COUT channels).  One
// block writes TAPS*CIN*COUT + COUT partials.
template <int CIN, int COUT, typename XT, typename ZT>
__global__ void __launch_bounds__(256) k_wgrad(const XT *__restrict__ x, const ZT *__restrict__ dz, int N, int H,
                                               int Wd, float *__restrict__ part) {
    constexpr int TRW = 4, TCW = 32;
    constexpr int NW = TAPS * CIN * COUT + COUT;
    __shared__ float sx[(TRW + 2) * (TCW + 2) * CIN];
    __shared__ float sz[TRW * TCW * COUT];
    const int tilesx = (Wd + TCW - 1) / TCW, tilesy = (H + TRW - 1) / TRW;
    const int bt = blockIdx.x;
    const int n = bt / (tilesx * tilesy), rem = bt % (tilesx * tilesy);
    const int r0 = (rem / tilesx) * TRW, c0 = (rem % tilesx) * TCW;
    for (int k = threadIdx.x; k < (TRW + 2) * (TCW + 2) * CIN; k += 256) {
        const int ci = k % CIN, pix = k / CIN, rr = pix / (TCW + 2), cc = pix % (TCW + 2);
        const int gr = r0 - 1 + rr, gc = c0 - 1 + cc;
        sx[k] = (gr >= 0 && gr < H && gc >= 0 && gc < Wd) ? (float)x[(((long long)n * H + gr) * Wd + gc) * CIN + ci] : 0.f;
    }
    for (int k = threadIdx.x; k < TRW * TCW * COUT; k += 256) {
        const int co = k % COUT, pix = k / COUT, rr = pix / TCW, cc = pix % TCW;
        const int gr = r0 + rr, gc = c0 + cc;
        sz[k] = (gr < H && gc < Wd) ? (float)dz[(((long long)n * H + gr) * Wd + gc) * COUT + co] : 0.f;
    }
    __syncthreads();
    for (int w = threadIdx.x; w < NW; w += 256) {
        float s = 0.f;
        if (w < TAPS * CIN * COUT) {
            const int co = w % COUT, ci = (w / COUT) % CIN, tap = w / (COUT * CIN);
            const int dy = tap_dy(tap), dx = tap_dx(tap);
            for (int rr = 0; rr < TRW; rr++)
                for (int cc = 0; cc < TCW; cc++)
                    s = fmaf(sx[((rr + 1 + dy) * (TCW + 2) + (cc + 1 + dx)) * CIN + ci], sz[(rr * TCW + cc) * COUT + co], s);
        } else {
            const int co = w - TAPS * CIN * COUT;
            for (int pix = 0; pix < TRW * TCW; pix++) s += sz[pix * COUT + co];
        }
        part[(long long)blockIdx.x * NW + w] = s;
    }
}

// grad[w] = sum over blocks of part[block][w] (fixed order)
__global__ void k_sum_partials(const float *__restrict__ part, int nblocks, int nw, float *__restrict__ grad) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; b++) s += part[(long long)b * nw + w];
    grad[w] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       long long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk, vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// BF16 copies of the mid-layer weights: forward [tap][co][ci] and
// backward-data [tap'][ci][co] = W[8 - tap'][ci][co]
__global__ void k_prep_weights(const float *__restrict__ W, __bf16 *__restrict__ wf, __bf16 *__restrict__ wb) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= TAPS * CH * CH) return;
    const int co = k % CH, ci = (k / CH) % CH, tap = k / (CH * CH);
    const __bf16 v = (__bf16)W[k];
    wf[(tap * CH + co) * CH + ci] = v;
    wb[((8 - tap) * CH + ci) * CH + co] = v;
}

// ---------------------------------------------------------------------------
// features from a context: log mobility, previous pressure, source
// ---------------------------------------------------------------------------
struct FeatArgs {
    const double *w;     // mobility quantity at (0,0), pitch ldw (mu/K harmonic, K/mu mean-conc)
    int ldw, harmonic;
    const double *p;     // pressure field whose features are formed, at (0,0), pitch ldp
    int ldp;
    const double *q;     // flat N1 x N2
    int N1, N2;
};

__device__ __forceinline__ double log_mob(const FeatArgs &f, int i, int j) {
    const double w = f.w[(long long)i * f.ldw + j];
    return f.harmonic ? -log(w) : log(w);
}

// partials: sum log(lambda), sum log(lambda)^2, sum p, sum p^2, max |q|
__global__ void k_feat_stats(FeatArgs f, double *__restrict__ part) {
    double s[4] = {0., 0., 0., 0.}, m[1] = {0.};
    const long long n = (long long)f.N1 * f.N2;
    for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(id / f.N2), j = (int)(id - (long long)i * f.N2);
        const double l = log_mob(f, i, j), p = f.p[(long long)i * f.ldp + j];
        s[0] += l; s[1] += l * l; s[2] += p; s[3] += p * p;
        m[0] = fmax(m[0], fabs(f.q[id]));
    }
    block_reduce<RED_SUM, 4>(s);
    block_reduce<RED_MAX, 1>(m);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) part[k * 1024 + blockIdx.x] = s[k];
    if (threadIdx.x == 0) part[4 * 1024 + blockIdx.x] = m[0];
}

// stat: [0] mean log(lambda), [1] 1/std log(lambda), [2] mean p, [3] std p, [4] 1/max|q|
__global__ void k_feat_finish(const double *__restrict__ part, int nb, long long n, double *__restrict__ stat) {
    double v[4];
    for (int k = 0; k < 4; k++) v[k] = strided_combine<RED_SUM>(part + k * 1024, nb);
    double m[1] = {strided_combine<RED_MAX>(part + 4 * 1024, nb)};
    block_reduce<RED_SUM, 4>(v);
    block_reduce<RED_MAX, 1>(m);
    if (threadIdx.x != 0) return;
    const double ml = v[0] / n, vl = fmax(v[1] / n - ml * ml, 0.), mp = v[2] / n, vp = fmax(v[3] / n - mp * mp, 0.);
    stat[0] = ml;
    stat[1] = vl > 1e-24 ? 1. / sqrt(vl) : 1.;
    stat[2] = mp;
    stat[3] = vp > 1e-24 ? sqrt(vp) : 1.;
    stat[4] = m[0] > 0 ? 1. / m[0] : 1.;
}

__device__ __forceinline__ void features_at(const FeatArgs &f, const double *stat, int i, int j, float out[3]) {
    out[0] = (float)((log_mob(f, i, j) - stat[0]) * stat[1]);
    out[1] = (float)((f.p[(long long)i * f.ldp + j] - stat[2]) / stat[3]);
    out[2] = (float)(f.q[(long long)i * f.N2 + j] * stat[4]);
}

// the whole field as one NHWC image (inference)
__global__ void k_features(FeatArgs f, const double *__restrict__ stat, float *__restrict__ x) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= f.N1 || j >= f.N2) return;
    float v[3];
    features_at(f, stat, i, j, v);
    float *o = x + ((long long)i * f.N2 + j) * 3;
    o[0] = v[0]; o[1] = v[1]; o[2] = v[2];
}

// front score for importance-sampled crops: |grad log(lambda)|^2 per cell
__global__ void k_front_score(FeatArgs f, float *__restrict__ score) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= f.N1 || j >= f.N2) return;
    const double l = log_mob(f, i, j);
    double s = 0.;
    if (i + 1 < f.N1) { const double d = log_mob(f, i + 1, j) - l; s += d * d; }
    if (j + 1 < f.N2) { const double d = log_mob(f, i, j + 1) - l; s += d * d; }
    score[(long long)i * f.N2 + j] = (float)s;
}

// training pairs: crops of (features of p_prev, (p_new - p_prev) / (std p_prev * tscale))
__global__ void k_collect(FeatArgs f, const double *__restrict__ stat, const double *__restrict__ pnew, int ldp,
                          const int *__restrict__ org, int ncrop, int crop, float tscale, float *__restrict__ pairs,
                          int first_slot, int capacity) {
    const int k = blockIdx.z;
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
    if (k >= ncrop || i >= crop || j >= crop) return;
    const int gi = org[2 * k] + i, gj = org[2 * k + 1] + j;
    float v[3];
    features_at(f, stat, gi, gj, v);
    const double dp = pnew[(long long)gi * ldp + gj] - f.p[(long long)gi * f.ldp + gj];
    const int slot = (first_slot + k) % capacity;
    float *o = pairs + (((long long)slot * crop + i) * crop + j) * 4;
    o[0] = v[0]; o[1] = v[1]; o[2] = v[2];
    o[3] = (float)(dp / (stat[3] * tscale));
}

// gather a batch of pairs into x0 (N x crop x crop x 3) and t (N x crop x crop)
__global__ void k_gather(const float *__restrict__ pairs, const int *__restrict__ idx, int nb, int crop,
                         float *__restrict__ x0, float *__restrict__ t) {
    const long long per = (long long)crop * crop;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb * per) return;
    const int b = (int)(k / per);
    const long long off = k - b * per;
    const float *src = pairs + ((long long)idx[b] * per + off) * 4;
    x0[k * 3 + 0] = src[0];
    x0[k * 3 + 1] = src[1];
    x0[k * 3 + 2] = src[2];
    t[k] = src[3];
}

// d = scale * y on the interior: the learned increment in pressure units
__global__ void k_increment(double *__restrict__ d, int ldd, const float *__restrict__ y, int N1, int N2,
                            const double *__restrict__ stat, double tscale) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= N1 || j >= N2) return;
    d[(long long)i * ldd + j] = (double)y[(long long)i * N2 + j] * stat[3] * tscale;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int cin_of(const impec_cnn *n, int l) { return l == 0 ? 3 : CH; }
static int cout_of(const impec_cnn *n, int l) { return l == n->L - 1 ? 1 : CH; }

// buffers for `pix` pixels: activations (all layers when training, two
// ping-pong buffers otherwise), inputs and outputs; when training also the
// backward buffers, the loss gradient and the targets
static int ensure_work(impec_cnn *n, long long pix, int batch, bool training) {
    const int nbuf = training ? n->L - 1 : 2;
    if (pix > n->act_pix || nbuf > n->act_nbuf) {
        const long long p2 = pix > n->act_pix ? pix : n->act_pix;
        const int b2 = nbuf > n->act_nbuf ? nbuf : n->act_nbuf;
        hipFree(n->act); hipFree(n->x0); hipFree(n->y);
        n->act = nullptr; n->x0 = n->y = nullptr;
        CNN_CHECK(n, hipMalloc(&n->act, sizeof(__bf16) * p2 * CH * b2));
        CNN_CHECK(n, hipMalloc(&n->x0, sizeof(float) * p2 * 3));
        CNN_CHECK(n, hipMalloc(&n->y, sizeof(float) * p2));
        n->act_pix = p2;
        n->act_nbuf = b2;
    }
    if (training && pix > n->train_pix) {
        hipFree(n->dz); hipFree(n->dy); hipFree(n->t);
        n->dz = nullptr; n->dy = n->t = nullptr;
        CNN_CHECK(n, hipMalloc(&n->dz, sizeof(__bf16) * pix * CH * 2));
        CNN_CHECK(n, hipMalloc(&n->dy, sizeof(float) * pix));
        CNN_CHECK(n, hipMalloc(&n->t, sizeof(float) * pix));
        n->train_pix = pix;
    }
    if (batch > n->idx_len) {
        hipFree(n->idx);
        CNN_CHECK(n, hipMalloc(&n->idx, sizeof(int) * batch));
        n->idx_len = batch;
    }
    return IMPEC_OK;
}

static long long wgrad_blocks(int N, int H, int Wd) {
    return (long long)N * ((H + 3) / 4) * ((Wd + 31) / 32);
}

static int ensure_part(impec_cnn *n, long long len) {
    if (len > n->part_len) {
        hipFree(n->part);
        n->part = nullptr;
        CNN_CHECK(n, hipMalloc(&n->part, sizeof(float) * len));
        n->part_len = len;
    }
    return IMPEC_OK;
}

static int refresh_bf16(impec_cnn *n) {
    for (int l = 1; l < n->L - 1; l++)
        hipLaunchKernelGGL(k_prep_weights, dim3((TAPS * CH * CH + 255) / 256), dim3(256), 0, n->stream,
                           n->param + n->woff[l], n->wf + (long long)(l - 1) * TAPS * CH * CH,
                           n->wb + (long long)(l - 1) * TAPS * CH * CH);
    CNN_CHECK(n, hipGetLastError());
    return IMPEC_OK;
}

static __bf16 *act_buf(impec_cnn *n, int l) { return n->act + (long long)l * n->act_pix * CH; }

// forward over x0 (N x H x W x 3) into y; activations kept in act[0 .. L-2]
// (keep_all = false: two ping-pong buffers are enough, for inference)
static int forward(impec_cnn *n, int N, int H, int Wd, bool keep_all, hipStream_t s) {
    const long long pix = (long long)N * H * Wd;
    const dim3 gp((unsigned)((pix + 255) / 256)), b256(256);
    hipLaunchKernelGGL(k_conv_first, gp, b256, 0, s, n->x0, n->param + n->woff[0], n->param + n->boff[0],
                       act_buf(n, 0), N, H, Wd);
    const dim3 gm((Wd + 31) / 32, (H + 4 * MID_TR - 1) / (4 * MID_TR), N);
    int cur = 0;
    for (int l = 1; l < n->L - 1; l++) {
        const int nxt = keep_all ? l : (cur ^ 1);
        hipLaunchKernelGGL(k_conv_mid<true>, gm, b256, 0, s, act_buf(n, cur),
                           n->wf + (long long)(l - 1) * TAPS * CH * CH, n->param + n->boff[l], (const __bf16 *)nullptr,
                           act_buf(n, nxt), N, H, Wd);
        cur = nxt;
    }
    hipLaunchKernelGGL(k_conv_last, gp, b256, 0, s, act_buf(n, cur), n->param + n->woff[n->L - 1],
                       n->param + n->boff[n->L - 1], n->y, N, H, Wd);
    CNN_CHECK(n, hipGetLastError());
    return IMPEC_OK;
}

// gradients of the mean squared error over (N, H, W) into n->grad; loss into *loss
static int backward(impec_cnn *n, int N, int H, int Wd, double *loss) {
    hipStream_t s = n->stream;
    const long long pix = (long long)N * H * Wd;
    const int L = n->L;
    const dim3 gp((unsigned)((pix + 255) / 256)), b256(256);
    int nb = (int)((pix + 255) / 256);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_loss, dim3(nb), b256, 0, s, n->y, n->t, n->dy, N, H, Wd, n->loss_kind, n->stat_part);
    const long long wb = wgrad_blocks(N, H, Wd);
    int st = ensure_part(n, wb * (TAPS * CH * CH + CH));
    if (st) return st;
    auto sum_into = [&](int l) {
        const int nw = TAPS * cin_of(n, l) * cout_of(n, l) + cout_of(n, l);
        hipLaunchKernelGGL(k_sum_partials, dim3((nw + 255) / 256), b256, 0, s, n->part, (int)wb, nw,
                           n->grad + n->woff[l]);
    };
    // output layer
    __bf16 *a_last = act_buf(n, L - 2);
    hipLaunchKernelGGL((k_wgrad<CH, 1, __bf16, float>), dim3((unsigned)wb), b256, 0, s, a_last, n->dy, N, H, Wd,
                       n->part);
    sum_into(L - 1);
    __bf16 *dz_cur = n->dz, *dz_nxt = n->dz + pix * CH;
    hipLaunchKernelGGL(k_last_back, gp, b256, 0, s, n->dy, n->param + n->woff[L - 1], a_last, dz_cur, N, H, Wd);
    const dim3 gm((Wd + 31) / 32, (H + 4 * MID_TR - 1) / (4 * MID_TR), N);
    for (int l = L - 2; l >= 1; l--) {
        hipLaunchKernelGGL((k_wgrad<CH, CH, __bf16, __bf16>), dim3((unsigned)wb), b256, 0, s, act_buf(n, l - 1), dz_cur,
                           N, H, Wd, n->part);
        sum_into(l);
        hipLaunchKernelGGL(k_conv_mid<false>, gm, b256, 0, s, dz_cur, n->wb + (long long)(l - 1) * TAPS * CH * CH,
                           (const float *)nullptr, act_buf(n, l - 1), dz_nxt, N, H, Wd);
        __bf16 *tmp = dz_cur;
        dz_cur = dz_nxt;
        dz_nxt = tmp;
    }
    hipLaunchKernelGGL((k_wgrad<3, CH, float, __bf16>), dim3((unsigned)wb), b256, 0, s, n->x0, dz_cur, N, H, Wd,
                       n->part);
    sum_into(0);
    CNN_CHECK(n, hipGetLastError());
    if (loss) {
        double *h = (double *)malloc(sizeof(double) * nb);
        CNN_CHECK(n, hipMemcpyAsync(h, n->stat_part, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
        CNN_CHECK(n, hipStreamSynchronize(s));
        double t = 0.;
        for (int b = 0; b < nb; b++) t += h[b];
        free(h);
        *loss = t / (double)pix;
    }
    return IMPEC_OK;
}

static FeatArgs feat_args(impec_ctx *ctx, const double *p) {
    FeatArgs f;
    f.w = ctx->w + (long long)ctx->B * ctx->lds + ctx->B;
    f.ldw = ctx->lds;
    f.harmonic = ctx->P.average_procedure == IMPEC_AVG_HARMONIC;
    f.p = p + (long long)ctx->B * ctx->lds + ctx->B;
    f.ldp = ctx->lds;
    f.q = ctx->q;
    f.N1 = ctx->N1;
    f.N2 = ctx->N2;
    return f;
}

static int field_stats(impec_cnn *n, const FeatArgs &f, hipStream_t s) {
    long long cells = (long long)f.N1 * f.N2;
    int nb = (int)((cells + 255) / 256);
    if (nb > 1024) nb = 1024;
    hipLaunchKernelGGL(k_feat_stats, dim3(nb), dim3(256), 0, s, f, n->stat_part);
    hipLaunchKernelGGL(k_feat_finish, dim3(1), dim3(1024), 0, s, n->stat_part, nb, cells, n->stat);
    CNN_CHECK(n, hipGetLastError());
    return IMPEC_OK;
}

// the network's pressure increment for the current mobility and p, into d
// (cell (0,0) at d, pitch ldd)
int cnn_increment(impec_ctx *ctx, double *d, int ldd) {
    impec_cnn *n = ctx->net;
    if (!n) return IMPEC_OK;
    const int N1 = ctx->N1, N2 = ctx->N2;
    int st = ensure_work(n, (long long)N1 * N2, 1, false);
    if (st) return set_error(ctx, st, "learned guess: %s", n->err);
    const FeatArgs f = feat_args(ctx, ctx->p);
    if ((st = field_stats(n, f, ctx->stream))) return set_error(ctx, st, "learned guess: %s", n->err);
    dim3 blk(64, 4), grd((N2 + 63) / 64, (N1 + 3) / 4);
    hipLaunchKernelGGL(k_features, grd, blk, 0, ctx->stream, f, n->stat, n->x0);
    if ((st = forward(n, 1, N1, N2, false, ctx->stream))) return set_error(ctx, st, "learned guess: %s", n->err);
    hipLaunchKernelGGL(k_increment, grd, blk, 0, ctx->stream, d, ldd, n->y, N1, N2, n->stat, (double)n->tscale);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

}  // namespace impec

using namespace impec;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int impec_cnn_config_default(impec_cnn_config *c) {
    if (!c) return IMPEC_EINVAL;
    memset(c, 0, sizeof(*c));
    c->layers = 8;
    c->channels = CH;
    c->capacity = 4096;
    c->crop = 64;
    c->seed = 1234;
    return IMPEC_OK;
}

int impec_cnn_create(const impec_cnn_config *c, impec_cnn **out) {
    if (!c || !out || c->layers < 3 || c->channels != CH || c->crop < 32 || c->crop % 32 || c->capacity < 1)
        return IMPEC_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return IMPEC_ENODEV;
    impec_cnn *n = new impec_cnn();
    n->L = c->layers;
    n->crop = c->crop;
    n->capacity = c->capacity;
    n->seed = c->seed;
    n->loss_kind = c->mse_loss == 1 ? 0 : 1;
    long long off = 0;
    for (int l = 0; l < n->L; l++) {
        n->woff.push_back(off);
        off += (long long)TAPS * cin_of(n, l) * cout_of(n, l);
        n->boff.push_back(off);
        off += cout_of(n, l);
    }
    n->nparam = off;
    // He initialisation (normal, std sqrt(2 / fan_in)); the output layer starts
    // near zero, so an untrained network leaves the previous pressure unchanged
    std::vector<float> h(off, 0.f);
    std::mt19937_64 rng(n->seed);
    for (int l = 0; l < n->L; l++) {
        const double fan = TAPS * cin_of(n, l);
        std::normal_distribution<double> nd(0., sqrt(2. / fan) * (l == n->L - 1 ? 0.01 : 1.));
        for (long long k = n->woff[l]; k < n->boff[l]; k++) h[k] = (float)nd(rng);
    }
    bool ok = hipStreamCreate(&n->stream) == hipSuccess;
    ok = ok && hipMalloc(&n->param, sizeof(float) * off) == hipSuccess;
    ok = ok && hipMalloc(&n->grad, sizeof(float) * off) == hipSuccess;
    ok = ok && hipMalloc(&n->adam_m, sizeof(float) * off) == hipSuccess;
    ok = ok && hipMalloc(&n->adam_v, sizeof(float) * off) == hipSuccess;
    const long long nw = (long long)(n->L - 2) * TAPS * CH * CH;
    ok = ok && hipMalloc(&n->wf, sizeof(__bf16) * (nw > 0 ? nw : 1)) == hipSuccess;
    ok = ok && hipMalloc(&n->wb, sizeof(__bf16) * (nw > 0 ? nw : 1)) == hipSuccess;
    ok = ok && hipMalloc(&n->pairs, sizeof(float) * (long long)n->capacity * n->crop * n->crop * 4) == hipSuccess;
    ok = ok && hipMalloc(&n->stat, sizeof(double) * 8) == hipSuccess;
    ok = ok && hipMalloc(&n->stat_part, sizeof(double) * 5 * 1024) == hipSuccess;
    ok = ok && hipMemcpy(n->param, h.data(), sizeof(float) * off, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemset(n->adam_m, 0, sizeof(float) * off) == hipSuccess;
    ok = ok && hipMemset(n->adam_v, 0, sizeof(float) * off) == hipSuccess;
    if (!ok || refresh_bf16(n) != IMPEC_OK || hipStreamSynchronize(n->stream) != hipSuccess) {
        impec_cnn_destroy(n);
        return IMPEC_ENOMEM;
    }
    *out = n;
    return IMPEC_OK;
}

int impec_cnn_destroy(impec_cnn *n) {
    if (!n) return IMPEC_EINVAL;
    if (n->stream) hipStreamSynchronize(n->stream);
    hipFree(n->param); hipFree(n->grad); hipFree(n->adam_m); hipFree(n->adam_v);
    hipFree(n->wf); hipFree(n->wb); hipFree(n->pairs);
    hipFree(n->act); hipFree(n->dz); hipFree(n->x0); hipFree(n->y); hipFree(n->dy); hipFree(n->t);
    hipFree(n->part); hipFree(n->stat); hipFree(n->stat_part); hipFree(n->idx);
    if (n->stream) hipStreamDestroy(n->stream);
    delete n;
    return IMPEC_OK;
}

const char *impec_cnn_last_error(const impec_cnn *n) { return n ? n->err : "null network"; }

int impec_cnn_num_params(const impec_cnn *n, long long *count) {
    if (!n || !count) return IMPEC_EINVAL;
    *count = n->nparam;
    return IMPEC_OK;
}

int impec_cnn_get_params(impec_cnn *n, float *h) {
    if (!n || !h) return IMPEC_EINVAL;
    CNN_CHECK(n, hipMemcpyAsync(h, n->param, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->stream));
    CNN_CHECK(n, hipStreamSynchronize(n->stream));
    return IMPEC_OK;
}

int impec_cnn_set_params(impec_cnn *n, const float *h) {
    if (!n || !h) return IMPEC_EINVAL;
    CNN_CHECK(n, hipMemcpyAsync(n->param, h, sizeof(float) * n->nparam, hipMemcpyHostToDevice, n->stream));
    int st = refresh_bf16(n);
    if (st) return st;
    CNN_CHECK(n, hipStreamSynchronize(n->stream));
    return IMPEC_OK;
}

int impec_cnn_get_scale(impec_cnn *n, float *tscale) {
    if (!n || !tscale) return IMPEC_EINVAL;
    *tscale = n->tscale;
    return IMPEC_OK;
}

int impec_cnn_set_scale(impec_cnn *n, float tscale) {
    if (!n || !(tscale > 0.f)) return IMPEC_EINVAL;
    n->tscale = tscale;
    return IMPEC_OK;
}

int impec_cnn_forward(impec_cnn *n, const float *x, int N, int H, int Wd, float *y) {
    if (!n || !x || !y || N < 1 || H < 1 || Wd < 1) return IMPEC_EINVAL;
    const long long pix = (long long)N * H * Wd;
    int st = ensure_work(n, pix, 1, false);
    if (st) return st;
    CNN_CHECK(n, hipMemcpyAsync(n->x0, x, sizeof(float) * pix * 3, hipMemcpyHostToDevice, n->stream));
    if ((st = forward(n, N, H, Wd, false, n->stream))) return st;
    CNN_CHECK(n, hipMemcpyAsync(y, n->y, sizeof(float) * pix, hipMemcpyDeviceToHost, n->stream));
    CNN_CHECK(n, hipStreamSynchronize(n->stream));
    return IMPEC_OK;
}

int impec_cnn_gradients(impec_cnn *n, const float *x, const float *t, int N, int H, int Wd, float *grads,
                        double *loss) {
    if (!n || !x || !t || !grads || N < 1 || H < 1 || Wd < 1) return IMPEC_EINVAL;
    const long long pix = (long long)N * H * Wd;
    int st = ensure_work(n, pix, 1, true);
    if (st) return st;
    CNN_CHECK(n, hipMemcpyAsync(n->x0, x, sizeof(float) * pix * 3, hipMemcpyHostToDevice, n->stream));
    CNN_CHECK(n, hipMemcpyAsync(n->t, t, sizeof(float) * pix, hipMemcpyHostToDevice, n->stream));
    if ((st = forward(n, N, H, Wd, true, n->stream))) return st;
    if ((st = backward(n, N, H, Wd, loss))) return st;
    CNN_CHECK(n, hipMemcpyAsync(grads, n->grad, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->stream));
    CNN_CHECK(n, hipStreamSynchronize(n->stream));
    return IMPEC_OK;
}

int impec_cnn_collect(impec_cnn *n, impec_ctx *ctx, int crops) {
    if (!n || !ctx || crops < 1) return IMPEC_EINVAL;
    if (!ctx->p_prev || !ctx->have_prev_p)
        return cnn_err(n, IMPEC_EINVAL, "collect: the context keeps no previous pressure (attach the network first)");
    if (ctx->N1 < n->crop || ctx->N2 < n->crop) return cnn_err(n, IMPEC_EINVAL, "collect: grid smaller than a crop");
    hipStream_t s = ctx->stream;
    const FeatArgs f = feat_args(ctx, ctx->p_prev);
    int st = field_stats(n, f, s);
    if (st) return st;
    if (n->tscale <= 0.f) {
        // target scale from this field: rms(p_new - p_prev) / std(p_prev) on the host
        const size_t cells = (size_t)(ctx->N1 + 2 * ctx->B) * ctx->lds;
        std::vector<double> pn(cells), pp(cells);
        double stat[8];
        CNN_CHECK(n, hipMemcpyAsync(pn.data(), ctx->p, sizeof(double) * cells, hipMemcpyDeviceToHost, s));
        CNN_CHECK(n, hipMemcpyAsync(pp.data(), ctx->p_prev, sizeof(double) * cells, hipMemcpyDeviceToHost, s));
        CNN_CHECK(n, hipMemcpyAsync(stat, n->stat, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
        CNN_CHECK(n, hipStreamSynchronize(s));
        double ss = 0.;
        for (int i = 0; i < ctx->N1; i++)
            for (int j = 0; j < ctx->N2; j++) {
                const size_t o = (size_t)(i + ctx->B) * ctx->lds + (j + ctx->B);
                ss += (pn[o] - pp[o]) * (pn[o] - pp[o]);
            }
        const double rms = sqrt(ss / ((double)ctx->N1 * ctx->N2));
        n->tscale = rms > 0 ? (float)(rms / stat[3]) : 1.f;
    }
    // crop origins: half uniform, half centred on cells drawn with probability
    // proportional to |grad log(lambda)|^2 (the displacement front, where the
    // pressure increment concentrates)
    const long long cells = (long long)ctx->N1 * ctx->N2;
    std::vector<float> score(cells);
    {
        float *dscore = nullptr;
        CNN_CHECK(n, hipMalloc(&dscore, sizeof(float) * cells));
        dim3 b2(64, 4), g2((ctx->N2 + 63) / 64, (ctx->N1 + 3) / 4);
        hipLaunchKernelGGL(k_front_score, g2, b2, 0, s, f, dscore);
        hipError_t e = hipMemcpyAsync(score.data(), dscore, sizeof(float) * cells, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        hipFree(dscore);
        CNN_CHECK(n, e);
    }
    std::vector<double> cdf(cells);
    double acc = 0.;
    for (long long k = 0; k < cells; k++) cdf[k] = (acc += score[k]);
    std::vector<int> org(2 * crops);
    std::mt19937_64 rng(n->seed + 7919ull * (unsigned long long)(n->next + 1));
    std::uniform_real_distribution<double> uni(0., 1.);
    for (int k = 0; k < crops; k++) {
        int oi, oj;
        if ((k & 1) && acc > 0.) {
            const long long cell = std::lower_bound(cdf.begin(), cdf.end(), uni(rng) * acc) - cdf.begin();
            const int ci = (int)(cell / ctx->N2), cj = (int)(cell % ctx->N2);
            oi = ci - n->crop / 2;
            oj = cj - n->crop / 2;
            oi = oi < 0 ? 0 : (oi > ctx->N1 - n->crop ? ctx->N1 - n->crop : oi);
            oj = oj < 0 ? 0 : (oj > ctx->N2 - n->crop ? ctx->N2 - n->crop : oj);
        } else {
            oi = (int)(rng() % (unsigned long long)(ctx->N1 - n->crop + 1));
            oj = (int)(rng() % (unsigned long long)(ctx->N2 - n->crop + 1));
        }
        org[2 * k] = oi;
        org[2 * k + 1] = oj;
    }
    int *dorg = nullptr;
    CNN_CHECK(n, hipMalloc(&dorg, sizeof(int) * 2 * crops));
    CNN_CHECK(n, hipMemcpyAsync(dorg, org.data(), sizeof(int) * 2 * crops, hipMemcpyHostToDevice, s));
    dim3 blk(32, 8), grd(n->crop / 32, n->crop / 8, crops);
    hipLaunchKernelGGL(k_collect, grd, blk, 0, s, f, n->stat, ctx->p + (long long)ctx->B * ctx->lds + ctx->B,
                       ctx->lds, dorg, crops, n->crop, n->tscale, n->pairs, n->next, n->capacity);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(dorg);
    CNN_CHECK(n, e);
    n->next = (n->next + crops) % n->capacity;
    n->count = n->count + crops < n->capacity ? n->count + crops : n->capacity;
    return IMPEC_OK;
}

int impec_cnn_pairs(const impec_cnn *n, int *count) {
    if (!n || !count) return IMPEC_EINVAL;
    *count = n->count;
    return IMPEC_OK;
}

int impec_cnn_train(impec_cnn *n, int iters, int batch, float lr, double *loss) {
    if (!n || iters < 0 || batch < 1 || !(lr > 0.f)) return IMPEC_EINVAL;
    if (n->count < 1) return cnn_err(n, IMPEC_EINVAL, "train: no training pairs collected");
    const int crop = n->crop;
    int st = ensure_work(n, (long long)batch * crop * crop, batch, true);
    if (st) return st;
    std::mt19937_64 rng(n->seed ^ (0x9e3779b97f4a7c15ull * (unsigned long long)(n->adam_t + 1)));
    std::vector<int> idx(batch);
    double last = 0.;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int it = 0; it < iters; it++) {
        for (int b = 0; b < batch; b++) idx[b] = (int)(rng() % (unsigned long long)n->count);
        CNN_CHECK(n, hipMemcpyAsync(n->idx, idx.data(), sizeof(int) * batch, hipMemcpyHostToDevice, n->stream));
        const long long pix = (long long)batch * crop * crop;
        hipLaunchKernelGGL(k_gather, dim3((unsigned)((pix + 255) / 256)), dim3(256), 0, n->stream, n->pairs, n->idx,
                           batch, crop, n->x0, n->t);
        if ((st = forward(n, batch, crop, crop, true, n->stream))) return st;
        if ((st = backward(n, batch, crop, crop, it == iters - 1 ? &last : nullptr))) return st;
        n->adam_t++;
        const float c1 = 1.f - powf(b1, (float)n->adam_t), c2 = 1.f - powf(b2, (float)n->adam_t);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((n->nparam + 255) / 256)), dim3(256), 0, n->stream, n->param,
                           n->grad, n->adam_m, n->adam_v, n->nparam, lr, b1, b2, eps, c1, c2);
        if ((st = refresh_bf16(n))) return st;
    }
    CNN_CHECK(n, hipStreamSynchronize(n->stream));
    if (loss) *loss = last;
    return IMPEC_OK;
}

int impec_set_warmstart(impec_ctx *ctx, impec_cnn *n) {
    if (!ctx) return IMPEC_EINVAL;
    ctx->net = n;
    const size_t bytes = sizeof(double) * (size_t)(ctx->N1 + 2 * ctx->B) * ctx->lds;
    if (n && !ctx->p_prev) {
        IMPEC_HIP_CHECK(ctx, hipMalloc(&ctx->p_prev, bytes));
        IMPEC_HIP_CHECK(ctx, hipMemset(ctx->p_prev, 0, bytes));
    }
    // the learned guess also steps along the last step's change p_n - p_(n-1)
    if (n && ctx->P.initial_guess == IMPEC_GUESS_LEARNED && !ctx->p_old) {
        IMPEC_HIP_CHECK(ctx, hipMalloc(&ctx->p_old, bytes));
        IMPEC_HIP_CHECK(ctx, hipMemset(ctx->p_old, 0, bytes));
    }
    return IMPEC_OK;
}

}  // extern "C"
