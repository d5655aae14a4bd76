// tp_mlp.hip -- the learned part: an MLP surrogate for log|B_kappa(l1, l2, l3; cosmology)|.
//
// Network: 8 inputs (the triangle's three sides as sorted logs -- equivalent to (ln l1, ln l2,
// cos phi) up to B_kappa's permutation symmetry, and resolving squeezed shapes -- and Omega_m,
// sigma8, h, w, n_s; each mapped to [-1, 1] over the training box) -> 4 hidden layers of 256 SiLU units -> 1 output (the
// standardised log B_kappa).  199 937 parameters, FP32 master copies, BF16 copies of the
// three 256x256 weight matrices for the matrix cores.
//
//   layer 1 (8 -> 256) and the output layer (256 -> 1): FP32 VALU kernels;
//   layers 2-4 (256 -> 256): v_mfma_f32_32x32x8_bf16 GEMMs with FP32 accumulation, BF16
//   activations, bias + SiLU fused into the epilogue (forward) and SiLU' fused into the
//   input-gradient epilogue (backward); weight gradients by split-K MFMA GEMMs over the batch.
//   Adam (beta1 0.9, beta2 0.999, eps 1e-8) with cosine learning-rate decay.
//
// Training data are generated on the GPU by the exact FP64 B_kappa kernel over random
// (l1, l2, phi) and random cosmologies of the training box, each with its own GPU-built tables.
//
// The network learns ln B_kappa - ln B_8, where B_8 is the same Limber integral on 8 Gauss-Legendre
// z-nodes in FP32 (about 8% of the FP32 exact cost): a small, smooth residual, fitted far more
// precisely than ln B_kappa itself (which spans ~40 e-folds over the box); B_surr = B_8 exp(MLP).
//
// Deployment (TP_MODE_SURROGATE in tp_map3_batch): every cubature point of a round goes
// through the MLP; points outside the training box and a seeded random fraction (check_frac)
// of the rest are evaluated exactly.  The exact points give, per region, an unbiased
// correction  vol * sum_p w_p (f_exact - f_surr) / check_frac  (w_p the degree-7 rule weight),
// kept only on the final (leaf) regions.  A round whose checked points show a weighted
// relative surrogate error above tol is re-evaluated exactly (fallback).
#include <hip/hip_bf16.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "tp_device.h"

namespace {

constexpr int H = 256;   // hidden width
constexpr int NIN = 8;   // inputs
constexpr int NLAYER_MFMA = 3;
constexpr int NFRAG = 3 * 16 * 8 * 64;  // 16-byte fragments of the hidden weights in k_mlp_reg's order

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__device__ __forceinline__ unsigned short f2bf(float f)
{
    // round-to-nearest-even FP32 -> BF16
    unsigned int u = __float_as_uint(f);
    unsigned int lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
// inference form: hardware reciprocal (1 ulp) instead of the IEEE division sequence
__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z)
{
    float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// parameter layout (FP32): W1[256][8] b1[256] W2[256][256] b2 W3 b3 W4 b4 W5[256] b5[1]
constexpr size_t OFF_W1 = 0, OFF_B1 = OFF_W1 + H * NIN;
constexpr size_t OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H;
constexpr size_t OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * H;
constexpr size_t OFF_W4 = OFF_B3 + H, OFF_B4 = OFF_W4 + H * H;
constexpr size_t OFF_W5 = OFF_B4 + H, OFF_B5 = OFF_W5 + H;
constexpr size_t NPARAM = OFF_B5 + 1;
static_assert(NPARAM == 199937, "parameter count");
constexpr size_t OFF_W[3] = {OFF_W2, OFF_W3, OFF_W4};
constexpr size_t OFF_B[3] = {OFF_B2, OFF_B3, OFF_B4};

// --------------------------------------------------------------- MFMA GEMM
// C[M][N] (+)= sum_k A(m,k) B(k,n), A and B BF16 addressed through strides, FP32 accumulation.
// Block 256 threads = 4 waves; block tile 64x64; wave tile 32x32 (one 32x32x8 MFMA accumulator);
// K staged through LDS in steps of 32 ([m][k] and [n][k] order, k contiguous for the MFMA).
// Requires M % 64 == 0, N % 64 == 0, K % 32 == 0 (checked on the host).
constexpr int BM = 64, BN = 64, BK = 32, KPAD = 8;

struct GemmArgs {
    const unsigned short *A;
    long long sam, sak;
    const unsigned short *B;
    long long sbk, sbn;
    int M, N, K;
    int kchunk;       // K range per blockIdx.z (split-K)
    int epi;          // 0: FP32 partial C[z][M][N]; 1: forward bias+SiLU (+Z); 2: backward SiLU'; 3: inference
    float *C;         // epi 0: partial outputs
    const float *bias;
    float *Z;         // epi 1: pre-activation out (FP32, [M][N]); epi 2: pre-activation in
    unsigned short *Aout;  // epi 1: BF16 activation; epi 2: BF16 gradient w.r.t. pre-activation
    int ldo;
};

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g)
{
    __shared__ unsigned short sA[BM][BK + KPAD];
    __shared__ unsigned short sB[BN][BK + KPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int kbeg = blockIdx.z * g.kchunk, kend = min(g.K, kbeg + g.kchunk);
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x16 acc;
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    const bool a_kfast = (g.sak == 1), b_kfast = (g.sbk == 1);
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        // global -> LDS, consecutive threads on the contiguous dimension
#pragma unroll
        for (int j = 0; j < (BM * BK) / 256; ++j) {
            int e = tid + 256 * j;
            int mm, kk;
            if (a_kfast) { mm = e / BK; kk = e % BK; } else { kk = e / BM; mm = e % BM; }
            sA[mm][kk] = g.A[(long long)(m0 + mm) * g.sam + (long long)(k0 + kk) * g.sak];
        }
#pragma unroll
        for (int j = 0; j < (BN * BK) / 256; ++j) {
            int e = tid + 256 * j;
            int nn, kk;
            if (b_kfast) { nn = e / BK; kk = e % BK; } else { kk = e / BN; nn = e % BN; }
            sB[nn][kk] = g.B[(long long)(k0 + kk) * g.sbk + (long long)(n0 + nn) * g.sbn];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 8) {
            // A fragment: row wm + lane%32, k = kk + 4*(lane/32) .. +3 ; B fragment: col wn + lane%32
            const int r = lane & 31, kq = kk + 4 * (lane >> 5);
            bf16x4 a, b;
            for (int i = 0; i < 4; ++i) {
                a[i] = (short)sA[wm + r][kq + i];
                b[i] = (short)sB[wn + r][kq + i];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // accumulator element acc[4j+i]: row 8j + 4*(lane/32) + i, column lane % 32
    const int col = n0 + wn + (lane & 31);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = m0 + wm + 8 * j + 4 * (lane >> 5) + i;
            float v = acc[4 * j + i];
            if (g.epi == 0) {
                g.C[((long long)blockIdx.z * g.M + row) * g.N + col] = v;
            } else if (g.epi == 1) {
                float z = v + g.bias[col];
                g.Z[(long long)row * g.ldo + col] = z;
                g.Aout[(long long)row * g.ldo + col] = f2bf(silu(z));
            } else if (g.epi == 2) {
                float z = g.Z[(long long)row * g.ldo + col];
                g.Aout[(long long)row * g.ldo + col] = f2bf(v * silu_grad(z));
            } else {
                g.Aout[(long long)row * g.ldo + col] = f2bf(silu(v + g.bias[col]));
            }
        }
}

// sum of split-K partials: out[i] = sum_z part[z][i]
__global__ void k_reduce_parts(const float *part, int nz, size_t n, float *out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int z = 0; z < nz; ++z) s += part[(size_t)z * n + i];
    out[i] = s;
}

// column sums of a BF16 [M][N] matrix (bias gradients): out[n] = sum_m X[m][n]
__global__ void k_colsum_bf16(const unsigned short *X, int M, int N, float *out)
{
    int n = blockIdx.x;
    float s = 0.f;
    for (int m = threadIdx.x; m < M; m += blockDim.x) s += bf2f(X[(size_t)m * N + n]);
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[n] = red[0] + red[1] + red[2] + red[3];
}

// layer 1 forward: Z1 = X W1^T + b1 (FP32), A1 = bf16(silu(Z1)); one thread per (row, unit)
__global__ void k_layer1_fwd(const float *X, int M, const float *P, float *Z1, unsigned short *A1)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)M * H) return;
    int m = (int)(t / H), n = (int)(t % H);
    const float *w = P + OFF_W1 + n * NIN;
    const float *x = X + (size_t)m * NIN;
    float z = P[OFF_B1 + n];
#pragma unroll
    for (int k = 0; k < NIN; ++k) z += w[k] * x[k];
    Z1[t] = z;
    A1[t] = f2bf(silu(z));
}

// output layer forward: y = A4 w5 + b5 (one wave per row); optional loss gradient and dZ4
__global__ void __launch_bounds__(256) k_out_fwd(const unsigned short *A4, const float *Z4, int M, const float *P,
                                                 float *y, const float *target, float inv_batch, float *dy,
                                                 unsigned short *dZ4, double *loss)
{
    int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wave >= M) return;
    const unsigned short *a = A4 + (size_t)wave * H;
    float s = 0.f;
    for (int k = lane; k < H; k += 64) s += bf2f(a[k]) * P[OFF_W5 + k];
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    float yy = s + P[OFF_B5];
    if (lane == 0 && y) y[wave] = yy;
    if (target) {
        float d = (yy - target[wave]) * inv_batch;  // d/dy of mean (y-t)^2/2
        if (lane == 0) {
            dy[wave] = d;
            atomicAdd(loss, 0.5 * (double)(yy - target[wave]) * (double)(yy - target[wave]) * inv_batch);
        }
        for (int k = lane; k < H; k += 64) {
            float z = Z4[(size_t)wave * H + k];
            dZ4[(size_t)wave * H + k] = f2bf(d * P[OFF_W5 + k] * silu_grad(z));
        }
    }
}

// output-layer gradients: gW5[k] = sum_m dy[m] A4[m][k], gb5 = sum dy
__global__ void k_out_grad(const unsigned short *A4, const float *dy, int M, float *G)
{
    int k = blockIdx.x;  // 0..H (H -> bias)
    float s = 0.f;
    for (int m = threadIdx.x; m < M; m += blockDim.x) s += dy[m] * (k < H ? bf2f(A4[(size_t)m * H + k]) : 1.0f);
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) G[k < H ? OFF_W5 + k : OFF_B5] = red[0] + red[1] + red[2] + red[3];
}

// layer-1 gradients: gW1[n][k] = sum_m dZ1[m][n] X[m][k], gb1[n] = sum_m dZ1[m][n]
__global__ void k_layer1_grad(const unsigned short *dZ1, const float *X, int M, float *G)
{
    int n = blockIdx.x;
    float s[NIN + 1];
    for (int k = 0; k <= NIN; ++k) s[k] = 0.f;
    for (int m = threadIdx.x; m < M; m += blockDim.x) {
        float d = bf2f(dZ1[(size_t)m * H + n]);
        for (int k = 0; k < NIN; ++k) s[k] += d * X[(size_t)m * NIN + k];
        s[NIN] += d;
    }
    __shared__ float red[NIN + 1][4];
    for (int k = 0; k <= NIN; ++k) {
        float v = s[k];
        for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x <= NIN) {
        float v = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
        if (threadIdx.x < NIN) G[OFF_W1 + n * NIN + threadIdx.x] = v;
        else G[OFF_B1 + n] = v;
    }
}

__global__ void k_adam(float *P, const float *G, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = G[i];
    float mi = b1 * m[i] + (1 - b1) * g;
    float vi = b2 * v[i] + (1 - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// FP32 -> BF16 copies of W2..W4 ([out][in] row-major)
__global__ void k_weights_bf16(const float *P, unsigned short *Wb)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)NLAYER_MFMA * H * H) return;
    int l = (int)(t / (H * H));
    size_t e = t % (H * H);
    const size_t off[3] = {OFF_W2, OFF_W3, OFF_W4};
    Wb[t] = f2bf(P[off[l] + e]);
}

// --------------------------------------------------------------- data generation
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;  // splitmix64 finaliser
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(unsigned long long key, unsigned long long ctr)
{
    return (double)(mix64(key + 0x9e3779b97f4a7c15ULL * (ctr + 1)) >> 11) * (1.0 / 9007199254740992.0);
}

struct SurBox {
    double lnl_min, lnl_max;
    double lo[5], hi[5];  // om, s8, h, w, ns
};

// Triangle encoding: B_kappa(l1, l2, l3) is symmetric under any permutation of the sides
// (classed_version/bispectrum.cpp:166-236 is), so the three sides enter sorted, as logs, each
// mapped from [lnl_min, lnl_max] to [-1, 1]; the cosmology parameters map their box to [-1, 1].
__device__ __forceinline__ void make_input(const SurBox &B, double l1, double l2, double l3, const double cp[5],
                                           float *x)
{
    double a = l1, b = l2, c = l3, t;
    if (a < b) { t = a; a = b; b = t; }
    if (b < c) { t = b; b = c; c = t; }
    if (a < b) { t = a; a = b; b = t; }
    double mid = 0.5 * (B.lnl_min + B.lnl_max), half = 0.5 * (B.lnl_max - B.lnl_min);
    x[0] = (float)((log(a) - mid) / half);
    x[1] = (float)((log(b) - mid) / half);
    x[2] = (float)((log(c) - mid) / half);
    for (int k = 0; k < 5; ++k) x[3 + k] = (float)(2.0 * (cp[k] - B.lo[k]) / (B.hi[k] - B.lo[k]) - 1.0);
}

// one thread per sample: random cosmology of the set, a random triangle with all sides in the box
// (half the samples with the opening angle uniform, half with ln l3 uniform, so squeezed shapes
// are covered), exact FP64 B_kappa and the FP32 8-node baseline B_8; target ln B - ln B_8
__global__ void __launch_bounds__(256) k_gen(const TpNode64 *__restrict__ nodes, const TpEH64 *__restrict__ eh,
                                             const TpNode32 *__restrict__ base, const TpEH32 *__restrict__ eh32,
                                             const TpCosmoConst *__restrict__ cc, int ncosmo, SurBox box,
                                             unsigned long long key, size_t n, float *X, float *T, int *nonpos)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long c0 = (unsigned long long)i * 8;
    int ic = min(ncosmo - 1, (int)(u01(key, c0) * ncosmo));
    const double lmin = exp(box.lnl_min), lmax = exp(box.lnl_max);
    double l1 = lmin, l2 = lmin, l3 = lmin;
    bool ok = false;
    for (int tries = 0; tries < 16 && !ok; ++tries) {  // bounded rejection into the box
        unsigned long long c1 = c0 + 4 * (unsigned long long)tries;
        l1 = exp(box.lnl_min + (box.lnl_max - box.lnl_min) * u01(key ^ 0x1111, c1));
        l2 = exp(box.lnl_min + (box.lnl_max - box.lnl_min) * u01(key ^ 0x2222, c1 + 1));
        double u3 = u01(key ^ 0x3333, c1 + 2);
        if (u01(key ^ 0x4444, c1 + 3) < 0.5) {  // opening angle uniform
            l3 = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * cos(tpd::kPi * u3));
        } else {                                   // ln l3 uniform over the triangle's range
            double lo = fmax(fabs(l1 - l2), lmin), hi = fmin(l1 + l2, lmax);
            if (!(hi > lo)) continue;
            l3 = exp(log(lo) + (log(hi) - log(lo)) * u3);
        }
        ok = (l3 >= lmin && l3 <= lmax);
    }
    if (!ok) l1 = l2 = l3 = exp(0.5 * (box.lnl_min + box.lnl_max));
    double b = tp_bkappa_sum<double>(nodes + (size_t)ic * TP_NNODES, eh[ic], l1, l2, l3);
    const TpCosmoConst &C = cc[ic];
    double cp[5] = {C.om, C.sigma8, C.h, C.w, C.ns};
    make_input(box, l1, l2, l3, cp, X + i * NIN);
    double bb = tp_bkappa_sum_n<float, TP_NBASE>(base + (size_t)ic * TP_NBASE, eh32[ic], l1, l2, l3);
    if (!(b > 0) || !(bb > 0)) {
        atomicAdd(nonpos, 1);
        b = bb = 1.0;
    }
    T[i] = (float)(log(b) - log(bb));  // residual over the 8-node baseline
}

__global__ void k_standardise(float *T, size_t n, float mean, float inv_std)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) T[i] = (T[i] - mean) * inv_std;
}

__global__ void k_strided(const float *T, size_t n, size_t stride, float *out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = T[i * stride];
}

__global__ void k_gather_rows(const float *X, const float *T, const unsigned int *idx, int M, float *Xb, float *Tb)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    unsigned int j = idx[t];
    for (int k = 0; k < NIN; ++k) Xb[(size_t)t * NIN + k] = X[(size_t)j * NIN + k];
    Tb[t] = T[j];
}

int grid(size_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

// ================================================================ host side
// Buffers of one group's surrogate rounds (each group of tp_map3_batch has its own, so the groups'
// rounds can be in flight together on their streams).
struct TpSurRoundBufs {
    DevBuf X, F, Bb, flag, logb, Bex, list, seg, rpart, misc;
    HostPinned stats;  // the round's check statistics, copied back at the end of the round
    // events around the round's network launches (one pair per chunk) and those launches' points
    std::vector<hipEvent_t> ev;
    std::vector<long long> ev_points;
    int nev = 0;
    void release()
    {
        for (DevBuf *b : {&X, &F, &Bb, &flag, &logb, &Bex, &list, &seg, &rpart, &misc}) b->release();
        stats.release();
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        ev.clear();
        ev_points.clear();
        nev = 0;
    }
};

struct TpSurrogate {
    TpSurRoundBufs rb[TP_NLANES];
    float *P = nullptr, *G = nullptr, *Am = nullptr, *Av = nullptr;
    unsigned short *Wb = nullptr;  // [3][H][H]
    uint4 *Wf = nullptr;           // [NFRAG] hidden weights, BF16, in k_mlp_reg's fragment order
    uint4 *Wh = nullptr;           // [NFRAG] the same BF16 weights as FP16 (exact), for k_mlp_h
    int infer = TP_SUR_INFER_FP16; // inference kernel (tp_surrogate_inference)
    // activations / gradients for a batch of capacity cap rows
    int cap = 0;
    float *X = nullptr, *T = nullptr, *Y = nullptr, *dy = nullptr;
    float *Z[4] = {nullptr, nullptr, nullptr, nullptr};
    unsigned short *A[4] = {nullptr, nullptr, nullptr, nullptr};
    unsigned short *dZ[4] = {nullptr, nullptr, nullptr, nullptr};
    float *parts = nullptr;
    int nsplit = 0;
    double *d_loss = nullptr;
    SurBox box{};
    float t_mean = 0, t_std = 1;
    long long step = 0;
    bool trained = false;
};

static int sur_alloc(TpSurrogate *s)
{
    TP_HIP(hipMalloc(&s->P, NPARAM * sizeof(float)));
    TP_HIP(hipMalloc(&s->G, NPARAM * sizeof(float)));
    TP_HIP(hipMalloc(&s->Am, NPARAM * sizeof(float)));
    TP_HIP(hipMalloc(&s->Av, NPARAM * sizeof(float)));
    TP_HIP(hipMalloc(&s->Wb, 3 * H * H * sizeof(unsigned short)));
    TP_HIP(hipMalloc(&s->Wf, NFRAG * sizeof(uint4)));
    TP_HIP(hipMalloc(&s->Wh, NFRAG * sizeof(uint4)));
    TP_HIP(hipMalloc(&s->d_loss, sizeof(double)));
    TP_HIP(hipMemset(s->G, 0, NPARAM * sizeof(float)));
    TP_HIP(hipMemset(s->Am, 0, NPARAM * sizeof(float)));
    TP_HIP(hipMemset(s->Av, 0, NPARAM * sizeof(float)));
    return 0;
}

static int sur_ensure_batch(TpSurrogate *s, int rows)
{
    if (rows <= s->cap) return 0;
    int cap = ((rows + 63) / 64) * 64;
    auto F = [](void *p) { if (p) (void)hipFree(p); };
    F(s->X); F(s->T); F(s->Y); F(s->dy); F(s->parts);
    for (int i = 0; i < 4; ++i) { F(s->Z[i]); F(s->A[i]); F(s->dZ[i]); }
    TP_HIP(hipMalloc(&s->X, (size_t)cap * NIN * sizeof(float)));
    TP_HIP(hipMalloc(&s->T, (size_t)cap * sizeof(float)));
    TP_HIP(hipMalloc(&s->Y, (size_t)cap * sizeof(float)));
    TP_HIP(hipMalloc(&s->dy, (size_t)cap * sizeof(float)));
    for (int i = 0; i < 4; ++i) {
        TP_HIP(hipMalloc(&s->Z[i], (size_t)cap * H * sizeof(float)));
        TP_HIP(hipMalloc(&s->A[i], (size_t)cap * H * sizeof(unsigned short)));
        TP_HIP(hipMalloc(&s->dZ[i], (size_t)cap * H * sizeof(unsigned short)));
    }
    s->nsplit = std::max(1, std::min(64, cap / 256));
    TP_HIP(hipMalloc(&s->parts, (size_t)s->nsplit * H * H * sizeof(float)));
    s->cap = cap;
    return 0;
}

static void sur_free(TpSurrogate *s)
{
    if (!s) return;
    auto F = [](void *p) { if (p) (void)hipFree(p); };
    F(s->P); F(s->G); F(s->Am); F(s->Av); F(s->Wb); F(s->Wf); F(s->Wh); F(s->d_loss);
    F(s->X); F(s->T); F(s->Y); F(s->dy); F(s->parts);
    for (int i = 0; i < 4; ++i) { F(s->Z[i]); F(s->A[i]); F(s->dZ[i]); }
    for (auto &b : s->rb) b.release();
    delete s;
}

void tp_surrogate_free(tp_ctx *ctx)
{
    if (ctx && ctx->sur) {
        sur_free(ctx->sur);
        ctx->sur = nullptr;
    }
}

static void launch_gemm(hipStream_t st, const GemmArgs &g, int nsplit)
{
    dim3 gr(g.M / BM, g.N / BN, nsplit);
    hipLaunchKernelGGL(k_gemm_bf16, gr, dim3(256), 0, st, g);
}

// forward over rows [0, M) of s->X (M % 64 == 0); leaves Y (standardised log B)
static int sur_forward(hipStream_t st, TpSurrogate *s, int M, const float *target, double *loss_accum)
{
    hipLaunchKernelGGL(k_layer1_fwd, dim3(grid((size_t)M * H, 256)), dim3(256), 0, st, s->X, M, s->P, s->Z[0],
                       s->A[0]);
    for (int l = 0; l < 3; ++l) {
        GemmArgs g{};
        g.A = s->A[l];
        g.sam = H;
        g.sak = 1;
        g.B = s->Wb + (size_t)l * H * H;  // B(k, n) = W[n][k]
        g.sbk = 1;
        g.sbn = H;
        g.M = M;
        g.N = H;
        g.K = H;
        g.kchunk = H;
        g.epi = 1;
        g.bias = s->P + OFF_B[l];
        g.Z = s->Z[l + 1];
        g.Aout = s->A[l + 1];
        g.ldo = H;
        launch_gemm(st, g, 1);
    }
    hipLaunchKernelGGL(k_out_fwd, dim3(grid((size_t)M * 64, 256)), dim3(256), 0, st, s->A[3], s->Z[3], M, s->P, s->Y,
                       target, 1.0f / (float)M, s->dy, s->dZ[3], loss_accum);
    TP_HIP(hipGetLastError());
    return 0;
}

// backward after sur_forward with a target: gradients into s->G
static int sur_backward(hipStream_t st, TpSurrogate *s, int M)
{
    hipLaunchKernelGGL(k_out_grad, dim3(H + 1), dim3(256), 0, st, s->A[3], s->dy, M, s->G);
    for (int l = 2; l >= 0; --l) {
        // weight gradient: gW[n][k] = sum_m dZ[l+1][m][n] A[l][m][k]  (split-K over the batch)
        GemmArgs g{};
        g.A = s->dZ[l + 1];
        g.sam = 1;
        g.sak = H;  // A(m=n_out, k=row) = dZ[row][n_out]
        g.B = s->A[l];
        g.sbk = H;
        g.sbn = 1;  // B(k=row, n=k_in) = A[row][k_in]
        g.M = H;
        g.N = H;
        g.K = M;
        int ns = std::max(1, std::min(s->nsplit, M / BK));
        int kc = ((M / ns + BK - 1) / BK) * BK;
        ns = (M + kc - 1) / kc;
        g.kchunk = kc;
        g.epi = 0;
        g.C = s->parts;
        launch_gemm(st, g, ns);
        hipLaunchKernelGGL(k_reduce_parts, dim3(grid((size_t)H * H, 256)), dim3(256), 0, st, s->parts, ns,
                           (size_t)H * H, s->G + OFF_W[l]);
        hipLaunchKernelGGL(k_colsum_bf16, dim3(H), dim3(256), 0, st, s->dZ[l + 1], M, H, s->G + OFF_B[l]);
        // input gradient: dZ[l][m][k] = (sum_n dZ[l+1][m][n] W[n][k]) * silu'(Z[l][m][k])
        GemmArgs h{};
        h.A = s->dZ[l + 1];
        h.sam = H;
        h.sak = 1;
        h.B = s->Wb + (size_t)l * H * H;  // B(k=n_out, n=k_in) = W[n_out][k_in]
        h.sbk = H;
        h.sbn = 1;
        h.M = M;
        h.N = H;
        h.K = H;
        h.kchunk = H;
        h.epi = 2;
        h.Z = s->Z[l];
        h.Aout = s->dZ[l];
        h.ldo = H;
        launch_gemm(st, h, 1);
    }
    hipLaunchKernelGGL(k_layer1_grad, dim3(H), dim3(256), 0, st, s->dZ[0], s->X, M, s->G);
    TP_HIP(hipGetLastError());
    return 0;
}

namespace {
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <bool FP16>
__global__ void k_weights_frag(const float *__restrict__ P, u32x4 *__restrict__ Wf);
}
// BF16 copies of the hidden weights: Wb for the training GEMMs and, unless only training
// continues, Wf and Wh (the same values in FP16) for the inference kernels
static int sur_refresh_bf16(hipStream_t st, TpSurrogate *s, bool for_inference = true)
{
    hipLaunchKernelGGL(k_weights_bf16, dim3(grid((size_t)3 * H * H, 256)), dim3(256), 0, st, s->P, s->Wb);
    if (for_inference) {
        hipLaunchKernelGGL(k_weights_frag<false>, dim3(grid((size_t)NFRAG, 256)), dim3(256), 0, st, s->P,
                           (u32x4 *)s->Wf);
        hipLaunchKernelGGL(k_weights_frag<true>, dim3(grid((size_t)NFRAG, 256)), dim3(256), 0, st, s->P,
                           (u32x4 *)s->Wh);
    }
    TP_HIP(hipGetLastError());
    return 0;
}

static void init_params(std::vector<float> &p, unsigned long long seed)
{
    // He-style uniform init: U(-sqrt(6/fan_in), sqrt(6/fan_in)) for weights, zero biases
    p.assign(NPARAM, 0.f);
    unsigned long long s = seed;
    auto rnd = [&]() {
        s += 0x9e3779b97f4a7c15ULL;
        unsigned long long z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        return (double)(z >> 11) / 9007199254740992.0;
    };
    auto fill = [&](size_t off, int fan_out, int fan_in) {
        double a = std::sqrt(6.0 / fan_in);
        for (size_t i = 0; i < (size_t)fan_out * fan_in; ++i) p[off + i] = (float)((2 * rnd() - 1) * a);
    };
    fill(OFF_W1, H, NIN);
    fill(OFF_W2, H, H);
    fill(OFF_W3, H, H);
    fill(OFF_W4, H, H);
    fill(OFF_W5, 1, H);
}

static int ensure_sur(tp_ctx *ctx)
{
    if (ctx->sur) return 0;
    TpSurrogate *s = new TpSurrogate();
    int rc = sur_alloc(s);
    if (rc) {
        sur_free(s);
        return rc;
    }
    ctx->sur = s;
    return 0;
}

extern "C" int tp_mlp_num_params(size_t *n)
{
    if (!n) return tp_set_error(TP_EINVAL, "NULL");
    *n = NPARAM;
    return 0;
}

extern "C" int tp_mlp_get_params(tp_ctx *ctx, float *p)
{
    if (!ctx || !p || !ctx->sur) return tp_set_error(TP_EINVAL, "tp_mlp_get_params: no model");
    TP_HIP(hipMemcpy(p, ctx->sur->P, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int tp_mlp_set_params(tp_ctx *ctx, const float *p)
{
    if (!ctx || !p) return tp_set_error(TP_EINVAL, "tp_mlp_set_params: NULL");
    TP_HIP(hipSetDevice(ctx->device));
    int rc = ensure_sur(ctx);
    if (rc) return rc;
    TP_HIP(hipMemcpy(ctx->sur->P, p, NPARAM * sizeof(float), hipMemcpyHostToDevice));
    return sur_refresh_bf16(ctx->stream, ctx->sur);
}

extern "C" int tp_mlp_loss_grad(tp_ctx *ctx, size_t n, const float *x, const float *t, double *loss, float *grad)
{
    if (!ctx || !x || !t || !loss || !grad || n == 0) return tp_set_error(TP_EINVAL, "tp_mlp_loss_grad: bad args");
    if (n % 64) return tp_set_error(TP_EINVAL, "tp_mlp_loss_grad: n must be a multiple of 64");
    if (!ctx->sur) return tp_set_error(TP_EINVAL, "tp_mlp_loss_grad: set parameters first");
    TP_HIP(hipSetDevice(ctx->device));
    TpSurrogate *s = ctx->sur;
    int rc = sur_ensure_batch(s, (int)n);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    TP_HIP(hipMemcpyAsync(s->X, x, n * NIN * sizeof(float), hipMemcpyHostToDevice, st));
    TP_HIP(hipMemcpyAsync(s->T, t, n * sizeof(float), hipMemcpyHostToDevice, st));
    TP_HIP(hipMemsetAsync(s->d_loss, 0, sizeof(double), st));
    if ((rc = sur_forward(st, s, (int)n, s->T, s->d_loss))) return rc;
    if ((rc = sur_backward(st, s, (int)n))) return rc;
    TP_HIP(hipMemcpyAsync(grad, s->G, NPARAM * sizeof(float), hipMemcpyDeviceToHost, st));
    TP_HIP(hipMemcpyAsync(loss, s->d_loss, sizeof(double), hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    return 0;
}

namespace {
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;

// ---------------------------------------------------------------- register-resident inference
// The network (8 -> 4 x 256 SiLU -> 1) evaluated for the surrogate, one persistent workgroup per
// CU, 128 points per chunk (32 per wave).  The layers are computed transposed, Z^T = W A^T: the
// MFMA's A operand is a 32-unit tile of W and its B operand the activations of 32 points, so the
// 32 x 32 FP32 result of one layer lands in the lane layout the next layer's B operand needs
// (lane = (point r, half h), accumulator element 4 j + i = unit 8 j + 4 h + i of the tile).
// Feeding 8 consecutive accumulator elements as one 16-wide K slice pairs K index k with unit
// u(k) = k with bits 2 and 3 swapped, so the weights are stored with their input columns in that
// order.  Activations never leave the registers: hidden activations enter the BF16 MFMAs as
// hi + lo pairs (a = hi + lo to ~16 significant bits, two MFMAs per weight fragment), which keeps
// the surrogate a smooth function of its inputs -- one BF16 rounding of the activations leaves a
// ~1e-3 staircase that an adaptive cubature cannot converge through.  Layer 1 (8 inputs) uses the
// FP32 MFMA on unrounded inputs.  The weights stream through two 32 KB LDS slots, four K-blocks
// (16 inputs x 256 units each) per slot, loaded straight into LDS one stage ahead.  Two
// accumulator sets alternate between layers so that the previous layer's bias/SiLU/split work
// can issue between the current layer's MFMAs.
constexpr int RB = 128;     // points per chunk: 4 waves x 32
constexpr int KPS = 4;      // K-blocks per LDS stage (32 KB)
constexpr int NSTAGE = 12;  // stages per chunk: 3 layers x 16 K-blocks / KPS
constexpr int SU4 = KPS * 512;  // 16-byte fragments per stage
constexpr int SP_B1 = H * NIN, SP_W5 = SP_B1 + 4 * H, SP_N = SP_W5 + H;  // LDS copy: W1, b1..b4, W5

// FP16: the BF16-rounded weight (the value the network was trained with), exact in FP16 down to
// |w| = 2^-14 (subnormal below; 2^-24 of a weight is far below the activations' rounding)
template <bool FP16>
__global__ void k_weights_frag(const float *__restrict__ P, u32x4 *__restrict__ Wf)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NFRAG) return;
    const int lane = e & 63, t = (e >> 6) & 7, kb = (e >> 9) & 15, l = e >> 13;
    const int r = lane & 31, h = lane >> 5;
    const float *W = P + OFF_W[l];
    unsigned int w[4];
    for (int q2 = 0; q2 < 4; ++q2) {
        unsigned int v[2];
        for (int m = 0; m < 2; ++m) {
            const int k = 16 * kb + 8 * h + 2 * q2 + m;
            const int u = (k & ~12) | ((k & 4) << 1) | ((k & 8) >> 1);
            const unsigned short b = f2bf(W[(size_t)(32 * t + r) * H + u]);
            v[m] = FP16 ? __builtin_bit_cast(unsigned short, (_Float16)bf2f(b)) : b;
        }
        w[q2] = v[0] | (v[1] << 16);
    }
    Wf[e] = u32x4{w[0], w[1], w[2], w[3]};
}

// Bias, SiLU and the hi/lo BF16 split of half JJ of tile T of a layer's FP32 result: the B
// operands of the next layer's K slice 2T+JJ.
template <int T, int JJ>
__device__ __forceinline__ void mlp_epi_half(const f32x16 (&acc)[8], const float *__restrict__ bias, int h,
                                             bf16x8 (&Bh)[16], bf16x8 (&Bl)[16])
{
    const float4 b0 = *(const float4 *)(bias + 32 * T + 16 * JJ + 4 * h);
    const float4 b1 = *(const float4 *)(bias + 32 * T + 16 * JJ + 8 + 4 * h);
    const float b[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float v = silu_fast(acc[T][8 * JJ + q] + b[q]);
        const __bf16 hi = (__bf16)v;
        Bh[2 * T + JJ][q] = hi;
        Bl[2 * T + JJ][q] = (__bf16)(v - (float)hi);
    }
}

// One 16-wide K slice: 8 weight fragments times the hi and the lo activations (KB is a template
// argument so that every register array index is a compile-time constant).
template <int KB>
__device__ __forceinline__ void mlp_kslice(const u32x4 *src, f32x16 (&acc)[8], const bf16x8 (&Bh)[16],
                                           const bf16x8 (&Bl)[16], int lane)
{
    bf16x8 a[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) a[t] = __builtin_bit_cast(bf16x8, src[t * 64 + lane]);
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[t], Bh[KB], acc[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[t], Bl[KB], acc[t], 0, 0, 0);
}

// Stage ST of a hidden layer: the next stage of weights starts loading straight into the other
// LDS slot (global_load_lds, no registers; the end-of-stage barrier waits for it); then the
// stage's KPS K-slices from this slot, accumulating into out.  Stage ST needs the previous
// layer's tiles 2ST and 2ST+1 (K-slices 4ST..4ST+3); the epilogue of tiles 2ST+2 and 2ST+3 is
// placed in the same stage so that its VALU work can issue while the stage's MFMAs execute
// (in and out are separate accumulator sets).  gs = the layer's first stage within NSTAGE.
__device__ __forceinline__ void mlp_stage_load(u32x4 *slot, const u32x4 *__restrict__ Wf, int stage, int wave,
                                               int lane)
{
#pragma unroll
    for (int m = 0; m < 8; ++m)
        __builtin_amdgcn_global_load_lds((const void *)(Wf + stage * SU4 + m * 256 + wave * 64 + lane),
                                         (__attribute__((address_space(3))) void *)(slot + m * 256 + wave * 64), 16,
                                         0, 0);
}

// Scheduling pattern of one K slice (masks of sched_group_barrier: 0x100 DS read, 0x8 MFMA,
// 0x2 VALU)
__device__ __forceinline__ void mlp_interleave()
{
    __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
    }
}

template <int ST>
__device__ __forceinline__ void mlp_stages(const f32x16 (&in)[8], f32x16 (&out)[8], const float *__restrict__ bias_in,
                                           bf16x8 (&Bh)[16], bf16x8 (&Bl)[16], u32x4 (*ring)[SU4],
                                           const u32x4 *__restrict__ Wf, int gs, int wave, int lane, int h)
{
    if constexpr (ST < 16 / KPS) {
        mlp_stage_load(ring[(ST + 1) & 1], Wf, (gs + ST + 1) % NSTAGE, wave, lane);
        const u32x4 *src = ring[ST & 1];
        // K slice k with a quarter of the next tiles' epilogue, scheduled as 8 fragment reads and
        // then 16 x (1 MFMA + up to 5 VALU) so that the VALU work issues in the MFMAs' shadow;
        // sched_barrier keeps each slice's reads from being hoisted into the previous slice
        mlp_kslice<KPS * ST + 0>(src, out, Bh, Bl, lane);
        if constexpr (ST < 3) mlp_epi_half<2 * ST + 2, 0>(in, bias_in, h, Bh, Bl);
        mlp_interleave();
        __builtin_amdgcn_sched_barrier(0);
        mlp_kslice<KPS * ST + 1>(src + 512, out, Bh, Bl, lane);
        if constexpr (ST < 3) mlp_epi_half<2 * ST + 2, 1>(in, bias_in, h, Bh, Bl);
        mlp_interleave();
        __builtin_amdgcn_sched_barrier(0);
        mlp_kslice<KPS * ST + 2>(src + 1024, out, Bh, Bl, lane);
        if constexpr (ST < 3) mlp_epi_half<2 * ST + 3, 0>(in, bias_in, h, Bh, Bl);
        mlp_interleave();
        __builtin_amdgcn_sched_barrier(0);
        mlp_kslice<KPS * ST + 3>(src + 1536, out, Bh, Bl, lane);
        if constexpr (ST < 3) mlp_epi_half<2 * ST + 3, 1>(in, bias_in, h, Bh, Bl);
        mlp_interleave();
        __syncthreads();
        mlp_stages<ST + 1>(in, out, bias_in, Bh, Bl, ring, Wf, gs, wave, lane, h);
    }
}

// One hidden layer: out = W_L (SiLU(in + bias_in) as hi/lo BF16); tiles 0 and 1 of the
// input epilogue run before the first stage, the rest inside the stages.
template <int L>
__device__ __forceinline__ void mlp_layer(const f32x16 (&in)[8], f32x16 (&out)[8], const float *__restrict__ bias_in,
                                          bf16x8 (&Bh)[16], bf16x8 (&Bl)[16], u32x4 (*ring)[SU4],
                                          const u32x4 *__restrict__ Wf, int wave, int lane, int h)
{
    mlp_epi_half<0, 0>(in, bias_in, h, Bh, Bl);
    mlp_epi_half<0, 1>(in, bias_in, h, Bh, Bl);
    mlp_epi_half<1, 0>(in, bias_in, h, Bh, Bl);
    mlp_epi_half<1, 1>(in, bias_in, h, Bh, Bl);
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int i = 0; i < 16; ++i) out[t][i] = 0.f;
    mlp_stages<0>(in, out, bias_in, Bh, Bl, ring, Wf, L * (16 / KPS), wave, lane, h);
}

__global__ void __launch_bounds__(256, 1) k_mlp_reg(const float *__restrict__ X, int M, const float *__restrict__ P,
                                                    const u32x4 *__restrict__ Wf, float mean, float sd,
                                                    double *__restrict__ logb)
{
    __shared__ u32x4 ring[2][SU4];
    // W1, b1..b4 and W5 in LDS: reads of them after a barrier stay inside the chunk loop
    // (loads straight from P are loop-invariant and would be hoisted into 800 live registers)
    __shared__ __attribute__((aligned(16))) float sP[SP_N];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = lane & 31, h = lane >> 5;
    const int nchunk = (M + RB - 1) / RB;
    for (int e = tid; e < SP_N; e += 256) {
        size_t src;
        if (e < SP_B1) {
            src = OFF_W1 + e;
        } else if (e < SP_W5) {
            const int which = (e - SP_B1) / H, k = (e - SP_B1) % H;  // b1, b2, b3, b4
            src = (which == 0 ? OFF_B1 : OFF_B[which - 1]) + k;
        } else {
            src = OFF_W5 + (e - SP_W5);
        }
        sP[e] = P[src];
    }
    mlp_stage_load(ring[0], Wf, 0, wave, lane);
    __syncthreads();
#pragma unroll 1
    for (int c = blockIdx.x; c < nchunk; c += gridDim.x) {
        const int row = c * RB + wave * 32 + r;
        const bool valid = row < M;
        f32x16 acc[8];
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[t][i] = 0.f;
#pragma unroll
        for (int cc = 0; cc < NIN / 2; ++cc) {
            const float xb = valid ? X[(size_t)row * NIN + 2 * cc + h] : 0.f;
#pragma unroll
            for (int t = 0; t < 8; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(sP[(32 * t + r) * NIN + 2 * cc + h], xb, acc[t], 0, 0, 0);
        }
        bf16x8 Bh[16], Bl[16];
        f32x16 acc2[8];
        // an opaque copy of the weight pointer per chunk: the 96 per-lane stage addresses derived
        // from it are recomputed in the loop instead of being hoisted out of it and spilled
        const u32x4 *wf = Wf;
        asm volatile("" : "+s"(wf));
        mlp_layer<0>(acc, acc2, sP + SP_B1, Bh, Bl, ring, wf, wave, lane, h);
        mlp_layer<1>(acc2, acc, sP + SP_B1 + H, Bh, Bl, ring, wf, wave, lane, h);
        mlp_layer<2>(acc, acc2, sP + SP_B1 + 2 * H, Bh, Bl, ring, wf, wave, lane, h);
        // output layer on the FP32 activations of layer 4
        float o = 0.f;
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float4 b = *(const float4 *)(sP + SP_B1 + 3 * H + 32 * t + 8 * j + 4 * h);
                const float4 w = *(const float4 *)(sP + SP_W5 + 32 * t + 8 * j + 4 * h);
                o += w.x * silu_fast(acc2[t][4 * j] + b.x) + w.y * silu_fast(acc2[t][4 * j + 1] + b.y) +
                     w.z * silu_fast(acc2[t][4 * j + 2] + b.z) + w.w * silu_fast(acc2[t][4 * j + 3] + b.w);
            }
        o += __shfl_xor(o, 32, 64);
        if (h == 0 && valid) logb[row] = (double)(o + P[OFF_B5]) * (double)sd + (double)mean;
    }
}

// ---------------------------------------------------------------- k_mlp_h: FP16 activations, one pass
// The same transposed layers, weight order and LDS weight stream as k_mlp_reg, with the hidden
// activations rounded once to FP16 (11 significant bits; round to nearest even) and the weights in
// FP16 (the trained BF16 values, exact): v_mfma_f32_32x32x16_f16, one MFMA per weight fragment
// instead of two.  The rounding leaves a relative staircase of ~1e-5 in the output, which the 1e-4
// cubature converges through with about the same point counts (one BF16 rounding, 2^-8, did not).
// One accumulator set: a layer's SiLU/FP16 epilogue runs between its MFMA phases, and two workgroups
// per CU (256 VGPRs per lane, 77 KB of LDS each) keep the matrix cores busy through it.
//
// Scaled pre-activations: the accumulators hold u = -log2(e) z instead of z (layer 1's weights and
// every bias are scaled by -log2(e) in LDS, and each layer's accumulators start at its scaled bias),
// and the activations carry the same factor, a' = -log2(e) SiLU(z) = u / (1 + 2^u): the hidden
// weights are unchanged (the factor passes through them) and the output weights are divided by
// -log2(e).  SiLU then costs v_exp_f32, one add, v_rcp_f32 and one multiply.
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(2 * sizeof(float)))) float f32x2;
typedef __attribute__((__vector_size__(2 * sizeof(_Float16)))) _Float16 f16x2;
constexpr float kNegLog2e = -1.44269504088896340736f;

__device__ __forceinline__ float silu_scaled(float u) { return u * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(u)); }

// SiLU + FP16 of tile T of a layer's scaled result: the B operands of K slices 2T and 2T+1
template <int T>
__device__ __forceinline__ void mlph_epi(const f32x16 (&acc)[8], f16x8 (&B)[16])
{
#pragma unroll
    for (int JJ = 0; JJ < 2; ++JJ)
#pragma unroll
        for (int q = 0; q < 8; q += 2) {
            const f32x2 v = {silu_scaled(acc[T][8 * JJ + q]), silu_scaled(acc[T][8 * JJ + q + 1])};
            const f16x2 hv = __builtin_convertvector(v, f16x2);
            B[2 * T + JJ][q] = hv[0];
            B[2 * T + JJ][q + 1] = hv[1];
        }
}

// accumulators := the layer's scaled bias (element i of tile t is unit 32 t + 8 (i / 4) + 4 h + i % 4)
__device__ __forceinline__ void mlph_bias_init(f32x16 (&acc)[8], const float *__restrict__ bias, int h)
{
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float4 b = *(const float4 *)(bias + 32 * t + 8 * j + 4 * h);
            acc[t][4 * j] = b.x;
            acc[t][4 * j + 1] = b.y;
            acc[t][4 * j + 2] = b.z;
            acc[t][4 * j + 3] = b.w;
        }
}

// one stage (4 K slices x 8 weight fragments) as 8 groups of 4 fragments (K slice 4ST + g/2, tiles
// 4(g%2)..4(g%2)+3): group g+1 is read from LDS while group g's MFMAs issue, so that an MFMA waits for a
// read issued four MFMAs earlier.  Left to itself the compiler reads each fragment just before its MFMA
// and waits for it there; the sched_group_barrier sequence (0x100 DS read, 0x008 MFMA) fixes the order.
template <int ST>
__device__ __forceinline__ void mlph_stage(const u32x4 *src, f32x16 (&acc)[8], const f16x8 (&B)[16], int lane)
{
    static_assert(KPS == 4, "8 groups of 4 fragments = 4 K slices of 8 fragments");
    f16x8 a[2][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) a[0][t] = __builtin_bit_cast(f16x8, src[t * 64 + lane]);
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        if (g < 7) {
#pragma unroll
            for (int t = 0; t < 4; ++t)
                a[(g + 1) & 1][t] =
                    __builtin_bit_cast(f16x8, src[((g + 1) >> 1) * 512 + (4 * ((g + 1) & 1) + t) * 64 + lane]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
            acc[4 * (g & 1) + t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[g & 1][t], B[KPS * ST + (g >> 1)],
                                                                          acc[4 * (g & 1) + t], 0, 0, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
#pragma unroll
    for (int g = 0; g < 7; ++g) {
        __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
}

template <int ST>
__device__ __forceinline__ void mlph_stages(f32x16 (&acc)[8], const f16x8 (&B)[16], u32x4 (*ring)[SU4],
                                            const u32x4 *__restrict__ Wf, int gs, int wave, int lane)
{
    if constexpr (ST < 16 / KPS) {
        mlp_stage_load(ring[(ST + 1) & 1], Wf, (gs + ST + 1) % NSTAGE, wave, lane);
        mlph_stage<ST>(ring[ST & 1], acc, B, lane);
        __syncthreads();
        mlph_stages<ST + 1>(acc, B, ring, Wf, gs, wave, lane);
    }
}

// one hidden layer in place: acc = W_L a' + bias_out (scaled), from acc = the previous layer's scaled result
template <int L>
__device__ __forceinline__ void mlph_layer(f32x16 (&acc)[8], const float *__restrict__ bias_out, u32x4 (*ring)[SU4],
                                           const u32x4 *__restrict__ Wf, int wave, int lane, int h)
{
    f16x8 B[16];
    // the epilogue and the bias reads stay here (hoisted into the previous layer's MFMA phase they only spill)
    __builtin_amdgcn_sched_barrier(0);
    mlph_epi<0>(acc, B);
    mlph_epi<1>(acc, B);
    mlph_epi<2>(acc, B);
    mlph_epi<3>(acc, B);
    mlph_epi<4>(acc, B);
    mlph_epi<5>(acc, B);
    mlph_epi<6>(acc, B);
    mlph_epi<7>(acc, B);
    mlph_bias_init(acc, bias_out, h);
    __builtin_amdgcn_sched_barrier(0);
    mlph_stages<0>(acc, B, ring, Wf, L * (16 / KPS), wave, lane);
}

__global__ void __launch_bounds__(256, 2) k_mlp_h(const float *__restrict__ X, int M, const float *__restrict__ P,
                                                  const u32x4 *__restrict__ Wf, float mean, float sd,
                                                  double *__restrict__ logb)
{
    __shared__ u32x4 ring[2][SU4];
    // W1 and b1..b4 times -log2(e), W5 divided by it
    __shared__ __attribute__((aligned(16))) float sP[SP_N];
    // the wave index in a scalar register: the LDS slot addresses of the weight stream (M0) stay scalar
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r = lane & 31, h = lane >> 5;
    const int nchunk = (M + RB - 1) / RB;
    for (int e = tid; e < SP_N; e += 256) {
        size_t src;
        float scale = kNegLog2e;
        if (e < SP_B1) {
            src = OFF_W1 + e;
        } else if (e < SP_W5) {
            const int which = (e - SP_B1) / H, k = (e - SP_B1) % H;  // b1, b2, b3, b4
            src = (which == 0 ? OFF_B1 : OFF_B[which - 1]) + k;
        } else {
            src = OFF_W5 + (e - SP_W5);
            scale = 1.0f / kNegLog2e;
        }
        sP[e] = P[src] * scale;
    }
    mlp_stage_load(ring[0], Wf, 0, wave, lane);
    __syncthreads();
#pragma unroll 1
    for (int c = blockIdx.x; c < nchunk; c += gridDim.x) {
        const int row = c * RB + wave * 32 + r;
        const bool valid = row < M;
        f32x16 acc[8];
        mlph_bias_init(acc, sP + SP_B1, h);
#pragma unroll
        for (int cc = 0; cc < NIN / 2; ++cc) {
            const float xb = valid ? X[(size_t)row * NIN + 2 * cc + h] : 0.f;
#pragma unroll
            for (int t = 0; t < 8; ++t)
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(sP[(32 * t + r) * NIN + 2 * cc + h], xb, acc[t], 0, 0, 0);
        }
        const u32x4 *wf = Wf;
        asm volatile("" : "+s"(wf));
        mlph_layer<0>(acc, sP + SP_B1 + H, ring, wf, wave, lane, h);
        mlph_layer<1>(acc, sP + SP_B1 + 2 * H, ring, wf, wave, lane, h);
        mlph_layer<2>(acc, sP + SP_B1 + 3 * H, ring, wf, wave, lane, h);
        float o = 0.f;
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float4 w = *(const float4 *)(sP + SP_W5 + 32 * t + 8 * j + 4 * h);
                o += w.x * silu_scaled(acc[t][4 * j]) + w.y * silu_scaled(acc[t][4 * j + 1]) +
                     w.z * silu_scaled(acc[t][4 * j + 2]) + w.w * silu_scaled(acc[t][4 * j + 3]);
            }
        o += __shfl_xor(o, 32, 64);
        if (h == 0 && valid) logb[row] = (double)(o + P[OFF_B5]) * (double)sd + (double)mean;
    }
}
}  // namespace

// log B for rows [0, M) of X (M % 64 == 0) into logb, using two ping-pong activation buffers
static int sur_infer(hipStream_t st, TpSurrogate *s, const float *X, int M, unsigned short *buf0, unsigned short *buf1,
                     double *logb)
{
    (void)buf0;
    (void)buf1;
    if (M <= 0) return 0;
    static int ncu = 0;  // one persistent workgroup per CU
    if (!ncu) {
        int dev = 0;
        TP_HIP(hipGetDevice(&dev));
        TP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        if (ncu < 1) ncu = 1;
    }
    const int nchunk = (M + RB - 1) / RB;
    if (s->infer == TP_SUR_INFER_FP16)  // two workgroups per CU
        hipLaunchKernelGGL(k_mlp_h, dim3(std::min(nchunk, 2 * ncu)), dim3(256), 0, st, X, M, s->P,
                           (const u32x4 *)s->Wh, s->t_mean, s->t_std, logb);
    else
        hipLaunchKernelGGL(k_mlp_reg, dim3(std::min(nchunk, ncu)), dim3(256), 0, st, X, M, s->P,
                           (const u32x4 *)s->Wf, s->t_mean, s->t_std, logb);
    TP_HIP(hipGetLastError());
    return 0;
}

// ================================================================ training
namespace {
__global__ void k_sample_idx(unsigned long long key, unsigned long long step, int M, unsigned int ntrain,
                             unsigned int *idx)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    idx[t] = (unsigned int)(u01(key, step * 1048576ULL + t) * ntrain) % ntrain;
}
__global__ void k_val_err(const float *Y, const float *T, int M, float sd, double *sse, float *maxerr)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    float e = fabsf(Y[t] - T[t]) * sd;
    atomicAdd(sse, (double)e * (double)e);
    atomicMax((int *)maxerr, __float_as_int(e));  // e >= 0: int order = float order
}
}  // namespace

extern "C" int tp_surrogate_train(tp_ctx *ctx, const tp_train_config *cfg, tp_train_report *rep)
{
    if (!ctx || !cfg || !cfg->nz || cfg->nsamples < 1024 || cfg->batch < 64 || cfg->batch % 64 || cfg->epochs < 1)
        return tp_set_error(TP_EINVAL, "tp_surrogate_train: bad configuration");
    TP_HIP(hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    tp_surrogate_free(ctx);
    int rc = ensure_sur(ctx);
    if (rc) return rc;
    TpSurrogate *s = ctx->sur;
    hipStream_t st = ctx->stream;
    SurBox box;
    box.lnl_min = cfg->lnl_min;
    box.lnl_max = cfg->lnl_max;
    const double *lo[5] = {cfg->om, cfg->s8, cfg->h, cfg->w, cfg->ns};
    for (int k = 0; k < 5; ++k) {
        box.lo[k] = lo[k][0];
        box.hi[k] = lo[k][1];
    }
    s->box = box;
    // cosmologies of the training box, each with its own FP64 GPU tables
    const int ncos = 2048;
    std::vector<tp_cosmology> cps(ncos);
    unsigned long long sd = cfg->seed ^ 0xC05u;
    auto rnd = [&]() {
        sd += 0x9e3779b97f4a7c15ULL;
        unsigned long long z = sd;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        return (double)(z >> 11) / 9007199254740992.0;
    };
    for (int i = 0; i < ncos; ++i) {
        cps[i].om = box.lo[0] + (box.hi[0] - box.lo[0]) * rnd();
        cps[i].sigma8 = box.lo[1] + (box.hi[1] - box.lo[1]) * rnd();
        cps[i].h = box.lo[2] + (box.hi[2] - box.lo[2]) * rnd();
        cps[i].w = box.lo[3] + (box.hi[3] - box.lo[3]) * rnd();
        cps[i].ns = box.lo[4] + (box.hi[4] - box.lo[4]) * rnd();
        cps[i].omb = cfg->omb;
        cps[i].zmax = cfg->z_max;
    }
    tp_cosmo *cos = nullptr;
    if ((rc = tp_cosmo_build(ctx, ncos, cps.data(), cfg->nz, cfg->n_z, 0, cfg->z_max, &cos))) return rc;
    size_t N = (size_t)cfg->nsamples;
    float *dX = nullptr, *dT = nullptr;
    int *dnp = nullptr;
    unsigned int *didx = nullptr;
    auto cleanup = [&]() {
        if (dX) (void)hipFree(dX);
        if (dT) (void)hipFree(dT);
        if (dnp) (void)hipFree(dnp);
        if (didx) (void)hipFree(didx);
        tp_cosmo_destroy(cos);
    };
#define TRY(call)                                                                                             \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess) {                                                                               \
            cleanup();                                                                                        \
            return tp_set_error(TP_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));                  \
        }                                                                                                     \
    } while (0)
    TRY(hipMalloc(&dX, N * NIN * sizeof(float)));
    TRY(hipMalloc(&dT, N * sizeof(float)));
    TRY(hipMalloc(&dnp, sizeof(int)));
    TRY(hipMalloc(&didx, (size_t)cfg->batch * sizeof(unsigned int)));
    TRY(hipMemsetAsync(dnp, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_gen, dim3(grid(N, 256)), dim3(256), 0, st, cos->d_node64, cos->d_eh64, cos->d_base32,
                       cos->d_eh32, cos->d_const, ncos, box,
                       (unsigned long long)(cfg->seed * 0x2545F4914F6CDD1DULL + 17), N, dX, dT, dnp);
    TRY(hipGetLastError());
    // standardisation of log B from a strided host subsample
    std::vector<float> sub(std::min<size_t>(N, 1 << 20));
    size_t stride = std::max<size_t>(1, N / sub.size());
    TRY(hipFree(didx));
    didx = nullptr;
    TRY(hipMalloc(&didx, std::max((size_t)cfg->batch, sub.size()) * sizeof(unsigned int)));
    hipLaunchKernelGGL(k_strided, dim3(grid(sub.size(), 256)), dim3(256), 0, st, dT, sub.size(), stride,
                       (float *)didx);
    TRY(hipMemcpyAsync(sub.data(), didx, sub.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    int nonpos = 0;
    TRY(hipMemcpyAsync(&nonpos, dnp, sizeof(int), hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
    double mean = 0, var = 0;
    for (float v : sub) mean += v;
    mean /= sub.size();
    for (float v : sub) var += (v - mean) * (v - mean);
    var /= sub.size();
    s->t_mean = (float)mean;
    s->t_std = (float)std::sqrt(std::max(var, 1e-12));
    hipLaunchKernelGGL(k_standardise, dim3(grid(N, 256)), dim3(256), 0, st, dT, N, s->t_mean, 1.0f / s->t_std);
    auto t1 = std::chrono::steady_clock::now();
    // parameters
    std::vector<float> p0;
    init_params(p0, cfg->seed);
    TRY(hipMemcpyAsync(s->P, p0.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice, st));
    TRY(hipMemsetAsync(s->Am, 0, NPARAM * sizeof(float), st));
    TRY(hipMemsetAsync(s->Av, 0, NPARAM * sizeof(float), st));
    if ((rc = sur_refresh_bf16(st, s)) || (rc = sur_ensure_batch(s, cfg->batch))) {
        cleanup();
        return rc;
    }
    size_t nval = std::max<size_t>(64, (N / 100) / 64 * 64);
    unsigned int ntrain = (unsigned int)(N - nval);
    long long steps = (long long)cfg->epochs * (long long)(ntrain / cfg->batch);
    const int M = cfg->batch;
    double last_loss = 0;
    const unsigned long long key = cfg->seed * 0x9E3779B97F4A7C15ULL + 99;
    for (long long it = 0; it < steps; ++it) {
        double frac = (double)it / (double)std::max<long long>(1, steps - 1);
        double lr = cfg->lr_final + 0.5 * (cfg->lr - cfg->lr_final) * (1 + std::cos(M_PI * frac));
        hipLaunchKernelGGL(k_sample_idx, dim3(grid(M, 256)), dim3(256), 0, st, key, (unsigned long long)it, M, ntrain,
                           didx);
        hipLaunchKernelGGL(k_gather_rows, dim3(grid(M, 256)), dim3(256), 0, st, dX, dT, didx, M, s->X, s->T);
        bool report = (it == steps - 1) || (it % 500 == 0);
        if (report) TRY(hipMemsetAsync(s->d_loss, 0, sizeof(double), st));
        if ((rc = sur_forward(st, s, M, s->T, s->d_loss)) || (rc = sur_backward(st, s, M))) {
            cleanup();
            return rc;
        }
        s->step++;
        float c1 = 1.0f - (float)std::pow(0.9, (double)s->step), c2 = 1.0f - (float)std::pow(0.999, (double)s->step);
        hipLaunchKernelGGL(k_adam, dim3(grid(NPARAM, 256)), dim3(256), 0, st, s->P, s->G, s->Am, s->Av, NPARAM,
                           (float)lr, 0.9f, 0.999f, 1e-8f, c1, c2);
        if ((rc = sur_refresh_bf16(st, s, it == steps - 1))) {
            cleanup();
            return rc;
        }
        if (report) {
            TRY(hipMemcpyAsync(&last_loss, s->d_loss, sizeof(double), hipMemcpyDeviceToHost, st));
            TRY(hipStreamSynchronize(st));
            std::fprintf(stderr, "surrogate train step %lld/%lld lr %.2e loss %.4e\n", it, steps, lr, last_loss);
        }
    }
    // validation on the held-out tail, in batch-sized chunks
    double *dsse = nullptr;
    float *dmax = nullptr;
    TRY(hipMalloc(&dsse, sizeof(double)));
    TRY(hipMalloc(&dmax, sizeof(float)));
    TRY(hipMemsetAsync(dsse, 0, sizeof(double), st));
    TRY(hipMemsetAsync(dmax, 0, sizeof(float), st));
    for (size_t off = ntrain; off + M <= N; off += M) {
        TRY(hipMemcpyAsync(s->X, dX + off * NIN, (size_t)M * NIN * sizeof(float), hipMemcpyDeviceToDevice, st));
        if ((rc = sur_forward(st, s, M, nullptr, nullptr))) {
            cleanup();
            return rc;
        }
        hipLaunchKernelGGL(k_val_err, dim3(grid(M, 256)), dim3(256), 0, st, s->Y, dT + off, M, s->t_std, dsse, dmax);
    }
    double sse = 0;
    float vmax = 0;
    TRY(hipMemcpyAsync(&sse, dsse, sizeof(double), hipMemcpyDeviceToHost, st));
    TRY(hipMemcpyAsync(&vmax, dmax, sizeof(float), hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
    (void)hipFree(dsse);
    (void)hipFree(dmax);
    auto t2 = std::chrono::steady_clock::now();
    size_t nvchk = ((N - ntrain) / M) * M;
    if (rep) {
        rep->final_loss = last_loss;
        rep->val_rmse = nvchk ? std::sqrt(sse / nvchk) : 0;
        rep->val_maxerr = vmax;
        rep->gen_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        rep->train_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        rep->steps = steps;
    }
    if (nonpos) std::fprintf(stderr, "surrogate: %d training samples had B_kappa <= 0\n", nonpos);
    cleanup();
    s->trained = true;
    return 0;
#undef TRY
}

// ---------------------------------------------------------------- persistence
static const char kMagic[8] = {'T', 'P', 'M', 'L', 'P', '0', '0', '2'};  // residual over the 8-node baseline

extern "C" int tp_surrogate_save(tp_ctx *ctx, const char *path)
{
    if (!ctx || !ctx->sur || !path) return tp_set_error(TP_EINVAL, "tp_surrogate_save: no model");
    std::vector<float> p(NPARAM);
    TP_HIP(hipSetDevice(ctx->device));
    TP_HIP(hipMemcpy(p.data(), ctx->sur->P, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    std::ofstream f(path, std::ios::binary);
    if (!f) return tp_set_error(TP_EINVAL, std::string("tp_surrogate_save: cannot open ") + path);
    f.write(kMagic, 8);
    f.write((const char *)&ctx->sur->box, sizeof(SurBox));
    f.write((const char *)&ctx->sur->t_mean, sizeof(float));
    f.write((const char *)&ctx->sur->t_std, sizeof(float));
    f.write((const char *)p.data(), NPARAM * sizeof(float));
    return f ? 0 : tp_set_error(TP_EINVAL, "tp_surrogate_save: write failed");
}

extern "C" int tp_surrogate_load(tp_ctx *ctx, const char *path)
{
    if (!ctx || !path) return tp_set_error(TP_EINVAL, "tp_surrogate_load: NULL");
    std::ifstream f(path, std::ios::binary);
    if (!f) return tp_set_error(TP_EINVAL, std::string("tp_surrogate_load: cannot open ") + path);
    char mg[8];
    SurBox box;
    float mean, sd;
    std::vector<float> p(NPARAM);
    f.read(mg, 8);
    f.read((char *)&box, sizeof(SurBox));
    f.read((char *)&mean, sizeof(float));
    f.read((char *)&sd, sizeof(float));
    f.read((char *)p.data(), NPARAM * sizeof(float));
    if (!f || std::memcmp(mg, kMagic, 8) != 0)
        return tp_set_error(TP_EINVAL, std::string("tp_surrogate_load: not a surrogate file: ") + path);
    for (float v : p)
        if (!std::isfinite(v)) return tp_set_error(TP_EINVAL, "tp_surrogate_load: non-finite parameter");
    if (!(sd > 0) || !std::isfinite(mean) || !(box.lnl_max > box.lnl_min))
        return tp_set_error(TP_EINVAL, "tp_surrogate_load: corrupted header");
    TP_HIP(hipSetDevice(ctx->device));
    int rc = ensure_sur(ctx);
    if (rc) return rc;
    ctx->sur->box = box;
    ctx->sur->t_mean = mean;
    ctx->sur->t_std = sd;
    TP_HIP(hipMemcpy(ctx->sur->P, p.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
    ctx->sur->trained = true;
    return sur_refresh_bf16(ctx->stream, ctx->sur);
}

extern "C" int tp_surrogate_inference(tp_ctx *ctx, int kind)
{
    if (!ctx || (kind != TP_SUR_INFER_HILO_BF16 && kind != TP_SUR_INFER_FP16))
        return tp_set_error(TP_EINVAL, "tp_surrogate_inference: kind must be TP_SUR_INFER_HILO_BF16 or _FP16");
    int rc;
    if ((rc = ensure_sur(ctx))) return rc;
    ctx->sur->infer = kind;
    return 0;
}

extern "C" int tp_surrogate_config(tp_ctx *ctx, double check_frac, double tol, uint64_t seed)
{
    if (!ctx || !(check_frac > 0) || check_frac > 1 || !(tol > 0))
        return tp_set_error(TP_EINVAL, "tp_surrogate_config: check_frac in (0,1], tol > 0");
    ctx->sur_check_frac = check_frac;
    ctx->sur_tol = tol;
    ctx->sur_seed = seed;
    return 0;
}

namespace {
// raw (l1, l2, l3, om, s8, h, w, ns) -> normalised inputs; cos phi from l3^2 = l1^2 + l2^2 + 2 l1 l2 cos phi
__global__ void k_pred_inputs(const double *in, size_t n, size_t npad, SurBox box, float *X)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    if (i >= n) {
        for (int k = 0; k < NIN; ++k) X[i * NIN + k] = 0.f;
        return;
    }
    const double *r = in + 8 * i;
    double cp[5] = {r[3], r[4], r[5], r[6], r[7]};
    make_input(box, r[0], r[1], r[2], cp, X + i * NIN);
}
}  // namespace

extern "C" int tp_surrogate_predict(tp_ctx *ctx, size_t n, const double *in8, double *y)
{
    if (!ctx || !ctx->sur || !ctx->sur->trained) return tp_set_error(TP_EINVAL, "tp_surrogate_predict: no model");
    if (n == 0) return 0;
    if (!in8 || !y) return tp_set_error(TP_EINVAL, "tp_surrogate_predict: NULL");
    TP_HIP(hipSetDevice(ctx->device));
    TpSurrogate *s = ctx->sur;
    hipStream_t st = ctx->stream;
    const size_t chunk = 1 << 18;
    int rc;
    if ((rc = ctx->d_in.ensure(chunk * 8 * sizeof(double))) || (rc = ctx->d_out.ensure(chunk * sizeof(double))) ||
        (rc = ctx->d_misc.ensure(chunk * (NIN * sizeof(float) + 2 * H * sizeof(unsigned short)))))
        return rc;
    float *X = (float *)ctx->d_misc.p;
    unsigned short *b0 = (unsigned short *)(X + chunk * NIN), *b1 = b0 + chunk * H;
    for (size_t off = 0; off < n; off += chunk) {
        size_t m = std::min(chunk, n - off), mp = (m + 63) / 64 * 64;
        TP_HIP(hipMemcpyAsync(ctx->d_in.p, in8 + 8 * off, m * 8 * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_pred_inputs, dim3(grid(mp, 256)), dim3(256), 0, st, (const double *)ctx->d_in.p, m, mp,
                           s->box, X);
        if ((rc = sur_infer(st, s, X, (int)mp, b0, b1, (double *)ctx->d_out.p))) return rc;
        TP_HIP(hipMemcpyAsync(y + off, ctx->d_out.p, m * sizeof(double), hipMemcpyDeviceToHost, st));
        TP_HIP(hipStreamSynchronize(st));
    }
    return 0;
}

// ================================================================ deployment in the cubature
namespace {

// degree-7 Genz-Malik weight of point p of a 3-D region (without the volume factor)
__device__ __forceinline__ double gm_weight3(int p)
{
    int nz = 0, mx = 0;
    for (int d = 0; d < 3; ++d) {
        int a = c_gm3_code[p][d] < 0 ? -c_gm3_code[p][d] : c_gm3_code[p][d];
        if (a) ++nz;
        mx = a > mx ? a : mx;
    }
    if (nz == 0) return -10936.0 / 19683.0;         // centre
    if (mx == 1) return 980. / 6561.;               // lambda2 axis points
    if (mx == 3) return 6859.0 / 19683.0 / 8.0;     // lambda5 corners
    return nz == 1 ? 620.0 / 19683.0 : 200. / 19683.;  // lambda4 axis points / (lambda4, lambda4) pairs
}

// The network's input row of a cubature point and its box test (every side's log inside the training range,
// every cosmology parameter inside its box) from one set of FP32 logarithms of the sorted sides; make_input
// (training data, tp_surrogate_predict) takes FP64 logs: the rows agree to ~2e-7, far inside the network's
// own precision.  l3 > 0.
__device__ __forceinline__ bool make_input_in_box(const SurBox &B, double l1, double l2, double l3, const double cp[5],
                                                  float *x)
{
    double a = l1, b = l2, c = l3, t;
    if (a < b) { t = a; a = b; b = t; }
    if (b < c) { t = b; b = c; c = t; }
    if (a < b) { t = a; a = b; b = t; }
    const float ga = logf((float)a), gb = logf((float)b), gc = logf((float)c);
    const float mid = (float)(0.5 * (B.lnl_min + B.lnl_max)), ihalf = (float)(2.0 / (B.lnl_max - B.lnl_min));
    x[0] = (ga - mid) * ihalf;
    x[1] = (gb - mid) * ihalf;
    x[2] = (gc - mid) * ihalf;
    bool in = gc >= (float)B.lnl_min && ga <= (float)B.lnl_max;
    for (int k = 0; k < 5; ++k) {
        x[3 + k] = (float)(2.0 * (cp[k] - B.lo[k]) / (B.hi[k] - B.lo[k]) - 1.0);
        const double tol = 1e-9 * (B.hi[k] - B.lo[k]);
        in = in && cp[k] >= B.lo[k] - tol && cp[k] <= B.hi[k] + tol;
    }
    return in;
}

// flag: 0 surrogate, 1 surrogate + exact check, 2 exact only (outside the training box), 3 zero (an l = 0)
// a point to evaluate exactly, as k_sur_points formed it: its sides, its index in the chunk and its cosmology
// (k_sur_exact reads this one record instead of the region, the integral's descriptor and the point rule)
struct SurPt {
    double l[3];
    int t, ic;
};

__global__ void __launch_bounds__(256) k_sur_points(const TpRegion *__restrict__ reg, size_t r0, size_t nr,
                                                    const TpIntegralDesc *__restrict__ desc,
                                                    const TpCosmoConst *__restrict__ cc, const TpNode32 *__restrict__ base,
                                                    const TpEH32 *__restrict__ eh32, SurBox box, double frac,
                                                    unsigned long long key, float *X, double *F, double *Bb, int *flag,
                                                    SurPt *list, int *seg, int *err)
{
    __shared__ int s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int f = 3;
    SurPt pt{{0.0, 0.0, 0.0}, (int)t, 0};
    if (t < nr * TP_NPTS_GM3) {  // every thread reaches the barriers below
        size_t r = r0 + t / TP_NPTS_GM3;
        int p = (int)(t % TP_NPTS_GM3);
        const TpRegion R = reg[r];
        const TpIntegralDesc D = desc[R.integral];
        double l1, l2, phi;
        gm_point3(R, p, &l1, &l2, &phi);
        double l3 = tp_ell3(l1, l2, phi);
        if (!isfinite(l1) || !isfinite(l2) || !isfinite(l3)) atomicOr(err, 1);
        pt.l[0] = l1;
        pt.l[1] = l2;
        pt.l[2] = l3;
        pt.ic = D.icosmo;
        // l1 l2 u(l1 t1) u(l2 t2) u(l3 t3), u(x) = x^2/2 exp(-x^2/2), with one exponential for the three filters
        const double a = l1 * D.th[0], b = l2 * D.th[1], e = l3 * D.th[2];
        const double a2 = a * a, b2 = b * b, e2 = e * e;
        F[t] = l1 * l2 * (0.125 * a2 * b2 * e2) * exp(-0.5 * (a2 + b2 + e2));
        const TpCosmoConst &C = cc[D.icosmo];
        double cp[5] = {C.om, C.sigma8, C.h, C.w, C.ns};
        const bool inb = make_input_in_box(box, l1, l2, l3 > 0 ? l3 : 1.0, cp, X + t * NIN);
        if (l1 == 0 || l2 == 0 || l3 == 0)
            f = 3;
        else if (!inb)
            f = 2;
        else
            f = (u01(key, r * TP_NPTS_GM3 + p) < frac) ? 1 : 0;
        flag[t] = f;
        Bb[t] = (f <= 1) ? tp_bkappa_sum_n<float, TP_NBASE>(base + (size_t)D.icosmo * TP_NBASE, eh32[D.icosmo], l1, l2,
                                                            l3)
                         : 0.0;
    }
    // the points to evaluate exactly (checked or outside the box) go to this block's own segment of the
    // list: an LDS counter, no device-wide atomic (atomics on one address serialise)
    const bool want = f == 1 || f == 2;
    int slot = 0;
    if (want) slot = atomicAdd(&s_cnt, 1);
    __syncthreads();
    if (want) list[(size_t)blockIdx.x * blockDim.x + slot] = pt;
    if (threadIdx.x == 0) seg[blockIdx.x] = s_cnt;
}

// exact B_kappa of the listed points in the FP32 form (FP32 bispectrum, FP64 z-sum: the arithmetic of the
// exact regions of a surrogate round, so the checks correct the network toward the same values): one
// workgroup per list segment (one block of k_sur_points: typically 1% of its 256 points plus the out-of-box
// ones, which cluster, so a whole workgroup shares a full segment), 16 lanes per point, 6 of the 96 z-nodes
// each.
constexpr int XL = 16;
__global__ void __launch_bounds__(256) k_sur_exact(const SurPt *__restrict__ list, const int *seg,
                                                   const TpNode32 *__restrict__ nodes, const TpEH32 *__restrict__ eh,
                                                   double *Bex, int *err)
{
    const int n = seg[blockIdx.x];
    const int sub = threadIdx.x % XL;
    // the XL lanes of a group share i, so a group runs the loop (and the shuffle below) with all its lanes
    for (int i = threadIdx.x / XL; i < n; i += blockDim.x / XL) {
        const SurPt q0 = list[(size_t)blockIdx.x * blockDim.x + i];
        const int t = q0.t;
        const double l1 = q0.l[0], l2 = q0.l[1], l3 = q0.l[2];
        TpPointPre P;
        tp_point_pre<double>(l1, l2, l3, &P);
        const float L[3] = {(float)l1, (float)l2, (float)l3};
        const float LL[3] = {(float)P.ll[0], (float)P.ll[1], (float)P.ll[2]};
        const float F2T[3] = {(float)P.f2t[0], (float)P.f2t[1], (float)P.f2t[2]};
        const float LR1 = (float)P.lr1, R2 = (float)P.r2;
        const TpNode32 *nd = nodes + (size_t)q0.ic * TP_NNODES;
        const TpEH32 E = eh[q0.ic];
        double part = 0.0;
        for (int k = sub; k < TP_NNODES; k += XL) {
            const TpNode32 &N = nd[k];
            float q[3] = {L[0] * N.ks, L[1] * N.ks, L[2] * N.ks};
            float lq[3] = {LL[0] + N.lnks, LL[1] + N.lnks, LL[2] + N.lnks};
            part += (double)(N.wz * tp_bispec_node<float>(N, E, q, lq, F2T, LR1, R2));
        }
        for (int o = XL / 2; o >= 1; o >>= 1) part += __shfl_xor(part, o, XL);
        if (sub == 0) {
            if (!isfinite(part)) atomicOr(err, 2);
            Bex[t] = part;
        }
    }
}

// The end of a surrogate chunk, per region: each point's value f = F B_8 e^net (F B_exact where the point was
// evaluated exactly), the Genz-Malik rule over the region's 33 values, and the Horvitz-Thompson correction
// from its checked points.  A workgroup holds kSurFinRegions whole regions (33 threads each), so the values
// stay in LDS and every global read is coalesced; the round's check sums go per workgroup into part
// (k_sur_count adds them up; atomics on one address serialise).
constexpr int kSurFinRegions = 7;  // 7 x 33 = 231 of 256 threads
__global__ void __launch_bounds__(256) k_sur_fin(const TpRegion *__restrict__ reg, size_t r0, size_t nr,
                                                 const int *__restrict__ flag, const double *__restrict__ F,
                                                 const double *__restrict__ logb, const double *__restrict__ Bb,
                                                 const double *__restrict__ Bex, double inv_frac, TpRegionOut *out,
                                                 float *maxrel, double *part)
{
    __shared__ double sv[kSurFinRegions][TP_NPTS_GM3];
    __shared__ double sx[4][kSurFinRegions][TP_NPTS_GM3];  // w r, (w r)^2, |w| |F B_exact|, checked
    __shared__ double red[kSurFinRegions][4];
    const int rl = threadIdx.x / TP_NPTS_GM3, p = threadIdx.x % TP_NPTS_GM3;
    const size_t rr = (size_t)blockIdx.x * kSurFinRegions + rl;
    const bool in = rl < kSurFinRegions && rr < nr;  // every thread reaches the barriers below
    float e = 0.f;
    if (in) {
        const size_t t = rr * TP_NPTS_GM3 + p;
        const int f = flag[t];
        // unconditional loads (Bex is read for every point and used only where it was computed)
        const double Fv = F[t], lb = logb[t], bb = Bb[t], bx = Bex[t];
        const double vol = reg[r0 + rr].vol;
        // baseline x learned correction; the network's output carries FP32 precision, so its exp is FP32
        const double bs = bb * (double)__expf((float)lb);
        double v = 0, r = 0, a = 0;
        if (f <= 1) v = Fv * bs;
        if (f == 1) {
            r = Fv * (bx - bs);
            a = fabs(Fv * bx);
            e = (float)(fabs(bx - bs) / fmax(fabs(bx), 1e-300));
        }
        if (f == 2) v = Fv * bx;
        const double w = gm_weight3(p) * vol, x = w * r;
        sv[rl][p] = v;
        sx[0][rl][p] = x;
        sx[1][rl][p] = x * x;
        sx[2][rl][p] = fabs(w) * a;
        sx[3][rl][p] = f == 1 ? 1.0 : 0.0;
    }
    // the worst checked error: a wave that cannot raise the running maximum takes no atomic
    for (int o = 32; o >= 1; o >>= 1) e = fmaxf(e, __shfl_xor(e, o, 64));
    if ((threadIdx.x & 63) == 0 && e > __int_as_float(__atomic_load_n((const int *)maxrel, __ATOMIC_RELAXED)))
        atomicMax((int *)maxrel, __float_as_int(e));
    __syncthreads();
    if (threadIdx.x < kSurFinRegions) {
        const int q = threadIdx.x;
        const size_t rq = (size_t)blockIdx.x * kSurFinRegions + q;
        double sum[4] = {0, 0, 0, 0};
        if (rq < nr) {
            for (int k = 0; k < TP_NPTS_GM3; ++k)
                for (int c = 0; c < 4; ++c) sum[c] += sx[c][q][k];
            TpRegionOut o;
            gm_rule3(reg[r0 + rq].vol, sv[q], &o);
            // Horvitz-Thompson: corr = sum w r / pi, Var ~ (1 - pi) sum (w r / pi)^2
            o.corr = sum[0] * inv_frac;
            o.cvar = (1.0 - 1.0 / inv_frac) * sum[1] * inv_frac * inv_frac;
            out[r0 + rq] = o;
        }
        red[q][0] = sum[0];
        red[q][1] = sum[2];
        red[q][2] = sum[1];
        red[q][3] = sum[3];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double acc = 0;
        for (int q = 0; q < kSurFinRegions; ++q) acc += red[q][threadIdx.x];
        part[(size_t)blockIdx.x * 4 + threadIdx.x] = acc;
    }
}

}  // namespace


// layout of misc / stats: [0] sum w r, [1] sum |w F B|, [2] sum (w r)^2, [3] checks, [4] exactly evaluated
// points (as a double), then the worst checked relative error (float)
constexpr int kSurStatsD = 5;

namespace {
// a chunk's statistics into the round's: the exactly evaluated points (the sum of the list segments) and the
// rule blocks' check sums
__global__ void __launch_bounds__(256) k_sur_count(const int *seg, int nseg, const double *part, int nrb, double *stats)
{
    __shared__ double red[4][5];
    double c[5] = {0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < nseg; i += 256) c[4] += (double)seg[i];
    for (int i = threadIdx.x; i < nrb; i += 256)
        for (int k = 0; k < 4; ++k) c[k] += part[(size_t)i * 4 + k];
    for (int k = 0; k < 5; ++k)
        for (int o = 32; o >= 1; o >>= 1) c[k] += __shfl_xor(c[k], o, 64);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 5; ++k) red[threadIdx.x >> 6][k] = c[k];
    __syncthreads();
    if (threadIdx.x < 5) stats[threadIdx.x] += red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                                              red[3][threadIdx.x];
}
}  // namespace

int tp_surrogate_queue(tp_ctx *ctx, const tp_cosmo *c, int lane, size_t nreg, const TpRegion *d_reg,
                       const TpIntegralDesc *d_desc, TpRegionOut *d_out, int *d_flag, hipStream_t st,
                       unsigned long long round_key)
{
    TpSurrogate *s = ctx->sur;
    if (!s || !s->trained) return tp_set_error(TP_EINVAL, "surrogate mode without a trained model");
    if (lane < 0 || lane >= TP_NLANES) return tp_set_error(TP_EINVAL, "tp_surrogate_queue: bad group");
    if (nreg == 0) return 0;
    const size_t RC = 32768;  // regions per chunk (1 081 344 points): bounds the per-point buffers
    const size_t PC = RC * TP_NPTS_GM3, PCP = (PC + 63) / 64 * 64;
    const size_t NSEG = (PC + 255) / 256;  // k_sur_points blocks per chunk: one list segment of 256 each
    int rc;
    TpSurRoundBufs &B = s->rb[lane];
    if ((rc = B.X.ensure(PCP * NIN * sizeof(float))) || (rc = B.F.ensure(PC * sizeof(double))) ||
        (rc = B.Bb.ensure(PC * sizeof(double))) || (rc = B.flag.ensure(PC * sizeof(int))) ||
        (rc = B.logb.ensure(PCP * sizeof(double))) || (rc = B.Bex.ensure(PC * sizeof(double))) ||
        (rc = B.list.ensure(NSEG * 256 * sizeof(SurPt))) ||
        (rc = B.seg.ensure(NSEG * sizeof(int))) || (rc = B.rpart.ensure((RC + kSurFinRegions - 1) / kSurFinRegions * 4 * sizeof(double))) ||
        (rc = B.misc.ensure(64)) || (rc = B.stats.ensure(64)))
        return rc;
    double *dstats = (double *)B.misc.p;
    float *dmaxrel = (float *)(dstats + kSurStatsD);
    TP_HIP(hipMemsetAsync(B.misc.p, 0, 64, st));
    const double frac = ctx->sur_check_frac;
    const unsigned long long key = ctx->sur_seed * 0x9E3779B97F4A7C15ULL + round_key * 7919ULL;
    // the chunks follow each other on the stream; nothing here waits for the device
    B.nev = 0;
    B.ev_points.clear();
    for (size_t r0 = 0; r0 < nreg; r0 += RC) {
        size_t nr = std::min(RC, nreg - r0), np = nr * TP_NPTS_GM3, npp = (np + 63) / 64 * 64;
        const int nseg = grid(np, 256);
        hipLaunchKernelGGL(k_sur_points, dim3(nseg), dim3(256), 0, st, d_reg, r0, nr, d_desc, c->d_const,
                           c->d_base32, c->d_eh32, s->box, frac, key, (float *)B.X.p, (double *)B.F.p,
                           (double *)B.Bb.p, (int *)B.flag.p, (SurPt *)B.list.p, (int *)B.seg.p, d_flag);
        if (npp > np) TP_HIP(hipMemsetAsync((float *)B.X.p + np * NIN, 0, (npp - np) * NIN * sizeof(float), st));
        while ((int)B.ev.size() < 2 * (B.nev + 1)) {
            hipEvent_t e;
            TP_HIP(hipEventCreate(&e));
            B.ev.push_back(e);
        }
        TP_HIP(hipEventRecord(B.ev[2 * B.nev], st));
        if ((rc = sur_infer(st, s, (const float *)B.X.p, (int)npp, nullptr, nullptr, (double *)B.logb.p))) return rc;
        TP_HIP(hipEventRecord(B.ev[2 * B.nev + 1], st));
        B.ev_points.push_back((long long)npp);
        B.nev += 1;
        hipLaunchKernelGGL(k_sur_exact, dim3(nseg), dim3(256), 0, st, (const SurPt *)B.list.p, (const int *)B.seg.p,
                           c->d_node32, c->d_eh32, (double *)B.Bex.p, d_flag);
        const int nrb = (int)((nr + kSurFinRegions - 1) / kSurFinRegions);
        hipLaunchKernelGGL(k_sur_fin, dim3(nrb), dim3(256), 0, st, d_reg, r0, nr, (const int *)B.flag.p,
                           (const double *)B.F.p, (const double *)B.logb.p, (const double *)B.Bb.p,
                           (const double *)B.Bex.p, 1.0 / frac, d_out, dmaxrel, (double *)B.rpart.p);
        hipLaunchKernelGGL(k_sur_count, dim3(1), dim3(256), 0, st, (const int *)B.seg.p, nseg,
                           (const double *)B.rpart.p, nrb, dstats);
        TP_HIP(hipGetLastError());
    }
    // read back by a kernel, not a DMA copy waiting on the round (tp_copy_to_pinned)
    return tp_copy_to_pinned(B.misc.p, B.stats.dp, 64, st);
}

int tp_surrogate_result(tp_ctx *ctx, int lane, size_t nreg, bool *fallback)
{
    *fallback = false;
    if (nreg == 0) return 0;
    TpSurRoundBufs &B = ctx->sur->rb[lane];
    for (int i = 0; i < B.nev; ++i) {  // the network's launches of the round (event-timed on their stream)
        float ms = 0;
        TP_HIP(hipEventElapsedTime(&ms, B.ev[2 * i], B.ev[2 * i + 1]));
        ctx->stats.mlp_ms += ms;
        ctx->stats.mlp_launches += 1;
        ctx->stats.mlp_points += B.ev_points[i];
    }
    const double *hs = (const double *)B.stats.p;
    const float hmax = *(const float *)(hs + kSurStatsD);
    ctx->stats.exact_checks += (long long)hs[4];
    ctx->stats.max_check_relerr = std::max(ctx->stats.max_check_relerr, (double)hmax);
    // relative standard error of this round's correction: sqrt((1 - pi) sum (w r)^2) / sum |w F B|
    const double relse = hs[1] > 0 ? std::sqrt((1.0 - ctx->sur_check_frac) * hs[2]) / hs[1] : 0.0;
    *fallback = relse > ctx->sur_tol;
    if (!*fallback) ctx->stats.surrogate_points += (long long)(nreg * TP_NPTS_GM3);
    return 0;
}

namespace {
__global__ void k_surb_inputs(const double *ell, size_t n, size_t npad, const TpCosmoConst *__restrict__ cc, int ic,
                              const TpNode32 *__restrict__ base, const TpEH32 *__restrict__ eh32, SurBox box, float *X,
                              double *Bb)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad) return;
    if (i >= n) {
        for (int k = 0; k < NIN; ++k) X[i * NIN + k] = 0.f;
        return;
    }
    double l1 = ell[3 * i], l2 = ell[3 * i + 1], l3 = ell[3 * i + 2];
    const TpCosmoConst &C = cc[ic];
    double cp[5] = {C.om, C.sigma8, C.h, C.w, C.ns};
    make_input(box, l1, l2, l3, cp, X + i * NIN);
    Bb[i] = (l1 > 0 && l2 > 0 && l3 > 0)
                ? tp_bkappa_sum_n<float, TP_NBASE>(base + (size_t)ic * TP_NBASE, eh32[ic], l1, l2, l3)
                : 0.0;
}
__global__ void k_surb_combine(size_t n, const double *Bb, double *y)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = Bb[i] * exp(y[i]);
}
}  // namespace

extern "C" int tp_surrogate_bkappa(tp_ctx *ctx, const tp_cosmo *c, int icosmo, size_t npts, const double *ell123,
                                   double *out)
{
    if (c && c->pk_nk) return tp_set_error(TP_EINVAL, "tp_surrogate_bkappa: tables from a tabulated P(k)");
    if (!ctx || !ctx->sur || !ctx->sur->trained) return tp_set_error(TP_EINVAL, "tp_surrogate_bkappa: no model");
    if (!c || icosmo < 0 || icosmo >= c->ncosmo) return tp_set_error(TP_EINVAL, "tp_surrogate_bkappa: bad cosmology");
    if (npts == 0) return 0;
    if (!ell123 || !out) return tp_set_error(TP_EINVAL, "tp_surrogate_bkappa: NULL");
    TP_HIP(hipSetDevice(ctx->device));
    TpSurrogate *s = ctx->sur;
    hipStream_t st = ctx->stream;
    const size_t chunk = 1 << 18;
    int rc;
    if ((rc = ctx->d_in.ensure(chunk * 3 * sizeof(double))) || (rc = ctx->d_out.ensure(2 * chunk * sizeof(double))) ||
        (rc = ctx->d_misc.ensure(chunk * NIN * sizeof(float))))
        return rc;
    float *X = (float *)ctx->d_misc.p;
    double *Y = (double *)ctx->d_out.p, *Bb = Y + chunk;
    for (size_t off = 0; off < npts; off += chunk) {
        size_t m = std::min(chunk, npts - off), mp = (m + 63) / 64 * 64;
        TP_HIP(hipMemcpyAsync(ctx->d_in.p, ell123 + 3 * off, m * 3 * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_surb_inputs, dim3(grid(mp, 256)), dim3(256), 0, st, (const double *)ctx->d_in.p, m, mp,
                           c->d_const, icosmo, c->d_base32, c->d_eh32, s->box, X, Bb);
        if ((rc = sur_infer(st, s, X, (int)mp, nullptr, nullptr, Y))) return rc;
        hipLaunchKernelGGL(k_surb_combine, dim3(grid(m, 256)), dim3(256), 0, st, m, Bb, Y);
        TP_HIP(hipMemcpyAsync(out + off, Y, m * sizeof(double), hipMemcpyDeviceToHost, st));
        TP_HIP(hipStreamSynchronize(st));
    }
    return 0;
}
