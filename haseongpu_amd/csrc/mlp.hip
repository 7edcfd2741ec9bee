// mlp.hip - the learned proposal's network on gfx950: FP32 training kernels
// (forward, backward, AdamW) and the BF16-MFMA inference kernel that scores
// (sample, reflection slice, prism cluster) rows for the proposal.
//
// Network: 10 inputs -> 4 x 128 SiLU -> 1 (the log transfer factor).
//
// Training runs layer by layer through one tiled FP32 GEMM kernel whose
// epilogues fuse bias + SiLU (forward), the SiLU derivative (backward) or
// nothing; weight gradients reduce over the batch in split-K partials summed
// in a fixed order, so a training run is reproducible.
//
// Inference keeps the whole network in LDS as BF16 operand fragments
// (100 KB) and runs each 32-row tile through v_mfma_f32_32x32x16_bf16 with
// the neurons as the M dimension and rows as N: a layer's FP32 accumulators
// then already sit in the lane layout of the next layer's B operand, up to a
// fixed permutation of the K index (bits 2 and 3 of each 16-block swapped)
// that is folded into the packed weights, so activations never leave the
// registers between layers.  Each wavefront takes 32 rows at a time; a
// 512-thread work-group (8 wavefronts, 2 per SIMD) shares one LDS copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/hase_ase.h"
#include "common.h"
#include "context.h"
#include "prism_walk.h"
#include "proposal.h"

namespace hase {

// ------------------------------------------------------------ FP32 training
namespace {

__device__ __forceinline__ float silu_exact(float z) { return z / (1.0f + expf(-z)); }

__device__ __forceinline__ float silu_grad(float z)
{
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// C(i, j) = sum_k A(i, k) B(k, j) over one K chunk, A(i, k) = A[i*sai + k*sak],
// B(k, j) = B[k*sbk + j*sbj]; blockIdx.z selects the K chunk and the partial
// slice C + z * split_stride.
struct GemmArgs {
    const float* A;
    long long sai, sak;
    const float* B;
    long long sbk, sbj;
    float* C;
    long long ldc, split_stride;
    int M, N, K, kchunk;
    const float* bias;   // EPI_BIAS, EPI_BIAS_SILU: + bias[j]
    float* act;          // EPI_BIAS_SILU: act(i, j) = silu(C(i, j))
    const float* gate;   // EPI_GATE: C(i, j) *= silu'(gate(i, j))
    int epi;
};

enum { EPI_NONE = 0, EPI_BIAS_SILU = 1, EPI_GATE = 2, EPI_BIAS = 3 };

constexpr int kTM = 64, kTN = 64, kTK = 16;

__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g)
{
    __shared__ float As[kTK][kTM + 4];
    __shared__ float Bs[kTK][kTN + 4];
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int i0 = blockIdx.y * kTM, j0 = blockIdx.x * kTN;
    const int kb = blockIdx.z * g.kchunk;
    const int ke = min(g.K, kb + g.kchunk);
    float acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = 0.0f;
    for (int k0 = kb; k0 < ke; k0 += kTK) {
        for (int e = tid; e < kTK * kTM; e += 256) {
            int kk, ii;
            if (g.sak == 1) { kk = e % kTK; ii = e / kTK; } else { ii = e % kTM; kk = e / kTM; }
            const int i = i0 + ii, k = k0 + kk;
            As[kk][ii] = (i < g.M && k < ke) ? g.A[(long long)i * g.sai + (long long)k * g.sak] : 0.0f;
        }
        for (int e = tid; e < kTK * kTN; e += 256) {
            int kk, jj;
            if (g.sbj == 1) { jj = e % kTN; kk = e / kTN; } else { kk = e % kTK; jj = e / kTK; }
            const int j = j0 + jj, k = k0 + kk;
            Bs[kk][jj] = (j < g.N && k < ke) ? g.B[(long long)k * g.sbk + (long long)j * g.sbj] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = As[kk][ty * 4 + r];
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = Bs[kk][tx * 4 + c];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = fmaf(a[r], b[c], acc[r][c]);
        }
        __syncthreads();
    }
    float* C = g.C + (long long)blockIdx.z * g.split_stride;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = i0 + ty * 4 + r;
        if (i >= g.M) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int j = j0 + tx * 4 + c;
            if (j >= g.N) continue;
            const long long at = (long long)i * g.ldc + j;
            float v = acc[r][c];
            if (g.epi == EPI_BIAS || g.epi == EPI_BIAS_SILU) v += g.bias[j];
            if (g.epi == EPI_GATE) v *= silu_grad(g.gate[at]);
            C[at] = v;
            if (g.epi == EPI_BIAS_SILU) g.act[at] = silu_exact(v);
        }
    }
}

// out[i] = sum_z part[z*stride + i], fixed order
__global__ void split_sum_kernel(const float* part, int splits, long long stride, int count, float* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += part[(long long)z * stride + i];
    out[i] = s;
}

// dy_i = 2 w_i (y_i - t_i) / W over the batch; loss += sum w r^2 / W
__global__ void loss_kernel(const float* y, const float* t, const float* w, int n, float inv_wsum, float* dy,
                            double* loss)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        const float wi = w ? w[i] : 1.0f;
        const float r = y[i] - t[i];
        dy[i] = 2.0f * wi * r * inv_wsum;
        part = (double)wi * (double)r * (double)r * (double)inv_wsum;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if ((threadIdx.x & 63) == 0 && part != 0.0) atomicAdd(loss, part);
}

__global__ void gather_kernel(const float* X, const float* t, const float* w, const uint32_t* perm, int b0, int nb,
                              float* Xb, float* tb, float* wb)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nb) return;
    const uint32_t src = perm[b0 + r];
#pragma unroll
    for (int f = 0; f < kMlpIn; ++f) Xb[(size_t)r * kMlpIn + f] = X[(size_t)src * kMlpIn + f];
    tb[r] = t[src];
    wb[r] = w ? w[src] : 1.0f;
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                            float eps, float wd, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    p[i] -= lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
}

template <class T>
T* dptr(DevBuf<T>& b) { return b.p; }

// Device state of the FP32 network for batches of up to `cap` rows.
struct Trainer {
    hipStream_t st = nullptr;
    int cap = 0;
    DevBuf<float> params, grad, m, v;
    DevBuf<float> X, t, w;          // batch rows
    DevBuf<float> Z[4], A[4];       // pre- and post-activation per hidden layer
    DevBuf<float> y, dy, dZa, dZb;
    DevBuf<float> part;             // split-K partials
    DevBuf<float> one;
    DevBuf<double> loss;
    int splits = 1;

    explicit Trainer(int capacity)
    {
        HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        cap = std::max(1, capacity);
        params.alloc(kMlpParams);
        grad.alloc(kMlpParams);
        X.alloc((size_t)cap * kMlpIn);
        t.alloc(cap);
        w.alloc(cap);
        for (int l = 0; l < 4; ++l) {
            Z[l].alloc((size_t)cap * kMlpH);
            A[l].alloc((size_t)cap * kMlpH);
        }
        y.alloc(cap);
        dy.alloc(cap);
        dZa.alloc((size_t)cap * kMlpH);
        dZb.alloc((size_t)cap * kMlpH);
        splits = std::max(1, std::min(64, (cap + 255) / 256));
        part.alloc((size_t)splits * kMlpH * kMlpH);
        one.alloc(1);
        const float f1 = 1.0f;
        HIP_CHECK(hipMemcpyAsync(one.p, &f1, sizeof(float), hipMemcpyHostToDevice, st));
        loss.alloc(1);
    }
    ~Trainer()
    {
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    }

    void gemm(GemmArgs g, int splits_used)
    {
        dim3 grid((g.N + kTN - 1) / kTN, (g.M + kTM - 1) / kTM, splits_used);
        hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, st, g);
        HIP_CHECK(hipGetLastError());
    }

    // out[M x N] = op over the batch dimension K = n, reduced in split-K partials
    void batch_reduce(GemmArgs g, float* out)
    {
        const int n = g.K;
        const int sp = std::max(1, std::min(splits, (n + 255) / 256));
        int chunk = (n + sp - 1) / sp;
        chunk = (chunk + kTK - 1) / kTK * kTK;
        const int used = (n + chunk - 1) / chunk;
        g.C = part.p;
        g.ldc = g.N;
        g.split_stride = (long long)g.M * g.N;
        g.kchunk = chunk;
        g.epi = EPI_NONE;
        gemm(g, used);
        const int count = g.M * g.N;
        hipLaunchKernelGGL(split_sum_kernel, dim3((count + 255) / 256), dim3(256), 0, st, (const float*)part.p, used,
                           (long long)count, count, out);
        HIP_CHECK(hipGetLastError());
    }

    GemmArgs base() const
    {
        GemmArgs g{};
        g.kchunk = 1 << 30;
        return g;
    }

    void forward(int n, const float* Xin)
    {
        const float* P = params.p;
        const float* in = Xin;
        int in_dim = kMlpIn;
        const size_t offW[4] = {kOffW1, kOffW2, kOffW3, kOffW4};
        const size_t offB[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
        for (int l = 0; l < 4; ++l) {
            GemmArgs g = base();
            g.A = in; g.sai = in_dim; g.sak = 1;
            g.B = P + offW[l]; g.sbk = 1; g.sbj = in_dim;
            g.C = Z[l].p; g.ldc = kMlpH;
            g.M = n; g.N = kMlpH; g.K = in_dim;
            g.bias = P + offB[l];
            g.act = A[l].p;
            g.epi = EPI_BIAS_SILU;
            gemm(g, 1);
            in = A[l].p;
            in_dim = kMlpH;
        }
        GemmArgs g = base();
        g.A = A[3].p; g.sai = kMlpH; g.sak = 1;
        g.B = P + kOffW5; g.sbk = 1; g.sbj = 0;
        g.C = y.p; g.ldc = 1;
        g.M = n; g.N = 1; g.K = kMlpH;
        g.bias = P + kOffB5;
        g.epi = EPI_BIAS;
        gemm(g, 1);
    }

    // gradient of the weighted squared error of the last forward() into grad
    void backward(int n, const float* Xin, const float* tin, const float* win, float inv_wsum)
    {
        const float* P = params.p;
        float* G = grad.p;
        hipLaunchKernelGGL(loss_kernel, dim3((n + 255) / 256), dim3(256), 0, st, (const float*)y.p, tin, win, n,
                           inv_wsum, dy.p, loss.p);
        HIP_CHECK(hipGetLastError());
        {   // dw5 = A4^T dy, db5 = sum dy
            GemmArgs g = base();
            g.A = dy.p; g.sai = 0; g.sak = 1;
            g.B = A[3].p; g.sbk = kMlpH; g.sbj = 1;
            g.M = 1; g.N = kMlpH; g.K = n;
            batch_reduce(g, G + kOffW5);
            GemmArgs b = base();
            b.A = dy.p; b.sai = 0; b.sak = 1;
            b.B = one.p; b.sbk = 0; b.sbj = 0;
            b.M = 1; b.N = 1; b.K = n;
            batch_reduce(b, G + kOffB5);
        }
        {   // dZ4 = (dy w5^T) * silu'(Z4)
            GemmArgs g = base();
            g.A = dy.p; g.sai = 1; g.sak = 0;
            g.B = P + kOffW5; g.sbk = 0; g.sbj = 1;
            g.C = dZa.p; g.ldc = kMlpH;
            g.M = n; g.N = kMlpH; g.K = 1;
            g.gate = Z[3].p;
            g.epi = EPI_GATE;
            gemm(g, 1);
        }
        const size_t offW[4] = {kOffW1, kOffW2, kOffW3, kOffW4};
        const size_t offB[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
        float* dZ = dZa.p;
        float* dZn = dZb.p;
        for (int l = 3; l >= 0; --l) {
            const float* prev = l > 0 ? A[l - 1].p : Xin;
            const int in_dim = l > 0 ? kMlpH : kMlpIn;
            GemmArgs g = base();   // dW_l[o][i] = sum_r dZ[r][o] prev[r][i]
            g.A = dZ; g.sai = 1; g.sak = kMlpH;
            g.B = prev; g.sbk = in_dim; g.sbj = 1;
            g.M = kMlpH; g.N = in_dim; g.K = n;
            batch_reduce(g, G + offW[l]);
            GemmArgs b = base();   // db_l = column sums of dZ
            b.A = one.p; b.sai = 0; b.sak = 0;
            b.B = dZ; b.sbk = kMlpH; b.sbj = 1;
            b.M = 1; b.N = kMlpH; b.K = n;
            batch_reduce(b, G + offB[l]);
            if (l > 0) {       // dZ_{l-1} = (dZ W_l) * silu'(Z_{l-1})
                GemmArgs d = base();
                d.A = dZ; d.sai = kMlpH; d.sak = 1;
                d.B = P + offW[l]; d.sbk = kMlpH; d.sbj = 1;
                d.C = dZn; d.ldc = kMlpH;
                d.M = n; d.N = kMlpH; d.K = kMlpH;
                d.gate = Z[l - 1].p;
                d.epi = EPI_GATE;
                gemm(d, 1);
                std::swap(dZ, dZn);
            }
        }
    }
};

void check_rows(const float* inputs, uint32_t n)
{
    if (n == 0) throw HaseError(HASE_E_INVALID, "n must be positive");
    if (!inputs) throw HaseError(HASE_E_INVALID, "inputs must not be NULL");
    if ((uint64_t)n > (1ull << 28)) throw HaseError(HASE_E_INVALID, "too many rows");
}

void use_device(int32_t device)
{
    int count = 0;
    HIP_CHECK(hipGetDeviceCount(&count));
    if (count <= 0) throw HaseError(HASE_E_DEVICE, "no HIP device is visible");
    if (device < 0 || device >= count) throw HaseError(HASE_E_INVALID, "device ordinal out of range");
    HIP_CHECK(hipSetDevice(device));
}

// ------------------------------------------------------------ BF16 inference
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kInferBlock = 1024;
constexpr int kInferWaves = kInferBlock / 64;

__device__ __forceinline__ float silu_fast(float z)
{
    return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

// Neuron held by accumulator v of output block mb in a lane of half h
// (v_mfma_f32_32x32x16 D layout: row (v%4) + 4h + 8(v/4) of the block).
__device__ __forceinline__ int acc_neuron(int mb, int v, int h) { return 32 * mb + (v & 3) + 4 * h + 8 * (v >> 2); }

struct InferParams {
    const uint4* frags;     // [kFrags * 64]
    const float* vec;       // [kMlpVec]
    uint32_t rows;
    float* out;             // [rows]
    // explicit rows (already normalised)
    const float* x;         // [rows * kMlpIn]
    // cluster rows: row = (slot * R + k) * C + c
    const float2* pts;
    const float4* clusters; // (x, y, level, beta)
    uint32_t np, sample_base, R, C, layers;
    float h, log_sigma;
    float shift[kMlpIn], scale[kMlpIn];
};

template <bool kClusters>
__device__ __forceinline__ bf16x8 row_inputs(const InferParams& p, uint32_t row, int h)
{
    float f[kMlpIn];
    if (row >= p.rows) {
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) f[i] = 0.0f;
    } else if (kClusters) {
        const uint32_t c = row % p.C, sk = row / p.C;
        const uint32_t k = sk % p.R, slot = sk / p.R;
        const uint32_t sample = p.sample_base + slot;
        const uint32_t pt = sample % p.np, lv = sample / p.np;
        const float2 xy = p.pts[pt];
        const int plane = image_plane((int)lv, (int)p.layers, (k + 1u) >> 1, k & 1u);
        const float4 cl = p.clusters[c];
        proposal_inputs(xy.x, xy.y, (float)plane * p.h, cl.x, cl.y, cl.z, cl.w, p.h, p.log_sigma, f);
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) f[i] = (f[i] - p.shift[i]) * p.scale[i];
    } else {
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) f[i] = p.x[(size_t)row * kMlpIn + i];
    }
    bf16x8 b;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int i = 8 * h + m;
        b[m] = (__bf16)(i < kMlpIn ? f[i] : 0.0f);
    }
    return b;
}

template <bool kClusters>
__global__ void __launch_bounds__(kInferBlock) mlp_infer_kernel(InferParams p)
{
    __shared__ uint4 sfrag[kFrags * 64];
    __shared__ float svec[kMlpVec];
    for (int i = threadIdx.x; i < kFrags * 64; i += kInferBlock) sfrag[i] = p.frags[i];
    for (int i = threadIdx.x; i < kMlpVec; i += kInferBlock) svec[i] = p.vec[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int h = lane >> 5, n = lane & 31;
    const uint32_t tiles = (p.rows + 31u) / 32u;
    const float* w5 = svec + 4 * kMlpH;
    const float b5 = svec[5 * kMlpH];
    for (uint32_t g = blockIdx.x * kInferWaves + wave; g < tiles; g += gridDim.x * kInferWaves) {
        const uint32_t row = g * 32u + (uint32_t)n;
        f32x16 acc[4];
        {
            const bf16x8 in = row_inputs<kClusters>(p, row, h);
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                f32x16 c;
#pragma unroll
                for (int v = 0; v < 16; ++v) c[v] = svec[acc_neuron(mb, v, h)];
                const bf16x8 a = __builtin_bit_cast(bf16x8, sfrag[mb * 64 + lane]);
                acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, in, c, 0, 0, 0);
            }
        }
        for (int layer = 1; layer < 4; ++layer) {
            bf16x8 bt[8];
#pragma unroll
            for (int kb = 0; kb < 8; ++kb)
#pragma unroll
                for (int m = 0; m < 8; ++m) bt[kb][m] = (__bf16)silu_fast(acc[kb >> 1][8 * (kb & 1) + m]);
            const float* bias = svec + layer * kMlpH;
            const uint4* fr = sfrag + (kFragsL1 + (layer - 1) * kFragsHidden) * 64;
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                f32x16 c;
#pragma unroll
                for (int v = 0; v < 16; ++v) c[v] = bias[acc_neuron(mb, v, h)];
#pragma unroll
                for (int kb = 0; kb < 8; ++kb) {
                    const bf16x8 a = __builtin_bit_cast(bf16x8, fr[(mb * 8 + kb) * 64 + lane]);
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bt[kb], c, 0, 0, 0);
                }
                acc[mb] = c;
                // keep the fragment loads of later output blocks from being hoisted here
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        float s = 0.0f;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int v = 0; v < 16; ++v) s = fmaf(w5[acc_neuron(mb, v, h)], silu_fast(acc[mb][v]), s);
        s += __shfl_xor(s, 32, 64);
        if (h == 0 && row < p.rows) p.out[row] = s + b5;
    }
}

uint16_t to_bf16(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return 0x7FC0;
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

int infer_grid(int device)
{
    int cus = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    return std::max(1, cus);
}

}  // namespace

void pack_mlp_bf16(const float* params, uint4* frags, float* vec)
{
    std::vector<uint16_t> bf((size_t)kFrags * 64 * 8);
    for (int layer = 0; layer < 4; ++layer) {
        const int nkb = layer == 0 ? 1 : 8;
        const float* W = params + (layer == 0 ? kOffW1 : layer == 1 ? kOffW2 : layer == 2 ? kOffW3 : kOffW4);
        const int in_dim = layer == 0 ? kMlpIn : kMlpH;
        for (int mb = 0; mb < 4; ++mb)
            for (int kb = 0; kb < nkb; ++kb) {
                const int f = layer == 0 ? mb : kFragsL1 + (layer - 1) * kFragsHidden + mb * 8 + kb;
                for (int lane = 0; lane < 64; ++lane) {
                    const int row = 32 * mb + (lane & 31), hh = lane >> 5;
                    for (int m = 0; m < 8; ++m) {
                        const int s = 16 * kb + 8 * hh + m;
                        // hidden layers read the previous layer's accumulators, whose K order
                        // is the slot index with bits 2 and 3 swapped
                        const int col = layer == 0 ? s : ((s & ~12) | ((s & 4) << 1) | ((s & 8) >> 1));
                        const float val = col < in_dim ? W[(size_t)row * in_dim + col] : 0.0f;
                        bf[((size_t)f * 64 + lane) * 8 + m] = to_bf16(val);
                    }
                }
            }
    }
    std::memcpy(frags, bf.data(), bf.size() * sizeof(uint16_t));
    const size_t offB[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
    for (int l = 0; l < 4; ++l)
        for (int i = 0; i < kMlpH; ++i) vec[l * kMlpH + i] = params[offB[l] + i];
    for (int i = 0; i < kMlpH; ++i) vec[4 * kMlpH + i] = params[kOffW5 + i];
    vec[5 * kMlpH] = params[kOffB5];
}

void build_clusters(hase_context* ctx)
{
    const uint32_t T = ctx->T, layers = ctx->L - 1;
    const uint32_t ct = ctx->cluster_t, cl = ctx->cluster_l;
    ctx->TG = (T + ct - 1) / ct;
    ctx->LG = (layers + cl - 1) / cl;
    ctx->C = ctx->TG * ctx->LG;
    ctx->clusters_host.assign(ctx->C, make_float4(0, 0, 0, 0));
    for (uint32_t lg = 0; lg < ctx->LG; ++lg)
        for (uint32_t tg = 0; tg < ctx->TG; ++tg) {
            double a = 0.0, x = 0.0, y = 0.0, lev = 0.0, bvs = 0.0, nl = 0.0;
            const uint32_t t1 = std::min(T, (tg + 1) * ct), l1 = std::min(layers, (lg + 1) * cl);
            for (uint32_t t = tg * ct; t < t1; ++t) {
                const double ar = ctx->tri_area[t];
                a += ar;
                x += ar * ctx->centroid_host[t].x;
                y += ar * ctx->centroid_host[t].y;
            }
            for (uint32_t l = lg * cl; l < l1; ++l) {
                lev += l;
                nl += 1.0;
                for (uint32_t t = tg * ct; t < t1; ++t)
                    bvs += ctx->tri_area[t] * ctx->beta_host[ctx->dev2user[t] + (size_t)l * T];
            }
            ctx->clusters_host[lg * ctx->TG + tg] =
                make_float4((float)(x / a), (float)(y / a), (float)(lev / nl), (float)(bvs / (a * nl)));
        }
    ctx->clusters.upload(ctx->clusters_host.data(), ctx->C, ctx->stream);
    std::vector<uint32_t> ctri(T), clev(layers);
    for (uint32_t t = 0; t < T; ++t) ctri[t] = t / ct;
    for (uint32_t l = 0; l < layers; ++l) clev[l] = (l / cl) * ctx->TG;
    ctx->cl_tri.upload(ctri.data(), T, ctx->stream);
    ctx->cl_lev.upload(clev.data(), layers, ctx->stream);
    ctx->pstats.clusters = ctx->C;
}

void launch_proposal_inference(hase_context* ctx, uint32_t sample_base, uint32_t slots, float* logits)
{
    const uint64_t rows = (uint64_t)slots * slices(ctx) * ctx->C;
    if (rows >= (1ull << 32)) throw HaseError(HASE_E_INVALID, "too many proposal rows in one batch");
    InferParams p{};
    p.frags = ctx->mlp_frags.p;
    p.vec = ctx->mlp_vec.p;
    p.rows = (uint32_t)rows;
    p.out = logits;
    p.pts = ctx->pts.p;
    p.clusters = ctx->clusters.p;
    p.np = ctx->np;
    p.sample_base = sample_base;
    p.R = slices(ctx);
    p.C = ctx->C;
    p.layers = ctx->L - 1;
    p.h = (float)ctx->h;
    p.log_sigma = ctx->log_sigma;
    for (int i = 0; i < kMlpIn; ++i) {
        p.shift[i] = ctx->in_shift[i];
        p.scale[i] = ctx->in_scale[i];
    }
    const uint32_t tiles = (uint32_t)((rows + 31) / 32);
    const uint32_t grid = std::min<uint32_t>((tiles + kInferWaves - 1) / kInferWaves,
                                             (uint32_t)infer_grid(ctx->device));
    hipLaunchKernelGGL(mlp_infer_kernel<true>, dim3(std::max(1u, grid)), dim3(kInferBlock), 0, ctx->stream, p);
    HIP_CHECK(hipGetLastError());
    ctx->pstats.evaluations += rows;
}

}  // namespace hase

using namespace hase;

extern "C" {

uint32_t hase_mlp_param_count(void) { return (uint32_t)kMlpParams; }

int hase_mlp_forward(const float* params, uint32_t n, const float* inputs, float* outputs, int32_t device)
{
    return guarded([&] {
        check_rows(inputs, n);
        if (!params || !outputs) throw HaseError(HASE_E_INVALID, "params and outputs must not be NULL");
        use_device(device);
        const int chunk = (int)std::min<uint32_t>(n, 65536u);
        Trainer tr(chunk);
        HIP_CHECK(hipMemcpyAsync(tr.params.p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, tr.st));
        for (uint32_t r0 = 0; r0 < n; r0 += (uint32_t)chunk) {
            const int nb = (int)std::min<uint32_t>((uint32_t)chunk, n - r0);
            HIP_CHECK(hipMemcpyAsync(tr.X.p, inputs + (size_t)r0 * kMlpIn, (size_t)nb * kMlpIn * sizeof(float),
                                     hipMemcpyHostToDevice, tr.st));
            tr.forward(nb, tr.X.p);
            HIP_CHECK(hipMemcpyAsync(outputs + r0, tr.y.p, nb * sizeof(float), hipMemcpyDeviceToHost, tr.st));
        }
        HIP_CHECK(hipStreamSynchronize(tr.st));
    });
}

int hase_mlp_loss_gradient(const float* params, uint32_t n, const float* inputs, const float* targets,
                           const float* weights, double* loss, float* gradient, int32_t device)
{
    return guarded([&] {
        check_rows(inputs, n);
        if (!params || !targets || !loss || !gradient)
            throw HaseError(HASE_E_INVALID, "params, targets, loss and gradient must not be NULL");
        if (n > (1u << 20)) throw HaseError(HASE_E_INVALID, "at most 2^20 rows per gradient call");
        double wsum = 0.0;
        for (uint32_t i = 0; i < n; ++i) wsum += weights ? (double)weights[i] : 1.0;
        if (!(wsum > 0.0)) throw HaseError(HASE_E_INVALID, "weights must have a positive sum");
        use_device(device);
        Trainer tr((int)n);
        HIP_CHECK(hipMemcpyAsync(tr.params.p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, tr.st));
        HIP_CHECK(hipMemcpyAsync(tr.X.p, inputs, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, tr.st));
        HIP_CHECK(hipMemcpyAsync(tr.t.p, targets, n * sizeof(float), hipMemcpyHostToDevice, tr.st));
        if (weights) HIP_CHECK(hipMemcpyAsync(tr.w.p, weights, n * sizeof(float), hipMemcpyHostToDevice, tr.st));
        HIP_CHECK(hipMemsetAsync(tr.loss.p, 0, sizeof(double), tr.st));
        tr.forward((int)n, tr.X.p);
        tr.backward((int)n, tr.X.p, tr.t.p, weights ? tr.w.p : nullptr, (float)(1.0 / wsum));
        HIP_CHECK(hipMemcpyAsync(gradient, tr.grad.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, tr.st));
        HIP_CHECK(hipMemcpyAsync(loss, tr.loss.p, sizeof(double), hipMemcpyDeviceToHost, tr.st));
        HIP_CHECK(hipStreamSynchronize(tr.st));
    });
}

int hase_mlp_train(float* params, uint32_t n, const float* inputs, const float* targets, const float* weights,
                   const hase_train_t* cfg, float* epoch_loss, int32_t device)
{
    return guarded([&] {
        check_rows(inputs, n);
        if (!params || !targets || !cfg) throw HaseError(HASE_E_INVALID, "params, targets and config must not be NULL");
        if (cfg->batch == 0 || cfg->epochs == 0) throw HaseError(HASE_E_INVALID, "batch and epochs must be positive");
        if (!(cfg->learning_rate > 0.0f)) throw HaseError(HASE_E_INVALID, "learning_rate must be positive");
        use_device(device);
        const int B = (int)std::min<uint32_t>(cfg->batch, n);
        Trainer tr(B);
        tr.m.alloc(kMlpParams);
        tr.v.alloc(kMlpParams);
        HIP_CHECK(hipMemsetAsync(tr.m.p, 0, kMlpParams * sizeof(float), tr.st));
        HIP_CHECK(hipMemsetAsync(tr.v.p, 0, kMlpParams * sizeof(float), tr.st));
        HIP_CHECK(hipMemcpyAsync(tr.params.p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, tr.st));
        DevBuf<float> X, t, w;
        DevBuf<uint32_t> perm;
        X.upload(inputs, (size_t)n * kMlpIn, tr.st);
        t.upload(targets, n, tr.st);
        if (weights) w.upload(weights, n, tr.st);
        perm.alloc(n);
        std::vector<uint32_t> order(n);
        std::iota(order.begin(), order.end(), 0u);
        std::mt19937_64 rng(cfg->seed);
        uint64_t step = 0;
        const uint32_t batches = n / (uint32_t)B;   // whole batches; the tail rows rotate in by shuffling
        for (uint32_t ep = 0; ep < cfg->epochs; ++ep) {
            std::shuffle(order.begin(), order.end(), rng);
            HIP_CHECK(hipMemcpyAsync(perm.p, order.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, tr.st));
            HIP_CHECK(hipMemsetAsync(tr.loss.p, 0, sizeof(double), tr.st));
            for (uint32_t bi = 0; bi < std::max(1u, batches); ++bi) {
                const int b0 = (int)(bi * (uint32_t)B);
                double wsum = 0.0;
                for (int r = 0; r < B; ++r) wsum += weights ? (double)weights[order[b0 + r]] : 1.0;
                if (!(wsum > 0.0)) continue;
                hipLaunchKernelGGL(gather_kernel, dim3((B + 255) / 256), dim3(256), 0, tr.st, (const float*)X.p,
                                   (const float*)t.p, weights ? (const float*)w.p : nullptr, (const uint32_t*)perm.p,
                                   b0, B, tr.X.p, tr.t.p, tr.w.p);
                HIP_CHECK(hipGetLastError());
                tr.forward(B, tr.X.p);
                tr.backward(B, tr.X.p, tr.t.p, tr.w.p, (float)(1.0 / wsum));
                ++step;
                const float c1 = 1.0f - std::pow(cfg->beta1, (float)step);
                const float c2 = 1.0f - std::pow(cfg->beta2, (float)step);
                hipLaunchKernelGGL(adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, tr.st, tr.params.p,
                                   (const float*)tr.grad.p, tr.m.p, tr.v.p, (int)kMlpParams, cfg->learning_rate,
                                   cfg->beta1, cfg->beta2, cfg->epsilon, cfg->weight_decay, c1, c2);
                HIP_CHECK(hipGetLastError());
            }
            double l = 0.0;
            HIP_CHECK(hipMemcpyAsync(&l, tr.loss.p, sizeof(double), hipMemcpyDeviceToHost, tr.st));
            HIP_CHECK(hipStreamSynchronize(tr.st));
            if (epoch_loss) epoch_loss[ep] = (float)(l / std::max(1u, batches));
        }
        HIP_CHECK(hipMemcpyAsync(params, tr.params.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, tr.st));
        HIP_CHECK(hipStreamSynchronize(tr.st));
    });
}

int hase_mlp_forward_bf16(const float* params, uint32_t n, const float* inputs, float* outputs, int32_t device)
{
    return guarded([&] {
        check_rows(inputs, n);
        if (!params || !outputs) throw HaseError(HASE_E_INVALID, "params and outputs must not be NULL");
        use_device(device);
        hipStream_t st;
        HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
        } guard{st};
        std::vector<uint4> frags((size_t)kFrags * 64);
        std::vector<float> vec(kMlpVec);
        pack_mlp_bf16(params, frags.data(), vec.data());
        DevBuf<uint4> df;
        DevBuf<float> dv, dx, dy;
        df.upload(frags.data(), frags.size(), st);
        dv.upload(vec.data(), vec.size(), st);
        dx.upload(inputs, (size_t)n * kMlpIn, st);
        dy.alloc(n);
        InferParams p{};
        p.frags = df.p;
        p.vec = dv.p;
        p.rows = n;
        p.out = dy.p;
        p.x = dx.p;
        const uint32_t tiles = (n + 31u) / 32u;
        const uint32_t grid = std::min<uint32_t>((tiles + kInferWaves - 1) / kInferWaves, (uint32_t)infer_grid(device));
        hipLaunchKernelGGL(mlp_infer_kernel<false>, dim3(std::max(1u, grid)), dim3(kInferBlock), 0, st, p);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(outputs, dy.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
    });
}

}  // extern "C"
