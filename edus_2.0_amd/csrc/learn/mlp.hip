// Per-k MLP kernels: BF16 GEMMs on the matrix cores (v_mfma_f32_16x16x16_bf16,
// FP32 accumulation) with fused bias / tanh / tanh-derivative epilogues, a
// split-K weight-gradient GEMM with a deterministic reduction, the MSE loss
// and Adam on FP32 master weights.
//
// Fragment layout of the 16x16x16 BF16 instruction (lane l, q = l / 16,
// i = l % 16): A row i, columns 4q .. 4q+3; B rows 4q .. 4q+3, column i; the
// FP32 result row 4q + v, column i in slot v.
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>

#include "../engine/engine_internal.hpp"
#include "mlp.hpp"

#define HIP_CHECK EDUS_HIP_CHECK

namespace edus {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;   // rows and columns of C per work-group
constexpr int kKStep = 16;  // K per MFMA

struct GemmArgs {
    const uint16_t* A;
    long sam, sak;
    const uint16_t* B;
    long sbk, sbn;
    int M, N, K;
    int kchunk;        // K range per blockIdx.z (split-K)
    int ones_row;      // row of A that reads 1.0 (bias gradient), -1: none
    const float* bias;
    int act;           // 1: tanh
    float* C;          // FP32 out (split-K: partial [z][M][N])
    int ldc;
    uint16_t* Cb;      // BF16 out
    int ldcb;
    const uint16_t* mul;  // BF16 [M][ldm]: multiply by (1 - mul^2)
    int ldm;
};

__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g)
{
    __shared__ uint16_t sA[kTile][kKStep + 4];
    __shared__ uint16_t sB[kKStep][kTile + 4];
    const int m0 = blockIdx.x * kTile, n0 = blockIdx.y * kTile;
    const int kb = blockIdx.z * g.kchunk;
    const int ke = min(g.K, kb + g.kchunk);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = lane & 15, q = lane >> 4;
    flt4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = flt4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = kb; k0 < ke; k0 += kKStep) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = threadIdx.x + e * 256;  // 64 x 16 tile of A
            const int r = idx >> 4, c = idx & 15;
            const int m = m0 + r, k = k0 + c;
            uint16_t v = 0;
            if (m < g.M && k < ke) v = (m == g.ones_row) ? uint16_t(0x3f80) : g.A[m * g.sam + k * g.sak];
            sA[r][c] = v;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = threadIdx.x + e * 256;  // 16 x 64 tile of B
            const int r = idx >> 6, c = idx & 63;
            const int k = k0 + r, n = n0 + c;
            sB[r][c] = (k < ke && n < g.N) ? g.B[k * g.sbk + n * g.sbn] : uint16_t(0);
        }
        __syncthreads();
        bf16x4 a;
#pragma unroll
        for (int e = 0; e < 4; ++e) a[e] = short(sA[16 * wave + i][4 * q + e]);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf16x4 b;
#pragma unroll
            for (int e = 0; e < 4; ++e) b[e] = short(sB[4 * q + e][16 * t + i]);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
        __syncthreads();
    }

    float* C = g.C ? g.C + size_t(blockIdx.z) * g.M * g.N : nullptr;  // split-K partials
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = m0 + 16 * wave + 4 * q + v, n = n0 + 16 * t + i;
            if (m >= g.M || n >= g.N) continue;
            float x = acc[t][v];
            if (g.bias) x += g.bias[n];
            if (g.act) x = tanhf(x);
            if (g.mul) {
                const float h = bf2f(g.mul[size_t(m) * g.ldm + n]);
                x *= 1.0f - h * h;
            }
            if (C) C[size_t(m) * g.ldc + n] = x;
            if (g.Cb) g.Cb[size_t(m) * g.ldcb + n] = f2bf(x);
        }
}

// sum of split-K partials [nz][rows][N]: rows < nrw -> dW [rows][N], row nrw -> db
__global__ void reduce_partials_kernel(const float* __restrict__ part, int nz, int nrw, int N, float* __restrict__ dW,
                                       float* __restrict__ db)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = (nrw + 1) * N;
    if (idx >= total) return;
    float s = 0.f;
    for (int z = 0; z < nz; ++z) s += part[size_t(z) * total + idx];
    if (idx < nrw * N)
        dW[idx] = s;
    else
        db[idx - nrw * N] = s;
}

// dY = 2 (Y - T) / (M dout) as BF16, and per-block partial sums of (Y - T)^2
__global__ void mse_kernel(const float* __restrict__ Y, const float* __restrict__ T, long n, float scale,
                           uint16_t* __restrict__ dYb, float* __restrict__ part)
{
    __shared__ float sh[256];
    float acc = 0.f;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += long(gridDim.x) * 256) {
        const float d = Y[i] - T[i];
        acc += d * d;
        dYb[i] = f2bf(2.0f * d * scale);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void sum_kernel(const float* __restrict__ part, int n, float scale, float* __restrict__ out)
{
    __shared__ float sh[256];
    float acc = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) acc += part[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0] * scale;
}

__global__ void adam_kernel(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ m,
                            float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// W [rows][cols] FP32 -> BF16 W and its transpose
__global__ void to_bf16_kernel(const float* __restrict__ W, int rows, int cols, uint16_t* __restrict__ Wb,
                               uint16_t* __restrict__ WbT)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= rows * cols) return;
    const int r = idx / cols, c = idx - r * cols;
    const uint16_t b = f2bf(W[idx]);
    Wb[idx] = b;
    WbT[size_t(c) * rows + r] = b;
}

int cdiv(long a, long b) { return int((a + b - 1) / b); }

}  // namespace

void launch_gemm_bf16(const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, int M, int N,
                      int K, const float* bias, int act, float* C, int ldc, uint16_t* Cb, int ldcb,
                      const uint16_t* mul, hipStream_t s)
{
    GemmArgs g{};
    g.A = A;
    g.sam = sam;
    g.sak = sak;
    g.B = B;
    g.sbk = sbk;
    g.sbn = sbn;
    g.M = M;
    g.N = N;
    g.K = K;
    g.kchunk = K;
    g.ones_row = -1;
    g.bias = bias;
    g.act = act;
    g.C = C;
    g.ldc = ldc;
    g.Cb = Cb;
    g.ldcb = ldcb;
    g.mul = mul;
    g.ldm = N;
    hipLaunchKernelGGL(gemm_bf16_kernel, dim3(cdiv(M, kTile), cdiv(N, kTile), 1), dim3(256), 0, s, g);
}

void* Mlp::alloc(size_t bytes)
{
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, bytes));
    owned.push_back(p);
    HIP_CHECK(hipMemsetAsync(p, 0, bytes, s));
    return p;
}

Mlp::Mlp(int din_, int dh_, int dout_, uint64_t seed, hipStream_t st) : din(din_), dh(dh_), dout(dout_), s(st)
{
    if (din < 1 || dh < 1 || dout < 1 || din > 4096 || dh > 4096 || dout > 4096)
        throw EngineError(EDUS_ERR_ARG, "MLP sizes must be in 1..4096");
    const int in[3] = {din, dh, dh}, out[3] = {dh, dh, dout};
    size_t o = 0;
    for (int l = 0; l < 3; ++l) {
        off_W[l] = o;
        o += size_t(in[l]) * out[l];
        off_b[l] = o;
        o += out[l];
    }
    nparam = o;
    P = static_cast<float*>(alloc(nparam * sizeof(float)));
    G = static_cast<float*>(alloc(nparam * sizeof(float)));
    Am = static_cast<float*>(alloc(nparam * sizeof(float)));
    Av = static_cast<float*>(alloc(nparam * sizeof(float)));
    for (int l = 0; l < 3; ++l) {
        Wb[l] = static_cast<uint16_t*>(alloc(size_t(in[l]) * out[l] * sizeof(uint16_t)));
        WbT[l] = static_cast<uint16_t*>(alloc(size_t(in[l]) * out[l] * sizeof(uint16_t)));
    }
    // Glorot-uniform weights, zero biases
    std::vector<float> h(nparam, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < 3; ++l) {
        const float a = std::sqrt(6.0f / float(in[l] + out[l]));
        std::uniform_real_distribution<float> u(-a, a);
        for (size_t k = 0; k < size_t(in[l]) * out[l]; ++k) h[off_W[l] + k] = u(rng);
    }
    set_params(h.data());
}

Mlp::~Mlp()
{
    (void)hipStreamSynchronize(s);
    for (void* p : owned) (void)hipFree(p);
}

void Mlp::set_params(const float* host)
{
    HIP_CHECK(hipMemcpyAsync(P, host, nparam * sizeof(float), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemsetAsync(Am, 0, nparam * sizeof(float), s));
    HIP_CHECK(hipMemsetAsync(Av, 0, nparam * sizeof(float), s));
    adam_t = 0;
    refresh_bf16();
    HIP_CHECK(hipStreamSynchronize(s));
}

void Mlp::get_params(float* host) const
{
    HIP_CHECK(hipMemcpyAsync(host, P, nparam * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

void Mlp::get_grads(float* host) const
{
    HIP_CHECK(hipMemcpyAsync(host, G, nparam * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

void Mlp::refresh_bf16()
{
    const int in[3] = {din, dh, dh}, out[3] = {dh, dh, dout};
    for (int l = 0; l < 3; ++l)
        hipLaunchKernelGGL(to_bf16_kernel, dim3(cdiv(long(in[l]) * out[l], 256)), dim3(256), 0, s, P + off_W[l],
                           in[l], out[l], Wb[l], WbT[l]);
}

void Mlp::ensure_rows(int M)
{
    if (M <= cap_rows) return;
    H1 = static_cast<uint16_t*>(alloc(size_t(M) * dh * sizeof(uint16_t)));
    H2 = static_cast<uint16_t*>(alloc(size_t(M) * dh * sizeof(uint16_t)));
    dZ1 = static_cast<uint16_t*>(alloc(size_t(M) * dh * sizeof(uint16_t)));
    dZ2 = static_cast<uint16_t*>(alloc(size_t(M) * dh * sizeof(uint16_t)));
    dYb = static_cast<uint16_t*>(alloc(size_t(M) * dout * sizeof(uint16_t)));
    Ybuf = static_cast<float*>(alloc(size_t(M) * dout * sizeof(float)));
    cap_rows = M;
}

void Mlp::ensure_partial(size_t floats)
{
    if (floats <= partial_cap) return;
    partial = static_cast<float*>(alloc(floats * sizeof(float)));
    partial_cap = floats;
}

void Mlp::forward(const uint16_t* X, int M, float* Y, bool keep)
{
    if (M <= 0) return;
    ensure_rows(M);
    (void)keep;  // the activations always land in H1, H2
    launch_gemm_bf16(X, din, 1, Wb[0], dh, 1, M, dh, din, P + off_b[0], 1, nullptr, 0, H1, dh, nullptr, s);
    launch_gemm_bf16(H1, dh, 1, Wb[1], dh, 1, M, dh, dh, P + off_b[1], 1, nullptr, 0, H2, dh, nullptr, s);
    launch_gemm_bf16(H2, dh, 1, Wb[2], dout, 1, M, dout, dh, P + off_b[2], 0, Y ? Y : Ybuf, dout, nullptr, 0,
                     nullptr, s);
}

// dW [K][N] = A^T D over M rows, db [N] = column sums of D; A [M][lda], D [M][ldd]
void Mlp::weight_grad(const uint16_t* A, int lda, const uint16_t* D, int ldd, int M, int K, int N, float* dW,
                      float* db)
{
    const int nz = std::max(1, std::min(64, M / 1024));
    const int kchunk = ((cdiv(M, nz) + kKStep - 1) / kKStep) * kKStep;
    const int nzz = cdiv(M, kchunk);
    ensure_partial(size_t(nzz) * (K + 1) * N);
    GemmArgs g{};
    g.A = A;
    g.sam = 1;
    g.sak = lda;
    g.B = D;
    g.sbk = ldd;
    g.sbn = 1;
    g.M = K + 1;  // the extra row reads ones: the bias gradient
    g.N = N;
    g.K = M;
    g.kchunk = kchunk;
    g.ones_row = K;
    g.C = partial;
    g.ldc = N;
    hipLaunchKernelGGL(gemm_bf16_kernel, dim3(cdiv(K + 1, kTile), cdiv(N, kTile), nzz), dim3(256), 0, s, g);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(cdiv(long(K + 1) * N, 256)), dim3(256), 0, s, partial, nzz, K, N,
                       dW, db);
}

void Mlp::backward(const uint16_t* X, const float* T, int M, float* loss_dev)
{
    if (M <= 0) return;
    ensure_rows(M);
    const long n = long(M) * dout;
    const int nblk = std::min(256, cdiv(n, 256));
    ensure_partial(size_t(nblk));
    hipLaunchKernelGGL(mse_kernel, dim3(nblk), dim3(256), 0, s, Ybuf, T, n, 1.0f / float(n), dYb, partial);
    if (loss_dev) hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, partial, nblk, 1.0f / float(n), loss_dev);
    // layer 3
    weight_grad(H2, dh, dYb, dout, M, dh, dout, G + off_W[2], G + off_b[2]);
    launch_gemm_bf16(dYb, dout, 1, WbT[2], dh, 1, M, dh, dout, nullptr, 0, nullptr, 0, dZ2, dh, H2, s);
    // layer 2
    weight_grad(H1, dh, dZ2, dh, M, dh, dh, G + off_W[1], G + off_b[1]);
    launch_gemm_bf16(dZ2, dh, 1, WbT[1], dh, 1, M, dh, dh, nullptr, 0, nullptr, 0, dZ1, dh, H1, s);
    // layer 1
    weight_grad(X, din, dZ1, dh, M, din, dh, G + off_W[0], G + off_b[0]);
}

void Mlp::adam_step(float lr, float beta1, float beta2, float eps)
{
    ++adam_t;
    const float c1 = 1.0f - std::pow(beta1, float(adam_t));
    const float c2 = 1.0f - std::pow(beta2, float(adam_t));
    hipLaunchKernelGGL(adam_kernel, dim3(cdiv(long(nparam), 256)), dim3(256), 0, s, P, G, Am, Av, long(nparam), lr,
                       beta1, beta2, eps, c1, c2);
    refresh_bf16();
}

}  // namespace edus
