// Per-k MLP of the learned coarse propagator: three dense layers
// din -> dh -> dh -> dout with tanh after the first two, BF16 weights and
// activations, FP32 accumulation on the matrix cores, FP32 master weights
// trained with Adam.  No counterpart in the reference.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../engine/engine_internal.hpp"

namespace edus {

// round-to-nearest-even FP32 -> BF16 bits (the top 16 bits of the FP32
// pattern after adding half an ulp of the kept part; NaN kept quiet)
__host__ __device__ inline uint16_t f2bf(float x)
{
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return uint16_t((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b)
{
    const uint32_t u = uint32_t(b) << 16;
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}

class Mlp {
  public:
    Mlp(int din, int dh, int dout, uint64_t seed, hipStream_t s);
    ~Mlp();
    Mlp(const Mlp&) = delete;
    Mlp& operator=(const Mlp&) = delete;

    int din, dh, dout;
    hipStream_t s;
    // parameters, FP32 master, one contiguous block: W1 b1 W2 b2 W3 b3
    // (W_l row-major [in][out])
    float* P = nullptr;
    float* G = nullptr;       // gradients, same layout
    float* Am = nullptr;      // Adam first moments
    float* Av = nullptr;      // Adam second moments
    size_t nparam = 0;
    long adam_t = 0;
    size_t off_W[3], off_b[3];
    // BF16 copies: W_l [in][out] and its transpose [out][in]
    uint16_t* Wb[3] = {nullptr, nullptr, nullptr};
    uint16_t* WbT[3] = {nullptr, nullptr, nullptr};

    // forward for M rows of X (BF16 [M][din]) into Y (FP32 [M][dout]); the
    // activations are kept for a backward pass when keep is set
    void forward(const uint16_t* X, int M, float* Y, bool keep);
    // gradients of L = mean over rows and outputs of (Y - T)^2 (after forward
    // with keep) into G; returns nothing, loss via loss_dev (one float) if set
    void backward(const uint16_t* X, const float* T, int M, float* loss_dev);
    void adam_step(float lr, float beta1 = 0.9f, float beta2 = 0.999f, float eps = 1e-8f);
    void refresh_bf16();  // BF16 copies from the FP32 master
    void set_params(const float* host);
    void get_params(float* host) const;
    void get_grads(float* host) const;

  private:
    std::vector<void*> owned;
    int cap_rows = 0;
    uint16_t *H1 = nullptr, *H2 = nullptr, *dYb = nullptr, *dZ2 = nullptr, *dZ1 = nullptr;
    float* Ybuf = nullptr;
    float* partial = nullptr;  // split-K partial sums
    size_t partial_cap = 0;
    void* alloc(size_t bytes);
    void ensure_rows(int M);
    void ensure_partial(size_t floats);
    void weight_grad(const uint16_t* A, int lda, const uint16_t* D, int ldd, int M, int K, int N, float* dW, float* db);
};

// GEMM launchers (BF16 operands, FP32 accumulation):
//   C[m][n] = sum_k A(m, k) B(k, n), A(m, k) = A[m * sam + k * sak], B(k, n) = B[k * sbk + n * sbn]
// epilogue: + bias[n] (if bias), tanh (if act), then FP32 out (ldc) and/or
// BF16 out (ldcb); mul (BF16 [M][N], if set) multiplies by (1 - mul^2) before
// the BF16 store: the tanh derivative for the backward pass.
void launch_gemm_bf16(const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, int M, int N,
                      int K, const float* bias, int act, float* C, int ldc, uint16_t* Cb, int ldcb,
                      const uint16_t* mul, hipStream_t s);

// owner of an Mlp and its stream (the C ABI handle), plus scratch buffers
struct MlpHandle {
    hipStream_t s = nullptr;
    Mlp* m = nullptr;
    std::vector<void*> bufs;
    ~MlpHandle()
    {
        if (s) (void)hipStreamSynchronize(s);
        delete m;
        for (void* p : bufs) (void)hipFree(p);
        if (s) (void)hipStreamDestroy(s);
    }
    template <class T>
    T* dev(size_t n)
    {
        void* p = nullptr;
        EDUS_HIP_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        bufs.push_back(p);
        return static_cast<T*>(p);
    }
};

}  // namespace edus

struct edus_mlp {
    edus::MlpHandle h;
};
