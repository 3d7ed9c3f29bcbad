// C ABI of the per-k MLP (edus_mlp_* in include/edus_gpu.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

#include "../engine/engine_internal.hpp"
#include "mlp.hpp"

#define HIP_CHECK EDUS_HIP_CHECK

namespace edus {

namespace {

// rows idx[r] of X [*][dx] (BF16) and T [*][dt] -> Xb, Tb
__global__ void gather_rows_kernel(const uint16_t* __restrict__ X, const float* __restrict__ T,
                                   const int* __restrict__ idx, int rows, int dx, int dt, uint16_t* __restrict__ Xb,
                                   float* __restrict__ Tb)
{
    const int r = blockIdx.x;
    if (r >= rows) return;
    const long src = idx[r];
    for (int c = threadIdx.x; c < dx; c += blockDim.x) Xb[long(r) * dx + c] = X[src * dx + c];
    for (int c = threadIdx.x; c < dt; c += blockDim.x) Tb[long(r) * dt + c] = T[src * dt + c];
}

}  // namespace

}  // namespace edus


using edus::guarded;

namespace {

std::vector<uint16_t> to_bf16(const float* x, size_t n)
{
    std::vector<uint16_t> b(n);
    for (size_t i = 0; i < n; ++i) b[i] = edus::f2bf(x[i]);
    return b;
}

}  // namespace

extern "C" {

edus_mlp* edus_mlp_create(int din, int dh, int dout, unsigned long long seed)
{
    edus_mlp* out = nullptr;
    const int rc = guarded([&] {
        auto* m = new edus_mlp();
        try {
            HIP_CHECK(hipStreamCreateWithFlags(&m->h.s, hipStreamNonBlocking));
            m->h.m = new edus::Mlp(din, dh, dout, seed, m->h.s);
        } catch (...) {
            delete m;
            throw;
        }
        out = m;
    });
    return rc == EDUS_OK ? out : nullptr;
}

int edus_mlp_destroy(edus_mlp* m)
{
    if (!m) return EDUS_OK;
    const int rc = guarded([&] { delete m; });
    return rc;
}

int edus_mlp_num_params(edus_mlp* m, long* n)
{
    EDUS_REQUIRE(m && n);
    *n = long(m->h.m->nparam);
    return EDUS_OK;
}

int edus_mlp_set_params(edus_mlp* m, const float* params)
{
    EDUS_REQUIRE(m && params);
    return guarded([&] { m->h.m->set_params(params); });
}

int edus_mlp_get_params(edus_mlp* m, float* params)
{
    EDUS_REQUIRE(m && params);
    return guarded([&] { m->h.m->get_params(params); });
}

int edus_mlp_forward(edus_mlp* m, int rows, const float* X, float* Y)
{
    EDUS_REQUIRE(m && X && Y && rows >= 0);
    return guarded([&] {
        edus::Mlp& net = *m->h.m;
        const auto xb = to_bf16(X, size_t(rows) * net.din);
        edus::MlpHandle tmp;  // scratch for this call
        uint16_t* dX = tmp.dev<uint16_t>(xb.size());
        float* dY = tmp.dev<float>(size_t(rows) * net.dout);
        HIP_CHECK(hipMemcpyAsync(dX, xb.data(), xb.size() * sizeof(uint16_t), hipMemcpyHostToDevice, net.s));
        net.forward(dX, rows, dY, false);
        HIP_CHECK(hipMemcpyAsync(Y, dY, size_t(rows) * net.dout * sizeof(float), hipMemcpyDeviceToHost, net.s));
        HIP_CHECK(hipStreamSynchronize(net.s));
    });
}

int edus_mlp_gradients(edus_mlp* m, int rows, const float* X, const float* T, float* grads, float* loss)
{
    EDUS_REQUIRE(m && X && T && grads && rows > 0);
    return guarded([&] {
        edus::Mlp& net = *m->h.m;
        const auto xb = to_bf16(X, size_t(rows) * net.din);
        edus::MlpHandle tmp;
        uint16_t* dX = tmp.dev<uint16_t>(xb.size());
        float* dT = tmp.dev<float>(size_t(rows) * net.dout);
        float* dL = tmp.dev<float>(1);
        HIP_CHECK(hipMemcpyAsync(dX, xb.data(), xb.size() * sizeof(uint16_t), hipMemcpyHostToDevice, net.s));
        HIP_CHECK(hipMemcpyAsync(dT, T, size_t(rows) * net.dout * sizeof(float), hipMemcpyHostToDevice, net.s));
        net.forward(dX, rows, nullptr, true);
        net.backward(dX, dT, rows, dL);
        float l = 0.f;
        HIP_CHECK(hipMemcpyAsync(&l, dL, sizeof(float), hipMemcpyDeviceToHost, net.s));
        net.get_grads(grads);
        if (loss) *loss = l;
    });
}

int edus_mlp_train(edus_mlp* m, int rows, const float* X, const float* T, int epochs, int batch, float lr,
                   unsigned long long seed, float* loss_per_epoch)
{
    EDUS_REQUIRE(m && X && T && rows > 0);
    if (epochs < 0 || batch < 1) {
        edus::set_error("edus_mlp_train: epochs >= 0 and batch >= 1 required");
        return EDUS_ERR_ARG;
    }
    return guarded([&] {
        edus::Mlp& net = *m->h.m;
        batch = std::min(batch, rows);
        const auto xb = to_bf16(X, size_t(rows) * net.din);
        edus::MlpHandle tmp;
        uint16_t* dX = tmp.dev<uint16_t>(xb.size());
        float* dT = tmp.dev<float>(size_t(rows) * net.dout);
        uint16_t* bX = tmp.dev<uint16_t>(size_t(batch) * net.din);
        float* bT = tmp.dev<float>(size_t(batch) * net.dout);
        const int nb = rows / batch;  // full batches per epoch
        float* dL = tmp.dev<float>(size_t(std::max(nb, 1)));
        int* didx = tmp.dev<int>(size_t(rows));
        HIP_CHECK(hipMemcpyAsync(dX, xb.data(), xb.size() * sizeof(uint16_t), hipMemcpyHostToDevice, net.s));
        HIP_CHECK(hipMemcpyAsync(dT, T, size_t(rows) * net.dout * sizeof(float), hipMemcpyHostToDevice, net.s));
        std::vector<int> perm(rows);
        std::iota(perm.begin(), perm.end(), 0);
        std::mt19937_64 rng(seed);
        std::vector<float> losses(std::max(nb, 1));
        for (int ep = 0; ep < epochs; ++ep) {
            std::shuffle(perm.begin(), perm.end(), rng);
            HIP_CHECK(hipMemcpyAsync(didx, perm.data(), rows * sizeof(int), hipMemcpyHostToDevice, net.s));
            for (int b = 0; b < nb; ++b) {
                hipLaunchKernelGGL(edus::gather_rows_kernel, dim3(batch), dim3(64), 0, net.s, dX, dT, didx + size_t(b) * batch,
                                   batch, net.din, net.dout, bX, bT);
                net.forward(bX, batch, nullptr, true);
                net.backward(bX, bT, batch, dL + b);
                net.adam_step(lr);
            }
            HIP_CHECK(hipMemcpyAsync(losses.data(), dL, nb * sizeof(float), hipMemcpyDeviceToHost, net.s));
            HIP_CHECK(hipStreamSynchronize(net.s));  // perm is rewritten next epoch
            if (loss_per_epoch) {
                double acc = 0.0;
                for (int b = 0; b < nb; ++b) acc += losses[b];
                loss_per_epoch[ep] = float(acc / std::max(nb, 1));
            }
        }
        HIP_CHECK(hipGetLastError());
    });
}

}  // extern "C"
