// HBM copy probe: streams a buffer of known size through a plain copy kernel.
// Used to calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters (a copy of
// B bytes must read B and write B) and to measure achievable HBM bandwidth.
//   usage: hbm_copy [MiB] [iterations]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const double2* __restrict__ in, double2* __restrict__ out, long n)
{
    for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) out[i] = in[i];
}

#define CHECK(x)                                                                      \
    do {                                                                              \
        hipError_t e = (x);                                                           \
        if (e != hipSuccess) {                                                        \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e));               \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

int main(int argc, char** argv)
{
    const long mib = argc > 1 ? std::atol(argv[1]) : 1024;
    const int iters = argc > 2 ? std::atoi(argv[2]) : 20;
    const long bytes = mib << 20;
    const long n = bytes / long(sizeof(double2));
    double2 *a, *b;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 0, bytes));
    CHECK(hipMemset(b, 0, bytes));
    const int blocks = 256 * 8 * 4, threads = 256;
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(threads), 0, 0, a, b, n);
    CHECK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    CHECK(hipEventRecord(e0, 0));
    for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(threads), 0, 0, a, b, n);
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double per = ms / iters / 1e3;
    std::printf("hbm_copy: %ld MiB read + %ld MiB written per launch, %.1f us per launch, %.1f GB/s (read+write)\n",
                mib, mib, per * 1e6, 2.0 * bytes / per / 1e9);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
