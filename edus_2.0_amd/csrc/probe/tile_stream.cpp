// Tile-stream probe: the HBM access pattern of the commutator kernel without
// its arithmetic.  Every work-group reads a tile of TK consecutive k-points
// from each of the P planes of six arrays and writes two, so every plane is
// touched in runs of TK * 16 bytes.  Reports the attainable bandwidth of this
// pattern for TK = 4 .. 64 (the commutator uses TK = 4 or 8), to compare with
// a plain copy (hbm_copy).  With chunked = 1 the planes are stored k-chunked
// ((p, k) at (k / CH) * P * CH + p * CH + k % CH, CH = 8 grid rows), so a
// tile's planes lie within P * CH * 16 bytes instead of spanning the array:
// the same runs, a far smaller address range per work-group (fewer pages).
//   usage: tile_stream [P] [N] [chunked]   (default 196 planes of 120 x 120, 0)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

__global__ void __launch_bounds__(256) stream6(const double2* __restrict__ a0, const double2* __restrict__ a1,
                                               const double2* __restrict__ a2, const double2* __restrict__ a3,
                                               const double2* __restrict__ a4, const double2* __restrict__ a5,
                                               double2* __restrict__ o0, double2* __restrict__ o1, int P, int Nk,
                                               int TK, int CH)
{
    const int k0 = blockIdx.x * TK;
    for (int e = threadIdx.x; e < P * TK; e += blockDim.x) {
        const int p = e / TK, k = k0 + e % TK;
        if (k >= Nk) continue;
        const long o = CH ? long(k / CH) * P * CH + long(p) * CH + (k % CH) : long(p) * Nk + k;
        const double2 v0 = a0[o], v1 = a1[o], v2 = a2[o], v3 = a3[o], v4 = a4[o], v5 = a5[o];
        o0[o] = make_double2(v0.x + v1.x + v2.x, v0.y + v1.y + v2.y);
        o1[o] = make_double2(v3.x + v4.x + v5.x, v3.y + v4.y + v5.y);
    }
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
    const int P = argc > 1 ? std::atoi(argv[1]) : 196;
    const int N = argc > 2 ? std::atoi(argv[2]) : 120;
    const int chunked = argc > 3 ? std::atoi(argv[3]) : 0;
    const int Nk = N * N;
    const long n = long(P) * Nk;
    const int CH = chunked ? 8 * N : 0;
    if (chunked && Nk % CH != 0) return 2;
    std::vector<double2*> buf(8);
    for (auto& b : buf) {
        CHECK(hipMalloc(&b, n * sizeof(double2)));
        CHECK(hipMemset(b, 0, n * sizeof(double2)));
    }
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    for (int TK : {4, 8, 16, 32, 64}) {
        const int blocks = (Nk + TK - 1) / TK;
        hipLaunchKernelGGL(stream6, dim3(blocks), dim3(256), 0, 0, buf[0], buf[1], buf[2], buf[3], buf[4], buf[5],
                           buf[6], buf[7], P, Nk, TK, CH);  // warm-up
        CHECK(hipEventRecord(e0, 0));
        for (int it = 0; it < 20; ++it)
            hipLaunchKernelGGL(stream6, dim3(blocks), dim3(256), 0, 0, buf[0], buf[1], buf[2], buf[3], buf[4], buf[5],
                               buf[6], buf[7], P, Nk, TK, CH);
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double per = ms / 20 / 1e3;
        std::printf("tile_stream: P %d, %dx%d, %s, TK %2d (%4d-byte runs): %7.1f us per launch, %6.0f GB/s\n", P, N,
                    N, chunked ? "k-chunked" : "plane-major", TK, TK * 16, per * 1e6, 8.0 * n * 16 / per / 1e9);
    }
    for (auto* b : buf) CHECK(hipFree(b));
    return 0;
}
