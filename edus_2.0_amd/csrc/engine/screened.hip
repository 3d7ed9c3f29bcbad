// Lindhard response of the ScreenedPotential application on the device:
// ScreenedPotential::ResponseFunction (src/Screened_potential/Screened_potential.cpp:17-58)
//
//   X_ab(q) = 1/Nk sum_k sum_{v < nf <= c}
//             U_k(a,v) U_k'(b,c) conj(U_k'(a,c)) conj(U_k(b,v)) / (e_v(k) - e_c(k'))
//
// with k' = k + q on the periodic mesh (MeshGrid::find, MeshGrid.cpp:282-297),
// U_k the eigenvectors in columns (ascending bands) and e the band energies.
// Writing A_v(k)_ab = U_k(a,v) conj(U_k(b,v)) and B_c(k')_ab = conj(U_k'(a,c)) U_k'(b,c),
//
//   X_ab(q) = 1/Nk sum_k sum_c B_c(k+q)_ab sum_v r_vc(k, k+q) A_v(k)_ab,
//   r_vc(k, k') = 1 / (e_v(k) - e_c(k')).
//
// The sum is elementwise in (a, b): no contraction that matrix cores could
// take, and O(Nk^2) pairs (q, k) with O(nb^2 nv nc) FP64 work each, so the
// kernel is bound by the FP64 vector pipes, not by HBM.  Mapping:
//   * a split pass stores the valence columns of every U_k as Uv[k][v][a] and
//     the conduction columns as Uc[k][c][a] (one array each), so the columns a
//     work-group needs are contiguous;
//   * one work-group = one wave of 64 lanes; each lane owns one element pair
//     (a, b) (the upper triangle when that packs the lanes better, the lower
//     triangle then follows from X_ba(q) = conj X_ab(q)), for G q-groups per
//     wave when the pairs fit G times in 64 lanes, or one part of a q when
//     they need W waves;
//   * the work-group's GT = G * TQ q lie on one grid row at consecutive q1,
//     so for every k their k + q are consecutive on one row (two runs where
//     the row wraps): per k the wave copies Uv(k) and the Uc(k + q) runs, with
//     the band energies, straight from memory into LDS (global_load_lds, no
//     registers), double-buffered: step k + 1 is in flight while step k is
//     computed;
//   * the reciprocals r_vc of each (q, k) are formed once per wave and read by
//     all lanes of that q from LDS (one address for all: broadcast); A_v(k) is
//     built once per k in registers and used for all TQ q of a lane;
//   * work-groups of neighbouring q go to the same XCD: their Uc(k + q)
//     windows overlap and stay in that XCD's L2.
// The summation order differs from the reference's (k outermost over pairs
// of bands, r multiplied instead of divided), so results agree to rounding,
// not bit for bit.
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../../../include/edus_gpu.h"
#include "common.hpp"
#include "engine_internal.hpp"
#include "kernels.hpp"

namespace edus {

namespace {

constexpr int kLindStageR = 4;  // reciprocals formed per lane per k

struct LindhardArgs {
    const double2* Uv;  // [Nk][nv][nb]: U_k(a, v) at (k nv + v) nb + a
    const double2* Uc;  // [Nk][nc][nb]: U_k(a, nv + c) at (k nc + c) nb + a
    const double* eps;  // [Nk][nb]
    double2* X;         // planes [nb * nb][Nk]
    int nb, nv, nc, N0, N1, Nk;
    int Pg;             // element pairs per q
    int G;              // q groups per wave (Pg <= 64), else 1
    int W;              // waves per q (Pg > 64), else 1
    int tri;            // pairs are the upper triangle
    int nqb_row;        // q-blocks per grid row
    int nR;             // reciprocals per k (whole wave)
    int nvp;            // nv rounded up to even (reciprocal row length)
    int buf;            // one LDS buffer, in double2
    double inv_nk;
};

// pair index i -> (a, b): row-major full list, or the upper triangle a <= b
__device__ inline void pair_of(int i, int nb, int tri, int& a, int& b)
{
    if (!tri) {
        a = i / nb;
        b = i % nb;
        return;
    }
    a = 0;
    while (i >= nb - a) {
        i -= nb - a;
        ++a;
    }
    b = a + i;
}

// U[k][a][n] -> Uv[k][v][a], Uc[k][c][a]
__global__ void lindhard_split_kernel(const double2* __restrict__ U, double2* __restrict__ Uv,
                                      double2* __restrict__ Uc, int nb, int nv, long total)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int M = nb * nb, nc = nb - nv;
    const long k = i / M;
    const int r = int(i - k * M), al = r / nb, n = r % nb;
    const double2 v = U[i];
    if (n < nv)
        Uv[(k * nv + n) * nb + al] = v;
    else
        Uc[(k * nc + (n - nv)) * nb + al] = v;
}

// copy count 16-byte elements src[0..count) to LDS dst (wave-uniform base)
__device__ inline void copy16(const double2* src, double2* dst, int count, int lane)
{
    for (int i = 0; i < count; i += 64)
        if (i + lane < count) __builtin_amdgcn_global_load_lds(src + i + lane, dst + i, 16, 0, 0);
}

// copy count 4-byte words to LDS
__device__ inline void copy4(const unsigned* src, unsigned* dst, int count, int lane)
{
    for (int i = 0; i < count; i += 64)
        if (i + lane < count) __builtin_amdgcn_global_load_lds(src + i + lane, dst + i, 4, 0, 0);
}

template <int NV, int TQ>
__global__ void __launch_bounds__(64) lindhard_kernel(LindhardArgs a)
{
    extern __shared__ double2 lds[];
    const int nb = a.nb, nc = a.nc, N1 = a.N1, GT = a.G * TQ;
    // buffer layout (double2 units): Uv(k) [NV][nb], Uc runs [GT][nc][nb],
    // energies (doubles) [1 + GT][nb]; the reciprocals [GT][nc][nvp] after both buffers
    const int oUc = NV * nb, oE = oUc + GT * nc * nb;
    double* sR = reinterpret_cast<double*>(lds + 2 * a.buf);

    // XCD-aware order: consecutive logical blocks on one XCD
    long L;
    {
        const long nblk = gridDim.x, q = nblk / 8, r = nblk % 8;
        const long x = blockIdx.x % 8, i = blockIdx.x / 8;
        L = x * q + min(x, r) + i;
    }
    const int qbi = int(L / a.W), wpart = int(L % a.W);
    const int q0 = qbi / a.nqb_row, q1f = (qbi % a.nqb_row) * GT;  // row and first q1 of the block
    const int nvalid = min(GT, N1 - q1f);
    const int lane = threadIdx.x;
    int s, pi;
    if (a.W > 1) {
        s = 0;
        pi = lane + 64 * wpart;
    } else {
        s = lane / a.Pg;
        pi = lane % a.Pg;
    }
    const bool active = s < a.G && pi < a.Pg;
    int al = 0, be = 0;
    if (active) pair_of(pi, nb, a.tri, al, be);

    // reciprocal entries of this lane: energy offsets and destination
    int rv[kLindStageR], rc[kLindStageR], rd[kLindStageR];
#pragma unroll
    for (int t = 0; t < kLindStageR; ++t) {
        const int e = lane + 64 * t;
        rv[t] = rc[t] = rd[t] = -1;
        if (e < a.nR) {
            const int slot = e / (nc * NV), c = (e / NV) % nc, v = e % NV;
            rv[t] = v;
            rc[t] = (1 + slot) * nb + NV + c;
            rd[t] = (slot * nc + c) * a.nvp + v;
        }
    }

    // issue the direct copies of step k into buffer b
    auto issue = [&](int k, int b) {
        double2* buf = lds + b * a.buf;
        const int k0 = k / N1, k1 = k - k0 * N1;
        int r0 = k0 + q0;
        if (r0 >= a.N0) r0 -= a.N0;
        int s1 = k1 + q1f;
        if (s1 >= N1) s1 -= N1;
        const int n1 = min(nvalid, N1 - s1), n2 = nvalid - n1;
        const long row = long(r0) * N1;
        copy16(a.Uv + long(k) * NV * nb, buf, NV * nb, lane);
        copy16(a.Uc + (row + s1) * nc * nb, buf + oUc, n1 * nc * nb, lane);
        if (n2 > 0) copy16(a.Uc + row * nc * nb, buf + oUc + n1 * nc * nb, n2 * nc * nb, lane);
        unsigned* eb = reinterpret_cast<unsigned*>(buf + oE);
        const unsigned* ep = reinterpret_cast<const unsigned*>(a.eps);
        copy4(ep + 2 * long(k) * nb, eb, 2 * nb, lane);
        copy4(ep + 2 * (row + s1) * nb, eb + 2 * nb, 2 * n1 * nb, lane);
        if (n2 > 0) copy4(ep + 2 * row * nb, eb + 2 * nb * (1 + n1), 2 * n2 * nb, lane);
    };

    double2 acc[TQ];
#pragma unroll
    for (int j = 0; j < TQ; ++j) acc[j] = make_double2(0.0, 0.0);

    issue(0, 0);
    for (int k = 0; k < a.Nk; ++k) {
        const double2* buf = lds + (k & 1) * a.buf;
        __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of step k are in LDS
        __syncthreads();                // and every lane is done with step k - 1
        if (k + 1 < a.Nk) issue(k + 1, (k + 1) & 1);
        const double* eb = reinterpret_cast<const double*>(buf + oE);
#pragma unroll
        for (int t = 0; t < kLindStageR; ++t)
            if (rd[t] >= 0) sR[rd[t]] = 1.0 / (eb[rv[t]] - eb[rc[t]]);
        __syncthreads();
        if (active) {
            double2 A[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) A[v] = cmul(buf[v * nb + al], cconj(buf[v * nb + be]));
#pragma unroll
            for (int j = 0; j < TQ; ++j) {
                const int slot = s * TQ + j;
                const double2* ucs = buf + oUc + slot * nc * nb;
                const double* rs = sR + slot * nc * a.nvp;
                for (int c = 0; c < nc; ++c) {
                    const double2* uc = ucs + c * nb;
                    const double2 B = cmul(cconj(uc[al]), uc[be]);
                    const double* r = rs + c * a.nvp;
                    double2 S = make_double2(0.0, 0.0);
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        S.x += r[v] * A[v].x;
                        S.y += r[v] * A[v].y;
                    }
                    acc[j] = cadd(acc[j], cmul(S, B));
                }
            }
        }
    }

    if (!active) return;
#pragma unroll
    for (int j = 0; j < TQ; ++j) {
        const int q1 = q1f + s * TQ + j;
        if (q1 >= N1) continue;
        const int q = q0 * N1 + q1;
        const double2 x = cscale(acc[j], a.inv_nk);
        a.X[long(al * nb + be) * a.Nk + q] = x;
        if (a.tri && al != be) a.X[long(be * nb + al) * a.Nk + q] = cconj(x);
    }
}

template <int NV, int TQ>
void launch_nv(const LindhardArgs& a, int blocks, size_t lds, hipStream_t s)
{
    hipLaunchKernelGGL((lindhard_kernel<NV, TQ>), dim3(blocks), dim3(64), lds, s, a);
}

template <int TQ>
void launch_tq(const LindhardArgs& a, int blocks, size_t lds, hipStream_t s)
{
    switch (a.nv) {
#define EDUS_NV(n) \
    case n: launch_nv<n, TQ>(a, blocks, lds, s); break;
        EDUS_NV(1) EDUS_NV(2) EDUS_NV(3) EDUS_NV(4) EDUS_NV(5) EDUS_NV(6) EDUS_NV(7) EDUS_NV(8)
        EDUS_NV(9) EDUS_NV(10) EDUS_NV(11) EDUS_NV(12) EDUS_NV(13) EDUS_NV(14) EDUS_NV(15)
#undef EDUS_NV
        default: throw EngineError(EDUS_ERR_ARG, "lindhard: filled bands out of range");
    }
}

// waves per q for L element pairs (fractional when several q share a wave)
double waves_per_q(int L) { return L <= 64 ? 1.0 / double(64 / L) : double((L + 63) / 64); }

}  // namespace

LindhardPlan lindhard_plan(int nb, int nf, int N0, int N1)
{
    if (nb < 2 || nb > kMaxNb) throw EngineError(EDUS_ERR_ARG, "lindhard: nb must be in [2, 16]");
    if (nf < 0 || nf > nb) throw EngineError(EDUS_ERR_ARG, "lindhard: filledbands out of range");
    if (N0 < 1 || N1 < 1) throw EngineError(EDUS_ERR_ARG, "lindhard: grid sizes must be positive");
    LindhardPlan p{};
    p.nb = nb;
    p.nv = nf;
    p.nc = nb - nf;
    p.N0 = N0;
    p.N1 = N1;
    p.Nk = N0 * N1;
    const int P = nb * (nb + 1) / 2, M = nb * nb;
    p.tri = waves_per_q(P) < waves_per_q(M) ? 1 : 0;
    p.Pg = p.tri ? P : M;
    p.G = p.Pg <= 64 ? 64 / p.Pg : 1;
    p.W = p.Pg <= 64 ? 1 : (p.Pg + 63) / 64;
    // four q per lane when that keeps >= 4 work-groups per CU and wastes
    // under 10% of the slots at the row ends, else one
    auto row_blocks = [&](int tq) { return (N1 + p.G * tq - 1) / (p.G * tq); };
    const double fill4 = double(N1) / double(row_blocks(4) * p.G * 4);
    p.TQ = (long(N0) * row_blocks(4) * p.W >= 1024 && fill4 >= 0.9) ? 4 : 1;
    const int GT = p.G * p.TQ;
    p.nqb_row = row_blocks(p.TQ);
    p.nR = GT * p.nc * p.nv;
    p.nvp = (p.nv + 1) & ~1;
    p.buf = p.nv * nb + GT * p.nc * nb + ((1 + GT) * nb + 1) / 2;
    p.blocks = N0 * p.nqb_row * p.W;
    p.lds = size_t(2 * p.buf) * sizeof(double2) + size_t(GT) * p.nc * p.nvp * sizeof(double);
    if (p.nv > 0 && p.nc > 0 && p.nR > 64 * kLindStageR)
        throw EngineError(EDUS_ERR_UNSUPPORTED, "lindhard: more reciprocals per step than the lanes hold");
    return p;
}

size_t lindhard_work_bytes(const LindhardPlan& p) { return size_t(p.Nk) * p.nb * p.nb * sizeof(double2); }

void launch_lindhard(const LindhardPlan& p, const double2* U, const double* eps, double2* work, double2* X,
                     hipStream_t s)
{
    const size_t bytes = size_t(p.nb) * p.nb * p.Nk * sizeof(double2);
    if (p.nv == 0 || p.nc == 0) {  // no valence-conduction pair: X = 0
        EDUS_HIP_CHECK(hipMemsetAsync(X, 0, bytes, s));
        return;
    }
    double2* Uv = work;
    double2* Uc = work + size_t(p.Nk) * p.nv * p.nb;
    const long total = long(p.Nk) * p.nb * p.nb;
    hipLaunchKernelGGL(lindhard_split_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, U, Uv, Uc, p.nb,
                       p.nv, total);
    LindhardArgs a{};
    a.Uv = Uv;
    a.Uc = Uc;
    a.eps = eps;
    a.X = X;
    a.nb = p.nb;
    a.nv = p.nv;
    a.nc = p.nc;
    a.N0 = p.N0;
    a.N1 = p.N1;
    a.Nk = p.Nk;
    a.Pg = p.Pg;
    a.G = p.G;
    a.W = p.W;
    a.tri = p.tri;
    a.nqb_row = p.nqb_row;
    a.nR = p.nR;
    a.nvp = p.nvp;
    a.buf = p.buf;
    a.inv_nk = 1.0 / double(p.Nk);
    if (p.TQ == 4)
        launch_tq<4>(a, p.blocks, p.lds, s);
    else
        launch_tq<1>(a, p.blocks, p.lds, s);
    EDUS_HIP_CHECK(hipGetLastError());
}

}  // namespace edus

extern "C" int edus_lindhard_device(int nb, int N0, int N1, int filledbands, const double* U, const double* eps,
                                    double* work, double* X, void* stream)
{
    EDUS_REQUIRE(U && eps && work && X);
    return edus::guarded([&] {
        const edus::LindhardPlan p = edus::lindhard_plan(nb, filledbands, N0, N1);
        edus::launch_lindhard(p, reinterpret_cast<const double2*>(U), eps, reinterpret_cast<double2*>(work),
                              reinterpret_cast<double2*>(X), static_cast<hipStream_t>(stream));
    });
}

extern "C" int edus_lindhard(int nb, int N0, int N1, int filledbands, const double* U, const double* eps, double* X)
{
    EDUS_REQUIRE(U && eps && X);
    return edus::guarded([&] {
        const edus::LindhardPlan p = edus::lindhard_plan(nb, filledbands, N0, N1);
        const size_t M = size_t(nb) * nb, Nk = size_t(p.Nk);
        double2 *dU = nullptr, *dX = nullptr, *dW = nullptr;
        double* de = nullptr;
        auto release = [&] {
            (void)hipFree(dU);
            (void)hipFree(dX);
            (void)hipFree(dW);
            (void)hipFree(de);
        };
        try {
            EDUS_HIP_CHECK(hipMalloc(&dU, Nk * M * sizeof(double2)));
            EDUS_HIP_CHECK(hipMalloc(&dX, Nk * M * sizeof(double2)));
            EDUS_HIP_CHECK(hipMalloc(&dW, edus::lindhard_work_bytes(p)));
            EDUS_HIP_CHECK(hipMalloc(&de, Nk * nb * sizeof(double)));
            EDUS_HIP_CHECK(hipMemcpy(dU, U, Nk * M * sizeof(double2), hipMemcpyHostToDevice));
            EDUS_HIP_CHECK(hipMemcpy(de, eps, Nk * nb * sizeof(double), hipMemcpyHostToDevice));
            edus::launch_lindhard(p, dU, de, dW, dX, nullptr);
            std::vector<double2> planes(Nk * M);
            EDUS_HIP_CHECK(hipMemcpy(planes.data(), dX, Nk * M * sizeof(double2), hipMemcpyDeviceToHost));
            // planes [M][Nk] -> [Nk][nb][nb]
            for (size_t h = 0; h < M; ++h)
                for (size_t q = 0; q < Nk; ++q) {
                    X[2 * (q * M + h)] = planes[h * Nk + q].x;
                    X[2 * (q * M + h) + 1] = planes[h * Nk + q].y;
                }
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}
