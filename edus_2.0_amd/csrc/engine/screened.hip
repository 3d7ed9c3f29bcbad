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
//   * a split pass stores the valence columns of every U_k as Uv[k][v][a]
//     and the conduction columns as planes Uc[c nb + a][k], so that a step's
//     valence block is one contiguous run and the conduction elements of the
//     consecutive k + q of a work-group are adjacent;
//   * one work-group = W waves; each lane owns one element pair (a, b) (the
//     upper triangle when that packs the lanes better, the lower triangle
//     then follows from X_ba(q) = conj X_ab(q)), for G q-groups per wave when
//     the pairs fit G times in 64 lanes (W = 1), or spread over the W waves of
//     the work-group when they need more than 64 lanes (G = 1); the waves
//     share the staged columns and reciprocals;
//   * the work-group's GT = G * TQ q lie on one grid row at consecutive q1,
//     so for every k their k + q are consecutive on one row: per k the
//     work-group copies Uv(k) and the conduction columns of the k + q
//     straight from memory into LDS (global_load_lds, no registers),
//     double-buffered: step k + 1 is in flight while step k is computed;
//   * the reciprocals r_vc of each (q, k) are formed once per work-group,
//     from energies loaded into registers one step ahead, into one of two
//     tables, and read by all lanes of that q (one address for all:
//     broadcast); one barrier per k publishes both the columns and the
//     table; A_v(k) is built once per k in registers and used for all TQ q
//     of a lane;
//   * work-groups of neighbouring q go to the same XCD: their Uc(k + q)
//     windows overlap and stay in that XCD's L2;
//   * when the q-blocks would leave the last round of resident work-groups
//     mostly empty, each block's k range is split in two and the partial sums
//     are added into a zeroed X (0 + p1 + p2 is the same in either order).
// The summation order differs from the reference's (k outermost over pairs
// of bands, r multiplied instead of divided), so results agree to rounding,
// not bit for bit.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <vector>

#include "../../../include/edus_gpu.h"
#include "common.hpp"
#include "engine_internal.hpp"
#include "kernels.hpp"

namespace edus {

namespace {

constexpr int kLindStageR = 4;  // reciprocals formed per thread per k
constexpr int kLindGatherU = 8;  // conduction-column elements gathered per thread per k

struct LindhardArgs {
    const double2* Uv;  // [Nk][nv][nb]: U_k(a, v) at (k nv + v) nb + a
    const double2* Uc;  // planes [nc * nb][Nk]: U_k(a, nv + c) at (c nb + a) Nk + k
    const double* eps;  // [Nk][nb]
    double2* X;         // planes [nb * nb][Nk]
    int nb, nv, nc, N0, N1, Nk;
    int Pg;             // element pairs per q
    int G;              // q groups per wave (Pg <= 64), else 1
    int W;              // waves per work-group (Pg > 64), else 1
    int tri;            // pairs are the upper triangle
    int nqb_row;        // q-blocks per grid row
    int KS;             // k-range parts (work-groups per q-block); > 1: partial sums added into X
    int nR;             // reciprocals per k (whole work-group)
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

// U[k][a][n] -> Uv[k][v][a] and the conduction planes Uc[(c nb + a)][k]
__global__ void lindhard_zero_kernel(double2* __restrict__ X, long n)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) X[i] = make_double2(0.0, 0.0);
}

__global__ void lindhard_split_kernel(const double2* __restrict__ U, double2* __restrict__ Uv,
                                      double2* __restrict__ Uc, int nb, int nv, int Nk)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long total = long(Nk) * nb * nb;
    if (i >= total) return;
    const int M = nb * nb;
    const long k = i / M;
    const int r = int(i - k * M), al = r / nb, n = r % nb;
    const double2 v = U[i];
    if (n < nv)
        Uv[(k * nv + n) * nb + al] = v;
    else
        Uc[long((n - nv) * nb + al) * Nk + k] = v;
}

// copy count 16-byte elements src[0..count) to LDS dst: wave w of nw takes
// the 64-element chunks w, w + nw, ... (wave-uniform LDS base per chunk)
__device__ inline void copy16(const double2* src, double2* dst, int count, int lane, int w, int nw)
{
    for (int i = 64 * w; i < count; i += 64 * nw)
        if (i + lane < count) __builtin_amdgcn_global_load_lds(src + i + lane, dst + i, 16, 0, 0);
}

template <int NV, int TQ>
__global__ void __launch_bounds__(256) lindhard_kernel(LindhardArgs a)
{
    extern __shared__ double2 lds[];
    constexpr int NVP = (NV + 1) & ~1;
    const int nb = a.nb, nc = a.nc, N1 = a.N1, GT = a.G * TQ;
    // buffer layout (double2 units): Uv(k) [NV][nb]; the conduction columns
    // of the GT slots, slot fastest: [nc][nb][GT].  With the slot fastest,
    // the TQ q of a lane sit at fixed offsets from one address.  After both
    // buffers: two reciprocal tables [nc][GT][NVP] (doubles), one per step
    // parity.
    const int oUc = NV * nb;
    const int nRt = nc * GT * NVP;
    double* sR0 = reinterpret_cast<double*>(lds + 2 * a.buf);

    // XCD-aware order: consecutive logical blocks on one XCD
    long L;
    {
        const long nblk = gridDim.x, q = nblk / 8, r = nblk % 8;
        const long x = blockIdx.x % 8, i = blockIdx.x / 8;
        L = x * q + min(x, r) + i;
    }
    const int qbi = int(L / a.KS), ks = int(L % a.KS);
    const int kb = int(long(a.Nk) * ks / a.KS), ke = int(long(a.Nk) * (ks + 1) / a.KS);
    const int q0 = qbi / a.nqb_row, q1f = (qbi % a.nqb_row) * GT;  // row and first q1 of the block
    const int nvalid = min(GT, N1 - q1f);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = a.W, nthr = 64 * a.W;
    int s, pi;
    if (a.W > 1) {
        s = 0;
        pi = tid;
    } else {
        s = lane / a.Pg;
        pi = lane % a.Pg;
    }
    const bool active = s < a.G && pi < a.Pg;
    int al = 0, be = 0;
    if (active) pair_of(pi, nb, a.tri, al, be);

    // per-thread work lists, fixed for the whole k loop:
    //  reciprocal entries: valence band, slot, conduction band (as a column
    //  of the energy row) and destination in the table
    int rv[kLindStageR], rs[kLindStageR], rc[kLindStageR], rd[kLindStageR];
#pragma unroll
    for (int t = 0; t < kLindStageR; ++t) {
        const int e = tid + nthr * t;
        rv[t] = rs[t] = rc[t] = rd[t] = -1;
        if (e < a.nR && (e / NV) % GT < nvalid) {
            const int c = e / (GT * NV), slot = (e / NV) % GT, v = e % NV;
            rv[t] = v;
            rs[t] = slot;
            rc[t] = NV + c;
            rd[t] = (c * GT + slot) * NVP + v;
        }
    }
    //  gathered conduction-column elements: offset of plane (c nb + a) and slot
    const int nU = nc * nb * GT;
    int gp[kLindGatherU], gs[kLindGatherU];
#pragma unroll
    for (int t = 0; t < kLindGatherU; ++t) {
        const int e = tid + nthr * t;
        gp[t] = gs[t] = -1;
        if (e < nU && e % GT < nvalid) {
            gp[t] = (e / GT) * a.Nk;  // < nb^2 Nk, fits an int for any grid the engine takes
            gs[t] = e % GT;
        }
    }

    // first grid point of row r0 + shift s1 of step (k0, k1)
    auto shift = [&](int k0, int k1, int& row, int& s1) {
        int r0 = k0 + q0;
        if (r0 >= a.N0) r0 -= a.N0;
        s1 = k1 + q1f;
        if (s1 >= N1) s1 -= N1;
        row = r0 * N1;
    };
    // issue the direct copies of step k into buffer b
    auto issue = [&](int k, int k0, int k1, int b) {
        double2* buf = lds + b * a.buf;
        int row, s1;
        shift(k0, k1, row, s1);
        copy16(a.Uv + long(k) * NV * nb, buf, NV * nb, lane, wave, nw);
        // element e = tid + nthr t lands at LDS slot oUc + e: wave-uniform base
        // oUc + nthr t + 64 wave, lane offset 16 bytes
#pragma unroll
        for (int t = 0; t < kLindGatherU; ++t) {
            if (nthr * t >= nU) break;
            if (gp[t] >= 0) {
                int kk = s1 + gs[t];
                if (kk >= N1) kk -= N1;
                __builtin_amdgcn_global_load_lds(a.Uc + (gp[t] + row + kk),
                                                 buf + oUc + nthr * t + 64 * wave, 16, 0, 0);
            }
        }
    };
    // the energies of this thread's reciprocal entries for step k, into registers
    double ev[kLindStageR], ec[kLindStageR];
    auto load_energies = [&](int k, int k0, int k1) {
        int row, s1;
        shift(k0, k1, row, s1);
#pragma unroll
        for (int t = 0; t < kLindStageR; ++t) {
            if (rd[t] >= 0) {
                int kk = s1 + rs[t];
                if (kk >= N1) kk -= N1;
                ev[t] = a.eps[long(k) * nb + rv[t]];
                ec[t] = a.eps[long(row + kk) * nb + rc[t]];
            }
        }
    };
    auto store_reciprocals = [&](double* sR) {
#pragma unroll
        for (int t = 0; t < kLindStageR; ++t)
            if (rd[t] >= 0) sR[rd[t]] = 1.0 / (ev[t] - ec[t]);
    };

    double2 acc[TQ];
#pragma unroll
    for (int j = 0; j < TQ; ++j) acc[j] = make_double2(0.0, 0.0);

    // grid coordinates of the next step to issue, advanced without divisions
    int n0 = kb / N1, n1 = kb - (kb / N1) * N1;
    auto advance = [&]() {
        if (++n1 == N1) {
            n1 = 0;
            ++n0;
        }
    };
    issue(kb, n0, n1, 0);
    load_energies(kb, n0, n1);
    store_reciprocals(sR0);
    advance();
    for (int k = kb; k < ke; ++k) {
        const int par = (k - kb) & 1;
        const double2* buf = lds + par * a.buf;
        const double* sR = sR0 + par * nRt;
        __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of step k are in LDS
        __syncthreads();                // and every lane is done with step k - 1
        const bool next = k + 1 < ke;
        if (next) {
            issue(k + 1, n0, n1, par ^ 1);
            load_energies(k + 1, n0, n1);
            advance();
        }
        if (active) {
            double2 A[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) A[v] = cmul(buf[v * nb + al], cconj(buf[v * nb + be]));
            // c outer, the TQ q inner: TQ independent chains per iteration;
            // the TQ slots of a lane are adjacent in the Uc tile and the
            // reciprocals, so their reads share one address
            const double2* ua = buf + oUc + al * GT + s * TQ;
            const double2* ub = buf + oUc + be * GT + s * TQ;
            const double* rr = sR + s * TQ * NVP;
            for (int c = 0; c < nc; ++c) {
                double2 B[TQ], S[TQ];
#pragma unroll
                for (int j = 0; j < TQ; ++j) B[j] = cmul(cconj(ua[j]), ub[j]);
#pragma unroll
                for (int j = 0; j < TQ; ++j) {
                    const double2* r2 = reinterpret_cast<const double2*>(rr + j * NVP);  // 16-byte aligned
                    S[j] = make_double2(0.0, 0.0);
#pragma unroll
                    for (int v = 0; v < NV; v += 2) {
                        const double2 r = r2[v / 2];
                        S[j].x = fma(r.x, A[v].x, S[j].x);
                        S[j].y = fma(r.x, A[v].y, S[j].y);
                        if (v + 1 < NV) {
                            S[j].x = fma(r.y, A[v + 1].x, S[j].x);
                            S[j].y = fma(r.y, A[v + 1].y, S[j].y);
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < TQ; ++j) {
                    acc[j].x = fma(S[j].x, B[j].x, fma(-S[j].y, B[j].y, acc[j].x));
                    acc[j].y = fma(S[j].x, B[j].y, fma(S[j].y, B[j].x, acc[j].y));
                }
                ua += nb * GT;
                ub += nb * GT;
                rr += GT * NVP;
            }
        }
        // reciprocals of step k + 1 into the other table: its readers finished
        // step k - 1 before this step's barrier; the next barrier publishes it
        if (next) store_reciprocals(sR0 + (par ^ 1) * nRt);
    }

    if (!active) return;
#pragma unroll
    for (int j = 0; j < TQ; ++j) {
        const int q1 = q1f + s * TQ + j;
        if (q1 >= N1) continue;
        const int q = q0 * N1 + q1;
        const double2 x = cscale(acc[j], a.inv_nk);
        double2* x1 = a.X + long(al * nb + be) * a.Nk + q;
        double2* x2 = a.X + long(be * nb + al) * a.Nk + q;
        if (a.KS == 1) {
            *x1 = x;
            if (a.tri && al != be) *x2 = cconj(x);
        } else {
            // two parts onto a zeroed X: 0 + p1 + p2 is the same sum in
            // either order, so the result does not depend on timing
            atomicAdd(&x1->x, x.x);
            atomicAdd(&x1->y, x.y);
            if (a.tri && al != be) {
                atomicAdd(&x2->x, x.x);
                atomicAdd(&x2->y, -x.y);
            }
        }
    }
}

// Work-groups of one launch run in rounds of (resident work-groups per CU) x
// CUs; when the q-blocks leave the last round mostly empty, each q-block's k
// range is split in two (partial sums added into a zeroed X) if that fills
// the rounds better.
template <int NV, int TQ>
void launch_nv(const LindhardArgs& a0, int blocks, size_t lds, hipStream_t s)
{
    LindhardArgs a = a0;
    auto kern = &lindhard_kernel<NV, TQ>;
    int per_cu = 0, dev = 0, ncu = 0;
    EDUS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern),
                                                                64 * a.W, lds));
    EDUS_HIP_CHECK(hipGetDevice(&dev));
    EDUS_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const long cap = long(std::max(per_cu, 1)) * std::max(ncu, 1);
    auto fill = [&](long n) { return double(n) / double(((n + cap - 1) / cap) * cap); };
    a.KS = (a.Nk >= 64 && fill(2L * blocks) > fill(blocks) + 0.05) ? 2 : 1;
    if (a.KS > 1) EDUS_HIP_CHECK(hipMemsetAsync(a.X, 0, size_t(a.nb) * a.nb * a.Nk * sizeof(double2), s));
    hipLaunchKernelGGL(kern, dim3(blocks * a.KS), dim3(64 * a.W), lds, s, a);
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
    if (long(nb) * nb * N0 * N1 >= (1L << 31))
        throw EngineError(EDUS_ERR_UNSUPPORTED, "lindhard: nb^2 * Nk must stay below 2^31");
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
    // four q per lane when that keeps >= 1024 waves, wastes under 10% of the
    // slots at the row ends and the per-thread staging lists hold a step;
    // else one (eight measured slower at nb = 14: 220 VGPRs, 2 waves per SIMD)
    if (p.W > 4) throw EngineError(EDUS_ERR_UNSUPPORTED, "lindhard: more than 256 element pairs");
    auto row_blocks = [&](int tq) { return (N1 + p.G * tq - 1) / (p.G * tq); };
    const int nthr = 64 * p.W;
    auto fits = [&](int tq) {
        const int gt = p.G * tq;
        return gt * p.nc * p.nv <= nthr * kLindStageR && p.nc * nb * gt <= nthr * kLindGatherU;
    };
    auto good = [&](int tq) {
        const double fill = double(N1) / double(row_blocks(tq) * p.G * tq);
        return long(N0) * row_blocks(tq) * p.W >= 1024 && fill >= 0.9 && fits(tq);
    };
    p.TQ = good(4) ? 4 : 1;
    const int GT = p.G * p.TQ;
    p.nqb_row = row_blocks(p.TQ);
    p.nR = GT * p.nc * p.nv;
    p.nvp = (p.nv + 1) & ~1;
    p.buf = p.nv * nb + p.nc * nb * GT;
    p.blocks = N0 * p.nqb_row;
    p.lds = size_t(2 * p.buf) * sizeof(double2) + size_t(2 * GT) * p.nc * p.nvp * sizeof(double);
    if (p.nv > 0 && p.nc > 0 && !fits(p.TQ))
        throw EngineError(EDUS_ERR_UNSUPPORTED, "lindhard: a step needs more per-thread staging than the kernel holds");
    return p;
}

size_t lindhard_work_bytes(const LindhardPlan& p) { return size_t(p.Nk) * p.nb * p.nb * sizeof(double2); }

void launch_lindhard(const LindhardPlan& p, const double2* U, const double* eps, double2* work, double2* X,
                     hipStream_t s)
{
    const size_t bytes = size_t(p.nb) * p.nb * p.Nk * sizeof(double2);
    if (p.nv == 0 || p.nc == 0) {  // no valence-conduction pair: X = 0 (a kernel: callers may capture this)
        const long n = long(bytes / sizeof(double2));
        hipLaunchKernelGGL(lindhard_zero_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, X, n);
        EDUS_HIP_CHECK(hipGetLastError());
        return;
    }
    double2* Uv = work;
    double2* Uc = work + size_t(p.Nk) * p.nv * p.nb;
    const long total = long(p.Nk) * p.nb * p.nb;
    hipLaunchKernelGGL(lindhard_split_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, s, U, Uv, Uc, p.nb,
                       p.nv, p.Nk);
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

// ---- the Lindhard response without allocation ----------------------------
// Callers that cannot allocate device memory (code under HIP graph capture,
// serving loops) use edus_lindhard_device_noalloc: the band-split copy of U
// that edus_lindhard_device keeps in its `work` argument lives here instead,
// in a fixed array of every GPU the library runs on, held until the process
// ends.  Its device address is taken once per device with
// hipGetSymbolAddress; host code passes that address, never the array's
// name (a host-side stand-in).
namespace edus {
constexpr size_t kLindhardScratchElems = size_t(3) << 20;  // 3,145,728 complex doubles = 48 MiB
static __device__ __align__(256) double2 scratch[kLindhardScratchElems];

namespace {
// address of `scratch` on the current device
double2* scratch_address()
{
    static std::mutex mu;
    static std::map<int, double2*> addr;
    int dev = 0;
    EDUS_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void* p = nullptr;
    EDUS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw EngineError(EDUS_ERR_HIP, "hipGetSymbolAddress returned no address for the Lindhard scratch");
    addr[dev] = static_cast<double2*>(p);
    return addr[dev];
}
}  // namespace
}  // namespace edus

extern "C" int edus_lindhard_device_noalloc(int nb, int N0, int N1, int filledbands, const double* U,
                                            const double* eps, double* X, void* stream)
{
    EDUS_REQUIRE(U && eps && X);
    return edus::guarded([&] {
        const edus::LindhardPlan p = edus::lindhard_plan(nb, filledbands, N0, N1);
        const size_t need = edus::lindhard_work_bytes(p);
        if (need > edus::kLindhardScratchElems * sizeof(double2))
            throw edus::EngineError(EDUS_ERR_ARG,
                                    "edus_lindhard_device_noalloc: Nk * nb^2 = " +
                                        std::to_string(size_t(p.Nk) * nb * nb) + " exceeds its fixed scratch of " +
                                        std::to_string(edus::kLindhardScratchElems) +
                                        " complex doubles (48 MiB); use edus_lindhard_device with a workspace");
        edus::launch_lindhard(p, reinterpret_cast<const double2*>(U), eps, edus::scratch_address(),
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
