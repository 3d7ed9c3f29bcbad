// Wannier <-> Bloch gauge rotation of an operator field on the matrix cores.
//
// Per k-point the reference forms (Operator::go_to_bloch / go_to_wannier,
// src/Operator/Operator.hpp:528-564)
//     O_b(k) = U(k)^dagger O_w(k) U(k)      (to Bloch)
//     O_w(k) = U(k) O_b(k) U(k)^dagger      (to Wannier)
// with two cblas_zgemm calls, U's columns being the eigenvectors of H0(k).
// (first U^dagger O or U O, then that times U or U^dagger, as the reference
// associates them).  Here a work-group holds the operator and U of 8
// consecutive k-points in LDS and each of its 8 waves forms both products of
// one k-point as real-split complex GEMMs on 16 x 16 tiles (nb <= 16, zero
// padded):
//   KIND kMfma64: v_mfma_f64_16x16x4 (FP64 throughout),
//   KIND kMfma32: v_mfma_f32_16x16x4 on FP32 operands with FP32 accumulation
//                 (the mixed-precision mode), results widened to FP64.
// The operator and U go straight from HBM into LDS (direct loads, no
// registers held); the intermediate product and then the result take the
// operator's place in LDS, so a work-group needs 57 KB at nb = 14 and two fit
// on a CU.  The result is written back through LDS so that the stores are
// coalesced, in one of three layouts:
// FULL planes, HERMITIAN planes (upper triangle) or k-major [Nk][nb][nb]
// (the reference's BlockMatrix order, for transfers to the host).
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace edus {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

constexpr int kRotThreads = 512;  // 8 waves, one k-point each
constexpr int kRotTK = 8;
constexpr int kRotPS = kRotTK + 1;  // plane stride in LDS, padded against bank conflicts

template <int NB, int LAYOUT>
constexpr int rot_npl()
{
    return LAYOUT == kFull ? NB * NB : NB * (NB + 1) / 2;
}

// LDS slots of a tile array of n planes, in whole 64-slot chunks (one chunk
// = one wave-wide direct load)
constexpr int rot_slots(int n) { return (n * kRotPS + 63) / 64 * 64; }

// the input operator's tile, later the intermediate product and the result
// (max of the two sizes), then U's tile: 57 KB at nb = 14, two work-groups
// per CU
template <int NB, int LAYOUT>
constexpr size_t rot_lds_bytes()
{
    constexpr int a = rot_slots(rot_npl<NB, LAYOUT>()), m = rot_slots(NB * NB);
    return size_t((a > m ? a : m) + m) * sizeof(double2);
}

// element (r, c) of the input tile (zero outside the nb x nb block)
template <int NB, int LAYOUT>
__device__ inline double2 in_elem(const double2* s, int r, int c, int kk)
{
    if (r >= NB || c >= NB) return make_double2(0.0, 0.0);
    if (LAYOUT == kFull) return s[(r * NB + c) * kRotPS + kk];
    if (r <= c) return s[tri_index(r, c, NB) * kRotPS + kk];
    const double2 v = s[tri_index(c, r, NB) * kRotPS + kk];
    return make_double2(v.x, -v.y);
}

// element (r, c) of a full nb x nb tile (U or the intermediate product)
template <int NB>
__device__ inline double2 full_elem(const double2* s, int r, int c, int kk)
{
    if (r >= NB || c >= NB) return make_double2(0.0, 0.0);
    return s[(r * NB + c) * kRotPS + kk];
}

// C += A B, complex, real-split over the matrix core of KIND.  Lane l gives
// A(i, c) and B(c, i) with i = l % 16, c = 4 s + l / 16 for K step s.
template <int KIND, class Acc>
__device__ inline void cgemm_step(double2 A, double2 B, Acc& cr, Acc& ci)
{
    if constexpr (KIND == kCommMfma64) {
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(A.x, B.x, cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-A.y, B.y, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(A.x, B.y, ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(A.y, B.x, ci, 0, 0, 0);
    } else {
        const float ax = float(A.x), ay = float(A.y), bx = float(B.x), by = float(B.y);
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(ax, bx, cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(-ay, by, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x4f32(ax, by, ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x4f32(ay, bx, ci, 0, 0, 0);
    }
}

// row of result slot v of lane group q: f64 16x16x4 puts row q + 4v in slot
// v, f32 16x16x4 row 4q + v (the column is the lane's i in both)
template <int KIND>
__device__ inline int result_row(int q, int v)
{
    return KIND == kCommMfma64 ? q + 4 * v : 4 * q + v;
}

template <int NB, int LAYOUT, int KIND>
__global__ void __launch_bounds__(kRotThreads) rotate_kernel(RotateArgs a)
{
    constexpr int NPL = rot_npl<NB, LAYOUT>();
    constexpr int M = NB * NB;
    constexpr int SA = rot_slots(NPL) > rot_slots(M) ? rot_slots(NPL) : rot_slots(M);
    using acc_t = typename std::conditional<KIND == kCommMfma64, dbl4, flt4>::type;
    extern __shared__ double2 lds[];
    double2* sO = lds;       // input operator, NPL planes; then the intermediate product; then the result
    double2* sU = lds + SA;  // U, M planes
    const int Nk = a.Nk;
    const int k0 = xcd_contiguous_block() * kRotTK;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;

    // every load of the tile goes straight from HBM into LDS: lane l of a
    // wave fills slot base + l (global_load_lds_dwordx4), slot q holds
    // element (q / PS, q % PS); the padding column and k-points past Nk are
    // not loaded (their results are not stored)
    for (int c = wave; c < rot_slots(NPL) / 64; c += kRotThreads / 64) {
        const int q = c * 64 + lane, p = q / kRotPS, kk = q - p * kRotPS;
        if (kk < kRotTK && p < NPL && k0 + kk < Nk)
            __builtin_amdgcn_global_load_lds(a.in + long(p) * Nk + k0 + kk, sO + c * 64, 16, 0, 0);
    }
    for (int c = wave; c < rot_slots(M) / 64; c += kRotThreads / 64) {
        const int q = c * 64 + lane, p = q / kRotPS, kk = q - p * kRotPS;
        if (kk < kRotTK && p < M && k0 + kk < Nk)
            __builtin_amdgcn_global_load_lds(a.U + long(p) * Nk + k0 + kk, sU + c * 64, 16, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's direct loads are in LDS
    __syncthreads();                // and every other wave's

    const int kk = wave, i = lane & 15, q = lane >> 4;
    const bool to_bloch = a.to_bloch != 0;
    acc_t cr{}, ci{};
    // first product T = Y O, Y = U^dagger (to Bloch) or U (to Wannier)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (4 * s >= NB) break;
        const int c = 4 * s + q;
        double2 A;
        if (to_bloch) {
            A = full_elem<NB>(sU, c, i, kk);
            A.y = -A.y;
        } else {
            A = full_elem<NB>(sU, i, c, kk);
        }
        const double2 B = in_elem<NB, LAYOUT>(sO, c, i, kk);
        cgemm_step<KIND>(A, B, cr, ci);
    }
    __syncthreads();  // every wave has read O before T replaces it
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = result_row<KIND>(q, v);
        if (r < NB && i < NB) sO[(r * NB + i) * kRotPS + kk] = make_double2(double(cr[v]), double(ci[v]));
    }
    __syncthreads();
    // second product R = T X, X = U (to Bloch) or U^dagger (to Wannier)
    cr = acc_t{};
    ci = acc_t{};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (4 * s >= NB) break;
        const int c = 4 * s + q;
        const double2 A = full_elem<NB>(sO, i, c, kk);
        double2 B;
        if (to_bloch) {
            B = full_elem<NB>(sU, c, i, kk);
        } else {
            B = full_elem<NB>(sU, i, c, kk);
            B.y = -B.y;
        }
        cgemm_step<KIND>(A, B, cr, ci);
    }
    __syncthreads();  // every wave has read its T before the results replace it
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = result_row<KIND>(q, v);
        if (r < NB && i < NB) sO[(r * NB + i) * kRotPS + kk] = make_double2(double(cr[v]), double(ci[v]));
    }
    __syncthreads();
    double2* sT = sO;

    if (a.out_layout == kRotKMajor) {
        // [Nk][nb][nb]: the tile's 8 k-points are one contiguous run
        for (int idx = threadIdx.x; idx < M * kRotTK; idx += kRotThreads) {
            const int t = idx / M, e = idx - t * M;
            if (k0 + t < Nk) a.out[long(k0 + t) * M + e] = sT[e * kRotPS + t];
        }
    } else if (a.out_layout == kFull) {
        for (int idx = threadIdx.x; idx < M * kRotTK; idx += kRotThreads) {
            const int p = idx / kRotTK, t = idx - p * kRotTK;
            if (k0 + t < Nk) a.out[long(p) * Nk + k0 + t] = sT[p * kRotPS + t];
        }
    } else {
        constexpr int NH = NB * (NB + 1) / 2;
        for (int idx = threadIdx.x; idx < NH * kRotTK; idx += kRotThreads) {
            const int p = idx / kRotTK, t = idx - p * kRotTK;
            int n = 0, r = p;
            while (r >= NB - n) {
                r -= NB - n;
                ++n;
            }
            if (k0 + t < Nk) a.out[long(p) * Nk + k0 + t] = sT[(n * NB + n + r) * kRotPS + t];
        }
    }
}

template <int NB, int LAYOUT, int KIND>
void launch_rot(const RotateArgs& a, hipStream_t s)
{
    constexpr size_t lds = rot_lds_bytes<NB, LAYOUT>();
    auto kern = &rotate_kernel<NB, LAYOUT, KIND>;
    static std::once_flag raised;  // contexts may be created from several host threads
    std::call_once(raised, [&] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(lds));
    });
    if (g_attributes_only) return;
    hipLaunchKernelGGL(kern, dim3((a.Nk + kRotTK - 1) / kRotTK), dim3(kRotThreads), lds, s, a);
}

template <int NB>
void launch_rot_nb(const RotateArgs& a, int in_layout, int kind, hipStream_t s)
{
    if (in_layout == kFull) {
        if (kind == kCommMfma32)
            launch_rot<NB, kFull, kCommMfma32>(a, s);
        else
            launch_rot<NB, kFull, kCommMfma64>(a, s);
    } else {
        if (kind == kCommMfma32)
            launch_rot<NB, kHermitian, kCommMfma32>(a, s);
        else
            launch_rot<NB, kHermitian, kCommMfma64>(a, s);
    }
}

}  // namespace

void launch_rotate(int nb, int in_layout, int kind, const RotateArgs& a, hipStream_t s)
{
    switch (nb) {
        case 1: launch_rot_nb<1>(a, in_layout, kind, s); break;
        case 2: launch_rot_nb<2>(a, in_layout, kind, s); break;
        case 3: launch_rot_nb<3>(a, in_layout, kind, s); break;
        case 4: launch_rot_nb<4>(a, in_layout, kind, s); break;
        case 5: launch_rot_nb<5>(a, in_layout, kind, s); break;
        case 6: launch_rot_nb<6>(a, in_layout, kind, s); break;
        case 7: launch_rot_nb<7>(a, in_layout, kind, s); break;
        case 8: launch_rot_nb<8>(a, in_layout, kind, s); break;
        case 9: launch_rot_nb<9>(a, in_layout, kind, s); break;
        case 10: launch_rot_nb<10>(a, in_layout, kind, s); break;
        case 11: launch_rot_nb<11>(a, in_layout, kind, s); break;
        case 12: launch_rot_nb<12>(a, in_layout, kind, s); break;
        case 13: launch_rot_nb<13>(a, in_layout, kind, s); break;
        case 14: launch_rot_nb<14>(a, in_layout, kind, s); break;
        case 15: launch_rot_nb<15>(a, in_layout, kind, s); break;
        case 16: launch_rot_nb<16>(a, in_layout, kind, s); break;
        default: break;  // rejected at context creation
    }
}

}  // namespace edus
