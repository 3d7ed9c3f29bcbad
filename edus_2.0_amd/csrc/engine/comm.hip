// Fused commutator + integrator-stage kernel.
//
// For a tile of TK consecutive k-points it forms, per k,
//     H(k) = H0(k) + sum_d E_d xi_d(k) [+ Sigma_HF(k)] [+ Hartree diagonal]
//     f(k) = D(k) - i [H(k), rho(k)]
// and applies the integrator's stage update to every element in the same
// pass.  The reference computes the same quantities as
//     Calculate_TDHamiltonian (Simulation.cpp:296-349),
//     Coulomb::EffectiveHamiltonian (Coulomb.cpp:134-180),
//     commutator(Output, -i, H, Input, false) (BlockMatrix_definitions.hpp:203-227),
//     SumWithProduct stage updates (DESolver.hpp:10-20,155-220)
// in separate sweeps over the whole grid.  Here H and rho of the tile sit in
// LDS, every global load of the tile is issued before the first barrier, and
// the product [H, rho] is formed by one of three engines (KIND):
//   kScalar   FP64 FMAs, one output element per thread:
//             C_nm = sum_l (H_nl rho_lm - rho_nl H_lm);
//   kMfma64   FP64 MFMA (v_mfma_f64_16x16x4): per k-point one wave forms the
//             real-split complex products H.rho and rho.H on 16x16 tiles
//             (nb <= 16, zero padded), 32 MFMAs per k;
//   kMfma32   the same with FP32 operands and FP32 accumulation
//             (v_mfma_f32_16x16x4): the mixed-precision mode.  D, the stage
//             updates and the state stay FP64.
// LAYOUT = kFull keeps every element; kHermitian keeps the upper triangle and
// expands the lower one by conjugation while reading LDS.  In kFull the
// Hermitian fields H0 and xi may come as upper triangles too (CommArgs::tri):
// the state keeps every element, and H0 and xi cost 0.54 of their FULL bytes
// at nb = 14.
#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace edus {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

constexpr int kCommThreads = 256;

// work-group size: 256 threads, or 512 for MFMA tile 2 (one k-point per wave)
template <int KIND, int TILE>
constexpr int comm_threads()
{
    return (KIND != kCommScalar && TILE == 2) ? 2 * kCommThreads : kCommThreads;
}

template <int NB, int LAYOUT>
__device__ inline void decode_plane(int p, int& n, int& m)
{
    if (LAYOUT == kFull) {
        n = p / NB;
        m = p - n * NB;
        return;
    }
    int nn = 0;
    while (p >= NB - nn) {
        p -= NB - nn;
        ++nn;
    }
    n = nn;
    m = nn + p;
}

// LDS slot of element (a, b) and whether it must be conjugated
template <int NB, int LAYOUT>
__device__ inline int slot(int a, int b, bool& conj)
{
    if (LAYOUT == kFull) {
        conj = false;
        return a * NB + b;
    }
    conj = a > b;
    const int lo = conj ? b : a, hi = conj ? a : b;
    return tri_index(lo, hi, NB);
}

template <int NB, int LAYOUT>
constexpr int npl_of()
{
    return LAYOUT == kFull ? NB * NB : NB * (NB + 1) / 2;
}

// scalar engine: keep the two LDS tiles (rho and H) under ~64 KB
template <int NB, int LAYOUT>
constexpr int scalar_tile()
{
    const int npl = npl_of<NB, LAYOUT>();
    int tk = 64;
    while (tk > 4 && 2 * npl * tk * 16 > 65536) tk /= 2;
    return tk;
}

// MFMA engines: tile 0 holds 8 k-points (two per wave; 8 x 16 B = one
// 128-byte line per plane), tile 1 holds 4 (one per wave; half the LDS, so
// more work-groups per CU), tile 2 holds 8 with 512 threads (one per wave;
// half the elements per thread); plane stride padded by one against LDS bank
// conflicts.  The per-k arithmetic does not depend on the tile.
template <int NB, int LAYOUT, int KIND, int TILE>
constexpr int tile_k()
{
    return KIND == kCommScalar ? scalar_tile<NB, LAYOUT>() : (TILE == 1 ? 4 : 8);
}

template <int KIND, int TK>
constexpr int plane_stride()
{
    return KIND == kCommScalar ? TK : TK + 1;
}

// Slots of one LDS tile array: NPL planes of PS slots, rounded up to whole
// 64-slot chunks (one chunk = one wave-wide direct load).
template <int NB, int LAYOUT, int KIND, int TILE>
constexpr int tile_slots()
{
    constexpr int TK = tile_k<NB, LAYOUT, KIND, TILE>();
    return (npl_of<NB, LAYOUT>() * plane_stride<KIND, TK>() + 63) / 64 * 64;
}

// Slots of the H tile: as the rho tile, or the upper triangle's planes when
// the fields come as upper triangles (TRI)
template <int NB, int LAYOUT, int KIND, int TILE, bool TRI>
constexpr int h_tile_slots()
{
    constexpr int TK = tile_k<NB, LAYOUT, KIND, TILE>();
    constexpr int np = TRI ? NB * (NB + 1) / 2 : npl_of<NB, LAYOUT>();
    return (np * plane_stride<KIND, TK>() + 63) / 64 * 64;
}

// the monitor's per-wave maxima (trace, Hermiticity) of up to 8 waves
constexpr int kMonRedSlots = 8;

template <int NB, int LAYOUT, int KIND, int TILE, bool TRI>
constexpr size_t comm_lds_bytes()
{
    // rho and H tiles, the Hartree diagonal and the monitor's reduction slots
    return (size_t(tile_slots<NB, LAYOUT, KIND, TILE>()) + h_tile_slots<NB, LAYOUT, KIND, TILE, TRI>() + kMaxNb +
            kMonRedSlots) *
           sizeof(double2);
}

// element (r, c) of a tile in LDS (zero outside the nb x nb block)
template <int NB, int LAYOUT, int PS>
__device__ inline double2 lds_elem(const double2* s, int r, int c, int kk)
{
    if (r >= NB || c >= NB) return make_double2(0.0, 0.0);
    bool cj;
    const int sl = slot<NB, LAYOUT>(r, c, cj);
    double2 v = s[sl * PS + kk];
    if (cj) v.y = -v.y;
    return v;
}

// LDS slot of element (a, b) of the H tile and whether it must be conjugated:
// as slot<NB, LAYOUT>, or with tri (FULL layout, upper-triangle fields) the
// HERMITIAN slot; selects rather than branches, so that both forms share one
// copy of the code that reads the tile
template <int NB, int LAYOUT>
__device__ inline int h_slot(int a, int b, bool tri, bool& conj)
{
    if (LAYOUT != kFull) return slot<NB, LAYOUT>(a, b, conj);
    conj = tri && a > b;
    const int lo = conj ? b : a, hi = conj ? a : b;
    return tri ? tri_index(lo, hi, NB) : a * NB + b;
}

// element (r, c) of the H tile (zero outside the nb x nb block)
template <int NB, int LAYOUT, int PS>
__device__ inline double2 lds_h(const double2* s, int r, int c, int kk, bool tri)
{
    if (r >= NB || c >= NB) return make_double2(0.0, 0.0);
    bool cj;
    const int sl = h_slot<NB, LAYOUT>(r, c, tri, cj);
    double2 v = s[sl * PS + kk];
    if (cj) v.y = -v.y;
    return v;
}

// [H, rho] of k-point kk of the tile on the matrix cores (the H tile as
// h_slot lays it out).  Lane l holds, per
// K step s, A-fragments (row i = l%16, column c = 4s + l/16) and B-fragments
// (row c, column i); the result element (l/16 + 4v, l%16) is in slot v of the
// f64 instruction, (4*(l/16) + v, l%16) in slot v of the f32 one.
//   Re C = Hr.Rr - Hi.Ri - Rr.Hr + Ri.Hi,   Im C = Hr.Ri + Hi.Rr - Rr.Hi - Ri.Hr
template <int NB, int LAYOUT, int PS>
__device__ inline void mfma_commutator64(const double2* sH, const double2* sR, int kk, int lane, bool tri, dbl4& cre,
                              dbl4& cim)
{
    const int i = lane & 15, q = lane >> 4;
    cre = dbl4{0.0, 0.0, 0.0, 0.0};
    cim = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int c = 4 * s + q;
        if (4 * s >= NB) break;
        const double2 hA = lds_h<NB, LAYOUT, PS>(sH, i, c, kk, tri), rA = lds_elem<NB, LAYOUT, PS>(sR, i, c, kk);
        const double2 hB = lds_h<NB, LAYOUT, PS>(sH, c, i, kk, tri), rB = lds_elem<NB, LAYOUT, PS>(sR, c, i, kk);
        cre = __builtin_amdgcn_mfma_f64_16x16x4f64(hA.x, rB.x, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-hA.y, rB.y, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-rA.x, hB.x, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f64_16x16x4f64(rA.y, hB.y, cre, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f64_16x16x4f64(hA.x, rB.y, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f64_16x16x4f64(hA.y, rB.x, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f64_16x16x4f64(-rA.x, hB.y, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f64_16x16x4f64(-rA.y, hB.x, cim, 0, 0, 0);
    }
}

template <int NB, int LAYOUT, int PS>
__device__ inline void mfma_commutator32(const double2* sH, const double2* sR, int kk, int lane, bool tri, flt4& cre,
                              flt4& cim)
{
    const int i = lane & 15, q = lane >> 4;
    cre = flt4{0.f, 0.f, 0.f, 0.f};
    cim = flt4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int c = 4 * s + q;
        if (4 * s >= NB) break;
        const double2 hA = lds_h<NB, LAYOUT, PS>(sH, i, c, kk, tri), rA = lds_elem<NB, LAYOUT, PS>(sR, i, c, kk);
        const double2 hB = lds_h<NB, LAYOUT, PS>(sH, c, i, kk, tri), rB = lds_elem<NB, LAYOUT, PS>(sR, c, i, kk);
        const float hAx = float(hA.x), hAy = float(hA.y), rAx = float(rA.x), rAy = float(rA.y);
        const float hBx = float(hB.x), hBy = float(hB.y), rBx = float(rB.x), rBy = float(rB.y);
        cre = __builtin_amdgcn_mfma_f32_16x16x4f32(hAx, rBx, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f32_16x16x4f32(-hAy, rBy, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f32_16x16x4f32(-rAx, hBx, cre, 0, 0, 0);
        cre = __builtin_amdgcn_mfma_f32_16x16x4f32(rAy, hBy, cre, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_16x16x4f32(hAx, rBy, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_16x16x4f32(hAy, rBx, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_16x16x4f32(-rAx, hBy, cim, 0, 0, 0);
        cim = __builtin_amdgcn_mfma_f32_16x16x4f32(-rAy, hBx, cim, 0, 0, 0);
    }
}

// AB: the stage is an Adams-Bashforth step (its history is prefetched, so the
// register budget of the other stages does not carry four history arrays).
// x1 += g1 (rho - x2) for a thread's elements, rho read from the tile
template <int PER>
__device__ inline void fold_g1(double g1, const double2* sR, const int (&slot_of)[PER], const bool (&ok)[PER],
                               const double2 (&x2)[PER], double2 (&x1)[PER])
{
#pragma unroll
    for (int j = 0; j < PER; ++j)
        if (ok[j]) {
            const double2 r = sR[slot_of[j]];
            x1[j].x += g1 * (r.x - x2[j].x);
            x1[j].y += g1 * (r.y - x2[j].y);
        }
}

template <int NB, int LAYOUT, int KIND, int TILE, bool AB, bool TRI>
__global__ void __launch_bounds__((comm_threads<KIND, TILE>())) comm_kernel(CommArgs a)
{
    static_assert(!TRI || (LAYOUT == kFull && KIND != kCommScalar), "upper-triangle fields: FULL, MFMA engines");
    constexpr int NT = comm_threads<KIND, TILE>();
    constexpr int NPL = npl_of<NB, LAYOUT>();
    constexpr int TK = tile_k<NB, LAYOUT, KIND, TILE>();
    constexpr int PS = plane_stride<KIND, TK>();
    constexpr int PER = (NPL * TK + NT - 1) / NT;  // elements per thread
    constexpr int NSLOT = tile_slots<NB, LAYOUT, KIND, TILE>();
    extern __shared__ double2 lds[];
    double2* sR = lds;
    double2* sH = lds + NSLOT;
    double2* sHd = sH + h_tile_slots<NB, LAYOUT, KIND, TILE, TRI>();
    double2* sMon = sHd + kMaxNb;  // (x, y) = (trace, Hermiticity) maximum of wave w
    static_assert(NT / 64 <= kMonRedSlots, "one monitor slot per wave");
    const int Nk = a.Nk;
    // each XCD takes a contiguous range of tiles, so that tiles sharing a
    // cache line (TK = 4) meet in the same L2; t covers the tiles of instance
    // 0, then instance 1, ...
    const int ntile = (Nk + TK - 1) / TK;
    const int t = xcd_contiguous_block();
    const int inst = t / ntile;
    const int k0 = (t - inst * ntile) * TK;
    const long ioff = long(inst) * NPL * Nk;
    double Ev[3];
#pragma unroll
    for (int dd = 0; dd < 3; ++dd) Ev[dd] = a.E_tab ? a.E_tab[inst * 3 + dd] : a.E[dd];
    const StageUpdate& u = a.up;

    // Load phase.  Element (p, kk) of the tile is owned by one thread, which
    // forms its H, later produces it and applies its update.
    //  1. rho and H0 go straight from HBM into the LDS tiles
    //     (global_load_lds_dwordx4: lane l of a wave writes 16 bytes at the
    //     wave-uniform base + 16 l, so one wave instruction fills 64
    //     consecutive slots; slot q holds element (q / PS, q % PS), the
    //     padding column kk = TK is not loaded).  No registers are held.
    //  2. The addends of H (E_d xi_d for the fields present, Sigma) come into
    //     registers two arrays at a time, issued with step 1; after one wait
    //     and a barrier every thread adds them to its elements of the H tile.
    //  3. The operands of the element's update (D, the stage arrays, the
    //     Adams-Bashforth history) are issued next and stay in flight through
    //     the commutator.
    // x1v holds the first stage operand (RK: x1, AB: y); ex holds the second
    // RK operand x2 in ex[0], or the AB history f_{n-1..n-4}.
    constexpr bool ab = AB;
    constexpr int NEX = AB ? 4 : 1;
    // the FP32 engine reads H0 and xi stored in FP32 (half the bytes) into
    // registers instead of the direct H0 load, when it is given them (not
    // with exchange: there H0 holds -Sigma[rho0], which cancels against
    // Sigma[rho] and must not be rounded first)
    const bool F32H = KIND == kCommMfma32 && a.n_addf > 0;
    // H0 and xi as upper-triangle planes (FULL layout, Hermitian fields): the
    // H tile then holds the upper triangle only, in HERMITIAN slot order, and
    // the products read the lower elements as conjugates; H0 and every
    // addend are loaded once per element pair
    constexpr int NPLU = NB * (NB + 1) / 2;
    // (a separate instantiation, TRI, so that the other kernels keep their
    // registers; the MFMA engines only: the FMA engine reads FULL fields)
    constexpr bool tri = TRI;
    const int nh = (tri ? NPLU : NPL) * TK;  // elements of the H tile
    const long dh = tri ? long(inst) * (NPLU - NPL) * Nk : 0;  // H0 / xi offset minus the state's
    const double2* x1p = ab ? u.y : u.x1;
    const bool ld_x1 = (ab || u.out1) && x1p != a.rho;
    const bool ld_x2 = !ab && u.out2 && u.x2 != a.rho;
    const double2 zero = make_double2(0., 0.);
    {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int c = wave; c < NSLOT / 64; c += NT / 64) {
            const int q = c * 64 + lane;
            const int p = q / PS;
            const int kk = q - p * PS;
            if (kk < TK && p < NPL && k0 + kk < Nk) {
                const long o = ioff + long(p) * Nk + (k0 + kk);
                __builtin_amdgcn_global_load_lds(a.rho + o, sR + c * 64, 16, 0, 0);
                if (!F32H && p * TK < nh) __builtin_amdgcn_global_load_lds(a.H0 + o + dh, sH + c * 64, 16, 0, 0);
            }
        }
    }
    // element idx = threadIdx.x + j NT of the tile: plane idx / TK, k-point
    // idx % TK, the same slot in the rho and H tiles; in the H tile (and in
    // H0 / xi, at off + dh) plane p is the upper plane p when tri
    long off[PER];
    bool in_tile[PER], ok[PER], h_in[PER], h_ok[PER];
    int slot_of[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int idx = threadIdx.x + j * NT;
        const int p = idx / TK;
        const int kk = idx - p * TK;
        in_tile[j] = idx < NPL * TK;
        ok[j] = in_tile[j] && k0 + kk < Nk;
        h_in[j] = idx < nh;
        h_ok[j] = h_in[j] && k0 + kk < Nk;
        off[j] = ioff + long(p) * Nk + (k0 + kk);
        slot_of[j] = p * PS + kk;
    }
    // addend coefficients: E_d of the instance for xi_d, 1 for Sigma
    double cf[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int dd = a.add_dd[r];
        cf[r] = dd == 0 ? Ev[0] : dd == 1 ? Ev[1] : dd == 2 ? Ev[2] : 1.0;
    }
    double2 dv[PER], x1v[PER], ex[NEX][PER];
    {
        if (F32H) {
            // H = H0 + sum_d E_d xi_d from the FP32 fields (addf[0] = H0,
            // coefficient 1; then the xi present), widened to FP64 in the tile
            double cff[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int dd = a.addf_dd[r];
                cff[r] = dd == 0 ? Ev[0] : dd == 1 ? Ev[1] : dd == 2 ? Ev[2] : 1.0;
            }
            float2 fa[PER], fb[PER];
            const float2 fz = make_float2(0.f, 0.f);
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                fa[j] = fb[j] = fz;
                if (h_ok[j]) fa[j] = a.addf[0][off[j] + dh];
            }
            if (a.n_addf > 1) {
#pragma unroll
                for (int j = 0; j < PER; ++j)
                    if (h_ok[j]) fb[j] = a.addf[1][off[j] + dh];
            }
            if (a.hsum && threadIdx.x < NB)
                sHd[threadIdx.x] = hartree_shift(a.hsum + long(inst) * NPL, a.hart + long(inst) * NB * NB, threadIdx.x,
                                                 NB, LAYOUT, a.inv_nk);
            __builtin_amdgcn_s_waitcnt(0);  // this wave's direct loads are in LDS
            __syncthreads();                // and every other wave's
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                if (!in_tile[j]) continue;
                const int sl = slot_of[j];
                if (h_in[j]) {
                    double2 h = make_double2(double(fa[j].x), double(fa[j].y));
                    if (a.n_addf > 1) {
                        h.x += cff[1] * double(fb[j].x);
                        h.y += cff[1] * double(fb[j].y);
                    }
                    sH[sl] = h;
                }
                x1v[j] = zero;
                ex[0][j] = zero;
                if (!ld_x1) x1v[j] = sR[sl];
                if (!ab && !ld_x2) ex[0][j] = sR[sl];
            }
            if (a.n_addf > 2 || a.sigma) {  // the other xi and Sigma: a second round
                double2 sg[PER];
#pragma unroll
                for (int j = 0; j < PER; ++j) {
                    fa[j] = fb[j] = fz;
                    sg[j] = zero;
                    if (h_ok[j] && a.n_addf > 2) fa[j] = a.addf[2][off[j] + dh];
                    if (h_ok[j] && a.n_addf > 3) fb[j] = a.addf[3][off[j] + dh];
                    if (ok[j] && a.sigma) sg[j] = a.sigma[off[j]];
                }
#pragma unroll
                for (int j = 0; j < PER; ++j) {
                    if (!h_in[j]) continue;
                    const int sl = slot_of[j];
                    double2 h = sH[sl];
                    h.x += cff[2] * double(fa[j].x) + cff[3] * double(fb[j].x) + sg[j].x;
                    h.y += cff[2] * double(fa[j].y) + cff[3] * double(fb[j].y) + sg[j].y;
                    sH[sl] = h;
                }
            }
        } else {
            double2 xa[PER], xb[PER];
    #pragma unroll
            for (int j = 0; j < PER; ++j) xa[j] = xb[j] = zero;
            if (a.n_add > 0) {
    #pragma unroll
                for (int j = 0; j < PER; ++j)
                    if (h_ok[j]) xa[j] = a.add[0][off[j] + dh];
            }
            if (a.n_add > 1) {
    #pragma unroll
                for (int j = 0; j < PER; ++j)
                    if (h_ok[j]) xb[j] = a.add[1][off[j] + dh];
            }
            if (a.hsum && threadIdx.x < NB)
                sHd[threadIdx.x] = hartree_shift(a.hsum + long(inst) * NPL, a.hart + long(inst) * NB * NB, threadIdx.x,
                                                 NB, LAYOUT, a.inv_nk);
            __builtin_amdgcn_s_waitcnt(0);  // this wave's direct loads are in LDS
            __syncthreads();                // and every other wave's
            // H = H0 + sum of the addends (in the order xi_0, xi_1, xi_2, Sigma)
            // + Hartree diagonal, in place in the H tile
    #pragma unroll
            for (int j = 0; j < PER; ++j) {
                if (!in_tile[j]) continue;
                const int sl = slot_of[j];
                if (h_in[j]) {
                    double2 h = sH[sl];
                    if (a.n_add > 0) {
                        h.x += cf[0] * xa[j].x;
                        h.y += cf[0] * xa[j].y;
                    }
                    if (a.n_add > 1) {
                        h.x += cf[1] * xb[j].x;
                        h.y += cf[1] * xb[j].y;
                    }
                    sH[sl] = h;
                }
                // stage operands that are the stage input itself
                // (if/else rather than ?: on double2 keeps the values out of scratch)
                x1v[j] = zero;
                ex[0][j] = zero;
                if (!ld_x1) x1v[j] = sR[sl];
                if (!ab && !ld_x2) ex[0][j] = sR[sl];
            }
            if (a.n_add > 2) {  // third and fourth addend: a second round
    #pragma unroll
                for (int j = 0; j < PER; ++j) {
                    xa[j] = xb[j] = zero;
                    if (h_ok[j]) xa[j] = a.add[2][off[j] + dh];
                }
                if (a.n_add > 3) {
    #pragma unroll
                    for (int j = 0; j < PER; ++j)
                        if (h_ok[j]) xb[j] = a.add[3][off[j] + dh];
                }
    #pragma unroll
                for (int j = 0; j < PER; ++j) {
                    if (!h_in[j]) continue;
                    const int sl = slot_of[j];
                    double2 h = sH[sl];
                    h.x += cf[2] * xa[j].x;
                    h.y += cf[2] * xa[j].y;
                    if (a.n_add > 3) {
                        h.x += cf[3] * xb[j].x;
                        h.y += cf[3] * xb[j].y;
                    }
                    sH[sl] = h;
                }
            }
        }
        if (!TRI && a.hsum) {  // (exchange only, never with upper-triangle fields)
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                if (!ok[j]) continue;
                int n, m;
                decode_plane<NB, LAYOUT>((threadIdx.x + j * NT) / TK, n, m);
                if (n == m) sH[slot_of[j]] = cadd(sH[slot_of[j]], sHd[n]);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        dv[j] = zero;
#pragma unroll
        for (int e = 1; e < NEX; ++e) ex[e][j] = zero;
    }
    if (a.D) {
#pragma unroll
        for (int j = 0; j < PER; ++j)
            if (ok[j]) dv[j] = a.D[off[j]];
    }
    if (ld_x1) {
#pragma unroll
        for (int j = 0; j < PER; ++j)
            if (ok[j]) x1v[j] = x1p[off[j]];
    }
    if (ld_x2) {
#pragma unroll
        for (int j = 0; j < PER; ++j)
            if (ok[j]) ex[0][j] = u.x2[off[j]];
    }
    if constexpr (AB) {
#pragma unroll
        for (int i = 1; i < 5; ++i)  // constant indices: the pointer table stays in SGPRs
            if (i < u.ab_order) {
#pragma unroll
                for (int j = 0; j < PER; ++j)
                    if (ok[j]) ex[i - 1][j] = u.hist[i][off[j]];
            }
    }
    if (u.mon) {
        // Trace and Hermiticity of the stage input (the state the step starts
        // from: RK4 stage 1, every AB step) over the tile's TK k-points, read
        // from the rho tile while the update operands are in flight; each
        // wave leaves its maxima in sMon, which thread 0 folds into the
        // monitor slots after a later barrier
        double mt = 0., mh = 0.;
        // trace and diagonal imaginary part: one thread per k-point of the tile
        for (int kk = threadIdx.x; kk < TK; kk += NT) {
            const int k = k0 + kk;
            if (k >= Nk) continue;
            double2 tr = make_double2(-u.nf, 0.);
            for (int n = 0; n < NB; ++n) {
                bool c;
                const double2 v = sR[slot<NB, LAYOUT>(n, n, c) * PS + kk];
                tr = cadd(tr, v);
                mh = mon_max(mh, 2.0 * fabs(v.y));
            }
            const double tt = sqrt(tr.x * tr.x + tr.y * tr.y);
            mt = mon_max(mt, tt);
        }
        // |rho_nm - conj(rho_mn)| over all pairs n < m, by the owner of (n, m)
        if constexpr (LAYOUT == kFull) {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int idx = threadIdx.x + j * NT;
                const int p = idx / TK;
                const int kk = idx - p * TK;
                const int n = p / NB, m = p - (p / NB) * NB;
                if (!ok[j] || n >= m) continue;
                const double2 x = sR[slot_of[j]];
                const double2 z = sR[(m * NB + n) * PS + kk];
                const double d = sqrt((x.x - z.x) * (x.x - z.x) + (x.y + z.y) * (x.y + z.y));
                mh = mon_max(mh, d);
            }
        }
        // across the wave by shuffles (no NaN is left: mon_max made it +inf)
#pragma unroll
        for (int w = 32; w > 0; w >>= 1) {
            mt = fmax(mt, __shfl_xor(mt, w));
            mh = fmax(mh, __shfl_xor(mh, w));
        }
        if ((threadIdx.x & 63) == 0) sMon[threadIdx.x >> 6] = make_double2(mt, mh);
    }
    __syncthreads();

    if constexpr (KIND != kCommScalar) {
        // every wave takes TK/4 k-points of the tile; the products wait in
        // registers until all waves are done reading rho and H, then replace
        // rho in LDS
        constexpr int KPW = TK / (NT / 64);
        static_assert(KPW >= 1 && KPW * (NT / 64) == TK, "every wave takes the same number of k-points");
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        using acc_t = typename std::conditional<KIND == kCommMfma64, dbl4, flt4>::type;
        acc_t cre[KPW], cim[KPW];
#pragma unroll
        for (int w = 0; w < KPW; ++w) {
            const int kk = wave + (NT / 64) * w;
            if constexpr (KIND == kCommMfma64)
                mfma_commutator64<NB, LAYOUT, PS>(sH, sR, kk, lane, tri, cre[w], cim[w]);
            else
                mfma_commutator32<NB, LAYOUT, PS>(sH, sR, kk, lane, tri, cre[w], cim[w]);
        }
        if (!ab && u.g1 != 0.0) fold_g1(u.g1, sR, slot_of, ok, ex[0], x1v);  // rho before the products replace it
        __syncthreads();
        const int i = lane & 15, q = lane >> 4;
#pragma unroll
        for (int w = 0; w < KPW; ++w) {
            const int kk = wave + (NT / 64) * w;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // result rows: f64 16x16x4 puts row q + 4v in slot v, f32 16x16x4 row 4q + v
                const int n = KIND == kCommMfma64 ? q + 4 * v : 4 * q + v, m = i;
                if (n < NB && m < NB && (LAYOUT == kFull || n <= m)) {
                    bool cj;
                    const int sl = slot<NB, LAYOUT>(n, m, cj);
                    sR[sl * PS + kk] = make_double2(double(cre[w][v]), double(cim[w][v]));
                }
            }
        }
        __syncthreads();
    }

    if constexpr (KIND == kCommScalar && !AB)  // rho stays in the tile
        if (u.g1 != 0.0) fold_g1(u.g1, sR, slot_of, ok, ex[0], x1v);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        if (!ok[j]) continue;
        const int o = threadIdx.x + j * NT;
        const int p = o / TK;
        const int kk = o - p * TK;
        double2 acc;
        if constexpr (KIND == kCommScalar) {
            int n, m;
            decode_plane<NB, LAYOUT>(p, n, m);
            acc = make_double2(0., 0.);
#pragma unroll
            for (int l = 0; l < NB; ++l) {
                bool cnl, clm;
                const int inl = slot<NB, LAYOUT>(n, l, cnl);
                const int ilm = slot<NB, LAYOUT>(l, m, clm);
                double2 Hnl = sH[inl * PS + kk], Rnl = sR[inl * PS + kk];
                double2 Rlm = sR[ilm * PS + kk], Hlm = sH[ilm * PS + kk];
                if (cnl) { Hnl.y = -Hnl.y; Rnl.y = -Rnl.y; }
                if (clm) { Rlm.y = -Rlm.y; Hlm.y = -Hlm.y; }
                acc.x += Hnl.x * Rlm.x - Hnl.y * Rlm.y - (Rnl.x * Hlm.x - Rnl.y * Hlm.y);
                acc.y += Hnl.x * Rlm.y + Hnl.y * Rlm.x - (Rnl.x * Hlm.y + Rnl.y * Hlm.x);
            }
        } else {
            acc = sR[p * PS + kk];
        }
        const long og = off[j];
        // f = D - i * [H, rho]
        const double2 f = make_double2(dv[j].x + acc.y, dv[j].y - acc.x);

        if (u.out_f) u.out_f[og] = f;
        if constexpr (AB) {
            // y <- y + dt beta_0 f, then + dt beta_i f_{n-i} in the reference's order
            double2 yv = x1v[j];
            yv.x += u.ab_c[0] * f.x;
            yv.y += u.ab_c[0] * f.y;
#pragma unroll
            for (int i = 1; i < 5; ++i) {
                if (i < u.ab_order) {
                    yv.x += u.ab_c[i] * ex[i - 1][j].x;
                    yv.y += u.ab_c[i] * ex[i - 1][j].y;
                }
            }
            u.y[og] = yv;
            x1v[j] = yv;  // kept for the monitor
        } else {
            if (u.out1) {
                const double2 o1 = make_double2(x1v[j].x + u.b1 * f.x, x1v[j].y + u.b1 * f.y);
                u.out1[og] = o1;
                x1v[j] = o1;  // kept for the monitor
            }
            if (u.out2) u.out2[og] = make_double2(ex[0][j].x + u.b2 * f.x, ex[0][j].y + u.b2 * f.y);
        }
    }

    if (u.mon) {
        // every wave wrote its maxima before the barrier ahead of the products
        if (threadIdx.x == 0) {
            double mt = 0., mh = 0.;
            for (int w = 0; w < NT / 64; ++w) {
                mt = fmax(mt, sMon[w].x);
                mh = fmax(mh, sMon[w].y);
            }
            mon_slot(u.mon, blockIdx.x, mt, mh);
        }
    }
}

template <int NB, int LAYOUT, int KIND, int TILE, bool AB, bool TRI>
void launch_one(const CommArgs& a, hipStream_t s)
{
    constexpr int TK = tile_k<NB, LAYOUT, KIND, TILE>();
    constexpr size_t lds = comm_lds_bytes<NB, LAYOUT, KIND, TILE, TRI>();
    auto kern = &comm_kernel<NB, LAYOUT, KIND, TILE, AB, TRI>;
    static std::once_flag raised;  // contexts may be created from several host threads
    std::call_once(raised, [&] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(lds));
    });
    if (g_attributes_only) return;
    hipLaunchKernelGGL(kern, dim3(((a.Nk + TK - 1) / TK) * std::max(a.B, 1)), dim3(comm_threads<KIND, TILE>()), lds,
                       s, a);
}

template <int NB, int LAYOUT, int KIND, int TILE, bool TRI>
void launch_ab(const CommArgs& a, hipStream_t s)
{
    if (a.up.ab_order > 0)
        launch_one<NB, LAYOUT, KIND, TILE, true, TRI>(a, s);
    else
        launch_one<NB, LAYOUT, KIND, TILE, false, TRI>(a, s);
}

template <int NB, int LAYOUT, int KIND, int TILE>
void launch_kind(const CommArgs& a, hipStream_t s)
{
    if constexpr (LAYOUT == kFull && KIND != kCommScalar) {
        if (a.tri) {
            launch_ab<NB, LAYOUT, KIND, TILE, true>(a, s);
            return;
        }
    }
    launch_ab<NB, LAYOUT, KIND, TILE, false>(a, s);
}

template <int NB, int LAYOUT>
void launch_layout(const CommArgs& a, int kind, int tile, hipStream_t s)
{
    if (kind == kCommMfma64) {
        if (tile == 1)
            launch_kind<NB, LAYOUT, kCommMfma64, 1>(a, s);
        else if (tile == 2)
            launch_kind<NB, LAYOUT, kCommMfma64, 2>(a, s);
        else
            launch_kind<NB, LAYOUT, kCommMfma64, 0>(a, s);
    } else if (kind == kCommMfma32) {
        if (tile == 1)
            launch_kind<NB, LAYOUT, kCommMfma32, 1>(a, s);
        else if (tile == 2)
            launch_kind<NB, LAYOUT, kCommMfma32, 2>(a, s);
        else
            launch_kind<NB, LAYOUT, kCommMfma32, 0>(a, s);
    } else {
        launch_kind<NB, LAYOUT, kCommScalar, 0>(a, s);
    }
}

template <int NB>
void launch_nb(const CommArgs& a, int layout, int kind, int tile, hipStream_t s)
{
    if (layout == kFull)
        launch_layout<NB, kFull>(a, kind, tile, s);
    else
        launch_layout<NB, kHermitian>(a, kind, tile, s);
}

}  // namespace

void launch_comm(int nb, int layout, int kind, int tile, const CommArgs& args, hipStream_t s)
{
    CommArgs a = args;
    a.n_add = 0;
    a.n_addf = 0;
    for (int r = 0; r < 4; ++r) {
        a.add[r] = nullptr;
        a.add_dd[r] = -1;
        a.addf[r] = nullptr;
        a.addf_dd[r] = -1;
    }
    if (args.tri && (layout != kFull || args.sigma || args.hsum || kind == kCommScalar))
        throw std::runtime_error("upper-triangle H0 and xi: FULL layout without exchange, MFMA engines only");
    if (args.hsum && !args.hart) throw std::runtime_error("the Hartree shift needs V with the plane sums");
    if (kind == kCommMfma32 && args.H0f) {
        if (args.sigma) throw std::runtime_error("FP32 H0 with exchange: H0 holds -Sigma[rho0] and stays FP64");
        a.addf[a.n_addf++] = args.H0f;
        for (int d = 0; d < 3; ++d)
            if (args.xif[d]) {
                a.addf[a.n_addf] = args.xif[d];
                a.addf_dd[a.n_addf++] = d;
            }
    } else {
        for (int d = 0; d < 3; ++d)
            if (args.xi[d]) {
                a.add[a.n_add] = args.xi[d];
                a.add_dd[a.n_add++] = d;
            }
        if (args.sigma) a.add[a.n_add++] = args.sigma;
    }
    switch (nb) {
        case 1: launch_nb<1>(a, layout, kind, tile, s); break;
        case 2: launch_nb<2>(a, layout, kind, tile, s); break;
        case 3: launch_nb<3>(a, layout, kind, tile, s); break;
        case 4: launch_nb<4>(a, layout, kind, tile, s); break;
        case 5: launch_nb<5>(a, layout, kind, tile, s); break;
        case 6: launch_nb<6>(a, layout, kind, tile, s); break;
        case 7: launch_nb<7>(a, layout, kind, tile, s); break;
        case 8: launch_nb<8>(a, layout, kind, tile, s); break;
        case 9: launch_nb<9>(a, layout, kind, tile, s); break;
        case 10: launch_nb<10>(a, layout, kind, tile, s); break;
        case 11: launch_nb<11>(a, layout, kind, tile, s); break;
        case 12: launch_nb<12>(a, layout, kind, tile, s); break;
        case 13: launch_nb<13>(a, layout, kind, tile, s); break;
        case 14: launch_nb<14>(a, layout, kind, tile, s); break;
        case 15: launch_nb<15>(a, layout, kind, tile, s); break;
        case 16: launch_nb<16>(a, layout, kind, tile, s); break;
        default: break;  // rejected at context creation
    }
}

}  // namespace edus
