// Fused commutator + integrator-stage kernel (FP64).
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
// LDS and each thread produces output elements
//     f_nm = D_nm - i sum_l (H_nl rho_lm - rho_nl H_lm).
// LAYOUT = kFull keeps every element; kHermitian keeps the upper triangle and
// expands the lower one by conjugation while reading LDS.
#include "common.hpp"
#include "kernels.hpp"

namespace edus {

namespace {

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
constexpr int tile_for()
{
    // keep the two LDS tiles (rho and H) of NPL*TK complex doubles under ~64 KB
    const int npl = LAYOUT == kFull ? NB * NB : NB * (NB + 1) / 2;
    int tk = 64;
    while (tk > 4 && 2 * npl * tk * 16 > 65536) tk /= 2;
    return tk;
}

template <int NB, int TK, int LAYOUT>
__global__ void __launch_bounds__(256) comm_kernel(CommArgs a)
{
    constexpr int NPL = LAYOUT == kFull ? NB * NB : NB * (NB + 1) / 2;
    constexpr int NT = 256;                        // launch block size
    constexpr int PER = (NPL * TK + NT - 1) / NT;  // elements per thread
    __shared__ double2 sR[NPL * TK];
    __shared__ double2 sH[NPL * TK];
    const int k0 = blockIdx.x * TK;
    const int Nk = a.Nk;
    const StageUpdate& u = a.up;

    // Load phase.  Element (p, kk) is loaded and later produced by the same
    // thread, so the operands of its update (D and the stage arrays) are
    // fetched here too: every global load of the tile is in flight before the
    // first barrier.
    const bool x1_in = (u.x1 == a.rho), x2_in = (u.x2 == a.rho), y_in = (u.y == a.rho);
    double2 dv[PER], x1v[PER], x2v[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int idx = threadIdx.x + j * NT;
        if (idx >= NPL * TK) break;
        const int p = idx / TK;
        const int kk = idx - p * TK;
        const int k = k0 + kk;
        double2 r = make_double2(0., 0.), h = r, d = r, x1 = r, x2 = r;
        if (k < Nk) {
            const long o = long(p) * Nk + k;
            r = a.rho[o];
            h = a.H0[o];
#pragma unroll
            for (int dd = 0; dd < 3; ++dd)
                if (a.xi[dd]) {
                    const double2 x = a.xi[dd][o];
                    h.x += a.E[dd] * x.x;
                    h.y += a.E[dd] * x.y;
                }
            if (a.sigma) h = cadd(h, a.sigma[o]);
            if (a.hdiag) {
                int n, m;
                decode_plane<NB, LAYOUT>(p, n, m);
                if (n == m) h = cadd(h, a.hdiag[n]);
            }
            if (a.D) d = a.D[o];
            if (u.out1) x1 = x1_in ? r : u.x1[o];
            if (u.out2) x2 = x2_in ? r : u.x2[o];
            if (u.ab_order) x1 = y_in ? r : u.y[o];
        }
        sR[idx] = r;
        sH[idx] = h;
        dv[j] = d;
        x1v[j] = x1;
        x2v[j] = x2;
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int o = threadIdx.x + j * NT;
        if (o >= NPL * TK) break;
        const int p = o / TK;
        const int kk = o - p * TK;
        const int k = k0 + kk;
        if (k >= Nk) continue;
        int n, m;
        decode_plane<NB, LAYOUT>(p, n, m);
        double2 acc = make_double2(0., 0.);
#pragma unroll
        for (int l = 0; l < NB; ++l) {
            bool cnl, clm;
            const int inl = slot<NB, LAYOUT>(n, l, cnl);
            const int ilm = slot<NB, LAYOUT>(l, m, clm);
            double2 Hnl = sH[inl * TK + kk], Rnl = sR[inl * TK + kk];
            double2 Rlm = sR[ilm * TK + kk], Hlm = sH[ilm * TK + kk];
            if (cnl) { Hnl.y = -Hnl.y; Rnl.y = -Rnl.y; }
            if (clm) { Rlm.y = -Rlm.y; Hlm.y = -Hlm.y; }
            acc.x += Hnl.x * Rlm.x - Hnl.y * Rlm.y - (Rnl.x * Hlm.x - Rnl.y * Hlm.y);
            acc.y += Hnl.x * Rlm.y + Hnl.y * Rlm.x - (Rnl.x * Hlm.y + Rnl.y * Hlm.x);
        }
        const long og = long(p) * Nk + k;
        // f = D - i * acc
        const double2 f = make_double2(dv[j].x + acc.y, dv[j].y - acc.x);

        if (u.out_f) u.out_f[og] = f;
        if (u.out1) u.out1[og] = make_double2(x1v[j].x + u.b1 * f.x, x1v[j].y + u.b1 * f.y);
        if (u.out2) u.out2[og] = make_double2(x2v[j].x + u.b2 * f.x, x2v[j].y + u.b2 * f.y);
        if (u.ab_order) {
            // y <- y + dt beta_0 f, then + dt beta_i f_{n-i} in the reference's order
            double2 yv = x1v[j];
            yv.x += u.ab_c[0] * f.x;
            yv.y += u.ab_c[0] * f.y;
            for (int i = 1; i < u.ab_order; ++i) {
                const double2 h = u.hist[i][og];
                yv.x += u.ab_c[i] * h.x;
                yv.y += u.ab_c[i] * h.y;
            }
            u.y[og] = yv;
        }
    }

    if (u.mon) {
        // The block's new state is in global memory (visible to the block after
        // the barrier, mostly still in cache): check trace and Hermiticity of
        // its TK k-points without another pass over HBM.
        constexpr int NT = 256;  // launch block size
        __shared__ double sred[2 * NT];
        __syncthreads();
        const double2* ys = u.mon_src;
        double mt = 0., mh = 0.;
        // trace and diagonal imaginary part: one thread per k-point of the tile
        for (int kk = threadIdx.x; kk < TK; kk += NT) {
            const int k = k0 + kk;
            if (k >= Nk) continue;
            double2 tr = make_double2(-u.nf, 0.);
            for (int n = 0; n < NB; ++n) {
                bool c;
                const double2 v = ys[long(slot<NB, LAYOUT>(n, n, c)) * Nk + k];
                tr = cadd(tr, v);
                mh = fmax(mh, 2.0 * fabs(v.y));
            }
            const double t = sqrt(tr.x * tr.x + tr.y * tr.y);
            mt = (t != t) ? __builtin_huge_val() : fmax(mt, t);  // NaN trips the monitor
        }
        // |rho_nm - conj(rho_mn)| over all pairs n < m, spread over the block
        if (LAYOUT == kFull) {
            constexpr int NPAIR = NB * (NB - 1) / 2;
            for (int item = threadIdx.x; item < TK * NPAIR; item += NT) {
                const int q = item / TK;
                const int kk = item - q * TK;
                const int k = k0 + kk;
                if (k >= Nk) continue;
                int n = 0, r = q;
                while (r >= NB - 1 - n) {
                    r -= NB - 1 - n;
                    ++n;
                }
                const int m = n + 1 + r;
                const double2 x = ys[long(n * NB + m) * Nk + k];
                const double2 z = ys[long(m * NB + n) * Nk + k];
                const double d = sqrt((x.x - z.x) * (x.x - z.x) + (x.y + z.y) * (x.y + z.y));
                mh = (d != d) ? __builtin_huge_val() : fmax(mh, d);
            }
        }
        sred[threadIdx.x] = mt;
        sred[NT + threadIdx.x] = mh;
        __syncthreads();
        for (int st = NT / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st) {
                sred[threadIdx.x] = fmax(sred[threadIdx.x], sred[threadIdx.x + st]);
                sred[NT + threadIdx.x] = fmax(sred[NT + threadIdx.x], sred[NT + threadIdx.x + st]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            atomicMax(&u.mon[0], static_cast<unsigned long long>(__double_as_longlong(sred[0])));
            atomicMax(&u.mon[1], static_cast<unsigned long long>(__double_as_longlong(sred[NT])));
        }
    }
}

template <int NB>
void launch_nb(const CommArgs& a, int layout, hipStream_t s)
{
    if (layout == kFull) {
        constexpr int TK = tile_for<NB, kFull>();
        hipLaunchKernelGGL((comm_kernel<NB, TK, kFull>), dim3((a.Nk + TK - 1) / TK), dim3(256), 0, s, a);
    } else {
        constexpr int TK = tile_for<NB, kHermitian>();
        hipLaunchKernelGGL((comm_kernel<NB, TK, kHermitian>), dim3((a.Nk + TK - 1) / TK), dim3(256), 0, s, a);
    }
}

}  // namespace

void launch_comm(int nb, int layout, const CommArgs& a, hipStream_t s)
{
    switch (nb) {
        case 1: launch_nb<1>(a, layout, s); break;
        case 2: launch_nb<2>(a, layout, s); break;
        case 3: launch_nb<3>(a, layout, s); break;
        case 4: launch_nb<4>(a, layout, s); break;
        case 5: launch_nb<5>(a, layout, s); break;
        case 6: launch_nb<6>(a, layout, s); break;
        case 7: launch_nb<7>(a, layout, s); break;
        case 8: launch_nb<8>(a, layout, s); break;
        case 9: launch_nb<9>(a, layout, s); break;
        case 10: launch_nb<10>(a, layout, s); break;
        case 11: launch_nb<11>(a, layout, s); break;
        case 12: launch_nb<12>(a, layout, s); break;
        case 13: launch_nb<13>(a, layout, s); break;
        case 14: launch_nb<14>(a, layout, s); break;
        case 15: launch_nb<15>(a, layout, s); break;
        case 16: launch_nb<16>(a, layout, s); break;
        default: break;  // rejected at context creation
    }
}

}  // namespace edus
