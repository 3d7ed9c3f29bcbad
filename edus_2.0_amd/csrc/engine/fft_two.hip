// Batched 1-D FFTs along one grid axis as two in-register steps around one
// LDS exchange ("two-step" kernel; same modes and results as the Stockham
// kernel in fft_axis.hip, which remains for line lengths not listed here).
//
// A line of length n = RA * RB, index j = ja*RB + jb, frequency
// k = ka + RA*kb:
//   step A  thread (line, jb) holds x[ja*RB + jb], ja < RA, does an RA-point
//           DFT in registers and multiplies by w_n^(jb*ka);
//   LDS     one exchange;
//   step B  thread (line, ka) holds the RB values of row ka, does an RB-point
//           DFT in registers: X[ka + RA*kb].
// The derivative sandwich (mode 0) multiplies by i coef R(k) and runs the
// inverse steps B^-1 (same thread, no exchange), LDS, A^-1, so one line costs
// two barriers and two LDS round trips.  Mode 3 multiplies by a real field
// instead (the Hartree-Fock convolution along this axis), mode 4 by
// scale exp(i coef R) (the Peierls phase along this axis, its input a sum
// in + sum_d c_d in_add_d), mode 1 stops after step B (forward transform),
// mode 2 starts at step B^-1 (inverse).  Mode 5 is mode 0 that also stores
// the forward transform after step B into out2, exactly as mode 1 stores it
// (the exchange term's first pass, taken from the same loads).
//
// Every thread owns at most one A task and one B task (L*max(RA,RB) <= NT,
// checked on the host), so each thread's global loads are all issued before
// the first arithmetic and the field loads of later phases (addend, multiplier)
// are issued one phase ahead.
#include <algorithm>
#include <mutex>

#include "common.hpp"
#include "fft_reg.hpp"
#include "kernels.hpp"

namespace edus {

namespace {

using reg::sfor;

constexpr int kTwoStepThreads = 256;

// v[j] = the kernel's input at global offset off(j), j < R (mode 4: a sum of
// fields; each field's R loads are issued together before any is used)
template <int MODE, int R, class Off>
__device__ inline void load_line(const AxisFFTArgs& a, double2 (&v)[R], Off off)
{
    sfor<R>([&](auto j_) {
        constexpr int j = decltype(j_)::value;
        v[j] = a.in[off(j)];
    });
    if constexpr (MODE == 4) {
        for (int d = 0; d < 3; ++d) {
            const double2* p = a.in_add[d];
            if (!p) continue;
            const double c = a.in_add_c[d];
            double2 w[R];
            sfor<R>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                w[j] = p[off(j)];
            });
            sfor<R>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                v[j].x += c * w[j].x;
                v[j].y += c * w[j].y;
            });
        }
    }
}

// mode 4 multiplier scale * exp(i coef R(j)) of frequency index j (line length n)
__device__ inline double2 phase(const AxisFFTArgs& a, int j, int n)
{
    if (a.ptab) return a.ptab[j];
    double sn, cs;
    sincos(a.coef * double(gamma_centred(j, n)), &sn, &cs);
    return make_double2(a.scale * cs, a.scale * sn);
}

template <int RA, int RB>
struct TwoStepShape {
    static constexpr int n = RA * RB;
    static constexpr int RBP = RB | 1;             // odd row stride in LDS
    static constexpr int LS = (RA * RBP) | 1;      // odd line stride in LDS
};

// EARLY: the addend of the last phase is loaded before the first exchange
// (latency hidden, more registers) or at the start of the last phase.
template <int RA, int RB, int NT, int MODE, bool EARLY>
__global__ void __launch_bounds__(NT) two_step_kernel(AxisFFTArgs a)
{
    using Sh = TwoStepShape<RA, RB>;
    constexpr int n = Sh::n, RBP = Sh::RBP, LS = Sh::LS;
    extern __shared__ double2 lds[];
    double2* twl = lds;      // w_n^e, e < n
    double2* buf = lds + n;  // L lines of LS values
    const int L = a.L;
    const int Nk = a.N0 * a.N1;
    const int tid = threadIdx.x;

    const int tb = xcd_contiguous_block();  // neighbouring line groups side by side in one XCD
    long base, cstride, rowstride;
    long row0 = 0, plane0 = 0;  // first global row (axis 1), plane (axis 0)
    int col0 = 0;               // first column (axis 0)
    int nlines;
    if (a.axis == 1) {
        row0 = long(tb) * L;
        nlines = int(min(long(L), long(a.P) * a.N0 - row0));
        base = row0 * a.N1;
        cstride = a.N1;
        rowstride = 1;
    } else {
        const int nblk = (a.N1 + L - 1) / L;
        plane0 = tb / nblk;
        const int first = int(tb - plane0 * nblk) * L;
        col0 = first;
        nlines = min(L, a.N1 - first);
        base = plane0 * Nk + first;
        cstride = 1;
        rowstride = a.N1;
    }
    // A task: (line, jb); B task: (line, ka).  Along rows the fastest thread
    // index runs inside a line, along columns across lines (adjacent columns
    // are adjacent in memory).
    int lineA, jb, lineB, ka;
    if (a.axis == 1) {
        lineA = tid / RB;
        jb = tid - lineA * RB;
        lineB = tid / RA;
        ka = tid - lineB * RA;
    } else {
        lineA = tid % L;
        jb = tid / L;
        lineB = lineA;
        ka = jb;
    }
    const bool doA = jb < RB && lineA < nlines;
    const bool doB = ka < RA && lineB < nlines;
    const long gA = base + lineA * cstride;
    const long gB = base + lineB * cstride;
    constexpr int mode = MODE;
    constexpr bool sandwich = mode == 0 || mode == 3 || mode == 4 || mode == 5;
    constexpr bool deriv = mode == 0 || mode == 5;

    if constexpr (RB == 1) {
        // the whole line in one thread: no exchange
        if (!doA) return;
        double2 v[RA];
        load_line<mode>(a, v, [&](int j) { return gA + long(j) * rowstride; });
        double mul[RA];
        if constexpr (mode == 3) {
            sfor<RA>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                mul[j] = a.mult[gA + long(j) * rowstride];
            });
        }
        if constexpr (mode == 2) {
            reg::rdft<RA, 1>(v);
        } else {
            reg::rdft<RA, -1>(v);
            if constexpr (mode == 5) {
                sfor<RA>([&](auto k_) {
                    constexpr int k = decltype(k_)::value;
                    a.out2[gA + long(k) * rowstride] = v[k];
                });
            }
            if constexpr (deriv) {
                const double coef = line_coef(a, a.axis == 1 ? (row0 + lineA) / a.N0 : plane0);
                sfor<RA>([&](auto k_) {
                    constexpr int k = decltype(k_)::value;
                    v[k] = cmul_i(v[k], coef * double(gamma_centred(k, RA)));
                });
            } else if constexpr (mode == 3) {
                if (a.r0 && a.axis == 0 && col0 + lineA == 0) a.r0[plane0] = v[0];
                sfor<RA>([&](auto k_) {
                    constexpr int k = decltype(k_)::value;
                    v[k] = cscale(v[k], a.scale * mul[k]);
                });
            } else if constexpr (mode == 4) {
                sfor<RA>([&](auto k_) {
                    constexpr int k = decltype(k_)::value;
                    v[k] = cmul(v[k], phase(a, k, RA));
                });
            }
            if constexpr (sandwich) reg::rdft<RA, 1>(v);
        }
        const double sc = sandwich ? 1.0 : a.scale;
        sfor<RA>([&](auto j_) {
            constexpr int j = decltype(j_)::value;
            double2 x = v[j];
            if constexpr (!sandwich) x = cscale(x, sc);
            if (a.addend) x = cadd(x, a.addend[gA + long(j) * rowstride]);
            a.out[gA + long(j) * rowstride] = x;
        });
        return;
    } else {
        double2* lb = buf + lineB * LS + ka * RBP;  // this B task's row in LDS
        double2* la = buf + lineA * LS + jb;        // this A task's column in LDS

        if constexpr (mode != 2) {
            // ---- load (A layout), step A forward
            double2 v[RA];
            if (doA) {
                load_line<mode>(a, v, [&](int j) { return gA + long(j * RB + jb) * rowstride; });
            }
            for (int i = tid; i < n; i += NT) twl[i] = a.tw[i];
            __syncthreads();
            if (doA) {
                reg::rdft<RA, -1>(v);
                sfor<RA>([&](auto k_) {
                    constexpr int k = decltype(k_)::value;
                    double2 x = v[k];
                    if (k > 0) x = cmul(x, twl[jb * k]);
                    la[k * RBP] = x;
                });
            }
        } else {
            // ---- load (B layout), step B inverse
            double2 u[RB];
            if (doB) {
                sfor<RB>([&](auto k_) {
                    constexpr int k = decltype(k_)::value;
                    u[k] = a.in[gB + long(ka) * rowstride + long(k * RA) * rowstride];
                });
            }
            for (int i = tid; i < n; i += NT) twl[i] = a.tw[i];
            __syncthreads();
            if (doB) {
                reg::rdft<RB, 1>(u);
                sfor<RB>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    double2 x = u[j];
                    if (j > 0) x = cmul(x, cconj(twl[j * ka]));
                    lb[j] = x;
                });
            }
        }

        // field loads of the next phases, issued before the exchange
        double mul[RB];
        if constexpr (mode == 3) {
            if (doB) {
                sfor<RB>([&](auto k_) {
                    constexpr int k = decltype(k_)::value;
                    mul[k] = a.mult[gB + long(ka) * rowstride + long(k * RA) * rowstride];
                });
            }
        }
        double2 add[RA];
        const bool has_add = mode != 1 && a.addend != nullptr;
        auto load_add = [&]() {
            if (has_add && doA) {
                sfor<RA>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    add[j] = a.addend[gA + long(j * RB) * rowstride + long(jb) * rowstride];
                });
            }
        };
        if constexpr (EARLY) load_add();
        __syncthreads();

        if constexpr (mode != 2) {
            // ---- step B forward (+ multiply + step B inverse)
            if (doB) {
                double2 u[RB];
                sfor<RB>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    u[j] = lb[j];
                });
                reg::rdft<RB, -1>(u);
                if constexpr (mode == 1) {
                    sfor<RB>([&](auto k_) {
                        constexpr int k = decltype(k_)::value;
                        double2 x = cscale(u[k], a.scale);
                        if (a.addend) x = cadd(x, a.addend[gB + long(ka) * rowstride + long(k * RA) * rowstride]);
                        a.out[gB + long(ka) * rowstride + long(k * RA) * rowstride] = x;
                    });
                } else {
                    if constexpr (mode == 5) {
                        sfor<RB>([&](auto k_) {
                            constexpr int k = decltype(k_)::value;
                            a.out2[gB + long(ka) * rowstride + long(k * RA) * rowstride] = u[k];
                        });
                    }
                    if constexpr (deriv) {
                        const double coef = line_coef(a, a.axis == 1 ? (row0 + lineB) / a.N0 : plane0);
                        sfor<RB>([&](auto k_) {
                            constexpr int k = decltype(k_)::value;
                            u[k] = cmul_i(u[k], coef * double(gamma_centred(ka + RA * k, n)));
                        });
                    } else if constexpr (mode == 4) {
                        sfor<RB>([&](auto k_) {
                            constexpr int k = decltype(k_)::value;
                            u[k] = cmul(u[k], phase(a, ka + RA * k, n));
                        });
                    } else {
                        // frequency ka + RA k: zero at ka = 0, k = 0
                        if (a.r0 && a.axis == 0 && ka == 0 && col0 + lineB == 0) a.r0[plane0] = u[0];
                        sfor<RB>([&](auto k_) {
                            constexpr int k = decltype(k_)::value;
                            u[k] = cscale(u[k], a.scale * mul[k]);
                        });
                    }
                    reg::rdft<RB, 1>(u);
                    sfor<RB>([&](auto j_) {
                        constexpr int j = decltype(j_)::value;
                        double2 x = u[j];
                        if (j > 0) x = cmul(x, cconj(twl[j * ka]));
                        lb[j] = x;
                    });
                }
            }
            if constexpr (mode == 1) return;
            __syncthreads();
        }

        // ---- step A inverse, store (A layout)
        if constexpr (!EARLY) load_add();
        if (doA) {
            double2 v[RA];
            sfor<RA>([&](auto k_) {
                constexpr int k = decltype(k_)::value;
                v[k] = la[k * RBP];
            });
            reg::rdft<RA, 1>(v);
            const double sc = a.scale;
            sfor<RA>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                double2 x = v[j];
                if constexpr (mode == 2) x = cscale(x, sc);
                if (has_add) x = cadd(x, add[j]);
                a.out[gA + long(j * RB) * rowstride + long(jb) * rowstride] = x;
            });
        }
    }
}

template <int RA, int RB, int NT, int MODE, bool EARLY>
void launch_mode(const AxisFFTArgs& a, int blocks, size_t lds, hipStream_t s)
{
    using Sh = TwoStepShape<RA, RB>;
    auto kern = &two_step_kernel<RA, RB, NT, MODE, EARLY>;
    static std::once_flag raised;  // contexts may be created from several host threads
    if (lds > 0)
        std::call_once(raised, [&] {
            const int maxl = NT / std::max(RA, RB);
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      int((Sh::n + maxl * Sh::LS) * sizeof(double2)));
        });
    if (g_attributes_only) return;
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(NT), lds, s, a);
}

// variant: bit 0 the early addend load (derivative mode), bit 1 work-groups
// of 512 threads (twice the lines, longer contiguous runs along columns)
template <int RA, int RB, int NT>
void launch_nt(const AxisFFTArgs& a, int variant, int blocks, size_t lds, hipStream_t s)
{
    switch (a.mode) {
        case 0:
            if (variant & 1)
                launch_mode<RA, RB, NT, 0, true>(a, blocks, lds, s);
            else
                launch_mode<RA, RB, NT, 0, false>(a, blocks, lds, s);
            break;
        case 5:
            if (variant & 1)
                launch_mode<RA, RB, NT, 5, true>(a, blocks, lds, s);
            else
                launch_mode<RA, RB, NT, 5, false>(a, blocks, lds, s);
            break;
        case 1: launch_mode<RA, RB, NT, 1, false>(a, blocks, lds, s); break;
        case 2: launch_mode<RA, RB, NT, 2, false>(a, blocks, lds, s); break;
        case 3: launch_mode<RA, RB, NT, 3, false>(a, blocks, lds, s); break;
        default: launch_mode<RA, RB, NT, 4, false>(a, blocks, lds, s); break;
    }
}

template <int RA, int RB>
void launch_two(const AxisFFTArgs& a, int variant, hipStream_t s)
{
    using Sh = TwoStepShape<RA, RB>;
    int blocks;
    if (a.axis == 1)
        blocks = int((long(a.P) * a.N0 + a.L - 1) / a.L);
    else
        blocks = a.P * ((a.N1 + a.L - 1) / a.L);
    const size_t lds = RB == 1 ? 0 : size_t(Sh::n + a.L * Sh::LS) * sizeof(double2);
    if (variant & 2)
        launch_nt<RA, RB, 2 * kTwoStepThreads>(a, variant, blocks, lds, s);
    else
        launch_nt<RA, RB, kTwoStepThreads>(a, variant, blocks, lds, s);
}

// Line lengths with a two-step kernel and their split n = RA * RB.
#define EDUS_TWO_STEP_SIZES(X)                                                                          \
    X(2, 2, 1) X(3, 3, 1) X(4, 4, 1) X(5, 5, 1) X(6, 6, 1) X(7, 7, 1) X(8, 8, 1) X(9, 9, 1) X(10, 10, 1)    \
    X(11, 11, 1) X(12, 12, 1) X(13, 13, 1) X(14, 14, 1) X(15, 15, 1) X(16, 16, 1) X(20, 5, 4)             \
    X(24, 8, 3) X(30, 6, 5) X(32, 8, 4) X(36, 6, 6) X(40, 8, 5) X(48, 8, 6) X(60, 10, 6) X(64, 8, 8)       \
    X(72, 8, 9) X(80, 8, 10) X(90, 10, 9) X(96, 8, 12) X(100, 10, 10) X(120, 8, 15) X(128, 8, 16)        \
    X(144, 12, 12) X(160, 16, 10) X(180, 12, 15) X(192, 16, 12) X(240, 16, 15) X(256, 16, 16)

}  // namespace

// both axes' tables in one launch: index j < n0 fills tab0, the next n1 tab1
__global__ void phase_table_kernel(double2* __restrict__ tab0, int n0, double coef0, double scale0,
                                   double2* __restrict__ tab1, int n1, double coef1, double scale1)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    double2* tab = tab0;
    int n = n0;
    double coef = coef0, scale = scale0;
    if (j >= n0) {
        j -= n0;
        tab = tab1;
        n = n1;
        coef = coef1;
        scale = scale1;
    }
    if (j >= n) return;
    double sn, cs;
    sincos(coef * double(gamma_centred(j, n)), &sn, &cs);
    tab[j] = make_double2(scale * cs, scale * sn);
}

void launch_phase_tables(double2* tab0, int n0, double coef0, double scale0, double2* tab1, int n1, double coef1,
                         double scale1, hipStream_t s)
{
    const int n = n0 + n1;
    if (n == 0) return;
    hipLaunchKernelGGL(phase_table_kernel, dim3((n + 255) / 256), dim3(256), 0, s, tab0, n0, coef0, scale0, tab1, n1,
                       coef1, scale1);
}

int two_step_max_lines(int n, int variant)
{
    const int nt = (variant & 2) ? 2 * kTwoStepThreads : kTwoStepThreads;
    switch (n) {
#define EDUS_X(N, A, B) \
    case N:             \
        return (B) == 1 ? nt : nt / std::max((A), (B));
        EDUS_TWO_STEP_SIZES(EDUS_X)
#undef EDUS_X
        default: return 0;
    }
}

void launch_two_step(const AxisFFTArgs& a, int variant, hipStream_t s)
{
    switch (a.n) {
#define EDUS_X(N, A, B)             \
    case N:                         \
        launch_two<(A), (B)>(a, variant, s); \
        break;
        EDUS_TWO_STEP_SIZES(EDUS_X)
#undef EDUS_X
        default: break;
    }
}

}  // namespace edus
