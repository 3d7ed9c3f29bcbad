// The spectral derivative E.grad_k rho in one pass per plane.
//
// The reference's gradient term (kGradient_definitions.hpp:35-41 inside the
// SourceTerm sandwich, Functional_SourceTerm.hpp:20-24) multiplies rho(R) by
// i(R.E); with R = R0 a0 + R1 a1 on the Gamma-centred grid that is
// i(c0 R0 + c1 R1), c_j = a_j.E, so the k-space result is the sum of two 1-D
// spectral derivatives:
//     D = F0^-1 diag(i c0 R0 / N0) F0 rho  +  F1^-1 diag(i c1 R1 / N1) F1 rho
// (F_j the transform along grid axis j).  The two-pass form (fft_two.hip) runs
// the column term into a work plane set G and the row term plus G into D:
// 5 plane sets of HBM traffic.  Here one work-group takes a whole plane and
// moves 2 (rho in, D out; rho is read a second time for the column term,
// from cache):
//   row phase     rows in batches of LR; each row goes through the two-step
//                 register DFT (RA1 x RB1 around one LDS exchange, the same
//                 arithmetic as the two-pass kernel); thread t keeps the row
//                 term of its A-task elements (row r0 + t / RB1, columns
//                 j RB1 + t % RB1) in registers, acc[batch][j];
//   column phase  columns in batches of LC through the same two-step DFT
//                 along axis 0; the column term is staged in LDS in (column,
//                 row) order, and every thread adds the elements its
//                 accumulators own and stores the finished D.
// The row and column terms go through the two-pass kernel's DFT arithmetic
// and D is their sum in the same order; the compiler contracts multiply-adds
// per kernel, so the two forms agree to rounding.
#include <algorithm>
#include <mutex>

#include "common.hpp"
#include "fft_reg.hpp"
#include "kernels.hpp"

namespace edus {

namespace {

using reg::sfor;

constexpr int kGradThreads = 512;

template <int N0, int RA0, int RB0, int N1, int RA1, int RB1>
struct FusedShape {
    static_assert(RA0 * RB0 == N0 && RA1 * RB1 == N1, "line splits");
    static constexpr int RBP0 = RB0 | 1, LS0 = (RA0 * RBP0) | 1;  // odd strides in LDS (as fft_two.hip)
    static constexpr int RBP1 = RB1 | 1, LS1 = (RA1 * RBP1) | 1;
    static constexpr int LR = kGradThreads / (RB1 > RA1 ? RB1 : RA1);  // rows per batch
    static constexpr int NBR = (N0 + LR - 1) / LR;                     // row batches
    static constexpr int LC = kGradThreads / (RB0 > RA0 ? RB0 : RA0);  // columns per batch
    static constexpr int NBC = (N1 + LC - 1) / LC;
    static constexpr int SP = N0 | 1;  // staging stride of one column (rows), odd
    static constexpr int BUF_ROW = LR * LS1;
    static constexpr int BUF_COL = LC * (LS0 > SP ? LS0 : SP);
    static constexpr int BUF = BUF_ROW > BUF_COL ? BUF_ROW : BUF_COL;
    // row batches whose row terms stay in registers; a longer list would
    // spill (120 x 120: four batches of 8 values), so the last batch then
    // keeps its row terms in LDS, in (row, column) order
    static constexpr int NREG = NBR * RA1 > 24 ? NBR - 1 : NBR;
    static constexpr int LAST_ROWS = NREG < NBR ? N0 - NREG * LR : 0;
    static constexpr size_t lds_bytes = size_t(N0 + N1 + BUF + LAST_ROWS * N1) * sizeof(double2);
    // the column batches' loads go one batch ahead where the registers allow
    // it (the accumulators plus the B step's line and the prefetch)
    static constexpr bool PREFETCH = NREG * RA1 + RB0 <= 36;
};

template <int N0, int RA0, int RB0, int N1, int RA1, int RB1>
__global__ void __launch_bounds__(kGradThreads) grad_fused_kernel(GradFusedArgs a)
{
    using Sh = FusedShape<N0, RA0, RB0, N1, RA1, RB1>;
    constexpr int RBP0 = Sh::RBP0, LS0 = Sh::LS0, RBP1 = Sh::RBP1, LS1 = Sh::LS1;
    constexpr int LR = Sh::LR, NBR = Sh::NBR, LC = Sh::LC, NBC = Sh::NBC, SP = Sh::SP;
    constexpr int NREG = Sh::NREG, LAST_ROWS = Sh::LAST_ROWS;
    constexpr long Nk = long(N0) * N1;
    extern __shared__ double2 lds[];
    double2* tw0 = lds;
    double2* tw1 = lds + N0;
    double2* buf = lds + N0 + N1;
    double2* accl = buf + Sh::BUF;  // row terms of the last batch when NREG < NBR
    const int tid = threadIdx.x;
    const long plane = xcd_contiguous_block();
    const double2* in = a.in + plane * Nk;
    double2* out = a.out + plane * Nk;
    const double c0 = a.coef_tab ? a.coef_tab[(plane / a.ppi) * a.coef_stride] : a.coef[0];
    const double c1 = a.coef_tab ? a.coef_tab[(plane / a.ppi) * a.coef_stride + 1] : a.coef[1];
    for (int i = tid; i < N0; i += kGradThreads) tw0[i] = a.tw0[i];
    for (int i = tid; i < N1; i += kGradThreads) tw1[i] = a.tw1[i];

    // ---- row phase: acc[b][j] = row term of (row r0_b + t / RB1, column j RB1 + t % RB1)
    const int lineA1 = tid / RB1, jb1 = tid - (tid / RB1) * RB1;
    const int lineB1 = tid / RA1, ka1 = tid - (tid / RA1) * RA1;
    // every row load is issued up front: the register batches' inputs go
    // straight into their accumulators (which the transforms then overwrite
    // with the row terms), the LDS batch's into its accl slots (each thread
    // later reads back and overwrites only its own)
    double2 acc[NREG][RA1];
    {
        double2 vl[RA1];
        sfor<NBR>([&](auto b_) {
            constexpr int b = decltype(b_)::value;
            const bool ok = lineA1 < LR && b * LR + lineA1 < N0;
            const double2* src = in + long(b * LR + lineA1) * N1 + jb1;
            sfor<RA1>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                double2 x = make_double2(0.0, 0.0);
                if (ok) x = src[j * RB1];
                if constexpr (b < NREG)
                    acc[b][j] = x;
                else
                    vl[j] = x;
            });
        });
        if constexpr (LAST_ROWS > 0) {
            if (lineA1 < LAST_ROWS)
                sfor<RA1>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    accl[lineA1 * N1 + j * RB1 + jb1] = vl[j];
                });
        }
    }
    sfor<NBR>([&](auto b_) {
        constexpr int b = decltype(b_)::value;
        const int r0 = b * LR;
        const bool doA = lineA1 < LR && r0 + lineA1 < N0;
        const bool doB = lineB1 < LR && r0 + lineB1 < N0;
        double2* la = buf + lineA1 * LS1 + jb1;
        double2* lb = buf + lineB1 * LS1 + ka1 * RBP1;
        double2 v[RA1];
        sfor<RA1>([&](auto j_) {
            constexpr int j = decltype(j_)::value;
            if constexpr (b < NREG) {
                v[j] = acc[b][j];
            } else {
                v[j] = make_double2(0.0, 0.0);
                if (doA) v[j] = accl[lineA1 * N1 + j * RB1 + jb1];
            }
        });
        __syncthreads();  // twiddles in LDS; the previous batch is done with buf
        if (doA) {
            reg::rdft<RA1, -1>(v);
            sfor<RA1>([&](auto k_) {
                constexpr int k = decltype(k_)::value;
                double2 x = v[k];
                if (k > 0) x = cmul(x, tw1[jb1 * k]);
                la[k * RBP1] = x;
            });
        }
        __syncthreads();
        if (doB) {
            double2 u[RB1];
            sfor<RB1>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                u[j] = lb[j];
            });
            reg::rdft<RB1, -1>(u);
            sfor<RB1>([&](auto k_) {
                constexpr int k = decltype(k_)::value;
                u[k] = cmul_i(u[k], c1 * double(gamma_centred(ka1 + RA1 * k, N1)));
            });
            reg::rdft<RB1, 1>(u);
            sfor<RB1>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                double2 x = u[j];
                if (j > 0) x = cmul(x, cconj(tw1[j * ka1]));
                lb[j] = x;
            });
        }
        __syncthreads();
        if (doA) {
            sfor<RA1>([&](auto k_) {
                constexpr int k = decltype(k_)::value;
                v[k] = la[k * RBP1];
            });
            reg::rdft<RA1, 1>(v);
        }
        if constexpr (b < NREG) {
            sfor<RA1>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                acc[b][j] = v[j];
            });
        } else if (doA) {
            sfor<RA1>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                accl[lineA1 * N1 + j * RB1 + jb1] = v[j];
            });
        }
    });

    // ---- column phase; each batch's loads are issued one batch ahead (vn)
    // where the registers allow it
    const int lineA = tid % LC, jb = tid / LC;  // adjacent threads: adjacent columns
    double2 vn[RA0];
    auto load_cols = [&](int cb) {
        const bool ok = jb < RB0 && cb + lineA < N1;
        const double2* src = in + long(jb) * N1 + cb + lineA;
        sfor<RA0>([&](auto j_) {
            constexpr int j = decltype(j_)::value;
            vn[j] = make_double2(0.0, 0.0);
            if (ok) vn[j] = src[long(j) * RB0 * N1];
        });
    };
    load_cols(0);
    sfor<NBC>([&](auto g_) {
        constexpr int g = decltype(g_)::value;
        const int cb = g * LC;
        const bool doA = jb < RB0 && cb + lineA < N1;
        const bool doB = jb < RA0 && cb + lineA < N1;  // B task (column lineA, ka = jb)
        double2* la = buf + lineA * LS0 + jb;
        double2* lb = buf + lineA * LS0 + jb * RBP0;
        if constexpr (!Sh::PREFETCH && g > 0) load_cols(cb);
        double2 v[RA0];
        sfor<RA0>([&](auto j_) { v[decltype(j_)::value] = vn[decltype(j_)::value]; });
        if constexpr (Sh::PREFETCH && g + 1 < NBC) load_cols(cb + LC);
        __syncthreads();  // the row phase / previous batch is done with buf
        if (doA) {
            reg::rdft<RA0, -1>(v);
            sfor<RA0>([&](auto k_) {
                constexpr int k = decltype(k_)::value;
                double2 x = v[k];
                if (k > 0) x = cmul(x, tw0[jb * k]);
                la[k * RBP0] = x;
            });
        }
        __syncthreads();
        if (doB) {
            double2 u[RB0];
            sfor<RB0>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                u[j] = lb[j];
            });
            reg::rdft<RB0, -1>(u);
            sfor<RB0>([&](auto k_) {
                constexpr int k = decltype(k_)::value;
                u[k] = cmul_i(u[k], c0 * double(gamma_centred(jb + RA0 * k, N0)));
            });
            reg::rdft<RB0, 1>(u);
            sfor<RB0>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                double2 x = u[j];
                if (j > 0) x = cmul(x, cconj(tw0[j * jb]));
                lb[j] = x;
            });
        }
        __syncthreads();
        if (doA) {
            sfor<RA0>([&](auto k_) {
                constexpr int k = decltype(k_)::value;
                v[k] = la[k * RBP0];
            });
            reg::rdft<RA0, 1>(v);
        }
        __syncthreads();  // every A task has read its inputs: stage the column terms
        if (doA) {
            double2* st = buf + lineA * SP + jb;
            sfor<RA0>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                st[j * RB0] = v[j];  // (column cb + lineA, row j RB0 + jb)
            });
        }
        __syncthreads();
        // D = row term + column term for the accumulator elements in this
        // batch's columns, stored as they are finished
        if constexpr (LAST_ROWS > 0) {
            constexpr int r0 = NREG * LR;
            for (int e = tid; e < LAST_ROWS * LC; e += kGradThreads) {
                const int rr = e / LC, cc = e - (e / LC) * LC;
                const int col = cb + cc;
                if (col < N1) out[long(r0 + rr) * N1 + col] = cadd(accl[rr * N1 + col], buf[cc * SP + r0 + rr]);
            }
        }
        sfor<NREG>([&](auto b_) {
            constexpr int b = decltype(b_)::value;
            const int row = b * LR + lineA1;
            if (lineA1 < LR && row < N0) {
                sfor<RA1>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    const int col = j * RB1 + jb1;
                    if (col >= cb && col < cb + LC) {
                        const double2 x = cadd(acc[b][j], buf[(col - cb) * SP + row]);
                        out[long(row) * N1 + col] = x;
                    }
                });
            }
        });
    });
}

template <int N0, int RA0, int RB0, int N1, int RA1, int RB1>
void launch_shape(const GradFusedArgs& a, hipStream_t s)
{
    using Sh = FusedShape<N0, RA0, RB0, N1, RA1, RB1>;
    auto kern = &grad_fused_kernel<N0, RA0, RB0, N1, RA1, RB1>;
    static std::once_flag raised;  // contexts may be created from several host threads
    std::call_once(raised, [&] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(Sh::lds_bytes));
    });
    if (g_attributes_only) return;
    hipLaunchKernelGGL(kern, dim3(a.P), dim3(kGradThreads), Sh::lds_bytes, s, a);
}

// Grid shapes with a fused kernel: the square grids of the benchmark
// workloads (each axis split as the two-step kernel splits it).
#define EDUS_FUSED_GRAD_SHAPES(X)                  \
    X(120, 8, 15, 120, 8, 15) X(96, 8, 12, 96, 8, 12) \
    X(90, 10, 9, 90, 10, 9) X(48, 8, 6, 48, 8, 6)

}  // namespace

bool grad_fused_available(int N0, int N1)
{
#define EDUS_X(A0, RA0, RB0, A1, RA1, RB1) \
    if (N0 == (A0) && N1 == (A1)) return true;
    EDUS_FUSED_GRAD_SHAPES(EDUS_X)
#undef EDUS_X
    return false;
}

void launch_grad_fused(const GradFusedArgs& a, hipStream_t s)
{
#define EDUS_X(A0, RA0, RB0, A1, RA1, RB1)                     \
    if (a.N0 == (A0) && a.N1 == (A1)) {                        \
        launch_shape<(A0), (RA0), (RB0), (A1), (RA1), (RB1)>(a, s); \
        return;                                                \
    }
    EDUS_FUSED_GRAD_SHAPES(EDUS_X)
#undef EDUS_X
}

}  // namespace edus
