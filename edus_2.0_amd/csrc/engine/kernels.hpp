// Kernel argument blocks and launch entry points of the EDUS engine.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"

namespace edus {

// While set (by an AttributesOnly guard), the launch functions only raise
// their kernels' dynamic-LDS limit (hipFuncSetAttribute, once per kernel) and
// launch nothing.  Engines do this for every kernel they may launch before
// any hipGraph capture: a limit first raised while a stream is being captured
// does not reach the captured launch.
extern thread_local bool g_attributes_only;
struct AttributesOnly {
    AttributesOnly() { g_attributes_only = true; }
    ~AttributesOnly() { g_attributes_only = false; }
};

struct FFTFactors {
    int count;
    int radix[kMaxFactors];
};

struct AxisFFTArgs {
    const double2* in;      // planes [P][N0][N1]
    const double2* addend;  // optional, same layout
    double2* out;
    int P, N0, N1;
    int axis;               // 0: columns (length N0), 1: rows (length N1)
    int n;                  // line length
    int L;                  // lines per work-group
    int mode;               // 0 derivative, 1 forward, 2 inverse, 3 convolution, 4 phase,
                            // 5 derivative that also stores the forward transform in out2
    double coef;            // derivative: c_axis / n
    // derivative of stacked instances: plane p uses coef_tab[(p / ppi) * coef_stride]
    // (instance p / ppi, ppi = planes per instance) instead of coef
    const double* coef_tab;
    int coef_stride, ppi;
    double scale;           // transforms: output scale; convolution: factor on mult
    const double* mult;     // convolution: real multiplier field, triangle planes [P][N0][N1]
    FFTFactors fac;
    const double2* tw;      // exp(-2 pi i j / n), j < n
    // mode 4 input: in + sum_d in_add_c[d] * in_add[d] (null terms skipped)
    const double2* in_add[3];
    double in_add_c[3];
    // mode 4 multiplier per frequency index j < n: scale * exp(i coef R(j)),
    // R(j) Gamma-centred (launch_phase_tables); null: formed per element
    const double2* ptab;
    // mode 5: the unscaled forward transform of in, as mode 1 with scale 1
    // and no addend would store it (two-step kernels only)
    double2* out2;
    // mode 3 along columns (axis 0): each plane's zero-frequency value of the
    // forward transform on column 0, r0[plane] (null: not stored); with the
    // unscaled forward row transform as input it is the plane's sum over k
    double2* r0;
};

// Per-stage update rule of the fused commutator kernel.  With f the
// right-hand side at this k-point and element:
//   out_f  = f                          (if out_f)
//   out1   = x1 + b1 * f                (if out1)
//   out2   = x2 + b2 * f                (if out2)
//   AB:    y  = y + sum_i c_i h_i  with h_0 = f, h_i = hist[i]  (if ab_order)
struct StageUpdate {
    double2* out_f;
    // RK: out1 = x1 + g1 (rho - x2) + b1 f and out2 = x2 + b2 f, rho the
    // stage input (g1 lets a stage recover an earlier stage's b f term from
    // rho - x2 instead of carrying it through another array)
    const double2* x1; double2* out1; double b1;
    const double2* x2; double2* out2; double b2;
    double g1;
    int ab_order;
    const double2* hist[5];  // hist[0] unused (= f)
    double ab_c[5];
    double2* y;               // AB target
    // invariant monitor on the state this stage writes (RK: out1, AB: y):
    // max over the tile of |Tr - nf| and of the Hermiticity defect into the
    // work-group's monitor slots (mon_slot)
    unsigned long long* mon;
    double nf;
};

struct CommArgs {
    int Nk, P;
    int B;                    // instances stacked as [B][P][Nk] in every per-k array (0 or 1: one)
    const double2* rho;       // stage input, planes
    const double2* D;         // E.grad rho part (may be null)
    const double2* H0;
    const double2* xi[3];     // null where E_d == 0 or operator absent
    double E[3];
    const double* E_tab;      // per-instance field [B][3] on device, replaces E when set
    const double2* sigma;     // Hartree-Fock self energy in k (may be null)
    // Hartree diagonal shift (may be null): every plane's sum over k, [B][P]
    // (AxisFFTArgs::r0 of the exchange pass), V [B][nb][nb] and 1 / Nk;
    // the kernel forms the shift with hartree_shift
    const double2* hsum;
    const double* hart;
    double inv_nk;
    StageUpdate up;
    // filled by launch_comm from xi and sigma: the addends of H in the order
    // xi_0, xi_1, xi_2, Sigma (present ones only); add_dd is the field
    // component of each (-1: Sigma, coefficient 1)
    const double2* add[4];
    int add_dd[4];
    int n_add;
    // the FP32 commutator engine (mixed precision) reads H0 and xi stored in
    // FP32 instead (H0f, xif; required for it); launch_comm fills addf with
    // H0f (addf_dd -1) and the xif present, and keeps Sigma in `sigma`
    const float2* H0f;
    const float2* xif[3];
    const float2* addf[4];
    int addf_dd[4];
    int n_addf;
    // FULL layout: H0, xi (and H0f, xif) hold only the upper-triangle planes
    // of Hermitian fields, nb (nb + 1) / 2 per instance in HERMITIAN order;
    // element (n, m) with n > m is read as the conjugate of (m, n).  Not with
    // Sigma (a FULL-layout exchange term is not Hermitian on even grids).
    int tri;
};

// derivative coefficient of the line in plane `plane`
__device__ inline double line_coef(const AxisFFTArgs& a, long plane)
{
    return a.coef_tab ? a.coef_tab[(plane / a.ppi) * a.coef_stride] : a.coef;
}

// Axis-FFT tile configurations (threads per work-group, values per work-group)
struct FFTConfig {
    int threads, values;
};
constexpr int kNumFFTConfigs = 5;
extern const FFTConfig kFFTConfigs[kNumFFTConfigs];
int fft_lines(int cfg, int n);  // lines per work-group for a line length n

void launch_axis_fft(const AxisFFTArgs& a, int cfg, hipStream_t s);

// Two-step (register DFT + one LDS exchange) axis FFT: lines per work-group
// it allows for a line length n (0: no kernel for this n), and its launch.
// variant 1 loads the derivative's addend one phase ahead.
// lines per work-group of the two-step kernel for a line of n (0: no such
// kernel); variant bit 1 doubles the work-group (512 threads)
int two_step_max_lines(int n, int variant = 0);
void launch_two_step(const AxisFFTArgs& a, int variant, hipStream_t s);

// The kernel family and tile used for one grid axis.  The family follows
// from the line length alone (two-step where it has a kernel), so results do
// not depend on timing; the autotuner picks the tile inside the family, and
// tiles of one family give bit-identical results.
struct FFTChoice {
    int two_step = 0;  // 1: two-step kernel, 0: Stockham
    int cfg = 0;       // Stockham configuration, or two-step variant
    int L = 1;         // lines per work-group
};
void launch_fft(const AxisFFTArgs& a, const FFTChoice& c, hipStream_t s);
// tab[j] = scale * exp(i coef R(j)), j < n, R(j) the Gamma-centred index, for
// two tables in one launch (n0 or n1 may be 0: that table is left alone)
void launch_phase_tables(double2* tab0, int n0, double coef0, double scale0, double2* tab1, int n1, double coef1,
                         double scale1, hipStream_t s);
// Engines for [H, rho] in the commutator kernel: FP64 FMAs, FP64 MFMA, or
// FP32 MFMA with FP32 accumulation (mixed precision; everything else FP64).
enum CommKind { kCommScalar = 0, kCommMfma64 = 1, kCommMfma32 = 2 };
// tile 0: 8 k-points per work-group, 1: 4 (MFMA engines; the FMA engine has one)
void launch_comm(int nb, int layout, int kind, int tile, const CommArgs& a, hipStream_t s);

// The spectral derivative of every plane in one pass (grad_fused.hip): out =
// F0^-1 i c0 R0 F0 in + F1^-1 i c1 R1 F1 in, the two-pass axis kernels'
// result to rounding; one work-group per plane, for the grid shapes
// grad_fused_available accepts.  coef = c_j / N_j, or per instance from
// coef_tab[(plane / ppi) * coef_stride + j].
struct GradFusedArgs {
    const double2* in;
    double2* out;
    int P, N0, N1;
    double coef[2];
    const double* coef_tab;
    int coef_stride, ppi;
    const double2* tw0;  // exp(-2 pi i j / N0), j < N0
    const double2* tw1;
};
bool grad_fused_available(int N0, int N1);
void launch_grad_fused(const GradFusedArgs& a, hipStream_t s);

// Gauge rotation per k (Operator::go_to_bloch / go_to_wannier): out =
// U^dagger in U (to_bloch) or U in U^dagger, in is planes [npl][Nk] of the
// launch's input layout, U full planes [nb*nb][Nk] (U(a, n) in plane a*nb + n,
// columns the eigenvectors), out in out_layout: kFull or kHermitian planes, or
// kRotKMajor ([Nk][nb][nb], the reference's BlockMatrix order).  kind:
// kCommMfma64 (FP64) or kCommMfma32 (FP32 operands and accumulation).
constexpr int kRotKMajor = 2;
struct RotateArgs {
    int Nk;
    const double2* in;
    const double2* U;
    double2* out;
    int out_layout;
    int to_bloch;
};
void launch_rotate(int nb, int in_layout, int kind, const RotateArgs& a, hipStream_t s);

// k-space finite-difference gradient (gradient_space "k"): one stencil point
// per mesh vector b of the shells in use: grid offset, shell weight, Cartesian b
struct KStencilPoint {
    int m0, m1;
    double w;
    double b[3];
    int shell;
};
std::vector<KStencilPoint> kgradient_stencil(const double A[3][3], int N0, int N1);
// out = sum_b w_b (b.E) in(k + b) for every plane of B stacked instances;
// E_tab ([B][3] on the device) replaces E when set
void launch_kgrad(const double2* in, double2* out, int N0, int N1, int P, int B, const KStencilPoint* st, int nst,
                  const double E[3], const double* E_tab, hipStream_t s);


// Lindhard response of the screened-potential application: launch shape for
// nb bands, nf filled, an N0 x N1 grid (element pairs per lane, q per
// work-group, LDS sizes), and the launch.  U: [Nk][nb][nb] eigenvectors in
// columns, eps: [Nk][nb]; work: lindhard_work_bytes of device scratch (the
// band-split copy of U); X: planes [nb * nb][Nk].
struct LindhardPlan {
    int nb, nv, nc, N0, N1, Nk;
    int Pg, G, W, tri, TQ;
    int nqb_row, nR, nvp, buf, blocks;
    size_t lds;
};
LindhardPlan lindhard_plan(int nb, int nf, int N0, int N1);
size_t lindhard_work_bytes(const LindhardPlan& p);
void launch_lindhard(const LindhardPlan& p, const double2* U, const double* eps, double2* work, double2* X,
                     hipStream_t s);

}  // namespace edus
