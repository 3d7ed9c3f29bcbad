// Batched 1-D FFTs along one axis of the k grid, staged in LDS (Stockham
// auto-sort, mixed radix 2/3/4/5/8 and register DFTs for the primes 7, 11, 13).
//
// Used for the E.grad_k rho term.  The reference transforms every plane to R
// with a 2-D FFT (sign -1, scaled 1/N), multiplies by i(R.E) on the
// Gamma-centred grid and transforms back (sign +1)
// (Functional_SourceTerm.hpp:15-20,49; kGradient_definitions.hpp:35-41;
// fftPair.cpp:89-107).  With R = R0 a0 + R1 a1 the multiplier is the SUM
// i(c0 R0 + c1 R1), c_j = a_j.E, so the 2-D sandwich splits exactly into two
// 1-D spectral derivatives:
//     D = c0 * F0^-1 diag(i R0 / N0) F0 rho  +  c1 * F1^-1 diag(i R1 / N1) F1 rho
// (F_a acts along grid axis a only).  Each line is one row or column of a
// plane, so a whole line lives in LDS and the transform never leaves the CU.
//
// The same kernel does plain forward / inverse transforms and the R-space
// point-wise product of the Hartree-Fock convolution along an axis.
//
// Work-group: NT threads, L lines of length n with L*n <= MAXV.  Global data
// moves through registers (every load of the tile is issued before the first
// LDS write); the passes ping-pong between two LDS buffers, one barrier each.
// (NT, MAXV) is picked per grid axis by the engine's autotuner.
#include <algorithm>
#include <mutex>

#include "common.hpp"
#include "kernels.hpp"

namespace edus {

namespace {

constexpr double kC5_1 = 0.30901699437494742;   // cos(2 pi/5)
constexpr double kC5_2 = -0.80901699437494742;  // cos(4 pi/5)
constexpr double kS5_1 = 0.95105651629515357;   // sin(2 pi/5)
constexpr double kS5_2 = 0.58778525229247313;   // sin(4 pi/5)
constexpr double kS3 = 0.86602540378443865;     // sin(2 pi/3)
constexpr double kR2 = 0.70710678118654752;     // 1/sqrt(2)

__device__ inline void dft4(const double2 x0, const double2 x1, const double2 x2, const double2 x3, double sign,
                            double2& y0, double2& y1, double2& y2, double2& y3)
{
    const double2 a02 = cadd(x0, x2), s02 = csub(x0, x2);
    const double2 a13 = cadd(x1, x3), s13 = csub(x1, x3);
    const double2 rot = cmul_i(s13, sign);
    y0 = cadd(a02, a13);
    y1 = cadd(s02, rot);
    y2 = csub(a02, a13);
    y3 = csub(s02, rot);
}

// In-register DFT of size R: y_s = sum_q v_q exp(sign 2 pi i q s / R).
// tw[e] = exp(-2 pi i e / n) is the line's table; n / R = nr.
template <int R>
__device__ inline void dft(double2 (&v)[R], double sign, const double2* __restrict__ tw, int nr)
{
    if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const double2 t1 = cadd(v[1], v[2]);
        const double2 u = csub(v[1], v[2]);
        const double2 base = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
        const double2 rot = cmul_i(u, sign * kS3);
        v[0] = cadd(v[0], t1);
        v[1] = cadd(base, rot);
        v[2] = csub(base, rot);
    } else if constexpr (R == 4) {
        double2 y0, y1, y2, y3;
        dft4(v[0], v[1], v[2], v[3], sign, y0, y1, y2, y3);
        v[0] = y0;
        v[1] = y1;
        v[2] = y2;
        v[3] = y3;
    } else if constexpr (R == 5) {
        const double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        const double2 u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
        const double2 b1 = make_double2(v[0].x + kC5_1 * t1.x + kC5_2 * t2.x, v[0].y + kC5_1 * t1.y + kC5_2 * t2.y);
        const double2 b2 = make_double2(v[0].x + kC5_2 * t1.x + kC5_1 * t2.x, v[0].y + kC5_2 * t1.y + kC5_1 * t2.y);
        const double2 r1 = cmul_i(make_double2(kS5_1 * u1.x + kS5_2 * u2.x, kS5_1 * u1.y + kS5_2 * u2.y), sign);
        const double2 r2 = cmul_i(make_double2(kS5_2 * u1.x - kS5_1 * u2.x, kS5_2 * u1.y - kS5_1 * u2.y), sign);
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(b1, r1);
        v[4] = csub(b1, r1);
        v[2] = cadd(b2, r2);
        v[3] = csub(b2, r2);
    } else if constexpr (R == 8) {
        // radix-2 step across stride 4, then two DFT4s: y_2m = DFT4(a)_m, y_2m+1 = DFT4(b)_m
        double2 a[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            a[k] = cadd(v[k], v[k + 4]);
            b[k] = csub(v[k], v[k + 4]);
        }
        // b_k *= w8^k, w8 = exp(sign 2 pi i / 8)
        b[1] = make_double2(kR2 * (b[1].x - sign * b[1].y), kR2 * (b[1].y + sign * b[1].x));
        b[2] = cmul_i(b[2], sign);
        b[3] = make_double2(kR2 * (-b[3].x - sign * b[3].y), kR2 * (-b[3].y + sign * b[3].x));
        dft4(a[0], a[1], a[2], a[3], sign, v[0], v[2], v[4], v[6]);
        dft4(b[0], b[1], b[2], b[3], sign, v[1], v[3], v[5], v[7]);
    } else {
        // direct DFT for the primes 7, 11, 13; W_R^e = tw[e * nr]
        double2 y[R];
#pragma unroll
        for (int s = 0; s < R; ++s) {
            double2 acc = v[0];
#pragma unroll
            for (int q = 1; q < R; ++q) {
                double2 w = tw[((q * s) % R) * nr];
                if (sign > 0) w.y = -w.y;
                acc = cadd(acc, cmul(v[q], w));
            }
            y[s] = acc;
        }
#pragma unroll
        for (int s = 0; s < R; ++s) v[s] = y[s];
    }
}

// One Stockham pass of radix R over `lines` lines of length n, src -> dst.
template <int R, int NT>
__device__ void pass(const double2* __restrict__ src, double2* __restrict__ dst, int lines, int n, int Ns,
                     const double2* __restrict__ tw, double sign)
{
    const int nr = n / R;
    const int tasks = lines * nr;
    const int tstride = n / (Ns * R);
    for (int t = threadIdx.x; t < tasks; t += NT) {
        const int line = t / nr;
        const int j = t - line * nr;
        const int k = j % Ns;
        const double2* s = src + line * n;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            double2 x = s[j + r * nr];
            if (r > 0) {
                double2 w = tw[r * k * tstride];  // forward uses w, inverse conj(w)
                if (sign > 0) w.y = -w.y;
                x = cmul(x, w);
            }
            v[r] = x;
        }
        dft<R>(v, sign, tw, nr);
        double2* d = dst + line * n + (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) d[r * Ns] = v[r];
    }
    __syncthreads();
}

// All passes, ping-ponging between a and b; returns the buffer with the result.
template <int NT>
__device__ double2* run_fft(double2* a, double2* b, int lines, int n, const FFTFactors& f, const double2* tw,
                            double sign)
{
    int Ns = 1;
    for (int i = 0; i < f.count; ++i) {
        switch (f.radix[i]) {
            case 2: pass<2, NT>(a, b, lines, n, Ns, tw, sign); break;
            case 3: pass<3, NT>(a, b, lines, n, Ns, tw, sign); break;
            case 4: pass<4, NT>(a, b, lines, n, Ns, tw, sign); break;
            case 5: pass<5, NT>(a, b, lines, n, Ns, tw, sign); break;
            case 7: pass<7, NT>(a, b, lines, n, Ns, tw, sign); break;
            case 8: pass<8, NT>(a, b, lines, n, Ns, tw, sign); break;
            case 11: pass<11, NT>(a, b, lines, n, Ns, tw, sign); break;
            case 13: pass<13, NT>(a, b, lines, n, Ns, tw, sign); break;
            default: break;  // rejected on the host
        }
        Ns *= f.radix[i];
        double2* t = a;
        a = b;
        b = t;
    }
    return a;
}

// One work-group handles `L` lines (L*n <= MAXV values).
//   axis 1: lines are consecutive rows (global row index p*N0 + i0): one
//           contiguous chunk of L*N1 values;
//   axis 0: lines are L consecutive columns of one plane.
// mode 0: derivative  out = addend? + F^-1 diag(i coef R) F in
// mode 1: forward     out = scale * F in          (sign -1)
// mode 2: inverse     out = scale * F^-1 in       (sign +1)
// mode 3: convolution out = F^-1 diag(scale * mult(R)) F in
// mode 4: phase       out = F^-1 diag(scale exp(i coef R)) F (in + sum_d c_d in_add_d)
template <int NT, int MAXV>
__global__ void __launch_bounds__(NT) axis_fft_kernel(AxisFFTArgs a)
{
    constexpr int PER = MAXV / NT;
    extern __shared__ double2 lds[];
    const int n = a.n;
    const int L = a.L;
    double2* twl = lds;                // n twiddles
    double2* bufA = lds + n;           // L*n values, line-major
    double2* bufB = bufA + L * n;
    const int Nk = a.N0 * a.N1;

    // global offset of value j of line c: base + c * cstride + j * rowstride
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
    // global enumeration of the tile: axis 1 walks the contiguous chunk line
    // after line, axis 0 walks row after row across the L columns
    const int span = a.axis == 1 ? nlines * n : n * L;

    {
        double2 v[PER];
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int idx = threadIdx.x + e * NT;
            const int c = a.axis == 1 ? idx / n : idx % L;
            const int j = a.axis == 1 ? idx - c * n : idx / L;
            if (idx < span && c < nlines) {
                v[e] = a.in[base + c * cstride + long(j) * rowstride];
            }
        }
        if (a.mode == 4) {  // input: a sum of fields, each field's loads issued together
            for (int d = 0; d < 3; ++d) {
                const double2* p = a.in_add[d];
                if (!p) continue;
                const double cd = a.in_add_c[d];
                double2 w[PER];
#pragma unroll
                for (int e = 0; e < PER; ++e) {
                    const int idx = threadIdx.x + e * NT;
                    const int c = a.axis == 1 ? idx / n : idx % L;
                    const int j = a.axis == 1 ? idx - c * n : idx / L;
                    if (idx < span && c < nlines) w[e] = p[base + c * cstride + long(j) * rowstride];
                }
#pragma unroll
                for (int e = 0; e < PER; ++e) {
                    const int idx = threadIdx.x + e * NT;
                    const int c = a.axis == 1 ? idx / n : idx % L;
                    if (idx < span && c < nlines) {
                        v[e].x += cd * w[e].x;
                        v[e].y += cd * w[e].y;
                    }
                }
            }
        }
        for (int i = threadIdx.x; i < n; i += NT) twl[i] = a.tw[i];
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int idx = threadIdx.x + e * NT;
            const int c = a.axis == 1 ? idx / n : idx % L;
            const int j = a.axis == 1 ? idx - c * n : idx / L;
            if (idx < span && c < nlines) bufA[c * n + j] = v[e];
        }
    }
    __syncthreads();

    double2* res;
    if (a.mode == 0 || a.mode == 3 || a.mode == 4) {
        double2* x = run_fft<NT>(bufA, bufB, nlines, n, a.fac, twl, -1.0);
        if (a.mode == 4) {  // scale exp(i coef R)
            for (int idx = threadIdx.x; idx < nlines * n; idx += NT) {
                const int j = idx % n;
                double2 ph;
                if (a.ptab) {
                    ph = a.ptab[j];
                } else {
                    double sn, cs;
                    sincos(a.coef * double(gamma_centred(j, n)), &sn, &cs);
                    ph = make_double2(a.scale * cs, a.scale * sn);
                }
                x[idx] = cmul(x[idx], ph);
            }
        } else if (a.mode == 0) {
            for (int idx = threadIdx.x; idx < nlines * n; idx += NT) {
                const int c = idx / n;
                const int j = idx - c * n;
                const double coef = line_coef(a, a.axis == 1 ? (row0 + c) / a.N0 : plane0);
                x[idx] = cmul_i(x[idx], coef * double(gamma_centred(j, n)));
            }
        } else {
            double mul[PER];
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                const int idx = threadIdx.x + e * NT;
                const int c = a.axis == 1 ? idx / n : idx % L;
                const int j = a.axis == 1 ? idx - c * n : idx / L;
                if (idx < span && c < nlines) mul[e] = a.mult[base + c * cstride + long(j) * rowstride];
            }
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                const int idx = threadIdx.x + e * NT;
                const int c = a.axis == 1 ? idx / n : idx % L;
                const int j = a.axis == 1 ? idx - c * n : idx / L;
                if (idx < span && c < nlines) {
                    if (a.r0 && a.axis == 0 && j == 0 && col0 + c == 0) a.r0[plane0] = x[c * n + j];
                    x[c * n + j] = cscale(x[c * n + j], a.scale * mul[e]);
                }
            }
        }
        __syncthreads();
        res = run_fft<NT>(x, x == bufA ? bufB : bufA, nlines, n, a.fac, twl, +1.0);
    } else {
        res = run_fft<NT>(bufA, bufB, nlines, n, a.fac, twl, a.mode == 1 ? -1.0 : +1.0);
    }
    const double sc = (a.mode == 1 || a.mode == 2) ? a.scale : 1.0;

    double2 add[PER];
    if (a.addend) {
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int idx = threadIdx.x + e * NT;
            const int c = a.axis == 1 ? idx / n : idx % L;
            const int j = a.axis == 1 ? idx - c * n : idx / L;
            if (idx < span && c < nlines) add[e] = a.addend[base + c * cstride + long(j) * rowstride];
        }
    }
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int idx = threadIdx.x + e * NT;
        const int c = a.axis == 1 ? idx / n : idx % L;
        const int j = a.axis == 1 ? idx - c * n : idx / L;
        if (idx < span && c < nlines) {
            double2 x = res[c * n + j];
            if (sc != 1.0) x = cscale(x, sc);
            if (a.addend) x = cadd(x, add[e]);
            a.out[base + c * cstride + long(j) * rowstride] = x;
        }
    }
}

}  // namespace

// Tile configurations: (threads per work-group, values per work-group).
// The engine's autotuner times them on the actual grid and picks one per axis.
const FFTConfig kFFTConfigs[kNumFFTConfigs] = {{256, 2048}, {256, 1024}, {128, 1024}, {128, 512}, {64, 512}};

int fft_lines(int cfg, int n)
{
    return std::max(1, std::min(32, kFFTConfigs[cfg].values / std::max(n, 1)));
}

template <int NT, int MAXV>
void launch_cfg(const AxisFFTArgs& a, int blocks, size_t lds, hipStream_t s)
{
    // dynamic LDS above 64 KB must be allowed explicitly, once per instantiation
    static std::once_flag raised;  // contexts may be created from several host threads
    std::call_once(raised, [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&axis_fft_kernel<NT, MAXV>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int((2048 + 2 * MAXV) * 16));
    });
    if (g_attributes_only) return;
    hipLaunchKernelGGL((axis_fft_kernel<NT, MAXV>), dim3(blocks), dim3(NT), lds, s, a);
}

void launch_axis_fft(const AxisFFTArgs& a, int cfg, hipStream_t s)
{
    int blocks;
    if (a.axis == 1)
        blocks = int((long(a.P) * a.N0 + a.L - 1) / a.L);
    else
        blocks = a.P * ((a.N1 + a.L - 1) / a.L);
    const size_t lds = size_t(a.n + 2 * a.L * a.n) * sizeof(double2);
    switch (cfg) {
        case 0: launch_cfg<256, 2048>(a, blocks, lds, s); break;
        case 1: launch_cfg<256, 1024>(a, blocks, lds, s); break;
        case 2: launch_cfg<128, 1024>(a, blocks, lds, s); break;
        case 3: launch_cfg<128, 512>(a, blocks, lds, s); break;
        default: launch_cfg<64, 512>(a, blocks, lds, s); break;
    }
}

}  // namespace edus
