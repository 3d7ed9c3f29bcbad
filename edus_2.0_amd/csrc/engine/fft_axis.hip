// Batched 1-D FFTs along one axis of the k grid, staged in LDS (Stockham
// auto-sort, mixed radix 2/3/4/5 and a direct DFT for other primes).
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
// The same kernel does plain forward / inverse transforms along an axis
// (used by the Hartree-Fock convolution and at set-up).
#include "common.hpp"
#include "kernels.hpp"

namespace edus {

namespace {

constexpr double kC5_1 = 0.30901699437494742;   // cos(2 pi/5)
constexpr double kC5_2 = -0.80901699437494742;  // cos(4 pi/5)
constexpr double kS5_1 = 0.95105651629515357;   // sin(2 pi/5)
constexpr double kS5_2 = 0.58778525229247313;   // sin(4 pi/5)
constexpr double kS3 = 0.86602540378443865;     // sin(2 pi/3)

// One Stockham pass of radix R over `lines` lines of length n held in LDS.
// sign = -1 forward, +1 inverse.  tw[e] = exp(-2 pi i e / n).
template <int R>
__device__ void stockham_pass(const double2* __restrict__ src, double2* __restrict__ dst, int lines, int n, int Ns,
                              const double2* __restrict__ tw, double sign)
{
    const int nr = n / R;
    const int tstride = n / (Ns * R);
    const int tasks = lines * nr;
    for (int t = threadIdx.x; t < tasks; t += blockDim.x) {
        const int line = t / nr;
        const int j = t - line * nr;
        const double2* s = src + line * n;
        double2* d = dst + line * n;
        const int k = j % Ns;
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
        double2 y[R];
        if constexpr (R == 2) {
            y[0] = cadd(v[0], v[1]);
            y[1] = csub(v[0], v[1]);
        } else if constexpr (R == 3) {
            double2 t1 = cadd(v[1], v[2]);
            double2 u = csub(v[1], v[2]);
            double2 base = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
            double2 rot = cmul_i(u, sign * kS3);
            y[0] = cadd(v[0], t1);
            y[1] = cadd(base, rot);
            y[2] = csub(base, rot);
        } else if constexpr (R == 4) {
            double2 a02 = cadd(v[0], v[2]), s02 = csub(v[0], v[2]);
            double2 a13 = cadd(v[1], v[3]), s13 = csub(v[1], v[3]);
            double2 rot = cmul_i(s13, sign);
            y[0] = cadd(a02, a13);
            y[1] = cadd(s02, rot);
            y[2] = csub(a02, a13);
            y[3] = csub(s02, rot);
        } else if constexpr (R == 5) {
            double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
            double2 u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
            double2 b1 = make_double2(v[0].x + kC5_1 * t1.x + kC5_2 * t2.x, v[0].y + kC5_1 * t1.y + kC5_2 * t2.y);
            double2 b2 = make_double2(v[0].x + kC5_2 * t1.x + kC5_1 * t2.x, v[0].y + kC5_2 * t1.y + kC5_1 * t2.y);
            double2 r1 = cmul_i(make_double2(kS5_1 * u1.x + kS5_2 * u2.x, kS5_1 * u1.y + kS5_2 * u2.y), sign);
            double2 r2 = cmul_i(make_double2(kS5_2 * u1.x - kS5_1 * u2.x, kS5_2 * u1.y - kS5_1 * u2.y), sign);
            y[0] = cadd(v[0], cadd(t1, t2));
            y[1] = cadd(b1, r1);
            y[4] = csub(b1, r1);
            y[2] = cadd(b2, r2);
            y[3] = csub(b2, r2);
        }
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) d[base + r * Ns] = y[r];
    }
}

// Direct DFT pass for a prime radix R > 5 (rare: grids with such factors).
__device__ void stockham_pass_generic(const double2* __restrict__ src, double2* __restrict__ dst, int lines, int n,
                                      int Ns, int R, const double2* __restrict__ tw, double sign)
{
    const int nr = n / R;
    const int tstride = n / (Ns * R);
    const int tasks = lines * nr * R;
    for (int t = threadIdx.x; t < tasks; t += blockDim.x) {
        const int line = t / (nr * R);
        const int rem = t - line * nr * R;
        const int j = rem / R;
        const int sidx = rem - j * R;  // output index within the butterfly
        const double2* s = src + line * n;
        const int k = j % Ns;
        double2 acc = make_double2(0., 0.);
        for (int q = 0; q < R; ++q) {
            double2 x = s[j + q * nr];
            // twiddle W_n^{q k tstride} * W_R^{q sidx} = W_n^{q (k tstride + sidx nr)}
            int e = (q * (k * tstride + sidx * nr)) % n;
            double2 w = tw[e];
            if (sign > 0) w.y = -w.y;
            acc = cadd(acc, cmul(x, w));
        }
        const int base = (j / Ns) * Ns * R + k;
        dst[line * n + base + sidx * Ns] = acc;
    }
}

// Run all passes; returns the buffer holding the result.
__device__ double2* run_fft(double2* a, double2* b, int lines, int n, const FFTFactors& f, const double2* tw,
                            double sign)
{
    int Ns = 1;
    double2* src = a;
    double2* dst = b;
    for (int i = 0; i < f.count; ++i) {
        const int R = f.radix[i];
        switch (R) {
            case 2: stockham_pass<2>(src, dst, lines, n, Ns, tw, sign); break;
            case 3: stockham_pass<3>(src, dst, lines, n, Ns, tw, sign); break;
            case 4: stockham_pass<4>(src, dst, lines, n, Ns, tw, sign); break;
            case 5: stockham_pass<5>(src, dst, lines, n, Ns, tw, sign); break;
            default: stockham_pass_generic(src, dst, lines, n, Ns, R, tw, sign); break;
        }
        __syncthreads();
        Ns *= R;
        double2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

}  // namespace

// One work-group handles `L` lines of one plane.
//   axis 1: lines are consecutive rows, one contiguous chunk of L*N1 values;
//   axis 0: lines are L consecutive columns of one plane.
// mode 0: derivative  out = addend? + coef * F^-1 diag(i R / n) F in
// mode 1: forward     out = scale * F in          (sign -1)
// mode 2: inverse     out = scale * F^-1 in       (sign +1, unnormalised if scale=1)
// mode 3: convolution out = F^-1 diag(scale * mult(R)) F in
__global__ void __launch_bounds__(256) axis_fft_kernel(AxisFFTArgs a)
{
    extern __shared__ double2 lds[];
    const int n = a.n;
    double2* twl = lds;                       // n twiddles
    double2* bufA = lds + n;                  // L*n
    double2* bufB = bufA + a.L * n;           // L*n
    const int Nk = a.N0 * a.N1;

    for (int i = threadIdx.x; i < n; i += blockDim.x) twl[i] = a.tw[i];

    // locate the lines
    int plane, first, nlines;
    if (a.axis == 1) {
        const long g0 = long(blockIdx.x) * a.L;  // global row index p*N0 + i0
        const long total = long(a.P) * a.N0;
        nlines = int(min(long(a.L), total - g0));
        plane = 0;
        first = 0;
        const double2* src = a.in + g0 * a.N1;
        for (int idx = threadIdx.x; idx < nlines * n; idx += blockDim.x) bufA[idx] = src[idx];
    } else {
        const int nblk = (a.N1 + a.L - 1) / a.L;
        plane = blockIdx.x / nblk;
        first = (blockIdx.x - plane * nblk) * a.L;
        nlines = min(a.L, a.N1 - first);
        const double2* src = a.in + long(plane) * Nk + first;
        for (int idx = threadIdx.x; idx < a.N0 * a.L; idx += blockDim.x) {
            const int row = idx / a.L;
            const int c = idx - row * a.L;
            if (c < nlines) bufA[c * n + row] = src[long(row) * a.N1 + c];
        }
    }
    __syncthreads();

    double2* res;
    if (a.mode == 0) {
        double2* x = run_fft(bufA, bufB, nlines, n, a.fac, twl, -1.0);
        for (int idx = threadIdx.x; idx < nlines * n; idx += blockDim.x) {
            const int j = idx % n;
            x[idx] = cmul_i(x[idx], a.coef * double(gamma_centred(j, n)));
        }
        __syncthreads();
        double2* other = (x == bufA) ? bufB : bufA;
        res = run_fft(x, other, nlines, n, a.fac, twl, +1.0);
    } else if (a.mode == 3) {
        // forward along the axis, multiply by scale * mult(R) point-wise, inverse
        double2* x = run_fft(bufA, bufB, nlines, n, a.fac, twl, -1.0);
        for (int idx = threadIdx.x; idx < nlines * n; idx += blockDim.x) {
            const int c = idx / n;
            const int j = idx - c * n;
            long off;
            if (a.axis == 1)
                off = (long(blockIdx.x) * a.L + c) * a.N1 + j;
            else
                off = long(plane) * Nk + long(j) * a.N1 + first + c;
            x[idx] = cscale(x[idx], a.scale * a.mult[off]);
        }
        __syncthreads();
        double2* other = (x == bufA) ? bufB : bufA;
        res = run_fft(x, other, nlines, n, a.fac, twl, +1.0);
    } else {
        res = run_fft(bufA, bufB, nlines, n, a.fac, twl, a.mode == 1 ? -1.0 : +1.0);
        if (a.scale != 1.0)
            for (int idx = threadIdx.x; idx < nlines * n; idx += blockDim.x) res[idx] = cscale(res[idx], a.scale);
        __syncthreads();
    }

    // epilogue: optional addend, store
    if (a.axis == 1) {
        const long g0 = long(blockIdx.x) * a.L;
        double2* dst = a.out + g0 * a.N1;
        const double2* add = a.addend ? a.addend + g0 * a.N1 : nullptr;
        for (int idx = threadIdx.x; idx < nlines * n; idx += blockDim.x) {
            double2 v = res[idx];
            if (add) v = cadd(v, add[idx]);
            dst[idx] = v;
        }
    } else {
        double2* dst = a.out + long(plane) * Nk + first;
        const double2* add = a.addend ? a.addend + long(plane) * Nk + first : nullptr;
        for (int idx = threadIdx.x; idx < a.N0 * a.L; idx += blockDim.x) {
            const int row = idx / a.L;
            const int c = idx - row * a.L;
            if (c < nlines) {
                double2 v = res[c * n + row];
                if (add) v = cadd(v, add[long(row) * a.N1 + c]);
                dst[long(row) * a.N1 + c] = v;
            }
        }
    }
}

}  // namespace edus

namespace edus {

void launch_axis_fft(const AxisFFTArgs& a, hipStream_t s)
{
    int blocks;
    if (a.axis == 1)
        blocks = int((long(a.P) * a.N0 + a.L - 1) / a.L);
    else
        blocks = a.P * ((a.N1 + a.L - 1) / a.L);
    const size_t lds = size_t(a.n + 2 * a.L * a.n) * sizeof(double2);
    static size_t configured = 0;
    if (lds > configured) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&axis_fft_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        configured = lds;
    }
    hipLaunchKernelGGL(axis_fft_kernel, dim3(blocks), dim3(256), lds, s, a);
}

}  // namespace edus
