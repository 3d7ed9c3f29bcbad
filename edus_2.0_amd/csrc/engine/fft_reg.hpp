// In-register DFTs of compile-time size with compile-time twiddles.
//
// rdft<N, S>(v) replaces v[0..N) by  y_k = sum_j v_j exp(S 2 pi i j k / N),
// S = -1 (forward) or +1 (inverse, unscaled), natural order in and out.
// Sizes 2, 3, 4, 5, 8 are written-out butterflies; other primes use the
// symmetric direct form; composite sizes split N = P*Q (Cooley-Tukey,
// j = jp*Q + jq, k = kp + P*kq) recursively.  All loops are unrolled at
// compile time (sfor) so every index and twiddle is a constant and the
// values stay in VGPRs.
//
// The twiddle constants come from a compile-time Taylor evaluation of sin and
// cos on [0, pi/4] after exact integer reduction of the angle 2 pi e / N to one
// octant.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

#include "common.hpp"

namespace edus {
namespace reg {

template <typename F, int... I>
__host__ __device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
// f(std::integral_constant<int, i>) for i = 0 .. N-1, unrolled
template <int N, typename F>
__host__ __device__ __forceinline__ void sfor(F&& f)
{
    sfor_impl(f, std::make_integer_sequence<int, N>{});
}

constexpr double kPi = 3.14159265358979323846264338327950288;

// Taylor series, |x| <= pi/4: the 2*24th term is far below double rounding
constexpr double taylor_sin(double x)
{
    double t = x, s = x;
    for (int i = 1; i < 24; ++i) {
        t *= -x * x / double((2 * i) * (2 * i + 1));
        s += t;
    }
    return s;
}
constexpr double taylor_cos(double x)
{
    double t = 1.0, s = 1.0;
    for (int i = 1; i < 24; ++i) {
        t *= -x * x / double((2 * i - 1) * (2 * i));
        s += t;
    }
    return s;
}

struct cd {
    double re, im;
};

// exp(S 2 pi i e / n)
constexpr cd root(long e, long n, int S)
{
    const long m = ((e % n) + n) % n;
    const long q = (4 * m) / n;      // quadrant
    const long r = 4 * m - q * n;    // angle inside the quadrant = (pi/2) r / n
    double c = 0.0, s = 0.0;
    if (2 * r <= n) {
        const double x = kPi / 2 * double(r) / double(n);
        c = taylor_cos(x);
        s = taylor_sin(x);
    } else {
        const double x = kPi / 2 * double(n - r) / double(n);
        c = taylor_sin(x);
        s = taylor_cos(x);
    }
    for (long i = 0; i < q; ++i) {  // times i^q
        const double t = c;
        c = -s;
        s = t;
    }
    return cd{c, S * s};
}

// x * exp(S 2 pi i E / NN), with the exact cases (1, -1, +-i) as moves
template <long E, long NN, int S>
__host__ __device__ __forceinline__ double2 rot(double2 x)
{
    constexpr long m = ((E % NN) + NN) % NN;
    if constexpr (m == 0) {
        return x;
    } else if constexpr (2 * m == NN) {
        return make_double2(-x.x, -x.y);
    } else if constexpr (4 * m == NN) {  // S i
        return S > 0 ? make_double2(-x.y, x.x) : make_double2(x.y, -x.x);
    } else if constexpr (4 * m == 3 * NN) {  // -S i
        return S > 0 ? make_double2(x.y, -x.x) : make_double2(-x.y, x.x);
    } else {
        constexpr cd w = root(E, NN, S);
        return make_double2(x.x * w.re - x.y * w.im, x.x * w.im + x.y * w.re);
    }
}

constexpr int split_of(int N)
{
    if (N % 8 == 0 && N != 8) return 8;
    if (N % 4 == 0 && N != 4) return 4;
    for (int p = 2; p < N; ++p)
        if (N % p == 0) return p;
    return N;
}

// times S i
template <int S>
__host__ __device__ __forceinline__ double2 mul_si(double2 x)
{
    return S > 0 ? make_double2(-x.y, x.x) : make_double2(x.y, -x.x);
}

template <int N, int S>
__host__ __device__ __forceinline__ void rdft(double2 (&v)[N])
{
    static_assert(S == 1 || S == -1, "sign");
    if constexpr (N == 1) {
    } else if constexpr (N == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (N == 3) {
        constexpr double h = 0.86602540378443864676;  // sin(2 pi/3)
        const double2 t1 = cadd(v[1], v[2]);
        const double2 u = csub(v[1], v[2]);
        const double2 b = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
        const double2 r = mul_si<S>(cscale(u, h));
        v[0] = cadd(v[0], t1);
        v[1] = cadd(b, r);
        v[2] = csub(b, r);
    } else if constexpr (N == 4) {
        const double2 a02 = cadd(v[0], v[2]), s02 = csub(v[0], v[2]);
        const double2 a13 = cadd(v[1], v[3]), s13 = mul_si<S>(csub(v[1], v[3]));
        v[0] = cadd(a02, a13);
        v[1] = cadd(s02, s13);
        v[2] = csub(a02, a13);
        v[3] = csub(s02, s13);
    } else if constexpr (N == 8) {
        double2 a[4], b[4];
        sfor<4>([&](auto k_) {
            constexpr int k = decltype(k_)::value;
            a[k] = cadd(v[k], v[k + 4]);
            b[k] = rot<k, 8, S>(csub(v[k], v[k + 4]));
        });
        rdft<4, S>(a);
        rdft<4, S>(b);
        sfor<4>([&](auto k_) {
            constexpr int k = decltype(k_)::value;
            v[2 * k] = a[k];
            v[2 * k + 1] = b[k];
        });
    } else if constexpr (split_of(N) == N) {
        // prime N: y_s = v0 + sum_{q<=(N-1)/2} (v_q + v_{N-q}) cos + S i (v_q - v_{N-q}) sin
        constexpr int H = (N - 1) / 2;
        double2 p[H + 1], m[H + 1];
        sfor<H>([&](auto q_) {
            constexpr int q = decltype(q_)::value + 1;
            p[q] = cadd(v[q], v[N - q]);
            m[q] = csub(v[q], v[N - q]);
        });
        double2 y[N];
        double2 sum = v[0];
        sfor<H>([&](auto q_) {
            constexpr int q = decltype(q_)::value + 1;
            sum = cadd(sum, p[q]);
        });
        y[0] = sum;
        sfor<H>([&](auto s_) {
            constexpr int s = decltype(s_)::value + 1;
            double2 c = v[0], t = make_double2(0.0, 0.0);
            sfor<H>([&](auto q_) {
                constexpr int q = decltype(q_)::value + 1;
                constexpr cd w = root(long(q) * s, N, 1);  // cos, sin of 2 pi q s / N
                c = make_double2(c.x + w.re * p[q].x, c.y + w.re * p[q].y);
                t = make_double2(t.x + w.im * m[q].x, t.y + w.im * m[q].y);
            });
            const double2 r = mul_si<S>(t);
            y[s] = cadd(c, r);
            y[N - s] = csub(c, r);
        });
        sfor<N>([&](auto s_) {
            constexpr int s = decltype(s_)::value;
            v[s] = y[s];
        });
    } else {
        constexpr int P = split_of(N), Q = N / P;
        double2 t[N];  // t[kp*Q + jq]
        sfor<Q>([&](auto jq_) {
            constexpr int jq = decltype(jq_)::value;
            double2 u[P];
            sfor<P>([&](auto jp_) {
                constexpr int jp = decltype(jp_)::value;
                u[jp] = v[jp * Q + jq];
            });
            rdft<P, S>(u);
            sfor<P>([&](auto kp_) {
                constexpr int kp = decltype(kp_)::value;
                t[kp * Q + jq] = rot<long(jq) * kp, N, S>(u[kp]);
            });
        });
        sfor<P>([&](auto kp_) {
            constexpr int kp = decltype(kp_)::value;
            double2 u[Q];
            sfor<Q>([&](auto jq_) {
                constexpr int jq = decltype(jq_)::value;
                u[jq] = t[kp * Q + jq];
            });
            rdft<Q, S>(u);
            sfor<Q>([&](auto kq_) {
                constexpr int kq = decltype(kq_)::value;
                v[kp + P * kq] = u[kq];
            });
        });
    }
}

}  // namespace reg
}  // namespace edus
