// Shared device-side helpers of the EDUS MI355X engine.
//
// Storage: an operator field over the k grid (nb x nb matrix per k-point) is
// held as "planes", one per matrix element, each plane a contiguous N0 x N1
// array of complex doubles in the reference's k order (index i1 + N1*i0,
// MeshGrid.cpp:117-127).  Two layouts:
//   FULL       nb*nb planes, p(n,m) = n*nb + m.  Same arithmetic as the
//              reference, which propagates every element (BlockMatrix.hpp:48).
//   HERMITIAN  upper triangle only, P = nb(nb+1)/2 planes,
//              p(n,m) = n*nb - n(n-1)/2 + (m-n) for n <= m; the lower triangle
//              is the conjugate.  Exact for Hermitian dynamics; see
//              DESIGN notes on the reference's even-grid Nyquist term.
#pragma once
#include <hip/hip_runtime.h>

namespace edus {

constexpr int kMaxNb = 16;
constexpr int kMaxFactors = 12;

enum Layout { kFull = 0, kHermitian = 1 };

__host__ __device__ inline int tri_index(int n, int m, int nb)
{
    // n <= m
    return n * nb - (n * (n - 1)) / 2 + (m - n);
}

__host__ __device__ inline int num_planes(int nb, int layout) { return layout == kFull ? nb * nb : nb * (nb + 1) / 2; }

// plane of the diagonal element n
__host__ __device__ inline int diag_plane(int n, int nb, int layout) { return layout == kFull ? n * (nb + 1) : tri_index(n, n, nb); }

__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// a * (i s)
__host__ __device__ inline double2 cmul_i(double2 a, double s) { return make_double2(-a.y * s, a.x * s); }

// element (r, c) of a field stored in `layout` at grid point k
__device__ inline double2 element(const double2* __restrict__ f, int r, int c, int nb, int layout, long Nk, int k)
{
    if (layout == kFull) return f[long(r * nb + c) * Nk + k];
    if (r <= c) return f[long(tri_index(r, c, nb)) * Nk + k];
    double2 v = f[long(tri_index(c, r, nb)) * Nk + k];
    v.y = -v.y;
    return v;
}

// Gamma-centred integer coordinate of FFT index j on an axis of length n
// (MeshGrid.cpp:458-486 with reduce() at 345-371: j >= n/2 maps to j - n).
__host__ __device__ inline int gamma_centred(int j, int n) { return (2 * j >= n) ? j - n : j; }

// Monitor maxima (bit patterns of non-negative doubles, combined with integer
// atomicMax) are spread over kMonSlots slots, each on its own 128-byte line, so
// that the work-groups of one launch do not queue on one address: slot s holds
// the trace defect at mon[kMonStride s] and the Hermiticity defect at
// mon[kMonHalf + kMonStride s].  The reader takes the maximum over all slots.
constexpr int kMonSlots = 256, kMonStride = 16, kMonHalf = kMonSlots * kMonStride;
constexpr int kMonWords = 2 * kMonHalf;

// Work-groups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// This index gives each XCD a contiguous range of [0, gridDim.x), so that
// neighbouring tiles run side by side in one XCD (and its L2).
__device__ inline int xcd_contiguous_block()
{
    const int nblk = gridDim.x, q = nblk / 8, r = nblk % 8;
    const int x = blockIdx.x % 8, i = blockIdx.x / 8;
    return x * q + min(x, r) + i;
}

// running maximum of a monitor: a NaN value becomes +inf, so it trips the
// monitor instead of being dropped by fmax
__device__ inline double mon_max(double m, double v) { return (v != v) ? __builtin_huge_val() : fmax(m, v); }

// Hartree shift of band n (Coulomb.cpp:150-159): sum_m V_nm rho_mm(R = 0),
// rho(R = 0) being the k-mean of the diagonal planes.  hsum holds every plane's
// plain sum over k (the zero-frequency value of its forward transform, stored
// by the exchange convolution pass), V = hart [nb][nb] (real).
__device__ inline double2 hartree_shift(const double2* hsum, const double* hart, int n, int nb, int layout,
                                        double inv_nk)
{
    double2 h = make_double2(0., 0.);
    for (int m = 0; m < nb; ++m) h = cadd(h, cscale(cscale(hsum[diag_plane(m, nb, layout)], inv_nk), hart[n * nb + m]));
    return h;
}

__device__ inline void mon_slot(unsigned long long* mon, unsigned block, double trace_dev, double herm_dev)
{
    const int o = int(block % kMonSlots) * kMonStride;
    atomicMax(&mon[o], static_cast<unsigned long long>(__double_as_longlong(trace_dev)));
    atomicMax(&mon[kMonHalf + o], static_cast<unsigned long long>(__double_as_longlong(herm_dev)));
}

}  // namespace edus
