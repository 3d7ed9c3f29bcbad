// Layout conversion at the C-ABI boundary, observable reductions and the
// small element-wise kernels of the Hartree-Fock term.
#include "common.hpp"
#include "obs.hpp"

namespace edus {

namespace {

__device__ inline void decode_tri(int p, int nb, int& n, int& m)
{
    int nn = 0;
    while (p >= nb - nn) {
        p -= nb - nn;
        ++nn;
    }
    n = nn;
    m = nn + p;
}

// [Nk][nb][nb] -> planes.  FULL: straight transpose.  HERMITIAN: upper
// triangle of the Hermitian projection (X + X^dagger)/2.
__global__ void to_planes_kernel(const double2* __restrict__ full, double2* __restrict__ planes, int nb, int Nk,
                                 int layout)
{
    const int npl = num_planes(nb, layout);
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= long(npl) * Nk) return;
    const int p = int(idx / Nk);
    const int k = int(idx - long(p) * Nk);
    if (layout == kFull) {
        planes[idx] = full[long(k) * nb * nb + p];
        return;
    }
    int n, m;
    decode_tri(p, nb, n, m);
    const double2 a = full[(long(k) * nb + n) * nb + m];
    const double2 b = full[(long(k) * nb + m) * nb + n];
    planes[idx] = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
}

// planes -> [Nk][nb][nb]
__global__ void from_planes_kernel(const double2* __restrict__ planes, double2* __restrict__ full, int nb, int Nk,
                                   int layout)
{
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long nb2 = long(nb) * nb;
    if (idx >= nb2 * Nk) return;
    const int k = int(idx / nb2);
    const int r = int(idx - long(k) * nb2);
    const int n = r / nb, m = r - (r / nb) * nb;
    full[idx] = element(planes, n, m, nb, layout, Nk, k);
}

// [Nk][nb][nb] complex -> real part as planes (real-valued W(R))
__global__ void to_planes_real_kernel(const double2* __restrict__ full, double* __restrict__ planes, int nb, int Nk,
                                      int layout)
{
    const int npl = num_planes(nb, layout);
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= long(npl) * Nk) return;
    const int p = int(idx / Nk);
    const int k = int(idx - long(p) * Nk);
    int n, m;
    if (layout == kFull) {
        n = p / nb;
        m = p - n * nb;
    } else {
        decode_tri(p, nb, n, m);
    }
    planes[idx] = full[(long(k) * nb + n) * nb + m].x;
}

// [Nk][nb][nb] -> full planes [nb*nb][Nk] (non-Hermitian fields, e.g. U)
__global__ void full_to_planes_kernel(const double2* __restrict__ full, double2* __restrict__ planes, int nb2,
                                      int Nk)
{
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= long(nb2) * Nk) return;
    const int q = int(idx / Nk);
    const int k = int(idx - long(q) * Nk);
    planes[idx] = full[long(k) * nb2 + q];
}

constexpr int kObsBlock = 128;  // keeps the (2nb+6) x block LDS tile of the observables kernel <= 38 KB

// Per-block partial sums of the observables over k:
//   vals[0..nb)       sum_k Re (U^dagger rho U)_nn      (Bloch population, from the rotated field)
//   vals[nb..2nb)     sum_k Re rho_nn                    (Wannier population)
//   vals[2nb..2nb+6)  sum_k Tr(V_d rho)  (re, im per d)
__global__ void __launch_bounds__(kObsBlock) observables_kernel(ObsArgs a)
{
    const int nb = a.nb, Nk = a.Nk, L = a.layout;
    const int nv = 2 * nb + 6;
    extern __shared__ double sh[];  // [nv][kObsBlock]
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    for (int v = 0; v < nv; ++v) sh[v * kObsBlock + threadIdx.x] = 0.;
    if (k < Nk) {
        for (int n = 0; n < nb; ++n) sh[(nb + n) * kObsBlock + threadIdx.x] = element(a.rho, n, n, nb, L, Nk, k).x;
        if (a.rho_bloch)
            for (int n = 0; n < nb; ++n) sh[n * kObsBlock + threadIdx.x] = a.rho_bloch[long(n * nb + n) * Nk + k].x;
        for (int d = 0; d < 3; ++d) {
            if (!a.vel[d]) continue;
            double2 tr = make_double2(0., 0.);
            for (int i = 0; i < nb; ++i)
                for (int j = 0; j < nb; ++j)
                    tr = cadd(tr, cmul(element(a.vel[d], i, j, nb, L, Nk, k), element(a.rho, j, i, nb, L, Nk, k)));
            sh[(2 * nb + 2 * d) * kObsBlock + threadIdx.x] = tr.x;
            sh[(2 * nb + 2 * d + 1) * kObsBlock + threadIdx.x] = tr.y;
        }
    }
    __syncthreads();
    for (int s = kObsBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int v = 0; v < nv; ++v) sh[v * kObsBlock + threadIdx.x] += sh[v * kObsBlock + threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int v = 0; v < nv; ++v) a.partial[long(blockIdx.x) * nv + v] = sh[v * kObsBlock];
}

// Deterministic final reduction of per-block partials (fixed order).
__global__ void reduce_partials_kernel(const double* __restrict__ partial, int nblocks, int nv, double scale,
                                       double* __restrict__ out)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double s = 0.;
    for (int b = 0; b < nblocks; ++b) s += partial[long(b) * nv + v];
    out[v] = s * scale;
}

// hdiag = Hartree shift of B stacked instances from their plane sums over k
// (set-up: the shift of the reference state, folded into H0); the step
// kernels form the same shift with the same function.
__global__ void hartree_sums_kernel(const double2* __restrict__ hsum, const double* __restrict__ hart, int nb,
                                    int layout, int P, double inv_nk, double2* __restrict__ hdiag)
{
    const int inst = blockIdx.x, n = threadIdx.x;
    if (n < nb)
        hdiag[long(inst) * nb + n] = hartree_shift(hsum + long(inst) * P, hart + long(inst) * nb * nb, n, nb, layout,
                                                   inv_nk);
}

__global__ void axpy_planes_kernel(double2* __restrict__ y, const double2* __restrict__ x, double a, long n)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = make_double2(y[i].x + a * x[i].x, y[i].y + a * x[i].y);
}

__global__ void to_f32_kernel(const double2* __restrict__ in, float2* __restrict__ out, long n)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2(float(in[i].x), float(in[i].y));
}

// H_nn(k) += scale * d_n for every k
__global__ void add_diag_kernel(double2* __restrict__ H, const double2* __restrict__ d, double scale, int nb, int Nk,
                                int layout)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= long(nb) * Nk) return;
    const int n = int(i / Nk);
    const int k = int(i - long(n) * Nk);
    double2& h = H[long(diag_plane(n, nb, layout)) * Nk + k];
    h = make_double2(h.x + scale * d[n].x, h.y + scale * d[n].y);
}

inline unsigned grid1d(long n, int bs) { return unsigned((n + bs - 1) / bs); }

// Trace and Hermiticity monitors of B stacked states ([B][P][Nk] planes in
// `layout`), one thread per (instance, k): max |Tr rho - nf|, and the largest
// |rho_nm - conj(rho_mn)| (FULL) or 2 |Im rho_nn| (HERMITIAN), into the
// monitor slots as the commutator kernel folds its own checks
__global__ void __launch_bounds__(kObsBlock) state_monitor_kernel(const double2* __restrict__ rho, int nb, int Nk,
                                                                  int layout, int P, int B, double nf,
                                                                  unsigned long long* __restrict__ mon)
{
    const long i = long(blockIdx.x) * kObsBlock + threadIdx.x;
    double mt = 0., mh = 0.;
    if (i < long(B) * Nk) {
        const int inst = int(i / Nk);
        const int k = int(i - long(inst) * Nk);
        const double2* r = rho + long(inst) * P * Nk + k;
        double2 tr = make_double2(-nf, 0.);
        for (int n = 0; n < nb; ++n) {
            const double2 v = r[long(diag_plane(n, nb, layout)) * Nk];
            tr = cadd(tr, v);
            if (layout != kFull) mh = mon_max(mh, 2.0 * fabs(v.y));
        }
        mt = mon_max(mt, sqrt(tr.x * tr.x + tr.y * tr.y));
        if (layout == kFull)
            for (int n = 0; n < nb; ++n)
                for (int m = n; m < nb; ++m) {
                    const double2 x = r[long(n * nb + m) * Nk], z = r[long(m * nb + n) * Nk];
                    mh = mon_max(mh, sqrt((x.x - z.x) * (x.x - z.x) + (x.y + z.y) * (x.y + z.y)));
                }
    }
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) {
        mt = fmax(mt, __shfl_xor(mt, w));
        mh = fmax(mh, __shfl_xor(mh, w));
    }
    __shared__ double sm[2 * (kObsBlock / 64)];
    if ((threadIdx.x & 63) == 0) {
        sm[threadIdx.x >> 6] = mt;
        sm[kObsBlock / 64 + (threadIdx.x >> 6)] = mh;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kObsBlock / 64; ++w) {
            mt = fmax(mt, sm[w]);
            mh = fmax(mh, sm[kObsBlock / 64 + w]);
        }
        mon_slot(mon, blockIdx.x, mt, mh);
    }
}

}  // namespace

void launch_state_monitor(const double2* rho, int nb, int Nk, int layout, int P, int B, double nf,
                          unsigned long long* mon, hipStream_t s)
{
    const long n = long(B) * Nk;
    hipLaunchKernelGGL(state_monitor_kernel, dim3(unsigned((n + kObsBlock - 1) / kObsBlock)), dim3(kObsBlock), 0, s,
                       rho, nb, Nk, layout, P, B, nf, mon);
}

void launch_to_planes(const double2* full, double2* planes, int nb, int Nk, int layout, hipStream_t s)
{
    const long n = long(num_planes(nb, layout)) * Nk;
    hipLaunchKernelGGL(to_planes_kernel, dim3(grid1d(n, 256)), dim3(256), 0, s, full, planes, nb, Nk, layout);
}
void launch_from_planes(const double2* planes, double2* full, int nb, int Nk, int layout, hipStream_t s)
{
    hipLaunchKernelGGL(from_planes_kernel, dim3(grid1d(long(nb) * nb * Nk, 256)), dim3(256), 0, s, planes, full, nb,
                       Nk, layout);
}
void launch_to_planes_real(const double2* full, double* planes, int nb, int Nk, int layout, hipStream_t s)
{
    const long n = long(num_planes(nb, layout)) * Nk;
    hipLaunchKernelGGL(to_planes_real_kernel, dim3(grid1d(n, 256)), dim3(256), 0, s, full, planes, nb, Nk, layout);
}
void launch_full_to_planes(const double2* full, double2* planes, int nb, int Nk, hipStream_t s)
{
    hipLaunchKernelGGL(full_to_planes_kernel, dim3(grid1d(long(nb) * nb * Nk, 256)), dim3(256), 0, s, full, planes,
                       nb * nb, Nk);
}
int observables_blocks(int Nk) { return int(grid1d(Nk, kObsBlock)); }
void launch_observables(const ObsArgs& a, double* out, hipStream_t s)
{
    const int nv = 2 * a.nb + 6;
    const int blocks = observables_blocks(a.Nk);
    hipLaunchKernelGGL(observables_kernel, dim3(blocks), dim3(kObsBlock), size_t(nv) * kObsBlock * sizeof(double), s,
                       a);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(64), 0, s, a.partial, blocks, nv, 1.0 / double(a.Nk),
                       out);
}
void launch_hartree_sums(const double2* hsum, const double* hart, int nb, int layout, int P, int B, double inv_nk,
                         double2* hdiag, hipStream_t s)
{
    hipLaunchKernelGGL(hartree_sums_kernel, dim3(B), dim3(64), 0, s, hsum, hart, nb, layout, P, inv_nk, hdiag);
}
void launch_axpy(double2* y, const double2* x, double a, long n, hipStream_t s)
{
    hipLaunchKernelGGL(axpy_planes_kernel, dim3(grid1d(n, 256)), dim3(256), 0, s, y, x, a, n);
}
void launch_to_f32(const double2* in, float2* out, long n, hipStream_t s)
{
    hipLaunchKernelGGL(to_f32_kernel, dim3(grid1d(n, 256)), dim3(256), 0, s, in, out, n);
}
void launch_add_diag(double2* H, const double2* d, double scale, int nb, int Nk, int layout, hipStream_t s)
{
    hipLaunchKernelGGL(add_diag_kernel, dim3(grid1d(long(nb) * Nk, 256)), dim3(256), 0, s, H, d, scale, nb, Nk,
                       layout);
}

}  // namespace edus
