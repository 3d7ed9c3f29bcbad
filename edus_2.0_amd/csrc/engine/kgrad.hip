// k-space finite-difference gradient (gradient_space "k"), the reference's
// non-default alternative to the spectral derivative:
//   Derivative(k) += sum_s w_s sum_{b in shell s} (b.u) F(k + b)
// (kGradient_definitions.hpp:19-31), with the shells of the Gamma-centred k
// mesh and their weights from kGradient::initialize (kGradient.cpp:25-210).
//
// Set-up (host): mesh vectors sorted into shells by Cartesian norm; the
// smallest number of shells whose weights w = A^+ q satisfy A w = q (the
// B1 condition sum_s w_s sum_b b_a b_b = delta_ab on the active directions);
// every b of those shells kept as a grid offset (m0, m1) with its Cartesian
// vector and weight.  The pseudo-inverse follows Matrix<double>::pseudoinv
// (Geometry/Matrix.cpp:130-161), including its treatment of singular values
// <= 1e-8 (multiplied by s instead of 1/s).
//
// Kernel: one thread per output element; the stencil is read uniformly (a
// handful of entries), neighbours come through the caches, so one launch
// reads F and writes the derivative once (2 plane sets of HBM traffic,
// against 5 for the two-pass spectral derivative).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "../../../include/edus_gpu.h"
#include "common.hpp"
#include "kernels.hpp"

namespace edus {

namespace {

constexpr double kThreshold = 1.e-08;  // Constants.hpp:23

// MeshGrid::reduce (MeshGrid.cpp:345-371) on one crystal coordinate, period 1
double fold(double x, double lo, double up)
{
    while (x < lo && std::abs(x - lo) > kThreshold) x += 1.0;
    while (x > up || std::abs(x - up) < kThreshold) x -= 1.0;
    return x;
}

// P = A^+ (n x m) for a real m x n matrix A (row-major) by one-sided Jacobi:
// A V = [a_1 .. a_n] with orthogonal columns, s_j = |a_j|, u_j = a_j / s_j,
// P = sum_j v_j sinv_j u_j^T with sinv = 1/s (s > 1e-8) or s (otherwise), i.e.
// v_j a_j^T / s_j^2 or v_j a_j^T.
std::vector<double> pinv(std::vector<double> A, int m, int n)
{
    std::vector<double> V(size_t(n) * n, 0.0);
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                double al = 0, be = 0, ga = 0;
                for (int r = 0; r < m; ++r) {
                    al += A[r * n + i] * A[r * n + i];
                    be += A[r * n + j] * A[r * n + j];
                    ga += A[r * n + i] * A[r * n + j];
                }
                if (std::abs(ga) <= 1e-15 * std::sqrt(al * be) || ga == 0.0) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::abs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                for (int r = 0; r < m; ++r) {
                    const double x = A[r * n + i], y = A[r * n + j];
                    A[r * n + i] = c * x - sn * y;
                    A[r * n + j] = sn * x + c * y;
                }
                for (int r = 0; r < n; ++r) {
                    const double x = V[r * n + i], y = V[r * n + j];
                    V[r * n + i] = c * x - sn * y;
                    V[r * n + j] = sn * x + c * y;
                }
            }
        if (!rotated) break;
    }
    std::vector<double> P(size_t(n) * m, 0.0);
    for (int j = 0; j < n; ++j) {
        double s2 = 0;
        for (int r = 0; r < m; ++r) s2 += A[r * n + j] * A[r * n + j];
        const double sv = std::sqrt(s2);
        const double f = sv > 1.e-8 ? 1.0 / s2 : 1.0;
        for (int i = 0; i < n; ++i)
            for (int r = 0; r < m; ++r) P[i * m + r] += V[i * n + j] * f * A[r * n + j];
    }
    return P;
}

// One work-group per (row i0, plane): the row index, the instance and with it
// E and every stencil coefficient w_b (b.E) are uniform over the group; the
// threads run along the row (contiguous k), so each stencil tap is a shifted
// coalesced row read served mostly from cache.  Work-groups are dealt to the
// 8 XCDs round-robin; the (row, plane) order is remapped so that each XCD
// gets a contiguous range of rows and the neighbour rows a tap reads sit in
// its own L2 (read once from HBM, not once per XCD).  Taps are taken kTaps at
// a time: all their loads are issued before the first one is used.
constexpr int kTaps = 8;

__global__ void __launch_bounds__(128) kgrad_kernel(const double2* __restrict__ in, double2* __restrict__ out, int N0,
                                                    int N1, int P, const KStencilPoint* __restrict__ st,
                                                    int nst, double e0, double e1, double e2,
                                                    const double* __restrict__ E_tab)
{
    long L;
    {
        const long nblk = gridDim.x, q = nblk / 8, r = nblk % 8;
        const long x = blockIdx.x % 8, i = blockIdx.x / 8;
        L = x * q + min(x, r) + i;
    }
    const int i0 = int(L % N0);
    {
        const int plane = int(L / N0);  // instance * P + p
        const int inst = plane / P;
        double E0 = e0, E1 = e1, E2 = e2;
        if (E_tab) {
            E0 = E_tab[inst * 3 + 0];
            E1 = E_tab[inst * 3 + 1];
            E2 = E_tab[inst * 3 + 2];
        }
        const long Nk = long(N0) * N1;
        const double2* pin = in + long(plane) * Nk;
        double2* pout = out + long(plane) * Nk + long(i0) * N1;
        for (int i1 = threadIdx.x; i1 < N1; i1 += blockDim.x) {
            double2 acc = make_double2(0.0, 0.0);
            for (int b0 = 0; b0 < nst; b0 += kTaps) {
                double2 v[kTaps];
                double c[kTaps];
#pragma unroll
                for (int t = 0; t < kTaps; ++t) {
                    v[t] = make_double2(0.0, 0.0);
                    c[t] = 0.0;
                    if (b0 + t < nst) {
                        const KStencilPoint& q = st[b0 + t];
                        int j0 = i0 + q.m0, j1 = i1 + q.m1;
                        if (j0 < 0) j0 += N0;
                        if (j0 >= N0) j0 -= N0;
                        if (j1 < 0) j1 += N1;
                        if (j1 >= N1) j1 -= N1;
                        const double bdotu = q.b[0] * E0 + q.b[1] * E1 + q.b[2] * E2;
                        c[t] = q.w * bdotu;
                        v[t] = pin[long(j0) * N1 + j1];
                    }
                }
#pragma unroll
                for (int t = 0; t < kTaps; ++t)
                    if (b0 + t < nst) {
                        acc.x += c[t] * v[t].x;
                        acc.y += c[t] * v[t].y;
                    }
            }
            pout[i1] = acc;
        }
    }
}

}  // namespace

std::vector<KStencilPoint> kgradient_stencil(const double A[3][3], int N0, int N1)
{
    if (N0 < 1 || N1 < 1) throw std::invalid_argument("grid sizes must be positive");
    // reciprocal basis B = 2 pi (A^-1)^T (Model.hpp:46-47)
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                       A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                       A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    if (det == 0.0) throw std::invalid_argument("singular lattice");
    double inv[3][3];
    inv[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
    inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
    inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
    inv[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
    inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
    inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
    inv[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
    inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
    inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
    const double kPi = 3.14159265358979323846;
    double Bm[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Bm[i][j] = 2.0 * kPi * inv[j][i];

    // Gamma-centred mesh in [-1/2, 1/2) (MeshGrid.cpp:458-486), Cartesian norms
    const int Nk = N0 * N1;
    std::vector<double> c0(Nk), c1(Nk), norm(Nk);
    std::vector<std::array<double, 3>> cart(Nk);
    for (int i0 = 0; i0 < N0; ++i0)
        for (int i1 = 0; i1 < N1; ++i1) {
            const int ik = i1 + N1 * i0;
            c0[ik] = fold(double(i0) / N0, -0.5, 0.5);
            c1[ik] = fold(double(i1) / N1, -0.5, 0.5);
            for (int i = 0; i < 3; ++i) cart[ik][i] = Bm[i][0] * c0[ik] + Bm[i][1] * c1[ik];
            norm[ik] = std::sqrt(cart[ik][0] * cart[ik][0] + cart[ik][1] * cart[ik][1] + cart[ik][2] * cart[ik][2]);
        }
    // SortInShells (kGradient.cpp:44-85): distinct norms (within 1e-8), sorted;
    // each point goes to the first norm within 1e-7; the zero shell dropped
    std::vector<double> norms;
    for (int ik = 0; ik < Nk; ++ik) {
        bool found = false;
        for (double v : norms)
            if (std::abs(v - norm[ik]) < kThreshold) {
                found = true;
                break;
            }
        if (!found) norms.push_back(norm[ik]);
    }
    std::sort(norms.begin(), norms.end());
    std::vector<std::vector<int>> shells(norms.size());
    for (int ik = 0; ik < Nk; ++ik) {
        size_t s = 0;
        while (s < norms.size() && !(std::abs(norms[s] - norm[ik]) < 1.e-07)) ++s;
        shells[s].push_back(ik);
    }
    if (!(std::abs(norms[0]) < kThreshold)) throw std::runtime_error("k mesh without k = 0");
    shells.erase(shells.begin());

    // Calculate_nshellsAndweights (kGradient.cpp:88-125), GradientMatrix (:161-180)
    const int al[6] = {0, 1, 2, 0, 0, 1}, be[6] = {0, 1, 2, 1, 2, 2};
    const double q[6] = {double(N0 > 1), double(N1 > 1), 0.0, 0.0, 0.0, 0.0};
    std::vector<double> w;
    int nshells = 0;
    for (nshells = 1;; ++nshells) {
        if (nshells > int(shells.size()))
            throw std::runtime_error("k-space gradient: no set of shells satisfies the B1 condition");
        std::vector<double> Am(size_t(6) * nshells, 0.0);
        for (int j = 0; j < 6; ++j)
            for (int s = 0; s < nshells; ++s)
                for (int ik : shells[s]) Am[j * nshells + s] += cart[ik][al[j]] * cart[ik][be[j]];
        const std::vector<double> P = pinv(Am, 6, nshells);
        w.assign(nshells, 0.0);
        for (int s = 0; s < nshells; ++s)
            for (int j = 0; j < 6; ++j) w[s] += P[s * 6 + j] * q[j];
        double res = 0.0;
        for (int j = 0; j < 6; ++j) {
            double g = 0.0;
            for (int s = 0; s < nshells; ++s) g += Am[j * nshells + s] * w[s];
            res += (g - q[j]) * (g - q[j]);
        }
        if (std::sqrt(res) < kThreshold) break;
    }
    // Find_kpb (kGradient.cpp:182-210): k + b on the periodic mesh = offset (m0, m1)
    std::vector<KStencilPoint> st;
    for (int s = 0; s < nshells; ++s)
        for (int ik : shells[s]) {
            KStencilPoint p{};
            p.m0 = int(std::lround(c0[ik] * N0));
            p.m1 = int(std::lround(c1[ik] * N1));
            p.w = w[s];
            p.shell = s;
            for (int d = 0; d < 3; ++d) p.b[d] = cart[ik][d];
            st.push_back(p);
        }
    return st;
}

void launch_kgrad(const double2* in, double2* out, int N0, int N1, int P, int B, const KStencilPoint* st, int nst,
                  const double E[3], const double* E_tab, hipStream_t s)
{
    const long nblk = long(N0) * P * std::max(B, 1);
    hipLaunchKernelGGL(kgrad_kernel, dim3(unsigned(nblk)), dim3(128), 0, s, in, out, N0, N1, P, st, nst,
                       E ? E[0] : 0.0, E ? E[1] : 0.0, E ? E[2] : 0.0, E_tab);
}

}  // namespace edus

extern "C" int edus_kgrad_stencil(const double A[3][3], int N0, int N1, double* rows, int cap)
{
    try {
        const std::vector<edus::KStencilPoint> st = edus::kgradient_stencil(A, N0, N1);
        if (rows && cap >= int(st.size()))
            for (size_t i = 0; i < st.size(); ++i) {
                double* r = rows + 7 * i;
                r[0] = st[i].shell;
                r[1] = st[i].m0;
                r[2] = st[i].m1;
                r[3] = st[i].b[0];
                r[4] = st[i].b[1];
                r[5] = st[i].b[2];
                r[6] = st[i].w;
            }
        return int(st.size());
    } catch (const std::exception&) {
        return EDUS_ERR_UNSUPPORTED;
    }
}
