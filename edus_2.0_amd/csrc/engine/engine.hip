// Engine context and the low-level C ABI (include/edus_gpu.h).
//
// Device memory (FP64 planes, `npl` of them per field -- nb*nb in the FULL
// layout, nb(nb+1)/2 in the HERMITIAN one; each plane a contiguous N0 x N1
// grid, see common.hpp):
//   y, acc, aux        integrator state: rho_n, RK4 accumulator, stage input
//   G, D               axis-0 derivative, full E.grad rho
//   f                  right-hand side output of edus_rhs
//   T                  Hartree-Fock work planes (R-space convolution)
//   H0                 H0(k) (+ the constant rho0 part of the HSEX term)
//   xi[3]              position operator r_d(k)
//   W                  screened potential W(R), real planes
//   U                  eigenvectors, full planes [nb*nb][Nk]
//   vel[3]             velocity operator
//   hist[5]            Adams-Bashforth right-hand-side history
//   stage              host transfer staging, [Nk][nb][nb]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../../include/edus_gpu.h"
#include "batch.hpp"
#include "engine_internal.hpp"
#include "kernels.hpp"
#include "obs.hpp"

#define HIP_CHECK EDUS_HIP_CHECK

namespace edus {

thread_local std::string g_last_error;
thread_local bool g_attributes_only = false;

void set_error(const std::string& s) { g_last_error = s; }

namespace {

// Whether a host field [Nk][nb][nb] is Hermitian at every k: the largest
// |a_nm - conj(a_mn)| within 16 ulp of the largest |a| (the DFT of a
// Hermitian H(R) leaves 1-2 ulp; r(k) is made exactly Hermitian).  Reading
// the lower triangle as the conjugate of the upper one then changes H by at
// most that defect.
bool hermitian_field(const double* f, int nb, int Nk)
{
    double dev = 0.0, mx = 0.0;
    for (long k = 0; k < Nk; ++k) {
        const double* a = f + 2 * k * nb * nb;
        for (int n = 0; n < nb; ++n)
            for (int m = 0; m < nb; ++m) {
                const double* x = a + 2 * (n * nb + m);
                mx = std::max(mx, std::hypot(x[0], x[1]));
                if (m > n) {
                    const double* y = a + 2 * (m * nb + n);
                    dev = std::max(dev, std::hypot(x[0] - y[0], x[1] + y[1]));
                }
            }
    }
    return dev <= 16.0 * std::numeric_limits<double>::epsilon() * mx;
}

FFTFactors factorize(int n)
{
    // radix 8 first (fewest passes), then 4, 2, 3, 5 and the primes 7, 11, 13
    FFTFactors f{};
    int x = n;
    for (int r : {8, 4, 2, 3, 5, 7, 11, 13})
        while (x % r == 0) {
            if (f.count >= kMaxFactors) throw EngineError(EDUS_ERR_UNSUPPORTED, "grid size has too many factors");
            f.radix[f.count++] = r;
            x /= r;
        }
    if (x != 1)
        throw EngineError(EDUS_ERR_UNSUPPORTED,
                          "grid size " + std::to_string(n) + " has a prime factor above 13 (FFT radices 2..13)");
    return f;
}

std::vector<double2> twiddles(int n)
{
    std::vector<double2> t(n);
    for (int j = 0; j < n; ++j) {
        // symmetric index so |angle| <= pi; long double for the rounding
        const int jj = (2 * j > n) ? j - n : j;
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * jj / n;
        t[j] = make_double2(double(cosl(a)), double(-sinl(a)));
    }
    return t;
}

// shapes already tuned in this process: (axis length, other axis, planes) -> config
std::map<std::tuple<int, int, int, int>, FFTChoice>& tuned_shapes()
{
    static std::map<std::tuple<int, int, int, int>, FFTChoice> m;
    return m;
}

// commutator engine and tile per (nb, layout, k-points of all instances,
// precision, kept engine, exchange, upper-triangle fields): the last two
// change the bytes per launch
using CommKey = std::tuple<int, int, int, int, int, int, int>;
std::map<CommKey, std::tuple<int, int, int>>& tuned_comm()
{
    static std::map<CommKey, std::tuple<int, int, int>> m;
    return m;
}

// one-pass gradient or not per (N0, N1, planes of all instances)
std::map<std::tuple<int, int, int>, bool>& tuned_grad()
{
    static std::map<std::tuple<int, int, int>, bool> m;
    return m;
}

// Guards the caches: contexts may be created from several host threads.
// Held for a whole tuning pass, so one shape is timed once.
std::mutex& tuning_mutex()
{
    static std::mutex m;
    return m;
}

// The HERMITIAN layout stores the upper triangle of the exchange self energy
// and conjugates it into the lower one.  Sigma(k) is Hermitian for Hermitian
// rho only if the real W satisfies W_mn(R) = W_nm(-R) with -R taken on the
// periodic FFT grid (index (N - j) mod N, so the even-grid Nyquist index is
// its own mirror).  W: [Nk][nb][nb] interleaved complex (real part used).
bool exchange_is_hermitian(const double* W, int nb, int N0, int N1)
{
    const size_t M = size_t(nb) * nb;
    double asym = 0., scale = 0.;
    for (int j0 = 0; j0 < N0; ++j0)
        for (int j1 = 0; j1 < N1; ++j1) {
            const size_t iR = size_t(j0) * N1 + j1;
            const size_t mR = size_t((N0 - j0) % N0) * N1 + (N1 - j1) % N1;
            for (int n = 0; n < nb; ++n)
                for (int m = 0; m < nb; ++m) {
                    const double a = W[2 * (iR * M + size_t(m) * nb + n)], b = W[2 * (mR * M + size_t(n) * nb + m)];
                    asym = std::max(asym, std::fabs(a - b));
                    scale = std::max(scale, std::fabs(a));
                }
        }
    return asym <= 1e-12 * std::max(scale, 1e-300);
}

}  // namespace

void Engine::alloc(void** p, size_t bytes)
{
    HIP_CHECK(hipMalloc(p, bytes));
    owned.push_back(*p);
    HIP_CHECK(hipMemsetAsync(*p, 0, bytes, s));
}

Engine::Engine(const edus_desc& d, const EngineOptions& opt)
{
    nb = d.nb;
    N0 = d.N0;
    N1 = d.N1;
    Nk = N0 * N1;
    layout = d.layout;
    prec = d.precision;
    nf = d.filledbands;
    coulomb = d.coulomb != 0;
    peierls = d.peierls != 0;
    if (d.gradient_space != 0 && d.gradient_space != 1) throw EngineError(EDUS_ERR_ARG, "gradient_space must be 0 (R) or 1 (k)");
    kgrad = d.gradient_space == 1;
    std::memcpy(A, d.A, sizeof(A));
    {
        std::array<double, 9> a;
        std::memcpy(a.data(), d.A, sizeof(d.A));
        Ainst.push_back(a);
    }
    if (nb < 1 || nb > kMaxNb) throw EngineError(EDUS_ERR_UNSUPPORTED, "nb must be in 1..16");
    if (N0 < 1 || N1 < 1) throw EngineError(EDUS_ERR_ARG, "grid sizes must be positive");
    if (N0 > 2048 || N1 > 2048) throw EngineError(EDUS_ERR_UNSUPPORTED, "grid axes longer than 2048 are not supported");
    if (layout != EDUS_LAYOUT_FULL && layout != EDUS_LAYOUT_HERMITIAN) throw EngineError(EDUS_ERR_ARG, "unknown layout");
    if (!d.H0k) throw EngineError(EDUS_ERR_ARG, "H0k is required");
    if (prec != EDUS_PREC_FP64 && prec != EDUS_PREC_MIXED) throw EngineError(EDUS_ERR_ARG, "unknown precision");
    if (coulomb && (!d.W || !d.hartree || (!d.rho0k && opt.initial_filled < 0)))
        throw EngineError(EDUS_ERR_ARG, "coulomb requires W, hartree and rho0k");
    if (coulomb && layout == EDUS_LAYOUT_HERMITIAN && !exchange_is_hermitian(d.W, nb, N0, N1))
        throw EngineError(EDUS_ERR_UNSUPPORTED,
                          "HERMITIAN layout with exchange needs W_mn(R) = W_nm(-R) on the FFT grid; this W breaks it "
                          "(even grids put the unpaired Nyquist R in it, and so does the reference's Wannier-centre "
                          "accumulation): use the FULL layout");
    npl = num_planes(nb, layout);

    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const size_t pb = plane_bytes();
    const size_t full = size_t(nb) * nb * Nk * sizeof(double2);
    for (double2** p : {&y, &acc, &aux, &G, &D, &f, &H0}) alloc(reinterpret_cast<void**>(p), pb);
    alloc(reinterpret_cast<void**>(&stage), full);
    alloc(reinterpret_cast<void**>(&mon), kMonWords * sizeof(unsigned long long));
    const int nv = 2 * nb + 6;
    alloc(reinterpret_cast<void**>(&obs_partial), size_t(observables_blocks(Nk)) * nv * sizeof(double));
    alloc(reinterpret_cast<void**>(&obs_out), size_t(nv) * sizeof(double));

    upload(d.H0k, H0);
    for (int k = 0; k < 3; ++k) {
        if (d.xik[k]) {
            alloc(reinterpret_cast<void**>(&xi[k]), pb);
            upload(d.xik[k], xi[k]);
        }
        if (d.vel[k]) {
            alloc(reinterpret_cast<void**>(&vel[k]), pb);
            upload(d.vel[k], vel[k]);
        }
    }
    if (d.U) {
        alloc(reinterpret_cast<void**>(&U), full);
        HIP_CHECK(hipMemcpyAsync(stage, d.U, full, hipMemcpyHostToDevice, s));
        launch_full_to_planes(stage, U, nb, Nk, s);
        HIP_CHECK(hipStreamSynchronize(s));
    }

    if (kgrad) {
        try {
            kst = kgradient_stencil(A, N0, N1);
        } catch (const std::exception& e) {
            throw EngineError(EDUS_ERR_UNSUPPORTED, e.what());
        }
        alloc(reinterpret_cast<void**>(&kst_dev), kst.size() * sizeof(KStencilPoint));
        HIP_CHECK(hipMemcpyAsync(kst_dev, kst.data(), kst.size() * sizeof(KStencilPoint), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    fac[0] = factorize(N0);
    fac[1] = factorize(N1);
    for (int ax = 0; ax < 2; ++ax) {
        const int n = ax == 0 ? N0 : N1;
        auto t = twiddles(n);
        alloc(reinterpret_cast<void**>(&tw[ax]), n * sizeof(double2));
        HIP_CHECK(hipMemcpyAsync(tw[ax], t.data(), n * sizeof(double2), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }

    tri_ok = layout == kFull && !coulomb && !peierls && hermitian_field(d.H0k, nb, Nk);
    for (int k = 0; k < 3; ++k)
        if (d.xik[k]) tri_ok = tri_ok && hermitian_field(d.xik[k], nb, Nk);
    tri = tri_ok;
    alloc_f32_fields();
    make_tri_fields();  // H0 is final here: the exchange fold below excludes them
    autotune();
    if (peierls) {
        alloc(reinterpret_cast<void**>(&Hp), pb);
        alloc(reinterpret_cast<void**>(&Tq), pb);
        alloc(reinterpret_cast<void**>(&ptab[0]), size_t(N0) * sizeof(double2));
        alloc(reinterpret_cast<void**>(&ptab[1]), size_t(N1) * sizeof(double2));
    }
    if (opt.velocity_det > 0.0) build_velocity(opt.velocity_det);
    if (opt.initial_filled >= 0) {
        if (!U) throw EngineError(EDUS_ERR_ARG, "the initial condition needs the eigenvectors U");
        initial_condition(opt.initial_filled);
    }

    if (coulomb) {
        alloc(reinterpret_cast<void**>(&T), pb);
        alloc(reinterpret_cast<void**>(&W), size_t(npl) * Nk * sizeof(double));
        HIP_CHECK(hipMemcpyAsync(stage, d.W, full, hipMemcpyHostToDevice, s));
        launch_to_planes_real(stage, W, nb, Nk, layout, s);
        std::vector<double> hr(nb * nb);
        for (int i = 0; i < nb * nb; ++i) hr[i] = d.hartree[2 * i];  // V is real (ModelCoulomb.cpp:168-176)
        alloc(reinterpret_cast<void**>(&hart), nb * nb * sizeof(double));
        alloc(reinterpret_cast<void**>(&hdiag), nb * sizeof(double2));
        alloc(reinterpret_cast<void**>(&hsum), size_t(npl) * sizeof(double2));
        HIP_CHECK(hipMemcpyAsync(hart, hr.data(), nb * nb * sizeof(double), hipMemcpyHostToDevice, s));
        // H_ee = -W o (rho - rho0) + Hartree(rho - rho0) (Coulomb.cpp:150-171).  The rho0
        // part does not change in time: fold it into H0 once, H0 <- H0 - HF[rho0].
        if (d.rho0k)
            upload(d.rho0k, aux);
        else  // the reference state is the initial condition computed above
            HIP_CHECK(hipMemcpyAsync(aux, y, pb, hipMemcpyDeviceToDevice, s));
        fock(aux);  // T = Sigma[rho0], hsum = the plane sums of rho0
        launch_hartree_sums(hsum, hart, nb, layout, npl, 1, 1.0 / double(Nk), hdiag, s);  // Hartree[rho0]
        launch_axpy(H0, T, -1.0, long(npl) * Nk, s);
        launch_add_diag(H0, hdiag, -1.0, nb, Nk, layout, s);
        HIP_CHECK(hipMemsetAsync(aux, 0, pb, s));
    }
    make_f32_fields();  // from the final H0
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipGetLastError());
}

// FP32 copies of H0 and xi for the mixed-precision commutator, which then
// reads 8 B instead of 16 B per element of each.  Not with exchange (H0 holds
// -Sigma[rho0], which cancels against Sigma[rho]: rounding it first would cost
// that cancellation) nor in the Peierls gauge (H is rebuilt every stage; a
// conversion pass would move more bytes than it saves).
bool Engine::uses_f32_fields() const { return prec == EDUS_PREC_MIXED && !coulomb && !peierls; }

void Engine::alloc_f32_fields()
{
    if (!uses_f32_fields()) return;
    const size_t fb = size_t(B) * npl * Nk * sizeof(float2);
    alloc(reinterpret_cast<void**>(&H0f), fb);
    for (int d = 0; d < 3; ++d)
        if (xi[d]) alloc(reinterpret_cast<void**>(&xif[d]), fb);
}

void Engine::make_f32_fields()
{
    if (!uses_f32_fields()) return;
    const long n = long(B) * npl * Nk;
    launch_to_f32(H0, H0f, n, s);
    for (int d = 0; d < 3; ++d)
        if (xi[d]) launch_to_f32(xi[d], xif[d], n, s);
    if (H0uf) {
        const long nu = long(B) * num_planes(nb, kHermitian) * Nk;
        launch_to_f32(H0u, H0uf, nu, s);
        for (int d = 0; d < 3; ++d)
            if (xiu[d]) launch_to_f32(xiu[d], xiuf[d], nu, s);
    }
}

// Upper-triangle copies of H0 and xi (and their FP32 versions, filled by
// make_f32_fields) from the FULL planes: row n's planes (n, n..nb-1) are
// consecutive in both layouts.
void Engine::make_tri_fields()
{
    if (!tri_ok) return;
    const int nu = num_planes(nb, kHermitian);
    const size_t ub = size_t(B) * nu * Nk * sizeof(double2);
    auto upper = [&](const double2* full, double2* up) {
        for (int i = 0; i < B; ++i)
            for (int n = 0; n < nb; ++n)
                HIP_CHECK(hipMemcpyAsync(up + (size_t(i) * nu + tri_index(n, n, nb)) * Nk,
                                         full + (size_t(i) * npl + size_t(n) * (nb + 1)) * Nk,
                                         size_t(nb - n) * Nk * sizeof(double2), hipMemcpyDeviceToDevice, s));
    };
    alloc(reinterpret_cast<void**>(&H0u), ub);
    upper(H0, H0u);
    for (int d = 0; d < 3; ++d)
        if (xi[d]) {
            alloc(reinterpret_cast<void**>(&xiu[d]), ub);
            upper(xi[d], xiu[d]);
        }
    if (uses_f32_fields()) {
        const size_t fb = size_t(B) * nu * Nk * sizeof(float2);
        alloc(reinterpret_cast<void**>(&H0uf), fb);
        for (int d = 0; d < 3; ++d)
            if (xi[d]) alloc(reinterpret_cast<void**>(&xiuf[d]), fb);
    }
}

void Engine::field_args(CommArgs& c, const bool use_xi[3], int kind) const
{
    const bool t = tri && kind != kCommScalar;  // the FMA engine reads every element
    c.tri = t ? 1 : 0;
    c.H0 = t ? H0u : H0;
    c.H0f = t ? H0uf : H0f;
    for (int d = 0; d < 3; ++d) {
        const bool on = use_xi[d] && xi[d];
        c.xi[d] = on ? (t ? xiu[d] : xi[d]) : nullptr;
        c.xif[d] = on ? (t ? xiuf[d] : xif[d]) : nullptr;
    }
}

// V_d = (-i[r_d, H0] + e_d . grad_k H0) / det, evaluated with the same
// kernels as the right-hand side: f = D - i[H, in] with H = H0 (no field
// terms), in = -r_d and D = e_d . grad_k H0 (Simulation.cpp:538-575).
void Engine::build_velocity(double det)
{
    const size_t pb = plane_bytes();
    for (int d = 0; d < 3; ++d) {
        if (!vel[d]) alloc(reinterpret_cast<void**>(&vel[d]), pb);
        HIP_CHECK(hipMemsetAsync(aux, 0, pb, s));
        if (xi[d]) launch_axpy(aux, xi[d], -1.0, long(npl) * Nk, s);  // aux = -r_d
        const double cu[2] = {A[d][0] / N0, A[d][1] / N1};  // unit field along d
        const double ud[3] = {double(d == 0), double(d == 1), double(d == 2)};
        const double2* Dp = kgrad ? kgradient(H0, ud, nullptr) : gradient(H0, cu, nullptr, cu[0] != 0.0, cu[1] != 0.0);
        HIP_CHECK(hipMemsetAsync(f, 0, pb, s));  // zero base for the scaled store
        CommArgs c{};
        c.Nk = Nk;
        c.P = npl;
        c.rho = aux;
        c.D = Dp;
        c.H0 = H0;
        c.up.x1 = f;
        c.up.out1 = vel[d];
        c.up.b1 = 1.0 / det;
        launch_comm(nb, layout, comm_kind64, comm_tile64, c, s);
    }
    HIP_CHECK(hipMemsetAsync(aux, 0, pb, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

void Engine::observables(double* pop_bloch, double* pop_wannier, double* vel_out, int inst,
                         const double2* rho_override)
{
    const size_t io = size_t(inst) * inst_elems();
    ObsArgs a{};
    a.nb = nb;
    a.Nk = Nk;
    a.layout = layout;
    a.rho = rho_override ? rho_override : y + io;
    a.rho_bloch = nullptr;
    if (U) {  // Bloch populations from U^dagger rho U (Simulation.cpp:439-470), staged in full planes
        rotate(a.rho, stage, kFull, true, inst);
        a.rho_bloch = stage;
    }
    for (int d = 0; d < 3; ++d) a.vel[d] = vel[d] ? vel[d] + io : nullptr;
    a.partial = obs_partial;
    launch_observables(a, obs_out, s);
    const int nv = 2 * nb + 6;
    std::vector<double> h(nv);
    HIP_CHECK(hipMemcpyAsync(h.data(), obs_out, nv * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int n = 0; n < nb; ++n) {
        if (pop_bloch) pop_bloch[n] = U ? ((n < nf) ? 1. - h[n] : h[n]) : 0.0;  // Simulation.cpp:459-465
        if (pop_wannier) pop_wannier[n] = h[nb + n];
    }
    if (vel_out)
        for (int c = 0; c < 6; ++c) vel_out[c] = h[2 * nb + c];
}

// out = U^dagger in U (to_bloch) or U in U^dagger for every k-point of
// instance `inst` (Operator::go_to_bloch / go_to_wannier, Operator.hpp:528-564);
// in: planes in this engine's layout, out: out_layout (kFull, kHermitian or
// kRotKMajor).  Mixed-precision engines rotate on FP32 matrix cores unless
// fp64 is asked for.
void Engine::rotate(const double2* in, double2* out, int out_layout, bool to_bloch, int inst, bool fp64)
{
    if (!U) throw EngineError(EDUS_ERR_STATE, "no eigenvectors: the context was created without U");
    RotateArgs r{};
    r.Nk = Nk;
    r.in = in;
    r.U = U + size_t(inst) * nb * nb * Nk;
    r.out = out;
    r.out_layout = out_layout;
    r.to_bloch = to_bloch ? 1 : 0;
    const int kind = (prec == EDUS_PREC_MIXED && !fp64) ? kCommMfma32 : kCommMfma64;
    TimedLaunch t;
    tbegin("rotate", t);
    launch_rotate(nb, layout, kind, r, s);
    tend(t);
}

// rho = U diag(1 for n < filled, else 0) U^dagger per k-point into y
// (Functional_InitialCondition.hpp:12-32), rotated on the FP64 matrix cores
void Engine::initial_condition(int filled)
{
    const size_t pb = plane_bytes();
    HIP_CHECK(hipMemsetAsync(aux, 0, pb, s));
    std::vector<double2> one(Nk, make_double2(1.0, 0.0));
    for (int n = 0; n < std::min(filled, nb); ++n)
        HIP_CHECK(hipMemcpyAsync(aux + size_t(diag_plane(n, nb, layout)) * Nk, one.data(), Nk * sizeof(double2),
                                 hipMemcpyHostToDevice, s));
    rotate(aux, y, layout, false, 0, true);
    HIP_CHECK(hipStreamSynchronize(s));  // the host vector is read by the copies
    HIP_CHECK(hipMemsetAsync(aux, 0, pb, s));
}

Engine::~Engine()
{
    if (self_batch) delete_batch(self_batch);  // its graph reads this engine's arrays
    if (s) (void)hipStreamSynchronize(s);
    for (auto& e : pending) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (void* p : owned) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
}

void Engine::upload(const double* host_full, double2* dst)
{
    HIP_CHECK(hipMemcpyAsync(stage, host_full, size_t(nb) * nb * Nk * sizeof(double2), hipMemcpyHostToDevice, s));
    launch_to_planes(stage, dst, nb, Nk, layout, s);
    HIP_CHECK(hipStreamSynchronize(s));  // the staging buffer is reused right away
}

void Engine::download(const double2* src, double* host_full)
{
    launch_from_planes(src, stage, nb, Nk, layout, s);
    HIP_CHECK(hipMemcpyAsync(host_full, stage, size_t(nb) * nb * Nk * sizeof(double2), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

// U^dagger src U of instance 0 -> host [Nk][nb][nb] (Operator::go_to_bloch)
void Engine::download_bloch(const double2* src, double* host_full)
{
    rotate(src, stage, kRotKMajor, true);
    HIP_CHECK(hipMemcpyAsync(host_full, stage, size_t(nb) * nb * Nk * sizeof(double2), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

void Engine::tbegin(const char* cls, TimedLaunch& t)
{
    t.cls = nullptr;
    if (!timing) return;
    t.cls = cls;
    HIP_CHECK(hipEventCreate(&t.a));
    HIP_CHECK(hipEventCreate(&t.b));
    HIP_CHECK(hipEventRecord(t.a, s));
}

void Engine::tend(TimedLaunch& t)
{
    if (!t.cls) return;
    HIP_CHECK(hipEventRecord(t.b, s));
    pending.push_back(t);
}

void Engine::collect_timing()
{
    HIP_CHECK(hipStreamSynchronize(s));
    for (auto& t : pending) {
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, t.a, t.b));
        auto& acc_ = tsum[t.cls];
        acc_.first += ms;
        acc_.second += 1;
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    pending.clear();
}

AxisFFTArgs Engine::axis_args(int axis, const FFTChoice& c, int mode, const double2* in, const double2* addend,
                              double2* out, double coef, double scale, const double* mult) const
{
    AxisFFTArgs a{};
    a.in = in;
    a.addend = addend;
    a.out = out;
    a.P = B * npl;
    a.N0 = N0;
    a.N1 = N1;
    a.axis = axis;
    a.n = axis == 0 ? N0 : N1;
    a.L = c.L;
    a.mode = mode;
    a.coef = coef;
    a.scale = scale;
    a.mult = mult;
    a.fac = fac[axis];
    a.tw = tw[axis];
    return a;
}

void Engine::axis_fft(int axis, int mode, const double2* in, const double2* addend, double2* out, double coef,
                      double scale, const double* mult, const char* cls, const double* coef_tab, double2* out2,
                      double2* r0)
{
    if (mode == 5 && (!fft_choice[axis].two_step || !out2))
        throw EngineError(EDUS_ERR_STATE, "mode 5 needs the two-step kernels and a second output");
    AxisFFTArgs a = axis_args(axis, fft_choice[axis], mode, in, addend, out, coef, scale, mult);
    a.out2 = out2;
    a.r0 = r0;
    if (coef_tab) {
        a.coef_tab = coef_tab;
        a.coef_stride = 2;
        a.ppi = npl;
    }
    TimedLaunch t;
    tbegin(cls, t);
    launch_fft(a, fft_choice[axis], s);
    tend(t);
}

namespace {

// Tile candidates for one axis.  Columns (axis 0) are read in row segments
// of L adjacent columns, so whole 128-byte lines (multiples of 8 columns) and
// divisors of the row length (no ragged last group) are tried.
std::vector<FFTChoice> fft_candidates(int axis, int n, int other)
{
    std::vector<FFTChoice> out;
    auto add = [&](FFTChoice c) {
        if (axis == 0) c.L = std::min(c.L, other);
        if (c.L < 1) return;
        for (const FFTChoice& o : out)
            if (o.two_step == c.two_step && o.cfg == c.cfg && o.L == c.L) return;
        out.push_back(c);
    };
    if (two_step_max_lines(n) > 0) {
        for (int variant : {0, 1, 2, 3}) {
            const int maxl = two_step_max_lines(n, variant);
            std::vector<int> Ls{maxl};
            if (axis == 1) {
                // fewer lines per work-group give small grids (90^2 x 64
                // planes: 230 work-groups at the most) more work-groups
                for (int div : {2, 4, 8}) Ls.push_back(std::max(1, maxl / div));
            } else {
                for (int L : {8 * (maxl / 8), 32, 16, 8, 4})
                    if (L >= 1 && L <= maxl) Ls.push_back(L);
                for (int L = std::min(maxl, other); L >= 1; --L) {
                    if (other % L == 0) {
                        Ls.push_back(L);
                        break;
                    }
                }
            }
            for (int L : Ls) {
                FFTChoice c;
                c.two_step = 1;
                c.cfg = variant;
                c.L = L;
                add(c);
            }
        }
    } else {
        for (int cfg = 0; cfg < kNumFFTConfigs; ++cfg) {
            if (fft_lines(cfg, n) * n > kFFTConfigs[cfg].values) continue;  // line longer than the tile
            FFTChoice c;
            c.cfg = cfg;
            c.L = fft_lines(cfg, n);
            add(c);
        }
    }
    return out;
}

}  // namespace

void launch_fft(const AxisFFTArgs& a, const FFTChoice& c, hipStream_t s)
{
    if (c.two_step)
        launch_two_step(a, c.cfg, s);
    else
        launch_axis_fft(a, c.cfg, s);
}

// Tile autotuner: time every candidate tile of the axis' kernel family on this
// engine's own arrays (derivative mode, the hot use) and keep the fastest.
// Results are cached per shape for the life of the process.
void Engine::autotune()
{
    std::lock_guard<std::mutex> lock(tuning_mutex());
    for (int axis = 0; axis < 2; ++axis) {
        const int n = axis == 0 ? N0 : N1;
        const int other = axis == 0 ? N1 : N0;
        if (n < 2) continue;
        const auto key = std::make_tuple(axis, n, other, B * npl);
        auto it = tuned_shapes().find(key);
        if (it != tuned_shapes().end()) {
            fft_choice[axis] = it->second;
            continue;
        }
        const std::vector<FFTChoice> cands = fft_candidates(axis, n, other);
        if (cands.empty()) throw EngineError(EDUS_ERR_UNSUPPORTED, "no FFT tile fits a line of " + std::to_string(n));
        FFTChoice best_c = cands[0];
        if (cands.size() > 1) {
            hipEvent_t e0, e1;
            HIP_CHECK(hipEventCreate(&e0));
            HIP_CHECK(hipEventCreate(&e1));
            float best = 1e30f;
            for (const FFTChoice& c : cands) {
                const AxisFFTArgs a = axis_args(axis, c, 0, y, nullptr, G, 1e-3, 1.0, nullptr);
                launch_fft(a, c, s);  // warm-up
                HIP_CHECK(hipEventRecord(e0, s));
                for (int r = 0; r < 3; ++r) launch_fft(a, c, s);
                HIP_CHECK(hipEventRecord(e1, s));
                HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    best_c = c;
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            HIP_CHECK(hipGetLastError());
        }
        fft_choice[axis] = best_c;
        tuned_shapes()[key] = best_c;
    }

    // commutator engine and tile, whichever is fastest here: for FP64 the FMA
    // engine or the matrix cores (they differ in summation order only), for
    // mixed precision the FP32 matrix cores; tiles give identical results.
    // The Adams-Bashforth kernel (history prefetched, more registers) gets a
    // tile of its own.
    const CommKey ckey = std::make_tuple(nb, layout, Nk * B, prec, force_kind, int(coulomb), int(tri));
    auto cit = tuned_comm().find(ckey);
    if (cit != tuned_comm().end()) {
        comm_kind = std::get<0>(cit->second);
        comm_tile = std::get<1>(cit->second);
        comm_tile_ab = std::get<2>(cit->second);
    } else {
        std::vector<std::pair<int, int>> cands;
        if (prec == EDUS_PREC_MIXED)
            cands = {{kCommMfma32, 0}, {kCommMfma32, 1}, {kCommMfma32, 2}};
        else
            cands = {{kCommScalar, 0}, {kCommMfma64, 0}, {kCommMfma64, 1}, {kCommMfma64, 2}};
        if (force_kind >= 0) {  // keep the members' engine: only the tile is free
            std::vector<std::pair<int, int>> keep;
            for (const auto& kt : cands)
                if (kt.first == force_kind) keep.push_back(kt);
            cands = keep.empty() ? std::vector<std::pair<int, int>>{{force_kind, 0}} : keep;
        }
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        auto time_of = [&](int kind, int tile, const CommArgs& c) {
            launch_comm(nb, layout, kind, tile, c, s);  // warm-up
            HIP_CHECK(hipEventRecord(e0, s));
            for (int r = 0; r < 3; ++r) launch_comm(nb, layout, kind, tile, c, s);
            HIP_CHECK(hipEventRecord(e1, s));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            return ms;
        };
        // the operand sets of the hot uses, so that the candidates see the
        // step's real HBM traffic, not a cache-resident subset.
        // Adams-Bashforth step of order 4: state (read and written in place),
        // H0, one xi, D and three history arrays read, the new history entry
        // written (over one of them: each element is read and written by the
        // same thread)
        auto ab_args = [&](int kind) {
            CommArgs cab = tuning_stage_args(kind);
            cab.up = StageUpdate{};
            cab.up.ab_order = 4;
            cab.up.y = aux;
            cab.up.out_f = G;
            cab.up.hist[1] = acc;
            cab.up.hist[2] = G;
            cab.up.hist[3] = f;
            return cab;
        };
        float best = 1e30f;
        for (const auto& kt : cands) {
            const float ms = time_of(kt.first, kt.second, tuning_stage_args(kt.first));
            if (ms < best) {
                best = ms;
                comm_kind = kt.first;
                comm_tile = kt.second;
            }
        }
        best = 1e30f;
        for (const auto& kt : cands) {
            if (kt.first != comm_kind) continue;
            const float ms = time_of(kt.first, kt.second, ab_args(kt.first));
            if (ms < best) {
                best = ms;
                comm_tile_ab = kt.second;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIP_CHECK(hipGetLastError());
        tuned_comm()[ckey] = std::make_tuple(comm_kind, comm_tile, comm_tile_ab);
    }
    // the spectral derivative: one pass per plane or two axis passes, whichever
    // gives the faster stage (timed with the commutator that consumes D: the
    // one-pass form leaves different data in the caches); they agree to rounding
    grad_fused = false;
    if (grad_fused_available(N0, N1) && !kgrad) {
        const auto gkey = std::make_tuple(N0, N1, B * npl);
        auto git = tuned_grad().find(gkey);
        if (force_grad >= 0) {  // a batch group keeps its members' form (they differ in rounding)
            grad_fused = force_grad == 1;
        } else if (git != tuned_grad().end()) {
            grad_fused = git->second;
        } else {
            hipEvent_t e0, e1;
            HIP_CHECK(hipEventCreate(&e0));
            HIP_CHECK(hipEventCreate(&e1));
            const double cg[2] = {1e-3, 1e-3};
            CommArgs c = tuning_stage_args(comm_kind);
            float best[2] = {1e30f, 1e30f};
            for (int trial = 0; trial < 2; ++trial)
                for (int f = 0; f < 2; ++f) {
                    grad_fused = f == 1;
                    c.D = gradient(c.rho, cg, nullptr, true, true);  // warm-up
                    launch_comm(nb, layout, comm_kind, comm_tile, c, s);
                    HIP_CHECK(hipEventRecord(e0, s));
                    for (int r = 0; r < 4; ++r) {
                        gradient(c.rho, cg, nullptr, true, true);
                        launch_comm(nb, layout, comm_kind, comm_tile, c, s);
                    }
                    HIP_CHECK(hipEventRecord(e1, s));
                    HIP_CHECK(hipEventSynchronize(e1));
                    float ms = 0.f;
                    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                    best[f] = std::min(best[f], ms);
                }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            HIP_CHECK(hipGetLastError());
            grad_fused = best[1] < best[0];
            tuned_grad()[gkey] = grad_fused;
        }
    }

    if (prec == EDUS_PREC_MIXED) {
        // set-up products (the velocity operator) stay FP64
        comm_kind64 = kCommMfma64;
        comm_tile64 = 0;
    } else {
        comm_kind64 = comm_kind;
        comm_tile64 = comm_tile;
    }

    for (double2* p : {G, f, acc, aux}) HIP_CHECK(hipMemsetAsync(p, 0, plane_bytes(), s));
    HIP_CHECK(hipStreamSynchronize(s));
    prime_launch_attributes();
}

// The commutator's operands in the RK4 step's heaviest stage (the third),
// for timing candidates on this engine's own arrays (they hold no state yet;
// every coefficient is 0): stage input, H0, one xi, D and two stage arrays
// read, two written.
CommArgs Engine::tuning_stage_args(int kind) const
{
    CommArgs c{};
    c.Nk = Nk;
    c.P = npl;
    c.B = B;
    c.rho = aux;
    bool use_xi[3] = {false, false, false};
    for (int d = 0; d < 3; ++d)
        if (xi[d]) {
            use_xi[d] = true;
            break;
        }
    field_args(c, use_xi, kind);
    c.D = D;
    c.up.x1 = acc;
    c.up.out1 = acc;
    c.up.g1 = 2. / 3.;
    c.up.x2 = y;
    c.up.out2 = f;
    return c;
}

// Raise the dynamic-LDS limit of every kernel this engine may launch with the
// choices just made (every FFT mode of each axis' tile, every tile of the
// commutator engines for both integrators and both H0 / xi storages, the
// rotation, the one-pass gradient) while no stream is being captured: a graph
// captured later then launches them with their full LDS.  Nothing is
// launched.  Also run after edus_set_tuning.
void Engine::prime_launch_attributes()
{
    AttributesOnly guard;
    for (int axis = 0; axis < 2; ++axis) {
        if ((axis == 0 ? N0 : N1) < 2) continue;
        for (int mode = 0; mode <= 5; ++mode) {
            if (mode == 5 && (axis != 1 || !fft_choice[axis].two_step)) continue;
            const AxisFFTArgs a = axis_args(axis, fft_choice[axis], mode, y, nullptr, G, 0.0, 1.0, nullptr);
            launch_fft(a, fft_choice[axis], s);
        }
    }
    CommArgs c{};
    c.Nk = Nk;
    c.P = npl;
    c.B = B;
    c.H0 = H0;
    c.H0f = H0f ? H0f : reinterpret_cast<const float2*>(H0);  // not launched
    for (int t = 0; t < (tri_ok ? 2 : 1); ++t)  // both H0 / xi storages (separate kernels)
        for (int ab = 0; ab < 2; ++ab) {
            c.up = StageUpdate{};
            c.up.ab_order = ab ? 4 : 0;
            for (int kind : {comm_kind, comm_kind64}) {
                c.tri = t == 1 && kind != kCommScalar;
                if (t == 1 && !c.tri) continue;
                for (int tile : {0, 1, 2})
                    if (tile == 0 || kind != kCommScalar) launch_comm(nb, layout, kind, tile, c, s);
            }
        }
    if (U) {
        RotateArgs r{};
        r.Nk = Nk;
        launch_rotate(nb, layout, kCommMfma64, r, s);
        launch_rotate(nb, layout, kCommMfma32, r, s);
    }
    if (grad_fused_available(N0, N1)) {
        GradFusedArgs g{};
        g.N0 = N0;
        g.N1 = N1;
        launch_grad_fused(g, s);
    }
}

// D = E . grad_k rho  (null when the field has no component along the grid):
// c_j = a_j . E for lattice vector j, spectral derivative along grid axis j
// (Functional_SourceTerm.hpp:20-24, kGradient_definitions.hpp:35-41).
const double2* Engine::gradient(const double2* in, const double c[2], const double* ctab, bool use0, bool use1,
                                double2* fwd1, bool* fwd1_done)
{
    if (fwd1_done) *fwd1_done = false;
    use0 = use0 && N0 > 1;
    use1 = use1 && N1 > 1;
    if (!use0 && !use1) return nullptr;
    const double* t0 = ctab ? ctab + 0 : nullptr;
    const double* t1 = ctab ? ctab + 1 : nullptr;
    // the row pass also stores the forward row transform of `in` (the
    // exchange term's first pass) when asked to and the row kernels are the
    // two-step ones
    const bool fuse = fwd1 && use1 && fft_choice[1].two_step && !(use0 && grad_fused);
    const int row_mode = fuse ? 5 : 0;
    if (fuse && fwd1_done) *fwd1_done = true;
    if (use0 && use1 && grad_fused) {
        GradFusedArgs g{};
        g.in = in;
        g.out = D;
        g.P = B * npl;
        g.N0 = N0;
        g.N1 = N1;
        g.coef[0] = c[0];
        g.coef[1] = c[1];
        g.coef_tab = ctab;
        g.coef_stride = 2;
        g.ppi = npl;
        g.tw0 = tw[0];
        g.tw1 = tw[1];
        TimedLaunch t;
        tbegin("grad_fused", t);
        launch_grad_fused(g, s);
        tend(t);
    } else if (use0 && use1) {
        axis_fft(0, 0, in, nullptr, G, c[0], 1.0, nullptr, "grad_col", t0);
        axis_fft(1, row_mode, in, G, D, c[1], 1.0, nullptr, "grad_row", t1, fuse ? fwd1 : nullptr);
    } else if (use0) {
        axis_fft(0, 0, in, nullptr, D, c[0], 1.0, nullptr, "grad_col", t0);
    } else {
        axis_fft(1, row_mode, in, nullptr, D, c[1], 1.0, nullptr, "grad_row", t1, fuse ? fwd1 : nullptr);
    }
    return D;
}

const double2* Engine::kgradient(const double2* in, const double E[3], const double* E_tab)
{
    TimedLaunch t;
    tbegin("kgrad", t);
    launch_kgrad(in, D, N0, N1, npl, B, kst_dev, int(kst.size()), E, E_tab, s);
    tend(t);
    return D;
}

// T = F^-1[ -W o (F rho)/N ]  (screened exchange in k).  The column pass also
// stores each plane's zero-frequency value, its sum over k, in hsum: the
// commutator forms the Hartree shift from it (hartree_shift), so the Hartree
// term costs no pass of its own.
void Engine::fock(const double2* in, bool row_done)
{
    if (!row_done) axis_fft(1, 1, in, nullptr, T, 0.0, 1.0, nullptr, "hf_row");
    axis_fft(0, 3, T, nullptr, T, 0.0, -1.0 / double(Nk), W, "hf_col", nullptr, nullptr, hsum);
    axis_fft(1, 2, T, nullptr, T, 0.0, 1.0, nullptr, "hf_row");
}

void Engine::transform_to_R(const double2* in, double2* out)
{
    if (layout != kFull || B != 1) throw EngineError(EDUS_ERR_UNSUPPORTED, "transform_to_R needs one FULL-layout context");
    if (N0 < 2 || N1 < 2) throw EngineError(EDUS_ERR_UNSUPPORTED, "transform_to_R needs a 2-D grid");
    axis_fft(1, 1, in, nullptr, G, 0.0, 1.0, nullptr, "to_R_row");
    axis_fft(0, 1, G, nullptr, out, 0.0, 1.0 / double(Nk), nullptr, "to_R_col");
}

void Engine::transform_to_k(const double2* in, double2* out)
{
    if (layout != kFull || B != 1) throw EngineError(EDUS_ERR_UNSUPPORTED, "transform_to_k needs one FULL-layout context");
    if (N0 < 2 || N1 < 2) throw EngineError(EDUS_ERR_UNSUPPORTED, "transform_to_k needs a 2-D grid");
    axis_fft(1, 2, in, nullptr, G, 0.0, 1.0, nullptr, "to_k_row");
    axis_fft(0, 2, G, nullptr, out, 0.0, 1.0, nullptr, "to_k_col");
}

void Engine::evaluate(const double2* in, const double E[3], const StageUpdate& up)
{
    double cg[2];
    for (int j = 0; j < 2; ++j) cg[j] = (A[0][j] * E[0] + A[1][j] * E[1] + A[2][j] * E[2]) / (j == 0 ? N0 : N1);
    bool hf_row = false;  // the exchange term's forward row pass done by the gradient
    const double2* Dp = kgrad ? ((cg[0] != 0.0 || cg[1] != 0.0) ? kgradient(in, E, nullptr) : nullptr)
                              : gradient(in, cg, nullptr, cg[0] != 0.0, cg[1] != 0.0,
                                         coulomb && hf_fuse ? T : nullptr, &hf_row);
    if (coulomb) fock(in, hf_row);
    CommArgs c{};
    c.Nk = Nk;
    c.P = npl;
    c.rho = in;
    c.D = Dp;
    const bool use_xi[3] = {E[0] != 0.0, E[1] != 0.0, E[2] != 0.0};
    field_args(c, use_xi, comm_kind);
    for (int d = 0; d < 3; ++d) c.E[d] = E[d];
    c.sigma = coulomb ? T : nullptr;
    c.hsum = coulomb ? hsum : nullptr;
    c.hart = hart;
    c.inv_nk = 1.0 / double(Nk);
    c.up = up;
    TimedLaunch t;
    tbegin("comm", t);
    launch_comm(nb, layout, comm_kind, tile_for(c.up), c, s);
    tend(t);
}


// The four stage updates of one classical RK4 step (DESolver.hpp:155-187):
//   k1 = f(t, y);        aux = y + dt/2 k1;  acc = y + dt/6 k1     (:166-171)
//   k2 = f(t+dt/2, aux); aux = y + dt/2 k2                         (:174-177)
//   k3 = f(t+dt/2, aux); aux = y + dt k3;
//                        acc += 2/3 (aux_in - y) + dt/3 k3         (:174-182)
//   k4 = f(t+dt, aux);   y = acc + dt/6 k4                         (:184-185)
// The reference adds dt/3 k2 to acc in stage 2; here stage 3 recovers it as
// 2/3 (aux_in - y) = dt/3 k2 from the two arrays it reads anyway, so stage 2
// neither reads nor writes acc (a step moves 26 plane sets instead of 28).
// Stage 1 carries the per-step trace / Hermiticity monitor of the y the step
// starts from (the y it writes is checked by the next step, or by
// check_final_state).
void Engine::rk4_stages(double dt, StageUpdate u[4]) const
{
    for (int i = 0; i < 4; ++i) u[i] = StageUpdate{};
    u[0].x1 = y; u[0].out1 = acc; u[0].b1 = dt / 6.;
    u[0].x2 = y; u[0].out2 = aux; u[0].b2 = dt / 2.;
    u[1].x2 = y; u[1].out2 = aux; u[1].b2 = dt / 2.;
    u[2].x1 = acc; u[2].out1 = acc; u[2].b1 = dt / 3.; u[2].g1 = 2. / 3.;
    u[2].x2 = y; u[2].out2 = aux; u[2].b2 = dt;
    u[3].x1 = acc; u[3].out1 = y; u[3].b1 = dt / 6.;
    u[0].mon = mon; u[0].nf = double(nf);
}

void Engine::rk4_step(double dt, const double* E3x3, int record_order)
{
    state_dirty = true;  // the state it writes is checked by the next step or check_final_state
    StageUpdate u[4];
    rk4_stages(dt, u);
    if (record_order > 0) {
        ensure_hist();
        u[0].out_f = hist[abidx[0]];
    }
    evaluate(y, E3x3 + 0, u[0]);
    evaluate(aux, E3x3 + 3, u[1]);
    evaluate(aux, E3x3 + 3, u[2]);
    evaluate(aux, E3x3 + 6, u[3]);
    if (record_order > 0) {  // the same rotation as ab_step
        const int tmp = abidx[record_order - 1];
        for (int i = record_order - 1; i > 0; --i) abidx[i] = abidx[i - 1];
        abidx[0] = tmp;
    }
}

Engine::Engine(const std::vector<const Engine*>& members)
{
    if (members.empty()) throw EngineError(EDUS_ERR_ARG, "empty batch group");
    const Engine& m0 = *members[0];
    nb = m0.nb;
    N0 = m0.N0;
    N1 = m0.N1;
    Nk = m0.Nk;
    layout = m0.layout;
    prec = m0.prec;
    nf = m0.nf;
    coulomb = m0.coulomb;
    npl = m0.npl;
    B = int(members.size());
    std::memcpy(A, m0.A, sizeof(A));
    for (const Engine* m : members) {
        if (m->nb != nb || m->N0 != N0 || m->N1 != N1 || m->layout != layout || m->prec != prec ||
            m->coulomb != coulomb || m->nf != nf || m->B != 1)
            throw EngineError(EDUS_ERR_ARG, "batch group members differ in shape");
        if (m->peierls) throw EngineError(EDUS_ERR_UNSUPPORTED, "Peierls-gauge contexts cannot be batched");
        if (m->kgrad != m0.kgrad) throw EngineError(EDUS_ERR_ARG, "batch group members differ in gradient space");
        if (m->kgrad) {
            bool same = m->kst.size() == m0.kst.size();
            for (size_t i = 0; same && i < m->kst.size(); ++i)
                same = std::memcmp(&m->kst[i], &m0.kst[i], sizeof(KStencilPoint)) == 0;
            if (!same)
                throw EngineError(EDUS_ERR_UNSUPPORTED,
                                  "k-space gradient contexts with different stencils (lattices) cannot share a group");
        }
        std::array<double, 9> a;
        std::memcpy(a.data(), m->A, sizeof(m->A));
        Ainst.push_back(a);
        HIP_CHECK(hipStreamSynchronize(m->s));
    }

    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const size_t pb = plane_bytes();
    const size_t ib = inst_elems() * sizeof(double2);
    const size_t full = size_t(nb) * nb * Nk * sizeof(double2);
    for (double2** p : {&y, &acc, &aux, &G, &D, &f, &H0}) alloc(reinterpret_cast<void**>(p), pb);
    alloc(reinterpret_cast<void**>(&stage), full);
    alloc(reinterpret_cast<void**>(&mon), kMonWords * sizeof(unsigned long long));
    const int nv = 2 * nb + 6;
    alloc(reinterpret_cast<void**>(&obs_partial), size_t(observables_blocks(Nk)) * nv * sizeof(double));
    alloc(reinterpret_cast<void**>(&obs_out), size_t(nv) * sizeof(double));

    // stacked copies of every member's arrays (zero where a member has none)
    auto stack = [&](double2* dst, auto get, size_t bytes) {
        for (int i = 0; i < B; ++i) {
            const void* src = get(members[i]);
            if (src) HIP_CHECK(hipMemcpyAsync(reinterpret_cast<char*>(dst) + i * bytes, src, bytes,
                                              hipMemcpyDeviceToDevice, s));
        }
    };
    stack(y, [](const Engine* m) { return m->y; }, ib);
    stack(H0, [](const Engine* m) { return m->H0; }, ib);
    for (int d = 0; d < 3; ++d) {
        bool any_xi = false, any_vel = false;
        for (const Engine* m : members) {
            any_xi |= m->xi[d] != nullptr;
            any_vel |= m->vel[d] != nullptr;
        }
        if (any_xi) {
            alloc(reinterpret_cast<void**>(&xi[d]), pb);
            stack(xi[d], [d](const Engine* m) { return m->xi[d]; }, ib);
        }
        if (any_vel) {
            alloc(reinterpret_cast<void**>(&vel[d]), pb);
            stack(vel[d], [d](const Engine* m) { return m->vel[d]; }, ib);
        }
    }
    bool all_u = true;
    for (const Engine* m : members) all_u &= m->U != nullptr;
    if (all_u) {
        alloc(reinterpret_cast<void**>(&U), size_t(B) * full);
        stack(U, [](const Engine* m) { return m->U; }, full);
    }
    kgrad = m0.kgrad;
    if (kgrad) {
        kst = m0.kst;
        alloc(reinterpret_cast<void**>(&kst_dev), kst.size() * sizeof(KStencilPoint));
        HIP_CHECK(hipMemcpyAsync(kst_dev, kst.data(), kst.size() * sizeof(KStencilPoint), hipMemcpyHostToDevice, s));
    }
    fac[0] = m0.fac[0];
    fac[1] = m0.fac[1];
    for (int ax = 0; ax < 2; ++ax) {
        const int n = ax == 0 ? N0 : N1;
        alloc(reinterpret_cast<void**>(&tw[ax]), n * sizeof(double2));
        HIP_CHECK(hipMemcpyAsync(tw[ax], m0.tw[ax], n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    }
    if (coulomb) {
        alloc(reinterpret_cast<void**>(&T), pb);
        alloc(reinterpret_cast<void**>(&W), size_t(B) * npl * Nk * sizeof(double));
        alloc(reinterpret_cast<void**>(&hart), size_t(B) * nb * nb * sizeof(double));
        alloc(reinterpret_cast<void**>(&hsum), size_t(B) * npl * sizeof(double2));
        for (int i = 0; i < B; ++i) {
            HIP_CHECK(hipMemcpyAsync(W + size_t(i) * npl * Nk, members[i]->W, size_t(npl) * Nk * sizeof(double),
                                     hipMemcpyDeviceToDevice, s));
            HIP_CHECK(hipMemcpyAsync(hart + size_t(i) * nb * nb, members[i]->hart, size_t(nb) * nb * sizeof(double),
                                     hipMemcpyDeviceToDevice, s));
        }
    }
    // upper-triangle fields when every member has them; the group reads them
    // as its members do (same_group keeps members with different choices apart)
    tri_ok = true;
    for (const Engine* m : members) tri_ok = tri_ok && m->tri_ok;
    tri = tri_ok && m0.tri;
    hf_fuse = m0.hf_fuse;  // the two exchange row-pass forms agree only to rounding
    alloc_f32_fields();
    make_tri_fields();
    make_f32_fields();
    HIP_CHECK(hipStreamSynchronize(s));
    // the commutator engines differ in summation order and the gradient forms
    // in rounding: a group uses its members' choices so that every member
    // ends exactly where it ends alone
    force_kind = m0.comm_kind;
    force_grad = m0.grad_fused ? 1 : 0;
    autotune();
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipGetLastError());
}

void Engine::ensure_tables(int steps)
{
    if (steps <= tab_steps) return;
    // grown tables replace the old ones; the old allocations stay owned
    // (freed with the engine) because queued work may still read them
    alloc(reinterpret_cast<void**>(&E_dev), size_t(steps) * 3 * B * 3 * sizeof(double));
    alloc(reinterpret_cast<void**>(&coef_dev), size_t(steps) * 3 * B * 2 * sizeof(double));
    tab_steps = steps;
}

void Engine::evaluate_tab(const double2* in, int ts, int j, const bool use_axis[2], const bool use_xi[3],
                          const StageUpdate& up)
{
    const size_t row = size_t(ts) * 3 + j;
    const double c0[2] = {0.0, 0.0};
    bool hf_row = false;  // the exchange term's forward row pass done by the gradient
    const double2* Dp = kgrad ? ((use_axis[0] || use_axis[1]) ? kgradient(in, nullptr, E_dev + row * B * 3) : nullptr)
                              : gradient(in, c0, coef_dev + row * B * 2, use_axis[0], use_axis[1],
                                         coulomb && hf_fuse ? T : nullptr, &hf_row);
    if (coulomb) fock(in, hf_row);
    CommArgs c{};
    c.Nk = Nk;
    c.P = npl;
    c.B = B;
    c.rho = in;
    c.D = Dp;
    field_args(c, use_xi, comm_kind);
    c.E_tab = E_dev + row * B * 3;
    c.sigma = coulomb ? T : nullptr;
    c.hsum = coulomb ? hsum : nullptr;
    c.hart = hart;
    c.inv_nk = 1.0 / double(Nk);
    c.up = up;
    TimedLaunch t;
    tbegin("comm", t);
    launch_comm(nb, layout, comm_kind, tile_for(c.up), c, s);
    tend(t);
}

void Engine::rk4_step_tab(double dt, int ts, const bool use_axis[2], const bool use_xi[3])
{
    state_dirty = true;  // the state it writes is checked by the next step or check_final_state
    StageUpdate u[4];
    rk4_stages(dt, u);
    evaluate_tab(y, ts, 0, use_axis, use_xi, u[0]);
    evaluate_tab(aux, ts, 1, use_axis, use_xi, u[1]);
    evaluate_tab(aux, ts, 1, use_axis, use_xi, u[2]);
    evaluate_tab(aux, ts, 2, use_axis, use_xi, u[3]);
}

// F^-1 diag(exp(i sign A.R)) F applied to (in + sum_d addc_d add_d), one
// axis after the other: A.R = R0 (A.a0) + R1 (A.a1) on the Gamma-centred grid
// is a sum, so the phase factorises (Simulation.cpp:876-899 applies it to the
// whole R-space array at once).
void Engine::peierls_sandwich(const double2* in, const double2* const add[3], const double addc[3],
                              const double Av[3], int sign, double2* out)
{
    double th[2];
    for (int j = 0; j < 2; ++j) th[j] = sign * (Av[0] * A[0][j] + Av[1] * A[1][j] + Av[2] * A[2][j]);
    const double2* src = in;
    const bool two = N0 > 1 && N1 > 1;
    // the phase of each frequency once per sandwich, not once per element
    launch_phase_tables(ptab[0], N0 > 1 ? N0 : 0, th[0], 1.0 / N0, ptab[1], N1 > 1 ? N1 : 0, th[1], 1.0 / N1, s);
    for (int axis = 0; axis < 2; ++axis) {
        const int n = axis == 0 ? N0 : N1;
        if (n < 2) continue;
        AxisFFTArgs a = axis_args(axis, fft_choice[axis], 4, src, nullptr, (two && axis == 0) ? Tq : out, th[axis],
                                  1.0 / n, nullptr);
        a.ptab = ptab[axis];
        if (src == in)
            for (int d = 0; d < 3; ++d) {  // a zero field component reads nothing
                const bool use = add && add[d] && addc && addc[d] != 0.0;
                a.in_add[d] = use ? add[d] : nullptr;
                a.in_add_c[d] = use ? addc[d] : 0.0;
            }
        TimedLaunch t;
        tbegin(axis == 0 ? "peierls_col" : "peierls_row", t);
        launch_fft(a, fft_choice[axis], s);
        tend(t);
        src = (two && axis == 0) ? Tq : out;
    }
    if (N0 < 2 && N1 < 2) {  // one k-point: the phase at R = 0 is 1
        HIP_CHECK(hipMemcpyAsync(out, in, plane_bytes(), hipMemcpyDeviceToDevice, s));
        for (int d = 0; d < 3; ++d)
            if (add && add[d] && addc[d] != 0.0) launch_axpy(out, add[d], addc[d], long(npl) * Nk, s);
    }
}

void Engine::evaluate_peierls(const double2* in, const double E[3], const double Av[3], const StageUpdate& up)
{
    if (!peierls) throw EngineError(EDUS_ERR_STATE, "context was not created for the Peierls gauge");
    // H(k) = S_A[H0 + E.xi] (+ exchange of the current state, whose phases
    // cancel; H0 already holds the rho0 terms, inside the phase as the
    // reference has them)
    const double2* add[3] = {xi[0], xi[1], xi[2]};
    // S_A[H0 + E.xi] depends on E and A only: RK4 stages 2 and 3 share them
    // (t + dt/2), and so do a step's stage 4 and the next step's stage 1 when
    // the host's times agree to the bit.  Hp is then reused; the sandwich
    // would recompute the same values bit for bit.
    const double key[6] = {E[0], E[1], E[2], Av[0], Av[1], Av[2]};
    if (!hp_valid || std::memcmp(key, hp_key, sizeof(key)) != 0) {
        peierls_sandwich(H0, add, E, Av, +1, Hp);
        std::memcpy(hp_key, key, sizeof(key));
        hp_valid = true;
    }
    if (coulomb) fock(in);
    CommArgs c{};
    c.Nk = Nk;
    c.P = npl;
    c.rho = in;
    c.D = nullptr;
    c.H0 = Hp;
    c.sigma = coulomb ? T : nullptr;
    c.hsum = coulomb ? hsum : nullptr;
    c.hart = hart;
    c.inv_nk = 1.0 / double(Nk);
    c.up = up;
    TimedLaunch t;
    tbegin("comm", t);
    launch_comm(nb, layout, comm_kind, tile_for(c.up), c, s);
    tend(t);
}

void Engine::rk4_step_peierls(double dt, const double* E3x3, const double* A3x3)
{
    state_dirty = true;  // the state it writes is checked by the next step or check_final_state
    StageUpdate u[4];
    rk4_stages(dt, u);
    evaluate_peierls(y, E3x3 + 0, A3x3 + 0, u[0]);
    evaluate_peierls(aux, E3x3 + 3, A3x3 + 3, u[1]);
    evaluate_peierls(aux, E3x3 + 3, A3x3 + 3, u[2]);
    evaluate_peierls(aux, E3x3 + 6, A3x3 + 6, u[3]);
}

void Engine::ab_step_peierls(double dt, int order, const double* E3, const double* A3)
{
    state_dirty = true;  // the state it writes is checked by the next step or check_final_state
    ensure_hist();
    static const double b4[5] = {55. / 24., -59. / 24., 37. / 24., -3. / 8., 0.};
    static const double b5[5] = {1901. / 720., -2774. / 720., 2616. / 720., -1274. / 720., 251. / 720.};
    const double* beta = order == 4 ? b4 : b5;
    StageUpdate u{};
    u.out_f = hist[abidx[0]];
    u.ab_order = order;
    u.y = y;
    for (int i = 0; i < order; ++i) {
        u.ab_c[i] = dt * beta[i];
        u.hist[i] = hist[abidx[i]];
    }
    u.mon = mon; u.nf = double(nf);
    evaluate_peierls(y, E3, A3, u);
    const int tmp = abidx[order - 1];
    for (int i = order - 1; i > 0; --i) abidx[i] = abidx[i - 1];
    abidx[0] = tmp;
}

void Engine::ensure_hist()
{
    for (int i = 0; i < 5; ++i)
        if (!hist[i]) alloc(reinterpret_cast<void**>(&hist[i]), plane_bytes());
}

void Engine::ab_step(double dt, int order, const double* E3)
{
    state_dirty = true;  // the state it writes is checked by the next step or check_final_state
    ensure_hist();
    // DESolver.hpp:110-126 coefficients, 202-220 update
    static const double b4[5] = {55. / 24., -59. / 24., 37. / 24., -3. / 8., 0.};
    static const double b5[5] = {1901. / 720., -2774. / 720., 2616. / 720., -1274. / 720., 251. / 720.};
    const double* beta = order == 4 ? b4 : b5;
    StageUpdate u{};
    u.out_f = hist[abidx[0]];
    u.ab_order = order;
    u.y = y;
    for (int i = 0; i < order; ++i) {
        u.ab_c[i] = dt * beta[i];
        u.hist[i] = hist[abidx[i]];
    }
    u.mon = mon; u.nf = double(nf);
    evaluate(y, E3, u);
    const int tmp = abidx[order - 1];
    for (int i = order - 1; i > 0; --i) abidx[i] = abidx[i - 1];
    abidx[0] = tmp;
}

// The step kernels check the state each step starts from (RK4 stage 1, every
// AB step); the state the last step wrote is checked here, once, before the
// monitors are read.
void Engine::check_final_state()
{
    if (!state_dirty) return;
    launch_state_monitor(y, nb, Nk, layout, npl, B, double(nf), mon, s);
    state_dirty = false;
}

void Engine::peek_monitor(double out[2])
{
    check_final_state();
    std::vector<unsigned long long> m(kMonWords);
    HIP_CHECK(hipMemcpyAsync(m.data(), mon, kMonWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    unsigned long long best[2] = {0, 0};
    for (int h = 0; h < 2; ++h)
        for (int i = 0; i < kMonSlots; ++i) best[h] = std::max(best[h], m[size_t(h) * kMonHalf + size_t(i) * kMonStride]);
    std::memcpy(out, best, sizeof(best));
}

// With limits set: the trace and Hermiticity defects of every state the steps
// since the last edus_check wrote must stay within them (a limit <= 0 is not
// checked); a NaN or inf anywhere (the monitors hold +inf then) always fails.
void Engine::enforce_monitor_limits()
{
    if (!mon_limits_on()) return;
    double v[2];
    peek_monitor(v);
    const char* name[2] = {"trace", "Hermiticity"};
    for (int i = 0; i < 2; ++i) {
        const bool bad = !std::isfinite(v[i]) || (mon_limit[i] > 0.0 && v[i] > mon_limit[i]);
        if (bad)
            throw EngineError(EDUS_ERR_MONITOR, std::string(name[i]) + " monitor " + std::to_string(v[i]) +
                                                    " exceeds its limit " + std::to_string(mon_limit[i]) +
                                                    " (or is not finite)");
    }
}

// maxima over every monitor slot since the last read; the slots are cleared
void Engine::read_monitor(double out[2])
{
    check_final_state();
    std::vector<unsigned long long> m(kMonWords);
    HIP_CHECK(hipMemcpyAsync(m.data(), mon, kMonWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemsetAsync(mon, 0, kMonWords * sizeof(unsigned long long), s));
    HIP_CHECK(hipStreamSynchronize(s));
    unsigned long long best[2] = {0, 0};
    for (int h = 0; h < 2; ++h)
        for (int i = 0; i < kMonSlots; ++i) best[h] = std::max(best[h], m[size_t(h) * kMonHalf + size_t(i) * kMonStride]);
    std::memcpy(out, best, sizeof(best));
}

}  // namespace edus

using edus::Engine;
using edus::EngineError;

using edus::guarded;

extern "C" {

const char* edus_last_error(void) { return edus::g_last_error.c_str(); }

edus_ctx* edus_create(const edus_desc* desc)
{
    if (!desc) {
        edus::set_error("null descriptor");
        return nullptr;
    }
    edus_ctx* c = nullptr;
    int rc = guarded([&] { c = new edus_ctx{new Engine(*desc)}; });
    return rc == EDUS_OK ? c : nullptr;
}

edus_ctx* edus_clone(edus_ctx* src)
{
    if (!src || !src->e) {
        edus::set_error("null context");
        return nullptr;
    }
    edus_ctx* c = nullptr;
    const std::vector<const Engine*> one{src->e};
    int rc = guarded([&] { c = new edus_ctx{new Engine(one)}; });
    return rc == EDUS_OK ? c : nullptr;
}

int edus_destroy(edus_ctx* ctx)
{
    if (!ctx) return EDUS_OK;
    int rc = guarded([&] { delete ctx->e; });
    delete ctx;
    return rc;
}

int edus_set_state(edus_ctx* ctx, const double* rho_k)
{
    EDUS_REQUIRE(ctx && rho_k);
    return guarded([&] {
        ctx->e->check_final_state();  // the state being replaced is checked first
        ctx->e->upload(rho_k, ctx->e->y);
    });
}

int edus_get_state(edus_ctx* ctx, double* rho_k)
{
    EDUS_REQUIRE(ctx && rho_k);
    return guarded([&] { ctx->e->download(ctx->e->y, rho_k); });
}

int edus_rotate(edus_ctx* ctx, int to_bloch, const double* in, double* out)
{
    EDUS_REQUIRE(ctx && in && out);
    return guarded([&] {
        Engine& e = *ctx->e;
        if (e.B != 1) throw EngineError(EDUS_ERR_ARG, "edus_rotate applies to single contexts");
        e.upload(in, e.aux);
        e.rotate(e.aux, e.stage, edus::kRotKMajor, to_bloch != 0);
        HIP_CHECK(hipMemcpyAsync(out, e.stage, size_t(e.nb) * e.nb * e.Nk * sizeof(double2), hipMemcpyDeviceToHost,
                                 e.s));
        HIP_CHECK(hipStreamSynchronize(e.s));
    });
}

int edus_get_state_bloch(edus_ctx* ctx, double* rho_bloch)
{
    EDUS_REQUIRE(ctx && rho_bloch);
    return guarded([&] { ctx->e->download_bloch(ctx->e->y, rho_bloch); });
}

int edus_get_state_bloch_device(edus_ctx* ctx, double* rho_bloch_dev)
{
    EDUS_REQUIRE(ctx && rho_bloch_dev);
    return guarded([&] {
        Engine& e = *ctx->e;
        if (e.B != 1) throw EngineError(EDUS_ERR_ARG, "edus_get_state_bloch_device applies to single contexts");
        e.rotate(e.y, reinterpret_cast<double2*>(rho_bloch_dev), edus::kRotKMajor, true);
    });
}

int edus_fourier(edus_ctx* ctx, int sign, const double* in, double* out)
{
    EDUS_REQUIRE(ctx && in && out);
    return guarded([&] {
        Engine& e = *ctx->e;
        if (sign != 1 && sign != -1) throw EngineError(EDUS_ERR_ARG, "edus_fourier: sign must be +1 or -1");
        if (e.layout != EDUS_LAYOUT_FULL || e.B != 1)
            throw EngineError(EDUS_ERR_UNSUPPORTED, "edus_fourier needs a single FULL-layout context");
        if (e.N0 < 2 || e.N1 < 2) throw EngineError(EDUS_ERR_UNSUPPORTED, "edus_fourier needs a 2-D grid");
        e.upload(in, e.aux);  // synchronous: in may equal out
        if (sign < 0)
            e.transform_to_R(e.aux, e.D);
        else
            e.transform_to_k(e.aux, e.D);
        e.download(e.D, out);
    });
}

int edus_ctx_info(edus_ctx* ctx, int* nb, int* N0, int* N1, int* layout)
{
    EDUS_REQUIRE(ctx);
    const Engine& e = *ctx->e;
    if (nb) *nb = e.nb;
    if (N0) *N0 = e.N0;
    if (N1) *N1 = e.N1;
    if (layout) *layout = e.layout;
    return EDUS_OK;
}

int edus_rhs(edus_ctx* ctx, const double E[3], const double* in, double* out)
{
    EDUS_REQUIRE(ctx && E && in && out);
    return guarded([&] {
        Engine& e = *ctx->e;
        e.upload(in, e.aux);
        edus::StageUpdate u{};
        u.out_f = e.f;
        e.evaluate(e.aux, E, u);
        e.download(e.f, out);
    });
}

int edus_rk4_steps(edus_ctx* ctx, double dt, int nsteps, const double* E)
{
    EDUS_REQUIRE(ctx && (nsteps <= 0 || E));
    if (nsteps < 0) { edus::set_error("nsteps < 0"); return EDUS_ERR_ARG; }
    return guarded([&] {
        ctx->e->rk4_steps(dt, nsteps, E);
        HIP_CHECK(hipGetLastError());
        ctx->e->enforce_monitor_limits();
    });
}

int edus_ab_steps(edus_ctx* ctx, double dt, int order, int nsteps, const double* E)
{
    EDUS_REQUIRE(ctx && (nsteps <= 0 || E));
    if (order != 4 && order != 5) {
        edus::set_error("order " + std::to_string(order) + "not implemented for AB");  // DESolver.hpp:120-124
        return EDUS_ERR_UNSUPPORTED;
    }
    if (nsteps < 0) { edus::set_error("nsteps < 0"); return EDUS_ERR_ARG; }
    return guarded([&] {
        for (int s = 0; s < nsteps; ++s) ctx->e->ab_step(dt, order, E + 3 * size_t(s));
        HIP_CHECK(hipGetLastError());
        ctx->e->enforce_monitor_limits();
    });
}

int edus_ab_seed(edus_ctx* ctx, double dt, int order, const double* E)
{
    EDUS_REQUIRE(ctx && E);
    if (order != 4 && order != 5) {
        edus::set_error("order " + std::to_string(order) + "not implemented for AB");
        return EDUS_ERR_UNSUPPORTED;
    }
    return guarded([&] {
        for (int s = 0; s < order - 1; ++s) ctx->e->rk4_step(dt, E + 9 * size_t(s), order);
        HIP_CHECK(hipGetLastError());
    });
}

int edus_rhs_peierls(edus_ctx* ctx, const double E[3], const double Av[3], const double* in, double* out)
{
    EDUS_REQUIRE(ctx && E && Av && in && out);
    return guarded([&] {
        Engine& e = *ctx->e;
        e.upload(in, e.aux);
        edus::StageUpdate u{};
        u.out_f = e.f;
        e.evaluate_peierls(e.aux, E, Av, u);
        e.download(e.f, out);
    });
}

int edus_rk4_steps_peierls(edus_ctx* ctx, double dt, int nsteps, const double* E, const double* Av)
{
    EDUS_REQUIRE(ctx && (nsteps <= 0 || (E && Av)));
    if (nsteps < 0) { edus::set_error("nsteps < 0"); return EDUS_ERR_ARG; }
    return guarded([&] {
        for (int s = 0; s < nsteps; ++s) ctx->e->rk4_step_peierls(dt, E + 9 * size_t(s), Av + 9 * size_t(s));
        HIP_CHECK(hipGetLastError());
        ctx->e->enforce_monitor_limits();
    });
}

int edus_ab_steps_peierls(edus_ctx* ctx, double dt, int order, int nsteps, const double* E, const double* Av)
{
    EDUS_REQUIRE(ctx && (nsteps <= 0 || (E && Av)));
    if (order != 4 && order != 5) {
        edus::set_error("order " + std::to_string(order) + "not implemented for AB");
        return EDUS_ERR_UNSUPPORTED;
    }
    if (nsteps < 0) { edus::set_error("nsteps < 0"); return EDUS_ERR_ARG; }
    return guarded([&] {
        for (int s = 0; s < nsteps; ++s) ctx->e->ab_step_peierls(dt, order, E + 3 * size_t(s), Av + 3 * size_t(s));
        HIP_CHECK(hipGetLastError());
        ctx->e->enforce_monitor_limits();
    });
}

int edus_observables_peierls(edus_ctx* ctx, const double Av[3], double* pop_bloch, double* pop_wannier,
                             double* vel)
{
    EDUS_REQUIRE(ctx && Av);
    return guarded([&] {
        Engine& e = *ctx->e;
        if (!e.peierls) throw EngineError(EDUS_ERR_STATE, "context was not created for the Peierls gauge");
        e.peierls_sandwich(e.y, nullptr, nullptr, Av, -1, e.G);  // G is free between steps
        e.observables(pop_bloch, pop_wannier, vel, 0, e.G);
    });
}

int edus_ab_reset(edus_ctx* ctx)
{
    EDUS_REQUIRE(ctx);
    return guarded([&] {
        Engine& e = *ctx->e;
        for (int i = 0; i < 5; ++i) {
            if (e.hist[i]) HIP_CHECK(hipMemsetAsync(e.hist[i], 0, e.plane_bytes(), e.s));
            e.abidx[i] = i;
        }
        HIP_CHECK(hipStreamSynchronize(e.s));
    });
}

int edus_observables(edus_ctx* ctx, double* pop_bloch, double* pop_wannier, double* vel)
{
    EDUS_REQUIRE(ctx);
    return guarded([&] { ctx->e->observables(pop_bloch, pop_wannier, vel); });
}

int edus_check(edus_ctx* ctx, double* trace_dev, double* herm_dev)
{
    EDUS_REQUIRE(ctx);
    return guarded([&] {
        double v[2];
        ctx->e->read_monitor(v);
        if (trace_dev) *trace_dev = v[0];
        if (herm_dev) *herm_dev = v[1];
    });
}

int edus_monitor_limits(edus_ctx* ctx, double trace_limit, double herm_limit)
{
    EDUS_REQUIRE(ctx);
    if (!(trace_limit >= 0.0) || !(herm_limit >= 0.0)) {  // also refuses NaN
        edus::set_error("monitor limits must be >= 0 (0: not checked)");
        return EDUS_ERR_ARG;
    }
    ctx->e->mon_limit[0] = trace_limit;
    ctx->e->mon_limit[1] = herm_limit;
    return EDUS_OK;
}

int edus_graph(edus_ctx* ctx, int steps_per_graph)
{
    EDUS_REQUIRE(ctx);
    if (steps_per_graph < 0) {
        edus::set_error("steps_per_graph < 0");
        return EDUS_ERR_ARG;
    }
    if (ctx->e->B != 1) {
        edus::set_error("graph replay applies to single contexts");
        return EDUS_ERR_ARG;
    }
    ctx->e->graph_steps = steps_per_graph;
    return EDUS_OK;
}

int edus_synchronize(edus_ctx* ctx)
{
    EDUS_REQUIRE(ctx);
    return guarded([&] { HIP_CHECK(hipStreamSynchronize(ctx->e->s)); });
}

int edus_get_tuning(edus_ctx* ctx, edus_tuning* out)
{
    EDUS_REQUIRE(ctx && out);
    const Engine& e = *ctx->e;
    for (int a = 0; a < 2; ++a) {
        out->fft_family[a] = e.fft_choice[a].two_step;
        out->fft_cfg[a] = e.fft_choice[a].cfg;
        out->fft_lines[a] = e.fft_choice[a].L;
    }
    out->comm_kind = e.comm_kind;
    out->comm_tile = e.comm_tile;
    out->comm_tile_ab = e.comm_tile_ab;
    out->grad_fused = e.grad_fused ? 1 : 0;
    out->tri_fields = e.tri ? 1 : 0;
    out->hf_fused = e.hf_fuse ? 1 : 0;
    return EDUS_OK;
}

int edus_set_tuning(edus_ctx* ctx, const edus_tuning* in)
{
    EDUS_REQUIRE(ctx && in);
    return guarded([&] {
        Engine& e = *ctx->e;
        e.hp_valid = false;  // the phased H is formed again with the new kernels
        edus::FFTChoice next[2];
        for (int a = 0; a < 2; ++a) {
            const int n = a == 0 ? e.N0 : e.N1;
            const int maxl = edus::two_step_max_lines(n, in->fft_cfg[a]);
            edus::FFTChoice c;
            c.two_step = in->fft_family[a];
            c.cfg = in->fft_cfg[a];
            c.L = in->fft_lines[a];
            if (c.two_step != (maxl > 0 ? 1 : 0)) throw EngineError(EDUS_ERR_ARG, "fft_family is fixed by the axis length");
            int lmax;
            if (c.two_step) {
                if (c.cfg < 0 || c.cfg > 3) throw EngineError(EDUS_ERR_ARG, "two-step variant must be 0..3");
                lmax = maxl;
            } else {
                if (c.cfg < 0 || c.cfg >= edus::kNumFFTConfigs || edus::fft_lines(c.cfg, n) * n > edus::kFFTConfigs[c.cfg].values)
                    throw EngineError(EDUS_ERR_ARG, "Stockham configuration does not fit this axis");
                lmax = edus::fft_lines(c.cfg, n);
            }
            if (a == 0) lmax = std::min(lmax, e.N1);
            if (c.L < 1 || c.L > lmax) throw EngineError(EDUS_ERR_ARG, "fft_lines out of range for this axis");
            next[a] = c;
        }
        if (e.prec == EDUS_PREC_MIXED ? in->comm_kind != edus::kCommMfma32
                                      : (in->comm_kind != edus::kCommScalar && in->comm_kind != edus::kCommMfma64))
            throw EngineError(EDUS_ERR_ARG, "comm_kind does not match the context precision");
        for (int t : {in->comm_tile, in->comm_tile_ab})
            if (t != 0 && (t < 1 || t > 2 || in->comm_kind == edus::kCommScalar))
                throw EngineError(EDUS_ERR_ARG, "comm_tile and comm_tile_ab must be 0, or 1 or 2 for the MFMA engines");
        if (in->grad_fused != 0 && (in->grad_fused != 1 || !edus::grad_fused_available(e.N0, e.N1) || e.kgrad))
            throw EngineError(EDUS_ERR_ARG, "grad_fused must be 0, or 1 on the grids with a one-pass gradient kernel");
        if (in->tri_fields != 0 && (in->tri_fields != 1 || !e.tri_ok))
            throw EngineError(EDUS_ERR_ARG, "tri_fields must be 0, or 1 where H0 and r are Hermitian (FULL layout, "
                                            "no exchange, no Peierls gauge)");
        if (in->hf_fused != 0 && in->hf_fused != 1) throw EngineError(EDUS_ERR_ARG, "hf_fused must be 0 or 1");
        e.fft_choice[0] = next[0];
        e.fft_choice[1] = next[1];
        e.comm_kind = in->comm_kind;
        e.comm_tile = in->comm_tile;
        e.comm_tile_ab = in->comm_tile_ab;
        e.grad_fused = in->grad_fused != 0;
        e.tri = in->tri_fields != 0;
        e.hf_fuse = in->hf_fused != 0;
        if (e.prec != EDUS_PREC_MIXED) {
            e.comm_kind64 = in->comm_kind;
            e.comm_tile64 = in->comm_tile;
        }
        e.prime_launch_attributes();
        // a captured graph holds the old kernels: capture again on next use
        if (e.self_batch) e.self_batch->invalidate();
    });
}

int edus_timing_enable(edus_ctx* ctx, int on)
{
    EDUS_REQUIRE(ctx);
    return guarded([&] {
        Engine& e = *ctx->e;
        e.collect_timing();
        e.tsum.clear();
        e.timing = on != 0;
    });
}

int edus_timing_get(edus_ctx* ctx, const char* cls, double* total_ms, long* count)
{
    EDUS_REQUIRE(ctx && cls);
    return guarded([&] {
        Engine& e = *ctx->e;
        e.collect_timing();
        auto it = e.tsum.find(cls);
        if (total_ms) *total_ms = it == e.tsum.end() ? 0.0 : it->second.first;
        if (count) *count = it == e.tsum.end() ? 0 : it->second.second;
    });
}

}  // extern "C"
