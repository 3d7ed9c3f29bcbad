// The C++ seam classes of include/edus_operator.hpp (Operator<T>,
// FourierTransform) on the reference's hBN ci-test set-up, after 30 RK4 steps:
//   - go_to_R against a direct DFT on the host (every R, every k);
//   - go_to_k after go_to_R returns the k-space block;
//   - go_to_bloch equals edus_get_state_bloch bit for bit, go_to_wannier
//     returns the block;
//   - FourierTransform alone: a k-independent operator M goes to M at R = 0
//     and zero elsewhere;
//   - a copied Operator transforms its own blocks, not the original's.
// Usage: seams_hbn <tb seedname>.  Prints one line per check; exit status 1
// on the first failed check.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "edus_gpu.h"
#include "edus_operator.hpp"
#include "hbn_params.hpp"

namespace {

using cplx = std::complex<double>;
using edus::seam::Operator;

void expect(bool ok, const char* what, double value, double bar)
{
    std::printf("seams: %-52s %.3e (bar %.1e)\n", what, value, bar);
    if (!ok) {
        std::fprintf(stderr, "seams: FAILED %s\n", what);
        std::exit(1);
    }
}

double max_abs_diff(const std::vector<cplx>& a, const std::vector<cplx>& b)
{
    double m = 0.0;
    for (size_t i = 0; i < a.size(); ++i) m = std::fmax(m, std::abs(a[i] - b[i]));
    return m;
}

double max_abs(const std::vector<cplx>& a)
{
    double m = 0.0;
    for (const cplx& x : a) m = std::fmax(m, std::abs(x));
    return m;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <tb seedname>\n", argv[0]);
        return 2;
    }
    const edus_sim_params p = hbn_params(argv[1]);
    edus_sim* sim = edus_sim_create(&p);
    if (!sim) {
        std::fprintf(stderr, "edus_sim_create failed: %s\n", edus_last_error());
        return 1;
    }
    edus::seam::check(edus_sim_steps(sim, 30));
    edus_ctx* ctx = edus_sim_engine(sim);
    int nb = 0, N0 = 0, N1 = 0;
    edus::seam::check(edus_ctx_info(ctx, &nb, &N0, &N1, nullptr));
    const int Nk = N0 * N1, M = nb * nb;
    const double kPi = std::acos(-1.0);

    Operator<cplx> rho;
    rho.initialize_fft(ctx);
    rho.lock_space(edus::seam::k);
    rho.lock_gauge(edus::seam::wannier);
    edus::seam::check(edus_get_state(ctx, reinterpret_cast<double*>(rho.get_Operator(edus::seam::k).data())));
    const std::vector<cplx> orig = rho.get_Operator(edus::seam::k);
    const double scale = max_abs(orig);

    // Bloch gauge first, on the untouched block: the same rotation as DMk_bloch
    std::vector<cplx> bl(size_t(Nk) * M);
    edus::seam::check(edus_get_state_bloch(ctx, reinterpret_cast<double*>(bl.data())));
    Operator<cplx> rb = rho;
    rb.go_to_bloch();
    const double db = max_abs_diff(rb.get_Operator(edus::seam::k), bl);
    expect(db == 0.0, "go_to_bloch vs edus_get_state_bloch, max|d|", db, 0.0);
    rb.go_to_wannier();
    const double dw = max_abs_diff(rb.get_Operator(edus::seam::k), orig) / scale;
    expect(dw <= 1e-13, "go_to_bloch then go_to_wannier, max|d| / max|rho|", dw, 1e-13);

    // k -> R against the DFT written out: 1/Nk sum_k rho(k) e^{-2 pi i k.R / N}
    rho.go_to_R();
    const std::vector<cplx>& got = rho.get_Operator(edus::seam::R);
    std::vector<cplx> ref(size_t(Nk) * M, cplx(0.0));
    for (int r0 = 0; r0 < N0; ++r0)
        for (int r1 = 0; r1 < N1; ++r1) {
            cplx* out = &ref[(size_t(r0) * N1 + r1) * M];
            for (int k0 = 0; k0 < N0; ++k0)
                for (int k1 = 0; k1 < N1; ++k1) {
                    const double ph = -2.0 * kPi * (double(k0 * r0 % N0) / N0 + double(k1 * r1 % N1) / N1);
                    const cplx e(std::cos(ph), std::sin(ph));
                    const cplx* in = &orig[(size_t(k0) * N1 + k1) * M];
                    for (int h = 0; h < M; ++h) out[h] += e * in[h];
                }
            for (int h = 0; h < M; ++h) out[h] /= double(Nk);
        }
    const double dR = max_abs_diff(got, ref) / max_abs(ref);
    expect(dR <= 1e-12, "go_to_R vs direct DFT, max|d| / max|rho(R)|", dR, 1e-12);
    expect(rho.get_space() == edus::seam::R, "space after go_to_R is R", 0.0, 0.0);

    Operator<cplx> copy = rho;  // R space; transforms its own blocks
    copy.go_to_k();
    const double dcopy = max_abs_diff(rho.get_Operator(edus::seam::R), got);
    expect(dcopy == 0.0 && rho.get_space() == edus::seam::R, "a copy's go_to_k leaves the original", dcopy, 0.0);

    rho.go_to_k();
    const double dk = max_abs_diff(rho.get_Operator(edus::seam::k), orig) / scale;
    expect(dk <= 1e-14, "go_to_R then go_to_k, max|d| / max|rho|", dk, 1e-14);
    const double dck = max_abs_diff(copy.get_Operator(edus::seam::k), rho.get_Operator(edus::seam::k));
    expect(dck == 0.0, "the copy's go_to_k equals the original's", dck, 0.0);

    // FourierTransform alone: a k-independent operator is M delta_{R,0}
    std::vector<cplx> xk(size_t(Nk) * M), xR(size_t(Nk) * M);
    for (int ik = 0; ik < Nk; ++ik)
        for (int h = 0; h < M; ++h) xk[size_t(ik) * M + h] = cplx(0.5 + h, -0.25 * h);
    edus::seam::FourierTransform ft;
    ft.initialize(ctx, xk.data(), xR.data());
    ft.fft(-1);
    double dd = 0.0;
    for (int iR = 0; iR < Nk; ++iR)
        for (int h = 0; h < M; ++h) {
            const cplx want = iR == 0 ? cplx(0.5 + h, -0.25 * h) : cplx(0.0);
            dd = std::fmax(dd, std::abs(xR[size_t(iR) * M + h] - want));
        }
    expect(dd <= 1e-14, "FourierTransform fft(-1) of a constant: M delta_R0", dd, 1e-14);
    bool threw = false;
    try {
        ft.fft(2);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    expect(threw, "fft(2) throws runtime_error", 0.0, 0.0);

    edus::seam::check(edus_sim_destroy(sim));
    std::printf("seams_hbn: nb %d, %d x %d k: all checks passed\n", nb, N0, N1);
    return 0;
}
