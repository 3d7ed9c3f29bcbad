// EDUS FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
// Each block cites the reference (gcistaro/EDUS_2.0) source it restates.
#include "oracle.hpp"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <functional>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace oracle {

// ------------------------------------------------------------------ units
// ConvertUnits.cpp:7-37
const Unit Angstrom{LENGTH, 1.e-10};
static const Unit NanoMeters{LENGTH, 1.e-09};
const Unit ElectronVolt{ENERGY, ElectronVolt_value};
static const Unit Joule{ENERGY, 1.};
static const Unit Wcm2{INTENSITY, 1.e+04};
static const Unit FemtoSeconds{TIME, 1.e-15};
const Unit AuIntensity{INTENSITY, AtomicUnitOfIntensity};
const Unit AuTime{TIME, AtomicUnitOfTime};
const Unit AuLength{LENGTH, Bohr_value};
const Unit AuEnergy{ENERGY, 2. * Rydberg_value};
static const Unit NullUnit{NULLTYPE, 0.};

Unit unit(const std::string& s)
{
    if (s == "angstrom") return Angstrom;
    if (s == "nanometers" || s == "nm") return NanoMeters;
    if (s == "electronvolt") return ElectronVolt;
    if (s == "joule") return Joule;
    if (s == "wcm2") return Wcm2;
    if (s == "femtoseconds" || s == "fs") return FemtoSeconds;
    if (s == "auintensity") return AuIntensity;
    if (s == "autime") return AuTime;
    if (s == "aulength") return AuLength;
    if (s == "auenergy") return AuEnergy;
    return NullUnit;  // the schema defaults "eV", "Wcm2", "au" land here
}

// ConvertUnits.hpp:26-31: value / out.value * in.value
double convert(double v, const Unit& in, const Unit& out) { return v / out.value * in.value; }

// ------------------------------------------------------------------ tb reader
// ReadWannier.cpp:5-153 (ndegen parsed, never applied).
static std::vector<std::vector<std::string>> read_words(const std::string& path)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<std::vector<std::string>> out;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::vector<std::string> w;
        std::string tok;
        while (ss >> tok) w.push_back(tok);
        out.push_back(w);
    }
    out.push_back({});  // sentinel empty line
    out.push_back({});
    return out;
}

TBModel read_tb(const std::string& path)
{
    auto L = read_words(path);
    TBModel m;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m.unitcell[i][j] = std::atof(L[1 + i].at(j).c_str());
    m.nb = std::atoi(L[4].at(0).c_str());
    m.nR = std::atoi(L[5].at(0).c_str());
    size_t li = 6;
    while (li < L.size() && !L[li].empty()) {
        for (auto& w : L[li]) m.degeneracy.push_back(int(std::atof(w.c_str())));
        ++li;
    }
    while (li < L.size() && L[li].empty()) ++li;
    const int nb = m.nb, nb2 = nb * nb;
    m.R.assign(m.nR, {0, 0, 0});
    m.H.assign(size_t(m.nR) * nb2, 0.);
    for (int d = 0; d < 3; ++d) m.r[d].assign(size_t(m.nR) * nb2, 0.);
    // Hamiltonian blocks: R line followed by "m n re im" lines, blank line after.
    int idx = 0;
    while (li + 1 < L.size() && L[li + 1].size() == 4 && idx < m.nR) {
        for (int d = 0; d < 3; ++d) m.R[idx][d] = std::atof(L[li][d].c_str());
        size_t j = li + 1;
        while (!L[j].empty()) {
            int a = std::atoi(L[j][0].c_str()) - 1, b = std::atoi(L[j][1].c_str()) - 1;
            m.H[size_t(idx) * nb2 + a * nb + b] = cplx(std::atof(L[j][2].c_str()), std::atof(L[j][3].c_str()));
            ++j;
        }
        li = j + 1;
        ++idx;
    }
    while (li < L.size() && L[li].empty()) ++li;
    // Position blocks: R line followed by "m n rex imx rey imy rez imz".
    idx = 0;
    while (li + 1 < L.size() && L[li + 1].size() == 8 && idx < m.nR) {
        for (int d = 0; d < 3; ++d) m.R[idx][d] = std::atof(L[li][d].c_str());
        size_t j = li + 1;
        while (L[j].size() == 8) {
            int a = std::atoi(L[j][0].c_str()) - 1, b = std::atoi(L[j][1].c_str()) - 1;
            for (int d = 0; d < 3; ++d)
                m.r[d][size_t(idx) * nb2 + a * nb + b] =
                    cplx(std::atof(L[j][2 + 2 * d].c_str()), std::atof(L[j][3 + 2 * d].c_str()));
            ++j;
        }
        li = j + 1;
        ++idx;
    }
    return m;
}

// ------------------------------------------------------------------ FFT
FFTPlan1D make_plan(int n)
{
    FFTPlan1D p;
    p.n = n;
    int x = n;
    for (int r : {5, 3, 2}) {
        while (x % r == 0) { p.factors.push_back(r); x /= r; }
    }
    for (int r = 7; x > 1; r += 2) {
        while (x % r == 0) { p.factors.push_back(r); x /= r; }
    }
    p.tw_fwd.resize(n);
    p.tw_bwd.resize(n);
    for (int j = 0; j < n; ++j) {
        // angle reduced to [-pi, pi] for accuracy
        int jj = (2 * j > n) ? j - n : j;
        double a = 2. * pi * double(jj) / double(n);
        p.tw_fwd[j] = cplx(std::cos(a), -std::sin(a));
        p.tw_bwd[j] = cplx(std::cos(a), std::sin(a));
    }
    return p;
}

// Decimation-in-time mixed-radix recursion: X[k] = sum_j x[j*stride] w^{jk}.
static void fft_rec(const cplx* x, int stride, cplx* out, int N, int fi, const FFTPlan1D& p,
                    const std::vector<cplx>& tw, int twstep)
{
    if (N == 1) { out[0] = x[0]; return; }
    const int r = p.factors[fi];
    const int m = N / r;
    for (int q = 0; q < r; ++q) fft_rec(x + q * stride, stride * r, out + q * m, m, fi + 1, p, tw, twstep * r);
    const int n = p.n;
    cplx y[64];
    std::vector<cplx> ybig;
    cplx* yy = y;
    if (r > 64) { ybig.resize(r); yy = ybig.data(); }
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < r; ++q) {
            long e = (long(q) * k * twstep) % n;
            yy[q] = out[q * m + k] * tw[e];
        }
        for (int s = 0; s < r; ++s) {
            cplx acc = 0.;
            for (int q = 0; q < r; ++q) {
                long e = (long(m) * q * s * twstep) % n;
                acc += yy[q] * tw[e];
            }
            out[k + s * m] = acc;
        }
    }
}

void fft1d(const FFTPlan1D& p, cplx* x, int stride, int sign, cplx* scratch)
{
    const int n = p.n;
    if (n == 1) return;
    cplx* tin = scratch;
    cplx* tout = scratch + n;
    for (int j = 0; j < n; ++j) tin[j] = x[j * stride];
    fft_rec(tin, 1, tout, n, 0, p, sign < 0 ? p.tw_fwd : p.tw_bwd, 1);
    for (int j = 0; j < n; ++j) x[j * stride] = tout[j];
}

void fft2d(const FFTPlan1D& p0, const FFTPlan1D& p1, cplx* plane, int sign, cplx* scratch)
{
    const int n0 = p0.n, n1 = p1.n;
    for (int i = 0; i < n0; ++i) fft1d(p1, plane + i * n1, 1, sign, scratch);
    for (int j = 0; j < n1; ++j) fft1d(p0, plane + j, n1, sign, scratch);
}

// ------------------------------------------------------------------ eigen
void eigh(int n, const cplx* Ain, double* w, cplx* V)
{
    std::vector<cplx> A(Ain, Ain + n * n);
    for (int i = 0; i < n * n; ++i) V[i] = 0.;
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.;
    for (int i = 0; i < n; ++i) A[i * n + i] = A[i * n + i].real();
    for (int sweep = 0; sweep < 100; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < n; ++p) {
            for (int q = p + 1; q < n; ++q) {
                cplx b = A[p * n + q];
                double ab = std::abs(b);
                double app = A[p * n + p].real(), aqq = A[q * n + q].real();
                if (ab <= 1e-18 * (std::abs(app) + std::abs(aqq)) || ab < 1e-300) {
                    A[p * n + q] = A[q * n + p] = 0.;
                    continue;
                }
                rotated = true;
                cplx e = b / ab;
                double tau = (aqq - app) / (2. * ab);
                double t = (tau >= 0 ? 1. : -1.) / (std::abs(tau) + std::sqrt(1. + tau * tau));
                double c = 1. / std::sqrt(1. + t * t), s = t * c;
                // G = [[c, s], [-s conj(e), c conj(e)]] on (p,q); A <- G^H A G, V <- V G
                cplx gqp = -s * std::conj(e), gqq = c * std::conj(e);
                for (int i = 0; i < n; ++i) {
                    cplx aip = A[i * n + p], aiq = A[i * n + q];
                    A[i * n + p] = aip * c + aiq * gqp;
                    A[i * n + q] = aip * s + aiq * gqq;
                    cplx vip = V[i * n + p], viq = V[i * n + q];
                    V[i * n + p] = vip * c + viq * gqp;
                    V[i * n + q] = vip * s + viq * gqq;
                }
                for (int j = 0; j < n; ++j) {
                    cplx apj = A[p * n + j], aqj = A[q * n + j];
                    A[p * n + j] = c * apj + std::conj(gqp) * aqj;
                    A[q * n + j] = s * apj + std::conj(gqq) * aqj;
                }
                A[p * n + q] = A[q * n + p] = 0.;
                A[p * n + p] = app - t * ab;
                A[q * n + q] = aqq + t * ab;
            }
        }
        if (!rotated) break;
    }
    for (int i = 0; i < n; ++i) w[i] = A[i * n + i].real();
    for (int i = 0; i < n; ++i) {  // ascending order
        int kmin = i;
        for (int k = i + 1; k < n; ++k)
            if (w[k] < w[kmin]) kmin = k;
        if (kmin != i) {
            std::swap(w[i], w[kmin]);
            for (int r = 0; r < n; ++r) std::swap(V[r * n + i], V[r * n + kmin]);
        }
    }
}

// ------------------------------------------------------------------ laser
// Laser.cpp:35-39 (wave), 96-104 (sin^2 envelope), 131-134 (field)
void Laser::E(double t, double out[3]) const
{
    double tt = t - t0;
    double env = (tt <= 1.e-07 || tt > duration) ? 0. : std::pow(std::sin(pi * tt / duration), 2.);
    double wave = std::sin(omega * (t - t0) + phase);
    double a = amplitude * env * wave;
    for (int d = 0; d < 3; ++d) out[d] = a * pol[d];
}

// ------------------------------------------------------------------ struve
// ModelCoulomb.cpp:7-44 (series / asymptotic form, Zhang & Jin, "Computation
// of Special Functions" (1996), routine STVH0).
double struve0(double x)
{
    double s = 1.0, r = 1.0, sh0;
    if (x <= 20.0) {
        double a0 = 2.0 * x / pi;
        for (int k = 1; k < 61; ++k) {
            r = -r * x / (2.0 * k + 1.0) * x / (2.0 * k + 1.0);
            s = s + r;
            if (std::fabs(r) < std::fabs(s) * 1.0e-12) break;
        }
        sh0 = a0 * s;
    } else {
        int km = int(0.5 * (x + 1.0));
        if (x >= 50.0) km = 25;
        for (int k = 1; k <= km; ++k) {
            r = -r * std::pow((2.0 * k - 1.0) / x, 2);
            s = s + r;
            if (std::fabs(r) < std::fabs(s) * 1.0e-12) break;
        }
        double t = 4.0 / x, t2 = t * t;
        double p0 = ((((-.37043e-5 * t2 + .173565e-4) * t2 - .487613e-4) * t2 + .17343e-3) * t2 - 0.1753062e-2) * t2 +
                    .3989422793;
        double q0 = t * (((((.32312e-5 * t2 - 0.142078e-4) * t2 + 0.342468e-4) * t2 - 0.869791e-4) * t2 + 0.4564324e-3) *
                             t2 -
                         0.0124669441);
        double ta0 = x - 0.25 * pi;
        double by0 = 2.0 / std::sqrt(x) * (p0 * std::sin(ta0) + q0 * std::cos(ta0));
        sh0 = 2.0 / (pi * x) * s + by0;
    }
    return sh0;
}

// ------------------------------------------------------------------ helpers
static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// MeshGrid.cpp:345-371 reduce(): fold x into [lo, up) on a grid of period n
static double reduce_coord(double x, double lo, double up, double period)
{
    while (x < lo && std::abs(x - lo) > threshold) x += period;
    while (x > up || std::abs(x - up) < threshold) x -= period;
    return x;
}

// out = alpha * A*B + beta * out  (row-major nb x nb; cblas_zgemm NoTrans)
static void zgemm(int n, cplx alpha, const cplx* A, const cplx* B, cplx beta, cplx* C)
{
    cplx tmp[32 * 32];
    std::vector<cplx> big;
    cplx* T = tmp;
    if (n > 32) { big.resize(n * n); T = big.data(); }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            cplx acc = 0.;
            for (int l = 0; l < n; ++l) acc += A[i * n + l] * B[l * n + j];
            T[i * n + j] = acc;
        }
    for (int i = 0; i < n * n; ++i) C[i] = (beta == cplx(0.) ? cplx(0.) : beta * C[i]) + alpha * T[i];
}

// ------------------------------------------------------------------ Simulation
// Setup order follows Simulation.cpp:10-233.
Simulation::Simulation(const SimParams& p) : par(p)
{
    // times in a.u. (Simulation.cpp:18-26)
    double ti = convert(p.initialtime, unit(p.initialtime_units), AuTime);
    double tf = convert(p.finaltime, unit(p.finaltime_units), AuTime);
    dt = convert(p.dt, unit(p.dt_units), AuTime);
    if (par.printresolution_pulse == 0) par.printresolution_pulse = par.printresolution;
    nsteps = int((tf - ti) / dt) + 2;  // Simulation.cpp:355

    // material (Model.hpp:34-80)
    TBModel tb = read_tb(p.tb_file + "_tb.dat");
    nb = tb.nb;
    M = nb * nb;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = convert(tb.unitcell[j][i], Angstrom, AuLength);
    for (auto& h : tb.H) h = convert(h.real(), ElectronVolt, AuEnergy) + cplx(0, 1) * convert(h.imag(), ElectronVolt, AuEnergy);
    for (int d = 0; d < 3; ++d)
        for (auto& v : tb.r[d]) v = convert(v.real(), Angstrom, AuLength) + cplx(0, 1) * convert(v.imag(), Angstrom, AuLength);

    for (int d = 0; d < 3; ++d) N[d] = p.grid[d];
    Nk = N[0] * N[1] * N[2];
    if (N[2] != 1) throw std::runtime_error("oracle: only 2-D grids (grid[2]==1) are restated");
    plan[0] = make_plan(N[0]);
    plan[1] = make_plan(N[1]);

    // Gamma-centred R grid (MeshGrid.cpp:458-486)
    for (int d = 0; d < 3; ++d) Rgc[d].resize(Nk);
    for (int i0 = 0; i0 < N[0]; ++i0)
        for (int i1 = 0; i1 < N[1]; ++i1) {
            int ik = i1 + N[1] * i0;
            double c[3] = {double(i0), double(i1), 0.};
            for (int d = 0; d < 3; ++d) Rgc[d][ik] = reduce_coord(c[d], -N[d] / 2., N[d] / 2., N[d]);
        }

    // Wannier centres (Model.hpp:72-78); tb.r is in bohr here
    rwann.assign(nb, {0., 0., 0.});
    for (int iR = 0; iR < tb.nR; ++iR)
        if (std::abs(tb.R[iR][0]) + std::abs(tb.R[iR][1]) + std::abs(tb.R[iR][2]) < 1e-6)
            for (int n = 0; n < nb; ++n)
                for (int d = 0; d < 3; ++d) rwann[n][d] = tb.r[d][size_t(iR) * M + n * nb + n].real();

    if (p.gradient_space == "k") setup_kgradient();
    else if (p.gradient_space != "R") throw std::runtime_error("gradient_space must be R or k");

    // explicit DFT H(R)->H(k), sign +1, over the Gamma-centred tb R list (fftPair.cpp:110-143)
    std::vector<std::array<double, 3>> Rtb(tb.nR);
    for (int iR = 0; iR < tb.nR; ++iR)
        for (int d = 0; d < 3; ++d)
            Rtb[iR][d] = reduce_coord(tb.R[iR][d], -N[d] / 2., (N[d] / 2. == 0 ? 1 : N[d] / 2.), N[d]);
    H0k.assign(size_t(Nk) * M, 0.);
    for (int d = 0; d < 3; ++d) rk[d].assign(size_t(Nk) * M, 0.);
#pragma omp parallel for schedule(static)
    for (int ik = 0; ik < Nk; ++ik) {
        int i0 = ik / N[1], i1 = ik % N[1];
        double kc[3] = {double(i0) / N[0], double(i1) / N[1], 0.};
        for (int h = 0; h < M; ++h) {
            cplx aH = 0., ar[3] = {0., 0., 0.};
            for (int iR = 0; iR < tb.nR; ++iR) {
                double dp = kc[0] * Rtb[iR][0] + kc[1] * Rtb[iR][1] + kc[2] * Rtb[iR][2];
                cplx ph = std::exp(cplx(0, 1) * 2. * pi * dp);
                aH += ph * tb.H[size_t(iR) * M + h];
                for (int d = 0; d < 3; ++d) ar[d] += ph * tb.r[d][size_t(iR) * M + h];
            }
            H0k[size_t(ik) * M + h] = aH;
            for (int d = 0; d < 3; ++d) rk[d][size_t(ik) * M + h] = ar[d];
        }
        // make_hermitian on r(k) (BlockMatrix_definitions.hpp:258-284)
        for (int d = 0; d < 3; ++d) {
            cplx* blk = &rk[d][size_t(ik) * M];
            for (int a = 0; a < nb; ++a)
                for (int b = a; b < nb; ++b) {
                    cplx term = (blk[a * nb + b] + std::conj(blk[b * nb + a])) / 2.;
                    blk[a * nb + b] = term;
                    blk[b * nb + a] = std::conj(term);
                }
        }
    }
    // H0_, r_ on the FFT grid in R (Simulation.cpp:99-118)
    H0R = H0k;
    fft_planes(H0R, -1);
    for (int d = 0; d < 3; ++d) { rR[d] = rk[d]; fft_planes(rR[d], -1); }

    // lasers (Simulation.cpp:121-144)
    for (auto& lp : p.lasers) {
        Laser L;
        L.t0 = convert(lp.t0, unit(lp.t0_units), AuTime);
        L.amplitude = std::sqrt(convert(lp.intensity, unit(lp.intensity_units), AuIntensity));
        if (std::abs(lp.frequency) > 1.e-07) {
            L.omega = convert(lp.frequency, unit(lp.frequency_units), AuEnergy);
        } else {
            if (!lp.wavelength) throw std::runtime_error("You must specify (nonzero) frequency *xor* wavelength!");
            double lam = convert(lp.wavelength, unit(lp.wavelength_units), AuLength);
            L.omega = 2 * pi * SpeedOfLight / lam;
        }
        double period = 2. * pi / L.omega;
        L.duration = lp.cycles * period;
        L.phase = lp.phase;
        // normalised twice, as Simulation.cpp:139-142 then Laser.cpp:197-200
        double p1[3];
        double nrm = std::sqrt(dot3(lp.polarization, lp.polarization));
        for (int d = 0; d < 3; ++d) p1[d] = lp.polarization[d] / nrm;
        double nrm2 = std::sqrt(dot3(p1, p1));
        for (int d = 0; d < 3; ++d) L.pol[d] = p1[d] / nrm2;
        lasers.push_back(L);
    }

    // eigensystem (Simulation.cpp:263-286)
    U.assign(size_t(Nk) * M, 0.);
    Udag.assign(size_t(Nk) * M, 0.);
    eps.assign(size_t(Nk) * nb, 0.);
#pragma omp parallel for schedule(static)
    for (int ik = 0; ik < Nk; ++ik) {
        eigh(nb, &H0k[size_t(ik) * M], &eps[size_t(ik) * nb], &U[size_t(ik) * M]);
        for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; ++b) Udag[size_t(ik) * M + a * nb + b] = std::conj(U[size_t(ik) * M + b * nb + a]);
    }

    // initial condition: rho_bloch = diag(filled) -> wannier U rho U^dagger (Functional_InitialCondition.hpp:12-32)
    rho.assign(size_t(Nk) * M, 0.);
#pragma omp parallel for schedule(static)
    for (int ik = 0; ik < Nk; ++ik) {
        std::vector<cplx> rb(M, 0.), tmp(M, 0.);
        for (int n = 0; n < par.filledbands; ++n) rb[n * nb + n] = 1.;
        zgemm(nb, 1., &U[size_t(ik) * M], rb.data(), 0., tmp.data());
        zgemm(nb, 1., tmp.data(), &Udag[size_t(ik) * M], 0., &rho[size_t(ik) * M]);
    }
    for (auto& a : aux) a.assign(size_t(Nk) * M, 0.);
    if (par.solver == "AB") {  // DESolver.hpp:110-126
        if (par.order == 4) { double b[5] = {55. / 24., -59. / 24., 37. / 24., -3. / 8., 0.}; std::copy(b, b + 5, beta); }
        else if (par.order == 5) {
            double b[5] = {1901. / 720., -2774. / 720., 2616. / 720., -1274. / 720., 251. / 720.};
            std::copy(b, b + 5, beta);
        } else throw std::runtime_error("order not implemented for AB");
    }

    // coulomb (Simulation.cpp:177-181)
    std::vector<cplx> rhoR = rho;
    fft_planes(rhoR, -1);
    rho0R = rhoR;
    if (par.coulomb) setup_coulomb(tb);
    // DensityMatrix_.go_to_k() after Calculate_Velocity (Simulation.cpp:193): round trip
    fft_planes(rhoR, +1);
    rho = rhoR;

    // velocity operator (Simulation.cpp:523-576)
    double det;
    {
        int dim = 0;
        for (int d = 0; d < 3; ++d) dim += (par.grid[d] > 1);
        if (dim == 3) throw std::runtime_error("oracle: 3-D jacobian not restated");
        else if (dim == 2) {
            if (par.grid[2] > 1) throw std::runtime_error("perpendicular != 2 -> implement me!");
            det = std::abs(A[0][0] * A[1][1] - A[1][0] * A[0][1]);
        } else throw std::runtime_error("Implement me!");
    }
    for (int d = 0; d < 3; ++d) {
        Vel[d].assign(size_t(Nk) * M, 0.);
#pragma omp parallel for schedule(static)
        for (int ik = 0; ik < Nk; ++ik) {
            const cplx* r = &rk[d][size_t(ik) * M];
            const cplx* h = &H0k[size_t(ik) * M];
            cplx* v = &Vel[d][size_t(ik) * M];
            zgemm(nb, cplx(0, -1), r, h, 0., v);
            zgemm(nb, cplx(0, 1), h, r, 1., v);
        }
        if (kgrad) {  // V += grad_k H0 along d, in k (Simulation.cpp:554-556)
            const double u[3] = {double(d == 0), double(d == 1), double(d == 2)};
            kgradient(Vel[d], H0k, u);
            for (auto& v : Vel[d]) v /= det;
            continue;
        }
        fft_planes(Vel[d], -1);
        for (int iR = 0; iR < Nk; ++iR) {
            double Rc[3] = {0, 0, 0};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Rc[i] += A[i][j] * (j == 0 ? Rgc[0][iR] : j == 1 ? Rgc[1][iR] : Rgc[2][iR]);
            double Rd = Rc[d];
            for (int h = 0; h < M; ++h) Vel[d][size_t(iR) * M + h] += cplx(0, 1) * Rd * H0R[size_t(iR) * M + h];
        }
        fft_planes(Vel[d], +1);
        for (auto& v : Vel[d]) v /= det;
    }
}

void Simulation::setup_coulomb(const TBModel& tb)
{
    // ModelCoulomb.cpp:89-134
    int iorig = -1;
    for (int iR = 0; iR < tb.nR; ++iR)
        if (std::abs(tb.R[iR][0]) + std::abs(tb.R[iR][1]) + std::abs(tb.R[iR][2]) < 1e-6) iorig = iR;
    if (iorig < 0) throw std::runtime_error("tb model has no R=0 block");
    // Wannier centres = diagonal of r(R=0); tb.r is already in bohr here
    std::vector<std::array<double, 3>> c(nb);
    for (int n = 0; n < nb; ++n)
        for (int d = 0; d < 3; ++d) c[n][d] = tb.r[d][size_t(iorig) * M + n * nb + n].real();
    std::array<double, 3> mind = {100, 0, 0};
    double mindn = 100;
    for (int n = 0; n < nb; ++n)
        for (int m = n + 1; m < nb; ++m) {
            std::array<double, 3> dd = {c[n][0] - c[m][0], c[n][1] - c[m][1], c[n][2] - c[m][2]};
            double dn = std::sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
            if (dn < mindn && dn > 1.e-05) { mind = dd; mindn = dn; }
        }
    // r0 (ModelCoulomb.cpp:187-215), converted to a.u. (Simulation.cpp:29-33)
    std::vector<double> r0au;
    for (double v : par.r0) r0au.push_back(convert(v, unit(par.r0_units), AuLength));
    double r0[3];
    r0[0] = r0au[0];
    if (r0au.size() == 1) { r0[1] = r0au[0]; r0[2] = r0au[0]; }
    else if (r0au.size() == 2) { r0[1] = r0au[1]; r0[2] = (r0au[0] + r0au[1]) / 2.; }
    else { r0[1] = r0au[1]; r0[2] = r0au[2]; }
    double r0avg = (r0[0] + r0[1] + r0[2]) / 3.;
    double eps_ = par.epsilon;
    auto Wf = [&](const double* r) -> double {
        double rr[3] = {r[0] / r0[0], r[1] / r0[1], r[2] / r0[2]};
        double n = std::sqrt(dot3(rr, rr));
        if (n < threshold) {
            double mr[3] = {mind[0] / r0[0], mind[1] / r0[1], mind[2] / r0[2]};
            n = std::sqrt(dot3(mr, mr));
        }
        return pi / (2. * r0avg * eps_) * (struve0(n) - ::y0(n));
    };
    auto Vf = [&](const double* r) -> double {
        double n = std::sqrt(dot3(r, r));
        return (n < mindn) ? 2. / mindn : 2. / n;
    };
    Wscr.assign(size_t(Nk) * M, 0.);
    Vbare.assign(size_t(Nk) * M, 0.);
    for (int iR = 0; iR < Nk; ++iR) {
        double Rc[3] = {0, 0, 0};
        double Rcr[3] = {Rgc[0][iR], Rgc[1][iR], Rgc[2][iR]};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Rc[i] += A[i][j] * Rcr[j];
        for (int n = 0; n < nb; ++n) {
            double r[3] = {c[n][0] - Rc[0], c[n][1] - Rc[1], c[n][2] - Rc[2]};
            for (int m = 0; m < nb; ++m) {
                double rr[3];
                if (par.coulomb_center_quirk) {  // ModelCoulomb.cpp:76-79: r accumulates
                    for (int d = 0; d < 3; ++d) r[d] -= c[m][d];
                    for (int d = 0; d < 3; ++d) rr[d] = r[d];
                } else {
                    for (int d = 0; d < 3; ++d) rr[d] = c[n][d] - Rc[d] - c[m][d];
                }
                Vbare[size_t(iR) * M + n * nb + m] = Vf(rr);
                Wscr[size_t(iR) * M + n * nb + m] = Wf(rr);
            }
        }
    }
    // Hartree = sum_R V(R), symmetrised (Coulomb.cpp:41-58)
    hartree.assign(M, 0.);
    for (int iR = 0; iR < Nk; ++iR)
        for (int h = 0; h < M; ++h) hartree[h] += Vbare[size_t(iR) * M + h];
    for (int a = 0; a < nb; ++a)
        for (int b = a + 1; b < nb; ++b) {
            cplx v = (hartree[a * nb + b] + hartree[b * nb + a]) / 2.;
            hartree[a * nb + b] = v;
            hartree[b * nb + a] = v;
        }
}

// ------------------------------------------------------------------ k-space gradient
// Pseudo-inverse of a real m x n matrix (row-major) as Matrix<double>::pseudoinv
// (Geometry/Matrix.cpp:130-161): V S^+ U^T, where singular values <= 1e-8 are
// multiplied by s instead of 1/s.  Here through the eigen-decomposition of
// A^T A (eigenvalues s^2, eigenvectors V) and U = A V / s.
static std::vector<double> pseudoinverse(const std::vector<double>& A, int m, int n)
{
    std::vector<cplx> AtA(size_t(n) * n, 0.);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double acc = 0.;
            for (int r = 0; r < m; ++r) acc += A[r * n + i] * A[r * n + j];
            AtA[i * n + j] = acc;
        }
    std::vector<double> w(n);
    std::vector<cplx> V(size_t(n) * n);
    eigh(n, AtA.data(), w.data(), V.data());
    std::vector<double> P(size_t(n) * m, 0.);
    for (int c = 0; c < n; ++c) {
        const double sv = std::sqrt(std::max(w[c], 0.));
        if (sv <= 1.e-8) continue;  // U undefined; its term is O(s) <= 1e-8
        // u_c = A v_c / s, P += v_c (1/s) u_c^T
        std::vector<double> u(m, 0.);
        for (int r = 0; r < m; ++r)
            for (int i = 0; i < n; ++i) u[r] += A[r * n + i] * V[i * n + c].real();
        for (int r = 0; r < m; ++r) u[r] /= sv;
        for (int i = 0; i < n; ++i)
            for (int r = 0; r < m; ++r) P[i * m + r] += V[i * n + c].real() / sv * u[r];
    }
    return P;
}

// kGradient::initialize for a k mesh (kGradient.cpp:25-41): shells of the
// Gamma-centred mesh by Cartesian norm (SortInShells, :44-85), the smallest
// number of shells whose weights w = A^+ q satisfy A w = q (B1 condition,
// Calculate_nshellsAndweights :88-125, GradientMatrix :161-180), and the
// k+b neighbours (Find_kpb :182-210) as grid offsets.
void Simulation::setup_kgradient()
{
    kgrad = true;
    // reciprocal basis B = 2 pi (A^-1)^T, columns b_j (Model.hpp:46-47)
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                       A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
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
    double Bm[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Bm[i][j] = 2. * pi * inv[j][i];
    // Gamma-centred k mesh, crystal coordinates in [-1/2, 1/2) (MeshGrid.cpp:458-486)
    std::vector<std::array<double, 3>> kc(Nk), kcart(Nk);
    std::vector<double> knorm(Nk);
    for (int i0 = 0; i0 < N[0]; ++i0)
        for (int i1 = 0; i1 < N[1]; ++i1) {
            const int ik = i1 + N[1] * i0;
            const double c[3] = {double(i0) / N[0], double(i1) / N[1], 0.};
            for (int d = 0; d < 3; ++d) kc[ik][d] = reduce_coord(c[d], -0.5, 0.5, 1.0);
            for (int i = 0; i < 3; ++i) kcart[ik][i] = Bm[i][0] * kc[ik][0] + Bm[i][1] * kc[ik][1] + Bm[i][2] * kc[ik][2];
            knorm[ik] = std::sqrt(dot3(kcart[ik].data(), kcart[ik].data()));
        }
    std::vector<double> norms;
    for (int ik = 0; ik < Nk; ++ik) {
        bool found = false;
        for (double v : norms) found = found || std::abs(v - knorm[ik]) < threshold;
        if (!found) norms.push_back(knorm[ik]);
    }
    std::sort(norms.begin(), norms.end());
    std::vector<std::vector<int>> shells(norms.size());
    for (int ik = 0; ik < Nk; ++ik) {
        size_t s = 0;
        while (s < norms.size() && !(std::abs(norms[s] - knorm[ik]) < 1.e-07)) ++s;
        shells[s].push_back(ik);
    }
    if (!(std::abs(norms[0]) < threshold)) throw std::runtime_error("kgradient: no k = 0 shell");
    shells.erase(shells.begin());
    // weights (alpha, beta) pairs xx yy zz xy xz yz
    const int al[6] = {0, 1, 2, 0, 0, 1}, be[6] = {0, 1, 2, 1, 2, 2};
    const double q[6] = {double(N[0] > 1), double(N[1] > 1), double(N[2] > 1), 0., 0., 0.};
    std::vector<double> w;
    for (nshells = 1;; ++nshells) {
        if (nshells > int(shells.size())) throw std::runtime_error("kgradient: not enough shells for the B1 condition");
        std::vector<double> Am(size_t(6) * nshells, 0.);
        for (int j = 0; j < 6; ++j)
            for (int s = 0; s < nshells; ++s)
                for (int ik : shells[s]) Am[j * nshells + s] += kcart[ik][al[j]] * kcart[ik][be[j]];
        const std::vector<double> P = pseudoinverse(Am, 6, nshells);
        w.assign(nshells, 0.);
        for (int s = 0; s < nshells; ++s)
            for (int j = 0; j < 6; ++j) w[s] += P[s * 6 + j] * q[j];
        double res = 0.;
        for (int j = 0; j < 6; ++j) {
            double qg = 0.;
            for (int s = 0; s < nshells; ++s) qg += Am[j * nshells + s] * w[s];
            res += (qg - q[j]) * (qg - q[j]);
        }
        if (std::sqrt(res) < threshold) break;
    }
    shell_weight = w;
    kstencil.clear();
    for (int s = 0; s < nshells; ++s)
        for (int ik : shells[s]) {
            KStencil st;
            st.shell = s;
            st.m0 = int(std::lround(kc[ik][0] * N[0]));
            st.m1 = int(std::lround(kc[ik][1] * N[1]));
            for (int d = 0; d < 3; ++d) st.b[d] = kcart[ik][d];
            kstencil.push_back(st);
        }
}

void Simulation::kgradient(std::vector<cplx>& out, const std::vector<cplx>& F, const double u[3]) const
{
#pragma omp parallel for schedule(static)
    for (int ik = 0; ik < Nk; ++ik) {
        const int i0 = ik / N[1], i1 = ik % N[1];
        for (const KStencil& st : kstencil) {
            const int j0 = ((i0 + st.m0) % N[0] + N[0]) % N[0];
            const int j1 = ((i1 + st.m1) % N[1] + N[1]) % N[1];
            const int jk = j1 + N[1] * j0;
            const double bdotu = st.b[0] * u[0] + st.b[1] * u[1] + st.b[2] * u[2];
            const cplx c = cplx(1., 0.) * shell_weight[st.shell] * bdotu;
            for (int h = 0; h < M; ++h) out[size_t(ik) * M + h] += c * F[size_t(jk) * M + h];
        }
    }
}

void Simulation::fft_planes(std::vector<cplx>& a, int sign) const
{
    const int n0 = N[0], n1 = N[1];
    const double scale = 1. / double(Nk);
#pragma omp parallel
    {
        std::vector<cplx> plane(Nk), scratch(2 * std::max(n0, n1) + 2);
#pragma omp for schedule(static)
        for (int h = 0; h < M; ++h) {
            for (int ik = 0; ik < Nk; ++ik) plane[ik] = a[size_t(ik) * M + h];
            fft2d(plan[0], plan[1], plane.data(), sign, scratch.data());
            if (sign < 0)
                for (int ik = 0; ik < Nk; ++ik) a[size_t(ik) * M + h] = plane[ik] * scale;  // fftPair.cpp:98-105
            else
                for (int ik = 0; ik < Nk; ++ik) a[size_t(ik) * M + h] = plane[ik];
        }
    }
}

void Simulation::E_of_t(double t, double E[3]) const
{
    E[0] = E[1] = E[2] = 0.;
    for (auto& L : lasers) {
        double e[3];
        L.E(t, e);
        for (int d = 0; d < 3; ++d) E[d] += e[d];
    }
}

void Simulation::vector_potential(double t, double out[3])
{
    out[0] = out[1] = out[2] = 0.;
    for (auto& L : lasers) {
        const double dT = t - A_prev;
        if (!(dT < 1.e-07)) {
            double E1[3], E2[3], E3[3];
            L.E(A_prev, E1);
            L.E(A_prev + dT / 2., E2);
            L.E(t, E3);
            for (int d = 0; d < 3; ++d) Apot[d] -= dT / 6. * (E1[d] + 4. * E2[d] + E3[d]);
            A_prev = t;
        }
        for (int d = 0; d < 3; ++d) out[d] += Apot[d];
    }
}

// Simulation::Apply_Peierls_phase (Simulation.cpp:876-899)
void Simulation::peierls_phase(std::vector<cplx>& aR, const double Av[3], int sign) const
{
#pragma omp parallel for schedule(static)
    for (int iR = 0; iR < Nk; ++iR) {
        double Rcr[3] = {Rgc[0][iR], Rgc[1][iR], Rgc[2][iR]};
        double Rc[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Rc[i] += A[i][j] * Rcr[j];
        const cplx ph = std::exp(cplx(0, 1) * double(sign) * dot3(Av, Rc));
        for (int h = 0; h < M; ++h) aR[size_t(iR) * M + h] *= ph;
    }
}

// SourceTerm (Functional_SourceTerm.hpp:11-57)
void Simulation::rhs(std::vector<cplx>& out, double t, const std::vector<cplx>& in)
{
    double E[3];
    E_of_t(t, E);
    if (par.peierls) {
        double Av[3];
        vector_potential(t, Av);
        rhs_EA(out, E, Av, in);
        return;
    }
    rhs_E(out, E, in);
}

// Peierls branch of SourceTerm: no gradient term; the exchange input is
// rho(R) exp(-i A.R); the whole H(R) is multiplied by exp(+i A.R) before
// going to k (Functional_SourceTerm.hpp:20-44)
void Simulation::rhs_EA(std::vector<cplx>& out, const double E[3], const double Av[3], const std::vector<cplx>& in)
{
    std::vector<cplx> HR(size_t(Nk) * M);
#pragma omp parallel for schedule(static)
    for (size_t o = 0; o < HR.size(); ++o) HR[o] = 0. + (H0R[o] + E[0] * rR[0][o] + E[1] * rR[1][o] + E[2] * rR[2][o]);
    if (par.coulomb) {
        std::vector<cplx> pR = in;
        fft_planes(pR, -1);
        peierls_phase(pR, Av, -1);
        for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; ++b)
                HR[a * nb + a] += hartree[a * nb + b] * (pR[b * nb + b] - rho0R[b * nb + b]);
#pragma omp parallel for schedule(static)
        for (size_t o = 0; o < HR.size(); ++o) HR[o] -= Wscr[o] * (pR[o] - rho0R[o]);
    }
    peierls_phase(HR, Av, +1);
    fft_planes(HR, +1);
    out.assign(size_t(Nk) * M, 0.);
#pragma omp parallel for schedule(static)
    for (int ik = 0; ik < Nk; ++ik) {
        const cplx* h = &HR[size_t(ik) * M];
        const cplx* r = &in[size_t(ik) * M];
        cplx* o = &out[size_t(ik) * M];
        zgemm(nb, cplx(0, -1), h, r, 1., o);
        zgemm(nb, cplx(0, 1), r, h, 1., o);
    }
}

void Simulation::rhs_E(std::vector<cplx>& out, const double E[3], const std::vector<cplx>& in)
{
    std::vector<cplx> inR = in;
    fft_planes(inR, -1);
    out.assign(size_t(Nk) * M, 0.);
    std::vector<cplx> HR(size_t(Nk) * M);
    if (kgrad) {
        // gradient in k: Output(k) += (E.grad) rho(k) by finite differences;
        // H(R) as below; the commutator adds to the k-space output
        kgradient(out, in, E);
#pragma omp parallel for schedule(static)
        for (size_t o = 0; o < HR.size(); ++o) HR[o] = 0. + (H0R[o] + E[0] * rR[0][o] + E[1] * rR[1][o] + E[2] * rR[2][o]);
        if (par.coulomb) {
            for (int a = 0; a < nb; ++a)
                for (int b = 0; b < nb; ++b)
                    HR[a * nb + a] += hartree[a * nb + b] * (inR[b * nb + b] - rho0R[b * nb + b]);
#pragma omp parallel for schedule(static)
            for (size_t o = 0; o < HR.size(); ++o) HR[o] -= Wscr[o] * (inR[o] - rho0R[o]);
        }
        fft_planes(HR, +1);
#pragma omp parallel for schedule(static)
        for (int ik = 0; ik < Nk; ++ik) {
            const cplx* h = &HR[size_t(ik) * M];
            const cplx* r = &in[size_t(ik) * M];
            cplx* o = &out[size_t(ik) * M];
            zgemm(nb, cplx(0, -1), h, r, 1., o);
            zgemm(nb, cplx(0, 1), r, h, 1., o);
        }
        return;
    }
#pragma omp parallel for schedule(static)
    for (int iR = 0; iR < Nk; ++iR) {
        // gradient: Output(R) += i (R.E) rho(R)   (kGradient_definitions.hpp:35-41)
        double Rcr[3] = {Rgc[0][iR], Rgc[1][iR], Rgc[2][iR]};
        double Rc[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Rc[i] += A[i][j] * Rcr[j];
        double RE = dot3(Rc, E);
        for (int h = 0; h < M; ++h) {
            size_t o = size_t(iR) * M + h;
            out[o] += 1. * cplx(0, 1) * RE * inR[o];
            // H(R) = H0(R) + E.r(R)   (Simulation.cpp:335-347)
            HR[o] = 0. + (H0R[o] + E[0] * rR[0][o] + E[1] * rR[1][o] + E[2] * rR[2][o]);
        }
    }
    if (par.coulomb) {  // Coulomb.cpp:134-180
        for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; ++b)
                HR[a * nb + a] += hartree[a * nb + b] * (inR[b * nb + b] - rho0R[b * nb + b]);
#pragma omp parallel for schedule(static)
        for (int iR = 0; iR < Nk; ++iR)
            for (int h = 0; h < M; ++h) {
                size_t o = size_t(iR) * M + h;
                HR[o] -= Wscr[o] * (inR[o] - rho0R[o]);
            }
    }
    fft_planes(out, +1);
    fft_planes(HR, +1);
#pragma omp parallel for schedule(static)
    for (int ik = 0; ik < Nk; ++ik) {
        const cplx* h = &HR[size_t(ik) * M];
        const cplx* r = &in[size_t(ik) * M];
        cplx* o = &out[size_t(ik) * M];
        zgemm(nb, cplx(0, -1), h, r, 1., o);
        zgemm(nb, cplx(0, 1), r, h, 1., o);
    }
}

static void sum_with_product(std::vector<cplx>& out, double a, const std::vector<cplx>& x, double b,
                             const std::vector<cplx>& y)
{
    const size_t n = out.size();
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) out[i] = a * x[i] + b * y[i];  // DESolver.hpp:10-20
}

void Simulation::step()
{
    if (par.solver == "RK") {  // DESolver.hpp:155-187
        auto& k = aux[0];
        auto& yaux = aux[1];
        auto& red = aux[2];
        rhs(k, t_current, rho);
        sum_with_product(yaux, 1., rho, dt / 2., k);
        sum_with_product(red, 1., rho, dt / 6., k);
        rhs(k, t_current + dt / 2., yaux);
        sum_with_product(yaux, 1., rho, dt / 2., k);
        sum_with_product(red, 1., red, dt / 3., k);
        rhs(k, t_current + dt / 2., yaux);
        sum_with_product(yaux, 1., rho, dt, k);
        sum_with_product(red, 1., red, dt / 3., k);
        rhs(k, t_current + dt, yaux);
        sum_with_product(rho, 1., red, dt / 6., k);
        t_current += dt;
    } else {  // DESolver.hpp:202-220
        rhs(aux[abindex[0]], t_current, rho);
        for (int i = 0; i < par.order; ++i) sum_with_product(rho, 1., rho, dt * beta[i], aux[abindex[i]]);
        int tmp = abindex[par.order - 1];
        for (int i = par.order - 1; i > 0; --i) abindex[i] = abindex[i - 1];
        abindex[0] = tmp;
        t_current += dt;
    }
}

bool Simulation::print_now(double t) const
{
    return (int(std::round(t / dt)) % par.printresolution_pulse == 0);  // Simulation.cpp:240-257
}

// Operator::go_to_bloch / go_to_wannier (Operator.hpp:528-564), in the
// reference's association: to Bloch (U^dagger O) U, to Wannier (U O) U^dagger
void Simulation::rotate(bool to_bloch, const std::vector<cplx>& in, std::vector<cplx>& out) const
{
    out.assign(size_t(Nk) * M, 0.);
#pragma omp parallel for schedule(static)
    for (int ik = 0; ik < Nk; ++ik) {
        std::vector<cplx> tmp(M);
        const cplx* u = &U[size_t(ik) * M];
        const cplx* ud = &Udag[size_t(ik) * M];
        zgemm(nb, 1., to_bloch ? ud : u, &in[size_t(ik) * M], 0., tmp.data());
        zgemm(nb, 1., tmp.data(), to_bloch ? u : ud, 0., &out[size_t(ik) * M]);
    }
}

void Simulation::populations(std::vector<double>& bloch, std::vector<double>& wannier) const
{
    // Simulation.cpp:439-470: go_to(bloch) = U^dagger rho U, then R=0 element = k-mean
    std::vector<cplx> sb(nb, 0.), sw(nb, 0.);
    std::vector<cplx> tmp(M), rb(M);
    for (int ik = 0; ik < Nk; ++ik) {
        const std::vector<cplx>& r = par.peierls ? rho_obs : rho;
        zgemm(nb, 1., &Udag[size_t(ik) * M], &r[size_t(ik) * M], 0., tmp.data());
        zgemm(nb, 1., tmp.data(), &U[size_t(ik) * M], 0., rb.data());
        for (int n = 0; n < nb; ++n) {
            sb[n] += rb[n * nb + n];
            sw[n] += r[size_t(ik) * M + n * nb + n];
        }
    }
    bloch.resize(nb);
    wannier.resize(nb);
    for (int n = 0; n < nb; ++n) {
        double pb = (sb[n] / double(Nk)).real();
        bloch[n] = (n < par.filledbands) ? 1. - pb : pb;
        wannier[n] = (sw[n] / double(Nk)).real();
    }
}

void Simulation::velocity(double v[6]) const
{
    // Simulation.cpp:583-622: v_d = (1/Nk) sum_k Tr(V_d(k) rho(k))
    for (int d = 0; d < 3; ++d) {
        cplx acc = 0.;
        for (int ik = 0; ik < Nk; ++ik) {
            const cplx* V = &Vel[d][size_t(ik) * M];
            const cplx* r = &(par.peierls ? rho_obs : rho)[size_t(ik) * M];
            for (int n = 0; n < nb; ++n)
                for (int l = 0; l < nb; ++l) acc += V[n * nb + l] * r[l * nb + n];
        }
        acc /= double(Nk);
        v[2 * d] = acc.real();
        v[2 * d + 1] = acc.imag();
    }
}

void Simulation::rho_R(std::vector<cplx>& out) const
{
    out = rho;
    fft_planes(out, -1);
}

void Simulation::propagate(Observables* obs, int nsteps_override)
{
    int n = nsteps_override >= 0 ? nsteps_override : nsteps;
    for (int it = 0; it < n; ++it) {
        if (obs && print_now(t_current)) {  // Simulation.cpp:420-429
            // Laser_A.txt: the vector potential is evaluated at every print,
            // in either gauge, which advances its integration state (:425)
            double Av[3];
            vector_potential(t_current, Av);
            obs->laser_A.push_back({Av[0], Av[1], Av[2]});
            if (par.peierls) {
                // the printed observables read rho phased by exp(-i A.R) (Simulation.cpp:447-449,593-595)
                rho_obs = rho;
                fft_planes(rho_obs, -1);
                peierls_phase(rho_obs, Av, -1);
                fft_planes(rho_obs, +1);
            }
            obs->time.push_back(t_current);
            double E[3];
            E_of_t(t_current, E);
            obs->laser.push_back({E[0], E[1], E[2]});
            std::vector<double> pb, pw;
            populations(pb, pw);
            obs->pop_bloch.push_back(pb);
            obs->pop_wannier.push_back(pw);
            std::array<double, 6> v;
            velocity(v.data());
            obs->velocity.push_back(v);
        }
        step();
    }
}


std::vector<cplx> lindhard(int nb, int N0, int N1, int filled, const cplx* U, const double* eps)
{
    const int Nk = N0 * N1, M = nb * nb;
    std::vector<cplx> Udag(size_t(Nk) * M);
    for (int k = 0; k < Nk; ++k)
        for (int a = 0; a < nb; ++a)
            for (int n = 0; n < nb; ++n) Udag[size_t(k) * M + n * nb + a] = std::conj(U[size_t(k) * M + a * nb + n]);
    std::vector<cplx> X(size_t(Nk) * M, 0.);
#pragma omp parallel for schedule(dynamic)
    for (int iq = 0; iq < Nk; ++iq) {
        const int q0 = iq / N1, q1 = iq % N1;
        for (int ik = 0; ik < Nk; ++ik) {
            // MeshGrid::find(k + q) on the periodic mesh
            const int k0 = ik / N1, k1 = ik % N1;
            const int ikp = ((k0 + q0) % N0) * N1 + (k1 + q1) % N1;
            const cplx* Uk = U + size_t(ik) * M;
            const cplx* Ukp = U + size_t(ikp) * M;
            const cplx* Udk = Udag.data() + size_t(ik) * M;
            const cplx* Udkp = Udag.data() + size_t(ikp) * M;
            for (int a = 0; a < nb; ++a)
                for (int b = 0; b < nb; ++b)
                    for (int c = filled; c < nb; ++c)
                        for (int v = 0; v < filled; ++v)
                            X[size_t(iq) * M + a * nb + b] +=
                                (Uk[a * nb + v] * Ukp[b * nb + c] * Udkp[c * nb + a] * Udk[v * nb + b]) /
                                (double(Nk) * (eps[size_t(ik) * nb + v] - eps[size_t(ikp) * nb + c]));
        }
    }
    return X;
}

void Simulation::screened_potential(std::vector<cplx>& XR, std::vector<cplx>& VR, std::vector<cplx>& EpsR,
                                    std::vector<cplx>& InvEpsR, std::vector<cplx>& WR) const
{
    // ResponseFunction, then X_.go_to_R() (sign -1, 1/N)
    XR = lindhard(nb, N[0], N[1], par.filledbands, U.data(), eps.data());
    fft_planes(XR, -1);
    VR.assign(size_t(Nk) * M, 0.);
    EpsR.assign(size_t(Nk) * M, 0.);
    InvEpsR.assign(size_t(Nk) * M, 0.);
    WR.assign(size_t(Nk) * M, 0.);
    for (int iR = 0; iR < Nk; ++iR) {
        double Rc[3];
        for (int i = 0; i < 3; ++i) Rc[i] = A[i][0] * Rgc[0][iR] + A[i][1] * Rgc[1][iR] + A[i][2] * Rgc[2][iR];
        for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; ++b) {
                const size_t o = size_t(iR) * M + a * nb + b;
                // BareCoulomb (:79-104)
                double d2 = 0.;
                for (int i = 0; i < 3; ++i) {
                    const double x = rwann[a][i] + Rc[i] - rwann[b][i];
                    d2 += x * x;
                }
                const double d = std::sqrt(d2);
                VR[o] = d < 1.e-05 ? cplx(0.) : cplx(1. / d);
                // Epsilon (:122-132), inverse (:149-157), W (:168-174)
                EpsR[o] = -VR[o] * XR[o] + 1.;
                InvEpsR[o] = std::abs(EpsR[o]) < 1.e-03 ? cplx(0.) : 1. / EpsR[o];
                WR[o] = InvEpsR[o] * VR[o];
            }
    }
}

}  // namespace oracle
