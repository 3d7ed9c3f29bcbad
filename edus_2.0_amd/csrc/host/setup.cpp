// Host FP64 set-up pieces (see setup.hpp).
#include "setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace edus {
namespace host {

// ------------------------------------------------------------ units
const Unit Angstrom{UnitType::Length, 1.e-10};
static const Unit NanoMeters{UnitType::Length, 1.e-09};
const Unit ElectronVolt{UnitType::Energy, kElectronVolt};
static const Unit Joule{UnitType::Energy, 1.};
static const Unit Wcm2{UnitType::Intensity, 1.e+04};
static const Unit FemtoSeconds{UnitType::Time, 1.e-15};
const Unit AuIntensity{UnitType::Intensity, kAuIntensity};
const Unit AuTime{UnitType::Time, kAuTime};
const Unit AuLength{UnitType::Length, kBohr};
const Unit AuEnergy{UnitType::Energy, 2. * kRydberg};
static const Unit NullUnit{UnitType::Null, 0.};

// src/ConvertUnits.cpp:22-37 -- lower-case names only; anything else is the
// null unit (so e.g. "eV" converts every value to 0, as in the reference).
Unit unit(const std::string& s)
{
    static const std::pair<const char*, const Unit*> table[] = {
        {"angstrom", &Angstrom}, {"nanometers", &NanoMeters}, {"nm", &NanoMeters},
        {"electronvolt", &ElectronVolt}, {"joule", &Joule}, {"wcm2", &Wcm2},
        {"femtoseconds", &FemtoSeconds}, {"fs", &FemtoSeconds}, {"auintensity", &AuIntensity},
        {"autime", &AuTime}, {"aulength", &AuLength}, {"auenergy", &AuEnergy},
    };
    for (auto& e : table)
        if (s == e.first) return *e.second;
    return NullUnit;
}

double Convert(double v, const Unit& in, const Unit& out) { return v / out.value * in.value; }

// ------------------------------------------------------------ tb reader
namespace {
std::vector<std::vector<std::string>> tokenize_lines(const std::string& path)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open tight-binding file " + path);
    std::vector<std::vector<std::string>> lines;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::vector<std::string> w;
        for (std::string t; ss >> t;) w.push_back(t);
        lines.push_back(std::move(w));
    }
    lines.emplace_back();
    lines.emplace_back();
    return lines;
}
double num(const std::string& s) { return std::strtod(s.c_str(), nullptr); }
}  // namespace

// Wannier90 "_tb.dat": comment, 3 lattice rows (angstrom), num_wann, nrpts,
// ndegen list, then H blocks (R line + "m n Re Im") and r blocks
// (R line + "m n Re x Im x Re y Im y Re z Im z"), blank-line separated
// (src/Wannier/ReadWannier.cpp:110-153).
Material read_material(const std::string& seedname)
{
    auto L = tokenize_lines(seedname + "_tb.dat");
    if (L.size() < 8) throw std::runtime_error("tight-binding file too short: " + seedname);
    Material m;
    double uc[3][3];
    for (int i = 0; i < 3; ++i) {
        if (L[1 + i].size() < 3) throw std::runtime_error("bad lattice line in " + seedname);
        for (int j = 0; j < 3; ++j) uc[i][j] = num(L[1 + i][j]);
    }
    m.nb = int(num(L[4].at(0)));
    m.nR = int(num(L[5].at(0)));
    if (m.nb < 1 || m.nR < 1) throw std::runtime_error("bad num_wann / nrpts in " + seedname);
    size_t li = 6;
    for (; li < L.size() && !L[li].empty(); ++li)
        for (auto& w : L[li]) m.ndegen.push_back(int(num(w)));
    while (li < L.size() && L[li].empty()) ++li;
    const int nb = m.nb, nb2 = nb * nb;
    m.R.assign(m.nR, {0, 0, 0});
    m.H.assign(size_t(m.nR) * nb2, 0.);
    for (auto& v : m.r) v.assign(size_t(m.nR) * nb2, 0.);
    for (int block = 0; block < m.nR && li + 1 < L.size() && L[li + 1].size() == 4; ++block) {
        for (int d = 0; d < 3; ++d) m.R[block][d] = num(L[li][d]);
        size_t j = li + 1;
        for (; !L[j].empty(); ++j) {
            const int a = int(num(L[j][0])) - 1, b = int(num(L[j][1])) - 1;
            if (a < 0 || a >= nb || b < 0 || b >= nb) throw std::runtime_error("band index out of range");
            m.H[size_t(block) * nb2 + a * nb + b] = cplx(num(L[j][2]), num(L[j][3]));
        }
        li = j + 1;
    }
    while (li < L.size() && L[li].empty()) ++li;
    for (int block = 0; block < m.nR && li + 1 < L.size() && L[li + 1].size() == 8; ++block) {
        for (int d = 0; d < 3; ++d) m.R[block][d] = num(L[li][d]);
        size_t j = li + 1;
        for (; L[j].size() == 8; ++j) {
            const int a = int(num(L[j][0])) - 1, b = int(num(L[j][1])) - 1;
            if (a < 0 || a >= nb || b < 0 || b >= nb) throw std::runtime_error("band index out of range");
            for (int d = 0; d < 3; ++d)
                m.r[d][size_t(block) * nb2 + a * nb + b] = cplx(num(L[j][2 + 2 * d]), num(L[j][3 + 2 * d]));
        }
        li = j + 1;
    }
    // unit conversion (Model.hpp:41-57): lattice vectors are the columns of A
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m.A[i][j] = Convert(uc[j][i], Angstrom, AuLength);
    for (auto& h : m.H) h = cplx(Convert(h.real(), ElectronVolt, AuEnergy), Convert(h.imag(), ElectronVolt, AuEnergy));
    for (auto& v : m.r)
        for (auto& x : v) x = cplx(Convert(x.real(), Angstrom, AuLength), Convert(x.imag(), Angstrom, AuLength));
    return m;
}

std::array<double, 3> Material::wannier_centre(int n) const
{
    for (int iR = 0; iR < nR; ++iR)
        if (R[iR][0] == 0 && R[iR][1] == 0 && R[iR][2] == 0) {
            const size_t o = size_t(iR) * nb * nb + n * nb + n;
            return {r[0][o].real(), r[1][o].real(), r[2][o].real()};
        }
    throw std::runtime_error("tight-binding model has no R=0 block");
}

void write_tb(const std::string& path, const Material& m)
{
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fprintf(f, "written by edus host set-up\n");
    for (int i = 0; i < 3; ++i)
        std::fprintf(f, "%25.16f %25.16f %25.16f\n", Convert(m.A[0][i], AuLength, Angstrom),
                     Convert(m.A[1][i], AuLength, Angstrom), Convert(m.A[2][i], AuLength, Angstrom));
    std::fprintf(f, "%12d\n%12d\n", m.nb, m.nR);
    for (int i = 0; i < m.nR; ++i) std::fprintf(f, "%5d%s", 1, (i % 15 == 14 || i == m.nR - 1) ? "\n" : "");
    const int nb = m.nb;
    for (int iR = 0; iR < m.nR; ++iR) {
        std::fprintf(f, "\n%5d%5d%5d\n", int(m.R[iR][0]), int(m.R[iR][1]), int(m.R[iR][2]));
        for (int b = 0; b < nb; ++b)
            for (int a = 0; a < nb; ++a) {
                cplx h = m.H[size_t(iR) * nb * nb + a * nb + b];
                std::fprintf(f, "%5d%5d %25.17e %25.17e\n", a + 1, b + 1, Convert(h.real(), AuEnergy, ElectronVolt),
                             Convert(h.imag(), AuEnergy, ElectronVolt));
            }
    }
    for (int iR = 0; iR < m.nR; ++iR) {
        std::fprintf(f, "\n%5d%5d%5d\n", int(m.R[iR][0]), int(m.R[iR][1]), int(m.R[iR][2]));
        for (int b = 0; b < nb; ++b)
            for (int a = 0; a < nb; ++a) {
                std::fprintf(f, "%5d%5d", a + 1, b + 1);
                for (int d = 0; d < 3; ++d) {
                    cplx x = m.r[d][size_t(iR) * nb * nb + a * nb + b];
                    std::fprintf(f, " %25.17e %25.17e", Convert(x.real(), AuLength, Angstrom),
                                 Convert(x.imag(), AuLength, Angstrom));
                }
                std::fprintf(f, "\n");
            }
    }
    std::fclose(f);
}

// ------------------------------------------------------------ eigen
// Cyclic complex Jacobi: each rotation is a phase that makes the (p,q)
// element real followed by a real Jacobi rotation.  Converges to ~1e-16
// relative accuracy for the nb <= 16 blocks used here.
void hermitian_eigen(int n, const cplx* Ain, double* w, cplx* V)
{
    std::vector<cplx> A(Ain, Ain + size_t(n) * n);
    std::fill(V, V + size_t(n) * n, cplx(0.));
    for (int i = 0; i < n; ++i) {
        V[i * n + i] = 1.;
        A[i * n + i] = A[i * n + i].real();
    }
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0., diag = 0.;
        for (int p = 0; p < n; ++p) {
            diag += std::norm(A[p * n + p]);
            for (int q = p + 1; q < n; ++q) off += std::norm(A[p * n + q]);
        }
        if (off <= 1e-34 * diag || off == 0.) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const cplx apq = A[p * n + q];
                const double mag = std::abs(apq);
                if (mag == 0.) continue;
                const double app = A[p * n + p].real(), aqq = A[q * n + q].real();
                const cplx ph = apq / mag;
                const double theta = (aqq - app) / (2. * mag);
                const double t = std::copysign(1., theta) / (std::abs(theta) + std::sqrt(theta * theta + 1.));
                const double c = 1. / std::sqrt(t * t + 1.), s = t * c;
                // columns: G = [[c, s], [-s conj(ph), c conj(ph)]]
                const cplx g10 = -s * std::conj(ph), g11 = c * std::conj(ph);
                for (int i = 0; i < n; ++i) {
                    const cplx ap = A[i * n + p], aq = A[i * n + q];
                    A[i * n + p] = c * ap + g10 * aq;
                    A[i * n + q] = s * ap + g11 * aq;
                    const cplx vp = V[i * n + p], vq = V[i * n + q];
                    V[i * n + p] = c * vp + g10 * vq;
                    V[i * n + q] = s * vp + g11 * vq;
                }
                for (int j = 0; j < n; ++j) {
                    const cplx ap = A[p * n + j], aq = A[q * n + j];
                    A[p * n + j] = c * ap + std::conj(g10) * aq;
                    A[q * n + j] = s * ap + std::conj(g11) * aq;
                }
                A[p * n + q] = A[q * n + p] = 0.;
                A[p * n + p] = app - t * mag;
                A[q * n + q] = aqq + t * mag;
            }
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return A[a * n + a].real() < A[b * n + b].real(); });
    std::vector<cplx> Vs(size_t(n) * n);
    for (int j = 0; j < n; ++j) {
        w[j] = A[order[j] * n + order[j]].real();
        for (int i = 0; i < n; ++i) Vs[i * n + j] = V[i * n + order[j]];
    }
    std::copy(Vs.begin(), Vs.end(), V);
}

// ------------------------------------------------------------ laser
// Laser.cpp:96-104 (sin^2 envelope, open at t0, closed at t0+T), 35-39 (carrier),
// 131-134 (amplitude * envelope * carrier * polarisation)
void Laser::field(double t, double E[3]) const
{
    const double tt = t - t0;
    const double env = (tt <= 1.e-07 || tt > duration) ? 0. : std::pow(std::sin(kPi * tt / duration), 2.);
    const double a = amplitude * env * std::sin(omega * (t - t0) + phase);
    for (int d = 0; d < 3; ++d) E[d] = a * pol[d];
}

void SetOfLaser::field(double t, double E[3]) const
{
    E[0] = E[1] = E[2] = 0.;
    for (auto& l : lasers) {
        double e[3];
        l.field(t, e);
        for (int d = 0; d < 3; ++d) E[d] += e[d];
    }
}

void SetOfLaser::vector_potential(double t, double A[3])
{
    A[0] = A[1] = A[2] = 0.;
    for (const Laser& L : lasers) {
        const double dT = t - A_prev;
        if (!(dT < 1.e-07)) {
            double E1[3], E2[3], E3[3];
            L.field(A_prev, E1);
            L.field(A_prev + dT / 2., E2);
            L.field(t, E3);
            for (int d = 0; d < 3; ++d) Astate[d] -= dT / 6. * (E1[d] + 4. * E2[d] + E3[d]);
            A_prev = t;
        }
        for (int d = 0; d < 3; ++d) A[d] += Astate[d];
    }
}

// ------------------------------------------------------------ struve H0
// Zhang & Jin, "Computation of Special Functions" (1996), STVH0: power series
// for x <= 20, asymptotic series plus Y0 expansion above (the form the
// reference uses, ModelCoulomb.cpp:7-44).
double struve0(double x)
{
    double s = 1.0, r = 1.0;
    if (x <= 20.0) {
        for (int k = 1; k < 61; ++k) {
            r = -r * x / (2.0 * k + 1.0) * x / (2.0 * k + 1.0);
            s += r;
            if (std::fabs(r) < std::fabs(s) * 1.0e-12) break;
        }
        return 2.0 * x / kPi * s;
    }
    int km = int(0.5 * (x + 1.0));
    if (x >= 50.0) km = 25;
    for (int k = 1; k <= km; ++k) {
        r = -r * std::pow((2.0 * k - 1.0) / x, 2);
        s += r;
        if (std::fabs(r) < std::fabs(s) * 1.0e-12) break;
    }
    const double t = 4.0 / x, t2 = t * t;
    const double p0 =
        ((((-.37043e-5 * t2 + .173565e-4) * t2 - .487613e-4) * t2 + .17343e-3) * t2 - 0.1753062e-2) * t2 + .3989422793;
    const double q0 =
        t * (((((.32312e-5 * t2 - 0.142078e-4) * t2 + 0.342468e-4) * t2 - 0.869791e-4) * t2 + 0.4564324e-3) * t2 -
             0.0124669441);
    const double ta0 = x - 0.25 * kPi;
    const double by0 = 2.0 / std::sqrt(x) * (p0 * std::sin(ta0) + q0 * std::cos(ta0));
    return 2.0 / (kPi * x) * s + by0;
}

}  // namespace host
}  // namespace edus
