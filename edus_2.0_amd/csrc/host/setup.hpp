// Host-side FP64 set-up of an EDUS simulation: the part of the reference's
// Simulation constructor (src/Simulation/Simulation.cpp:10-233) that runs
// once, restated in C++ next to the GPU engine.  Class names follow the
// reference so a maintainer can map one onto the other.
#pragma once
#include <array>
#include <complex>
#include <string>
#include <vector>

namespace edus {
namespace host {

using cplx = std::complex<double>;

// physical constants, CODATA values as in src/Constants.hpp:7-23
constexpr double kPi = 3.14159265358979323846;
constexpr double kBohr = 5.2917721090380e-11;
constexpr double kElectronVolt = 1.602176634e-19;
constexpr double kRydberg = 2.1798723611035e-18;
constexpr double kAuTime = 2.4188843265857e-17;
constexpr double kEpsilon0 = 8.8541878188e-12;
constexpr double kAuField = 5.14220674763e+11;
constexpr double kFineStructure = 7.2973525693e-03;
constexpr double kSpeedOfLightSI = 299792458.;
constexpr double kSpeedOfLight = 1. / kFineStructure;
constexpr double kAuIntensity = .5 * kSpeedOfLightSI * kEpsilon0 * kAuField * kAuField;
constexpr double kThreshold = 1.e-08;

// src/ConvertUnits.{hpp,cpp}
enum class UnitType { Length, Time, Energy, Intensity, Null };
struct Unit {
    UnitType type;
    double value;
};
Unit unit(const std::string& name);
double Convert(double v, const Unit& in, const Unit& out);
extern const Unit Angstrom, ElectronVolt, AuTime, AuLength, AuEnergy, AuIntensity;

// src/Model/Model.hpp + src/Wannier/ReadWannier.cpp
struct Material {
    int nb = 0, nR = 0;
    double A[3][3];                        // A[i][j]: component i of lattice vector j, bohr
    std::vector<std::array<double, 3>> R;  // tb R list, crystal coordinates
    std::vector<cplx> H;                   // [nR][nb][nb], Hartree
    std::vector<cplx> r[3];                // [nR][nb][nb], bohr
    std::vector<int> ndegen;               // parsed, never applied (as the reference)
    std::array<double, 3> wannier_centre(int n) const;  // diag r(R=0)
};
Material read_material(const std::string& seedname);
void write_tb(const std::string& path, const Material& m);  // reverse of the parser (PrintWannier.cpp)

// Hermitian eigen-decomposition (ascending eigenvalues, eigenvectors in
// columns, row-major) -- the role of LAPACKE_zheevd in Matrix.cpp:9-23.
void hermitian_eigen(int n, const cplx* A, double* w, cplx* V);

// src/Laser/Laser.{hpp,cpp}
struct Laser {
    double amplitude = 0, omega = 0, t0 = 0, phase = 0, duration = 0;
    double pol[3] = {0, 0, 0};
    void field(double t, double E[3]) const;
};
struct SetOfLaser {
    std::vector<Laser> lasers;
    void field(double t, double E[3]) const;
    // SetOfLaser::VectorPotential (Laser.hpp:152-158) over Laser::VectorPotential
    // (Laser.cpp:136-152).  The reference keeps A and the previous call time in
    // function-level statics shared by every laser: the first laser integrates
    // its own field, A -= dT/6 (E1 + 4 E2 + E3), from the previous call time,
    // and every laser adds the same A.  Here the state lives in this object.
    void vector_potential(double t, double A[3]);
    double Astate[3] = {0, 0, 0}, A_prev = 0.0;
};

// src/ModelCoulomb/ModelCoulomb.cpp
double struve0(double x);

}  // namespace host
}  // namespace edus
