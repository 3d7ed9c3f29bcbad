// EDUS FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// This is a plain restatement of the reference's propagation algorithm
// (gcistaro/EDUS_2.0, C++/MKL/FFTW), written from its source behaviour so that
// the GPU engine can be checked against it.  Only tests/, the smoke check and
// the CPU-baseline leg of bench.py may load it; the product never does.
//
// It keeps the reference's own structure on purpose:
//   * full nb x nb matrices per k-point, layout [k][n][m] (BlockMatrix.hpp:48),
//   * three 2-D FFTs per right-hand side in SourceTerm order
//     (Functional_SourceTerm.hpp:11-57),
//   * two complex GEMMs per k for the commutator (BlockMatrix_definitions.hpp:203-227),
//   * RK4 / AB4 / AB5 exactly as DESolver.hpp:155-220 (AB history starts at zero).
//
// Parity pinning of every function is listed in the test files; the
// functions without a closed form or golden vector are marked there.
#pragma once
#include <complex>
#include <string>
#include <vector>
#include <array>

namespace oracle {

using cplx = std::complex<double>;

// Reference Constants.hpp:7-23 (CODATA values).
constexpr double pi = 3.14159265358979323846;
constexpr double Bohr_value = 5.2917721090380e-11;
constexpr double ElectronVolt_value = 1.602176634e-19;
constexpr double Rydberg_value = 2.1798723611035e-18;
constexpr double AtomicUnitOfTime = 2.4188843265857e-17;
constexpr double Epsilon0 = 8.8541878188e-12;
constexpr double E0field = 5.14220674763e+11;
constexpr double FineStructure = 7.2973525693e-03;
constexpr double SpeedOfLight_SI = 299792458.;
constexpr double SpeedOfLight = 1. / FineStructure;
constexpr double AtomicUnitOfIntensity = .5 * SpeedOfLight_SI * Epsilon0 * E0field * E0field;
constexpr double threshold = 1.e-08;

// Unit handling (ConvertUnits.cpp:7-37): a unit is (type, SI value); an
// unknown name gives the null unit whose value is 0.
enum UnitType { LENGTH, TIME, ENERGY, INTENSITY, NULLTYPE };
struct Unit { UnitType type; double value; };
Unit unit(const std::string& name);
double convert(double v, const Unit& in, const Unit& out);
extern const Unit AuTime, AuLength, AuEnergy, AuIntensity, Angstrom, ElectronVolt;

// ---------------------------------------------------------------- model
struct TBModel {
    int nb = 0, nR = 0;
    double unitcell[3][3];                 // rows = lattice vectors, angstrom (as in file)
    std::vector<int> degeneracy;           // parsed, never applied (ReadWannier.cpp:5-17)
    std::vector<std::array<double, 3>> R;  // crystal coordinates
    std::vector<cplx> H;                   // [nR][nb][nb], eV
    std::vector<cplx> r[3];                // [nR][nb][nb], angstrom
};
TBModel read_tb(const std::string& path_tb_dat);

// ---------------------------------------------------------------- FFT
// Own mixed-radix FFT (any length; radices 2,3,4,5 and a direct DFT for
// other primes).  Sign convention as FFTW: sign=-1 means exp(-2 pi i jx/N).
struct FFTPlan1D {
    int n = 0;
    std::vector<int> factors;
    std::vector<cplx> tw_fwd, tw_bwd;   // exp(-+2 pi i j / n)
};
FFTPlan1D make_plan(int n);
void fft1d(const FFTPlan1D& p, cplx* x, int stride, int sign, cplx* scratch);
// 2-D transform of one contiguous N0 x N1 plane (row-major).
void fft2d(const FFTPlan1D& p0, const FFTPlan1D& p1, cplx* plane, int sign, cplx* scratch);

// ---------------------------------------------------------------- linalg
// Complex Hermitian eigen-decomposition by cyclic Jacobi rotations.
// Eigenvalues ascending; eigenvectors are the columns of V (row-major
// V[i*n+j] = component i of vector j), matching LAPACKE_zheevd row-major
// output used at Matrix.cpp:9-23.
void eigh(int n, const cplx* A, double* w, cplx* V);

// ---------------------------------------------------------------- laser
struct LaserParams {
    double intensity = 0; std::string intensity_units = "Wcm2";
    double frequency = 0; std::string frequency_units = "eV";
    double wavelength = 0; std::string wavelength_units = "nm";
    double polarization[3] = {0, 0, 0};
    double cycles = 0;
    double t0 = 0; std::string t0_units = "fs";
    double phase = 0;
};
struct Laser {
    double amplitude = 0, omega = 0, t0 = 0, phase = 0, duration = 0;
    double pol[3] = {0, 0, 0};
    void E(double t, double out[3]) const;   // Laser.cpp:131-134
};

// ---------------------------------------------------------------- simulation
struct SimParams {
    std::string tb_file;                       // prefix, "_tb.dat" appended
    int grid[3] = {10, 10, 1};
    int filledbands = 0;
    double dt = 0.01; std::string dt_units = "autime";
    double initialtime = 0; std::string initialtime_units = "autime";
    double finaltime = 0; std::string finaltime_units = "autime";
    int printresolution = 1, printresolution_pulse = 0;
    std::string solver = "RK"; int order = 4;
    bool coulomb = false; double epsilon = 1.0;
    std::vector<double> r0 = {10.0}; std::string r0_units = "angstrom";
    bool coulomb_center_quirk = true;          // ModelCoulomb.cpp:76-79 behaviour
    bool peierls = false;                      // Peierls gauge (Functional_SourceTerm.hpp:20-44)
    std::string gradient_space = "R";          // "k": finite-difference gradient (kGradient.cpp:25-210)
    std::vector<LaserParams> lasers;
};

struct Observables {
    std::vector<double> time;                  // Time.txt
    std::vector<std::array<double, 3>> laser;  // Laser.txt
    std::vector<std::array<double, 3>> laser_A;  // Laser_A.txt
    std::vector<std::vector<double>> pop_bloch, pop_wannier;
    std::vector<std::array<double, 6>> velocity;
};

class Simulation {
  public:
    explicit Simulation(const SimParams& p);
    // reference dims
    int nb, N[3], Nk, M;
    double A[3][3];          // A[i][j]: Cartesian component i of lattice vector j (bohr)
    std::vector<cplx> H0k, rk[3];    // material H(k), r(k) after dft (+make_hermitian for r)
    std::vector<cplx> H0R, rR[3];    // H0_, r_ on the FFT R grid (after go_to_R)
    std::vector<cplx> U, Udag;       // eigenvectors per k
    std::vector<double> eps;         // eigenvalues per k
    std::vector<cplx> rho;           // density matrix, Wannier gauge, k space
    std::vector<cplx> Vel[3];        // velocity operator in k
    std::vector<cplx> rho0R;         // DM0 (R space) for Coulomb
    std::vector<cplx> Wscr, Vbare;   // screened / bare potential [R][n][m]
    std::vector<cplx> hartree;       // nb x nb
    std::vector<double> Rgc[3];      // Gamma-centred R in crystal coords, per FFT R index
    std::vector<Laser> lasers;
    SimParams par;
    double dt, t_current = 0.0;
    int nsteps;                      // int((tf-t0)/dt)+2
    // integrator state (DESolver.hpp:45-49)
    std::vector<cplx> aux[5];
    int abindex[5] = {0, 1, 2, 3, 4};
    double beta[5] = {0, 0, 0, 0, 0};

    void E_of_t(double t, double E[3]) const;
    // SetOfLaser::VectorPotential (Laser.hpp:152-158) over Laser::VectorPotential
    // (Laser.cpp:136-152): one integration state shared by every laser (the
    // reference keeps it in function-level statics): A -= dT/6 (E1 + 4 E2 + E3)
    // of the first laser from the previous call time, then every laser adds A.
    void vector_potential(double t, double out[3]);
    double Apot[3] = {0, 0, 0}, A_prev = 0.0;
    // Peierls-gauge right-hand side with field E and vector potential Av
    void rhs_EA(std::vector<cplx>& out, const double E[3], const double Av[3], const std::vector<cplx>& in);
    // O(R) *= exp(i sign Av.R) for every plane of an R-space array
    void peierls_phase(std::vector<cplx>& aR, const double Av[3], int sign) const;
    std::vector<cplx> rho_obs;  // the density matrix the observables read (phased in the Peierls gauge)
    std::vector<double> last_laser_A;  // Laser_A rows of the last propagate through the C API
    // k-space finite-difference gradient (gradient_space "k"): the shells of
    // the Gamma-centred k mesh and their weights (kGradient.cpp:44-125), and
    // for every b of the first nshells shells its grid offset and Cartesian b
    bool kgrad = false;
    int nshells = 0;
    std::vector<double> shell_weight;
    struct KStencil { int shell, m0, m1; double b[3]; };
    std::vector<KStencil> kstencil;
    // Derivative(k) += sum_s w_s sum_{b in s} (b.u) F(k+b)  (kGradient_definitions.hpp:19-31)
    void kgradient(std::vector<cplx>& out, const std::vector<cplx>& F, const double u[3]) const;
    void rhs(std::vector<cplx>& out, double t, const std::vector<cplx>& in);
    // same right-hand side with the field given instead of evaluated at t
    void rhs_E(std::vector<cplx>& out, const double E[3], const std::vector<cplx>& in);
    void step();                                    // DESolver::Propagate
    void propagate(Observables* obs, int nsteps_override = -1);
    void populations(std::vector<double>& bloch, std::vector<double>& wannier) const;
    void rotate(bool to_bloch, const std::vector<cplx>& in, std::vector<cplx>& out) const;
    void velocity(double v[6]) const;
    void rho_R(std::vector<cplx>& out) const;       // FFT k->R of rho (1/N)
    // Wannier centres (Model.hpp:72-78): the diagonal of r(R = 0), bohr
    std::vector<std::array<double, 3>> rwann;
    // ScreenedPotential::Calculate (Screened_potential.cpp:139-199) on this
    // simulation's eigenvectors and energies: X(R) (ResponseFunction then
    // go_to_R), the bare Coulomb V(R), Eps(R) = 1 - V X, its elementwise
    // inverse and W(R) = InvEps V, each [R][n][m] in grid order
    void screened_potential(std::vector<cplx>& XR, std::vector<cplx>& VR, std::vector<cplx>& EpsR,
                            std::vector<cplx>& InvEpsR, std::vector<cplx>& WR) const;
    bool print_now(double t) const;
    void fft_planes_public(std::vector<cplx>& a, int sign) const { fft_planes(a, sign); }

  private:
    FFTPlan1D plan[2];
    void fft_planes(std::vector<cplx>& a, int sign) const;   // all nb^2 planes of [k][n][m]
    void setup_coulomb(const TBModel& tb);
    void setup_kgradient();
};

double struve0(double x);                          // ModelCoulomb.cpp:7-44

// ScreenedPotential::ResponseFunction (Screened_potential.cpp:17-58): the
// Lindhard response X(q) [q][a][b] for eigenvectors U [k][a][n] (bands in
// columns, ascending) and energies eps [k][n] on an N0 x N1 grid, with the
// reference's loop order (q, k, a, b, c, v) and its per-term division.
std::vector<cplx> lindhard(int nb, int N0, int N1, int filled, const cplx* U, const double* eps);

}  // namespace oracle
