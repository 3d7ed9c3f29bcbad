// Reference-format simulation on top of the engine: the host C++ mirror of
// src/Simulation/Simulation.cpp and src/DESolver/DESolver.hpp.  Set-up runs
// in FP64 on the host (explicit DFT, eigensystem, lasers, Coulomb model);
// every per-step operation runs in the engine.
#include <hip/hip_runtime.h>

#include <fstream>
#include <iomanip>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/edus_gpu.h"
#include "../engine/engine_internal.hpp"
#include "h5out.hpp"
#include "setup.hpp"

namespace edus {
namespace host {

namespace {

void parallel_for(int n, const std::function<void(int)>& body)
{
    const int nt = std::max(1, std::min<int>(16, int(std::thread::hardware_concurrency())));
    if (n < 64 || nt == 1) {
        for (int i = 0; i < n; ++i) body(i);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            for (int i = t; i < n; i += nt) body(i);
        });
    for (auto& th : pool) th.join();
}

// MeshGrid.cpp:345-371 reduce(): fold x into [lo, up) with period `period`
double fold(double x, double lo, double up, double period)
{
    if (period <= 0) return x;
    while (x < lo && std::abs(x - lo) > kThreshold) x += period;
    while (x > up || std::abs(x - up) < kThreshold) x -= period;
    return x;
}

void matmul(int n, const cplx* A, const cplx* B, cplx* C)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            cplx s = 0.;
            for (int l = 0; l < n; ++l) s += A[i * n + l] * B[l * n + j];
            C[i * n + j] = s;
        }
}

}  // namespace

// DESolver mirror (DESolver.hpp:26-82): owns the time bookkeeping, the
// right-hand side and stage updates live in the engine.
class DESolver {
  public:
    enum SolverType { AB, RK };
    void initialize(Engine* e, SetOfLaser* l, SolverType t, int order, bool seed = false, bool peierls = false)
    {
        seed_ = seed;
        peierls_ = peierls;
        engine_ = e;
        lasers_ = l;
        type_ = t;
        order_ = order;
        if (t == RK && order != 4) throw EngineError(EDUS_ERR_UNSUPPORTED, "RK order must be 4");  // DESolver.hpp:157
        if (t == AB && order != 4 && order != 5)
            throw EngineError(EDUS_ERR_UNSUPPORTED, "order " + std::to_string(order) + "not implemented for AB");
    }
    const double& get_CurrentTime() const { return CurrentTime; }
    const double& get_ResolutionTime() const { return ResolutionTime; }
    void set_InitialTime(double t) { InitialTime = t; }  // CurrentTime is left at 0, as the reference does
    void set_ResolutionTime(double dt) { ResolutionTime = dt; }
    void Propagate() { Propagate(1); }
    void Propagate(int nstep)
    {
        if (nstep <= 0) return;
        const double dt = ResolutionTime;
        if (type_ == RK) {
            std::vector<double> E(size_t(nstep) * 9), A;
            double t = CurrentTime;
            if (peierls_) A.resize(E.size());
            for (int s = 0; s < nstep; ++s) {
                lasers_->field(t, &E[9 * size_t(s)]);
                lasers_->field(t + dt / 2., &E[9 * size_t(s) + 3]);
                lasers_->field(t + dt, &E[9 * size_t(s) + 6]);
                if (peierls_) {
                    // SourceTerm asks for A at every stage time, in this order
                    // (the third call repeats t + dt/2 and changes nothing)
                    lasers_->vector_potential(t, &A[9 * size_t(s)]);
                    lasers_->vector_potential(t + dt / 2., &A[9 * size_t(s) + 3]);
                    lasers_->vector_potential(t + dt, &A[9 * size_t(s) + 6]);
                }
                t += dt;
            }
            if (peierls_)
                for (int s = 0; s < nstep; ++s) engine_->rk4_step_peierls(dt, &E[9 * size_t(s)], &A[9 * size_t(s)]);
            else
                engine_->rk4_steps(dt, nstep, E.data());
            for (int s = 0; s < nstep; ++s) CurrentTime += dt;
        } else {
            std::vector<double> E(size_t(nstep) * 3);
            double t = CurrentTime;
            for (int s = 0; s < nstep; ++s) {
                lasers_->field(t, &E[3 * size_t(s)]);
                t += dt;
            }
            for (int s = 0; s < nstep; ++s) {
                if (seed_ && taken_ < order_ - 1) {
                    // RK4 start-up: its first-stage right-hand sides become the history
                    double E3[9];
                    lasers_->field(CurrentTime, E3);
                    lasers_->field(CurrentTime + dt / 2., E3 + 3);
                    lasers_->field(CurrentTime + dt, E3 + 6);
                    if (peierls_)
                        throw EngineError(EDUS_ERR_UNSUPPORTED, "RK4 start-up of AB in the Peierls gauge");
                    engine_->rk4_step(dt, E3, order_);
                } else if (peierls_) {
                    double A3[3];
                    lasers_->vector_potential(CurrentTime, A3);
                    engine_->ab_step_peierls(dt, order_, &E[3 * size_t(s)], A3);
                } else {
                    engine_->ab_step(dt, order_, &E[3 * size_t(s)]);
                }
                ++taken_;
                CurrentTime += dt;
            }
        }
    }

  private:
    Engine* engine_ = nullptr;
    SetOfLaser* lasers_ = nullptr;
    bool peierls_ = false;
    SolverType type_ = RK;
    int order_ = 4;
    bool seed_ = false;  // RK4 start-up of AB (not the reference's behaviour)
    long taken_ = 0;
    double InitialTime = 0., ResolutionTime = 0., CurrentTime = 0.;
};

class Simulation {
  public:
    explicit Simulation(const edus_sim_params& p);
    int nb, N0, N1, Nk, nsteps, filled, print_stride;
    bool peierls = false;
    double dt;
    Material material;
    SetOfLaser setoflaser;
    std::vector<cplx> H0k, rk[3], U, rho0, W, Vbare, hartree;
    std::vector<double> eps;
    std::unique_ptr<Engine> engine;
    DESolver desolver;

    bool PrintObservables(double t) const { return int(std::round(t / dt)) % print_stride == 0; }
    int propagate(int n, double* times, double* laser, double* popb, double* popw, double* vel,
                  double* laser_A = nullptr);

    // output.h5 (Simulation.cpp:215-229,385-420; GlobalFunctions.cpp:19-66)
    edus_sim_params par;               // the input keys, as given (the /json group)
    std::unique_ptr<H5Out> h5;         // open while edus_sim_write_outputs runs with write_h5
    int h5_counter = 0;                // get_it_sparse (Simulation.cpp:685-691): one node per sparse print
    bool PrintSparse(double t) const;  // PrintObservables(t, true) (Simulation.cpp:240-257)
    void h5_open(const std::string& path);
    void h5_record(double t);

  private:
    double jacobian() const;
    void setup_coulomb(const edus_sim_params& p);
};

double Simulation::jacobian() const
{
    // Simulation.cpp:476-508: 2-D slab, perpendicular axis must be z
    const double(*A)[3] = material.A;
    return std::abs(A[0][0] * A[1][1] - A[1][0] * A[0][1]);
}

Simulation::Simulation(const edus_sim_params& p) : par(p)
{
    // times in a.u. (Simulation.cpp:18-26); step count (:355)
    const double ti = Convert(p.initialtime, unit(p.initialtime_units), AuTime);
    const double tf = Convert(p.finaltime, unit(p.finaltime_units), AuTime);
    dt = Convert(p.dt, unit(p.dt_units), AuTime);
    if (!(dt > 0)) throw EngineError(EDUS_ERR_ARG, "dt must be positive (check dt_units)");
    nsteps = int((tf - ti) / dt) + 2;
    print_stride = p.printresolution_pulse != 0 ? p.printresolution_pulse : p.printresolution;
    if (print_stride <= 0) throw EngineError(EDUS_ERR_ARG, "printresolution must be positive");
    filled = p.filledbands;

    material = read_material(p.tb_file);
    nb = material.nb;
    if (p.grid[2] != 1) throw EngineError(EDUS_ERR_UNSUPPORTED, "only 2-D k grids (grid[2] == 1) are supported");
    int dim = (p.grid[0] > 1) + (p.grid[1] > 1);
    if (dim != 2) throw EngineError(EDUS_ERR_UNSUPPORTED, "the velocity normalisation needs a 2-D grid");
    N0 = p.grid[0];
    N1 = p.grid[1];
    Nk = N0 * N1;
    const int M = nb * nb;
    if (filled < 0 || filled > nb) throw EngineError(EDUS_ERR_ARG, "filledbands out of range");

    // explicit DFT over the tb R list folded onto the Gamma-centred grid (Operator.hpp:331-381,
    // fftPair.cpp:110-143), sign +1; r(k) made Hermitian (BlockMatrix_definitions.hpp:258-284)
    std::vector<std::array<double, 3>> Rf(material.nR);
    const int Nd[3] = {N0, N1, 1};
    for (int i = 0; i < material.nR; ++i)
        for (int d = 0; d < 3; ++d) Rf[i][d] = fold(material.R[i][d], -Nd[d] / 2., Nd[d] / 2., Nd[d]);
    H0k.assign(size_t(Nk) * M, 0.);
    for (auto& v : rk) v.assign(size_t(Nk) * M, 0.);
    parallel_for(Nk, [&](int ik) {
        const int i0 = ik / N1, i1 = ik % N1;
        const double kc[2] = {double(i0) / N0, double(i1) / N1};
        std::vector<cplx> ph(material.nR);
        for (int iR = 0; iR < material.nR; ++iR)
            ph[iR] = std::exp(cplx(0, 1) * 2. * kPi * (kc[0] * Rf[iR][0] + kc[1] * Rf[iR][1]));
        for (int h = 0; h < M; ++h) {
            cplx a = 0., b[3] = {0., 0., 0.};
            for (int iR = 0; iR < material.nR; ++iR) {
                a += ph[iR] * material.H[size_t(iR) * M + h];
                for (int d = 0; d < 3; ++d) b[d] += ph[iR] * material.r[d][size_t(iR) * M + h];
            }
            H0k[size_t(ik) * M + h] = a;
            for (int d = 0; d < 3; ++d) rk[d][size_t(ik) * M + h] = b[d];
        }
        for (int d = 0; d < 3; ++d) {
            cplx* m = &rk[d][size_t(ik) * M];
            for (int a = 0; a < nb; ++a)
                for (int b2 = a; b2 < nb; ++b2) {
                    const cplx t = (m[a * nb + b2] + std::conj(m[b2 * nb + a])) / 2.;
                    m[a * nb + b2] = t;
                    m[b2 * nb + a] = std::conj(t);
                }
        }
    });

    // lasers (Simulation.cpp:121-144)
    for (int i = 0; i < p.nlasers; ++i) {
        const edus_laser_params& lp = p.lasers[i];
        Laser L;
        L.t0 = Convert(lp.t0, unit(lp.t0_units), AuTime);
        L.amplitude = std::sqrt(Convert(lp.intensity, unit(lp.intensity_units), AuIntensity));
        if (std::abs(lp.frequency) > 1.e-07) {
            L.omega = Convert(lp.frequency, unit(lp.frequency_units), AuEnergy);
        } else {
            if (!lp.wavelength) throw EngineError(EDUS_ERR_ARG, "You must specify (nonzero) frequency *xor* wavelength!");
            L.omega = 2 * kPi * kSpeedOfLight / Convert(lp.wavelength, unit(lp.wavelength_units), AuLength);
        }
        L.duration = lp.cycles * (2. * kPi / L.omega);
        L.phase = lp.phase;
        double pol[3], n1 = 0.;
        for (int d = 0; d < 3; ++d) n1 += lp.polarization[d] * lp.polarization[d];
        n1 = std::sqrt(n1);
        if (!(n1 > 0)) throw EngineError(EDUS_ERR_ARG, "laser polarization must be non-zero");
        double n2 = 0.;
        for (int d = 0; d < 3; ++d) { pol[d] = lp.polarization[d] / n1; n2 += pol[d] * pol[d]; }
        n2 = std::sqrt(n2);
        for (int d = 0; d < 3; ++d) L.pol[d] = pol[d] / n2;  // normalised twice (Simulation.cpp:141, Laser.cpp:199)
        setoflaser.lasers.push_back(L);
    }

    // eigensystem (Simulation.cpp:263-286); the initial condition
    // (Functional_InitialCondition.hpp:12-32) is rotated on the device below
    U.assign(size_t(Nk) * M, 0.);
    eps.assign(size_t(Nk) * nb, 0.);
    parallel_for(Nk, [&](int ik) { hermitian_eigen(nb, &H0k[size_t(ik) * M], &eps[size_t(ik) * nb], &U[size_t(ik) * M]); });

    if (p.coulomb) setup_coulomb(p);

    edus_desc d{};
    d.nb = nb;
    d.N0 = N0;
    d.N1 = N1;
    d.precision = p.precision;
    d.layout = p.layout;
    d.H0k = reinterpret_cast<const double*>(H0k.data());
    for (int k = 0; k < 3; ++k) d.xik[k] = reinterpret_cast<const double*>(rk[k].data());
    std::memcpy(d.A, material.A, sizeof(d.A));
    d.coulomb = p.coulomb;
    if (p.coulomb) {
        d.W = reinterpret_cast<const double*>(W.data());
        d.hartree = reinterpret_cast<const double*>(hartree.data());
        d.rho0k = nullptr;  // the engine's initial condition (Coulomb::set_DM0 of the IC, Simulation.cpp:177-181)
    }
    d.U = reinterpret_cast<const double*>(U.data());
    d.filledbands = filled;
    d.peierls = p.peierls;
    d.gradient_space = p.gradient_space;
    peierls = p.peierls != 0;
    EngineOptions opt;
    opt.velocity_det = jacobian();  // velocity operator built on the device (Simulation.cpp:523-576)
    opt.initial_filled = filled;     // rho(t=0) = U diag(filled) U^dagger on the device
    engine = std::make_unique<Engine>(d, opt);
    rho0.assign(size_t(Nk) * M, 0.);  // host copy of the initial state (DM0.txt, edus_sim_get)
    engine->download(engine->y, reinterpret_cast<double*>(rho0.data()));

    const std::string solver = p.solver;
    desolver.initialize(engine.get(), &setoflaser, solver == "AB" ? DESolver::AB : DESolver::RK, p.order,
                        p.ab_seed != 0, p.peierls != 0);
    if (solver != "AB" && solver != "RK") throw EngineError(EDUS_ERR_ARG, "solver must be RK or AB");
    desolver.set_InitialTime(ti);
    desolver.set_ResolutionTime(dt);
}

bool Simulation::PrintSparse(double t) const
{
    // inside any pulse (t0, t0 + duration] the pulse resolution, else the
    // general one; a zero resolution (the reference divides by it) falls back
    // to the other
    int pr = par.printresolution;
    for (const Laser& L : setoflaser.lasers) {
        if (t > L.t0 + 1.e-07 && t < L.t0 + L.duration + 1.e-07) {
            pr = par.printresolution_pulse;
            break;
        }
        pr = par.printresolution;
    }
    if (pr <= 0) pr = par.printresolution > 0 ? par.printresolution : par.printresolution_pulse;
    if (pr <= 0) throw EngineError(EDUS_ERR_ARG, "printresolution must be positive");
    return int(std::round(t / dt)) % pr == 0;
}

void Simulation::h5_open(const std::string& path)
{
    h5 = std::make_unique<H5Out>(path);
    h5_counter = 0;
    H5Out& f = *h5;
    // dump_json_in_h5 (GlobalFunctions.cpp:19-66): the input's own values
    f.group("/json");
    f.write_bytes("/json/tb_file", par.tb_file);
    const int coul = par.coulomb != 0;
    f.write_ints("/json/coulomb", &coul, 1);
    f.write_ints("/json/grid", par.grid, 3);
    f.write_ints("/json/filledbands", &par.filledbands, 1);
    f.write_doubles("/json/dt", &par.dt, 1);
    f.write_ints("/json/printresolution", &par.printresolution, 1);
    f.write_doubles("/json/initialtime", &par.initialtime, 1);
    f.write_doubles("/json/finaltime", &par.finaltime, 1);
    f.write_ints("/json/num_bands", &nb, 1);
    f.write_ints("/json/num_kpoints", &Nk, 1);
    f.write_bytes("/json/solver", par.solver);
    const int one = 1;  // the k-pool size of the reference's MPI build, one process here
    f.write_ints("/json/comm_size", &one, 1);
    for (int i = 0; i < par.nlasers; ++i) {
        const edus_laser_params& L = par.lasers[i];
        const std::string g = "/json/laser" + std::to_string(i);
        f.group(g);
        f.write_doubles(g + "/intensity", &L.intensity, 1);
        f.write_doubles(g + "/frequency", &L.frequency, 1);
        f.write_doubles(g + "/wavelength", &L.wavelength, 1);
        const int cycles = int(L.cycles);  // json get<int>() truncates
        f.write_ints(g + "/cycles", &cycles, 1);
        // the reference stores the cycle count under "polarization" too (:58)
        f.write_ints(g + "/polarization", &cycles, 1);
    }
    // k points in the reciprocal basis, [Nk][3] (Simulation.cpp:208-214), and the
    // two bases as 3 x 3 matrices, column j = vector j, row-major
    std::vector<double> kp(size_t(Nk) * 3, 0.0);
    for (int ik = 0; ik < Nk; ++ik) {
        kp[3 * size_t(ik)] = double(ik / N1) / N0;
        kp[3 * size_t(ik) + 1] = double(ik % N1) / N1;
    }
    f.write_doubles("/json/kpoints", kp.data(), kp.size());
    double Am[9], Bm[9];
    const double(*A)[3] = material.A;
    const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                       A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    const double inv[3][3] = {{(A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det, (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det,
                               (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det},
                              {(A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det, (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det,
                               (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det},
                              {(A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det, (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det,
                               (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Am[3 * i + j] = A[i][j];
            Bm[3 * i + j] = 2. * kPi * inv[j][i];  // 2 pi (A^-1)^T (Model.hpp:46)
        }
    f.write_doubles("/json/A", Am, 9);
    f.write_doubles("/json/B", Bm, 9);
    // the groups of Simulation.cpp:225-229
    for (const char* g : {"/fullH", "/DensityMatrix_k", "/DensityMatrix_k_bloch", "/SelfEnergy", "/time_au"}) f.group(g);
}

void Simulation::h5_record(double t)
{
    H5Out& f = *h5;
    const std::string node = std::to_string(h5_counter++);
    const size_t M = size_t(nb) * nb;
    if (par.toprint_dmk_wannier || par.toprint_dmk_bloch) {
        std::vector<cplx> rho(size_t(Nk) * M);
        engine->download(engine->y, reinterpret_cast<double*>(rho.data()));
        if (par.toprint_dmk_wannier) {
            const std::string g = "/DensityMatrix_k/" + node;
            f.group(g);
            f.group(g + "/0");
            f.write_doubles(g + "/0/local", reinterpret_cast<const double*>(rho.data()), 2 * rho.size());
        }
        if (par.toprint_dmk_bloch) {
            // go_to_bloch (Operator.hpp:528-564): O_b = U^dagger O_w U per k, on the device
            std::vector<cplx> rb(rho.size());
            engine->download_bloch(engine->y, reinterpret_cast<double*>(rb.data()));
            const std::string g = "/DensityMatrix_k_bloch/" + node;
            f.group(g);
            f.group(g + "/0");
            f.write_doubles(g + "/0/local", reinterpret_cast<const double*>(rb.data()), 2 * rb.size());
        }
    }
    f.write_doubles("/time_au/" + node, &t, 1);
}

void Simulation::setup_coulomb(const edus_sim_params& p)
{
    // ModelCoulomb.cpp:89-134 and Coulomb.cpp:19-59
    const int M = nb * nb;
    std::vector<std::array<double, 3>> c(nb);
    for (int n = 0; n < nb; ++n) c[n] = material.wannier_centre(n);
    std::array<double, 3> mind = {100., 0., 0.};
    double mindn = 100.;
    for (int n = 0; n < nb; ++n)
        for (int m = n + 1; m < nb; ++m) {
            const double dx = c[n][0] - c[m][0], dy = c[n][1] - c[m][1], dz = c[n][2] - c[m][2];
            const double dn = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (dn < mindn && dn > 1.e-05) { mind = {dx, dy, dz}; mindn = dn; }
        }
    double r0[3];
    std::vector<double> r0in;
    for (int i = 0; i < std::max(1, std::min(3, p.nr0)); ++i) r0in.push_back(Convert(p.r0[i], unit(p.r0_units), AuLength));
    r0[0] = r0in[0];
    if (r0in.size() == 1) { r0[1] = r0[2] = r0in[0]; }
    else if (r0in.size() == 2) { r0[1] = r0in[1]; r0[2] = (r0in[0] + r0in[1]) / 2.; }
    else { r0[1] = r0in[1]; r0[2] = r0in[2]; }
    const double r0avg = (r0[0] + r0[1] + r0[2]) / 3.;
    const double epsr = p.epsilon;
    auto Wf = [&](const double* r) {
        double q[3] = {r[0] / r0[0], r[1] / r0[1], r[2] / r0[2]};
        double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        if (n < kThreshold) {
            const double m[3] = {mind[0] / r0[0], mind[1] / r0[1], mind[2] / r0[2]};
            n = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        }
        return kPi / (2. * r0avg * epsr) * (struve0(n) - ::y0(n));
    };
    auto Vf = [&](const double* r) {
        const double n = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        return (n < mindn) ? 2. / mindn : 2. / n;
    };
    W.assign(size_t(Nk) * M, 0.);
    Vbare.assign(size_t(Nk) * M, 0.);
    const double(*A)[3] = material.A;
    parallel_for(Nk, [&](int iR) {
        const int j0 = iR / N1, j1 = iR % N1;
        const double R0 = fold(j0, -N0 / 2., N0 / 2., N0), R1 = fold(j1, -N1 / 2., N1 / 2., N1);
        const double Rc[3] = {A[0][0] * R0 + A[0][1] * R1, A[1][0] * R0 + A[1][1] * R1, A[2][0] * R0 + A[2][1] * R1};
        for (int n = 0; n < nb; ++n) {
            double r[3] = {c[n][0] - Rc[0], c[n][1] - Rc[1], c[n][2] - Rc[2]};
            for (int m = 0; m < nb; ++m) {
                double rr[3];
                for (int d = 0; d < 3; ++d) {
                    if (p.coulomb_center_quirk) {
                        r[d] -= c[m][d];  // ModelCoulomb.cpp:76-79: r accumulates over m
                        rr[d] = r[d];
                    } else {
                        rr[d] = c[n][d] - Rc[d] - c[m][d];
                    }
                }
                W[size_t(iR) * M + n * nb + m] = Wf(rr);
                Vbare[size_t(iR) * M + n * nb + m] = Vf(rr);
            }
        }
    });
    hartree.assign(M, 0.);
    for (int iR = 0; iR < Nk; ++iR)
        for (int h = 0; h < M; ++h) hartree[h] += Vbare[size_t(iR) * M + h];
    for (int a = 0; a < nb; ++a)
        for (int b = a + 1; b < nb; ++b) {
            const cplx v = (hartree[a * nb + b] + hartree[b * nb + a]) / 2.;
            hartree[a * nb + b] = hartree[b * nb + a] = v;
        }
    // The engine stores Hermitian fields only.  The exchange term is Hermitian
    // iff W_mn(R) = W_nm(-R); the reference's centre accumulation breaks this
    // whenever more than one Wannier centre sits off the origin.
    double asym = 0., scale = 0.;
    for (int iR = 0; iR < Nk; ++iR) {
        const int j0 = iR / N1, j1 = iR % N1;
        const int mR = ((N0 - j0) % N0) * N1 + (N1 - j1) % N1;
        for (int n = 0; n < nb; ++n)
            for (int m = 0; m < nb; ++m) {
                asym = std::max(asym, std::abs(W[size_t(iR) * M + m * nb + n] - W[size_t(mR) * M + n * nb + m]));
                scale = std::max(scale, std::abs(W[size_t(iR) * M + n * nb + m]));
            }
    }
    if (p.layout == EDUS_LAYOUT_HERMITIAN && asym > 1e-12 * std::max(scale, 1e-300))
        throw EngineError(EDUS_ERR_UNSUPPORTED,
                          "screened exchange W_mn(R) != W_nm(-R): the reference's Wannier-centre accumulation "
                          "(coulomb_center_quirk) makes H non-Hermitian for this model; use the FULL layout "
                          "or set coulomb_center_quirk=0");
}

int Simulation::propagate(int n, double* times, double* laser, double* popb, double* popw, double* vel,
                          double* laser_A)
{
    // Simulation::Propagate / do_onestep (Simulation.cpp:352-432): print before stepping
    int printed = 0;
    int it = 0;
    while (it < n) {
        const double t = desolver.get_CurrentTime();
        if (h5 && PrintSparse(t)) h5_record(t);  // the HDF5 block comes first in do_onestep
        if (PrintObservables(t)) {
            if (times) times[printed] = t;
            double E[3];
            setoflaser.field(t, E);
            if (laser)
                for (int d = 0; d < 3; ++d) laser[3 * printed + d] = E[d];
            // the vector potential is evaluated at every print in either gauge
            // (Laser_A.txt, Simulation.cpp:425), which advances its integration state
            double A[3];
            setoflaser.vector_potential(t, A);
            if (laser_A)
                for (int d = 0; d < 3; ++d) laser_A[3 * printed + d] = A[d];
            const double2* rho_obs = nullptr;
            if (peierls) {  // the printed observables read rho(R) exp(-i A.R) (Simulation.cpp:447-449,593-595)
                engine->peierls_sandwich(engine->y, nullptr, nullptr, A, -1, engine->G);
                rho_obs = engine->G;
            }
            engine->observables(popb ? popb + size_t(printed) * nb : nullptr, popw ? popw + size_t(printed) * nb : nullptr,
                                vel ? vel + 6 * size_t(printed) : nullptr, 0, rho_obs);
            ++printed;
        }
        // advance to the next printed step in one batch
        int batch = 1;
        double tt = t + dt;
        while (it + batch < n && !PrintObservables(tt) && !(h5 && PrintSparse(tt))) {
            ++batch;
            tt += dt;
        }
        desolver.Propagate(batch);
        it += batch;
    }
    return printed;
}

}  // namespace host
}  // namespace edus

using edus::EngineError;
using edus::host::Simulation;

struct edus_sim {
    Simulation* s;
    edus_ctx ctx;
};

template <class F>
static int guarded_sim(F&& fn)
{
    try {
        fn();
        return EDUS_OK;
    } catch (EngineError& e) {
        edus::set_error(e.what());
        return e.code;
    } catch (std::exception& e) {
        edus::set_error(e.what());
        return EDUS_ERR_IO;
    }
}

// DM0.txt: Operator::print_Rdecay("DM", rwann) of the initial state
// (Operator.hpp:288-311, called at Simulation.cpp:190): for every R of the
// Gamma-centred FFT grid (grid order) and every (a, b), the distance
// |rwann_a + R - rwann_b| (bohr) and rho_ab(R) = (1/N) sum_k exp(-i k.R) rho_ab(k)
// (fftPair.cpp:89-107), printed as the reference's stream does: the distance
// with the stream's current precision, the complex value through
// ostream.cpp:3-8 (setw 25, 10 digits each).  The transform is a separable DFT
// on the host (set-up only).
static void write_dm0(const Simulation& S, const char* dir)
{
    const int nb = S.nb, N0 = S.N0, N1 = S.N1, Nk = N0 * N1, M = nb * nb;
    const double kPi = 3.14159265358979323846;
    std::vector<edus::host::cplx> tmp(size_t(Nk) * M), rR(size_t(Nk) * M);
    // along axis 1, then axis 0, sign -1
    edus::host::parallel_for(N0, [&](int i0) {
        for (int j1 = 0; j1 < N1; ++j1)
            for (int h = 0; h < M; ++h) {
                edus::host::cplx a = 0.;
                for (int i1 = 0; i1 < N1; ++i1) {
                    const double ph = -2. * kPi * double((long(i1) * j1) % N1) / N1;
                    a += std::polar(1.0, ph) * S.rho0[size_t(i0 * N1 + i1) * M + h];
                }
                tmp[size_t(i0 * N1 + j1) * M + h] = a;
            }
    });
    edus::host::parallel_for(N1, [&](int j1) {
        for (int j0 = 0; j0 < N0; ++j0)
            for (int h = 0; h < M; ++h) {
                edus::host::cplx a = 0.;
                for (int i0 = 0; i0 < N0; ++i0) {
                    const double ph = -2. * kPi * double((long(i0) * j0) % N0) / N0;
                    a += std::polar(1.0, ph) * tmp[size_t(i0 * N1 + j1) * M + h];
                }
                rR[size_t(j0 * N1 + j1) * M + h] = a / double(Nk);
            }
    });
    std::vector<std::array<double, 3>> c(nb);
    for (int n = 0; n < nb; ++n) c[n] = S.material.wannier_centre(n);
    auto fold = [](double x, double lo, double up, double period) {
        while (x < lo && std::abs(x - lo) > 1.e-08) x += period;
        while (x > up || std::abs(x - up) < 1.e-08) x -= period;
        return x;
    };
    std::ofstream f(std::string(dir) + "/DM0.txt");
    if (!f) throw EngineError(EDUS_ERR_IO, std::string("cannot write ") + dir + "/DM0.txt");
    for (int j0 = 0; j0 < N0; ++j0)
        for (int j1 = 0; j1 < N1; ++j1) {
            const double r0 = fold(j0, -N0 / 2., N0 / 2., N0), r1 = fold(j1, -N1 / 2., N1 / 2., N1);
            double Rc[3];
            for (int i = 0; i < 3; ++i) Rc[i] = S.material.A[i][0] * r0 + S.material.A[i][1] * r1;
            for (int a = 0; a < nb; ++a)
                for (int b = 0; b < nb; ++b) {
                    double d2 = 0.;
                    for (int i = 0; i < 3; ++i) {
                        const double x = c[a][i] + Rc[i] - c[b][i];
                        d2 += x * x;
                    }
                    const edus::host::cplx v = rR[size_t(j0 * N1 + j1) * M + a * nb + b];
                    f << std::sqrt(d2) << std::setw(25) << std::setprecision(10) << v.real() << std::setw(25)
                      << std::setprecision(10) << v.imag() << std::endl;
                }
        }
}

// ScreenedPotential::Calculate (src/Screened_potential/Screened_potential.cpp:139-199)
// on a set-up simulation: the application derives from Simulation and uses
// its eigenvectors, band energies, grid and Wannier centres.  All results are
// R-space fields [R][a][b] over the FFT grid (grid order, Gamma-centred R):
//   X       ResponseFunction (:17-58): the Lindhard sum on the device, then k -> R
//   V       BareCoulomb (:79-104): V_ab(R) = 1/|r_a + R - r_b|, 0 below 1e-5 bohr
//   Eps     Epsilon (:122-132): 1 - V_ab(R) X_ab(R), element by element
//   InvEps  (:149-157): 1/Eps where |Eps| >= 1e-3, else 0
//   W       (:168-174): InvEps_ab(R) V_ab(R)
struct ScreenedFields {
    std::vector<edus::host::cplx> X, V, Eps, InvEps, W;
    std::vector<edus::host::cplx> Xk;  // the response in k, before go_to_R
};

static ScreenedFields screened_potential(Simulation& S)
{
    using edus::host::cplx;
    edus::Engine& e = *S.engine;
    const int nb = S.nb, N0 = S.N0, N1 = S.N1, Nk = S.Nk, M = nb * nb;
    const edus::LindhardPlan plan = edus::lindhard_plan(nb, S.filled, N0, N1);
    ScreenedFields out;
    // ResponseFunction on the device, in the engine's per-step scratch arrays
    // (FULL layout: nb^2 planes each): U into G, the energies into D, the
    // band-split copy of U into f, X(k) into aux; then k -> R into D.
    if (e.layout != edus::kFull || e.B != 1)
        throw EngineError(EDUS_ERR_UNSUPPORTED, "the screened potential needs a simulation in the FULL layout");
    if (edus::lindhard_work_bytes(plan) > size_t(e.npl) * Nk * sizeof(double2))
        throw EngineError(EDUS_ERR_STATE, "screened potential: scratch arrays too small");
    EDUS_HIP_CHECK(hipMemcpyAsync(e.G, S.U.data(), size_t(Nk) * M * sizeof(double2), hipMemcpyHostToDevice, e.s));
    EDUS_HIP_CHECK(hipMemcpyAsync(e.D, S.eps.data(), size_t(Nk) * nb * sizeof(double), hipMemcpyHostToDevice, e.s));
    edus::launch_lindhard(plan, e.G, reinterpret_cast<const double*>(e.D), e.f, e.aux, e.s);
    out.Xk.resize(size_t(Nk) * M);
    e.download(e.aux, reinterpret_cast<double*>(out.Xk.data()));  // synchronises the stream
    e.transform_to_R(e.aux, e.D);
    out.X.resize(size_t(Nk) * M);
    e.download(e.D, reinterpret_cast<double*>(out.X.data()));

    std::vector<std::array<double, 3>> c(nb);
    for (int n = 0; n < nb; ++n) c[n] = S.material.wannier_centre(n);
    out.V.assign(size_t(Nk) * M, 0.);
    out.Eps.resize(out.V.size());
    out.InvEps.resize(out.V.size());
    out.W.resize(out.V.size());
    for (int j0 = 0; j0 < N0; ++j0)
        for (int j1 = 0; j1 < N1; ++j1) {
            const double r0 = edus::host::fold(j0, -N0 / 2., N0 / 2., N0), r1 = edus::host::fold(j1, -N1 / 2., N1 / 2., N1);
            double Rc[3];
            for (int i = 0; i < 3; ++i) Rc[i] = S.material.A[i][0] * r0 + S.material.A[i][1] * r1;
            for (int a = 0; a < nb; ++a)
                for (int b = 0; b < nb; ++b) {
                    const size_t o = size_t(j0 * N1 + j1) * M + a * nb + b;
                    double d2 = 0.;
                    for (int i = 0; i < 3; ++i) {
                        const double x = c[a][i] + Rc[i] - c[b][i];
                        d2 += x * x;
                    }
                    const double d = std::sqrt(d2);
                    out.V[o] = d < 1.e-05 ? 0. : 1. / d;
                    out.Eps[o] = -out.V[o] * out.X[o] + 1.;
                    out.InvEps[o] = std::abs(out.Eps[o]) < 1.e-03 ? cplx(0.) : 1. / out.Eps[o];
                    out.W[o] = out.InvEps[o] * out.V[o];
                }
        }
    return out;
}

// Operator::go_to_k (sign +1, no scaling) of an R-space field [R][a][b], on
// the device through the engine's FULL-layout scratch arrays
static std::vector<edus::host::cplx> field_to_k(Simulation& S, const std::vector<edus::host::cplx>& fR)
{
    edus::Engine& e = *S.engine;
    std::vector<edus::host::cplx> fk(fR.size());
    e.upload(reinterpret_cast<const double*>(fR.data()), e.aux);
    e.transform_to_k(e.aux, e.D);
    e.download(e.D, reinterpret_cast<double*>(fk.data()));
    return fk;
}

// BlockMatrix::write_h5(name, node, label) (BlockMatrix_definitions.hpp:345-380),
// one process: group node/label/0 holding "local", the interleaved doubles
static void write_blockmatrix(edus::host::H5Out& f, const std::string& path, const std::vector<edus::host::cplx>& a)
{
    f.group(path);
    f.group(path + "/0");
    f.write_doubles(path + "/0/local", reinterpret_cast<const double*>(a.data()), 2 * a.size());
}

// Operator::print_Rdecay(name, rwann) (Operator.hpp:288-311): file name + "0.txt"
// (one process), per R and (a, b) the distance |r_a + R - r_b| and the value,
// both through the stream's defaults (6 significant digits, "(re,im)").
static void write_rdecay(const Simulation& S, const std::vector<edus::host::cplx>& f, const std::string& path)
{
    const int nb = S.nb, N0 = S.N0, N1 = S.N1, M = nb * nb;
    std::ofstream os(path);
    if (!os) throw EngineError(EDUS_ERR_IO, "cannot write " + path);
    std::vector<std::array<double, 3>> c(nb);
    for (int n = 0; n < nb; ++n) c[n] = S.material.wannier_centre(n);
    for (int j0 = 0; j0 < N0; ++j0)
        for (int j1 = 0; j1 < N1; ++j1) {
            const double r0 = edus::host::fold(j0, -N0 / 2., N0 / 2., N0), r1 = edus::host::fold(j1, -N1 / 2., N1 / 2., N1);
            double Rc[3];
            for (int i = 0; i < 3; ++i) Rc[i] = S.material.A[i][0] * r0 + S.material.A[i][1] * r1;
            for (int a = 0; a < nb; ++a)
                for (int b = 0; b < nb; ++b) {
                    double d2 = 0.;
                    for (int i = 0; i < 3; ++i) {
                        const double x = c[a][i] + Rc[i] - c[b][i];
                        d2 += x * x;
                    }
                    os << std::sqrt(d2);
                    os << f[size_t(j0 * N1 + j1) * M + a * nb + b] << std::endl;
                }
        }
}

extern "C" {

edus_sim* edus_sim_create(const edus_sim_params* p)
{
    if (!p) {
        edus::set_error("null parameters");
        return nullptr;
    }
    edus_sim* out = nullptr;
    int rc = guarded_sim([&] {
        auto* s = new Simulation(*p);
        out = new edus_sim{s, edus_ctx{s->engine.get()}};
    });
    return rc == EDUS_OK ? out : nullptr;
}

int edus_sim_destroy(edus_sim* sim)
{
    if (!sim) return EDUS_OK;
    int rc = guarded_sim([&] { delete sim->s; });
    delete sim;
    return rc;
}

int edus_sim_info(edus_sim* sim, int* nb, int* N0, int* N1, int* nsteps, double* dt)
{
    if (!sim) { edus::set_error("null simulation"); return EDUS_ERR_ARG; }
    if (nb) *nb = sim->s->nb;
    if (N0) *N0 = sim->s->N0;
    if (N1) *N1 = sim->s->N1;
    if (nsteps) *nsteps = sim->s->nsteps;
    if (dt) *dt = sim->s->dt;
    return EDUS_OK;
}

edus_ctx* edus_sim_engine(edus_sim* sim) { return sim ? &sim->ctx : nullptr; }

int edus_sim_field(edus_sim* sim, double t, double E[3])
{
    if (!sim || !E) { edus::set_error("null argument"); return EDUS_ERR_ARG; }
    sim->s->setoflaser.field(t, E);
    return EDUS_OK;
}

double edus_sim_time(edus_sim* sim) { return sim ? sim->s->desolver.get_CurrentTime() : 0.0; }

int edus_sim_propagate(edus_sim* sim, int nsteps, double* times, double* laser, double* pop_bloch,
                       double* pop_wannier, double* vel, int* nprinted)
{
    if (!sim) { edus::set_error("null simulation"); return EDUS_ERR_ARG; }
    return guarded_sim([&] {
        const int n = nsteps < 0 ? sim->s->nsteps : nsteps;
        const int k = sim->s->propagate(n, times, laser, pop_bloch, pop_wannier, vel);
        if (nprinted) *nprinted = k;
        sim->s->engine->enforce_monitor_limits();
    });
}

int edus_sim_write_outputs(edus_sim* sim, const char* dir, int nsteps, int* nprinted)
{
    if (!sim || !dir) { edus::set_error("null argument"); return EDUS_ERR_ARG; }
    return guarded_sim([&] {
        Simulation& S = *sim->s;
        const int n = nsteps < 0 ? S.nsteps : nsteps;
        const int nb = S.nb;
        // at most one printed row per step
        std::vector<double> t(n + 1), las(3 * size_t(n + 1)), lasA(3 * size_t(n + 1)), pb(size_t(n + 1) * nb),
            pw(size_t(n + 1) * nb), v(6 * size_t(n + 1));
        // DM0.txt (print_Rdecay of the initial state, Simulation.cpp:190) before stepping
        write_dm0(S, dir);
        if (S.par.write_h5) S.h5_open(std::string(dir) + "/output.h5");
        struct CloseH5 {
            Simulation& s;
            ~CloseH5() { s.h5.reset(); }
        } close_h5{S};
        const int k = S.propagate(n, t.data(), las.data(), pb.data(), pw.data(), v.data(), lasA.data());
        if (nprinted) *nprinted = k;
        const std::string d(dir);
        auto open = [&](const char* name) {
            std::ofstream f(d + "/" + name);
            if (!f) throw EngineError(EDUS_ERR_IO, "cannot write " + d + "/" + name);
            return f;
        };
        // formats of Simulation.cpp:420-429 (Time: default stream, Laser: the
        // Cartesian vector's setw 24 / 15 digits), :459-467 (populations: setw
        // 30, 14 digits, blank after each band) and :614-620 (velocity: setw
        // 20, 8 digits)
        std::ofstream ft = open("Time.txt"), fl = open("Laser.txt"), fa = open("Laser_A.txt"),
                      fp = open("Population.txt"), fw = open("Population_wannier.txt"), fv = open("Velocity.txt");
        for (int r = 0; r < k; ++r) {
            ft << t[r] << std::endl;
            for (int c = 0; c < 3; ++c) fl << std::setw(24) << std::setprecision(15) << las[3 * size_t(r) + c];
            fl << std::endl;
            for (int c = 0; c < 3; ++c) fa << std::setw(24) << std::setprecision(15) << lasA[3 * size_t(r) + c];
            fa << std::endl;
            for (int b = 0; b < nb; ++b) fp << std::setw(30) << std::setprecision(14) << pb[size_t(r) * nb + b] << " ";
            fp << std::endl;
            for (int b = 0; b < nb; ++b) fw << std::setw(30) << std::setprecision(14) << pw[size_t(r) * nb + b] << " ";
            fw << std::endl;
            for (int c = 0; c < 6; ++c) fv << std::setw(20) << std::setprecision(8) << v[6 * size_t(r) + c];
            fv << std::endl;
        }
    });
}

int edus_sim_screened(edus_sim* sim, double* X, double* V, double* Eps, double* InvEps, double* W)
{
    if (!sim) { edus::set_error("null simulation"); return EDUS_ERR_ARG; }
    return guarded_sim([&] {
        const ScreenedFields f = screened_potential(*sim->s);
        auto put = [](const std::vector<edus::host::cplx>& v, double* dst) {
            if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(edus::host::cplx));
        };
        put(f.X, X);
        put(f.V, V);
        put(f.Eps, Eps);
        put(f.InvEps, InvEps);
        put(f.W, W);
    });
}

int edus_sim_write_screened(edus_sim* sim, const char* dir)
{
    if (!sim || !dir) { edus::set_error("null argument"); return EDUS_ERR_ARG; }
    return guarded_sim([&] {
        const Simulation& S = *sim->s;
        const ScreenedFields f = screened_potential(*sim->s);
        const std::string d(dir);
        // the order of ScreenedPotential::Calculate; the reference's
        // inverse-permittivity print goes to "InverseEpsilon0.txt"
        write_rdecay(S, f.X, d + "/X0.txt");
        write_rdecay(S, f.V, d + "/V0.txt");
        write_rdecay(S, f.Eps, d + "/Epsilon0.txt");
        write_rdecay(S, f.InvEps, d + "/InverseEpsilon0.txt");
        write_rdecay(S, f.W, d + "/W0.txt");
        if (S.par.write_h5) {
            // the HDF5 build's nodes under "-1" in output.h5, added to an
            // existing file (Screened_potential.cpp:60-70,104-109,137-142,192-196):
            // Xk, Uk, Vk, Epsilonk, Wk.  The inverse permittivity is written
            // under "Epsilonk" again there (:174), which HDF5 refuses (the
            // node exists), so the file keeps Epsilonk = eps(k).
            edus::host::H5Out h5(d + "/output.h5", true);
            h5.group("/-1");
            write_blockmatrix(h5, "/-1/Xk", f.Xk);
            write_blockmatrix(h5, "/-1/Uk", sim->s->U);
            write_blockmatrix(h5, "/-1/Vk", field_to_k(*sim->s, f.V));
            write_blockmatrix(h5, "/-1/Epsilonk", field_to_k(*sim->s, f.Eps));
            write_blockmatrix(h5, "/-1/Wk", field_to_k(*sim->s, f.W));
        }
    });
}

int edus_sim_steps(edus_sim* sim, int nsteps)
{
    if (!sim) { edus::set_error("null simulation"); return EDUS_ERR_ARG; }
    return guarded_sim([&] {
        sim->s->desolver.Propagate(nsteps);
        sim->s->engine->enforce_monitor_limits();
    });
}

long edus_sim_get(edus_sim* sim, const char* name, double* out)
{
    if (!sim || !name) return -1;
    Simulation& s = *sim->s;
    const std::string n = name;
    const std::vector<edus::host::cplx>* v = nullptr;
    if (n == "H0k") v = &s.H0k;
    else if (n == "rk0") v = &s.rk[0];
    else if (n == "rk1") v = &s.rk[1];
    else if (n == "rk2") v = &s.rk[2];
    else if (n == "U") v = &s.U;
    else if (n == "rho0") v = &s.rho0;
    else if (n == "W") v = &s.W;
    else if (n == "hartree") v = &s.hartree;
    else if (n == "eps") {
        if (out) std::memcpy(out, s.eps.data(), s.eps.size() * sizeof(double));
        return long(s.eps.size());
    } else if (n == "A") {  // lattice, A[i][j] = component i of lattice vector j (bohr)
        if (out) std::memcpy(out, s.material.A, 9 * sizeof(double));
        return 9;
    } else if (n == "Vel0" || n == "Vel1" || n == "Vel2") {
        const int d = n[3] - '0';
        const long cnt = long(s.Nk) * s.nb * s.nb;
        if (out) {
            int rc = guarded_sim([&] { s.engine->download(s.engine->vel[d], out); });
            if (rc) return -1;
        }
        return cnt;
    }
    if (!v) return -1;
    if (out) std::memcpy(out, v->data(), v->size() * sizeof(edus::host::cplx));
    return long(v->size());
}

}  // extern "C"
