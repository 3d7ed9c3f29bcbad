// C entry points of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.
// Loaded by tests/ (ctypes), by the smoke check and by bench.py's CPU-baseline
// leg; never by the product library.
#include <chrono>
#include <memory>
#include <cstring>
#include <exception>
#include <string>

#include "oracle.hpp"

using namespace oracle;

extern "C" {

// Same field meaning as the reference JSON keys (input_schema.json); unit
// strings are passed through and resolved with the reference's unit() rules.
struct oracle_laser {
    double intensity; char intensity_units[32];
    double frequency; char frequency_units[32];
    double wavelength; char wavelength_units[32];
    double polarization[3];
    double cycles;
    double t0; char t0_units[32];
    double phase;
};
struct oracle_params {
    char tb_file[1024];
    int grid[3];
    int filledbands;
    double dt; char dt_units[32];
    double initialtime; char initialtime_units[32];
    double finaltime; char finaltime_units[32];
    int printresolution, printresolution_pulse;
    char solver[8]; int order;
    int coulomb; double epsilon;
    int nr0; double r0[3]; char r0_units[32];
    int coulomb_center_quirk;
    int nlasers;
    oracle_laser lasers[10];
    int peierls;
    int gradient_k;  // 1: gradient_space "k"
};

static thread_local std::string g_err;
const char* oracle_last_error() { return g_err.c_str(); }

void* oracle_create(const oracle_params* p)
{
    try {
        SimParams s;
        s.tb_file = p->tb_file;
        for (int d = 0; d < 3; ++d) s.grid[d] = p->grid[d];
        s.filledbands = p->filledbands;
        s.dt = p->dt; s.dt_units = p->dt_units;
        s.initialtime = p->initialtime; s.initialtime_units = p->initialtime_units;
        s.finaltime = p->finaltime; s.finaltime_units = p->finaltime_units;
        s.printresolution = p->printresolution; s.printresolution_pulse = p->printresolution_pulse;
        s.solver = p->solver; s.order = p->order;
        s.coulomb = p->coulomb != 0; s.epsilon = p->epsilon;
        s.r0.assign(p->r0, p->r0 + (p->nr0 > 0 ? p->nr0 : 1));
        s.r0_units = p->r0_units;
        s.coulomb_center_quirk = p->coulomb_center_quirk != 0;
        s.peierls = p->peierls != 0;
        s.gradient_space = p->gradient_k ? "k" : "R";
        for (int i = 0; i < p->nlasers; ++i) {
            const oracle_laser& l = p->lasers[i];
            LaserParams lp;
            lp.intensity = l.intensity; lp.intensity_units = l.intensity_units;
            lp.frequency = l.frequency; lp.frequency_units = l.frequency_units;
            lp.wavelength = l.wavelength; lp.wavelength_units = l.wavelength_units;
            for (int d = 0; d < 3; ++d) lp.polarization[d] = l.polarization[d];
            lp.cycles = l.cycles; lp.t0 = l.t0; lp.t0_units = l.t0_units; lp.phase = l.phase;
            s.lasers.push_back(lp);
        }
        return new Simulation(s);
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void oracle_destroy(void* h) { delete static_cast<Simulation*>(h); }

void oracle_dims(void* h, int* nb, int* n0, int* n1, int* nsteps, double* dt)
{
    auto* s = static_cast<Simulation*>(h);
    *nb = s->nb; *n0 = s->N[0]; *n1 = s->N[1]; *nsteps = s->nsteps; *dt = s->dt;
}

double oracle_time(void* h) { return static_cast<Simulation*>(h)->t_current; }

static std::vector<cplx>* named(Simulation* s, const char* name)
{
    std::string n = name;
    if (n == "rho") return &s->rho;
    if (n == "H0k") return &s->H0k;
    if (n == "H0R") return &s->H0R;
    if (n == "rk0") return &s->rk[0];
    if (n == "rk1") return &s->rk[1];
    if (n == "rk2") return &s->rk[2];
    if (n == "U") return &s->U;
    if (n == "Vel0") return &s->Vel[0];
    if (n == "Vel1") return &s->Vel[1];
    if (n == "Vel2") return &s->Vel[2];
    if (n == "rho0R") return &s->rho0R;
    if (n == "Wscr") return &s->Wscr;
    if (n == "Vbare") return &s->Vbare;
    if (n == "hartree") return &s->hartree;
    // Adams-Bashforth history slots aux[i] (DESolver.hpp:45; the step reads
    // f_{n-i} from aux[abindex[i]], abindex = identity before the first step)
    if (n.size() == 5 && n.compare(0, 4, "hist") == 0 && n[4] >= '0' && n[4] <= '4') {
        auto& v = s->aux[n[4] - '0'];
        if (v.empty()) v.assign(size_t(s->Nk) * s->M, 0.);
        return &v;
    }
    return nullptr;
}

// copy a named complex array out ([k][n][m] interleaved re/im); returns element count
long oracle_get(void* h, const char* name, double* out)
{
    auto* v = named(static_cast<Simulation*>(h), name);
    if (!v) return -1;
    if (out) std::memcpy(out, v->data(), v->size() * sizeof(cplx));
    return long(v->size());
}

int oracle_set(void* h, const char* name, const double* in)
{
    auto* v = named(static_cast<Simulation*>(h), name);
    if (!v) return -1;
    std::memcpy(v->data(), in, v->size() * sizeof(cplx));
    return 0;
}

void oracle_eigenvalues(void* h, double* out)
{
    auto* s = static_cast<Simulation*>(h);
    std::memcpy(out, s->eps.data(), s->eps.size() * sizeof(double));
}

void oracle_rhs(void* h, double t, const double* in, double* out)
{
    auto* s = static_cast<Simulation*>(h);
    std::vector<cplx> vin(size_t(s->Nk) * s->M), vout;
    std::memcpy(vin.data(), in, vin.size() * sizeof(cplx));
    s->rhs(vout, t, vin);
    std::memcpy(out, vout.data(), vout.size() * sizeof(cplx));
}

void oracle_rhs_E(void* h, const double* E, const double* in, double* out)
{
    auto* s = static_cast<Simulation*>(h);
    std::vector<cplx> vin(size_t(s->Nk) * s->M), vout;
    std::memcpy(vin.data(), in, vin.size() * sizeof(cplx));
    s->rhs_E(vout, E, vin);
    std::memcpy(out, vout.data(), vout.size() * sizeof(cplx));
}

void oracle_rhs_EA(void* h, const double* E, const double* Av, const double* in, double* out)
{
    auto* s = static_cast<Simulation*>(h);
    std::vector<cplx> vin(size_t(s->Nk) * s->M), vout;
    std::memcpy(vin.data(), in, vin.size() * sizeof(cplx));
    s->rhs_EA(vout, E, Av, vin);
    std::memcpy(out, vout.data(), vout.size() * sizeof(cplx));
}

void oracle_vector_potential(void* h, double t, double* out)
{
    static_cast<Simulation*>(h)->vector_potential(t, out);
}

void oracle_lattice(void* h, double* A)
{
    auto* s = static_cast<Simulation*>(h);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[3 * i + j] = s->A[i][j];
}

void oracle_step(void* h, int n)
{
    auto* s = static_cast<Simulation*>(h);
    for (int i = 0; i < n; ++i) s->step();
}

void oracle_E(void* h, double t, double* E) { static_cast<Simulation*>(h)->E_of_t(t, E); }

void oracle_observables(void* h, double* popb, double* popw, double* vel6)
{
    auto* s = static_cast<Simulation*>(h);
    if (s->par.peierls) {  // as at a print step of propagate
        double Av[3];
        s->vector_potential(s->t_current, Av);
        s->rho_R(s->rho_obs);
        s->peierls_phase(s->rho_obs, Av, -1);
        s->fft_planes_public(s->rho_obs, +1);
    }
    std::vector<double> pb, pw;
    s->populations(pb, pw);
    for (int n = 0; n < s->nb; ++n) { popb[n] = pb[n]; popw[n] = pw[n]; }
    s->velocity(vel6);
}

// Lindhard response for given eigenvectors and energies (free function)
int oracle_lindhard(int nb, int N0, int N1, int filled, const double* U, const double* eps, double* X)
{
    try {
        const std::vector<cplx> x = lindhard(nb, N0, N1, filled, reinterpret_cast<const cplx*>(U), eps);
        std::memcpy(X, x.data(), x.size() * sizeof(cplx));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Wannier centres [nb][3] (bohr)
void oracle_rwann(void* h, double* out)
{
    auto* s = static_cast<Simulation*>(h);
    for (int n = 0; n < s->nb; ++n)
        for (int d = 0; d < 3; ++d) out[3 * n + d] = s->rwann[n][d];
}

// ScreenedPotential::Calculate on this simulation; null outputs skipped
void oracle_screened(void* h, double* X, double* V, double* Eps, double* InvEps, double* W)
{
    auto* s = static_cast<Simulation*>(h);
    std::vector<cplx> f[5];
    s->screened_potential(f[0], f[1], f[2], f[3], f[4]);
    double* out[5] = {X, V, Eps, InvEps, W};
    for (int i = 0; i < 5; ++i)
        if (out[i]) std::memcpy(out[i], f[i].data(), f[i].size() * sizeof(cplx));
}

void oracle_rhoR(void* h, double* out)
{
    auto* s = static_cast<Simulation*>(h);
    std::vector<cplx> r;
    s->rho_R(r);
    std::memcpy(out, r.data(), r.size() * sizeof(cplx));
}

// Full propagation with observables printed before each step as
// Simulation::do_onestep does.  Arrays are sized by the caller for nsteps rows.
int oracle_propagate(void* h, int nsteps, double* times, double* laser, double* popb, double* popw, double* vel)
{
    auto* s = static_cast<Simulation*>(h);
    Observables o;
    s->propagate(&o, nsteps);
    int n = int(o.time.size());
    for (int i = 0; i < n; ++i) {
        times[i] = o.time[i];
        for (int d = 0; d < 3; ++d) laser[3 * i + d] = o.laser[i][d];
        for (int b = 0; b < s->nb; ++b) {
            popb[i * s->nb + b] = o.pop_bloch[i][b];
            popw[i * s->nb + b] = o.pop_wannier[i][b];
        }
        for (int c = 0; c < 6; ++c) vel[6 * i + c] = o.velocity[i][c];
    }
    s->last_laser_A.clear();
    for (auto& a : o.laser_A) s->last_laser_A.insert(s->last_laser_A.end(), a.begin(), a.end());
    return n;
}

// the vector potential printed (Laser_A.txt) by the last oracle_propagate
int oracle_laser_A(void* h, double* out)
{
    auto* s = static_cast<Simulation*>(h);
    std::copy(s->last_laser_A.begin(), s->last_laser_A.end(), out);
    return int(s->last_laser_A.size() / 3);
}

// stand-alone building blocks for the closed-form tests
void oracle_eigh(int n, const double* A, double* w, double* V)
{
    eigh(n, reinterpret_cast<const cplx*>(A), w, reinterpret_cast<cplx*>(V));
}

void oracle_fft2d(int n0, int n1, double* data, int sign)
{
    auto p0 = make_plan(n0), p1 = make_plan(n1);
    std::vector<cplx> scratch(2 * (n0 > n1 ? n0 : n1) + 2);
    fft2d(p0, p1, reinterpret_cast<cplx*>(data), sign, scratch.data());
}

double oracle_struve0(double x) { return struve0(x); }

// k-space gradient stencil: n rows of (shell, m0, m1, b_x, b_y, b_z, w_shell);
// returns the number of rows (0 when gradient_space is "R").  rows may be null.
int oracle_kstencil(void* h, double* rows)
{
    auto* sim = static_cast<Simulation*>(h);
    const int n = int(sim->kstencil.size());
    if (rows)
        for (int i = 0; i < n; ++i) {
            const auto& st = sim->kstencil[i];
            double* r = rows + 7 * i;
            r[0] = st.shell; r[1] = st.m0; r[2] = st.m1;
            r[3] = st.b[0]; r[4] = st.b[1]; r[5] = st.b[2];
            r[6] = sim->shell_weight[st.shell];
        }
    return n;
}

// out = finite-difference (u.grad_k) of in ([k][n][m]), gradient_space "k" only
int oracle_kgradient(void* h, const double* u, const double* in, double* out)
{
    auto* sim = static_cast<Simulation*>(h);
    if (!sim->kgrad) return -1;
    const size_t n = size_t(sim->Nk) * sim->M;
    std::vector<cplx> a(reinterpret_cast<const cplx*>(in), reinterpret_cast<const cplx*>(in) + n), o(n, 0.);
    sim->kgradient(o, a, u);
    std::copy(o.begin(), o.end(), reinterpret_cast<cplx*>(out));
    return 0;
}

// Gauge rotation with the oracle's eigenvectors (Operator.hpp:528-564):
// to_bloch: out = U^dagger in U, else U in U^dagger; [k][n][m] interleaved
void oracle_rotate(void* h, int to_bloch, const double* in, double* out)
{
    auto* s = static_cast<Simulation*>(h);
    const size_t n = size_t(s->Nk) * s->M;
    std::vector<cplx> a(reinterpret_cast<const cplx*>(in), reinterpret_cast<const cplx*>(in) + n), o;
    s->rotate(to_bloch != 0, a, o);
    std::copy(o.begin(), o.end(), reinterpret_cast<cplx*>(out));
}

// Host memory bandwidth in bytes/s (bytes read + written) of a parallel copy
// between two arrays of `bytes` each, the best of `reps` passes after one
// untimed pass: the bandwidth context for the CPU baseline's byte model.
double oracle_host_bandwidth(long bytes, int reps)
{
    const size_t n = size_t(bytes) / sizeof(double);
    std::unique_ptr<double[]> pa(new double[n]), pb(new double[n]);  // left uninitialised here
    double* a = pa.get();
    double* b = pb.get();
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) a[i] = double(i & 1023), b[i] = 0.0;  // first touch by the copying thread
    double best = 0.0;
    for (int r = 0; r <= reps; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < n; ++i) b[i] = a[i];
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (r > 0 && s > 0) best = std::max(best, 2.0 * double(n) * sizeof(double) / s);
        std::swap(a, b);
    }
    return best;
}

}  // extern "C"
