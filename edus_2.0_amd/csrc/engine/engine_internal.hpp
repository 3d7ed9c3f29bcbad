// Engine class behind the C ABI.  Host-side only.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/edus_gpu.h"
#include "kernels.hpp"

namespace edus {

struct EngineError : std::runtime_error {
    int code;
    EngineError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string& s);
extern thread_local std::string g_last_error;

#define EDUS_HIP_CHECK(x)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess)                                                                          \
            throw ::edus::EngineError(EDUS_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));  \
    } while (0)

struct TimedLaunch {
    const char* cls = nullptr;
    hipEvent_t a{}, b{};
};

// Options used by the in-library host set-up (not part of the C ABI).
struct EngineOptions {
    // > 0: build the velocity operator V_d = (-i[r_d, H0] + d_k_d H0) / det on
    // the device from H0 and r (Simulation.cpp:523-576) instead of taking it
    // from the descriptor.
    double velocity_det = 0.0;
    // >= 0: set the state to the initial condition U diag(n < initial_filled) U^dagger
    // (Functional_InitialCondition.hpp:12-32) on the device; with exchange and no
    // rho0k in the descriptor, it is also the HSEX reference state
    int initial_filled = -1;
};

class Batch;

class Engine {
  public:
    explicit Engine(const edus_desc& d, const EngineOptions& opt = EngineOptions());
    // A batch group: the members' set-up arrays and states stacked along the
    // plane axis ([B][npl][Nk] in every per-k array).  All members share nb,
    // grid, layout, precision, exchange on/off and filledbands.
    explicit Engine(const std::vector<const Engine*>& members);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    int nb, N0, N1, Nk, npl, layout, prec, nf;
    int B = 1;  // stacked instances
    bool coulomb;
    bool peierls = false;
    // gradient_space "k": finite-difference stencil (host copy and device table)
    bool kgrad = false;
    std::vector<KStencilPoint> kst;
    KStencilPoint* kst_dev = nullptr;
    // E.grad_k in by the stencil into D (E or, for tables, E_tab on the device)
    const double2* kgradient(const double2* in, const double E[3], const double* E_tab);
    double2 *Hp = nullptr, *Tq = nullptr;  // Peierls: phased H(k) and its axis-0 half
    // E and A that the phased H in Hp was formed with (hp_valid: it was)
    bool hp_valid = false;
    double hp_key[6] = {};
    double2* ptab[2] = {nullptr, nullptr};   // Peierls: per-axis phase tables (N0, N1 entries)
    double A[3][3];
    std::vector<std::array<double, 9>> Ainst;  // lattice of each stacked instance

    // Field tables of stacked instances, device-resident for `tab_steps`
    // steps: E [step][j][B][3] and gradient coefficients c_axis / N_axis
    // [step][j][B][2] at the stage times t, t + dt/2, t + dt (j = 0, 1, 2).
    double* E_dev = nullptr;
    double* coef_dev = nullptr;
    int tab_steps = 0;
    hipStream_t s = nullptr;

    double2 *y = nullptr, *acc = nullptr, *aux = nullptr, *G = nullptr, *D = nullptr, *f = nullptr;
    double2 *H0 = nullptr, *xi[3] = {nullptr, nullptr, nullptr}, *vel[3] = {nullptr, nullptr, nullptr};
    // mixed precision: H0 and xi rounded to FP32 for the FP32 commutator engine
    // (the FP64 copies stay for the set-up products); Peierls: the phased H
    float2 *H0f = nullptr, *xif[3] = {nullptr, nullptr, nullptr};
    bool uses_f32_fields() const;
    void alloc_f32_fields();
    void make_f32_fields();
    // FULL layout without exchange or the Peierls gauge, with H0(k) and every
    // r_d(k) Hermitian to a few ulp: their upper triangles (nb (nb + 1) / 2
    // planes per instance, FP64 and for mixed precision FP32), which the
    // commutator reads instead when `tri` is set (CommArgs::tri).  The FULL
    // arrays stay for the set-up products.
    bool tri_ok = false;
    bool tri = false;
    double2 *H0u = nullptr, *xiu[3] = {nullptr, nullptr, nullptr};
    float2 *H0uf = nullptr, *xiuf[3] = {nullptr, nullptr, nullptr};
    void make_tri_fields();
    // H0, xi (where use_xi[d]), their FP32 copies and the tri flag of the
    // arguments of a commutator launch on engine `kind`
    void field_args(CommArgs& c, const bool use_xi[3], int kind) const;
    // the dynamic-LDS attribute of every kernel the current choices may launch
    void prime_launch_attributes();
    double2 *U = nullptr, *T = nullptr, *stage = nullptr;
    double *W = nullptr, *hart = nullptr;
    double2 *hdiag = nullptr, *hsum = nullptr;  // hdiag: set-up only
    int force_kind = -1;  // >= 0: the autotuner keeps this commutator engine (batch groups)
    double *obs_partial = nullptr, *obs_out = nullptr;
    unsigned long long* mon = nullptr;
    double2* hist[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int abidx[5] = {0, 1, 2, 3, 4};
    double2* tw[2] = {nullptr, nullptr};
    FFTFactors fac[2];
    FFTChoice fft_choice[2];  // autotuned kernel and tile per grid axis
    bool grad_fused = false;  // spectral derivative in one pass per plane (grad_fused.hip)
    int force_grad = -1;      // >= 0: the autotuner keeps this gradient form (batch groups)
    int comm_kind = kCommScalar;     // engine of the RHS commutator (precision-dependent)
    int comm_kind64 = kCommScalar;   // FP64 engine, also used for set-up products
    int comm_tile = 0, comm_tile64 = 0;  // tile of each
    int comm_tile_ab = 0;                 // tile of the Adams-Bashforth kernel
    bool hf_fuse = true;  // exchange: its forward row pass taken from the gradient's row pass
    void read_monitor(double out[2]);
    // steps since the last check_final_state: the state they wrote is unchecked
    bool state_dirty = false;
    void check_final_state();
    // the monitors' maxima since the last read, left in place
    void peek_monitor(double out[2]);
    // > 0: step calls check the monitors against these limits (edus_monitor_limits)
    double mon_limit[2] = {0.0, 0.0};
    bool mon_limits_on() const { return mon_limit[0] > 0.0 || mon_limit[1] > 0.0; }
    // throws EDUS_ERR_MONITOR when a limit is crossed or a value is not finite
    void enforce_monitor_limits();
    int tile_for(const StageUpdate& up) const { return up.ab_order > 0 ? comm_tile_ab : comm_tile; }

    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::map<std::string, std::pair<double, long>> tsum;

    size_t inst_elems() const { return size_t(npl) * Nk; }                 // one instance's field
    size_t plane_bytes() const { return size_t(B) * npl * Nk * sizeof(double2); }  // all instances
    void upload(const double* host_full, double2* dst);       // [Nk][nb][nb] -> planes
    void download(const double2* src, double* host_full);     // planes -> [Nk][nb][nb]
    void download_bloch(const double2* src, double* host_full);  // U^dagger src U -> [Nk][nb][nb]
    void evaluate(const double2* in, const double E[3], const StageUpdate& up);
    // Peierls gauge: the same stage with field E and vector potential Av
    void evaluate_peierls(const double2* in, const double E[3], const double Av[3], const StageUpdate& up);
    void rk4_step_peierls(double dt, const double* E3x3, const double* A3x3);
    void ab_step_peierls(double dt, int order, const double* E3, const double* A3);
    // out = planes of the phase sandwich F^-1 exp(i sign A.R) F in (both axes)
    void peierls_sandwich(const double2* in, const double2* const add[3], const double addc[3], const double Av[3],
                          int sign, double2* out);
    // record_order > 0: stage 1's right-hand side f(t_n, y_n) is also stored as
    // the newest Adams-Bashforth history entry of that order (RK4 start-up)
    void rk4_step(double dt, const double* E3x3, int record_order = 0);
    // the stage updates of one RK4 step on y / acc / aux
    void rk4_stages(double dt, StageUpdate u[4]) const;
    // stacked instances: fields from the device tables at table step `ts`;
    // use_axis / use_xi switch off gradient axes and E.xi terms that are zero
    // for every instance of the chunk
    void ensure_tables(int steps);
    void evaluate_tab(const double2* in, int ts, int j, const bool use_axis[2], const bool use_xi[3],
                      const StageUpdate& up);
    void rk4_step_tab(double dt, int ts, const bool use_axis[2], const bool use_xi[3]);
    // nsteps RK4 steps, E[(s * 3 + j) * 3 + d]; with graph_steps > 0 the steps
    // are replayed from a hipGraph of graph_steps steps (self_batch)
    void rk4_steps(double dt, int nsteps, const double* E);
    int graph_steps = 0;
    Batch* self_batch = nullptr;
    void ab_step(double dt, int order, const double* E3);
    void observables(double* pop_bloch, double* pop_wannier, double* vel, int inst = 0,
                     const double2* rho_override = nullptr);
    // gauge rotation of planes `in` (this layout) of instance inst into out
    // (kFull / kHermitian planes or kRotKMajor); FP32 matrix cores in mixed
    // precision unless fp64
    void rotate(const double2* in, double2* out, int out_layout, bool to_bloch, int inst = 0, bool fp64 = false);
    void initial_condition(int filled);
    void collect_timing();
    // k -> R of every plane of a FULL-layout field, sign -1 and 1/Nk
    // (Operator::go_to_R, Operator.hpp:566-606, fftPair.cpp:89-107); uses G
    void transform_to_R(const double2* in, double2* out);
    // R -> k of every plane, sign +1 and no scaling (Operator::go_to_k); uses G
    void transform_to_k(const double2* in, double2* out);

  private:
    std::vector<void*> owned;
    void alloc(void** p, size_t bytes);
    void axis_fft(int axis, int mode, const double2* in, const double2* addend, double2* out, double coef,
                  double scale, const double* mult, const char* cls, const double* coef_tab = nullptr,
                  double2* out2 = nullptr, double2* r0 = nullptr);
    // D = E.grad rho with c[j] = c_j / N_j, or per instance from ctab [B][2]
    // fwd1 (may be null): also store the forward row transform of in there
    // from the row pass (mode 5) where it can; *fwd1_done says whether it did
    const double2* gradient(const double2* in, const double c[2], const double* ctab, bool use0, bool use1,
                            double2* fwd1 = nullptr, bool* fwd1_done = nullptr);
    void ensure_hist();
    // exchange term of in into T (+ Hartree diagonal); row_done: T already
    // holds the forward row transform of in (from the gradient's row pass)
    void fock(const double2* in, bool row_done = false);
    void build_velocity(double det);
    void autotune();
    CommArgs tuning_stage_args(int kind) const;
    AxisFFTArgs axis_args(int axis, const FFTChoice& c, int mode, const double2* in, const double2* addend,
                          double2* out, double coef, double scale, const double* mult) const;
    void tbegin(const char* cls, TimedLaunch& t);
    void tend(TimedLaunch& t);
};

void delete_batch(Batch* b);

// Runs fn, turning exceptions into the C ABI's status codes and message.
template <class F>
int guarded(F&& fn)
{
    try {
        fn();
        return EDUS_OK;
    } catch (EngineError& e) {
        set_error(e.what());
        return e.code;
    } catch (std::exception& e) {
        set_error(e.what());
        return EDUS_ERR_STATE;
    }
}

}  // namespace edus

#define EDUS_REQUIRE(cond)                         \
    do {                                           \
        if (!(cond)) {                             \
            edus::set_error("null argument");      \
            return EDUS_ERR_ARG;                   \
        }                                          \
    } while (0)

// The opaque handle of the C ABI.
struct edus_ctx {
    edus::Engine* e;
};
