/*
 * edus_gpu.h -- C ABI of the MI355X EDUS propagation engine.
 *
 * The engine advances the one-body density matrix rho(k) of the reference
 * (gcistaro/EDUS_2.0) under
 *     d rho(k)/dt = -i [H0(k) + E(t).xi(k) + H_ee(k), rho(k)] + E(t).grad_k rho(k)
 * on one GPU, for the whole k grid.  It replaces the reference's hot path:
 *   - DESolver<Operator<complex<double>>>::Propagate_RK / Propagate_AB
 *     (src/DESolver/DESolver.hpp:155-220) and its SumWithProduct sweeps
 *     (DESolver.hpp:10-20),
 *   - the SourceTerm right-hand side (src/Simulation/Functional_SourceTerm.hpp:11-57)
 *     with everything it calls: Operator::go_to_R/go_to_k
 *     (src/Operator/Operator.hpp:566-606), FourierTransform::fft
 *     (src/fftPair/fftPair.cpp:89-107), kGradient::Calculate
 *     (src/kGradient/kGradient_definitions.hpp:7-41), Calculate_TDHamiltonian
 *     (src/Simulation/Simulation.cpp:296-349), Coulomb::EffectiveHamiltonian
 *     (src/Coulomb/Coulomb.cpp:134-180) and commutator/multiply
 *     (src/Operator/BlockMatrix_definitions.hpp:149-227),
 *   - the observable reductions Print_Population / Print_Velocity / TraceK
 *     (src/Simulation/Simulation.cpp:439-470,583-622; Simulation.hpp:125-142).
 *
 * Conventions at this boundary
 *   - Operator fields are passed in the reference's BlockMatrix layout
 *     [Nk][nb][nb], row-major, complex<double> as interleaved (re, im)
 *     doubles (src/Operator/BlockMatrix.hpp:48), k index i1 + N1*i0
 *     (src/MeshGrid/MeshGrid.cpp:117-127).  Pointers are host pointers.
 *   - All quantities in Hartree atomic units.
 *   - Every function returns 0 on success or a non-zero EDUS_ERR_* code;
 *     edus_last_error() gives a message (thread-local).  No C++ exception
 *     crosses this boundary.
 *   - One host thread drives a context; calls are stream-ordered on the
 *     context's own HIP stream; contexts are independent.
 */
#ifndef EDUS_GPU_H
#define EDUS_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define EDUS_OK 0
#define EDUS_ERR_ARG 1
#define EDUS_ERR_HIP 2
#define EDUS_ERR_STATE 3
#define EDUS_ERR_UNSUPPORTED 4
#define EDUS_ERR_MONITOR 5
#define EDUS_ERR_IO 6

#define EDUS_PREC_FP64 0  /* everything FP64 (parity path) */
#define EDUS_PREC_MIXED 1 /* FP32 MFMA commutator, FP64 state and updates */

/* Storage of operator fields on the device.
 * FULL: all nb*nb elements per k, the reference's arithmetic element for
 *   element (it propagates every element, BlockMatrix.hpp:48).  On even grids
 *   the reference's Gamma-centred Nyquist row (R = -N/2 has no +N/2 partner,
 *   MeshGrid.cpp:458-486) gives i(R.E)rho(R) a small anti-Hermitian part;
 *   FULL reproduces it.
 * HERMITIAN: upper triangle only (nb(nb+1)/2 planes), the lower triangle being
 *   the conjugate -- 1.9x less memory traffic at nb = 14.  Identical to FULL
 *   on odd grids; on even grids it drops the reference's anti-Hermitian
 *   Nyquist component (edus_check reports its size in FULL runs). */
#define EDUS_LAYOUT_FULL 0
#define EDUS_LAYOUT_HERMITIAN 1

typedef struct edus_ctx edus_ctx;
typedef struct edus_sim edus_sim;

/* Everything the propagation needs, set up on the host in FP64.
 * Replaces the state the reference's Simulation constructor builds
 * (src/Simulation/Simulation.cpp:10-233). */
typedef struct {
    int nb;                /* bands (Wannier functions), 1..16 */
    int N0, N1;            /* k grid; the third dimension must be 1 */
    int precision;         /* EDUS_PREC_* */
    int layout;            /* EDUS_LAYOUT_* */
    const double* H0k;     /* H0(k)   [Nk][nb][nb] complex (Hermitian) */
    const double* xik[3];  /* r_d(k)  [Nk][nb][nb] complex, NULL = zero */
    double A[3][3];        /* A[i][j] = Cartesian component i of lattice vector j (bohr) */
    int coulomb;           /* 1: Hartree + screened exchange (HSEX) on */
    const double* W;       /* screened potential W(R) on the FFT R grid [Nk][nb][nb] complex */
    const double* hartree; /* nb x nb complex: sum_R V(R), symmetrised */
    const double* rho0k;   /* rho(t=0) in k [Nk][nb][nb] complex: HSEX reference state (Coulomb::set_DM0) */
    const double* U;       /* eigenvectors per k, columns [Nk][nb][nb] complex; NULL = no Bloch populations */
    const double* vel[3];  /* velocity operator V_d(k) [Nk][nb][nb]; NULL = no current */
    int filledbands;
    int peierls;           /* 1: Peierls gauge (Functional_SourceTerm.hpp:20-44); step with the *_peierls calls */
    int gradient_space;    /* 0: "R" spectral derivative (default); 1: "k" finite differences over the
                              k-mesh shells (kGradient.cpp:25-210, kGradient_definitions.hpp:19-31) */
} edus_desc;

/* ---- context ----------------------------------------------------------- */
edus_ctx* edus_create(const edus_desc* desc);
/* A new context with a device-side copy of src's set-up arrays and state (no
 * counterpart in the reference; used to start many runs from one set-up). */
edus_ctx* edus_clone(edus_ctx* src);
int edus_destroy(edus_ctx* ctx);
const char* edus_last_error(void);

/* State transfer (Operator::get_Operator_k() data of the reference). */
int edus_set_state(edus_ctx* ctx, const double* rho_k);
int edus_get_state(edus_ctx* ctx, double* rho_k);

/* Gauge rotation of an operator field per k (Operator::go_to_bloch /
 * go_to_wannier, src/Operator/Operator.hpp:528-564): out = U^dagger in U
 * (to_bloch = 1) or U in U^dagger (to_bloch = 0) with the context's
 * eigenvectors U (columns, as the descriptor's U).  in/out host arrays
 * [Nk][nb][nb] (in may equal out); a HERMITIAN-layout context takes the
 * Hermitian part of in.
 * FP64 contexts rotate on the FP64 matrix cores, mixed-precision ones on the
 * FP32 matrix cores (FP32 operands and accumulation). */
int edus_rotate(edus_ctx* ctx, int to_bloch, const double* in, double* out);
/* The current state in the Bloch gauge, U^dagger rho U, [Nk][nb][nb]: the
 * reference's DMk_bloch output (Simulation.cpp:398-403 with Operator.hpp:528-564). */
int edus_get_state_bloch(edus_ctx* ctx, double* rho_bloch);
/* The same into a device buffer of Nk*nb*nb complex doubles ([Nk][nb][nb]),
 * queued on the context's stream (edus_synchronize waits for it): for callers
 * that keep the output on the GPU.  No counterpart in the reference. */
int edus_get_state_bloch_device(edus_ctx* ctx, double* rho_bloch_dev);

/* Fourier transform of an operator field between the context's k grid and
 * its R grid (Operator::go_to_R / go_to_k, src/Operator/Operator.hpp:566-606,
 * through FourierTransform::fft, src/fftPair/fftPair.cpp:89-107):
 *   sign -1:  out(R) = 1/Nk sum_k in(k) e^{-2 pi i (k0 R0 / N0 + k1 R1 / N1)}
 *   sign +1:  out(k) = sum_R in(R) e^{+2 pi i (k0 R0 / N0 + k1 R1 / N1)}
 * in/out host arrays [N0 N1][nb][nb] in FFT index order (index i of an axis
 * of length N is R = i for 2 i < N, i - N otherwise; MeshGrid.cpp:458-486);
 * in may equal out.  FULL-layout contexts with a 2-D grid; the state is left
 * alone. */
int edus_fourier(edus_ctx* ctx, int sign, const double* in, double* out);
/* nb, grid and layout of a context (any pointer may be NULL).  No
 * counterpart in the reference (its Operator carries its MeshGrid). */
int edus_ctx_info(edus_ctx* ctx, int* nb, int* N0, int* N1, int* layout);

/* One right-hand-side evaluation f = SourceTerm(t, rho) with the field E
 * given (the reference evaluates SetOfLaser at t).  in/out host arrays. */
int edus_rhs(edus_ctx* ctx, const double E[3], const double* in, double* out);

/* nsteps RK4 steps of size dt on the device-resident state
 * (DESolver.hpp:155-187).  E[(s*3 + j)*3 + d]: field component d at
 * t_s, t_s + dt/2, t_s + dt for j = 0, 1, 2. */
int edus_rk4_steps(edus_ctx* ctx, double dt, int nsteps, const double* E);

/* nsteps Adams-Bashforth steps of order 4 or 5 (DESolver.hpp:202-220), history
 * starting at zero as in the reference.  E[s*3 + d] at t_s. */
int edus_ab_steps(edus_ctx* ctx, double dt, int order, int nsteps, const double* E);
/* Peierls gauge (contexts created with peierls = 1), the SourceTerm branch of
 * Functional_SourceTerm.hpp:20-44: no E.grad_k term; H(R) = H0(R) + E.r(R) +
 * H_ee[rho(R) exp(-i A.R)] multiplied by exp(+i A.R) before the commutator.
 * A is the vector potential at the same times as E (the reference's
 * SetOfLaser::VectorPotential, Laser.hpp:152-158); observables read
 * rho(R) exp(-i A.R) (Simulation.cpp:447-449,593-595). */
int edus_rhs_peierls(edus_ctx* ctx, const double E[3], const double A[3], const double* in, double* out);
int edus_rk4_steps_peierls(edus_ctx* ctx, double dt, int nsteps, const double* E, const double* A);
int edus_ab_steps_peierls(edus_ctx* ctx, double dt, int order, int nsteps, const double* E, const double* A);
int edus_observables_peierls(edus_ctx* ctx, const double A[3], double* pop_bloch, double* pop_wannier,
                             double* vel);

/* Clear the Adams-Bashforth history (zero, as DESolver.cpp:26-29). */
int edus_ab_reset(edus_ctx* ctx);
/* Start-up of Adams-Bashforth by RK4 instead of the reference's zero history:
 * order-1 RK4 steps (E as in edus_rk4_steps) whose first-stage right-hand
 * sides become the history; edus_ab_steps continues from there.  No
 * counterpart in the reference (it starts AB from zero history). */
int edus_ab_seed(edus_ctx* ctx, double dt, int order, const double* E);

/* Observables of the current state (Simulation.cpp:439-470, 583-622):
 * pop_bloch[nb] (filled bands as 1 - P_n), pop_wannier[nb], vel[6] =
 * (Re vx, Im vx, Re vy, Im vy, Re vz, Im vz).  Any pointer may be NULL. */
int edus_observables(edus_ctx* ctx, double* pop_bloch, double* pop_wannier, double* vel);

/* Invariant monitors accumulated since the last call (and reset by it): the
 * maximum, over every state the step calls since then started from or wrote
 * and over k, of |Tr rho(k) - filledbands| and of the Hermiticity defect
 * max_{n<=m} |rho_nm - conj(rho_mn)| (HERMITIAN layout: twice the largest
 * |Im rho_nn|, the only part its storage leaves free).  The step kernels check
 * the state each step starts from; the state the last step wrote is checked
 * when the monitors are read, or before edus_set_state replaces it. */
int edus_check(edus_ctx* ctx, double* trace_dev, double* herm_dev);
/* Limits on those monitors (no counterpart in the reference, which checks
 * nothing while it propagates).  With a limit > 0 set, edus_rk4_steps,
 * edus_ab_steps, their Peierls versions and edus_sim_steps /
 * edus_sim_propagate wait for their steps and return EDUS_ERR_MONITOR when
 * the monitors (the maxima since the last edus_check, which they leave in
 * place) exceed a limit; a NaN or inf in any state then always fails.  Both
 * 0 (the default): not checked, and the step calls stay asynchronous. */
int edus_monitor_limits(edus_ctx* ctx, double trace_limit, double herm_limit);

int edus_synchronize(edus_ctx* ctx);

/* steps_per_graph > 0: edus_rk4_steps (and edus_sim stepping) replays hipGraphs
 * of that many RK4 steps instead of launching every kernel; 0 (default):
 * direct launches.  Same results: the captured kernels read the field from a
 * device table.  Per-kernel timing (below) sees direct launches only.  No
 * counterpart in the reference. */
int edus_graph(edus_ctx* ctx, int steps_per_graph);

/* Kernel timing with HIP events on the context stream.  When enabled, every
 * launch of the named kernel class is bracketed by events; edus_timing_get
 * returns the summed milliseconds and launch count since the last reset.
 * Classes: "grad_col", "grad_row", "grad_fused", "kgrad", "comm", "hf_row",
 * "hf_col", "peierls_col", "peierls_row", "rotate". */
int edus_timing_enable(edus_ctx* ctx, int on);
int edus_timing_get(edus_ctx* ctx, const char* kernel_class, double* total_ms, long* count);

/* Kernel and tile choices.  No counterpart in the reference (its FFTW plans
 * and MKL calls have no tiles); edus_create fills them by timing candidates on
 * the context's own arrays.  A caller may read them, and force others (for
 * instance to check that every choice gives the same results):
 *   fft_family[a]  1: two-step register FFT, 0: LDS Stockham; fixed by the
 *                  length of grid axis a (not settable)
 *   fft_cfg[a]     Stockham configuration 0..4, or two-step variant 0..3 (bit 0:
 *                  addend loaded early, bit 1: 512-thread work-groups)
 *   fft_lines[a]   lines per work-group
 *   comm_kind      0: FP64 FMAs, 1: FP64 MFMA (EDUS_PREC_FP64);
 *                  2: FP32 MFMA (EDUS_PREC_MIXED, not settable)
 *   comm_tile      commutator work-group: 0: 8 k-points, 256 threads;
 *                  1: 4 k-points, 256 threads; 2: 8 k-points, 512 threads
 *                  (1 and 2: MFMA engines only; the FMA engine takes 0)
 *   comm_tile_ab   the same for the Adams-Bashforth steps
 *   grad_fused     1: the spectral derivative in one pass per plane (one
 *                  work-group holds a plane: grids 120 x 120, 96 x 96,
 *                  90 x 90, 48 x 48); 0: two axis passes through a work plane
 *                  set.  Used when the field has components along both axes.
 *   tri_fields     1: the MFMA commutator engines read H0(k) and r_d(k) from
 *                  copies of their upper triangles, the lower elements as
 *                  conjugates (0.54 of the bytes at nb = 14); 0: from every
 *                  element (as the FMA engine always does).
 *                  Available (and the default) in the FULL layout without
 *                  exchange or the Peierls gauge when H0(k) and every r_d(k)
 *                  are Hermitian to 16 ulp of their largest element; for
 *                  exactly Hermitian fields both give identical results.
 *   hf_fused       1 (default): with exchange, the forward row transform of
 *                  the exchange term comes out of the gradient's row pass
 *                  (same loads; two-step row kernels, two-pass gradient,
 *                  field along grid axis 1); 0: a pass of its own.  The two
 *                  agree to rounding (the compiler contracts multiply-adds
 *                  per kernel).
 * Within one family, FFT tiles give bit-identical results, and so do the
 * commutator tiles; the two FP64 commutator engines differ in summation order
 * only, and the two gradient forms in rounding only (the compiler contracts
 * multiply-adds per kernel).
 * edus_set_tuning returns EDUS_ERR_ARG for a choice that does not fit. */
typedef struct {
    int fft_family[2];
    int fft_cfg[2];
    int fft_lines[2];
    int comm_kind;
    int comm_tile;
    int comm_tile_ab;
    int grad_fused;
    int tri_fields;
    int hf_fused;
} edus_tuning;
int edus_get_tuning(edus_ctx* ctx, edus_tuning* out);

/* The finite-difference stencil that gradient_space "k" uses for a lattice A
 * (A[i][j]: component i of lattice vector j, bohr) and an N0 x N1 k mesh: rows
 * (shell, m0, m1, b_x, b_y, b_z, w_shell), one per mesh vector b of the shells
 * in use, in the reference's order (kGradient.cpp:44-210).  Host only (no GPU
 * needed).  Returns the number of rows, or a negative EDUS_ERR_* code; rows
 * are written only if cap is large enough. */
int edus_kgrad_stencil(const double A[3][3], int N0, int N1, double* rows, int cap);
int edus_set_tuning(edus_ctx* ctx, const edus_tuning* in);

/* ---- batches ------------------------------------------------------------
 * Many contexts advanced together.  No counterpart in the reference (one
 * simulation per process).  Members with the same nb, grid, layout,
 * precision, exchange flag, filledbands, gradient space (and k-space stencil),
 * commutator engine and gradient form are stacked into one group whose
 * kernels advance all of them per launch; groups run on their own streams.
 * The batch copies each member's set-up arrays and state at creation;
 * members must outlive the batch (store/load copy states to and from them). */
typedef struct edus_batch edus_batch;
edus_batch* edus_batch_create(int n, edus_ctx* const* members);
int edus_batch_destroy(edus_batch* batch);
/* number of groups; member_group[n] (may be NULL) the group of each member */
int edus_batch_groups(edus_batch* batch, int* ngroups, int* member_group);
/* steps_per_graph > 0: capture that many RK4 steps of every group in one
 * hipGraph and replay it; 0: direct launches (default) */
int edus_batch_graph(edus_batch* batch, int steps_per_graph);
/* nsteps RK4 steps of every member (DESolver.hpp:155-187), member i's field
 * E[((s*3 + j)*n + i)*3 + d] at t_s, t_s + dt/2, t_s + dt (j = 0, 1, 2) */
int edus_batch_rk4_steps(edus_batch* batch, double dt, int nsteps, const double* E);
/* state of member i: batch -> member context (store) / member -> batch (load) */
int edus_batch_store(edus_batch* batch, int i);
int edus_batch_load(edus_batch* batch, int i);
/* observables of member i's state in the batch (as edus_observables) */
int edus_batch_observables(edus_batch* batch, int i, double* pop_bloch, double* pop_wannier, double* vel);
/* invariant monitors over all members (as edus_check) */
int edus_batch_check(edus_batch* batch, double* trace_dev, double* herm_dev);
int edus_batch_synchronize(edus_batch* batch);

/* ---- per-k MLP (learned coarse propagator) -------------------------------
 * No counterpart in the reference.  Three dense layers din -> dh -> dh -> dout,
 * tanh after the first two; BF16 weights and activations, FP32 accumulation
 * on the matrix cores, FP32 master weights.  Parameters are one FP32 block:
 * W1 [din][dh], b1 [dh], W2 [dh][dh], b2 [dh], W3 [dh][dout], b3 [dout].
 * Inputs are rounded to BF16 (nearest even) on entry. */
typedef struct edus_mlp edus_mlp;
edus_mlp* edus_mlp_create(int din, int dh, int dout, unsigned long long seed);
int edus_mlp_destroy(edus_mlp* mlp);
int edus_mlp_num_params(edus_mlp* mlp, long* n);
int edus_mlp_set_params(edus_mlp* mlp, const float* params);
int edus_mlp_get_params(edus_mlp* mlp, float* params);
/* Y [rows][dout] = net(X [rows][din]) */
int edus_mlp_forward(edus_mlp* mlp, int rows, const float* X, float* Y);
/* gradient of L = mean over rows and outputs of (net(X) - T)^2 w.r.t. the
 * parameters (same layout), and L */
int edus_mlp_gradients(edus_mlp* mlp, int rows, const float* X, const float* T, float* grads, float* loss);
/* Adam (beta 0.9, 0.999, eps 1e-8) over shuffled mini-batches of `batch`
 * rows for `epochs` passes; loss_per_epoch[epochs] (may be NULL) */
int edus_mlp_train(edus_mlp* mlp, int rows, const float* X, const float* T, int epochs, int batch, float lr,
                   unsigned long long seed, float* loss_per_epoch);

/* ---- learned parareal ------------------------------------------------------
 * No counterpart in the reference.  The run of nslices * fine_steps RK4 steps
 * of dt from a context's state is cut into nslices slices; F = RK4 over a
 * slice (fine_steps steps), G = RK4 with coarse_steps steps, plus the per-k
 * network's correction when one is set (evaluated in the first, coarse sweep
 * and reused unchanged by later iterations).  Parareal iterations
 *     U_{n+1} <- G(U_n new) + F(U_n old) - G(U_n old)
 * run every slice's F as one stacked batch; they stop when the largest change
 * of a slice boundary is <= tol, and after max_iter iterations without that
 * the run is redone as sequential RK4 (fallback = 1). */
typedef struct {
    int nslices, fine_steps, coarse_steps;
    double dt;        /* fine step; the coarse step is dt * fine_steps / coarse_steps */
    double tol;       /* max |change of U_n| over slices and elements */
    int max_iter;     /* parareal iterations before the sequential fallback */
    int graph_steps;  /* > 0: hipGraph replay of the fine batch (the coarse RK4 replays its whole slice) */
} edus_parareal_params;
typedef struct {
    int iterations, converged, fallback;
    double last_update;  /* max change of the last iteration */
    double seconds, fine_seconds, coarse_seconds;
    double updates[16];  /* max change of iterations 1 .. min(iterations, 16), in order */
} edus_parareal_result;
typedef struct edus_parareal edus_parareal;
/* clones `base` (its set-up; the run starts from its state unless rho0 is given) */
edus_parareal* edus_parareal_create(edus_ctx* base, const edus_parareal_params* params);
int edus_parareal_destroy(edus_parareal* pr);
/* fields: E_fine[((n*fine_steps + s)*3 + j)*3 + d] at t_s, t_s + dt/2, t_s + dt of
 * slice n; E_coarse likewise with coarse_steps; E_feat[(n*field_samples + i)*2 + c]
 * the x, y field samples of slice n given to the network */
int edus_parareal_set_fields(edus_parareal* pr, const double* E_fine, const double* E_coarse, const float* E_feat,
                             int field_samples);
/* the network (din = 2 nb^2 + 2 field_samples + 4, dout = nb^2) with its input
 * shift/scale mu, sig [din] and output scale/shift tsig, tmu [nb^2]; NULL:
 * plain coarse RK4 */
int edus_parareal_set_network(edus_parareal* pr, edus_mlp* net, const float* mu, const float* sig, const float* tmu,
                              const float* tsig);
int edus_parareal_run(edus_parareal* pr, const double* rho0, edus_parareal_result* result);
/* slice boundary n = 0 .. nslices of the last run */
int edus_parareal_get_state(edus_parareal* pr, int n, double* rho);
/* training data for the network from `runs` runs with their own fields (run r:
 * E_fine[r][...], E_coarse[r][...], E_feat[r][...] as above), started from
 * base's state: for k_samples random k-points per run and slice, the features
 * X[row][din] and targets T[row][nb^2] = param(F(U_n) - RK4_coarse(U_n)), rows
 * ordered [run][slice][sample] */
int edus_parareal_collect(edus_ctx* base, const edus_parareal_params* params, int runs, const double* E_fine,
                          const double* E_coarse, const float* E_feat, int field_samples, int k_samples,
                          unsigned long long seed, float* X, float* T);

/* ---- reference-format simulation (host set-up mirrored in C++) --------- */
typedef struct {
    double intensity; char intensity_units[32];
    double frequency; char frequency_units[32];
    double wavelength; char wavelength_units[32];
    double polarization[3];
    double cycles;
    double t0; char t0_units[32];
    double phase;
} edus_laser_params;

/* Keys of the reference's JSON input (src/InputVariables/input_schema.json);
 * unit strings are resolved with the reference's unit() (src/ConvertUnits.cpp:22-37). */
typedef struct {
    char tb_file[1024];     /* seedname; "_tb.dat" is appended */
    int grid[3];
    int filledbands;
    double dt; char dt_units[32];
    double initialtime; char initialtime_units[32];
    double finaltime; char finaltime_units[32];
    int printresolution, printresolution_pulse;
    char solver[8]; int order;
    int coulomb; double epsilon;
    int nr0; double r0[3]; char r0_units[32];
    int coulomb_center_quirk; /* 1: keep ModelCoulomb.cpp:76-79 accumulation */
    int nlasers;
    edus_laser_params lasers[10];
    int precision;            /* EDUS_PREC_* */
    int layout;               /* EDUS_LAYOUT_* */
    int ab_seed;              /* 1: start Adams-Bashforth with RK4 steps (edus_ab_seed), 0: zero history */
    int peierls;              /* the reference's "peierls" key */
    int gradient_space;       /* the reference's "gradient_space" key: 0 "R", 1 "k" */
    int write_h5;             /* 1: edus_sim_write_outputs also writes output.h5 (the HDF5 1.10 C
                                 library is loaded at run time: $EDUS_HDF5_LIB or libhdf5.so.103) */
    int toprint_dmk_wannier;  /* the reference's "toprint": {"DMk_wannier": "true"} */
    int toprint_dmk_bloch;    /* the reference's "toprint": {"DMk_bloch": "true"} */
} edus_sim_params;

edus_sim* edus_sim_create(const edus_sim_params* p);
int edus_sim_destroy(edus_sim* sim);
int edus_sim_info(edus_sim* sim, int* nb, int* N0, int* N1, int* nsteps, double* dt);
edus_ctx* edus_sim_engine(edus_sim* sim);
/* Field of the set of lasers at t (SetOfLaser::operator(), Laser.hpp:128-170). */
int edus_sim_field(edus_sim* sim, double t, double E[3]);
double edus_sim_time(edus_sim* sim);
/* Simulation::Propagate (Simulation.cpp:352-432): nsteps steps (-1: the
 * reference's int((tf-t0)/dt)+2), observables recorded before each step on
 * the print stride.  Output arrays hold one row per printed step:
 * times[n], laser[n][3], pop_bloch[n][nb], pop_wannier[n][nb], vel[n][6]. */
int edus_sim_propagate(edus_sim* sim, int nsteps, double* times, double* laser, double* pop_bloch,
                       double* pop_wannier, double* vel, int* nprinted);
/* Advance nsteps without observables (benchmark path). */
int edus_sim_steps(edus_sim* sim, int nsteps);
/* Propagate nsteps (< 0: the whole run) and write the reference's text
 * outputs into directory dir: DM0.txt (the initial state in R,
 * Simulation.cpp:190 with Operator.hpp:288-311) and Time.txt, Laser.txt,
 * Laser_A.txt, Population.txt, Population_wannier.txt, Velocity.txt, one row
 * per printed step in the reference's formats (Simulation.cpp:420-429,
 * 459-467,614-620).  With write_h5, also output.h5 as the reference's HDF5
 * build writes it (Simulation.cpp:215-229,385-420, GlobalFunctions.cpp:19-66):
 * the /json group, and at every sparse print step /time_au/<i> and, if asked
 * for by toprint, /DensityMatrix_k/<i>/0/local and
 * /DensityMatrix_k_bloch/<i>/0/local (interleaved complex [Nk][nb][nb]). */
int edus_sim_write_outputs(edus_sim* sim, const char* dir, int nsteps, int* nprinted);
/* Set-up arrays for inspection: "H0k", "rk0", "rk1", "rk2", "U", "Vel0",
 * "Vel1", "Vel2", "W", "rho0" (initial state in k), "hartree" ([nb][nb]),
 * "eps" (real [Nk][nb]), "A" (9 reals, the edus_desc lattice A[i][j]).
 * Returns the element count (complex elements; reals for "eps" and "A"). */
long edus_sim_get(edus_sim* sim, const char* name, double* out);

/* ---- screened-potential application (src/Screened_potential/) ---- */
/* ScreenedPotential::ResponseFunction (Screened_potential.cpp:17-58) on the
 * device: the Lindhard response in the point-like approximation,
 *   X_ab(q) = 1/Nk sum_k sum_{v < nf <= c} U_k(a,v) U_k+q(b,c) conj(U_k+q(a,c))
 *             conj(U_k(b,v)) / (e_v(k) - e_c(k+q)),
 * for the eigenvectors U ([Nk][nb][nb], U(a,n) at a*nb+n, bands ascending)
 * and band energies eps ([Nk][nb]) on an N0 x N1 grid (k order i1 + N1 i0);
 * 2 <= nb <= 16.  Host arrays; X out [Nk][nb][nb] (q, a, b), interleaved
 * complex.  Results agree with the reference's loop to rounding (another
 * summation order). */
int edus_lindhard(int nb, int N0, int N1, int filledbands, const double* U, const double* eps, double* X);
/* The same on device arrays in the stream `stream` (a hipStream_t, 0: the
 * null stream); work: device scratch of Nk*nb*nb complex doubles (the
 * band-split copy of U); X out as planes [nb*nb][Nk] (plane a*nb+b).
 * Allocates nothing. */
int edus_lindhard_device(int nb, int N0, int N1, int filledbands, const double* U, const double* eps, double* work,
                         double* X, void* stream);
/* edus_lindhard_device without the workspace argument: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc): it calls no
 * allocator.  The band-split copy of U lives in a fixed 48 MiB array that the
 * library holds on every GPU it runs on until the process ends, so Nk * nb^2
 * may be at most 3,145,728 (120 x 120 k at nb = 14); larger shapes return
 * EDUS_ERR_ARG.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (two streams, two threads) race and must be serialised by the
 * caller.  Results equal edus_lindhard_device's. */
int edus_lindhard_device_noalloc(int nb, int N0, int N1, int filledbands, const double* U, const double* eps,
                                 double* X, void* stream);
/* ScreenedPotential::Calculate (Screened_potential.cpp:139-199) on a set-up
 * simulation (the application derives from Simulation): X(R) (the response
 * transformed k -> R, Operator::go_to_R), the bare Coulomb V(R)
 * = 1/|r_a + R - r_b| (0 below 1e-5 bohr), Eps(R) = 1 - V X, its elementwise
 * inverse (0 where |Eps| < 1e-3) and W(R) = Eps^-1 V; each [Nk][nb][nb] over
 * the FFT R grid in grid order; null outputs are skipped.  The simulation must
 * use the FULL layout. */
int edus_sim_screened(edus_sim* sim, double* X, double* V, double* Eps, double* InvEps, double* W);
/* The same, written as the reference's Operator::print_Rdecay files
 * (Operator.hpp:288-311) X0.txt, V0.txt, Epsilon0.txt, InverseEpsilon0.txt and
 * W0.txt in directory dir; with write_h5 also the nodes -1/Xk, -1/Uk, -1/Vk,
 * -1/Epsilonk and -1/Wk (each <node>/0/local, interleaved complex) added to
 * dir/output.h5 (Screened_potential.cpp:60-70,104-109,137-142,192-196). */
int edus_sim_write_screened(edus_sim* sim, const char* dir);

#ifdef __cplusplus
}
#endif
#endif /* EDUS_GPU_H */
