// Learned parareal (edus_parareal_* in include/edus_gpu.h).  No counterpart in
// the reference.
//
// The run [0, N * fine_steps * dt] is cut into N slices.  F is the engine's
// RK4 over one slice (fine_steps steps of dt), G the same RK4 with
// coarse_steps steps, optionally corrected by the per-k MLP:
//     G(U) = R(U) + net(features(U, R(U) - U, E samples of the slice, k))
// with R the coarse RK4 (the network predicts F - R per k-point from local
// data; its weights are shared by all k).  Parareal iteration k:
//     U_{n+1}^k = G(U_n^k) + F(U_n^{k-1}) - G(U_n^{k-1})
// with all F solves of an iteration run as one stacked batch of N instances
// (one launch per kernel advances every slice) and the G sweep sequential.
// Iterations stop when max_n |U_n^k - U_n^{k-1}| <= tol; after max_iter
// iterations without that the run falls back to sequential RK4 from U_0.
//
// Hermitian parametrisation of an nb x nb matrix (nb^2 reals), used for the
// network's inputs and outputs: Re rho_nn (n < nb), then Re and Im of rho_nm
// for the pairs n < m in row order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "../engine/batch.hpp"
#include "../engine/engine_internal.hpp"
#include "mlp.hpp"

#define HIP_CHECK EDUS_HIP_CHECK

namespace edus {

namespace {

// the nb^2 reals of the Hermitian parametrisation of (a - b)(k) (b may be
// null), handed to put(index, value) in order
template <class Put>
__device__ inline void herm_param(const double2* a, const double2* b, int nb, int layout, long Nk, int k, int base,
                                  Put put)
{
    int o = base;
    for (int n = 0; n < nb; ++n) {
        double v = element(a, n, n, nb, layout, Nk, k).x;
        if (b) v -= element(b, n, n, nb, layout, Nk, k).x;
        put(o++, float(v));
    }
    for (int n = 0; n < nb; ++n)
        for (int m = n + 1; m < nb; ++m) {
            double2 v = element(a, n, m, nb, layout, Nk, k);
            if (b) {
                const double2 w = element(b, n, m, nb, layout, Nk, k);
                v.x -= w.x;
                v.y -= w.y;
            }
            put(o++, float(v.x));
            put(o++, float(v.y));
        }
}

struct FeatArgs {
    int nb, layout, N0, N1, ne, din;
    const double2* U;     // slice start state
    const double2* R;     // coarse RK4 result
    const double2* F;     // fine result (collection only)
    const float* E;       // [ne][2] field samples of the slice
    const float* mu;      // [din] feature shift and scale (inference)
    const float* sig;
    uint16_t* X;          // inference: [Nk][din] BF16, normalised
    const int* klist;     // collection: k-point of each row
    int rows;
    float* Xf;            // collection: [rows][din] raw features
    float* Tf;            // collection: [rows][nb^2] raw targets F - R
};

// features of one k-point: param(U), param(R - U), E samples, k on the torus
template <class Put>
__device__ inline void features_of(const FeatArgs& a, int k, Put put)
{
    const long Nk = long(a.N0) * a.N1;
    const int q = a.nb * a.nb;
    herm_param(a.U, nullptr, a.nb, a.layout, Nk, k, 0, put);
    herm_param(a.R, a.U, a.nb, a.layout, Nk, k, q, put);
    int o = 2 * q;
    for (int i = 0; i < 2 * a.ne; ++i) put(o++, a.E[i]);
    const int i0 = k / a.N1, i1 = k - i0 * a.N1;
    const double t0 = 2.0 * 3.14159265358979323846 * i0 / a.N0, t1 = 2.0 * 3.14159265358979323846 * i1 / a.N1;
    put(o++, float(cos(t0)));
    put(o++, float(sin(t0)));
    put(o++, float(cos(t1)));
    put(o++, float(sin(t1)));
}

constexpr int kMaxDin = 2 * 16 * 16 + 2 * 64 + 4;

__global__ void features_kernel(FeatArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.N0 * a.N1) return;
    uint16_t* x = a.X + long(k) * a.din;
    features_of(a, k, [&](int i, float v) { x[i] = f2bf((v - a.mu[i]) / a.sig[i]); });
}

__global__ void collect_kernel(FeatArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.rows) return;
    const int k = a.klist[r];
    float* x = a.Xf + long(r) * a.din;
    features_of(a, k, [&](int i, float v) { x[i] = v; });
    float* t = a.Tf + long(r) * a.nb * a.nb;
    herm_param(a.F, a.R, a.nb, a.layout, long(a.N0) * a.N1, k, 0, [&](int i, float v) { t[i] = v; });
}

// R += the network's correction at every k: y * tsig + tmu, unpacked from the
// Hermitian parametrisation (FULL layout: both triangles; HERMITIAN: upper)
__global__ void apply_kernel(double2* R, const float* __restrict__ Y, const float* __restrict__ tmu,
                             const float* __restrict__ tsig, int nb, int layout, long Nk)
{
    const long k = blockIdx.x * long(blockDim.x) + threadIdx.x;
    if (k >= Nk) return;
    const float* y = Y + k * nb * nb;
    int o = 0;
    for (int n = 0; n < nb; ++n) {
        const double v = double(y[o] * tsig[o] + tmu[o]);
        ++o;
        R[long(diag_plane(n, nb, layout)) * Nk + k].x += v;
    }
    for (int n = 0; n < nb; ++n)
        for (int m = n + 1; m < nb; ++m) {
            const double re = double(y[o] * tsig[o] + tmu[o]);
            const double im = double(y[o + 1] * tsig[o + 1] + tmu[o + 1]);
            o += 2;
            if (layout == kFull) {
                double2& u = R[long(n * nb + m) * Nk + k];
                u.x += re;
                u.y += im;
                double2& l = R[long(m * nb + n) * Nk + k];
                l.x += re;
                l.y -= im;
            } else {
                double2& u = R[long(tri_index(n, m, nb)) * Nk + k];
                u.x += re;
                u.y += im;
            }
        }
}

// U = G + F - Gp elementwise, and max |U_new - U_old| into mx (bit pattern)
__global__ void combine_kernel(double2* __restrict__ U, const double2* __restrict__ G, const double2* __restrict__ F,
                               const double2* __restrict__ Gp, long n, unsigned long long* mx)
{
    __shared__ double sh[256];
    double m = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += long(gridDim.x) * 256) {
        const double2 g = G[i], f = F[i], p = Gp[i], u = U[i];
        const double2 v = make_double2(g.x + f.x - p.x, g.y + f.y - p.y);
        const double d = fmax(fabs(v.x - u.x), fabs(v.y - u.y));
        m = (d != d) ? __builtin_huge_val() : fmax(m, d);
        U[i] = v;
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(mx, static_cast<unsigned long long>(__double_as_longlong(sh[0])));
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

class Parareal {
  public:
    Parareal(Engine* base, const edus_parareal_params& p);
    ~Parareal();

    Engine* base;
    edus_parareal_params prm;
    int N, q, din = 0, ne = 0;
    size_t ie;  // elements of one state
    std::unique_ptr<Batch> fine;
    std::unique_ptr<Engine> coarse;
    std::unique_ptr<Engine> seq;  // sequential fallback, made on first use
    std::vector<double> Efine;    // batch layout [s][j][n][3]
    std::vector<double> Ecoarse;  // [n][s][j][3]
    std::vector<double> Eseq;     // [n * fine_steps + s][j][3]
    std::vector<float> Efeat;     // [n][ne][2]
    Mlp* net = nullptr;           // borrowed
    std::vector<void*> owned;
    double2* U = nullptr;         // N + 1 slice boundaries
    double2* Gp = nullptr;        // N previous G results
    double2* Gt = nullptr;        // G result of the current slice
    float *mu = nullptr, *sig = nullptr, *tmu = nullptr, *tsig = nullptr, *Ed = nullptr;
    float* Ysweep = nullptr;      // the network's output of every slice in the first (coarse) sweep
    uint16_t* X = nullptr;
    unsigned long long* dmax = nullptr;
    double t_fine = 0.0, t_coarse = 0.0;

    void set_fields(const double* Ef, const double* Ec, const float* Ee, int ne_);
    void set_net(Mlp* m, const float* mu_, const float* sig_, const float* tmu_, const float* tsig_);
    void gop(const double2* in, int n, double2* out, bool first);
    void run(const double* rho0, edus_parareal_result* res);

  private:
    void* alloc(size_t bytes);
};

void* Parareal::alloc(size_t bytes)
{
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, bytes));
    owned.push_back(p);
    return p;
}

Parareal::Parareal(Engine* b, const edus_parareal_params& p) : base(b), prm(p)
{
    if (b->B != 1) throw EngineError(EDUS_ERR_ARG, "parareal needs a single context");
    if (p.nslices < 1 || p.fine_steps < 1 || p.coarse_steps < 1 || p.max_iter < 0 || !(p.dt > 0.0))
        throw EngineError(EDUS_ERR_ARG, "parareal: nslices, fine_steps, coarse_steps >= 1, max_iter >= 0, dt > 0");
    N = p.nslices;
    q = b->nb * b->nb;
    ie = b->inst_elems();
    HIP_CHECK(hipStreamSynchronize(b->s));
    fine = std::make_unique<Batch>(std::make_unique<Engine>(std::vector<const Engine*>(N, b)));
    fine->graph_steps = p.graph_steps;
    coarse = std::make_unique<Engine>(std::vector<const Engine*>{b});
    coarse->graph_steps = p.graph_steps > 0 ? p.coarse_steps : 0;
    U = static_cast<double2*>(alloc((N + 1) * ie * sizeof(double2)));
    Gp = static_cast<double2*>(alloc(N * ie * sizeof(double2)));
    Gt = static_cast<double2*>(alloc(ie * sizeof(double2)));
    dmax = static_cast<unsigned long long*>(alloc(sizeof(unsigned long long)));
}

Parareal::~Parareal()
{
    if (coarse) (void)hipStreamSynchronize(coarse->s);
    if (fine) fine->sync_all();
    for (void* p : owned) (void)hipFree(p);
}

void Parareal::set_fields(const double* Ef, const double* Ec, const float* Ee, int ne_)
{
    const int fs = prm.fine_steps, cs = prm.coarse_steps;
    Efine.assign(size_t(fs) * 3 * N * 3, 0.0);
    Eseq.assign(size_t(N) * fs * 9, 0.0);
    for (int n = 0; n < N; ++n)
        for (int s = 0; s < fs; ++s)
            for (int j = 0; j < 3; ++j)
                for (int d = 0; d < 3; ++d) {
                    const double v = Ef[((size_t(n) * fs + s) * 3 + j) * 3 + d];
                    Efine[((size_t(s) * 3 + j) * N + n) * 3 + d] = v;
                    Eseq[((size_t(n) * fs + s) * 3 + j) * 3 + d] = v;
                }
    Ecoarse.assign(Ec, Ec + size_t(N) * cs * 9);
    ne = ne_;
    if (Ee && ne > 0)
        Efeat.assign(Ee, Ee + size_t(N) * ne * 2);
    else
        Efeat.assign(size_t(N) * ne * 2, 0.f);
    din = 2 * q + 2 * ne + 4;
    if (din > kMaxDin) throw EngineError(EDUS_ERR_ARG, "too many field samples per slice");
    Ed = static_cast<float*>(alloc(std::max<size_t>(Efeat.size(), 1) * sizeof(float)));
    HIP_CHECK(hipMemcpyAsync(Ed, Efeat.data(), Efeat.size() * sizeof(float), hipMemcpyHostToDevice, coarse->s));
    HIP_CHECK(hipStreamSynchronize(coarse->s));
}

void Parareal::set_net(Mlp* m, const float* mu_, const float* sig_, const float* tmu_, const float* tsig_)
{
    net = m;
    if (!m) return;
    if (din == 0) throw EngineError(EDUS_ERR_STATE, "set the fields before the network");
    if (m->din != din || m->dout != q)
        throw EngineError(EDUS_ERR_ARG, "network sizes do not match: din must be 2 nb^2 + 2 ne + 4, dout nb^2");
    mu = static_cast<float*>(alloc(din * sizeof(float)));
    sig = static_cast<float*>(alloc(din * sizeof(float)));
    tmu = static_cast<float*>(alloc(q * sizeof(float)));
    tsig = static_cast<float*>(alloc(q * sizeof(float)));
    // on the stream the network's inputs are built on (see gop)
    const hipStream_t cs = coarse->s;
    HIP_CHECK(hipMemcpyAsync(mu, mu_, din * sizeof(float), hipMemcpyHostToDevice, cs));
    HIP_CHECK(hipMemcpyAsync(sig, sig_, din * sizeof(float), hipMemcpyHostToDevice, cs));
    HIP_CHECK(hipMemcpyAsync(tmu, tmu_, q * sizeof(float), hipMemcpyHostToDevice, cs));
    HIP_CHECK(hipMemcpyAsync(tsig, tsig_, q * sizeof(float), hipMemcpyHostToDevice, cs));
    HIP_CHECK(hipStreamSynchronize(cs));
    X = static_cast<uint16_t*>(alloc(size_t(base->Nk) * din * sizeof(uint16_t)));
    Ysweep = static_cast<float*>(alloc(size_t(N) * base->Nk * q * sizeof(float)));
}

// out = G(in) for slice n.  The network is evaluated in the first (coarse)
// sweep only, on that sweep's slice-start states, and its correction is
// reused unchanged by every later iteration: it then cancels exactly in
// G(U_new) - G(U_old), so the iterations see the coarse RK4's sensitivity
// and the network only moves the starting point.  (Re-evaluating it would
// add the jumps of its BF16-rounded inputs to that difference, at the scale
// of the tolerance.)
void Parareal::gop(const double2* in, int n, double2* out, bool first)
{
    const auto t0 = std::chrono::steady_clock::now();
    Engine& c = *coarse;
    const double dtc = prm.dt * prm.fine_steps / prm.coarse_steps;
    HIP_CHECK(hipMemcpyAsync(c.y, in, ie * sizeof(double2), hipMemcpyDeviceToDevice, c.s));
    HIP_CHECK(hipStreamSynchronize(c.s));
    c.rk4_steps(dtc, prm.coarse_steps, Ecoarse.data() + size_t(n) * prm.coarse_steps * 9);
    if (net) {
        float* Yn = Ysweep + size_t(n) * c.Nk * q;
        if (first) {
            FeatArgs a{};
            a.nb = c.nb;
            a.layout = c.layout;
            a.N0 = c.N0;
            a.N1 = c.N1;
            a.ne = ne;
            a.din = din;
            a.U = in;
            a.R = c.y;
            a.E = Ed + size_t(n) * ne * 2;
            a.mu = mu;
            a.sig = sig;
            a.X = X;
            hipLaunchKernelGGL(features_kernel, dim3((c.Nk + 127) / 128), dim3(128), 0, c.s, a);
            const hipStream_t own = net->s;
            net->s = c.s;  // keep the network's launches in order with the engine's
            net->forward(X, c.Nk, Yn, false);
            net->s = own;
        }
        hipLaunchKernelGGL(apply_kernel, dim3((c.Nk + 255) / 256), dim3(256), 0, c.s, c.y, Yn, tmu, tsig, c.nb,
                           c.layout, long(c.Nk));
    }
    HIP_CHECK(hipMemcpyAsync(out, c.y, ie * sizeof(double2), hipMemcpyDeviceToDevice, c.s));
    HIP_CHECK(hipStreamSynchronize(c.s));
    t_coarse += seconds_since(t0);
}

void Parareal::run(const double* rho0, edus_parareal_result* res)
{
    if (Efine.empty()) throw EngineError(EDUS_ERR_STATE, "parareal fields not set");
    const auto t_start = std::chrono::steady_clock::now();
    t_fine = t_coarse = 0.0;
    Engine& c = *coarse;
    // Every device-to-device copy below is stream-ordered: the engines run on
    // non-blocking streams, which do not wait for work on the null stream (a
    // plain hipMemcpy between device buffers may return before it lands).
    if (rho0) {
        c.upload(rho0, U);
    } else {
        HIP_CHECK(hipStreamSynchronize(base->s));
        HIP_CHECK(hipMemcpyAsync(U, base->y, ie * sizeof(double2), hipMemcpyDeviceToDevice, c.s));
    }
    // iteration 0: the coarse sweep (gop reads U_n on c.s, behind these copies)
    for (int n = 0; n < N; ++n) {
        gop(U + n * ie, n, Gp + n * ie, true);
        HIP_CHECK(hipMemcpyAsync(U + (n + 1) * ie, Gp + n * ie, ie * sizeof(double2), hipMemcpyDeviceToDevice, c.s));
    }
    HIP_CHECK(hipStreamSynchronize(c.s));
    int it = 0;
    bool conv = false;
    double diff = 0.0;
    double hist[16] = {};
    while (it < prm.max_iter && !conv) {
        ++it;
        // F(U_n) for every slice, one stacked batch
        const auto tf = std::chrono::steady_clock::now();
        for (int n = 0; n < N; ++n)
            HIP_CHECK(hipMemcpyAsync(fine->state(n), U + n * ie, ie * sizeof(double2), hipMemcpyDeviceToDevice,
                                     fine->bs));
        HIP_CHECK(hipStreamSynchronize(fine->bs));
        fine->rk4_steps(prm.dt, prm.fine_steps, Efine.data());
        t_fine += seconds_since(tf);
        // correction sweep; U_0 .. U_{it-1} are already exact
        HIP_CHECK(hipMemsetAsync(dmax, 0, sizeof(unsigned long long), c.s));
        for (int n = it - 1; n < N; ++n) {
            gop(U + n * ie, n, Gt, false);
            const long ne_ = long(ie);
            hipLaunchKernelGGL(combine_kernel, dim3(std::min<long>((ne_ + 255) / 256, 1024)), dim3(256), 0, c.s,
                               U + (n + 1) * ie, Gt, fine->state(n), Gp + n * ie, ne_, dmax);
            HIP_CHECK(hipMemcpyAsync(Gp + n * ie, Gt, ie * sizeof(double2), hipMemcpyDeviceToDevice, c.s));
        }
        unsigned long long m = 0;
        HIP_CHECK(hipMemcpyAsync(&m, dmax, sizeof(m), hipMemcpyDeviceToHost, c.s));
        HIP_CHECK(hipStreamSynchronize(c.s));
        std::memcpy(&diff, &m, sizeof(diff));
        if (it <= 16) hist[it - 1] = diff;
        conv = diff <= prm.tol;
    }
    bool fallback = false;
    if (!conv) {
        // sequential RK4 from U_0, slice by slice (the boundaries are kept)
        fallback = true;
        const auto tf = std::chrono::steady_clock::now();
        if (!seq) {
            seq = std::make_unique<Engine>(std::vector<const Engine*>{base});
            seq->graph_steps = prm.graph_steps;
        }
        HIP_CHECK(hipMemcpyAsync(seq->y, U, ie * sizeof(double2), hipMemcpyDeviceToDevice, seq->s));
        HIP_CHECK(hipStreamSynchronize(seq->s));
        for (int n = 0; n < N; ++n) {
            seq->rk4_steps(prm.dt, prm.fine_steps, Eseq.data() + size_t(n) * prm.fine_steps * 9);
            HIP_CHECK(hipStreamSynchronize(seq->s));
            // finished before the next slice's steps overwrite the state
            HIP_CHECK(hipMemcpyAsync(U + (n + 1) * ie, seq->y, ie * sizeof(double2), hipMemcpyDeviceToDevice,
                                     seq->s));
            HIP_CHECK(hipStreamSynchronize(seq->s));
        }
        t_fine += seconds_since(tf);
    }
    HIP_CHECK(hipDeviceSynchronize());
    HIP_CHECK(hipGetLastError());
    if (res) {
        res->iterations = it;
        res->converged = conv ? 1 : 0;
        res->fallback = fallback ? 1 : 0;
        res->last_update = diff;
        res->seconds = seconds_since(t_start);
        res->fine_seconds = t_fine;
        res->coarse_seconds = t_coarse;
        std::memcpy(res->updates, hist, sizeof(hist));
    }
}

// Training data: `runs` independent runs, each with its own fields, propagated
// slice by slice with F (the fine RK4) as one stacked batch; at every slice
// start U_n the coarse RK4 R(U_n) (also batched) and, after the slice, F(U_n).
// For `samples` k-points per run and slice: features(U_n, R, E, k) and the
// target F - R (raw, unnormalised), rows ordered [run][slice][sample].
void collect(Engine* base, const edus_parareal_params& p, int runs, const double* Ef, const double* Ec,
             const float* Ee, int ne, int samples, unsigned long long seed, float* Xout, float* Tout)
{
    if (runs < 1 || samples < 1) throw EngineError(EDUS_ERR_ARG, "collect: runs and samples must be >= 1");
    const int N = p.nslices, fs = p.fine_steps, cs = p.coarse_steps;
    const int q = base->nb * base->nb;
    const int din = 2 * q + 2 * ne + 4;
    if (din > kMaxDin) throw EngineError(EDUS_ERR_ARG, "too many field samples per slice");
    samples = std::min(samples, base->Nk);
    HIP_CHECK(hipStreamSynchronize(base->s));
    Batch fine(std::make_unique<Engine>(std::vector<const Engine*>(runs, base)));
    Batch crs(std::make_unique<Engine>(std::vector<const Engine*>(runs, base)));
    fine.graph_steps = p.graph_steps;
    crs.graph_steps = p.graph_steps > 0 ? cs : 0;
    const hipStream_t st = fine.bs;  // this function's own copies and kernels
    const size_t ie = base->inst_elems();
    std::vector<void*> bufs;
    auto dev = [&](size_t bytes) {
        void* ptr = nullptr;
        HIP_CHECK(hipMalloc(&ptr, std::max<size_t>(bytes, 1)));
        bufs.push_back(ptr);
        return ptr;
    };
    try {
        double2* Ustart = static_cast<double2*>(dev(size_t(runs) * ie * sizeof(double2)));
        float* Ed = static_cast<float*>(dev(size_t(runs) * N * ne * 2 * sizeof(float)));
        int* kl = static_cast<int*>(dev(size_t(samples) * sizeof(int)));
        float* Xd = static_cast<float*>(dev(size_t(samples) * din * sizeof(float)));
        float* Td = static_cast<float*>(dev(size_t(samples) * q * sizeof(float)));
        if (ne > 0) {
            HIP_CHECK(hipMemcpyAsync(Ed, Ee, size_t(runs) * N * ne * 2 * sizeof(float), hipMemcpyHostToDevice, st));
            HIP_CHECK(hipStreamSynchronize(st));
        }
        std::mt19937_64 rng(seed);
        std::vector<int> all(base->Nk);
        for (int k = 0; k < base->Nk; ++k) all[k] = k;
        std::vector<double> ef(size_t(fs) * 3 * runs * 3), ec(size_t(cs) * 3 * runs * 3);
        for (int n = 0; n < N; ++n) {
            // field tables of slice n for every run, batch layout [s][j][run][3]
            for (int r = 0; r < runs; ++r) {
                for (int s = 0; s < fs; ++s)
                    for (int j = 0; j < 3; ++j)
                        for (int d = 0; d < 3; ++d)
                            ef[((size_t(s) * 3 + j) * runs + r) * 3 + d] =
                                Ef[(((size_t(r) * N + n) * fs + s) * 3 + j) * 3 + d];
                for (int s = 0; s < cs; ++s)
                    for (int j = 0; j < 3; ++j)
                        for (int d = 0; d < 3; ++d)
                            ec[((size_t(s) * 3 + j) * runs + r) * 3 + d] =
                                Ec[(((size_t(r) * N + n) * cs + s) * 3 + j) * 3 + d];
            }
            // stream-ordered copies, finished before either batch steps on its own streams
            for (int r = 0; r < runs; ++r) {
                HIP_CHECK(hipMemcpyAsync(Ustart + r * ie, fine.state(r), ie * sizeof(double2), hipMemcpyDeviceToDevice,
                                         st));
                HIP_CHECK(hipMemcpyAsync(crs.state(r), fine.state(r), ie * sizeof(double2), hipMemcpyDeviceToDevice,
                                         st));
            }
            HIP_CHECK(hipStreamSynchronize(st));
            crs.rk4_steps(p.dt * fs / cs, cs, ec.data());
            fine.rk4_steps(p.dt, fs, ef.data());
            for (int r = 0; r < runs; ++r) {
                std::shuffle(all.begin(), all.end(), rng);
                HIP_CHECK(hipMemcpyAsync(kl, all.data(), samples * sizeof(int), hipMemcpyHostToDevice, st));
                FeatArgs a{};
                a.nb = base->nb;
                a.layout = base->layout;
                a.N0 = base->N0;
                a.N1 = base->N1;
                a.ne = ne;
                a.din = din;
                a.U = Ustart + r * ie;
                a.R = crs.state(r);
                a.F = fine.state(r);
                a.E = Ed + (size_t(r) * N + n) * ne * 2;
                a.klist = kl;
                a.rows = samples;
                a.Xf = Xd;
                a.Tf = Td;
                hipLaunchKernelGGL(collect_kernel, dim3((samples + 127) / 128), dim3(128), 0, st, a);
                const size_t row0 = (size_t(r) * N + n) * samples;
                HIP_CHECK(hipMemcpyAsync(Xout + row0 * din, Xd, size_t(samples) * din * sizeof(float),
                                         hipMemcpyDeviceToHost, st));
                HIP_CHECK(hipMemcpyAsync(Tout + row0 * q, Td, size_t(samples) * q * sizeof(float),
                                         hipMemcpyDeviceToHost, st));
                // the k list is reshuffled and the buffers refilled for the next run
                HIP_CHECK(hipStreamSynchronize(st));
            }
        }
        HIP_CHECK(hipDeviceSynchronize());
    } catch (...) {
        (void)hipDeviceSynchronize();
        for (void* ptr : bufs) (void)hipFree(ptr);
        throw;
    }
    for (void* ptr : bufs) (void)hipFree(ptr);
}

}  // namespace edus

struct edus_parareal {
    edus::Parareal* p;
};

using edus::guarded;

extern "C" {

edus_parareal* edus_parareal_create(edus_ctx* base, const edus_parareal_params* params)
{
    if (!base || !params) {
        edus::set_error("edus_parareal_create: null argument");
        return nullptr;
    }
    edus_parareal* out = nullptr;
    const int rc = guarded([&] { out = new edus_parareal{new edus::Parareal(base->e, *params)}; });
    return rc == EDUS_OK ? out : nullptr;
}

int edus_parareal_destroy(edus_parareal* pr)
{
    if (!pr) return EDUS_OK;
    const int rc = guarded([&] { delete pr->p; });
    delete pr;
    return rc;
}

int edus_parareal_set_fields(edus_parareal* pr, const double* E_fine, const double* E_coarse, const float* E_feat,
                             int field_samples)
{
    EDUS_REQUIRE(pr && E_fine && E_coarse && field_samples >= 0 && (field_samples == 0 || E_feat));
    return guarded([&] { pr->p->set_fields(E_fine, E_coarse, E_feat, field_samples); });
}

int edus_parareal_set_network(edus_parareal* pr, edus_mlp* net, const float* mu, const float* sig, const float* tmu,
                              const float* tsig)
{
    EDUS_REQUIRE(pr && (!net || (mu && sig && tmu && tsig)));
    return guarded([&] { pr->p->set_net(net ? net->h.m : nullptr, mu, sig, tmu, tsig); });
}

int edus_parareal_run(edus_parareal* pr, const double* rho0, edus_parareal_result* result)
{
    EDUS_REQUIRE(pr);
    return guarded([&] { pr->p->run(rho0, result); });
}

int edus_parareal_get_state(edus_parareal* pr, int n, double* rho)
{
    EDUS_REQUIRE(pr && rho);
    return guarded([&] {
        if (n < 0 || n > pr->p->N) throw edus::EngineError(EDUS_ERR_ARG, "slice boundary out of range");
        pr->p->coarse->download(pr->p->U + size_t(n) * pr->p->ie, rho);
    });
}

int edus_parareal_collect(edus_ctx* base, const edus_parareal_params* params, int runs, const double* E_fine,
                          const double* E_coarse, const float* E_feat, int field_samples, int k_samples,
                          unsigned long long seed, float* X, float* T)
{
    EDUS_REQUIRE(base && params && E_fine && E_coarse && X && T && field_samples >= 0 &&
                 (field_samples == 0 || E_feat));
    return guarded([&] {
        edus::collect(base->e, *params, runs, E_fine, E_coarse, E_feat, field_samples, k_samples, seed, X, T);
    });
}

}  // extern "C"
