// C++ caller of the C ABI, written as the reference-side binding would be:
// the set-up arrays come from the library's reference-format set-up
// (edus_sim_*, standing in for the reference's Operator data), and the
// propagation runs through the low-level context exactly as the binding
// sketch for Simulation::do_onestep does (Simulation.cpp:381-432): print the
// observables, then one RK4 step with the field at t, t + dt/2, t + dt.
//
// Input: the reference's ci-test hBN run (ci-test/inputs/hBN.json.in).
// Usage: abi_hbn <tb seedname> <output dir>
// Writes <dir>/Time.txt, Population.txt and Velocity.txt, one row per step.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "edus_gpu.h"
#include "hbn_params.hpp"

namespace {

void check(int rc, const char* what)
{
    if (rc != EDUS_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, edus_last_error());
        std::exit(1);
    }
}

std::vector<double> get(edus_sim* sim, const char* name)
{
    const long n = edus_sim_get(sim, name, nullptr);
    if (n < 0) {
        std::fprintf(stderr, "edus_sim_get(%s) failed\n", name);
        std::exit(1);
    }
    const bool real = std::strcmp(name, "A") == 0 || std::strcmp(name, "eps") == 0;
    std::vector<double> out(size_t(n) * (real ? 1 : 2));
    if (edus_sim_get(sim, name, out.data()) < 0) {
        std::fprintf(stderr, "edus_sim_get(%s) failed: %s\n", name, edus_last_error());
        std::exit(1);
    }
    return out;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <tb seedname> <output dir>\n", argv[0]);
        return 2;
    }
    const edus_sim_params p = hbn_params(argv[1]);  // ci-test/inputs/hBN.json.in

    edus_sim* sim = edus_sim_create(&p);
    if (!sim) {
        std::fprintf(stderr, "edus_sim_create failed: %s\n", edus_last_error());
        return 1;
    }
    int nb = 0, N0 = 0, N1 = 0, nsteps = 0;
    double dt = 0.0;
    check(edus_sim_info(sim, &nb, &N0, &N1, &nsteps, &dt), "edus_sim_info");

    // the binding sketch: a context from the set-up arrays
    const std::vector<double> H0 = get(sim, "H0k"), U = get(sim, "U"), rho0 = get(sim, "rho0"), A = get(sim, "A");
    const std::vector<double> r[3] = {get(sim, "rk0"), get(sim, "rk1"), get(sim, "rk2")};
    const std::vector<double> V[3] = {get(sim, "Vel0"), get(sim, "Vel1"), get(sim, "Vel2")};
    edus_desc d;
    std::memset(&d, 0, sizeof(d));
    d.nb = nb;
    d.N0 = N0;
    d.N1 = N1;
    d.precision = EDUS_PREC_FP64;
    d.layout = EDUS_LAYOUT_FULL;
    d.H0k = H0.data();
    for (int x = 0; x < 3; ++x) {
        d.xik[x] = r[x].data();
        d.vel[x] = V[x].data();
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) d.A[i][j] = A[3 * i + j];
    d.U = U.data();
    d.filledbands = 1;
    edus_ctx* ctx = edus_create(&d);
    if (!ctx) {
        std::fprintf(stderr, "edus_create failed: %s\n", edus_last_error());
        return 1;
    }
    check(edus_set_state(ctx, rho0.data()), "edus_set_state");

    const std::string dir = argv[2];
    FILE* ft = std::fopen((dir + "/Time.txt").c_str(), "w");
    FILE* fp = std::fopen((dir + "/Population.txt").c_str(), "w");
    FILE* fv = std::fopen((dir + "/Velocity.txt").c_str(), "w");
    if (!ft || !fp || !fv) {
        std::fprintf(stderr, "cannot write into %s\n", dir.c_str());
        return 1;
    }
    double t = 0.0;  // DESolver's CurrentTime starts at 0
    std::vector<double> pb(nb), pw(nb);
    for (int s = 0; s < nsteps; ++s) {
        double v[6];
        check(edus_observables(ctx, pb.data(), pw.data(), v), "edus_observables");
        std::fprintf(ft, "%.17g\n", t);
        for (int n = 0; n < nb; ++n) std::fprintf(fp, "%.17g%s", pb[n], n + 1 < nb ? " " : "\n");
        for (int c = 0; c < 6; ++c) std::fprintf(fv, "%.17g%s", v[c], c < 5 ? " " : "\n");
        double E[9];
        for (int j = 0; j < 3; ++j) check(edus_sim_field(sim, t + 0.5 * dt * j, &E[3 * j]), "edus_sim_field");
        check(edus_rk4_steps(ctx, dt, 1, E), "edus_rk4_steps");
        t += dt;
    }
    check(edus_synchronize(ctx), "edus_synchronize");
    std::fclose(ft);
    std::fclose(fp);
    std::fclose(fv);
    check(edus_destroy(ctx), "edus_destroy");
    check(edus_sim_destroy(sim), "edus_sim_destroy");
    std::printf("abi_hbn: %d steps, nb %d, %d x %d k\n", nsteps, nb, N0, N1);
    return 0;
}
