// Launchers of the conversion and observable kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

namespace edus {

struct ObsArgs {
    int nb, Nk, layout;
    const double2* rho;      // planes in `layout`
    const double2* rho_bloch;  // U^dagger rho U as full planes [nb*nb][Nk] (launch_rotate), or null
    const double2* vel[3];   // planes in `layout` or null
    double* partial;         // [blocks][2nb+6]
};

void launch_to_planes(const double2* full, double2* planes, int nb, int Nk, int layout, hipStream_t s);
void launch_from_planes(const double2* planes, double2* full, int nb, int Nk, int layout, hipStream_t s);
void launch_to_planes_real(const double2* full, double* planes, int nb, int Nk, int layout, hipStream_t s);
void launch_full_to_planes(const double2* full, double2* planes, int nb, int Nk, hipStream_t s);
int observables_blocks(int Nk);
void launch_observables(const ObsArgs& a, double* out, hipStream_t s);
// Hartree shift hdiag [B][nb] of B stacked instances from their plane sums
// over k, hsum [B][P] (hartree_shift)
void launch_hartree_sums(const double2* hsum, const double* hart, int nb, int layout, int P, int B, double inv_nk,
                         double2* hdiag, hipStream_t s);
void launch_axpy(double2* y, const double2* x, double a, long n, hipStream_t s);
// out = in rounded to FP32 (nearest), n complex values
void launch_to_f32(const double2* in, float2* out, long n, hipStream_t s);
void launch_add_diag(double2* H, const double2* d, double scale, int nb, int Nk, int layout, hipStream_t s);
// trace / Hermiticity monitors of B stacked states into the monitor slots
void launch_state_monitor(const double2* rho, int nb, int Nk, int layout, int P, int B, double nf,
                          unsigned long long* mon, hipStream_t s);

}  // namespace edus
