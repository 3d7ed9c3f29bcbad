/* C++ host classes with the reference's seam names over the C ABI of
 * edus_gpu.h (header only; nothing here is exported by the library).
 *
 *   edus::seam::FourierTransform   FourierTransform (src/fftPair/fftPair.hpp:41-60):
 *                                  two non-owning arrays and fft(sign)
 *   edus::seam::Operator<T>        Operator<T> (src/Operator/Operator.hpp): the k- and
 *                                  R-space blocks of one operator field, its space and
 *                                  band gauge, go_to_R / go_to_k (566-606) and
 *                                  go_to_bloch / go_to_wannier (528-564)
 *
 * The data stay host arrays [Nk][nb][nb] of std::complex<double> in the
 * reference's k order and FFT-order R grid; the transforms and rotations run on
 * the GPU through edus_fourier and edus_rotate with the context's grid and
 * eigenvectors.  Errors are thrown as std::runtime_error with
 * edus_last_error()'s text, as the reference throws runtime_error.  One host
 * thread per context, as for the C ABI. */
#pragma once

#include <complex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "edus_gpu.h"

namespace edus {
namespace seam {

enum Space { R, k };                 // Operator.hpp's Space
enum BandGauge { wannier, bloch };   // Operator.hpp's BandGauge

inline void check(int rc)
{
    if (rc != EDUS_OK) throw std::runtime_error(std::string("edus: ") + edus_last_error());
}

// fft(-1): x -> k, sign -1, scaled by 1/Nk; fft(+1): k -> x, sign +1, unscaled
// (fftPair.cpp:89-107).  The Operator binds x to its k-space block and k to
// its R-space block (Operator.hpp:247), so fft(-1) is go_to_R.  Copies are
// deleted and moves defaulted, as the reference's (fftPair.hpp:46-49).
class FourierTransform {
  public:
    FourierTransform() = default;
    FourierTransform(const FourierTransform&) = delete;
    FourierTransform& operator=(const FourierTransform&) = delete;
    FourierTransform(FourierTransform&&) = default;
    FourierTransform& operator=(FourierTransform&&) = default;

    void initialize(edus_ctx* ctx, std::complex<double>* x, std::complex<double>* k)
    {
        ctx_ = ctx;
        x_ = x;
        k_ = k;
    }

    void fft(int sign)
    {
        if (!ctx_ || !x_ || !k_) throw std::runtime_error("FourierTransform: not initialized");
        if (sign != 1 && sign != -1) throw std::runtime_error("FourierTransform: sign must be +1 or -1");
        if (sign == -1)
            check(edus_fourier(ctx_, -1, reinterpret_cast<const double*>(x_), reinterpret_cast<double*>(k_)));
        else
            check(edus_fourier(ctx_, +1, reinterpret_cast<const double*>(k_), reinterpret_cast<double*>(x_)));
    }

  private:
    edus_ctx* ctx_ = nullptr;
    std::complex<double>* x_ = nullptr;
    std::complex<double>* k_ = nullptr;
};

template <class T>
class Operator {
    static_assert(std::is_same<T, std::complex<double>>::value, "the engine's operators are complex<double>");

  public:
    Operator() = default;
    // Copies copy the data but not the transform, which is rebound to the
    // copy's own blocks (the reference copies the blocks and never the plan,
    // Operator.hpp:132-145).
    Operator(const Operator& o) { *this = o; }
    Operator& operator=(const Operator& o)
    {
        ctx_ = o.ctx_;
        nb_ = o.nb_;
        nk_ = o.nk_;
        blk_k_ = o.blk_k_;
        blk_R_ = o.blk_R_;
        space_ = o.space_;
        gauge_ = o.gauge_;
        locked_space_ = o.locked_space_;
        locked_gauge_ = o.locked_gauge_;
        if (ctx_) ft_.initialize(ctx_, blk_k_.data(), blk_R_.data());
        return *this;
    }

    // initialize_fft (Operator.hpp:192-252) on the context's grid: both
    // blocks zero, the space locked to R as the reference leaves it
    void initialize_fft(edus_ctx* ctx)
    {
        int layout = 0, n0 = 0, n1 = 0;
        check(edus_ctx_info(ctx, &nb_, &n0, &n1, &layout));
        if (layout != EDUS_LAYOUT_FULL) throw std::runtime_error("Operator: the context must use the FULL layout");
        ctx_ = ctx;
        nk_ = n0 * n1;
        blk_k_.assign(size_t(nk_) * nb_ * nb_, T(0.));
        blk_R_.assign(size_t(nk_) * nb_ * nb_, T(0.));
        ft_.initialize(ctx_, blk_k_.data(), blk_R_.data());
        space_ = R;
        locked_space_ = true;
    }

    std::vector<T>& get_Operator(Space s) { return s == k ? blk_k_ : blk_R_; }
    const std::vector<T>& get_Operator(Space s) const { return s == k ? blk_k_ : blk_R_; }
    // element (n, m) at grid point ik of the block of space s
    T& operator()(Space s, int ik, int n, int m) { return get_Operator(s)[(size_t(ik) * nb_ + n) * nb_ + m]; }
    Space get_space() const { return space_; }
    BandGauge get_bandgauge() const { return gauge_; }
    int get_nbnd() const { return nb_; }
    int get_nk() const { return nk_; }

    void lock_space(Space s)
    {
        space_ = s;
        locked_space_ = true;
    }
    void lock_gauge(BandGauge g)
    {
        gauge_ = g;
        locked_gauge_ = true;
    }

    // go_to_R / go_to_k (Operator.hpp:566-606): do_fft false only relabels
    void go_to_R(bool do_fft = true)
    {
        require_space();
        if (space_ == R) return;
        if (do_fft) ft_.fft(-1);
        space_ = R;
    }
    void go_to_k(bool do_fft = true)
    {
        require_space();
        if (space_ == k) return;
        if (do_fft) ft_.fft(+1);
        space_ = k;
    }

    // O_bloch = U^dagger O U and back, in k space (Operator.hpp:528-564)
    void go_to_bloch() { rotate(bloch); }
    void go_to_wannier() { rotate(wannier); }

  private:
    edus_ctx* ctx_ = nullptr;
    int nb_ = 0, nk_ = 0;
    std::vector<T> blk_k_, blk_R_;
    Space space_ = R;
    BandGauge gauge_ = wannier;
    bool locked_space_ = false, locked_gauge_ = false;
    FourierTransform ft_;

    void require_space() const
    {
        if (!locked_space_ || !ctx_) throw std::runtime_error("Operator: initialize_fft and lock_space first");
    }

    void rotate(BandGauge to)
    {
        if (!locked_gauge_) throw std::runtime_error("Operator: lock_gauge first");
        if (gauge_ == to) return;
        go_to_k();
        check(edus_rotate(ctx_, to == bloch ? 1 : 0, reinterpret_cast<const double*>(blk_k_.data()),
                          reinterpret_cast<double*>(blk_k_.data())));
        gauge_ = to;
    }
};

}  // namespace seam
}  // namespace edus
