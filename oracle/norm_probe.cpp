// Test infrastructure only (never linked into the product).
//
// The reference's state-vector sampling sums std::norm of each amplitude
// sequentially (src/cppsim/state.hpp:915-920) in a translation unit built with
// the reference's release flags (CMakeLists.txt:36, 281, 349: -std=gnu++14
// -O3 -march=native -mfpmath=both).  Whether GCC contracts norm's x*x + y*y
// into a fused multiply-add, and in which order, decides the last bit of each
// probability.  This file holds the same loop so that `make -C oracle
// norm-probe` compiles it with exactly those flags; tests/test_oracle.py checks
// its stacked probabilities against the oracle's fma(re, re, im*im) bit for bit.
#include <complex>
#include <cstddef>

extern "C" void norm_probe_stacked_prob(const std::complex<double>* ptr, std::size_t dim, double* stacked) {
    double sum = 0.;
    stacked[0] = 0.;
    for (std::size_t i = 0; i < dim; ++i) {
        sum += norm(ptr[i]);
        stacked[i + 1] = sum;
    }
}
