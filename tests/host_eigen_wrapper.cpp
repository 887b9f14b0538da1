// Host-only wrapper of the library's dense Hermitian eigensolver (csrc/sed_eigen.cpp) for
// the CPU tests: complex arrays as interleaved doubles, column-major.
#include "../realspace_exactdiagonalization_amd/csrc/sed_internal.hpp"

namespace sed {
void set_error(const std::string&) {}
int fail(int code, const std::string&) { return code; }
}  // namespace sed

extern "C" void host_hermitian_eigh(int n, const double* a, double* evals, double* evecs, int nvec) {
    std::vector<sed::cd> A((size_t)n * n);
    for (size_t i = 0; i < A.size(); ++i) A[i] = sed::cd(a[2 * i], a[2 * i + 1]);
    std::vector<double> e;
    std::vector<sed::cd> X;
    sed::hermitian_eigh(n, A, e, evecs ? &X : nullptr, nvec);
    for (int i = 0; i < n; ++i) evals[i] = e[i];
    if (evecs)
        for (size_t i = 0; i < X.size(); ++i) {
            evecs[2 * i] = X[i].real();
            evecs[2 * i + 1] = X[i].imag();
        }
}
