// Host-side dense Hermitian eigensolver: Householder reduction to a real symmetric
// tridiagonal matrix followed by the implicit-shift QL iteration.  Used for the
// Lanczos projected matrix (order <= a few hundred) and for sectors of dimension
// <= 500, which the reference diagonalises densely (src/symmetry_resolved_ed.jl:1004-1017).
#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <vector>

#include "sed_internal.hpp"

namespace sed {

// sqrt(x^2 + y^2): the plain form where it can neither overflow nor underflow, hypot
// elsewhere (hypot costs several times more, and QL calls this once per rotation).
static inline double pyth(double x, double y) {
    const double r = std::sqrt(x * x + y * y);
    return (r > 1e-150 && r < 1e150) ? r : std::hypot(x, y);
}

// Implicit QL on a real symmetric tridiagonal matrix (diagonal d, sub-diagonal e with
// e[i] coupling i and i+1).  The rotations are applied from the right to z (zrows x n,
// row-major) when z is non-null: with z = Q (n x n) it ends as the eigenvector matrix
// (column j = eigenvector j); with z = a few rows of Q it ends as those rows of it.
static void tridiag_ql(int n, std::vector<double>& d, std::vector<double>& e, std::vector<double>* z,
                       int zrows = -1) {
    if (zrows < 0) zrows = n;
    if (n <= 1) return;
    e.resize(n);
    e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        while (true) {
            int m = l;
            for (; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-300 || std::fabs(e[m]) <= 2.2e-16 * dd) break;
            }
            if (m == l) break;
            if (++iter > 200) break;
            // Wilkinson-type shift from the leading 2x2 block
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = pyth(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            for (; i >= l; --i) {
                double f = s * e[i];
                const double b = c * e[i];
                r = pyth(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (z) {
                    for (int k = 0; k < zrows; ++k) {
                        f = (*z)[(size_t)k * n + i + 1];
                        (*z)[(size_t)k * n + i + 1] = s * (*z)[(size_t)k * n + i] + c * f;
                        (*z)[(size_t)k * n + i] = c * (*z)[(size_t)k * n + i] - s * f;
                    }
                }
            }
            if (r == 0.0 && i >= l) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
}

void hermitian_eigh(int n, std::vector<cd>& A, std::vector<double>& evals, std::vector<cd>* evecs) {
    // A: column-major n x n Hermitian (only consistency of both triangles assumed).
    auto a = [&](int i, int j) -> cd& { return A[(size_t)j * n + i]; };
    std::vector<std::vector<cd>> vs;      // Householder vectors (length n, zeros above k+1)
    std::vector<double> taus;
    std::vector<cd> v(n), p(n), q(n);
    for (int k = 0; k + 2 < n; ++k) {
        double sigma2 = 0.0;
        for (int i = k + 1; i < n; ++i) sigma2 += std::norm(a(i, k));
        const double tail2 = sigma2 - std::norm(a(k + 1, k));
        if (tail2 <= 0.0) {
            vs.emplace_back();
            taus.push_back(0.0);
            continue;
        }
        const double sigma = std::sqrt(sigma2);
        const cd x0 = a(k + 1, k);
        const double ax0 = std::abs(x0);
        const cd ph = ax0 > 0 ? x0 / ax0 : cd(1.0, 0.0);
        const cd alpha = -ph * sigma;
        std::fill(v.begin(), v.end(), cd(0.0));
        for (int i = k + 1; i < n; ++i) v[i] = a(i, k);
        v[k + 1] -= alpha;
        const double vnorm2 = 2.0 * sigma * (sigma + ax0);
        const double tau = 2.0 / vnorm2;
        // p = tau * B v over the trailing block
        for (int i = k + 1; i < n; ++i) {
            cd s = 0.0;
            for (int j = k + 1; j < n; ++j) s += a(i, j) * v[j];
            p[i] = tau * s;
        }
        cd vhp = 0.0;
        for (int i = k + 1; i < n; ++i) vhp += std::conj(v[i]) * p[i];
        const double K = 0.5 * tau * vhp.real();
        for (int i = k + 1; i < n; ++i) q[i] = p[i] - K * v[i];
        for (int j = k + 1; j < n; ++j)
            for (int i = k + 1; i < n; ++i) a(i, j) -= v[i] * std::conj(q[j]) + q[i] * std::conj(v[j]);
        a(k + 1, k) = alpha;
        a(k, k + 1) = std::conj(alpha);
        for (int i = k + 2; i < n; ++i) {
            a(i, k) = 0.0;
            a(k, i) = 0.0;
        }
        vs.push_back(v);
        taus.push_back(tau);
    }
    std::vector<double> d(n), e(std::max(n, 1), 0.0);
    std::vector<cd> ph(n, cd(1.0, 0.0));
    for (int i = 0; i < n; ++i) d[i] = a(i, i).real();
    for (int i = 0; i + 1 < n; ++i) {
        const cd ei = a(i + 1, i);
        const double m = std::abs(ei);
        e[i] = m;
        ph[i + 1] = m > 0 ? ph[i] * ei / m : ph[i];
    }
    std::vector<double> z;
    if (evecs) {
        z.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
    }
    tridiag_ql(n, d, e, evecs ? &z : nullptr);
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
    evals.resize(n);
    for (int j = 0; j < n; ++j) evals[j] = d[order[j]];
    if (!evecs) return;
    // eigenvectors of A: Q P z, Q = H_0 H_1 ... H_{n-3}
    std::vector<cd>& X = *evecs;
    X.assign((size_t)n * n, cd(0.0));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) X[(size_t)j * n + i] = ph[i] * z[(size_t)i * n + order[j]];
    for (int k = (int)vs.size() - 1; k >= 0; --k) {
        if (taus[k] == 0.0) continue;
        const std::vector<cd>& vk = vs[k];
        for (int j = 0; j < n; ++j) {
            cd s = 0.0;
            for (int i = k + 1; i < n; ++i) s += std::conj(vk[i]) * X[(size_t)j * n + i];
            s *= taus[k];
            for (int i = k + 1; i < n; ++i) X[(size_t)j * n + i] -= vk[i] * s;
        }
    }
}

// Householder reduction of a real symmetric n x n matrix (row-major a, destroyed) to
// tridiagonal form d, e: A = Q T Q^T with Q = H_0 H_1 ..., H_k = I - tau_k v_k v_k^T
// (tau_k = 0 where column k was already reduced).
static void householder_tridiag(int n, std::vector<double>& a, std::vector<std::vector<double>>& vs,
                                std::vector<double>& taus, std::vector<double>& d, std::vector<double>& e) {
    auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
    std::vector<double> v(n), p(n), q(n);
    for (int k = 0; k + 2 < n; ++k) {
        double sigma2 = 0.0;
        for (int i = k + 1; i < n; ++i) sigma2 += A(i, k) * A(i, k);
        const double x0 = A(k + 1, k);
        if (sigma2 - x0 * x0 <= 0.0) {
            vs.emplace_back();
            taus.push_back(0.0);
            continue;
        }
        const double sigma = std::sqrt(sigma2);
        const double alpha = x0 >= 0 ? -sigma : sigma;
        std::fill(v.begin(), v.end(), 0.0);
        for (int i = k + 1; i < n; ++i) v[i] = A(i, k);
        v[k + 1] -= alpha;
        const double tau = 2.0 / (2.0 * sigma * (sigma + std::fabs(x0)));
        for (int i = k + 1; i < n; ++i) {
            double sm = 0.0;
            for (int j = k + 1; j < n; ++j) sm += A(i, j) * v[j];
            p[i] = tau * sm;
        }
        double vp = 0.0;
        for (int i = k + 1; i < n; ++i) vp += v[i] * p[i];
        const double K = 0.5 * tau * vp;
        for (int i = k + 1; i < n; ++i) q[i] = p[i] - K * v[i];
        for (int i = k + 1; i < n; ++i)
            for (int j = k + 1; j < n; ++j) A(i, j) -= v[i] * q[j] + q[i] * v[j];
        A(k + 1, k) = alpha;
        A(k, k + 1) = alpha;
        for (int i = k + 2; i < n; ++i) A(i, k) = A(k, i) = 0.0;
        vs.push_back(v);
        taus.push_back(tau);
    }
    d.assign(n, 0.0);
    e.assign(std::max(n, 1), 0.0);
    for (int i = 0; i < n; ++i) d[i] = A(i, i);
    for (int i = 0; i + 1 < n; ++i) e[i] = A(i + 1, i);
}

static std::vector<int> ascending(const std::vector<double>& d) {
    std::vector<int> order(d.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
    return order;
}

// Real symmetric n x n (row-major a, destroyed): Householder reduction to tridiagonal
// form with the orthogonal factor accumulated, then implicit QL on it.  Eigenvalues
// ascending; evecs row-major, column j = eigenvector j.
void jacobi_eigh(int n, std::vector<double>& a, std::vector<double>& evals, std::vector<double>& evecs) {
    std::vector<std::vector<double>> vs;
    std::vector<double> taus, d, e;
    householder_tridiag(n, a, vs, taus, d, e);
    // Q = H_0 H_1 ... applied to the identity (row-major)
    std::vector<double> z((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
    for (int k = (int)vs.size() - 1; k >= 0; --k) {
        if (taus[k] == 0.0) continue;
        const std::vector<double>& vk = vs[k];
        for (int j = 0; j < n; ++j) {
            double sm = 0.0;
            for (int i = k + 1; i < n; ++i) sm += vk[i] * z[(size_t)i * n + j];
            sm *= taus[k];
            for (int i = k + 1; i < n; ++i) z[(size_t)i * n + j] -= vk[i] * sm;
        }
    }
    tridiag_ql(n, d, e, &z);
    const std::vector<int> order = ascending(d);
    evals.resize(n);
    evecs.assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        evals[j] = d[order[j]];
        for (int i = 0; i < n; ++i) evecs[(size_t)i * n + j] = z[(size_t)i * n + order[j]];
    }
}

// Eigenvalues (ascending) of a real symmetric n x n matrix (row-major a, destroyed) and
// the last row of its eigenvector matrix, last[j] = component n-1 of eigenvector j (the
// Lanczos residual estimates): the same reduction and QL iteration as jacobi_eigh, with
// only the row e_{n-1}^T Q carried through the reflectors and rotations, O(n^2) per QL
// sweep instead of O(n^3).
void eigh_last_row(int n, std::vector<double>& a, std::vector<double>& evals, std::vector<double>& last) {
    std::vector<std::vector<double>> vs;
    std::vector<double> taus, d, e;
    householder_tridiag(n, a, vs, taus, d, e);
    std::vector<double> z((size_t)n, 0.0);   // e_{n-1}^T H_0 H_1 ...
    if (n > 0) z[(size_t)n - 1] = 1.0;
    for (size_t k = 0; k < vs.size(); ++k) {
        if (taus[k] == 0.0) continue;
        const std::vector<double>& vk = vs[k];
        double sm = 0.0;
        for (int i = (int)k + 1; i < n; ++i) sm += z[i] * vk[i];
        sm *= taus[k];
        for (int i = (int)k + 1; i < n; ++i) z[i] -= sm * vk[i];
    }
    tridiag_ql(n, d, e, &z, 1);
    const std::vector<int> order = ascending(d);
    evals.resize(n);
    last.resize(n);
    for (int j = 0; j < n; ++j) {
        evals[j] = d[order[j]];
        last[j] = z[order[j]];
    }
}

}  // namespace sed
