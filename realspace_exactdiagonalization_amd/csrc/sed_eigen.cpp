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

// Eigenvectors of the real symmetric tridiagonal T (diagonal d, off-diagonal e[i]
// coupling i and i+1) for its eigenvalues w[0..m-1] (ascending, from the QL iteration,
// accurate to a few ulps of |T|), by inverse iteration: T - sigma is factored once per
// vector with partial pivoting (U's pivots below eps |T| raised to eps |T|), and the
// solve is repeated from a fixed pseudo-random start until the growth of the solution
// shows a residual at rounding level, plus one more solve.  Vectors whose eigenvalues lie
// within 1e-3 |T| of each other are kept orthogonal by modified Gram-Schmidt after every
// solve, and equal shifts within such a cluster are separated by 10 eps |T| (the rules
// of LAPACK's xSTEIN).  z: n x m column-major.
static void tridiag_inverse_iteration(int n, const std::vector<double>& d, const std::vector<double>& e,
                                      const std::vector<double>& w, int m, std::vector<double>& z) {
    z.assign((size_t)n * m, 0.0);
    double tnorm = 0.0;
    for (int i = 0; i < n; ++i)
        tnorm = std::max(tnorm, std::fabs(d[i]) + (i > 0 ? std::fabs(e[i - 1]) : 0.0) +
                                    (i + 1 < n ? std::fabs(e[i]) : 0.0));
    if (!(tnorm > 0.0) || n == 1) {   // T = 0 (or 1 x 1): the unit vectors
        for (int j = 0; j < m; ++j) z[(size_t)j * n + j] = 1.0;
        return;
    }
    const double eps = 2.220446049250313e-16;
    const double pivmin = eps * tnorm;
    const double ortol = 1e-3 * tnorm;
    const double pertol = 10.0 * eps * tnorm;
    const double grow = 1.0 / (1e3 * n * eps * tnorm);   // |solution| that certifies convergence
    std::vector<double> dd(n), dl(n), du(n), du2(n), x(n);
    std::vector<char> swap(n);
    int c0 = 0;           // first vector of the current cluster
    double sig_prev = 0.0;
    for (int j = 0; j < m; ++j) {
        double sigma = w[j];
        if (j > 0 && w[j] - w[j - 1] > ortol) c0 = j;
        if (j > c0 && sigma - sig_prev < pertol) sigma = sig_prev + pertol;
        sig_prev = sigma;
        // LU of T - sigma with partial pivoting: L unit lower (multipliers dl, row swaps),
        // U upper with diagonal dd and two super-diagonals du, du2
        for (int i = 0; i < n; ++i) {
            dd[i] = d[i] - sigma;
            dl[i] = du[i] = (i + 1 < n) ? e[i] : 0.0;
            du2[i] = 0.0;
        }
        for (int i = 0; i + 1 < n; ++i) {
            if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
                swap[i] = 0;
                if (dd[i] != 0.0) {
                    const double f = dl[i] / dd[i];
                    dl[i] = f;
                    dd[i + 1] -= f * du[i];
                } else {
                    dl[i] = 0.0;
                }
            } else {
                swap[i] = 1;
                const double f = dd[i] / dl[i];
                dd[i] = dl[i];
                dl[i] = f;
                const double t = du[i];
                du[i] = dd[i + 1];
                dd[i + 1] = t - f * dd[i + 1];
                if (i + 2 < n) {
                    du2[i] = du[i + 1];
                    du[i + 1] = -f * du[i + 1];
                }
            }
        }
        for (int i = 0; i < n; ++i)
            if (std::fabs(dd[i]) < pivmin) dd[i] = dd[i] < 0.0 ? -pivmin : pivmin;
        uint64_t st = 0x9e3779b97f4a7c15ull * (uint64_t)(j + 1);
        double nx = 0.0;
        for (int i = 0; i < n; ++i) {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            x[i] = (double)(st >> 11) * (2.0 / 9007199254740992.0) - 1.0;
            nx += x[i] * x[i];
        }
        for (int i = 0; i < n; ++i) x[i] /= std::sqrt(nx);
        bool conv = false;
        for (int it = 0; it < 8; ++it) {
            for (int i = 0; i + 1 < n; ++i) {   // L, with the row swaps
                if (!swap[i]) {
                    x[i + 1] -= dl[i] * x[i];
                } else {
                    const double t = x[i];
                    x[i] = x[i + 1];
                    x[i + 1] = t - dl[i] * x[i];
                }
            }
            for (int i = n - 1; i >= 0; --i) {   // U, rescaling before an overflow
                double v = x[i];
                if (i + 1 < n) v -= du[i] * x[i + 1];
                if (i + 2 < n) v -= du2[i] * x[i + 2];
                v /= dd[i];
                if (std::fabs(v) > 1e150) {
                    const double sc = 1.0 / std::fabs(v);
                    for (int t = 0; t < n; ++t) x[t] *= sc;
                    v *= sc;
                }
                x[i] = v;
            }
            for (int c = c0; c < j; ++c) {
                const double* zc = &z[(size_t)c * n];
                double dot = 0.0;
                for (int i = 0; i < n; ++i) dot += zc[i] * x[i];
                for (int i = 0; i < n; ++i) x[i] -= dot * zc[i];
            }
            double nrm = 0.0;
            for (int i = 0; i < n; ++i) nrm += x[i] * x[i];
            nrm = std::sqrt(nrm);
            if (!(nrm > 0.0)) {   // the start lay in the span of the cluster: a fresh one
                for (int i = 0; i < n; ++i) {
                    st = st * 6364136223846793005ull + 1442695040888963407ull;
                    x[i] = (double)(st >> 11) * (2.0 / 9007199254740992.0) - 1.0;
                }
                continue;
            }
            for (int i = 0; i < n; ++i) x[i] /= nrm;
            if (conv) break;
            conv = nrm >= grow;
        }
        std::copy(x.begin(), x.end(), z.begin() + (size_t)j * n);
    }
}

// Complex Hermitian n x n (column-major A; only its lower triangle is read): all
// eigenvalues ascending, and the eigenvectors of the nvec lowest (n x nvec column-major)
// when evecs is non-null.  Householder reduction to a complex Hermitian tridiagonal form
// (lower triangle, re/im parts in separate columns so the inner loops run contiguously,
// complex products written out), a diagonal unitary P making it real, QL for the
// eigenvalues, inverse iteration on the real tridiagonal matrix for the wanted vectors,
// and the back-transformation Q P z of those vectors only.
void hermitian_eigh(int n, std::vector<cd>& A, std::vector<double>& evals, std::vector<cd>* evecs, int nvec) {
    nvec = std::max(0, std::min(nvec < 0 ? n : nvec, n));
    const size_t N = (size_t)n;
    // scaled by a power of two near 1 / max |a_ij| (exact), so that no sum of squares
    // below can overflow or underflow; the eigenvalues are scaled back at the end
    double amax = 0.0;
    for (size_t j = 0; j < N; ++j)
        for (size_t i = j; i < N; ++i) amax = std::max(amax, std::abs(A[j * N + i]));
    int ex = 0;
    if (amax > 0.0 && std::isfinite(amax)) std::frexp(amax, &ex);
    std::vector<double> ar(N * N), ai(N * N);
    for (size_t j = 0; j < N; ++j)
        for (size_t i = j; i < N; ++i) {
            ar[j * N + i] = std::ldexp(A[j * N + i].real(), -ex);
            ai[j * N + i] = i == j ? 0.0 : std::ldexp(A[j * N + i].imag(), -ex);
        }
    std::vector<double> hr(N * N, 0.0), hi(N * N, 0.0);   // Householder vector k in column k
    std::vector<double> taus(N, 0.0);
    std::vector<double> pr(N), pi(N), qr(N), qi(N);
    for (int k = 0; k + 2 < n; ++k) {
        const size_t c = (size_t)k * N;
        double sigma2 = 0.0;
        for (int i = k + 1; i < n; ++i) sigma2 += ar[c + i] * ar[c + i] + ai[c + i] * ai[c + i];
        const double x0r = ar[c + k + 1], x0i = ai[c + k + 1];
        const double ax0 = std::hypot(x0r, x0i);
        if (sigma2 - ax0 * ax0 <= 0.0) continue;   // column already reduced (tau = 0)
        const double sigma = std::sqrt(sigma2);
        const double phr = ax0 > 0 ? x0r / ax0 : 1.0, phi = ax0 > 0 ? x0i / ax0 : 0.0;
        const double alr = -phr * sigma, ali = -phi * sigma;   // new sub-diagonal entry
        double* vr = &hr[c];
        double* vi = &hi[c];
        for (int i = k + 1; i < n; ++i) {
            vr[i] = ar[c + i];
            vi[i] = ai[c + i];
        }
        vr[k + 1] -= alr;
        vi[k + 1] -= ali;
        const double tau = 1.0 / (sigma * (sigma + ax0));   // 2 / |v|^2
        // p = tau B v over the trailing block B (lower triangle: a(i,j), i > j, and
        // a(j,i) = conj(a(i,j)))
        for (int i = k + 1; i < n; ++i) pr[i] = pi[i] = 0.0;
        for (int j = k + 1; j < n; ++j) {
            const double* cr = &ar[(size_t)j * N];
            const double* ci = &ai[(size_t)j * N];
            const double vjr = vr[j], vji = vi[j];
            double sr = cr[j] * vjr, si = cr[j] * vji;
            for (int i = j + 1; i < n; ++i) {
                pr[i] += cr[i] * vjr - ci[i] * vji;
                pi[i] += cr[i] * vji + ci[i] * vjr;
                sr += cr[i] * vr[i] + ci[i] * vi[i];
                si += cr[i] * vi[i] - ci[i] * vr[i];
            }
            pr[j] += sr;
            pi[j] += si;
        }
        double vhp = 0.0;   // Re v^H p (v^H B v is real)
        for (int i = k + 1; i < n; ++i) {
            pr[i] *= tau;
            pi[i] *= tau;
            vhp += vr[i] * pr[i] + vi[i] * pi[i];
        }
        const double K = 0.5 * tau * vhp;
        for (int i = k + 1; i < n; ++i) {
            qr[i] = pr[i] - K * vr[i];
            qi[i] = pi[i] - K * vi[i];
        }
        // B -= v q^H + q v^H on the lower triangle
        for (int j = k + 1; j < n; ++j) {
            double* cr = &ar[(size_t)j * N];
            double* ci = &ai[(size_t)j * N];
            const double qjr = qr[j], qji = qi[j], vjr = vr[j], vji = vi[j];
            for (int i = j; i < n; ++i) {
                cr[i] -= vr[i] * qjr + vi[i] * qji + qr[i] * vjr + qi[i] * vji;
                ci[i] -= vi[i] * qjr - vr[i] * qji + qi[i] * vjr - qr[i] * vji;
            }
            ci[j] = 0.0;
        }
        ar[c + k + 1] = alr;
        ai[c + k + 1] = ali;
        taus[k] = tau;
    }
    // T = P T_r P^H with T_r real: e[i] = |t(i+1,i)|, P = diag(ph)
    std::vector<double> d(N), e(std::max<size_t>(N, 1), 0.0);
    std::vector<cd> ph(N, cd(1.0, 0.0));
    for (size_t i = 0; i < N; ++i) d[i] = ar[i * N + i];
    for (size_t i = 0; i + 1 < N; ++i) {
        const cd ei(ar[i * N + i + 1], ai[i * N + i + 1]);
        const double m = std::abs(ei);
        e[i] = m;
        ph[i + 1] = m > 0 ? ph[i] * (ei / m) : ph[i];
    }
    std::vector<double> w = d, ew = e;
    tridiag_ql(n, w, ew, nullptr);
    std::sort(w.begin(), w.end());
    evals.resize(N);
    for (size_t i = 0; i < N; ++i) evals[i] = std::ldexp(w[i], ex);
    if (!evecs || nvec == 0) return;
    std::vector<double> z;
    tridiag_inverse_iteration(n, d, e, w, nvec, z);
    std::vector<cd>& X = *evecs;
    X.assign(N * nvec, cd(0.0));
    for (int j = 0; j < nvec; ++j) {
        cd* x = &X[(size_t)j * N];
        for (size_t i = 0; i < N; ++i) x[i] = ph[i] * z[(size_t)j * N + i];
        for (int k = n - 3; k >= 0; --k) {   // x <- H_k x, H_k = I - tau_k v_k v_k^H
            if (taus[k] == 0.0) continue;
            const double* vr = &hr[(size_t)k * N];
            const double* vi = &hi[(size_t)k * N];
            double sr = 0.0, si = 0.0;
            for (int i = k + 1; i < n; ++i) {
                sr += vr[i] * x[i].real() + vi[i] * x[i].imag();
                si += vr[i] * x[i].imag() - vi[i] * x[i].real();
            }
            sr *= taus[k];
            si *= taus[k];
            for (int i = k + 1; i < n; ++i)
                x[i] -= cd(vr[i] * sr - vi[i] * si, vr[i] * si + vi[i] * sr);
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
