// The harmonic-Ritz eigenproblem of GCRO-DR on the host.
//
// The reference solves (AZ)^T AZ y = theta (AZ)^T Z y with LAPACK's dggev
// (gcrodr.hpp:30-78, 105-135).  Here the first matrix G = (AZ)^T AZ is
// symmetric positive definite whenever AZ has full column rank, so with its
// Cholesky factor G = L L^T the problem becomes the standard one
//
//     C w = mu w,   C = L^-1 M L^-T,   mu = 1 / theta,   y = L^-T w,
//
// for M = (AZ)^T Z.  Its eigenvalues come from Householder reduction to
// Hessenberg form and the shifted QR algorithm in complex arithmetic
// (Wilkinson shifts, deflation on negligible subdiagonals); each eigenvector
// comes from inverse iteration with the computed eigenvalue on C itself.
// Output follows dggev's convention: theta_j = (real_j + i imag_j) / beta_j,
// a complex pair as two consecutive entries with imag > 0 first, its
// eigenvector as (real part, imaginary part) columns; beta = 0 marks an
// infinite theta (mu = 0, singular M).
//
// When AZ has dependent columns G is singular and has no Cholesky factor;
// the problem is then solved over G's range (T = S V_r diag(l_r)^-1/2 from
// the Jacobi eigendecomposition of G equilibrated by S = diag(G)^-1/2,
// C = T^T M T, y = T w) and the remaining values, indeterminate 0/0 in
// dggev, come last with beta = 0.
#include "../owtk_hip_api.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <stdexcept>
#include <vector>

namespace owtk::host {

namespace {

using cplx = std::complex<double>;

// Column-major n x n access.
struct Mat {
    int n;
    std::vector<double> a;
    double& operator()(int i, int j) { return a[static_cast<std::size_t>(j) * n + i]; }
    double operator()(int i, int j) const { return a[static_cast<std::size_t>(j) * n + i]; }
};

// Upper Hessenberg form by Householder reflections (similarity, in place).
void hessenberg(Mat& h)
{
    const int n = h.n;
    std::vector<double> v(n);
    for (int k = 0; k < n - 2; ++k) {
        double alpha = 0.0;
        for (int i = k + 1; i < n; ++i) alpha = std::hypot(alpha, h(i, k));
        if (alpha == 0.0) continue;
        if (h(k + 1, k) > 0) alpha = -alpha;
        for (int i = 0; i < n; ++i) v[i] = 0.0;
        v[k + 1] = h(k + 1, k) - alpha;
        for (int i = k + 2; i < n; ++i) v[i] = h(i, k);
        double vn = 0.0;
        for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
        if (vn == 0.0) continue;
        // H = P H P with P = I - 2 v v^T / (v^T v)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int i = k + 1; i < n; ++i) s += v[i] * h(i, j);
            s = 2.0 * s / vn;
            for (int i = k + 1; i < n; ++i) h(i, j) -= s * v[i];
        }
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = k + 1; j < n; ++j) s += h(i, j) * v[j];
            s = 2.0 * s / vn;
            for (int j = k + 1; j < n; ++j) h(i, j) -= s * v[j];
        }
        for (int i = k + 2; i < n; ++i) h(i, k) = 0.0;
    }
}

// Eigenvalues of an upper Hessenberg matrix by shifted QR in complex
// arithmetic, one Givens sweep per step on the active block.
std::vector<cplx> hessenberg_eigenvalues(const Mat& h0)
{
    const int n = h0.n;
    std::vector<cplx> h(static_cast<std::size_t>(n) * n);
    auto H = [&](int i, int j) -> cplx& { return h[static_cast<std::size_t>(j) * n + i]; };
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) H(i, j) = h0(i, j);
    std::vector<cplx> ev(n);
    const double eps = std::numeric_limits<double>::epsilon();
    int hi = n - 1, iter = 0, total = 0;
    std::vector<cplx> cs(n), sn(n);
    while (hi >= 0) {
        int l = hi;
        while (l > 0) {
            const double s = std::abs(H(l - 1, l - 1)) + std::abs(H(l, l));
            if (std::abs(H(l, l - 1)) <= eps * (s > 0 ? s : 1.0)) break;
            --l;
        }
        if (l > 0) H(l, l - 1) = 0.0;
        if (l == hi) {
            ev[hi] = H(hi, hi);
            --hi;
            iter = 0;
            continue;
        }
        if (++total > 100 * n + 1000) throw std::runtime_error("harmonic-Ritz eigenproblem did not converge");
        ++iter;
        // Wilkinson shift: the eigenvalue of the trailing 2x2 block nearer its
        // last diagonal entry; an exceptional shift every 11th step
        cplx sigma;
        const cplx a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
        if (iter % 11 == 0) {
            sigma = d + cplx(std::abs(c), std::abs(H(hi - 1, std::max(hi - 2, l))));
        } else {
            const cplx half = 0.5 * (a - d);
            const cplx root = std::sqrt(half * half + b * c);
            const cplx m1 = d - b * c / (half + root), m2 = d - b * c / (half - root);
            const bool use1 = std::abs(half + root) >= std::abs(half - root);
            sigma = use1 ? m1 : m2;
            if (!(std::isfinite(sigma.real()) && std::isfinite(sigma.imag()))) sigma = d;
        }
        for (int i = l; i <= hi; ++i) H(i, i) -= sigma;
        // QR: rotations zeroing the subdiagonal, applied from the left
        for (int k = l; k < hi; ++k) {
            const cplx x = H(k, k), y = H(k + 1, k);
            const double r = std::hypot(std::abs(x), std::abs(y));
            cplx cc = 1.0, ss = 0.0;
            if (r != 0.0) {
                cc = x / r;
                ss = y / r;
            }
            cs[k] = cc;
            sn[k] = ss;
            for (int j = k; j <= hi; ++j) {
                const cplx u = H(k, j), w = H(k + 1, j);
                H(k, j) = std::conj(cc) * u + std::conj(ss) * w;
                H(k + 1, j) = -ss * u + cc * w;
            }
        }
        // RQ: the same rotations from the right
        for (int k = l; k < hi; ++k) {
            const cplx cc = cs[k], ss = sn[k];
            for (int i = l; i <= std::min(k + 1, hi); ++i) {
                const cplx u = H(i, k), w = H(i, k + 1);
                H(i, k) = u * cc + w * ss;
                H(i, k + 1) = -u * std::conj(ss) + w * std::conj(cc);
            }
        }
        for (int i = l; i <= hi; ++i) H(i, i) += sigma;
    }
    return ev;
}

// Solve (A - mu I) x = b in place by Gaussian elimination with partial
// pivoting (complex).  A is n x n real column-major.
void shifted_solve(int n, const std::vector<double>& a, cplx mu, std::vector<cplx>& x)
{
    std::vector<cplx> m(static_cast<std::size_t>(n) * n);
    auto M = [&](int i, int j) -> cplx& { return m[static_cast<std::size_t>(j) * n + i]; };
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) M(i, j) = a[static_cast<std::size_t>(j) * n + i] - (i == j ? mu : cplx(0.0));
    double scale = 0.0;
    for (const cplx& v : m) scale = std::max(scale, std::abs(v));
    const double tiny = std::numeric_limits<double>::epsilon() * (scale > 0 ? scale : 1.0);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::abs(M(i, k)) > std::abs(M(p, k))) p = i;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(M(k, j), M(p, j));
            std::swap(x[k], x[p]);
        }
        if (std::abs(M(k, k)) < tiny) M(k, k) = tiny;  // exactly singular: perturb the pivot
        for (int i = k + 1; i < n; ++i) {
            const cplx f = M(i, k) / M(k, k);
            if (f == cplx(0.0)) continue;
            for (int j = k + 1; j < n; ++j) M(i, j) -= f * M(k, j);
            x[i] -= f * x[k];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        cplx s = x[k];
        for (int j = k + 1; j < n; ++j) s -= M(k, j) * x[j];
        x[k] = s / M(k, k);
    }
}

// Eigenvector of A for the eigenvalue mu by inverse iteration.
std::vector<cplx> inverse_iteration(int n, const std::vector<double>& a, cplx mu)
{
    std::vector<cplx> x(n);
    for (int i = 0; i < n; ++i) x[i] = cplx(1.0 + 0.01 * i, 0.0);
    for (int it = 0; it < 4; ++it) {
        shifted_solve(n, a, mu, x);
        double nr = 0.0;
        for (const cplx& v : x) nr = std::hypot(nr, std::abs(v));
        if (!(nr > 0.0) || !std::isfinite(nr)) throw std::runtime_error("harmonic-Ritz eigenproblem did not converge");
        for (cplx& v : x) v /= nr;
    }
    // fix the phase so that the largest component is real and positive
    int big = 0;
    for (int i = 1; i < n; ++i)
        if (std::abs(x[i]) > std::abs(x[big])) big = i;
    const cplx phase = std::abs(x[big]) > 0 ? std::conj(x[big]) / std::abs(x[big]) : cplx(1.0);
    for (cplx& v : x) v *= phase;
    return x;
}

// Cholesky G = L L^T (lower, column-major).  False when a pivot is below
// 8 n eps of its diagonal entry: column j of AZ then lies, to working
// precision, in the span of the columns before it (the pivot is the squared
// distance), and the factor would carry no correct digits.
bool cholesky(int n, const double* g, std::vector<double>& L)
{
    L.assign(static_cast<std::size_t>(n) * n, 0.0);
    auto Lr = [&](int i, int j) -> double& { return L[static_cast<std::size_t>(j) * n + i]; };
    for (int j = 0; j < n; ++j) {
        const double gjj = g[static_cast<std::size_t>(j) * n + j];
        double d = gjj;
        for (int k = 0; k < j; ++k) d -= Lr(j, k) * Lr(j, k);
        if (!(d > 8.0 * n * std::numeric_limits<double>::epsilon() * gjj)) return false;
        Lr(j, j) = std::sqrt(d);
        for (int i = j + 1; i < n; ++i) {
            double s = g[static_cast<std::size_t>(j) * n + i];
            for (int k = 0; k < j; ++k) s -= Lr(i, k) * Lr(j, k);
            Lr(i, j) = s / Lr(j, j);
        }
    }
    return true;
}

// The whitening map of a singular G over its numerical range.  G is first
// equilibrated, S G S with S = diag(G)^-1/2 (columns of AZ scaled to unit
// norm, so a small column is not mistaken for a dependent one), then
// S G S = V diag(l) V^T by cyclic Jacobi rotations, and T = S V_r
// diag(l_r)^-1/2 (n x r, column-major) over the eigenvalues l > n eps max(l),
// so that T^T G T = I.
std::vector<double> range_whitening(int n, const double* g, int& r)
{
    std::vector<double> sc(n);
    for (int i = 0; i < n; ++i) {
        const double d = g[static_cast<std::size_t>(i) * n + i];
        sc[i] = d > 0.0 ? 1.0 / std::sqrt(d) : 0.0;  // a zero column of AZ is a null direction
    }
    Mat a{n, std::vector<double>(static_cast<std::size_t>(n) * n)};
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) a(i, j) = sc[i] * g[static_cast<std::size_t>(j) * n + i] * sc[j];
    Mat v{n, std::vector<double>(static_cast<std::size_t>(n) * n, 0.0)};
    for (int i = 0; i < n; ++i) v(i, i) = 1.0;
    double scale = 0.0;
    for (double x : a.a) scale = std::max(scale, std::abs(x));
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < j; ++i) off = std::max(off, std::abs(a(i, j)));
        if (off <= std::numeric_limits<double>::epsilon() * std::numeric_limits<double>::epsilon() * scale) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                if (a(p, q) == 0.0) continue;
                // rotation zeroing a(p, q) (the symmetric Schur step)
                const double tau = (a(q, q) - a(p, p)) / (2.0 * a(p, q));
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double x = a(k, p), y = a(k, q);
                    a(k, p) = c * x - sn * y;
                    a(k, q) = sn * x + c * y;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double x = a(p, k), y = a(q, k);
                    a(p, k) = c * x - sn * y;
                    a(q, k) = sn * x + c * y;
                }
                for (int k = 0; k < n; ++k) {
                    const double x = v(k, p), y = v(k, q);
                    v(k, p) = c * x - sn * y;
                    v(k, q) = sn * x + c * y;
                }
            }
    }
    double top = 0.0, bottom = 0.0;
    for (int i = 0; i < n; ++i) {
        top = std::max(top, a(i, i));
        bottom = std::min(bottom, a(i, i));
    }
    // a Gram matrix is semidefinite: beyond rounding, a negative eigenvalue
    // means the input is not (AZ)^T AZ
    if (bottom < -1e3 * n * std::numeric_limits<double>::epsilon() * top || !(top >= 0.0))
        throw std::invalid_argument("harmonic-Ritz eigenproblem: (AZ)^T AZ is not positive semidefinite");
    std::vector<double> t;
    r = 0;
    for (int i = 0; i < n; ++i) {
        if (!(a(i, i) > n * std::numeric_limits<double>::epsilon() * top)) continue;
        const double w = 1.0 / std::sqrt(a(i, i));
        for (int k = 0; k < n; ++k) t.push_back(sc[k] * v(k, i) * w);
        ++r;
    }
    return t;
}

} // namespace

// theta and right eigenvectors of harmonic y = theta metric y (harmonic
// symmetric positive semidefinite), in dggev's output convention (see the
// file comment).
void harmonic_ggev(int n, const double* harmonic, const double* metric, double* real, double* imag, double* beta,
                   double* right)
{
    if (n <= 0) return;
    for (std::size_t i = 0; i < static_cast<std::size_t>(n) * n; ++i)
        if (!std::isfinite(harmonic[i]) || !std::isfinite(metric[i]))
            throw std::invalid_argument("harmonic-Ritz eigenproblem: nonfinite projected matrix");
    // The whitening map T with T^T G T = I: L^-T from the Cholesky factor, or
    // when G is singular (AZ with dependent columns) the map over G's range.
    // A direction y with G y = 0 has AZ y = 0, so Z y = 0 and M y = 0: dggev
    // leaves such values indeterminate (0/0); they are reported last with
    // beta = 0 and a zero vector.
    std::vector<double> L, T;
    const bool definite = cholesky(n, harmonic, L);
    int r = n;
    if (!definite) T = range_whitening(n, harmonic, r);
    auto Lr = [&](int i, int j) -> double { return L[static_cast<std::size_t>(j) * n + i]; };
    auto Tr = [&](int i, int j) -> double { return T[static_cast<std::size_t>(j) * n + i]; };
    Mat C{r, std::vector<double>(static_cast<std::size_t>(r) * r, 0.0)};
    if (definite) {
        // C = L^-1 M L^-T: forward-substitute the columns of M, then the rows
        C.a.assign(metric, metric + static_cast<std::size_t>(n) * n);
        for (int j = 0; j < n; ++j)  // columns: L^-1 M
            for (int i = 0; i < n; ++i) {
                double s = C(i, j);
                for (int k = 0; k < i; ++k) s -= Lr(i, k) * C(k, j);
                C(i, j) = s / Lr(i, i);
            }
        for (int i = 0; i < n; ++i)  // rows: (L^-1 M) L^-T
            for (int j = 0; j < n; ++j) {
                double s = C(i, j);
                for (int k = 0; k < j; ++k) s -= C(i, k) * Lr(j, k);
                C(i, j) = s / Lr(j, j);
            }
    } else {
        // C = T^T M T
        std::vector<double> mt(static_cast<std::size_t>(n) * r, 0.0);
        for (int j = 0; j < r; ++j)
            for (int k = 0; k < n; ++k)
                for (int i = 0; i < n; ++i)
                    mt[static_cast<std::size_t>(j) * n + i] += metric[static_cast<std::size_t>(k) * n + i] * Tr(k, j);
        for (int j = 0; j < r; ++j)
            for (int i = 0; i < r; ++i) {
                double s = 0.0;
                for (int k = 0; k < n; ++k) s += Tr(k, i) * mt[static_cast<std::size_t>(j) * n + k];
                C(i, j) = s;
            }
    }
    double cnorm = 0.0;
    for (double v : C.a) cnorm = std::max(cnorm, std::abs(v));
    std::vector<cplx> mu;
    if (r > 0) {
        Mat Hs = C;
        hessenberg(Hs);
        mu = hessenberg_eigenvalues(Hs);
    }
    // real eigenvalues and conjugate pairs, in the order found
    const double tol = 1e3 * std::numeric_limits<double>::epsilon() * (cnorm > 0 ? cnorm : 1.0);
    std::vector<int> used(r, 0);
    int out = 0;
    auto store_vector = [&](int col, const std::vector<cplx>& w, bool imaginary_part) {
        // y = L^-T w (back substitution with L^T) or y = T w
        std::vector<double> y(n, 0.0);
        if (definite) {
            for (int i = 0; i < n; ++i) y[i] = imaginary_part ? w[i].imag() : w[i].real();
            for (int i = n - 1; i >= 0; --i) {
                double s = y[i];
                for (int k = i + 1; k < n; ++k) s -= Lr(k, i) * y[k];
                y[i] = s / Lr(i, i);
            }
        } else {
            for (int j = 0; j < r; ++j) {
                const double wj = imaginary_part ? w[j].imag() : w[j].real();
                for (int i = 0; i < n; ++i) y[i] += Tr(i, j) * wj;
            }
        }
        for (int i = 0; i < n; ++i) right[static_cast<std::size_t>(col) * n + i] = y[i];
    };
    auto set_theta = [&](int col, cplx m) {
        if (std::abs(m) <= tol * 1e-3 || m == cplx(0.0)) {  // theta infinite
            real[col] = 1.0;
            imag[col] = 0.0;
            beta[col] = 0.0;
            return;
        }
        const cplx theta = 1.0 / m;
        real[col] = theta.real();
        imag[col] = theta.imag();
        beta[col] = 1.0;
    };
    for (int i = 0; i < r; ++i) {
        if (used[i]) continue;
        used[i] = 1;
        if (std::abs(mu[i].imag()) <= tol) {
            const cplx m(mu[i].real(), 0.0);
            set_theta(out, m);
            store_vector(out, inverse_iteration(r, C.a, m), false);
            ++out;
            continue;
        }
        // complex: pair with the nearest unused conjugate
        int partner = -1;
        double best = std::numeric_limits<double>::infinity();
        for (int j = 0; j < r; ++j)
            if (!used[j] && std::abs(mu[j] - std::conj(mu[i])) < best) {
                best = std::abs(mu[j] - std::conj(mu[i]));
                partner = j;
            }
        if (partner < 0) throw std::runtime_error("harmonic-Ritz eigenproblem: unpaired complex eigenvalue");
        used[partner] = 1;
        // the member whose theta has positive imaginary part first
        // (theta = 1/mu, so that is the member with negative Im mu)
        const cplx m = mu[i].imag() < 0 ? mu[i] : std::conj(mu[i]);
        const std::vector<cplx> w = inverse_iteration(r, C.a, m);
        set_theta(out, m);
        set_theta(out + 1, std::conj(m));
        store_vector(out, w, false);
        store_vector(out + 1, w, true);
        out += 2;
    }
    for (; out < n; ++out) {  // indeterminate: G y = M y = 0
        real[out] = 0.0;
        imag[out] = 0.0;
        beta[out] = 0.0;
        for (int i = 0; i < n; ++i) right[static_cast<std::size_t>(out) * n + i] = 0.0;
    }
}

} // namespace owtk::host

extern "C" int owtk_harmonic_ggev(int n, const double* harmonic, const double* metric, double* real, double* imag,
                                  double* beta, double* right)
{
    return owtk::guarded([&] {
        if (n < 0) throw std::invalid_argument("negative dimension");
        if (n > 0 && (!harmonic || !metric || !real || !imag || !beta || !right))
            throw std::invalid_argument("null argument");
        owtk::host::harmonic_ggev(n, harmonic, metric, real, imag, beta, right);
    });
}
