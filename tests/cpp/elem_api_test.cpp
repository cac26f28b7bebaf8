// Drives the C++ drop-in (include/elemental_amd/elem.hpp) the way a caller of
// elem::HPDSolve / elem::Cholesky does, and checks the results on the host.
// Prints one line per check and exits nonzero on a failure.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>

#include "elemental_amd/elem.hpp"

static uint64_t sm(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double unif(uint64_t& s) { return ((double)(sm(s) >> 11) + 0.5) * 0x1.0p-53 - 0.5; }

int main() {
    int fails = 0;
    const int n = 700, k = 5;
    uint64_t seed = 42;
    elem::Matrix<double> G(n, n), A(n, n), B(n, k);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) G.Set(i, j, unif(seed));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int l = 0; l < n; ++l) s += G.Get(i, l) * G.Get(j, l);
            A.Set(i, j, s / n + (i == j ? 1.0 : 0.0));
        }
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < n; ++i) B.Set(i, j, unif(seed));
    elem::Matrix<double> A0(A), B0(B);

    // HPDSolve: X in B, L in A's lower triangle, upper triangle untouched
    elem::HPDSolve(elem::LOWER, elem::NORMAL, A, B);
    double rn = 0, bn = 0;
    for (int c = 0; c < k; ++c)
        for (int i = 0; i < n; ++i) {
            double s = B0.Get(i, c);
            for (int j = 0; j < n; ++j) s -= A0.Get(i, j) * B.Get(j, c);
            rn += s * s;
            bn += B0.Get(i, c) * B0.Get(i, c);
        }
    double rel = std::sqrt(rn / bn);
    bool upper_ok = true;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i) upper_ok &= (A.Get(i, j) == A0.Get(i, j));
    double llt = 0, an = 0;  // ||L L^T - A||_F / ||A||_F on the lower triangle
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) {
            double s = 0;
            for (int l = 0; l <= j; ++l) s += A.Get(i, l) * A.Get(j, l);
            llt += (s - A0.Get(i, j)) * (s - A0.Get(i, j));
            an += A0.Get(i, j) * A0.Get(i, j);
        }
    llt = std::sqrt(llt / an);
    std::printf("HPDSolve n=%d k=%d: rel residual %.3e, ||LL^T-A||/||A|| %.3e, upper untouched %d\n", n, k, rel, llt,
                (int)upper_ok);
    fails += !(rel < 1e-13 && llt < 1e-14 && upper_ok);

    // Cholesky on a view with a leading dimension larger than the height
    elem::Matrix<double> big(n + 9, n + 3);
    elem::Matrix<double> V;
    elem::View(V, big, 4, 2, n, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) V.Set(i, j, A0.Get(i, j));
    elem::Cholesky(elem::LOWER, V);
    double d = 0;
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) d = std::fmax(d, std::fabs(V.Get(i, j) - A.Get(i, j)));
    std::printf("Cholesky on a view: max |L_view - L| = %.3e, border untouched %d\n", d,
                (int)(big.Get(0, 0) == 0.0 && big.Get(n + 8, n + 2) == 0.0));
    fails += !(d < 1e-12 && big.Get(0, 0) == 0.0);

    // a non-HPD matrix throws std::logic_error("A was not numerically HPD")
    elem::Matrix<double> N(A0);
    N.Set(300, 300, -1.0);
    bool threw = false;
    try {
        elem::Cholesky(elem::LOWER, N);
    } catch (const std::logic_error& e) {
        threw = std::string(e.what()) == "A was not numerically HPD";
    }
    std::printf("non-HPD throws logic_error: %d\n", (int)threw);
    fails += !threw;

    // Trsm(RIGHT, LOWER, TRANSPOSE) panel and Trsv pair
    elem::Matrix<double> P(50, 64), L64(64, 64);
    for (int j = 0; j < 64; ++j)
        for (int i = 0; i < 64; ++i) L64.Set(i, j, i == j ? 2.0 + unif(seed) : (i > j ? unif(seed) : 0.0));
    for (int j = 0; j < 64; ++j)
        for (int i = 0; i < 50; ++i) P.Set(i, j, unif(seed));
    elem::Matrix<double> P0(P);
    elem::Trsm(elem::RIGHT, elem::LOWER, elem::TRANSPOSE, elem::NON_UNIT, 1.0, L64, P);
    double e2 = 0;  // P L^T == P0
    for (int j = 0; j < 64; ++j)
        for (int i = 0; i < 50; ++i) {
            double s = 0;
            for (int l = 0; l <= j; ++l) s += P.Get(i, l) * L64.Get(j, l);
            e2 = std::fmax(e2, std::fabs(s - P0.Get(i, j)));
        }
    std::printf("Trsm(R,L,T): max |X L^T - B| = %.3e\n", e2);
    fails += !(e2 < 1e-13);

    // HPDSolve(UPPER): U in the upper triangle with U^T U = A, lower untouched
    {
        elem::Matrix<double> Au(A0), Bu(B0);
        elem::HPDSolve(elem::UPPER, elem::NORMAL, Au, Bu);
        double r2 = 0, utu = 0;
        for (int c = 0; c < k; ++c)
            for (int i = 0; i < n; ++i) {
                double s = B0.Get(i, c);
                for (int j = 0; j < n; ++j) s -= A0.Get(i, j) * Bu.Get(j, c);
                r2 += s * s;
            }
        bool lower_ok = true;
        for (int j = 0; j < n; ++j)
            for (int i = j + 1; i < n; ++i) lower_ok &= (Au.Get(i, j) == A0.Get(i, j));
        for (int j = 0; j < n; ++j)
            for (int i = 0; i <= j; ++i) {
                double s = 0;
                for (int l = 0; l <= i; ++l) s += Au.Get(l, i) * Au.Get(l, j);
                utu += (s - A0.Get(i, j)) * (s - A0.Get(i, j));
            }
        const double relu = std::sqrt(r2 / bn), du = std::sqrt(utu / an);
        std::printf("HPDSolve(UPPER): rel residual %.3e, ||U^T U-A||/||A|| %.3e, lower untouched %d\n", relu, du,
                    (int)lower_ok);
        fails += !(relu < 1e-13 && du < 1e-14 && lower_ok);
    }

    // Matrix<float>: HPDSolve<float>
    {
        elem::Matrix<float> Af(n, n), Bf(n, k);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) Af.Set(i, j, (float)A0.Get(i, j));
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < n; ++i) Bf.Set(i, j, (float)B0.Get(i, j));
        elem::Matrix<float> Af0(Af), Bf0(Bf);
        elem::HPDSolve(elem::LOWER, elem::NORMAL, Af, Bf);
        double r2 = 0, b2 = 0;
        for (int c = 0; c < k; ++c)
            for (int i = 0; i < n; ++i) {
                double s = Bf0.Get(i, c);
                for (int j = 0; j < n; ++j) s -= (double)Af0.Get(i, j) * Bf.Get(j, c);
                r2 += s * s;
                b2 += (double)Bf0.Get(i, c) * Bf0.Get(i, c);
            }
        const double relf = std::sqrt(r2 / b2);
        std::printf("HPDSolve<float>: rel residual %.3e\n", relf);
        fails += !(relf < 1e-5);
    }

    // HPDInverse: inv(A) on the lower triangle, the upper one untouched
    {
        const int m = 300;
        elem::Matrix<double> H(m, m), H0;
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) H.Set(i, j, A0.Get(i, j));
        H0 = H;
        elem::HPDInverse(elem::LOWER, H);
        double worst = 0;
        bool upper_ok = true;
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                double s = 0;  // (inv(A) A)(i, j) from the lower triangle of inv(A)
                for (int l = 0; l < m; ++l) s += (i >= l ? H.Get(i, l) : H.Get(l, i)) * H0.Get(l, j);
                worst = std::fmax(worst, std::fabs(s - (i == j ? 1.0 : 0.0)));
                if (i < j && H.Get(i, j) != H0.Get(i, j)) upper_ok = false;
            }
        std::printf("HPDInverse n=%d: max |inv(A) A - I| = %.3e, upper untouched %d\n", m, worst, (int)upper_ok);
        if (!(worst < 1e-10) || !upper_ok) ++fails;
    }
    // HermitianGenDefiniteEig(AXBX): A x = lambda B x for every pair
    {
        const int m = 120;
        elem::Matrix<double> S(m, m), P(m, m), w, X;
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                S.Set(i, j, G.Get(i, j) + G.Get(j, i));
                P.Set(i, j, A0.Get(i, j));
            }
        elem::Matrix<double> S0(S), P0(P);
        elem::HermitianGenDefiniteEig(elem::AXBX, elem::LOWER, S, P, w, X);
        double worst = 0;
        for (int c = 0; c < w.Height(); ++c)
            for (int i = 0; i < m; ++i) {
                double ax = 0, bx = 0;
                for (int l = 0; l < m; ++l) {
                    ax += S0.Get(i, l) * X.Get(l, c);
                    bx += P0.Get(i, l) * X.Get(l, c);
                }
                worst = std::fmax(worst, std::fabs(ax - w.Get(c, 0) * bx));
            }
        std::printf("HermitianGenDefiniteEig n=%d: %d eigenpairs, max |Ax - wBx| = %.3e\n", m, w.Height(), worst);
        if (w.Height() != m || !(worst < 1e-9)) ++fails;
    }
    std::printf("%s\n", fails ? "FAILED" : "ALL OK");
    return fails ? 1 : 0;
}
