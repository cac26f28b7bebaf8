// Cholesky test driver with the command line and the report of Elemental's
// tests/lapack-like/Cholesky.cpp:21-191, run through the C++ drop-in
// (elem::Cholesky on one MI355X) so the two programs' reports can be diffed.
//
//   --gridHeight  process grid height (accepted; one GPU is one 1 x 1 grid)
//   --uplo        upper or lower storage: L/U
//   --height      height of matrix
//   --nb          algorithmic blocksize
//   --nbLocal     local blocksize (accepted; the local Trrk blocking of the
//                 distributed route, which one GPU does not take)
//   --correctness test correctness?
//   --print       print matrices?
//
// The matrix is HermitianUniformSpectrum(A, m, 1, 10)
// (include/elemental/matrices/HermitianUniformSpectrum.hpp): a diagonal with
// entries drawn from (1, 10] rotated by one random Householder reflection,
// here from the splitmix64 stream of the library's generator.  The
// correctness check applies A and L L^H (or U^H U) to 100 random vectors and
// reports the 1-, infinity- and Frobenius norms of the difference, as the
// reference does (TestCorrectness, Cholesky.cpp:21-88).
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <typeinfo>
#include <vector>

#include "../../../include/elemental_amd/elem.hpp"

namespace {

// ----------------------------------------------------------- arguments
// Input(name, desc, default) / ProcessInput / PrintInputReport with the
// reference's report format (core/imports/choice.hpp:153-270).
struct Args {
    int argc;
    char** argv;
    std::vector<bool> used;
    struct Opt {
        std::string name, desc, type, def, val;
        bool found;
    };
    std::vector<Opt> opts;
    Args(int c, char** v) : argc(c), argv(v), used(c, false) {}

    template <typename T>
    T input(const std::string& name, const std::string& desc, T def) {
        int at = -1;
        for (int i = 1; i < argc; ++i)
            if (name == argv[i]) {
                at = i;
                break;
            }
        if (at == argc - 1) throw std::runtime_error("Missing value for last command-line argument");
        std::ostringstream ds;
        ds << def;
        Opt o{name, desc, typeid(T).name(), ds.str(), at > 0 ? argv[at + 1] : ds.str(), at > 0};
        opts.push_back(o);
        if (at < 0) return def;
        used[at] = used[at + 1] = true;
        return parse<T>(argv[at + 1]);
    }
    template <typename T>
    static T parse(const std::string& s) {
        std::istringstream is(s);
        T v{};
        is >> v;
        return v;
    }
    bool help() const {
        for (int i = 1; i < argc; ++i)
            if (std::string(argv[i]) == "--help") return true;
        return false;
    }
    void report(std::ostream& os) const {
        if (!opts.empty()) os << "Optional arguments:\n";
        int notfound = 0;
        for (const Opt& o : opts) {
            notfound += !o.found;
            os << "  " << o.name << " [" << o.type << "," << o.def << "," << o.val << ","
               << (o.found ? "found" : "NOT found") << "]\n"
               << "    " << o.desc << "\n\n";
        }
        os << "Out of 0 required arguments, 0 were not specified." << std::endl;
        os << "Out of " << opts.size() << " optional arguments, " << notfound << " were not specified.\n" << std::endl;
    }
};
template <>
bool Args::parse<bool>(const std::string& s) {
    if (s == "true") return true;
    if (s == "false") return false;
    std::istringstream is(s);
    bool v = false;
    is >> v;
    return v;
}

// ----------------------------------------------------- the test matrix
uint64_t sm64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double u01(uint64_t& s) { return ((double)(sm64(s) >> 11) + 1.0) * 0x1.0p-53; }

// (I - 2uu^T) D (I - 2uu^T) = D - 2(u (Du)^T + (Du) u^T) + 4 (u^T D u) u u^T
void hermitian_uniform_spectrum(elem::Matrix<double>& A, int n, double lower, double upper, uint64_t seed) {
    std::vector<double> d(n), u(n), du(n);
    for (int j = 0; j < n; ++j) d[j] = lower + (upper - lower) * u01(seed);
    double nrm = 0.0;
    for (int j = 0; j < n; ++j) {
        u[j] = 2.0 * u01(seed) - 1.0;
        nrm += u[j] * u[j];
    }
    nrm = std::sqrt(nrm);
    double gamma = 0.0;
    for (int j = 0; j < n; ++j) {
        u[j] /= nrm;
        du[j] = d[j] * u[j];
        gamma += u[j] * du[j];
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            A.Set(i, j, (i == j ? d[i] : 0.0) - 2.0 * (u[i] * du[j] + du[i] * u[j]) + 4.0 * gamma * u[i] * u[j]);
}

// ------------------------------------------------------------- norms
double one_norm(const elem::Matrix<double>& A) {
    double m = 0.0;
    for (int j = 0; j < A.Width(); ++j) {
        double s = 0.0;
        for (int i = 0; i < A.Height(); ++i) s += std::fabs(A.Get(i, j));
        m = std::fmax(m, s);
    }
    return m;
}
double inf_norm(const elem::Matrix<double>& A) {
    std::vector<double> r(A.Height(), 0.0);
    for (int j = 0; j < A.Width(); ++j)
        for (int i = 0; i < A.Height(); ++i) r[i] += std::fabs(A.Get(i, j));
    double m = 0.0;
    for (double v : r) m = std::fmax(m, v);
    return m;
}
double frob_norm(const elem::Matrix<double>& A) {
    double s = 0.0;
    for (int j = 0; j < A.Width(); ++j)
        for (int i = 0; i < A.Height(); ++i) s += A.Get(i, j) * A.Get(i, j);
    return std::sqrt(s);
}
// the Hermitian matrix whose `lower` (or upper) triangle is stored
double sym_entry(const elem::Matrix<double>& A, bool lower, int i, int j) {
    return (lower ? i >= j : i <= j) ? A.Get(i, j) : A.Get(j, i);
}

// Y := X; Y := L^T Y; Y := L Y (LOWER) or Y := U Y; Y := U^T Y (UPPER);
// then Y := Y - A X (Trmm, Trmm, Hemm of Cholesky.cpp:37-41, 66-70).
void test_correctness(bool lower, const elem::Matrix<double>& F, const elem::Matrix<double>& A0, uint64_t seed) {
    const int m = A0.Height(), nv = 100;
    elem::Matrix<double> X(m, nv), Y(m, nv), Z(m, nv);
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < m; ++i) X.Set(i, j, 2.0 * u01(seed) - 1.0);
    // T(i, j) of the triangular factor as stored
    auto L = [&](int i, int j) { return lower ? (i >= j ? F.Get(i, j) : 0.0) : (i >= j ? F.Get(j, i) : 0.0); };
    for (int c = 0; c < nv; ++c) {
        for (int i = 0; i < m; ++i) {  // Z = L^T X (= U X)
            double s = 0.0;
            for (int l = i; l < m; ++l) s += L(l, i) * X.Get(l, c);
            Z.Set(i, c, s);
        }
        for (int i = 0; i < m; ++i) {  // Y = L Z (= U^T Z)
            double s = 0.0;
            for (int l = 0; l <= i; ++l) s += L(i, l) * Z.Get(l, c);
            Y.Set(i, c, s);
        }
        for (int i = 0; i < m; ++i) {  // Y -= A X
            double s = 0.0;
            for (int l = 0; l < m; ++l) s += sym_entry(A0, lower, i, l) * X.Get(l, c);
            Y.Set(i, c, Y.Get(i, c) - s);
        }
    }
    double an = 0.0, af = 0.0;
    for (int i = 0; i < m; ++i) {
        double r = 0.0;
        for (int j = 0; j < m; ++j) {
            const double v = sym_entry(A0, lower, i, j);
            r += std::fabs(v);
            af += v * v;
        }
        an = std::fmax(an, r);
    }
    af = std::sqrt(af);
    const char* prod = lower ? "L L^H" : "U^H U";
    std::cout << "||A||_1 = ||A||_oo   = " << an << "\n"
              << "||A||_F              = " << af << "\n"
              << "||X||_1              = " << one_norm(X) << "\n"
              << "||X||_oo             = " << inf_norm(X) << "\n"
              << "||X||_F              = " << frob_norm(X) << "\n"
              << "||A X - " << prod << " X||_1  = " << one_norm(Y) << "\n"
              << "||A X - " << prod << " X||_oo = " << inf_norm(Y) << "\n"
              << "||A X - " << prod << " X||_F  = " << frob_norm(Y) << std::endl;
}

void test_cholesky(bool correctness, bool print, elem::UpperOrLower uplo, int m, int nb) {
    elem::Matrix<double> A(m, m), A0;
    hermitian_uniform_spectrum(A, m, 1.0, 10.0, 0x5eed);
    if (correctness) {
        std::cout << "  Making copy of original matrix...";
        std::cout.flush();
        A0 = A;
        std::cout << "DONE" << std::endl;
    }
    if (print) elem::Print(A, "A");
    std::cout << "  Starting Cholesky factorization...";
    std::cout.flush();
    // the timed region: the drop-in call (host Matrix staged to HBM and back)
    const auto t0 = std::chrono::steady_clock::now();
    elem::Cholesky(uplo, A, nb);
    const double run = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double gflops = 1. / 3. * std::pow(double(m), 3.) / (1.e9 * run);
    std::cout << "DONE.\n"
              << "  Time = " << run << " seconds. GFlops = " << gflops << std::endl;
    if (print) elem::Print(A, "A after factorization");
    if (correctness) test_correctness(uplo == elem::LOWER, A, A0, 0xc0ffee);
}

}  // namespace

int main(int argc, char* argv[]) {
    try {
        Args a(argc, argv);
        const int r = a.input<int>("--gridHeight", "process grid height", 0);
        const char uploChar = a.input<char>("--uplo", "upper or lower storage: L/U", 'L');
        const int m = a.input<int>("--height", "height of matrix", 100);
        const int nb = a.input<int>("--nb", "algorithmic blocksize", 96);
        const int nbLocal = a.input<int>("--nbLocal", "local blocksize", 32);
        const bool testCorrectness = a.input<bool>("--correctness", "test correctness?", true);
        const bool printMatrices = a.input<bool>("--print", "print matrices?", false);
        if (a.help()) {
            a.report(std::cout);
            return 0;
        }
        a.report(std::cout);
        (void)r;
        (void)nbLocal;
        const elem::UpperOrLower uplo = (uploChar == 'U' || uploChar == 'u') ? elem::UPPER : elem::LOWER;
        if (nb < 1 || nb > 128) throw std::logic_error("--nb must be in 1..128 on this path");
        std::cout << "Will test Cholesky" << uploChar << std::endl;
        std::cout << "---------------------\n"
                  << "Testing with doubles:\n"
                  << "---------------------" << std::endl;
        test_cholesky(testCorrectness, printMatrices, uplo, m, nb);
        std::cout << "--------------------------------------\n"
                  << "Testing with double-precision complex:\n"
                  << "--------------------------------------" << std::endl;
        std::cout << "  (complex arithmetic is not on this path: skipped)" << std::endl;
    } catch (std::exception& e) {
        std::cout << "Process 0 caught error message:\n" << e.what() << std::endl;
        return 1;
    }
    return 0;
}
