// elem.hpp -- C++ drop-in for Elemental's HPD Cholesky / HPDSolve path
// on one MI355X.  Same call signatures and column-major Matrix<F> semantics as
// the reference (names below cite aterrel/Elemental); the work runs on the GPU
// through the C ABI in elemental_amd.h.
//
//   elem::Matrix<T>                  include/elemental/core/matrix.hpp:16-183
//   elem::Cholesky(uplo, A)          include/elemental/lapack-like/Cholesky.hpp:49-62 (LVar3 / UVar3)
//   elem::LocalCholesky(uplo, A)     include/elemental/lapack-like/Cholesky.hpp:29-37
//   elem::HPDSolve(uplo, orient, A, B)  include/elemental/lapack-like/HPDSolve.hpp:18-28
//   elem::cholesky::SolveAfter(...)  include/elemental/lapack-like/Cholesky/SolveAfter.hpp:20-75
//   elem::Trsm(...)                  include/elemental/blas-like/level3/Trsm.hpp:49-86
//   elem::Herk / elem::Syrk(...)     include/elemental/blas-like/level3/Herk.hpp:17-27, Syrk.hpp:20-57
//   elem::Trrk(...)                  src/blas-like/Trrk.cpp:19-38
//   elem::Trsv(...)                  include/elemental/blas-like/level2/Trsv.hpp:20-44
//   elem::Print / elem::Write        include/elemental/io/Print.hpp:15-35, io/Write.hpp:17-28
//   elem::Gemm(...)                  include/elemental/blas-like/level3/Gemm.hpp
//   elem::HPDInverse(uplo, A)        include/elemental/lapack-like/Inverse.hpp:29-40 (CholeskyLVar2)
//   elem::TriangularInverse(...)     include/elemental/lapack-like/TriangularInverse.hpp (LVar3)
//   elem::HermitianGenDefiniteEig    include/elemental/lapack-like/HermitianGenDefiniteEig.hpp:22-250
//
// A host Matrix is staged to device memory for the call and copied back, so
// existing callers keep their buffers.  Callers that keep data on the GPU use
// the C ABI directly (device pointers, a stream, no copies).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../elemental_amd.h"

namespace elem {

enum UpperOrLower { LOWER, UPPER };              // core/types_decl.hpp:152-156
enum Orientation { NORMAL, TRANSPOSE, ADJOINT };  // core/types_decl.hpp:129-134
enum LeftOrRight { LEFT, RIGHT };                 // core/types_decl.hpp:104-108
enum UnitOrNonUnit { NON_UNIT, UNIT };            // core/types_decl.hpp:141-145

// Thrown like the reference's SingularMatrixException (Trsm.hpp:76-82).
class SingularMatrixException : public std::logic_error {
   public:
    SingularMatrixException() : std::logic_error("Matrix was singular") {}
};

// Column-major matrix with a leading dimension; owning, or a view of a caller
// buffer (views never free).  Element (i, j) is Buffer()[i + j*LDim()].
template <typename T>
class Matrix {
   public:
    Matrix() : h_(0), w_(0), ld_(1), data_(nullptr) {}
    Matrix(int height, int width) : h_(height), w_(width), ld_(height > 1 ? height : 1) {
        own_.assign((size_t)ld_ * (w_ > 0 ? w_ : 0), T(0));
        data_ = own_.data();
    }
    Matrix(int height, int width, int ldim) : h_(height), w_(width), ld_(ldim) {
        if (ldim < (height > 1 ? height : 1)) throw std::logic_error("Leading dimension must be no less than height");
        own_.assign((size_t)ld_ * (w_ > 0 ? w_ : 0), T(0));
        data_ = own_.data();
    }
    // view of a caller buffer
    Matrix(int height, int width, T* buffer, int ldim) : h_(height), w_(width), ld_(ldim), data_(buffer) {}
    Matrix(const Matrix& A) : h_(A.h_), w_(A.w_), ld_(A.h_ > 1 ? A.h_ : 1) {
        own_.assign((size_t)ld_ * (w_ > 0 ? w_ : 0), T(0));
        data_ = own_.data();
        for (int j = 0; j < w_; ++j)
            for (int i = 0; i < h_; ++i) data_[i + (size_t)j * ld_] = A.Get(i, j);
    }
    Matrix& operator=(const Matrix& A) {
        if (this != &A) {
            Matrix tmp(A);
            std::swap(h_, tmp.h_);
            std::swap(w_, tmp.w_);
            std::swap(ld_, tmp.ld_);
            own_.swap(tmp.own_);
            data_ = own_.data();
        }
        return *this;
    }

    int Height() const { return h_; }
    int Width() const { return w_; }
    int LDim() const { return ld_; }
    T* Buffer() { return data_; }
    const T* LockedBuffer() const { return data_; }
    T* Buffer(int i, int j) { return data_ + i + (size_t)j * ld_; }
    const T* LockedBuffer(int i, int j) const { return data_ + i + (size_t)j * ld_; }
    // Attach (matrix.hpp: Attach) makes this Matrix a view of a caller buffer.
    void Attach(int height, int width, T* buffer, int ldim) {
        own_.clear();
        h_ = height;
        w_ = width;
        ld_ = ldim;
        data_ = buffer;
    }
    T Get(int i, int j) const { return data_[i + (size_t)j * ld_]; }
    void Set(int i, int j, T v) { data_[i + (size_t)j * ld_] = v; }
    void Update(int i, int j, T v) { data_[i + (size_t)j * ld_] += v; }

   private:
    int h_, w_, ld_;
    T* data_;
    std::vector<T> own_;
};

// View(B, A, i, j, height, width): B views a submatrix of A.
template <typename T>
inline void View(Matrix<T>& B, Matrix<T>& A, int i, int j, int height, int width) {
    if (i < 0 || j < 0 || i + height > A.Height() || j + width > A.Width())
        throw std::logic_error("Trying to view outside of a Matrix");
    B.Attach(height, width, A.Buffer(i, j), A.LDim());
}

// Print(A, title, os) (io/Print.hpp:15-35): the title line, then each row
// as its entries with the stream's current format, each followed by a
// space, then an empty line.
template <typename T>
inline void Print(const Matrix<T>& A, std::string title = "", std::ostream& os = std::cout) {
    if (title != "") os << title << std::endl;
    for (int i = 0; i < A.Height(); ++i) {
        for (int j = 0; j < A.Width(); ++j) os << A.Get(i, j) << " ";
        os << std::endl;
    }
    os << std::endl;
}

// Write(A, title, filename) (io/Write.hpp:17-28): Print into the file
// `filename` in scientific notation.
template <typename T>
inline void Write(const Matrix<T>& A, std::string title = "", std::string filename = "Matrix") {
    std::ofstream file(filename.c_str());
    file.setf(std::ios::scientific);
    Print(A, title, file);
    file.close();
}

namespace detail {

inline void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
inline void check(int rc, const char* what) {
    if (rc != ELEMAMD_OK) throw std::runtime_error(std::string(what) + " failed with status " + std::to_string(rc));
}
inline char uplo_char(UpperOrLower u) { return u == LOWER ? 'L' : 'U'; }
inline char orient_char(Orientation o) { return o == NORMAL ? 'N' : (o == TRANSPOSE ? 'T' : 'C'); }

// Device copy of a host Matrix (packed, ld = max(height, 1)); copies back on demand.
template <typename T>
struct Dev {
    T* p = nullptr;
    int h = 0, w = 0, ld = 1;
    explicit Dev(const Matrix<T>& A) : h(A.Height()), w(A.Width()), ld(A.Height() > 1 ? A.Height() : 1) {
        size_t bytes = sizeof(T) * (size_t)ld * (w > 0 ? w : 1);
        check_hip(hipMalloc(&p, bytes), "hipMalloc");
        if (h > 0 && w > 0)
            check_hip(hipMemcpy2D(p, sizeof(T) * ld, A.LockedBuffer(), sizeof(T) * A.LDim(), sizeof(T) * h, w,
                                  hipMemcpyHostToDevice),
                      "hipMemcpy2D");
    }
    void back(Matrix<T>& A) const {
        if (h > 0 && w > 0)
            check_hip(hipMemcpy2D(A.Buffer(), sizeof(T) * A.LDim(), p, sizeof(T) * ld, sizeof(T) * h, w,
                                  hipMemcpyDeviceToHost),
                      "hipMemcpy2D");
    }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
};

struct DevInfo {
    int* p = nullptr;
    DevInfo() {
        check_hip(hipMalloc(&p, sizeof(int)), "hipMalloc");
        check_hip(hipMemset(p, 0, sizeof(int)), "hipMemset");
    }
    ~DevInfo() {
        if (p) (void)hipFree(p);
    }
    // cholesky::LVar3Unb / UVar3Unb throw this at the failing column
    // (LVar3.hpp:37-38, UVar3.hpp:38-39)
    void raise_if_failed() const {
        int info = elemamd_info_read(p, nullptr);
        if (info > 0) throw std::logic_error("A was not numerically HPD");
        if (info < 0) throw std::runtime_error("reading the factorization status failed");
    }
};

// the C ABI entry points by field
inline int cholesky_abi(char u, int n, double* A, int64_t lda, int nb, int* info) {
    return elemamd_cholesky_f64(u, n, A, lda, nb, info, nullptr);
}
inline int cholesky_abi(char u, int n, float* A, int64_t lda, int nb, int* info) {
    return elemamd_cholesky_f32(u, n, A, lda, nb, info, nullptr);
}
inline int solve_after_abi(char u, char o, int n, int k, const double* A, int64_t lda, double* B, int64_t ldb) {
    return elemamd_solve_after_f64(u, o, n, k, A, lda, B, ldb, nullptr);
}
inline int solve_after_abi(char u, char o, int n, int k, const float* A, int64_t lda, float* B, int64_t ldb) {
    return elemamd_solve_after_f32(u, o, n, k, A, lda, B, ldb, nullptr);
}
inline int hpd_solve_abi(char u, char o, int n, int k, double* A, int64_t lda, double* B, int64_t ldb, int nb,
                         int* info) {
    return elemamd_hpd_solve_f64(u, o, n, k, A, lda, B, ldb, nb, info, nullptr);
}
inline int hpd_solve_abi(char u, char o, int n, int k, float* A, int64_t lda, float* B, int64_t ldb, int nb,
                         int* info) {
    return elemamd_hpd_solve_f32(u, o, n, k, A, lda, B, ldb, nb, info, nullptr);
}

}  // namespace detail

inline int Blocksize() { return 128; }  // src/core/global.cpp:194

// F = double or float (Cholesky.hpp:47-62 templates on the field).
template <typename F>
inline void Cholesky(UpperOrLower uplo, Matrix<F>& A, int nb = Blocksize()) {
    if (A.Height() != A.Width()) throw std::logic_error("A must be square");
    detail::Dev<F> d(A);
    detail::DevInfo info;
    detail::check(detail::cholesky_abi(detail::uplo_char(uplo), A.Height(), d.p, d.ld, nb, info.p),
                  "elemamd_cholesky");
    d.back(A);
    info.raise_if_failed();
}

template <typename F>
inline void LocalCholesky(UpperOrLower uplo, Matrix<F>& A) {
    Cholesky(uplo, A);
}

namespace cholesky {
template <typename F>
inline void SolveAfter(UpperOrLower uplo, Orientation orientation, const Matrix<F>& A, Matrix<F>& B) {
    if (A.Height() != A.Width()) throw std::logic_error("A must be square");
    if (A.Height() != B.Height()) throw std::logic_error("A and B must be the same height");
    detail::Dev<F> dA(A), dB(B);
    detail::check(detail::solve_after_abi(detail::uplo_char(uplo), detail::orient_char(orientation), A.Height(),
                                          B.Width(), dA.p, dA.ld, dB.p, dB.ld),
                  "elemamd_solve_after");
    dB.back(B);
}
}  // namespace cholesky

template <typename F>
inline void HPDSolve(UpperOrLower uplo, Orientation orientation, Matrix<F>& A, Matrix<F>& B, int nb = Blocksize()) {
    if (A.Height() != A.Width()) throw std::logic_error("A must be square");
    if (A.Height() != B.Height()) throw std::logic_error("A and B must be the same height");
    detail::Dev<F> dA(A), dB(B);
    detail::DevInfo info;
    detail::check(detail::hpd_solve_abi(detail::uplo_char(uplo), detail::orient_char(orientation), A.Height(),
                                        B.Width(), dA.p, dA.ld, dB.p, dB.ld, nb, info.p),
                  "elemamd_hpd_solve");
    dA.back(A);
    dB.back(B);
    info.raise_if_failed();
}

inline void Trsm(LeftOrRight side, UpperOrLower uplo, Orientation orientation, UnitOrNonUnit diag, double alpha,
                 const Matrix<double>& A, Matrix<double>& B, bool checkIfSingular = false) {
    if (A.Height() != A.Width()) throw std::logic_error("Triangular matrix must be square");
    if ((side == LEFT && A.Height() != B.Height()) || (side == RIGHT && A.Height() != B.Width()))
        throw std::logic_error("Nonconformal Trsm");
    if (checkIfSingular && diag != UNIT)
        for (int j = 0; j < A.Height(); ++j)
            if (A.Get(j, j) == 0.0) throw SingularMatrixException();
    detail::Dev<double> dA(A), dB(B);
    detail::check(elemamd_trsm_f64(side == LEFT ? 'L' : 'R', detail::uplo_char(uplo), detail::orient_char(orientation),
                                   diag == NON_UNIT ? 'N' : 'U', B.Height(), B.Width(), alpha, dA.p, dA.ld, dB.p, dB.ld,
                                   nullptr),
                  "elemamd_trsm_f64");
    dB.back(B);
}

// C := alpha op(A) op(A)^T + beta C on the uplo triangle; op(A) = A (NORMAL,
// A n x k) or A^T (A k x n).
inline void Herk(UpperOrLower uplo, Orientation orientation, double alpha, const Matrix<double>& A, double beta,
                 Matrix<double>& C) {
    const int n = orientation == NORMAL ? A.Height() : A.Width();
    const int k = orientation == NORMAL ? A.Width() : A.Height();
    if (n != C.Height() || C.Height() != C.Width()) throw std::logic_error("Nonconformal Syrk");
    detail::Dev<double> dA(A), dC(C);
    detail::check(elemamd_herk_f64(detail::uplo_char(uplo), detail::orient_char(orientation), n, k, alpha, dA.p,
                                   dA.ld, beta, dC.p, dC.ld, nullptr),
                  "elemamd_herk_f64");
    dC.back(C);
}

inline void Syrk(UpperOrLower uplo, Orientation orientation, double alpha, const Matrix<double>& A, double beta,
                 Matrix<double>& C) {
    Herk(uplo, orientation, alpha, A, beta, C);
}

// The uplo triangle of C := alpha op(A) op(B) + beta C, op(A) n x k, op(B) k x n.
inline void Trrk(UpperOrLower uplo, Orientation orientationOfA, Orientation orientationOfB, double alpha,
                 const Matrix<double>& A, const Matrix<double>& B, double beta, Matrix<double>& C) {
    const int n = C.Height();
    const int k = orientationOfA == NORMAL ? A.Width() : A.Height();
    if (C.Width() != n) throw std::logic_error("C must be square");
    detail::Dev<double> dA(A), dB(B), dC(C);
    detail::check(elemamd_trrk_f64(detail::uplo_char(uplo), detail::orient_char(orientationOfA),
                                   detail::orient_char(orientationOfB), n, k, alpha, dA.p, dA.ld, dB.p, dB.ld, beta,
                                   dC.p, dC.ld, nullptr),
                  "elemamd_trrk_f64");
    dC.back(C);
}

inline void Trsv(UpperOrLower uplo, Orientation orientation, UnitOrNonUnit diag, const Matrix<double>& A,
                 Matrix<double>& x) {
    if (x.Width() != 1) throw std::logic_error("x must be a column vector here");
    detail::Dev<double> dA(A), dx(x);
    detail::check(elemamd_trsv_f64(detail::uplo_char(uplo), detail::orient_char(orientation),
                                   diag == NON_UNIT ? 'N' : 'U', A.Height(), dA.p, dA.ld, dx.p, 1, nullptr),
                  "elemamd_trsv_f64");
    dx.back(x);
}

// C := alpha op(A) op(B) + beta C.
inline void Gemm(Orientation orientationOfA, Orientation orientationOfB, double alpha, const Matrix<double>& A,
                 const Matrix<double>& B, double beta, Matrix<double>& C) {
    const int k = orientationOfA == NORMAL ? A.Width() : A.Height();
    const int kb = orientationOfB == NORMAL ? B.Height() : B.Width();
    const int m = orientationOfA == NORMAL ? A.Height() : A.Width();
    const int n = orientationOfB == NORMAL ? B.Width() : B.Height();
    if (k != kb || m != C.Height() || n != C.Width()) throw std::logic_error("Nonconformal Gemm");
    detail::Dev<double> dA(A), dB(B), dC(C);
    detail::check(elemamd_gemm_f64(detail::orient_char(orientationOfA), detail::orient_char(orientationOfB), m, n, k,
                                   alpha, dA.p, dA.ld, dB.p, dB.ld, beta, dC.p, dC.ld, nullptr),
                  "elemamd_gemm_f64");
    dC.back(C);
}

// The uplo triangle of A := inv(A) for HPD A; throws like Cholesky.
inline void HPDInverse(UpperOrLower uplo, Matrix<double>& A, int nb = Blocksize()) {
    if (A.Height() != A.Width()) throw std::logic_error("Nonsquare matrices cannot be triangular");
    detail::Dev<double> d(A);
    detail::DevInfo info;
    detail::check(elemamd_hpd_inverse_f64(detail::uplo_char(uplo), A.Height(), d.p, d.ld, nb, info.p, nullptr),
                  "elemamd_hpd_inverse_f64");
    d.back(A);
    info.raise_if_failed();
}

inline void TriangularInverse(UpperOrLower uplo, UnitOrNonUnit diag, Matrix<double>& A, int nb = Blocksize()) {
    if (A.Height() != A.Width()) throw std::logic_error("Nonsquare matrices cannot be triangular");
    detail::Dev<double> d(A);
    detail::check(elemamd_triangular_inverse_f64(detail::uplo_char(uplo), diag == NON_UNIT ? 'N' : 'U', A.Height(),
                                                 d.p, d.ld, nb, nullptr),
                  "elemamd_triangular_inverse_f64");
    d.back(A);
}

// lapack-like_decl.hpp: hermitian_gen_definite_eig_type_wrapper
enum HermitianGenDefiniteEigType { AXBX = 1, ABX = 2, BAX = 3 };

namespace detail {
// w (nev x 1) and, when X is not null, X (n x nev); range as in the C ABI.
inline void gen_eig(HermitianGenDefiniteEigType type, UpperOrLower uplo, Matrix<double>& A, Matrix<double>& B,
                    Matrix<double>& w, Matrix<double>* X, int range, int il, int iu, double vl, double vu) {
    const int n = A.Height();
    if (A.Width() != n || B.Height() != B.Width() || B.Height() != n)
        throw std::logic_error("Hermitian matrices must be square.");
    Dev<double> dA(A), dB(B);
    Matrix<double> wfull(n > 0 ? n : 1, 1), Xfull(n, n > 0 ? n : 1);
    Dev<double> dw(wfull), dX(Xfull);
    int nev = 0, info = 0;
    check(elemamd_hermitian_gen_definite_eig_f64((int)type, uplo_char(uplo), n, dA.p, dA.ld, dB.p, dB.ld, dw.p,
                                                 X ? dX.p : nullptr, dX.ld, range, il, iu, vl, vu, &nev, &info,
                                                 nullptr),
          "elemamd_hermitian_gen_definite_eig_f64");
    if (info > 0) throw std::logic_error("A was not numerically HPD");
    dA.back(A);
    dB.back(B);
    dw.back(wfull);
    w = Matrix<double>(nev, 1);
    for (int i = 0; i < nev; ++i) w.Set(i, 0, wfull.Get(i, 0));
    if (X) {
        dX.back(Xfull);
        *X = Matrix<double>(n, nev);
        for (int j = 0; j < nev; ++j)
            for (int i = 0; i < n; ++i) X->Set(i, j, Xfull.Get(i, j));
    }
}
}  // namespace detail

// All eigenvalues, then all eigenpairs (HermitianGenDefiniteEig.hpp:22-126).
inline void HermitianGenDefiniteEig(HermitianGenDefiniteEigType type, UpperOrLower uplo, Matrix<double>& A,
                                    Matrix<double>& B, Matrix<double>& w) {
    detail::gen_eig(type, uplo, A, B, w, nullptr, ELEMAMD_EIG_ALL, 0, 0, 0.0, 0.0);
}
inline void HermitianGenDefiniteEig(HermitianGenDefiniteEigType type, UpperOrLower uplo, Matrix<double>& A,
                                    Matrix<double>& B, Matrix<double>& w, Matrix<double>& X) {
    detail::gen_eig(type, uplo, A, B, w, &X, ELEMAMD_EIG_ALL, 0, 0, 0.0, 0.0);
}
// Eigenvalues a..b (0-based, inclusive) and the eigenvalues in (a, b].
inline void HermitianGenDefiniteEig(HermitianGenDefiniteEigType type, UpperOrLower uplo, Matrix<double>& A,
                                    Matrix<double>& B, Matrix<double>& w, int a, int b) {
    detail::gen_eig(type, uplo, A, B, w, nullptr, ELEMAMD_EIG_INDEX, a, b, 0.0, 0.0);
}
inline void HermitianGenDefiniteEig(HermitianGenDefiniteEigType type, UpperOrLower uplo, Matrix<double>& A,
                                    Matrix<double>& B, Matrix<double>& w, Matrix<double>& X, int a, int b) {
    detail::gen_eig(type, uplo, A, B, w, &X, ELEMAMD_EIG_INDEX, a, b, 0.0, 0.0);
}
inline void HermitianGenDefiniteEig(HermitianGenDefiniteEigType type, UpperOrLower uplo, Matrix<double>& A,
                                    Matrix<double>& B, Matrix<double>& w, double a, double b) {
    detail::gen_eig(type, uplo, A, B, w, nullptr, ELEMAMD_EIG_VALUE, 0, 0, a, b);
}
inline void HermitianGenDefiniteEig(HermitianGenDefiniteEigType type, UpperOrLower uplo, Matrix<double>& A,
                                    Matrix<double>& B, Matrix<double>& w, Matrix<double>& X, double a, double b) {
    detail::gen_eig(type, uplo, A, B, w, &X, ELEMAMD_EIG_VALUE, 0, 0, a, b);
}

}  // namespace elem
