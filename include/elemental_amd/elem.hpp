// elem.hpp -- C++ drop-in for Elemental's lower HPD Cholesky / HPDSolve path
// on one MI355X.  Same call signatures and column-major Matrix<F> semantics as
// the reference (names below cite aterrel/Elemental); the work runs on the GPU
// through the C ABI in elemental_amd.h.
//
//   elem::Matrix<T>                  include/elemental/core/matrix.hpp:16-183
//   elem::Cholesky(uplo, A)          include/elemental/lapack-like/Cholesky.hpp:49-62
//   elem::LocalCholesky(uplo, A)     include/elemental/lapack-like/Cholesky.hpp:29-37
//   elem::HPDSolve(uplo, orient, A, B)  include/elemental/lapack-like/HPDSolve.hpp:18-28
//   elem::cholesky::SolveAfter(...)  include/elemental/lapack-like/Cholesky/SolveAfter.hpp:20-75
//   elem::Trsm(...)                  include/elemental/blas-like/level3/Trsm.hpp:49-86
//   elem::Herk / elem::Syrk(...)     include/elemental/blas-like/level3/Herk.hpp:17-27, Syrk.hpp:20-57
//   elem::Trrk(...)                  src/blas-like/Trrk.cpp:19-38
//   elem::Trsv(...)                  include/elemental/blas-like/level2/Trsv.hpp:20-44
//
// A host Matrix is staged to device memory for the call and copied back, so
// existing callers keep their buffers.  Callers that keep data on the GPU use
// the C ABI directly (device pointers, a stream, no copies).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
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
struct Dev {
    double* p = nullptr;
    int h = 0, w = 0, ld = 1;
    explicit Dev(const Matrix<double>& A) : h(A.Height()), w(A.Width()), ld(A.Height() > 1 ? A.Height() : 1) {
        size_t bytes = sizeof(double) * (size_t)ld * (w > 0 ? w : 1);
        check_hip(hipMalloc(&p, bytes), "hipMalloc");
        if (h > 0 && w > 0)
            check_hip(hipMemcpy2D(p, sizeof(double) * ld, A.LockedBuffer(), sizeof(double) * A.LDim(),
                                  sizeof(double) * h, w, hipMemcpyHostToDevice),
                      "hipMemcpy2D");
    }
    void back(Matrix<double>& A) const {
        if (h > 0 && w > 0)
            check_hip(hipMemcpy2D(A.Buffer(), sizeof(double) * A.LDim(), p, sizeof(double) * ld, sizeof(double) * h, w,
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
    DevInfo() { check_hip(hipMalloc(&p, sizeof(int)), "hipMalloc"); }
    ~DevInfo() {
        if (p) (void)hipFree(p);
    }
    // cholesky::LVar3Unb throws this at the failing column (LVar3.hpp:37-38)
    void raise_if_failed() const {
        int info = elemamd_info_read(p, nullptr);
        if (info > 0) throw std::logic_error("A was not numerically HPD");
        if (info < 0) throw std::runtime_error("reading the factorization status failed");
    }
};

}  // namespace detail

inline int Blocksize() { return 128; }  // src/core/global.cpp:194

inline void Cholesky(UpperOrLower uplo, Matrix<double>& A, int nb = Blocksize()) {
    if (A.Height() != A.Width()) throw std::logic_error("A must be square");
    if (uplo != LOWER) throw std::logic_error("only LOWER is supported on this path");
    detail::Dev d(A);
    detail::DevInfo info;
    detail::check(elemamd_cholesky_f64('L', A.Height(), d.p, d.ld, nb, info.p, nullptr), "elemamd_cholesky_f64");
    d.back(A);
    info.raise_if_failed();
}

inline void LocalCholesky(UpperOrLower uplo, Matrix<double>& A) { Cholesky(uplo, A); }

namespace cholesky {
inline void SolveAfter(UpperOrLower uplo, Orientation orientation, const Matrix<double>& A, Matrix<double>& B) {
    if (A.Height() != A.Width()) throw std::logic_error("A must be square");
    if (A.Height() != B.Height()) throw std::logic_error("A and B must be the same height");
    if (uplo != LOWER) throw std::logic_error("only LOWER is supported on this path");
    detail::Dev dA(A), dB(B);
    detail::check(elemamd_solve_after_f64('L', detail::orient_char(orientation), A.Height(), B.Width(), dA.p, dA.ld,
                                          dB.p, dB.ld, nullptr),
                  "elemamd_solve_after_f64");
    dB.back(B);
}
}  // namespace cholesky

inline void HPDSolve(UpperOrLower uplo, Orientation orientation, Matrix<double>& A, Matrix<double>& B,
                     int nb = Blocksize()) {
    if (A.Height() != A.Width()) throw std::logic_error("A must be square");
    if (A.Height() != B.Height()) throw std::logic_error("A and B must be the same height");
    if (uplo != LOWER) throw std::logic_error("only LOWER is supported on this path");
    detail::Dev dA(A), dB(B);
    detail::DevInfo info;
    detail::check(elemamd_hpd_solve_f64('L', detail::orient_char(orientation), A.Height(), B.Width(), dA.p, dA.ld,
                                        dB.p, dB.ld, nb, info.p, nullptr),
                  "elemamd_hpd_solve_f64");
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
    detail::Dev dA(A), dB(B);
    detail::check(elemamd_trsm_f64(side == LEFT ? 'L' : 'R', detail::uplo_char(uplo), detail::orient_char(orientation),
                                   diag == NON_UNIT ? 'N' : 'U', B.Height(), B.Width(), alpha, dA.p, dA.ld, dB.p, dB.ld,
                                   nullptr),
                  "elemamd_trsm_f64");
    dB.back(B);
}

inline void Herk(UpperOrLower uplo, Orientation orientation, double alpha, const Matrix<double>& A, double beta,
                 Matrix<double>& C) {
    if (A.Height() != C.Height() || C.Height() != C.Width()) throw std::logic_error("Nonconformal Syrk");
    detail::Dev dA(A), dC(C);
    detail::check(elemamd_herk_f64(detail::uplo_char(uplo), detail::orient_char(orientation), C.Height(), A.Width(),
                                   alpha, dA.p, dA.ld, beta, dC.p, dC.ld, nullptr),
                  "elemamd_herk_f64");
    dC.back(C);
}

inline void Syrk(UpperOrLower uplo, Orientation orientation, double alpha, const Matrix<double>& A, double beta,
                 Matrix<double>& C) {
    Herk(uplo, orientation, alpha, A, beta, C);
}

inline void Trrk(UpperOrLower uplo, Orientation orientationOfA, Orientation orientationOfB, double alpha,
                 const Matrix<double>& A, const Matrix<double>& B, double beta, Matrix<double>& C) {
    detail::Dev dA(A), dB(B), dC(C);
    detail::check(elemamd_trrk_f64(detail::uplo_char(uplo), detail::orient_char(orientationOfA),
                                   detail::orient_char(orientationOfB), C.Height(), A.Height(), alpha, dA.p, dA.ld,
                                   dB.p, dB.ld, beta, dC.p, dC.ld, nullptr),
                  "elemamd_trrk_f64");
    dC.back(C);
}

inline void Trsv(UpperOrLower uplo, Orientation orientation, UnitOrNonUnit diag, const Matrix<double>& A,
                 Matrix<double>& x) {
    if (x.Width() != 1) throw std::logic_error("x must be a column vector here");
    detail::Dev dA(A), dx(x);
    detail::check(elemamd_trsv_f64(detail::uplo_char(uplo), detail::orient_char(orientation),
                                   diag == NON_UNIT ? 'N' : 'U', A.Height(), dA.p, dA.ld, dx.p, 1, nullptr),
                  "elemamd_trsv_f64");
    dx.back(x);
}

}  // namespace elem
