// tlfea/types.h -- scalar and dense-matrix types of the drop-in API.
//
// The reference's public signatures use Eigen types through aliases in namespace tlfea
// (src/types.h:19-42).  Eigen is not a dependency here, so this header provides the
// subset those signatures and their callers use: a column-major dense Matrix with
// dynamic or fixed sizes, Map (a non-owning view), comma initialisation, element access,
// resize/setZero/min/max, element-wise +/-, scalar *, col(), block() copies and eval().
#pragma once

#include <algorithm>
#include <cassert>
#include <cstddef>
#include <initializer_list>
#include <stdexcept>
#include <vector>

namespace tlfea {

typedef double Real;

static constexpr int DynamicMatrix = -1;
static constexpr int RowMajorMatrix = 1;
static constexpr int ColMajorMatrix = 0;

template <typename Scalar, int Rows, int Cols, int Options = 0>
class Matrix;

namespace detail {
template <typename M>
class CommaInit {
  public:
    CommaInit(M& m, typename M::Scalar first) : m_(m), k_(0) { put(first); }
    CommaInit& operator,(typename M::Scalar v) {
        put(v);
        return *this;
    }
    M& finished() { return m_; }
    operator M&() { return m_; }

  private:
    // fills row by row, as Eigen's comma initialiser does
    void put(typename M::Scalar v) {
        const std::ptrdiff_t r = k_ / m_.cols(), c = k_ % m_.cols();
        if (r >= m_.rows()) throw std::out_of_range("too many coefficients in comma initializer");
        m_(r, c) = v;
        ++k_;
    }
    M& m_;
    std::ptrdiff_t k_;
};
}  // namespace detail

// Column-major dense matrix.  Rows/Cols may be DynamicMatrix (-1).  Options only
// records RowMajor for type compatibility; storage is always column-major.
template <typename Scalar_, int Rows, int Cols, int Options>
class Matrix {
  public:
    typedef Scalar_ Scalar;
    typedef std::ptrdiff_t Index;

    Matrix() : r_(Rows > 0 ? Rows : 0), c_(Cols > 0 ? Cols : 0), d_((size_t)r_ * c_, Scalar(0)) {}
    // vector constructor (one dynamic dimension)
    explicit Matrix(Index n) : Matrix() {
        static_assert(Rows == 1 || Cols == 1, "single-size constructor needs a vector type");
        if (Cols == 1)
            resize(n, 1);
        else
            resize(1, n);
    }
    Matrix(Index r, Index c) : r_(r), c_(c), d_((size_t)r * c, Scalar(0)) {}

    Index rows() const { return r_; }
    Index cols() const { return c_; }
    Index size() const { return r_ * c_; }
    Scalar* data() { return d_.data(); }
    const Scalar* data() const { return d_.data(); }

    Scalar& operator()(Index i, Index j) { return d_[(size_t)(i + j * r_)]; }
    const Scalar& operator()(Index i, Index j) const { return d_[(size_t)(i + j * r_)]; }
    Scalar& operator()(Index i) { return d_[(size_t)i]; }
    const Scalar& operator()(Index i) const { return d_[(size_t)i]; }
    Scalar& operator[](Index i) { return d_[(size_t)i]; }
    const Scalar& operator[](Index i) const { return d_[(size_t)i]; }

    void resize(Index r, Index c) {
        r_ = r;
        c_ = c;
        d_.assign((size_t)r * c, Scalar(0));
    }
    void resize(Index n) {
        if (Cols == 1)
            resize(n, 1);
        else
            resize(1, n);
    }
    Matrix& setZero() {
        std::fill(d_.begin(), d_.end(), Scalar(0));
        return *this;
    }
    Matrix& setConstant(Scalar v) {
        std::fill(d_.begin(), d_.end(), v);
        return *this;
    }
    Scalar minCoeff() const { return *std::min_element(d_.begin(), d_.end()); }
    Scalar maxCoeff() const { return *std::max_element(d_.begin(), d_.end()); }
    Scalar sum() const {
        Scalar s(0);
        for (auto v : d_) s += v;
        return s;
    }
    Scalar squaredNorm() const {
        Scalar s(0);
        for (auto v : d_) s += v * v;
        return s;
    }
    Scalar norm() const;

    detail::CommaInit<Matrix> operator<<(Scalar v) { return detail::CommaInit<Matrix>(*this, v); }

    Matrix eval() const { return *this; }

    Matrix<Scalar, DynamicMatrix, 1> col(Index j) const {
        Matrix<Scalar, DynamicMatrix, 1> v(r_, 1);
        for (Index i = 0; i < r_; i++) v(i) = (*this)(i, j);
        return v;
    }
    Matrix<Scalar, DynamicMatrix, DynamicMatrix> block(Index i0, Index j0, Index nr, Index nc) const {
        Matrix<Scalar, DynamicMatrix, DynamicMatrix> b(nr, nc);
        for (Index j = 0; j < nc; j++)
            for (Index i = 0; i < nr; i++) b(i, j) = (*this)(i0 + i, j0 + j);
        return b;
    }

    template <int R2, int C2, int O2>
    Matrix(const Matrix<Scalar, R2, C2, O2>& o) : r_(o.rows()), c_(o.cols()), d_(o.data(), o.data() + o.size()) {}

    Matrix operator-(const Matrix& o) const { return zip(o, [](Scalar a, Scalar b) { return a - b; }); }
    Matrix operator+(const Matrix& o) const { return zip(o, [](Scalar a, Scalar b) { return a + b; }); }
    Matrix operator*(Scalar s) const {
        Matrix m(*this);
        for (auto& v : m.d_) v *= s;
        return m;
    }
    Matrix operator/(Scalar s) const {
        Matrix m(*this);
        for (auto& v : m.d_) v /= s;
        return m;
    }

  private:
    template <typename F>
    Matrix zip(const Matrix& o, F f) const {
        if (o.r_ != r_ || o.c_ != c_) throw std::invalid_argument("matrix size mismatch");
        Matrix m(*this);
        for (size_t k = 0; k < d_.size(); k++) m.d_[k] = f(d_[k], o.d_[k]);
        return m;
    }
    Index r_, c_;
    std::vector<Scalar> d_;
};

}  // namespace tlfea

#include <cmath>

namespace tlfea {
template <typename S, int R, int C, int O>
S Matrix<S, R, C, O>::norm() const {
    return std::sqrt(squaredNorm());
}

// Non-owning column-major view over existing memory (the reference's tlfea::Map).
template <typename PlainObjectType, int MapOptions = 0, typename StrideType = void>
class Map {
  public:
    typedef typename PlainObjectType::Scalar Scalar;
    typedef std::ptrdiff_t Index;
    Map(Scalar* p, Index n) : p_(p), r_(n), c_(1) {}
    Map(Scalar* p, Index r, Index c) : p_(p), r_(r), c_(c) {}
    Index rows() const { return r_; }
    Index cols() const { return c_; }
    Index size() const { return r_ * c_; }
    Scalar* data() const { return p_; }
    Scalar& operator()(Index i, Index j) const { return p_[i + j * r_]; }
    Scalar& operator()(Index i) const { return p_[i]; }
    Scalar& operator[](Index i) const { return p_[i]; }
    PlainObjectType eval() const {
        PlainObjectType m(r_, c_);
        std::copy(p_, p_ + r_ * c_, m.data());
        return m;
    }
    operator PlainObjectType() const { return eval(); }

  private:
    Scalar* p_;
    Index r_, c_;
};

template <typename PlainObjectType, int MapOptions, typename StrideType>
class Map<const PlainObjectType, MapOptions, StrideType> {
  public:
    typedef typename PlainObjectType::Scalar Scalar;
    typedef std::ptrdiff_t Index;
    Map(const Scalar* p, Index n) : p_(p), r_(n), c_(1) {}
    Map(const Scalar* p, Index r, Index c) : p_(p), r_(r), c_(c) {}
    Index rows() const { return r_; }
    Index cols() const { return c_; }
    Index size() const { return r_ * c_; }
    const Scalar* data() const { return p_; }
    const Scalar& operator()(Index i, Index j) const { return p_[i + j * r_]; }
    const Scalar& operator()(Index i) const { return p_[i]; }
    PlainObjectType eval() const {
        PlainObjectType m(r_, c_);
        std::copy(p_, p_ + r_ * c_, m.data());
        return m;
    }
    operator PlainObjectType() const { return eval(); }

  private:
    const Scalar* p_;
    Index r_, c_;
};

typedef Matrix<Real, DynamicMatrix, DynamicMatrix> MatrixXR;
typedef Matrix<Real, DynamicMatrix, 1> VectorXR;
typedef Matrix<Real, 3, 3> Matrix3R;
typedef Matrix<Real, 3, 1> Vector3R;
typedef Matrix<int, DynamicMatrix, DynamicMatrix> MatrixXi;
typedef Matrix<int, DynamicMatrix, 1> VectorXi;

}  // namespace tlfea
