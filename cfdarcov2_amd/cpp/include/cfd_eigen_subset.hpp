// The subset of Eigen's dense API that cfdARCO drivers and the operator API
// use: dynamic column vectors and column-major matrices of arithmetic types,
// small fixed-size vectors, coefficient access, reductions, cwise products and
// `.array()` arithmetic.  Everything evaluates eagerly (no expression
// templates); scalars of another arithmetic type are converted to the
// matrix's scalar type, as Eigen does for literals such as `0.5 * v.array()`.
#pragma once

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstddef>
#include <initializer_list>
#include <limits>
#include <map>
#include <mutex>
#include <new>
#include <numeric>
#include <ostream>
#include <stdexcept>
#include <type_traits>
#include <vector>

namespace Eigen {

namespace detail {

// Optional page-locked host memory for large blocks, provided by the
// operator-API library once the device is enabled (weak: a program built
// without it, or a call before the device is up, gets ordinary memory).  The
// copies between the device and page-locked memory run at DMA speed instead of
// through the runtime's staging buffers.
extern "C" {
__attribute__((weak)) void* cfdarco_pinned_alloc(size_t bytes);
__attribute__((weak)) void cfdarco_pinned_free(void* p, size_t bytes);
}

// Process-wide reuse of large freed blocks.  Drivers that round-trip a field
// through the host each step (a host boundary callback copies the field,
// returns the copy, and the old storage is released) otherwise pay a fresh
// mapping, and a page fault per page, for every such vector.  Blocks of at
// least kMinBytes carry a 64-byte header (size and whether the memory is
// page-locked), so any copy of this code can release any block; they are kept
// by exact size, at most kMaxBlocks of them and kMaxBytes in all.  Smaller
// blocks go straight to operator new / delete.
class BlockCache {
public:
    static constexpr size_t kMinBytes = size_t(1) << 20;
    static constexpr size_t kMaxBlocks = 8;
    static constexpr size_t kMaxBytes = size_t(4) << 30;
    static constexpr size_t kHeader = 64;

    static BlockCache& get() {
        static BlockCache* c = new BlockCache;  // never destroyed: frees may run during exit
        return *c;
    }
    void* take(size_t bytes) {
        if (bytes < kMinBytes) return ::operator new(bytes);
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.find(bytes);
            if (it != free_.end()) {
                void* p = it->second;
                free_.erase(it);
                held_ -= bytes;
                return p;
            }
        }
        void* raw = nullptr;
        int pinned = 0;
        if (cfdarco_pinned_alloc != nullptr) {
            raw = cfdarco_pinned_alloc(bytes + kHeader);
            pinned = raw != nullptr;
        }
        if (raw == nullptr) raw = ::operator new(bytes + kHeader);
        Header* h = static_cast<Header*>(raw);
        h->bytes = bytes;
        h->pinned = pinned;
        return static_cast<char*>(raw) + kHeader;
    }
    void give(void* p, size_t bytes) {
        if (bytes < kMinBytes) {
            ::operator delete(p);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (free_.size() < kMaxBlocks && held_ + bytes <= kMaxBytes) {
                free_.emplace(bytes, p);
                held_ += bytes;
                return;
            }
        }
        void* raw = static_cast<char*>(p) - kHeader;
        if (!static_cast<Header*>(raw)->pinned)
            ::operator delete(raw);
        else if (cfdarco_pinned_free != nullptr)
            cfdarco_pinned_free(raw, bytes + kHeader);
        // (page-locked memory is never handed to operator delete)
    }

private:
    struct Header {
        size_t bytes;
        int pinned;
    };
    static_assert(sizeof(Header) <= kHeader, "block header");
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    size_t held_ = 0;
};

template <class T> struct BlockAllocator {
    using value_type = T;
    BlockAllocator() noexcept = default;
    template <class U> BlockAllocator(const BlockAllocator<U>&) noexcept {}
    T* allocate(size_t n) {
        if (n > std::numeric_limits<size_t>::max() / sizeof(T)) throw std::bad_array_new_length();
        return static_cast<T*>(BlockCache::get().take(n * sizeof(T)));
    }
    void deallocate(T* p, size_t n) noexcept { BlockCache::get().give(p, n * sizeof(T)); }
    template <class U> bool operator==(const BlockAllocator<U>&) const noexcept { return true; }
    template <class U> bool operator!=(const BlockAllocator<U>&) const noexcept { return false; }
};

// Sum of p[0..n): halves summed recursively down to blocks of at most 1024,
// each block summed in eight interleaved partial sums (which the compiler
// keeps in vector registers) combined pairwise.  A plain left-to-right float
// sum stalls once it reaches 2^24 times the terms' size (large meshes of O(1)
// values) and runs one dependent add per element; this one's error grows
// with log2(n) and it streams at memory speed.
template <class S> S pairwise_sum(const S* p, size_t n) {
    if (n <= 1024) {
        S acc[8] = {S(0), S(0), S(0), S(0), S(0), S(0), S(0), S(0)};
        size_t i = 0;
        for (; i + 8 <= n; i += 8)
            for (int k = 0; k < 8; ++k) acc[k] += p[i + k];
        for (int k = 0; i < n; ++i, ++k) acc[k] += p[i];
        return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    }
    const size_t h = n / 2;
    return pairwise_sum(p, h) + pairwise_sum(p + h, n - h);
}

// Largest (Greater = true) or smallest coefficient of p[0..n), n >= 1, in
// eight interleaved lanes.  A NaN is kept only as the first element, as
// std::max_element / std::min_element keep it.
template <bool Greater, class S> S extreme(const S* p, size_t n) {
    S m[8];
    for (int k = 0; k < 8; ++k) m[k] = p[0];
    size_t i = 1;
    for (; i + 8 <= n; i += 8)
        for (int k = 0; k < 8; ++k) {
            const S x = p[i + k];
            m[k] = (Greater ? m[k] < x : x < m[k]) ? x : m[k];
        }
    for (int k = 0; i < n; ++i, ++k) {
        const S x = p[i];
        m[k] = (Greater ? m[k] < x : x < m[k]) ? x : m[k];
    }
    S r = m[0];
    for (int k = 1; k < 8; ++k) r = (Greater ? r < m[k] : m[k] < r) ? m[k] : r;
    return r;
}

}  // namespace detail

// coefficient storage of the dense types
template <class S> using Buffer = std::vector<S, detail::BlockAllocator<S>>;

using Index = std::ptrdiff_t;
constexpr int Dynamic = -1;
enum StorageOptions { ColMajor = 0, RowMajor = 1, AutoAlign = 0 };

template <class S> class Array1;

template <class S, int Rows, int Cols, int Options = ColMajor, int MaxRows = Rows, int MaxCols = Cols>
class Matrix {
    static_assert(std::is_arithmetic<S>::value, "arithmetic scalars only");

public:
    using Scalar = S;

    Matrix() : rows_(Rows > 0 ? Rows : 0), cols_(Cols > 0 ? Cols : 0), v_(size_t(rows_ * cols_)) {}

    // Matrix<S,-1,1>{n}: a vector of n entries; Matrix<S,-1,-1>{r, c}
    template <class I, class = typename std::enable_if<std::is_integral<I>::value>::type>
    explicit Matrix(I n) {
        if (Cols == 1 || Rows == Dynamic) {
            rows_ = Index(n);
            cols_ = Cols > 0 ? Cols : 1;
        } else {
            rows_ = 1;
            cols_ = Index(n);
        }
        v_.assign(size_t(rows_ * cols_), S(0));
    }
    template <class I, class J,
              class = typename std::enable_if<std::is_integral<I>::value && std::is_integral<J>::value>::type>
    Matrix(I r, J c) : rows_(Index(r)), cols_(Index(c)), v_(size_t(Index(r) * Index(c)), S(0)) {}

    // fixed-size 3-vectors from coordinates
    template <class A, class B, class C,
              class = typename std::enable_if<std::is_floating_point<A>::value>::type>
    Matrix(A x, B y, C z) : rows_(3), cols_(1), v_{S(x), S(y), S(z)} {}

    // between vector/matrix shapes of the same data (N x 1 <-> dynamic N x M)
    template <int R2, int C2, int O2, int MR2, int MC2>
    Matrix(const Matrix<S, R2, C2, O2, MR2, MC2>& o) : rows_(o.rows()), cols_(o.cols()), v_(o.data(), o.data() + o.size()) {}

    Matrix(const Array1<S>& a);

    Index rows() const { return rows_; }
    Index cols() const { return cols_; }
    Index size() const { return rows_ * cols_; }
    S* data() { return v_.data(); }
    const S* data() const { return v_.data(); }

    void resize(Index n) {
        rows_ = n;
        cols_ = 1;
        v_.assign(size_t(n), S(0));
    }
    void resize(Index r, Index c) {
        rows_ = r;
        cols_ = c;
        v_.assign(size_t(r * c), S(0));
    }

    S& operator()(Index i) { return v_[size_t(i)]; }
    const S& operator()(Index i) const { return v_[size_t(i)]; }
    S& operator()(Index r, Index c) { return v_[size_t(c * rows_ + r)]; }
    const S& operator()(Index r, Index c) const { return v_[size_t(c * rows_ + r)]; }
    S& operator[](Index i) { return v_[size_t(i)]; }
    const S& operator[](Index i) const { return v_[size_t(i)]; }

    template <class T> Matrix& setConstant(T c) {
        std::fill(v_.begin(), v_.end(), S(c));
        return *this;
    }
    Matrix& setZero() { return setConstant(0); }
    static Matrix Zero(Index n) { return Matrix(n); }
    static Matrix Zero(Index r, Index c) { return Matrix(r, c); }
    template <class T> static Matrix Constant(Index n, T c) {
        Matrix m(n);
        m.setConstant(c);
        return m;
    }

    // sums are pairwise; maxima and minima in eight lanes
    S sum() const { return detail::pairwise_sum(v_.data(), v_.size()); }
    S maxCoeff() const { return detail::extreme<true>(v_.data(), v_.size()); }
    S minCoeff() const { return detail::extreme<false>(v_.data(), v_.size()); }
    S squaredNorm() const {
        S s = 0;
        for (const S& x : v_) s += x * x;
        return s;
    }
    S norm() const { return std::sqrt(squaredNorm()); }
    S dot(const Matrix& o) const {
        S s = 0;
        for (size_t i = 0; i < v_.size(); ++i) s += v_[i] * o.v_[i];
        return s;
    }

    Matrix cwiseProduct(const Matrix& o) const { return zip(o, [](S a, S b) { return a * b; }); }
    Matrix cwiseQuotient(const Matrix& o) const { return zip(o, [](S a, S b) { return a / b; }); }
    Matrix cwiseSqrt() const { return map([](S a) { return std::sqrt(a); }); }
    Matrix cwiseAbs() const { return map([](S a) { return std::abs(a); }); }
    Matrix cwiseInverse() const { return map([](S a) { return S(1) / a; }); }

    Matrix operator+(const Matrix& o) const { return zip(o, [](S a, S b) { return a + b; }); }
    Matrix operator-(const Matrix& o) const { return zip(o, [](S a, S b) { return a - b; }); }
    Matrix operator-() const { return map([](S a) { return -a; }); }
    Matrix& operator+=(const Matrix& o) { return *this = *this + o; }
    Matrix& operator-=(const Matrix& o) { return *this = *this - o; }
    template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
    Matrix operator*(T s) const {
        const S c = S(s);
        return map([c](S a) { return a * c; });
    }
    template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
    Matrix operator/(T s) const {
        const S c = S(s);
        return map([c](S a) { return a / c; });
    }
    template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
    Matrix& operator/=(T s) {
        return *this = *this / s;
    }
    template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
    Matrix& operator*=(T s) {
        return *this = *this * s;
    }

    Array1<S> array() const;

    Buffer<S>& storage() { return v_; }
    const Buffer<S>& storage() const { return v_; }

private:
    template <class F> Matrix zip(const Matrix& o, F f) const {
        if (o.size() != size()) throw std::invalid_argument("Eigen subset: size mismatch");
        Matrix r = *this;
        for (size_t i = 0; i < v_.size(); ++i) r.v_[i] = f(v_[i], o.v_[i]);
        return r;
    }
    template <class F> Matrix map(F f) const {
        Matrix r = *this;
        for (auto& x : r.v_) x = f(x);
        return r;
    }

    Index rows_ = 0, cols_ = 0;
    Buffer<S> v_;
};

template <class T, class S, int R, int C, int O, int MR, int MC,
          class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
Matrix<S, R, C, O, MR, MC> operator*(T s, const Matrix<S, R, C, O, MR, MC>& m) {
    const S c = S(s);
    Matrix<S, R, C, O, MR, MC> r = m;
    for (auto& x : r.storage()) x = c * x;
    return r;
}

// Owning one-dimensional array with coefficient-wise arithmetic (the value of
// `.array()`), convertible back to a column vector.
template <class S> class Array1 {
public:
    Array1() = default;
    explicit Array1(Buffer<S> v) : v_(std::move(v)) {}
    Index size() const { return Index(v_.size()); }
    S& operator()(Index i) { return v_[size_t(i)]; }
    const S& operator()(Index i) const { return v_[size_t(i)]; }
    S& operator[](Index i) { return v_[size_t(i)]; }
    const S& operator[](Index i) const { return v_[size_t(i)]; }
    const Buffer<S>& storage() const { return v_; }
    Buffer<S>& storage() { return v_; }
    Matrix<S, Dynamic, 1> matrix() const { return Matrix<S, Dynamic, 1>(*this); }
    S sum() const { return detail::pairwise_sum(v_.data(), v_.size()); }
    S maxCoeff() const { return detail::extreme<true>(v_.data(), v_.size()); }
    S minCoeff() const { return detail::extreme<false>(v_.data(), v_.size()); }

    template <class F> Array1 map(F f) const {
        Array1 r = *this;
        for (auto& x : r.v_) x = f(x);
        return r;
    }
    template <class F> Array1 zip(const Array1& o, F f) const {
        if (o.size() != size()) throw std::invalid_argument("Eigen subset: size mismatch");
        Array1 r = *this;
        for (size_t i = 0; i < v_.size(); ++i) r.v_[i] = f(v_[i], o.v_[i]);
        return r;
    }
    Array1 sqrt() const { return map([](S a) { return std::sqrt(a); }); }
    Array1 cwiseSqrt() const { return sqrt(); }
    Array1 exp() const { return map([](S a) { return std::exp(a); }); }
    Array1 abs() const { return map([](S a) { return std::abs(a); }); }
    Array1 inverse() const { return map([](S a) { return S(1) / a; }); }
    Array1 cwiseInverse() const { return inverse(); }
    Array1 cwiseQuotient(const Array1& o) const { return zip(o, [](S a, S b) { return a / b; }); }
    Array1 cwiseProduct(const Array1& o) const { return zip(o, [](S a, S b) { return a * b; }); }
    Array1 array() const { return *this; }

private:
    Buffer<S> v_;
};

template <class S, int R, int C, int O, int MR, int MC>
Matrix<S, R, C, O, MR, MC>::Matrix(const Array1<S>& a) : rows_(a.size()), cols_(1), v_(a.storage()) {}

template <class S, int R, int C, int O, int MR, int MC>
Array1<S> Matrix<S, R, C, O, MR, MC>::array() const {
    return Array1<S>(v_);
}

#define CFD_EIGEN_ARRAY_OP(OP)                                                                             \
    template <class S> Array1<S> operator OP(const Array1<S>& a, const Array1<S>& b) {                     \
        return a.zip(b, [](S x, S y) { return x OP y; });                                                  \
    }                                                                                                      \
    template <class S, class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>       \
    Array1<S> operator OP(const Array1<S>& a, T s) {                                                       \
        const S c = S(s);                                                                                  \
        return a.map([c](S x) { return x OP c; });                                                         \
    }                                                                                                      \
    template <class S, class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>       \
    Array1<S> operator OP(T s, const Array1<S>& a) {                                                       \
        const S c = S(s);                                                                                  \
        return a.map([c](S x) { return c OP x; });                                                         \
    }
CFD_EIGEN_ARRAY_OP(+)
CFD_EIGEN_ARRAY_OP(-)
CFD_EIGEN_ARRAY_OP(*)
CFD_EIGEN_ARRAY_OP(/)
#undef CFD_EIGEN_ARRAY_OP

template <class S> Array1<S> operator-(const Array1<S>& a) {
    return a.map([](S x) { return -x; });
}
template <class S> Array1<S> exp(const Array1<S>& a) { return a.exp(); }
template <class S> Array1<S> abs(const Array1<S>& a) { return a.abs(); }
template <class S> Array1<S> sqrt(const Array1<S>& a) { return a.sqrt(); }

template <class S, int R, int C, int O, int MR, int MC>
std::ostream& operator<<(std::ostream& os, const Matrix<S, R, C, O, MR, MC>& m) {
    for (Index r = 0; r < m.rows(); ++r) {
        for (Index c = 0; c < m.cols(); ++c) os << (c ? " " : "") << m(r, c);
        os << "\n";
    }
    return os;
}

template <class S, int Size = Dynamic> using Vector = Matrix<S, Size, 1>;
using VectorXf = Matrix<float, Dynamic, 1>;
using VectorXd = Matrix<double, Dynamic, 1>;
using MatrixXf = Matrix<float, Dynamic, Dynamic>;
using MatrixXd = Matrix<double, Dynamic, Dynamic>;
using Vector3f = Matrix<float, 3, 1>;

// Placeholder for Eigen::Tensor, which only appears in a type alias of the
// reference declarations.
template <class S, int Rank, int Options = ColMajor> class Tensor {};

}  // namespace Eigen
