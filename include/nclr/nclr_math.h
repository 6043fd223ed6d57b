// nclr_math.h - small fixed-size linear algebra and scene helpers of the drop-in API.
//
// Mirrors the reference's src/nclr_math.h (type aliases 5-11, diag 13-19, constmat/constvec
// 40-48, nclr_svd 50-74, nclr_polar 76-98, cube 100-129, pt_3d_to_2d 154-167) without Eigen:
// Vector<T, n> / Matrix<T, n> are small value types with the Eigen spellings callers use
// (operator(), x()/y()/z(), arithmetic, transpose(), determinant(), cast<U>(), Zero(), ...).
// The simulation itself runs on the GPU (nclr.h); the SVD/polar here are host utilities that
// follow the same two-sided Jacobi algorithm as the device kernels.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <initializer_list>
#include <limits>
#include <ostream>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

namespace nclr {

template <typename T, int R, int C>
class Mat {
  public:
    static constexpr int rows_ = R;
    static constexpr int cols_ = C;
    Mat() {
        for (auto& v : d_) v = T(0);
    }
    template <typename... Args, typename = std::enable_if_t<(sizeof...(Args) == R * C) && (R * C > 1) &&
                                                            (std::is_arithmetic_v<Args> && ...)>>
    Mat(Args... args) : d_{T(args)...} {}

    static Mat Zero() { return Mat(); }
    static Mat Constant(T v) {
        Mat m;
        for (auto& x : m.d_) x = v;
        return m;
    }
    static Mat Ones() { return Constant(T(1)); }
    static Mat Identity() {
        Mat m;
        for (int i = 0; i < std::min(R, C); ++i) m(i, i) = T(1);
        return m;
    }
    void setZero() { *this = Zero(); }
    void setIdentity() { *this = Identity(); }
    void setConstant(T v) { *this = Constant(v); }

    T& operator()(int r, int c) { return d_[r * C + c]; }
    const T& operator()(int r, int c) const { return d_[r * C + c]; }
    T& operator()(int i) { return d_[i]; }
    const T& operator()(int i) const { return d_[i]; }
    T& operator[](int i) { return d_[i]; }
    const T& operator[](int i) const { return d_[i]; }
    T& x() { return d_[0]; }
    T& y() { return d_[1]; }
    T& z() { return d_[2]; }
    const T& x() const { return d_[0]; }
    const T& y() const { return d_[1]; }
    const T& z() const { return d_[2]; }
    static constexpr int rows() { return R; }
    static constexpr int cols() { return C; }
    static constexpr int size() { return R * C; }
    T* data() { return d_; }
    const T* data() const { return d_; }

    Mat& operator+=(const Mat& o) {
        for (int i = 0; i < R * C; ++i) d_[i] += o.d_[i];
        return *this;
    }
    Mat& operator-=(const Mat& o) {
        for (int i = 0; i < R * C; ++i) d_[i] -= o.d_[i];
        return *this;
    }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
    Mat& operator*=(S s) {
        for (auto& v : d_) v = T(v * s);
        return *this;
    }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
    Mat& operator/=(S s) {
        for (auto& v : d_) v = T(v / s);
        return *this;
    }
    Mat operator-() const {
        Mat m;
        for (int i = 0; i < R * C; ++i) m.d_[i] = -d_[i];
        return m;
    }
    friend Mat operator+(Mat a, const Mat& b) { return a += b; }
    friend Mat operator-(Mat a, const Mat& b) { return a -= b; }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
    friend Mat operator*(Mat a, S s) {
        return a *= s;
    }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
    friend Mat operator*(S s, Mat a) {
        return a *= s;
    }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
    friend Mat operator/(Mat a, S s) {
        return a /= s;
    }
    template <int K>
    Mat<T, R, K> operator*(const Mat<T, C, K>& b) const {
        Mat<T, R, K> out;
        for (int i = 0; i < R; ++i)
            for (int j = 0; j < K; ++j) {
                T s = (*this)(i, 0) * b(0, j);
                for (int k = 1; k < C; ++k) s += (*this)(i, k) * b(k, j);
                out(i, j) = s;
            }
        return out;
    }
    Mat<T, C, R> transpose() const {
        Mat<T, C, R> t;
        for (int i = 0; i < R; ++i)
            for (int j = 0; j < C; ++j) t(j, i) = (*this)(i, j);
        return t;
    }
    Mat cwiseProduct(const Mat& o) const {
        Mat m;
        for (int i = 0; i < R * C; ++i) m.d_[i] = d_[i] * o.d_[i];
        return m;
    }
    template <typename U>
    Mat<U, R, C> cast() const {
        Mat<U, R, C> m;
        for (int i = 0; i < R * C; ++i) m.data()[i] = U(d_[i]);
        return m;
    }
    T sum() const {
        T s = T(0);
        for (auto v : d_) s += v;
        return s;
    }
    T squaredNorm() const {
        T s = T(0);
        for (auto v : d_) s += v * v;
        return s;
    }
    T norm() const { return std::sqrt(squaredNorm()); }
    T determinant() const {
        static_assert(R == C && (R == 2 || R == 3), "determinant of 2x2 or 3x3");
        const auto& a = *this;
        if constexpr (R == 2) {
            return a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0);
        } else {
            return a(0, 0) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) - a(0, 1) * (a(1, 0) * a(2, 2) - a(1, 2) * a(2, 0)) +
                   a(0, 2) * (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0));
        }
    }

    // Eigen-style text: rows on separate lines, entries right-aligned to a common width.
    friend std::ostream& operator<<(std::ostream& os, const Mat& m) {
        std::vector<std::string> cells;
        size_t width = 0;
        for (int i = 0; i < R * C; ++i) {
            std::ostringstream s;
            s.precision(os.precision());
            s << m.d_[i];
            cells.push_back(s.str());
            width = std::max(width, cells.back().size());
        }
        for (int r = 0; r < R; ++r) {
            for (int c = 0; c < C; ++c) {
                if (c) os << ' ';
                const std::string& s = cells[r * C + c];
                os << std::string(width - s.size(), ' ') << s;
            }
            if (r + 1 < R) os << '\n';
        }
        return os;
    }

  private:
    T d_[R * C];
};

template <typename T, int dim>
using Vector = Mat<T, dim, 1>;

template <typename T, int dim>
using Matrix = Mat<T, dim, dim>;

using real = float;

// diag<dim>(v): only (0,0) and (1,1) are set, so in 3D this is diag(v, v, 0) (src/nclr_math.h:13-19).
template <int dim>
inline auto diag(const float value) -> Matrix<real, dim> {
    Matrix<real, dim> m = Matrix<real, dim>::Zero();
    m(0, 0) = value;
    m(1, 1) = value;
    return m;
}

template <int dim>
inline auto constmat(const float value) -> Matrix<real, dim> {
    return Matrix<real, dim>::Constant(value);
}

template <int dim>
inline auto constvec(const float value) -> Vector<real, dim> {
    return Vector<real, dim>::Constant(value);
}

// Marsaglia xorshift128 with the reference's fixed seed words (src/nclr_math.h:22-31).
inline auto nc_rand_int() -> uint32_t {
    static uint32_t x = 123456789u, y = 362436069u, z = 521288629u, w = 88675123u;
    const uint32_t t = x ^ (x << 11);
    x = y;
    y = z;
    z = w;
    return (w = (w ^ (w >> 19)) ^ (t ^ (t >> 8)));
}

inline auto nc_rand() -> float { return float(nc_rand_int()) * (1.0f / 4294967296.0f); }

template <int dim>
inline auto randvec() -> Vector<real, dim> {
    Vector<real, dim> r;
    for (int i = 0; i < dim; ++i) r(i) = nc_rand();
    return r;
}

namespace detail {

template <typename T>
struct Rot {
    T c, s;
};

template <typename T>
inline Rot<T> jacobi(T x, T y, T z) {
    const T deno = T(2) * std::fabs(y);
    if (deno < std::numeric_limits<T>::min()) return {T(1), T(0)};
    const T tau = (x - z) / deno;
    const T w = std::sqrt(tau * tau + T(1));
    const T t = tau > T(0) ? T(1) / (tau + w) : T(1) / (tau - w);
    const T sign_t = t > T(0) ? T(1) : T(-1);
    const T n = T(1) / std::sqrt(t * t + T(1));
    return {n, -sign_t * (y / std::fabs(y)) * std::fabs(t) * n};
}

}  // namespace detail

// Singular value decomposition a = U diag(sig) V^T by two-sided cyclic Jacobi rotations,
// singular values sorted by magnitude; in 3D U, V are proper rotations with the sign of det(a)
// on the smallest singular value (src/nclr_math.h:50-74).
template <int dim>
inline auto nclr_svd(const Matrix<real, dim>& a, Matrix<real, dim>& U, Matrix<real, dim>& sig,
                     Matrix<real, dim>& V) -> void {
    using T = real;
    T scale = T(0);
    for (int i = 0; i < dim * dim; ++i) scale = std::max(scale, std::fabs(a.data()[i]));
    if (scale == T(0)) scale = T(1);
    Matrix<real, dim> W = a / scale;
    U.setIdentity();
    V.setIdentity();
    T max_diag = T(0);
    for (int i = 0; i < dim; ++i) max_diag = std::max(max_diag, std::fabs(W(i, i)));
    bool finished = false;
    for (int sweep = 0; !finished && sweep < 32; ++sweep) {
        finished = true;
        for (int p = 1; p < dim; ++p)
            for (int q = 0; q < p; ++q) {
                const T thr = std::max(std::numeric_limits<T>::min(), T(2) * std::numeric_limits<T>::epsilon() * max_diag);
                if (!(std::fabs(W(p, q)) > thr || std::fabs(W(q, p)) > thr)) continue;
                finished = false;
                const T m00 = W(p, p), m01 = W(p, q), m10 = W(q, p), m11 = W(q, q);
                detail::Rot<T> r1{T(1), T(0)};
                const T t = m00 + m11, d = m10 - m01;
                if (std::fabs(d) >= std::numeric_limits<T>::min()) {
                    const T u = t / d, tmp = std::sqrt(T(1) + u * u);
                    r1 = {u / tmp, T(1) / tmp};
                }
                const auto jr = detail::jacobi(r1.c * m00 + r1.s * m10, r1.c * m01 + r1.s * m11, -r1.s * m01 + r1.c * m11);
                const detail::Rot<T> jl{r1.c * jr.c + r1.s * jr.s, -r1.c * jr.s + r1.s * jr.c};
                for (int k = 0; k < dim; ++k) {
                    const T x = W(p, k), y = W(q, k);
                    W(p, k) = jl.c * x + jl.s * y;
                    W(q, k) = -jl.s * x + jl.c * y;
                    const T ux = U(k, p), uy = U(k, q);
                    U(k, p) = jl.c * ux + jl.s * uy;
                    U(k, q) = -jl.s * ux + jl.c * uy;
                }
                for (int k = 0; k < dim; ++k) {
                    const T x = W(k, p), y = W(k, q);
                    W(k, p) = jr.c * x - jr.s * y;
                    W(k, q) = jr.s * x + jr.c * y;
                    const T vx = V(k, p), vy = V(k, q);
                    V(k, p) = jr.c * vx - jr.s * vy;
                    V(k, q) = jr.s * vx + jr.c * vy;
                }
                max_diag = std::max(max_diag, std::max(std::fabs(W(p, p)), std::fabs(W(q, q))));
            }
    }
    T s[dim];
    for (int i = 0; i < dim; ++i) {
        s[i] = std::fabs(W(i, i)) * scale;
        if (W(i, i) < T(0))
            for (int k = 0; k < dim; ++k) U(k, i) = -U(k, i);
    }
    for (int i = 0; i < dim; ++i) {
        int pos = i;
        for (int k = i + 1; k < dim; ++k)
            if (s[k] > s[pos]) pos = k;
        if (s[pos] == T(0)) break;
        if (pos != i) {
            std::swap(s[i], s[pos]);
            for (int k = 0; k < dim; ++k) {
                std::swap(U(k, i), U(k, pos));
                std::swap(V(k, i), V(k, pos));
            }
        }
    }
    if constexpr (dim == 3) {
        if (U.determinant() < T(0)) {
            for (int k = 0; k < 3; ++k) U(k, 2) = -U(k, 2);
            s[2] = -s[2];
        }
        if (V.determinant() < T(0)) {
            for (int k = 0; k < 3; ++k) V(k, 2) = -V(k, 2);
            s[2] = -s[2];
        }
    }
    sig.setIdentity();
    for (int i = 0; i < dim; ++i) sig(i, i) = s[i];
}

// Polar decomposition m = R S (src/nclr_math.h:76-98): closed form in 2D, R = U V^T in 3D.
template <int dim>
inline auto nclr_polar(const Matrix<real, dim>& m, Matrix<real, dim>& R, Matrix<real, dim>& S) -> void {
    if constexpr (dim == 2) {
        const real x = m(0, 0) + m(1, 1);
        const real y = m(1, 0) - m(0, 1);
        const real scale = 1.0f / std::sqrt(x * x + y * y);
        const real c = x * scale, s = y * scale;
        R = Matrix<real, 2>(c, -s, s, c);
        S = R.transpose() * m;
    } else {
        Matrix<real, dim> U, sig, V;
        nclr_svd<dim>(m, U, sig, V);
        R = U * V.transpose();
        S = V * sig * V.transpose();
    }
}

// LinSpaced(n, low, high): low + i*step, the last value exactly high (built from the high end
// when |high| < |low|).
inline std::vector<real> linspaced(int n, real low, real high) {
    std::vector<real> out(size_t(std::max(n, 0)));
    if (n == 1) out[0] = high;
    if (n <= 1) return out;
    const int size1 = n - 1;
    const real step = (high - low) / real(size1);
    const bool flip = std::fabs(high) < std::fabs(low);
    for (int i = 0; i < n; ++i)
        out[size_t(i)] = flip ? (i == 0 ? low : real(high - real(size1 - i) * step))
                              : (i == size1 ? high : real(low + real(i) * step));
    return out;
}

// cube<dim>(res, min, max): res^dim lattice over [min, max]^dim, last axis fastest
// (src/nclr_math.h:100-129).
template <int dim>
inline auto cube(int res, real min, real max) -> std::vector<Vector<real, dim>> {
    static_assert(dim == 2 || dim == 3, "cube<2> or cube<3>");
    const auto t = linspaced(res, min, max);
    std::vector<Vector<real, dim>> pts;
    if constexpr (dim == 2) {
        for (int r = 0; r < res; ++r)
            for (int c = 0; c < res; ++c) pts.emplace_back(t[r], t[c]);
    } else {
        for (int l = 0; l < res; ++l)
            for (int r = 0; r < res; ++r)
                for (int c = 0; c < res; ++c) pts.emplace_back(t[l], t[r], t[c]);
    }
    return pts;
}

inline real to_radians(real degrees) { return degrees * real(M_PI / 180.0); }

// Oblique projection used by the reference's 3D viewer (src/nclr_math.h:154-167).
inline auto pt_3d_to_2d(const Vector<real, 3>& point) -> Vector<real, 2> {
    const real phi = to_radians(28), theta = to_radians(32);
    const real c = std::cos(phi), s = std::sin(phi), C = std::cos(theta), S = std::sin(theta);
    const real x = point(0) * c + point(2) * s;
    const real z = point(2) * c - point(0) * s;
    const real v = point(1) * C + z * S;
    return Vector<real, 2>(x + 0.5f, v + 0.5f);
}

}  // namespace nclr
