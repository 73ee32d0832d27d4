// recon/linalg.h — fixed-size double vectors and matrices used by the MVS
// interface.  Same type names and member functions as the reference's
// core/include/recon/linalg.h:29-287 (Vec<N>, Mat<R,C>, Vec3d, Mat3d ...), so
// code written against the reference compiles unchanged; products and norms
// sum left to right from 0.0 as the reference does (linalg.h:242-250).
#ifndef RECON_LINALG_H_
#define RECON_LINALG_H_

#include <cmath>
#include <cstddef>

namespace recon {

constexpr double kPi = 3.14159265358979323846;

template <int N>
class Vec {
 public:
  Vec() { for (double& d : d_) d = 0.0; }
  explicit Vec(double s) { for (double& d : d_) d = s; }
  Vec(double x, double y) { static_assert(N == 2, "Vec2 only"); d_[0] = x; d_[1] = y; }
  Vec(double x, double y, double z) { static_assert(N == 3, "Vec3 only"); d_[0] = x; d_[1] = y; d_[2] = z; }
  Vec(double x, double y, double z, double w) {
    static_assert(N == 4, "Vec4 only");
    d_[0] = x; d_[1] = y; d_[2] = z; d_[3] = w;
  }
  double& operator()(int i) { return d_[i]; }
  const double& operator()(int i) const { return d_[i]; }
  double& operator[](int i) { return d_[i]; }
  const double& operator[](int i) const { return d_[i]; }
  static Vec Zero() { return Vec(); }
  static Vec Ones() { return Vec(1.0); }
  double squaredNorm() const { double s = 0.0; for (double d : d_) s += d * d; return s; }
  double norm() const { return std::sqrt(squaredNorm()); }
  Vec normalized() const {
    const double n = norm();
    if (n < 1e-15) return Vec();
    Vec o; for (int i = 0; i < N; ++i) o.d_[i] = d_[i] / n; return o;
  }
  void normalize() { const double n = norm(); if (n < 1e-15) return; for (double& d : d_) d /= n; }
  Vec operator+(const Vec& b) const { Vec o; for (int i = 0; i < N; ++i) o.d_[i] = d_[i] + b.d_[i]; return o; }
  Vec operator-(const Vec& b) const { Vec o; for (int i = 0; i < N; ++i) o.d_[i] = d_[i] - b.d_[i]; return o; }
  Vec operator-() const { Vec o; for (int i = 0; i < N; ++i) o.d_[i] = -d_[i]; return o; }
  Vec operator*(double s) const { Vec o; for (int i = 0; i < N; ++i) o.d_[i] = d_[i] * s; return o; }
  Vec operator/(double s) const { Vec o; for (int i = 0; i < N; ++i) o.d_[i] = d_[i] / s; return o; }
  Vec& operator+=(const Vec& b) { for (int i = 0; i < N; ++i) d_[i] += b.d_[i]; return *this; }
  Vec& operator-=(const Vec& b) { for (int i = 0; i < N; ++i) d_[i] -= b.d_[i]; return *this; }
  Vec& operator*=(double s) { for (double& d : d_) d *= s; return *this; }
  Vec& operator/=(double s) { if (std::fabs(s) < 1e-300) return *this; for (double& d : d_) d /= s; return *this; }
  friend Vec operator*(double s, const Vec& v) { Vec o; for (int i = 0; i < N; ++i) o.d_[i] = s * v.d_[i]; return o; }
  double dot(const Vec& b) const { double s = 0.0; for (int i = 0; i < N; ++i) s += d_[i] * b.d_[i]; return s; }
  Vec cross(const Vec& b) const {
    Vec o;
    if constexpr (N == 3) {
      o.d_[0] = d_[1] * b.d_[2] - d_[2] * b.d_[1];
      o.d_[1] = d_[2] * b.d_[0] - d_[0] * b.d_[2];
      o.d_[2] = d_[0] * b.d_[1] - d_[1] * b.d_[0];
    }
    return o;
  }
  bool isFinite() const { for (double d : d_) if (!std::isfinite(d)) return false; return true; }
  double* data() { return d_; }
  const double* data() const { return d_; }
  static constexpr int size() { return N; }

 private:
  double d_[N];
};

template <int R, int C>
class Mat {
 public:
  Mat() { for (auto& r : m_) for (double& d : r) d = 0.0; }
  static Mat Zero() { return Mat(); }
  static Mat Identity() { Mat o; for (int i = 0; i < R && i < C; ++i) o.m_[i][i] = 1.0; return o; }
  double& operator()(int i, int j) { return m_[i][j]; }
  const double& operator()(int i, int j) const { return m_[i][j]; }
  Mat operator+(const Mat& b) const { Mat o; for (int i = 0; i < R; ++i) for (int j = 0; j < C; ++j) o.m_[i][j] = m_[i][j] + b.m_[i][j]; return o; }
  Mat operator-(const Mat& b) const { Mat o; for (int i = 0; i < R; ++i) for (int j = 0; j < C; ++j) o.m_[i][j] = m_[i][j] - b.m_[i][j]; return o; }
  Mat operator*(double s) const { Mat o; for (int i = 0; i < R; ++i) for (int j = 0; j < C; ++j) o.m_[i][j] = m_[i][j] * s; return o; }
  friend Mat operator*(double s, const Mat& a) { Mat o; for (int i = 0; i < R; ++i) for (int j = 0; j < C; ++j) o.m_[i][j] = s * a.m_[i][j]; return o; }
  Mat& operator+=(const Mat& b) { for (int i = 0; i < R; ++i) for (int j = 0; j < C; ++j) m_[i][j] += b.m_[i][j]; return *this; }
  Mat& operator*=(double s) { for (auto& r : m_) for (double& d : r) d *= s; return *this; }
  template <int K>
  Mat<R, K> operator*(const Mat<C, K>& b) const {
    Mat<R, K> o;
    for (int i = 0; i < R; ++i)
      for (int j = 0; j < K; ++j) { double s = 0.0; for (int k = 0; k < C; ++k) s += m_[i][k] * b(k, j); o(i, j) = s; }
    return o;
  }
  Vec<R> operator*(const Vec<C>& v) const {
    Vec<R> o;
    for (int i = 0; i < R; ++i) { double s = 0.0; for (int j = 0; j < C; ++j) s += m_[i][j] * v(j); o(i) = s; }
    return o;
  }
  Mat<C, R> transpose() const { Mat<C, R> o; for (int i = 0; i < R; ++i) for (int j = 0; j < C; ++j) o(j, i) = m_[i][j]; return o; }
  double trace() const { double s = 0.0; for (int i = 0; i < R && i < C; ++i) s += m_[i][i]; return s; }
  bool isFinite() const { for (auto& r : m_) for (double d : r) if (!std::isfinite(d)) return false; return true; }
  double* data() { return &m_[0][0]; }
  const double* data() const { return &m_[0][0]; }

 private:
  double m_[R][C];
};

using Vec2d = Vec<2>;
using Vec3d = Vec<3>;
using Vec4d = Vec<4>;
using Mat3d = Mat<3, 3>;
using Mat4d = Mat<4, 4>;

}  // namespace recon

#endif  // RECON_LINALG_H_
