// Vector types of the C++ API.  With Eigen installed (the reference's own
// dependency) the public structs use Eigen::Vector3d / Eigen::Matrix3d, so
// code written against the reference compiles unchanged.  Without Eigen the
// same names in namespace sdf are small value types with the members the
// API's callers use (x(), y(), z(), operator[], operator(), norm(), dot(),
// cross(), normalized(), cwiseProduct(), UnitX(), Zero()).
#pragma once

#include <cmath>

#if defined(__has_include)
#if __has_include(<Eigen/Dense>) && !defined(SDFO_NO_EIGEN)
#define SDFO_HAVE_EIGEN 1
#endif
#endif

#ifdef SDFO_HAVE_EIGEN
#include <Eigen/Dense>
namespace sdf {
using Vector3d = Eigen::Vector3d;
using Matrix3d = Eigen::Matrix3d;
} // namespace sdf
#else
namespace sdf {

struct Vector3d
{
    double v[3]{0.0, 0.0, 0.0};
    Vector3d() = default;
    Vector3d(double x, double y, double z) : v{x, y, z} {}
    static Vector3d Zero() { return {0.0, 0.0, 0.0}; }
    static Vector3d UnitX() { return {1.0, 0.0, 0.0}; }
    static Vector3d UnitY() { return {0.0, 1.0, 0.0}; }
    static Vector3d UnitZ() { return {0.0, 0.0, 1.0}; }
    double& x() { return v[0]; }
    double& y() { return v[1]; }
    double& z() { return v[2]; }
    double x() const { return v[0]; }
    double y() const { return v[1]; }
    double z() const { return v[2]; }
    double& operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
    double& operator()(int i) { return v[i]; }
    double operator()(int i) const { return v[i]; }
    double* data() { return v; }
    const double* data() const { return v; }
    Vector3d operator+(const Vector3d& o) const { return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]}; }
    Vector3d operator-(const Vector3d& o) const { return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]}; }
    Vector3d operator-() const { return {-v[0], -v[1], -v[2]}; }
    Vector3d operator*(double s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
    Vector3d operator/(double s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
    Vector3d& operator+=(const Vector3d& o) { return *this = *this + o; }
    Vector3d& operator-=(const Vector3d& o) { return *this = *this - o; }
    Vector3d& operator*=(double s) { return *this = *this * s; }
    Vector3d& operator/=(double s) { return *this = *this / s; }
    bool operator==(const Vector3d& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
    bool operator!=(const Vector3d& o) const { return !(*this == o); }
    double dot(const Vector3d& o) const { return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2]; }
    Vector3d cross(const Vector3d& o) const
    {
        return {v[1] * o.v[2] - v[2] * o.v[1], v[2] * o.v[0] - v[0] * o.v[2], v[0] * o.v[1] - v[1] * o.v[0]};
    }
    double squaredNorm() const { return dot(*this); }
    double norm() const { return std::sqrt(squaredNorm()); }
    Vector3d normalized() const
    {
        const double n = norm();
        return n > 0.0 ? *this / n : *this;
    }
    Vector3d cwiseProduct(const Vector3d& o) const { return {v[0] * o.v[0], v[1] * o.v[1], v[2] * o.v[2]}; }
    Vector3d cwiseQuotient(const Vector3d& o) const { return {v[0] / o.v[0], v[1] / o.v[1], v[2] / o.v[2]}; }
};

inline Vector3d operator*(double s, const Vector3d& a) { return a * s; }

struct Matrix3d
{
    double m[9]{0, 0, 0, 0, 0, 0, 0, 0, 0}; // row major
    static Matrix3d Zero() { return Matrix3d{}; }
    static Matrix3d Identity()
    {
        Matrix3d r;
        r.m[0] = r.m[4] = r.m[8] = 1.0;
        return r;
    }
    double& operator()(int i, int j) { return m[3 * i + j]; }
    double operator()(int i, int j) const { return m[3 * i + j]; }
    Vector3d operator*(const Vector3d& x) const
    {
        return {m[0] * x[0] + m[1] * x[1] + m[2] * x[2], m[3] * x[0] + m[4] * x[1] + m[5] * x[2],
                m[6] * x[0] + m[7] * x[1] + m[8] * x[2]};
    }
    Matrix3d transpose() const
    {
        Matrix3d r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.m[3 * i + j] = m[3 * j + i];
        return r;
    }
    bool operator==(const Matrix3d& o) const
    {
        for (int i = 0; i < 9; ++i)
            if (m[i] != o.m[i]) return false;
        return true;
    }
};

} // namespace sdf
#endif
