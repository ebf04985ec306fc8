// linalg.hpp -- the small vector, quaternion and matrix types of the solver
// facade's API.
//
// The reference's physics API (src/physics/solver.hpp, force.hpp, energy.hpp,
// helpers/geometry.hpp) takes Eigen types: Eigen::Vector3f, Quaternionf =
// Eigen::Quaternion<float> (math.hpp:7), Vector6f = Eigen::Matrix<float,6,1>
// (geometry.hpp:8).  Built with -DSKYEPHYX_WITH_EIGEN the facade uses those
// very types, so callers written against the reference compile unchanged.
// Without Eigen (this image has none) the value types below stand in with
// the part of Eigen's interface the facade and its callers use:
// construction from components, x()/y()/z()/w(), operator[] and (),
// Zero/Ones/Constant/Identity, +, -, scalar *, norm, squaredNorm,
// conjugate, normalized and the Hamilton product.  The facade itself reads
// and writes components only through operator[] and x()/y()/z()/w(), which
// both sets provide.
#pragma once

#include <cmath>

#ifdef SKYEPHYX_WITH_EIGEN
#include <Eigen/Dense>
#include <Eigen/Geometry>
using Vector2f = Eigen::Vector2f;
using Vector3f = Eigen::Vector3f;
using Matrix3f = Eigen::Matrix3f;
using Vector6f = Eigen::Matrix<float, 6, 1>;
using Quaternionf = Eigen::Quaternion<float>;
#else

struct Vector2f {
    float v[2] = {0.f, 0.f};
    Vector2f() = default;
    Vector2f(float x, float y) : v{x, y} {}
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    float& operator()(int i) { return v[i]; }
    float operator()(int i) const { return v[i]; }
    float& x() { return v[0]; }
    float& y() { return v[1]; }
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    static Vector2f Zero() { return Vector2f(); }
};

struct Vector3f {
    float v[3] = {0.f, 0.f, 0.f};
    Vector3f() = default;
    Vector3f(float x, float y, float z) : v{x, y, z} {}
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    float& operator()(int i) { return v[i]; }
    float operator()(int i) const { return v[i]; }
    float& x() { return v[0]; }
    float& y() { return v[1]; }
    float& z() { return v[2]; }
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float z() const { return v[2]; }
    static Vector3f Zero() { return Vector3f(); }
    static Vector3f Ones() { return Vector3f(1.f, 1.f, 1.f); }
    static Vector3f Constant(float c) { return Vector3f(c, c, c); }
    Vector3f operator+(const Vector3f& o) const { return Vector3f(v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]); }
    Vector3f operator-(const Vector3f& o) const { return Vector3f(v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]); }
    Vector3f operator*(float s) const { return Vector3f(v[0] * s, v[1] * s, v[2] * s); }
    friend Vector3f operator*(float s, const Vector3f& a) { return a * s; }
    float squaredNorm() const { return (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]; }
    float norm() const { return std::sqrt(squaredNorm()); }
};

struct Vector6f {
    float v[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    Vector6f() = default;
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    float& operator()(int i) { return v[i]; }
    float operator()(int i) const { return v[i]; }
    static Vector6f Zero() { return Vector6f(); }
    static Vector6f Constant(float c) {
        Vector6f r;
        for (float& x : r.v) x = c;
        return r;
    }
};

struct Matrix3f {
    float m[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    float& operator()(int r, int c) { return m[r][c]; }
    float operator()(int r, int c) const { return m[r][c]; }
    static Matrix3f Zero() { return Matrix3f(); }
    static Matrix3f Identity() {
        Matrix3f r;
        r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.f;
        return r;
    }
};

// Stored and constructed (w, x, y, z), like Eigen::Quaternion's constructor.
struct Quaternionf {
    float q[4] = {1.f, 0.f, 0.f, 0.f};
    Quaternionf() = default;
    Quaternionf(float w, float x, float y, float z) : q{w, x, y, z} {}
    float& w() { return q[0]; }
    float& x() { return q[1]; }
    float& y() { return q[2]; }
    float& z() { return q[3]; }
    float w() const { return q[0]; }
    float x() const { return q[1]; }
    float y() const { return q[2]; }
    float z() const { return q[3]; }
    static Quaternionf Identity() { return Quaternionf(); }
    Quaternionf conjugate() const { return Quaternionf(q[0], -q[1], -q[2], -q[3]); }
    Quaternionf operator*(const Quaternionf& b) const {
        const float aw = q[0], ax = q[1], ay = q[2], az = q[3];
        const float bw = b.q[0], bx = b.q[1], by = b.q[2], bz = b.q[3];
        return Quaternionf(((aw * bw - ax * bx) - ay * by) - az * bz, ((aw * bx + ax * bw) + ay * bz) - az * by,
                           ((aw * by - ax * bz) + ay * bw) + az * bx, ((aw * bz + ax * by) - ay * bx) + az * bw);
    }
    float squaredNorm() const { return ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]; }
    Quaternionf normalized() const {
        const float n = std::sqrt(squaredNorm());
        return n > 0.f ? Quaternionf(q[0] / n, q[1] / n, q[2] / n, q[3] / n) : *this;
    }
};

#endif
