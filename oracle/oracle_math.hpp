// oracle_math.hpp -- TEST INFRASTRUCTURE ONLY.
//
// Small fixed-size linear algebra for the CPU oracle of the AVBD step.  The
// oracle is the checker that the HIP path is compared against; nothing in the
// product links it.  Templated on the scalar so the same restatement runs in
// FP32 (parity) and FP64 (error reporting).
//
// Arithmetic conventions (the HIP kernels follow the same ones so that
// discrete decisions such as SAT axis choice and clipping agree bit for bit):
//   dot3(a,b)   = (a0*b0 + a1*b1) + a2*b2
//   norm(a)     = sqrt(dot3(a,a))
//   cross(a,b)  = (a1*b2 - a2*b1, a2*b0 - a0*b2, a0*b1 - a1*b0)
//   (M v)_i     = dot3(row_i(M), v);  (M^T v)_i = dot3(col_i(M), v)
//   (A B)_ij    = (A_i0 B_0j + A_i1 B_1j) + A_i2 B_2j
//   quaternion product: Hamilton product, terms summed left to right
//   |q|^2       = ((w*w + x*x) + y*y) + z*z
//   R(q)        = the standard unit-quaternion rotation matrix with
//                 t = 2*(x,y,z) products (Eigen's toRotationMatrix formula)
//   q * v       = R(q) v
// Sums over 6-vectors run left to right.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

namespace orc {

template <class R>
struct V3 {
    R x{0}, y{0}, z{0};
    V3() = default;
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    const R& operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> operator+(const V3<R>& a, const V3<R>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(const V3<R>& a, const V3<R>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(const V3<R>& a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(const V3<R>& a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, const V3<R>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(const V3<R>& a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(const V3<R>& a, const V3<R>& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> inline R sqnorm(const V3<R>& a) { return dot(a, a); }
template <class R> inline R norm(const V3<R>& a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> cross(const V3<R>& a, const V3<R>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline V3<R> cmul(const V3<R>& a, const V3<R>& b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }

template <class R>
struct Q {
    R w{1}, x{0}, y{0}, z{0};
    Q() = default;
    Q(R w_, R x_, R y_, R z_) : w(w_), x(x_), y(y_), z(z_) {}
    V3<R> vec() const { return {x, y, z}; }
};
template <class R> inline Q<R> qmul(const Q<R>& a, const Q<R>& b) {
    return {((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z,
            ((a.w * b.x + a.x * b.w) + a.y * b.z) - a.z * b.y,
            ((a.w * b.y - a.x * b.z) + a.y * b.w) + a.z * b.x,
            ((a.w * b.z + a.x * b.y) - a.y * b.x) + a.z * b.w};
}
template <class R> inline Q<R> qconj(const Q<R>& a) { return {a.w, -a.x, -a.y, -a.z}; }
template <class R> inline R qsqnorm(const Q<R>& a) { return ((a.w * a.w + a.x * a.x) + a.y * a.y) + a.z * a.z; }
template <class R> inline Q<R> qnormalized(const Q<R>& a) {
    R n = std::sqrt(qsqnorm(a));
    if (!(n > R(0))) return a;
    return {a.w / n, a.x / n, a.y / n, a.z / n};
}

template <class R>
struct M3 {
    R m[3][3]{};
    static M3 zero() { return M3(); }
    static M3 identity() { M3 r; r.m[0][0] = r.m[1][1] = r.m[2][2] = R(1); return r; }
    V3<R> col(int j) const { return {m[0][j], m[1][j], m[2][j]}; }
    V3<R> row(int i) const { return {m[i][0], m[i][1], m[i][2]}; }
    void set_col(int j, const V3<R>& v) { m[0][j] = v.x; m[1][j] = v.y; m[2][j] = v.z; }
    M3 transpose() const { M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = m[j][i]; return r; }
};
template <class R> inline V3<R> mv(const M3<R>& a, const V3<R>& v) { return {dot(a.row(0), v), dot(a.row(1), v), dot(a.row(2), v)}; }
template <class R> inline V3<R> mtv(const M3<R>& a, const V3<R>& v) { return {dot(a.col(0), v), dot(a.col(1), v), dot(a.col(2), v)}; }
template <class R> inline M3<R> mm(const M3<R>& a, const M3<R>& b) {
    M3<R> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
    return r;
}
template <class R> inline M3<R> madd(const M3<R>& a, const M3<R>& b) {
    M3<R> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j]; return r;
}
template <class R> inline M3<R> msub(const M3<R>& a, const M3<R>& b) {
    M3<R> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] - b.m[i][j]; return r;
}
template <class R> inline M3<R> mscale(const M3<R>& a, R s) {
    M3<R> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] * s; return r;
}
template <class R> inline M3<R> mneg(const M3<R>& a) {
    M3<R> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = -a.m[i][j]; return r;
}
template <class R> inline M3<R> outer(const V3<R>& a, const V3<R>& b) {
    M3<R> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a[i] * b[j]; return r;
}
// Cofactor expansion along the first row.
template <class R> inline R det3(const M3<R>& a) {
    return (a.m[0][0] * (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1])
            - a.m[0][1] * (a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0]))
           + a.m[0][2] * (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]);
}
// Cofactor matrix (cof(F) = det(F) F^-T), as written out in math.cpp:67-78.
template <class R> inline M3<R> cofactor(const M3<R>& F) {
    M3<R> c;
    c.m[0][0] = (F.m[1][1] * F.m[2][2] - F.m[1][2] * F.m[2][1]);
    c.m[0][1] = -(F.m[1][0] * F.m[2][2] - F.m[1][2] * F.m[2][0]);
    c.m[0][2] = (F.m[1][0] * F.m[2][1] - F.m[1][1] * F.m[2][0]);
    c.m[1][0] = -(F.m[0][1] * F.m[2][2] - F.m[0][2] * F.m[2][1]);
    c.m[1][1] = (F.m[0][0] * F.m[2][2] - F.m[0][2] * F.m[2][0]);
    c.m[1][2] = -(F.m[0][0] * F.m[2][1] - F.m[0][1] * F.m[2][0]);
    c.m[2][0] = (F.m[0][1] * F.m[1][2] - F.m[0][2] * F.m[1][1]);
    c.m[2][1] = -(F.m[0][0] * F.m[1][2] - F.m[0][2] * F.m[1][0]);
    c.m[2][2] = (F.m[0][0] * F.m[1][1] - F.m[0][1] * F.m[1][0]);
    return c;
}
template <class R> inline M3<R> skew(const V3<R>& v) {
    M3<R> s;
    s.m[0][0] = R(0); s.m[0][1] = -v.z; s.m[0][2] = v.y;
    s.m[1][0] = v.z; s.m[1][1] = R(0); s.m[1][2] = -v.x;
    s.m[2][0] = -v.y; s.m[2][1] = v.x; s.m[2][2] = R(0);
    return s;
}
template <class R> inline R fro2(const M3<R>& a) {
    R s = R(0);
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) s = s + a.m[i][j] * a.m[i][j];
    return s;
}

// Rotation matrix of a unit quaternion.
template <class R> inline M3<R> rotmat(const Q<R>& q) {
    const R tx = R(2) * q.x, ty = R(2) * q.y, tz = R(2) * q.z;
    const R twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const R txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const R tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    M3<R> r;
    r.m[0][0] = R(1) - (tyy + tzz); r.m[0][1] = txy - twz; r.m[0][2] = txz + twy;
    r.m[1][0] = txy + twz; r.m[1][1] = R(1) - (txx + tzz); r.m[1][2] = tyz - twx;
    r.m[2][0] = txz - twy; r.m[2][1] = tyz + twx; r.m[2][2] = R(1) - (txx + tyy);
    return r;
}
template <class R> inline V3<R> qrot(const Q<R>& q, const V3<R>& v) { return mv(rotmat(q), v); }

template <class R>
struct V6 {
    R v[6]{};
    R& operator[](int i) { return v[i]; }
    const R& operator[](int i) const { return v[i]; }
    void set(const V3<R>& a, const V3<R>& b) { v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = b.x; v[4] = b.y; v[5] = b.z; }
};
template <class R> inline R dot6(const V6<R>& a, const V6<R>& b) {
    R s = a[0] * b[0];
    for (int i = 1; i < 6; ++i) s = s + a[i] * b[i];
    return s;
}
template <class R>
struct M6 {
    R m[6][6]{};
};

// ---- rotation helpers (math.hpp:10-63) ----

// RotationDifference (math.hpp:10-28): rotation vector of q * conj(p),
// taken in the w >= 0 hemisphere.
template <class R> inline V3<R> rotation_difference(const Q<R>& q, const Q<R>& p) {
    Q<R> dq = qmul(q, qconj(p));
    if (dq.w < R(0)) dq = Q<R>(-dq.w, -dq.x, -dq.y, -dq.z);
    V3<R> v = dq.vec();
    R s = norm(v);
    if (s < R(1e-6f)) return R(2) * v;
    R half = std::atan2(s, dq.w);
    return (R(2) * half / s) * v;
}

// QuaternionFromDifference (math.hpp:31-48).
template <class R> inline Q<R> quaternion_from_difference(const V3<R>& v, R sign) {
    R angle = norm(v);
    Q<R> q;
    if (angle < R(1e-6f)) {
        V3<R> h = (sign * R(0.5)) * v;
        q = Q<R>(R(1), h.x, h.y, h.z);
    } else {
        R half = R(0.5) * angle;
        R c = std::cos(half);
        R s = std::sin(half);
        V3<R> h = (sign * s) * (v / angle);
        q = Q<R>(c, h.x, h.y, h.z);
    }
    return qnormalized(q);
}

// QuaternionToRotVec (math.hpp:51-63).
template <class R> inline V3<R> quaternion_to_rotvec(const Q<R>& q0) {
    Q<R> q = q0.w >= R(0) ? q0 : Q<R>(-q0.w, -q0.x, -q0.y, -q0.z);
    V3<R> v = q.vec();
    R s = norm(v);
    if (s < R(1e-6f)) return V3<R>(R(0), R(0), R(0));
    R half = std::atan2(s, q.w);
    return v * (R(2) * half / s);
}

// ---- LDL^T with symmetric diagonal pivoting on a 6x6 (lower triangle) ----
// Each step takes the remaining diagonal entry of largest magnitude as pivot
// (first one on ties) and swaps it into place; a zero pivot leaves its column
// of L at zero.  The solve uses the pseudo-inverse of D: components whose
// |D_i| <= numeric_limits<R>::min() are set to zero.  These are the semantics
// of the LDLT used by the reference (solver.hpp:91, solver.cpp:437-438,585);
// the elimination order is ordinary right-looking Gaussian elimination.
template <class R>
struct LDLT6 {
    R a[6][6];   // L below the diagonal, D on it
    int perm[6]; // perm[k] = original index placed at position k
    void factor(const M6<R>& M) {
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j <= i; ++j) { a[i][j] = M.m[i][j]; a[j][i] = M.m[i][j]; }
        for (int i = 0; i < 6; ++i) perm[i] = i;
        for (int k = 0; k < 6; ++k) {
            int p = k;
            R best = std::abs(a[k][k]);
            for (int i = k + 1; i < 6; ++i) {
                R v = std::abs(a[i][i]);
                if (v > best) { best = v; p = i; }
            }
            if (p != k) {
                for (int j = 0; j < 6; ++j) std::swap(a[k][j], a[p][j]);
                for (int i = 0; i < 6; ++i) std::swap(a[i][k], a[i][p]);
                std::swap(perm[k], perm[p]);
            }
            R d = a[k][k];
            if (std::abs(d) > R(0)) {
                R l[6];
                for (int i = k + 1; i < 6; ++i) l[i] = a[i][k] / d;
                for (int i = k + 1; i < 6; ++i)
                    for (int j = k + 1; j <= i; ++j) {
                        a[i][j] = a[i][j] - l[i] * a[j][k];
                        a[j][i] = a[i][j];
                    }
                for (int i = k + 1; i < 6; ++i) { a[i][k] = l[i]; a[k][i] = l[i]; }
            } else {
                for (int i = k + 1; i < 6; ++i) { a[i][k] = R(0); a[k][i] = R(0); }
            }
        }
    }
    V6<R> solve(const V6<R>& b) const {
        R y[6];
        for (int k = 0; k < 6; ++k) y[k] = b[perm[k]];
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < i; ++j) y[i] = y[i] - a[i][j] * y[j];
        const R tol = std::numeric_limits<R>::min();
        for (int i = 0; i < 6; ++i) y[i] = (std::abs(a[i][i]) > tol) ? y[i] / a[i][i] : R(0);
        for (int i = 5; i >= 0; --i)
            for (int j = i + 1; j < 6; ++j) y[i] = y[i] - a[j][i] * y[j];
        V6<R> x;
        for (int k = 0; k < 6; ++k) x[perm[k]] = y[k];
        return x;
    }
};

// ---- symmetric eigensolvers (cyclic Jacobi; ascending eigenvalues) ----

// 3x3 symmetric: eigenvalues ascending in w, eigenvectors in the columns of V.
template <class R> inline void sym_eig3(const M3<R>& A, V3<R>& w, M3<R>& V) {
    R a[3][3];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) a[i][j] = A.m[i][j];
    M3<R> v = M3<R>::identity();
    for (int sweep = 0; sweep < 16; ++sweep) {
        R off = std::abs(a[0][1]) + std::abs(a[0][2]) + std::abs(a[1][2]);
        R diag = std::abs(a[0][0]) + std::abs(a[1][1]) + std::abs(a[2][2]);
        if (off <= std::numeric_limits<R>::epsilon() * R(0.25) * diag || off == R(0)) break;
        static const int P[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (int k = 0; k < 3; ++k) {
            int p = P[k][0], q = P[k][1];
            if (a[p][q] == R(0)) continue;
            R theta = (a[q][q] - a[p][p]) / (R(2) * a[p][q]);
            R t = (theta >= R(0) ? R(1) : R(-1)) / (std::abs(theta) + std::sqrt(theta * theta + R(1)));
            R c = R(1) / std::sqrt(t * t + R(1));
            R s = t * c;
            for (int r = 0; r < 3; ++r) {  // A <- A J (columns p,q)
                R arp = a[r][p], arq = a[r][q];
                a[r][p] = c * arp - s * arq;
                a[r][q] = s * arp + c * arq;
            }
            for (int r = 0; r < 3; ++r) {  // A <- J^T A (rows p,q)
                R apr = a[p][r], aqr = a[q][r];
                a[p][r] = c * apr - s * aqr;
                a[q][r] = s * apr + c * aqr;
            }
            for (int r = 0; r < 3; ++r) {
                R vrp = v.m[r][p], vrq = v.m[r][q];
                v.m[r][p] = c * vrp - s * vrq;
                v.m[r][q] = s * vrp + c * vrq;
            }
        }
    }
    // stable ascending sort: compare-exchange (0,1), (1,2), then (0,1) again
    R d[3] = {a[0][0], a[1][1], a[2][2]};
    auto cx = [&](int i, int j) {
        if (d[j] < d[i]) {
            std::swap(d[i], d[j]);
            V3<R> ci = v.col(i), cj = v.col(j);
            v.set_col(i, cj);
            v.set_col(j, ci);
            return true;
        }
        return false;
    };
    cx(0, 1);
    if (cx(1, 2)) cx(0, 1);
    for (int k = 0; k < 3; ++k) w[k] = d[k];
    V = v;
}

// 2x2 symmetric: ascending eigenvalues, eigenvectors in columns.
template <class R> inline void sym_eig2(R a00, R a01, R a11, R w[2], R V[2][2]) {
    R c = R(1), s = R(0);
    if (a01 != R(0)) {
        R theta = (a11 - a00) / (R(2) * a01);
        R t = (theta >= R(0) ? R(1) : R(-1)) / (std::abs(theta) + std::sqrt(theta * theta + R(1)));
        c = R(1) / std::sqrt(t * t + R(1));
        s = t * c;
    }
    // eigenvalues of the rotated matrix
    R d0 = (c * c * a00 - R(2) * c * s * a01) + s * s * a11;
    R d1 = (s * s * a00 + R(2) * c * s * a01) + c * c * a11;
    // column 0 = (c, -s), column 1 = (s, c)
    if (d0 <= d1) {
        w[0] = d0; w[1] = d1;
        V[0][0] = c; V[1][0] = -s; V[0][1] = s; V[1][1] = c;
    } else {
        w[0] = d1; w[1] = d0;
        V[0][0] = s; V[1][0] = c; V[0][1] = c; V[1][1] = -s;
    }
}

// 3x3 SVD by one-sided (Hestenes) Jacobi: F V = U S.  Singular values are
// sorted in decreasing order; then the sign fix of math.cpp:22-32 makes U and
// V rotations by negating the third column (and S(2)) where det < 0.
template <class R> inline void svd3(const M3<R>& F, M3<R>& U, V3<R>& S, M3<R>& V) {
    M3<R> A = F;
    M3<R> W = M3<R>::identity();
    const R eps = std::numeric_limits<R>::epsilon();
    static const int P[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 30; ++sweep) {
        bool rotated = false;
        for (int k = 0; k < 3; ++k) {
            int p = P[k][0], q = P[k][1];
            V3<R> ap = A.col(p), aq = A.col(q);
            R alpha = dot(ap, ap), beta = dot(aq, aq), gamma = dot(ap, aq);
            if (std::abs(gamma) <= R(2) * eps * std::sqrt(alpha * beta) || gamma == R(0)) continue;
            rotated = true;
            R zeta = (beta - alpha) / (R(2) * gamma);
            R t = (zeta >= R(0) ? R(1) : R(-1)) / (std::abs(zeta) + std::sqrt(R(1) + zeta * zeta));
            R c = R(1) / std::sqrt(R(1) + t * t);
            R s = c * t;
            A.set_col(p, c * ap - s * aq);
            A.set_col(q, s * ap + c * aq);
            V3<R> wp = W.col(p), wq = W.col(q);
            W.set_col(p, c * wp - s * wq);
            W.set_col(q, s * wp + c * wq);
        }
        if (!rotated) break;
    }
    R sig[3] = {norm(A.col(0)), norm(A.col(1)), norm(A.col(2))};
    // stable descending sort: compare-exchange (0,1), (1,2), then (0,1) again
    auto cx = [&](int i, int j) {
        if (sig[j] > sig[i]) {
            std::swap(sig[i], sig[j]);
            V3<R> a = A.col(i), b = A.col(j);
            A.set_col(i, b); A.set_col(j, a);
            V3<R> c = W.col(i), d = W.col(j);
            W.set_col(i, d); W.set_col(j, c);
            return true;
        }
        return false;
    };
    cx(0, 1);
    if (cx(1, 2)) cx(0, 1);
    M3<R> Us = A;
    for (int k = 0; k < 3; ++k) S[k] = sig[k];
    V = W;
    // normalise U columns; complete any null column with a cross product
    const R tiny = std::numeric_limits<R>::min() * R(1e6);
    V3<R> u0 = Us.col(0), u1 = Us.col(1), u2 = Us.col(2);
    if (S[0] > tiny) u0 = u0 / S[0]; else u0 = V3<R>(R(1), R(0), R(0));
    if (S[1] > tiny) u1 = u1 / S[1];
    else {
        V3<R> e = std::abs(u0.x) < R(0.9) ? V3<R>(R(1), R(0), R(0)) : V3<R>(R(0), R(1), R(0));
        u1 = cross(u0, e); u1 = u1 / norm(u1);
    }
    if (S[2] > tiny) u2 = u2 / S[2]; else u2 = cross(u0, u1);
    U.set_col(0, u0); U.set_col(1, u1); U.set_col(2, u2);
    if (det3(U) < R(0)) { U.set_col(2, -U.col(2)); S[2] = -S[2]; }
    if (det3(V) < R(0)) { V.set_col(2, -V.col(2)); S[2] = -S[2]; }
}

}  // namespace orc
