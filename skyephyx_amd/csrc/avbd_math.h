// avbd_math.h -- device math for the AVBD kernels (gfx950).
//
// Arithmetic conventions (kept identical in the CPU oracle so that discrete
// decisions -- SAT axis choice, clipping, contact IDs -- agree bit for bit;
// the library is compiled with -ffp-contract=off):
//   dot3(a,b)  = (a0*b0 + a1*b1) + a2*b2,   norm = sqrt(dot3(a,a))
//   cross      = (a1*b2 - a2*b1, a2*b0 - a0*b2, a0*b1 - a1*b0)
//   (M v)_i    = dot3(row_i, v),  (M^T v)_i = dot3(col_i, v)
//   (A B)_ij   = (A_i0 B_0j + A_i1 B_1j) + A_i2 B_2j
//   Hamilton product summed left to right; |q|^2 = ((w*w + x*x) + y*y) + z*z
//   R(q) from t = 2*(x,y,z) products; q * v = R(q) v
//   6-vector sums left to right.
// Quaternions are (w, x, y, z).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define AVBD_DEV __device__ __forceinline__
#define AVBD_HD __host__ __device__ __forceinline__

struct v3 {
    float x, y, z;
};
struct qt {
    float w, x, y, z;
};
struct m3 {
    float m[3][3];
};

AVBD_HD v3 mk3(float a, float b, float c) { return v3{a, b, c}; }
AVBD_HD v3 add(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
AVBD_HD v3 sub(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
AVBD_HD v3 neg(v3 a) { return v3{-a.x, -a.y, -a.z}; }
AVBD_HD v3 mul(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
AVBD_HD v3 smul(float s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
AVBD_HD v3 vdiv(v3 a, float s) { return v3{a.x / s, a.y / s, a.z / s}; }
AVBD_HD v3 cmul(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
AVBD_HD float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
AVBD_HD float norm(v3 a) { return sqrtf(dot(a, a)); }
AVBD_HD v3 cross(v3 a, v3 b) { return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
AVBD_HD float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
AVBD_HD v3 unit_axis(int i) { return v3{i == 0 ? 1.f : 0.f, i == 1 ? 1.f : 0.f, i == 2 ? 1.f : 0.f}; }

AVBD_HD qt qmul(qt a, qt b) {
    return qt{((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z, ((a.w * b.x + a.x * b.w) + a.y * b.z) - a.z * b.y,
              ((a.w * b.y - a.x * b.z) + a.y * b.w) + a.z * b.x, ((a.w * b.z + a.x * b.y) - a.y * b.x) + a.z * b.w};
}
AVBD_HD qt qconj(qt a) { return qt{a.w, -a.x, -a.y, -a.z}; }
AVBD_HD qt qnormalized(qt a) {
    float n = sqrtf(((a.w * a.w + a.x * a.x) + a.y * a.y) + a.z * a.z);
    if (!(n > 0.f)) return a;
    return qt{a.w / n, a.x / n, a.y / n, a.z / n};
}

AVBD_HD v3 col(const m3& a, int j) { return v3{a.m[0][j], a.m[1][j], a.m[2][j]}; }
AVBD_HD v3 row(const m3& a, int i) { return v3{a.m[i][0], a.m[i][1], a.m[i][2]}; }
AVBD_HD v3 mv(const m3& a, v3 v) { return v3{dot(row(a, 0), v), dot(row(a, 1), v), dot(row(a, 2), v)}; }
AVBD_HD v3 mtv(const m3& a, v3 v) { return v3{dot(col(a, 0), v), dot(col(a, 1), v), dot(col(a, 2), v)}; }
AVBD_HD m3 mm(const m3& a, const m3& b) {
    m3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
    return r;
}
AVBD_HD m3 transpose(const m3& a) {
    m3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
AVBD_HD float det3(const m3& a) {
    return (a.m[0][0] * (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) -
            a.m[0][1] * (a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0])) +
           a.m[0][2] * (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]);
}
AVBD_HD m3 cofactor(const m3& F) {
    m3 c;
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
AVBD_HD m3 skew(v3 v) {
    m3 s;
    s.m[0][0] = 0.f; s.m[0][1] = -v.z; s.m[0][2] = v.y;
    s.m[1][0] = v.z; s.m[1][1] = 0.f; s.m[1][2] = -v.x;
    s.m[2][0] = -v.y; s.m[2][1] = v.x; s.m[2][2] = 0.f;
    return s;
}

AVBD_HD m3 rotmat(qt q) {
    const float tx = 2.f * q.x, ty = 2.f * q.y, tz = 2.f * q.z;
    const float twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
    const float txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
    const float tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
    m3 r;
    r.m[0][0] = 1.f - (tyy + tzz); r.m[0][1] = txy - twz; r.m[0][2] = txz + twy;
    r.m[1][0] = txy + twz; r.m[1][1] = 1.f - (txx + tzz); r.m[1][2] = tyz - twx;
    r.m[2][0] = txz - twy; r.m[2][1] = tyz + twx; r.m[2][2] = 1.f - (txx + tyy);
    return r;
}
AVBD_HD v3 qrot(qt q, v3 v) { return mv(rotmat(q), v); }

// Rotation vector of q * conj(p) in the w >= 0 hemisphere (math.hpp:10-28).
AVBD_DEV v3 rotation_difference(qt q, qt p) {
    qt d = qmul(q, qconj(p));
    if (d.w < 0.f) d = qt{-d.w, -d.x, -d.y, -d.z};
    v3 v{d.x, d.y, d.z};
    float s = norm(v);
    if (s < 1e-6f) return smul(2.f, v);
    float half = atan2f(s, d.w);
    return smul(2.f * half / s, v);
}

// math.hpp:31-48
AVBD_DEV qt quaternion_from_difference(v3 v, float sign) {
    float angle = norm(v);
    qt q;
    if (angle < 1e-6f) {
        v3 h = smul(sign * 0.5f, v);
        q = qt{1.f, h.x, h.y, h.z};
    } else {
        float half = 0.5f * angle;
        float c = cosf(half);
        float s = sinf(half);
        v3 h = smul(sign * s, vdiv(v, angle));
        q = qt{c, h.x, h.y, h.z};
    }
    return qnormalized(q);
}

// math.hpp:51-63
AVBD_DEV v3 quaternion_to_rotvec(qt q0) {
    qt q = q0.w >= 0.f ? q0 : qt{-q0.w, -q0.x, -q0.y, -q0.z};
    v3 v{q.x, q.y, q.z};
    float s = norm(v);
    if (s < 1e-6f) return v3{0.f, 0.f, 0.f};
    float half = atan2f(s, q.w);
    return mul(v, 2.f * half / s);
}

AVBD_HD float clampf(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

// ---- packed lower-triangular 6x6 (21 entries, row-major lower) ----
AVBD_HD int tri_idx(int i, int j) { return i * (i + 1) / 2 + j; }  // i >= j

// 3x3 version of ldlt6_solve below.  For a body whose system is
// blockdiag(A, 0) (a particle with only FEM terms) the 6x6 factorisation
// never pivots onto the zero block (pivots must be strictly larger), so it
// performs exactly these operations on A and returns 0 for the rotation.
AVBD_DEV void ldlt3_solve(const float* lower6, const float* b, float* x) {
    float a[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            float v = lower6[tri_idx(i, j)];
            a[i][j] = v;
            a[j][i] = v;
        }
    int perm[3] = {0, 1, 2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int p = k;
        float best = fabsf(a[k][k]);
#pragma unroll
        for (int i = k + 1; i < 3; ++i) {
            float v = fabsf(a[i][i]);
            if (v > best) { best = v; p = i; }
        }
#pragma unroll
        for (int q = k + 1; q < 3; ++q) {
            if (p == q) {
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    float t = a[k][j];
                    a[k][j] = a[q][j];
                    a[q][j] = t;
                }
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    float t = a[i][k];
                    a[i][k] = a[i][q];
                    a[i][q] = t;
                }
                int t = perm[k];
                perm[k] = perm[q];
                perm[q] = t;
            }
        }
        float d = a[k][k];
        if (fabsf(d) > 0.f) {
            float l[3];
#pragma unroll
            for (int i = k + 1; i < 3; ++i) l[i] = a[i][k] / d;
#pragma unroll
            for (int i = k + 1; i < 3; ++i)
#pragma unroll
                for (int j = k + 1; j <= i; ++j) {
                    a[i][j] = a[i][j] - l[i] * a[j][k];
                    a[j][i] = a[i][j];
                }
#pragma unroll
            for (int i = k + 1; i < 3; ++i) { a[i][k] = l[i]; a[k][i] = l[i]; }
        } else {
#pragma unroll
            for (int i = k + 1; i < 3; ++i) { a[i][k] = 0.f; a[k][i] = 0.f; }
        }
    }
    float y[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) y[k] = perm[k] == 0 ? b[0] : (perm[k] == 1 ? b[1] : b[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) y[i] = y[i] - a[i][j] * y[j];
#pragma unroll
    for (int i = 0; i < 3; ++i) y[i] = (fabsf(a[i][i]) > 1.17549435e-38f) ? y[i] / a[i][i] : 0.f;
#pragma unroll
    for (int i = 2; i >= 0; --i)
#pragma unroll
        for (int j = i + 1; j < 3; ++j) y[i] = y[i] - a[j][i] * y[j];
#pragma unroll
    for (int q = 0; q < 3; ++q) x[q] = perm[0] == q ? y[0] : (perm[1] == q ? y[1] : y[2]);
}

// LDL^T with symmetric diagonal pivoting (largest remaining |diagonal|, first
// on ties) and a pseudo-inverse solve (|D_i| <= FLT_MIN -> component 0): the
// semantics of the LDLT in the reference primal (solver.cpp:437-438).
AVBD_DEV void ldlt6_solve(const float* lower21, const float* b, float* x) {
    float a[6][6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            float v = lower21[tri_idx(i, j)];
            a[i][j] = v;
            a[j][i] = v;
        }
    int perm[6] = {0, 1, 2, 3, 4, 5};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        int p = k;
        float best = fabsf(a[k][k]);
#pragma unroll
        for (int i = k + 1; i < 6; ++i) {
            float v = fabsf(a[i][i]);
            if (v > best) { best = v; p = i; }
        }
        // swap row/column k with the pivot p; p is a run-time value, so every
        // candidate q is a compile-time index guarded by a predicate (keeps
        // the matrix in registers)
#pragma unroll
        for (int q = k + 1; q < 6; ++q) {
            if (p == q) {
#pragma unroll
                for (int j = 0; j < 6; ++j) {
                    float t = a[k][j];
                    a[k][j] = a[q][j];
                    a[q][j] = t;
                }
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    float t = a[i][k];
                    a[i][k] = a[i][q];
                    a[i][q] = t;
                }
                int t = perm[k];
                perm[k] = perm[q];
                perm[q] = t;
            }
        }
        float d = a[k][k];
        if (fabsf(d) > 0.f) {
            float l[6];
#pragma unroll
            for (int i = k + 1; i < 6; ++i) l[i] = a[i][k] / d;
#pragma unroll
            for (int i = k + 1; i < 6; ++i)
#pragma unroll
                for (int j = k + 1; j <= i; ++j) {
                    a[i][j] = a[i][j] - l[i] * a[j][k];
                    a[j][i] = a[i][j];
                }
#pragma unroll
            for (int i = k + 1; i < 6; ++i) { a[i][k] = l[i]; a[k][i] = l[i]; }
        } else {
#pragma unroll
            for (int i = k + 1; i < 6; ++i) { a[i][k] = 0.f; a[k][i] = 0.f; }
        }
    }
    float y[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        float v = b[0];
#pragma unroll
        for (int q = 1; q < 6; ++q)
            if (perm[k] == q) v = b[q];
        y[k] = v;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) y[i] = y[i] - a[i][j] * y[j];
#pragma unroll
    for (int i = 0; i < 6; ++i) y[i] = (fabsf(a[i][i]) > 1.17549435e-38f) ? y[i] / a[i][i] : 0.f;
#pragma unroll
    for (int i = 5; i >= 0; --i)
#pragma unroll
        for (int j = i + 1; j < 6; ++j) y[i] = y[i] - a[j][i] * y[j];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        float v = y[0];
#pragma unroll
        for (int k = 1; k < 6; ++k)
            if (perm[k] == q) v = y[k];
        x[q] = v;
    }
}
