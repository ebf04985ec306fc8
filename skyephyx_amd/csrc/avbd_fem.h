// avbd_fem.h -- per-vertex FEM terms on the device: Neo-Hookean tetrahedra
// (neohookeanfem.cpp:40-110, math.cpp:13-230) and StVK membrane triangles
// (stvkfem.cpp:38-86, math.cpp:237-356).
//
// The SVD is one-sided (Hestenes) Jacobi on F; the small symmetric
// eigenproblems are cyclic Jacobi.  Any SVD with the reference's sign fix
// yields the same projected vertex Hessian (it is invariant to the choice of
// singular vectors), so these are pinned to tolerance, not to the bit.
// Everything stays in registers: loops over matrix indices are unrolled and
// run-time permutations are applied as predicated compile-time swaps.
//
// This is the compute-heavy part of the step (one SVD and one 3x3
// eigensystem per tet visit), and nothing here feeds a discrete decision
// that must match the CPU bit for bit, so it is built for throughput: a
// multiply-add written in one expression may be fused and the Jacobi rotation
// parameters and normalisations use the hardware reciprocal /
// reciprocal-sqrt / sqrt (v_rcp_f32, v_rsq_f32, v_sqrt_f32, ~1 ulp) instead of
// the correctly rounded sequences.  Contraction stays per expression
// (contract(on), not fast): which operations fuse then follows from the
// source alone, not from how a kernel around it is scheduled, so every
// instantiation of the sweep (each lane width) rounds alike.
#pragma once

#include "avbd_math.h"

#define AVBD_FLT_EPS 1.1920929e-07f

AVBD_DEV float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
AVBD_DEV float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
AVBD_DEV float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Vector helpers for the FEM terms with multiply-adds fused (the shared
// helpers of avbd_math.h keep every product and sum separately rounded for
// the contact code).
namespace fc {
AVBD_DEV float dot(v3 a, v3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
AVBD_DEV float norm(v3 a) { return fsqrt(dot(a, a)); }
AVBD_DEV v3 lin2(float a, v3 u, float b, v3 v) {
    return v3{fmaf(b, v.x, a * u.x), fmaf(b, v.y, a * u.y), fmaf(b, v.z, a * u.z)};
}
AVBD_DEV v3 lin3(float a, v3 u, float b, v3 v, float c, v3 w) {
    return v3{fmaf(c, w.x, fmaf(b, v.x, a * u.x)), fmaf(c, w.y, fmaf(b, v.y, a * u.y)), fmaf(c, w.z, fmaf(b, v.z, a * u.z))};
}
AVBD_DEV v3 mv(const m3& a, v3 v) { return v3{dot(row(a, 0), v), dot(row(a, 1), v), dot(row(a, 2), v)}; }
AVBD_DEV m3 mm(const m3& a, const m3& b) {
    m3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = fmaf(a.m[i][2], b.m[2][j], fmaf(a.m[i][1], b.m[1][j], a.m[i][0] * b.m[0][j]));
    return r;
}
AVBD_DEV float det2(float a, float b, float c, float d) { return fmaf(a, d, -(b * c)); }  // a d - b c
AVBD_DEV m3 cofactor(const m3& F) {
    m3 c;
    c.m[0][0] = det2(F.m[1][1], F.m[1][2], F.m[2][1], F.m[2][2]);
    c.m[0][1] = -det2(F.m[1][0], F.m[1][2], F.m[2][0], F.m[2][2]);
    c.m[0][2] = det2(F.m[1][0], F.m[1][1], F.m[2][0], F.m[2][1]);
    c.m[1][0] = -det2(F.m[0][1], F.m[0][2], F.m[2][1], F.m[2][2]);
    c.m[1][1] = det2(F.m[0][0], F.m[0][2], F.m[2][0], F.m[2][2]);
    c.m[1][2] = -det2(F.m[0][0], F.m[0][1], F.m[2][0], F.m[2][1]);
    c.m[2][0] = det2(F.m[0][1], F.m[0][2], F.m[1][1], F.m[1][2]);
    c.m[2][1] = -det2(F.m[0][0], F.m[0][2], F.m[1][0], F.m[1][2]);
    c.m[2][2] = det2(F.m[0][0], F.m[0][1], F.m[1][0], F.m[1][1]);
    return c;
}
// det F from its first row and cofactors
AVBD_DEV float det_cof(const m3& F, const m3& cof) {
    return fmaf(F.m[0][2], cof.m[0][2], fmaf(F.m[0][1], cof.m[0][1], F.m[0][0] * cof.m[0][0]));
}
AVBD_DEV float det3(const m3& a) { return det_cof(a, cofactor(a)); }
AVBD_DEV v3 cross(v3 a, v3 b) { return v3{det2(a.y, a.z, b.y, b.z), det2(a.z, a.x, b.z, b.x), det2(a.x, a.y, b.x, b.y)}; }
}  // namespace fc

AVBD_DEV void swap_cols(m3& a, int p, int q) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float t = a.m[r][p];
        a.m[r][p] = a.m[r][q];
        a.m[r][q] = t;
    }
}

// Jacobi rotation parameters that zero g_pq of a 2x2 symmetric block.
AVBD_DEV void jacobi_cs(float app, float aqq, float apq, float& c, float& s) {
#pragma clang fp contract(on)
    float theta = (aqq - app) * frcp(2.f * apq);
    float t = (theta >= 0.f ? 1.f : -1.f) * frcp(fabsf(theta) + fsqrt(theta * theta + 1.f));
    c = frsq(t * t + 1.f);
    s = t * c;
}

// Symmetric 3x3 eigen-decomposition, ascending eigenvalues, vectors in columns.
AVBD_DEV void sym_eig3(const m3& A, v3& w, m3& V) {
#pragma clang fp contract(on)
    m3 a = A;
    m3 v;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) v.m[i][j] = (i == j) ? 1.f : 0.f;
    for (int sweep = 0; sweep < 16; ++sweep) {
        float off = (fabsf(a.m[0][1]) + fabsf(a.m[0][2])) + fabsf(a.m[1][2]);
        float diag = (fabsf(a.m[0][0]) + fabsf(a.m[1][1])) + fabsf(a.m[2][2]);
        if (off <= AVBD_FLT_EPS * 0.25f * diag || off == 0.f) break;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int p = (k == 2) ? 1 : 0;
            const int q = (k == 0) ? 1 : 2;
            if (a.m[p][q] == 0.f) continue;
            float c, s;
            jacobi_cs(a.m[p][p], a.m[q][q], a.m[p][q], c, s);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                float arp = a.m[r][p], arq = a.m[r][q];
                a.m[r][p] = c * arp - s * arq;
                a.m[r][q] = s * arp + c * arq;
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                float apr = a.m[p][r], aqr = a.m[q][r];
                a.m[p][r] = c * apr - s * aqr;
                a.m[q][r] = s * apr + c * aqr;
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                float vrp = v.m[r][p], vrq = v.m[r][q];
                v.m[r][p] = c * vrp - s * vrq;
                v.m[r][q] = s * vrp + c * vrq;
            }
        }
    }
    float d0 = a.m[0][0], d1 = a.m[1][1], d2 = a.m[2][2];
    // stable ascending sort of (d0,d1,d2) with their columns
    if (d1 < d0) { float t = d0; d0 = d1; d1 = t; swap_cols(v, 0, 1); }
    if (d2 < d1) {
        float t = d1; d1 = d2; d2 = t; swap_cols(v, 1, 2);
        if (d1 < d0) { t = d0; d0 = d1; d1 = t; swap_cols(v, 0, 1); }
    }
    w = v3{d0, d1, d2};
    V = v;
}

// Symmetric 2x2: ascending eigenvalues w0 <= w1, vectors (c0, c1) in columns.
AVBD_DEV void sym_eig2(float a00, float a01, float a11, float& w0, float& w1, float V[2][2]) {
    float c = 1.f, s = 0.f;
    if (a01 != 0.f) jacobi_cs(a00, a11, a01, c, s);
    float d0 = (c * c * a00 - 2.f * c * s * a01) + s * s * a11;
    float d1 = (s * s * a00 + 2.f * c * s * a01) + c * c * a11;
    if (d0 <= d1) {
        w0 = d0; w1 = d1;
        V[0][0] = c; V[1][0] = -s; V[0][1] = s; V[1][1] = c;
    } else {
        w0 = d1; w1 = d0;
        V[0][0] = s; V[1][0] = c; V[0][1] = c; V[1][1] = -s;
    }
}

// F V = U S by one-sided Jacobi; S descending; U, V proper rotations with the
// sign moved into S(2) (math.cpp:22-32).
AVBD_DEV void svd3(const m3& F, m3& U, v3& S, m3& V) {
#pragma clang fp contract(on)
    m3 A = F;
    m3 W;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) W.m[i][j] = (i == j) ? 1.f : 0.f;
    for (int sweep = 0; sweep < 30; ++sweep) {
        bool rotated = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int p = (k == 2) ? 1 : 0;
            const int q = (k == 0) ? 1 : 2;
            v3 ap = col(A, p), aq = col(A, q);
            float alpha = fc::dot(ap, ap), beta = fc::dot(aq, aq), gamma = fc::dot(ap, aq);
            if (gamma * gamma <= (4.f * AVBD_FLT_EPS * AVBD_FLT_EPS) * (alpha * beta) || gamma == 0.f) continue;
            rotated = true;
            float zeta = (beta - alpha) * frcp(2.f * gamma);
            float t = (zeta >= 0.f ? 1.f : -1.f) * frcp(fabsf(zeta) + fsqrt(1.f + zeta * zeta));
            float c = frsq(1.f + t * t);
            float s = c * t;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                float x = A.m[r][p], y = A.m[r][q];
                A.m[r][p] = c * x - s * y;
                A.m[r][q] = s * x + c * y;
                float wx = W.m[r][p], wy = W.m[r][q];
                W.m[r][p] = c * wx - s * wy;
                W.m[r][q] = s * wx + c * wy;
            }
        }
        if (!rotated) break;
    }
    float s0 = fc::norm(col(A, 0)), s1 = fc::norm(col(A, 1)), s2 = fc::norm(col(A, 2));
    // stable descending sort
    if (s1 > s0) { float t = s0; s0 = s1; s1 = t; swap_cols(A, 0, 1); swap_cols(W, 0, 1); }
    if (s2 > s1) {
        float t = s1; s1 = s2; s2 = t; swap_cols(A, 1, 2); swap_cols(W, 1, 2);
        if (s1 > s0) { t = s0; s0 = s1; s1 = t; swap_cols(A, 0, 1); swap_cols(W, 0, 1); }
    }
    const float tiny = 1.17549435e-38f * 1e6f;
    v3 u0 = col(A, 0), u1 = col(A, 1), u2 = col(A, 2);
    if (s0 > tiny) u0 = mul(u0, frcp(s0)); else u0 = v3{1.f, 0.f, 0.f};
    if (s1 > tiny) u1 = mul(u1, frcp(s1));
    else {
        v3 e = fabsf(u0.x) < 0.9f ? v3{1.f, 0.f, 0.f} : v3{0.f, 1.f, 0.f};
        u1 = fc::cross(u0, e);
        u1 = mul(u1, frsq(fc::dot(u1, u1)));
    }
    if (s2 > tiny) u2 = mul(u2, frcp(s2)); else u2 = fc::cross(u0, u1);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        U.m[r][0] = comp(u0, r);
        U.m[r][1] = comp(u1, r);
        U.m[r][2] = comp(u2, r);
    }
    V = W;
    if (fc::det3(U) < 0.f) {
#pragma unroll
        for (int r = 0; r < 3; ++r) U.m[r][2] = -U.m[r][2];
        s2 = -s2;
    }
    if (fc::det3(V) < 0.f) {
#pragma unroll
        for (int r = 0; r < 3; ++r) V.m[r][2] = -V.m[r][2];
        s2 = -s2;
    }
    S = v3{s0, s1, s2};
}

// ProjectEigenvalues (math.cpp:137-171) with rho = 0 (the reference never
// updates it, solver.cpp:451-455) and threshold 0.01, epsilon 1e-6.
AVBD_DEV float project_ev(float e, int mode) {
    const float eps = 1e-6f;
    if (mode == 0) return fmaxf(eps, e);  // CLAMP
    // ABSOLUTE, and ADAPTIVE with |rho - 1| = 1 > threshold
    e = fabsf(e);
    return e < eps ? eps : e;
}

AVBD_DEV void add_outer(float Hv[6], float l, v3 w) {
    // packed lower 3x3: (0,0) (1,0) (1,1) (2,0) (2,1) (2,2)
    const v3 lw = smul(l, w);
    Hv[0] = fmaf(lw.x, w.x, Hv[0]);
    Hv[1] = fmaf(lw.y, w.x, Hv[1]);
    Hv[2] = fmaf(lw.y, w.y, Hv[2]);
    Hv[3] = fmaf(lw.z, w.x, Hv[3]);
    Hv[4] = fmaf(lw.z, w.y, Hv[4]);
    Hv[5] = fmaf(lw.z, w.z, Hv[5]);
}

// Gradient g and projected Hessian H (packed lower 3x3) of tet energy with
// respect to the vertex whose shape gradient is gradN.
AVBD_DEV void neohookean_vertex_terms(v3 p0, v3 p1, v3 p2, v3 p3, const m3& DmInv, float V0, float mu, float lam,
                                      v3 gradN, int projMode, v3& g, float H[6]) {
#pragma clang fp contract(on)
    m3 Ds;
    v3 e1 = sub(p1, p0), e2 = sub(p2, p0), e3 = sub(p3, p0);
    Ds.m[0][0] = e1.x; Ds.m[1][0] = e1.y; Ds.m[2][0] = e1.z;
    Ds.m[0][1] = e2.x; Ds.m[1][1] = e2.y; Ds.m[2][1] = e2.z;
    Ds.m[0][2] = e3.x; Ds.m[1][2] = e3.y; Ds.m[2][2] = e3.z;
    const m3 F = fc::mm(Ds, DmInv);
    const m3 cof = fc::cofactor(F);
    const float J = fc::det_cof(F, cof);
    const v3 cg = fc::mv(cof, gradN);  // dJ/dx of the vertex
    if (J <= 1e-6f) {  // inverted element (neohookeanfem.cpp:86-110)
        float stiff = mu + lam;
        float pm = stiff * (1e-6f - J);
        g = smul(-V0 * pm, cg);
        float d = V0 * stiff;
        H[0] = d; H[1] = 0.f; H[2] = d; H[3] = 0.f; H[4] = 0.f; H[5] = d;
        return;
    }
    const float logJ = logf(J);
    const float invJ = frcp(J);
    const float dPhi = (lam * logJ - mu) * invJ;
    // P gradN = mu F gradN + dPhi cof gradN
    g = smul(V0, fc::lin2(mu, fc::mv(F, gradN), dPhi, cg));
    const float diagExtra = mu + lam * (1.f - logJ);
    // Whole Hessian positive definite (every eigenvalue >= 1e-6, so the
    // projection leaves it unchanged).  With dPhi <= 0 (lam log J <= mu) the
    // flip modes mu - dPhi s_k are >= mu, the scaling block
    // mu I + (mu - lam log J) S^-2 + lam S^-1 1 1^T S^-1 is >= mu, and the
    // twist modes mu + dPhi s_k are >= 1e-6 iff s_max < T = (mu - 1e-6) / -dPhi,
    // i.e. T^2 I - F^T F is positive definite.  The vertex block of the
    // unprojected Hessian mu I + c1 cof cof^T + dPhi d2J/dF2 is then
    // mu |gradN|^2 I + c1 (cof gradN)(cof gradN)^T, c1 = diagExtra / J^2:
    // the d2J term vanishes on a single vertex (dF = dx gradN^T is rank one
    // and det(F + t dx gradN^T) is affine in t).
    if (mu > 1e-6f && dPhi <= 0.f) {
        bool pd = dPhi == 0.f;
        if (!pd) {
            const float T = (mu - 1e-6f) * frcp(-dPhi);
            const float T2 = T * T;
            const v3 f0 = col(F, 0), f1 = col(F, 1), f2 = col(F, 2);
            const float m00 = T2 - fc::dot(f0, f0), m11 = T2 - fc::dot(f1, f1), m22 = T2 - fc::dot(f2, f2);
            const float m01 = -fc::dot(f0, f1), m02 = -fc::dot(f0, f2), m12 = -fc::dot(f1, f2);
            const float d01 = m00 * m11 - m01 * m01;
            const float dM = m00 * (m11 * m22 - m12 * m12) - m01 * (m01 * m22 - m12 * m02) + m02 * (m01 * m12 - m11 * m02);
            pd = m00 > 0.f && d01 > 0.f && dM > 0.f;
        }
        if (pd) {
            const float dg = mu * fc::dot(gradN, gradN);
            const float c1 = diagExtra * invJ * invJ;
            const v3 cc = smul(c1, cg);
            H[0] = V0 * fmaf(cc.x, cg.x, dg);
            H[1] = V0 * (cc.y * cg.x);
            H[2] = V0 * fmaf(cc.y, cg.y, dg);
            H[3] = V0 * (cc.z * cg.x);
            H[4] = V0 * (cc.z * cg.y);
            H[5] = V0 * fmaf(cc.z, cg.z, dg);
            return;
        }
    }
    m3 U, V;
    v3 S;
    svd3(F, U, S, V);
    const float s0 = S.x, s1 = S.y, s2 = S.z;
    float ev3 = mu + dPhi * s2, ev4 = mu + dPhi * s1, ev5 = mu + dPhi * s0;
    float ev6 = mu - dPhi * s2, ev7 = mu - dPhi * s1, ev8 = mu - dPhi * s0;
    m3 a;
    a.m[0][0] = mu + diagExtra * frcp(s0 * s0);
    a.m[1][1] = mu + diagExtra * frcp(s1 * s1);
    a.m[2][2] = mu + diagExtra * frcp(s2 * s2);
    const float lJ = lam * invJ;
    a.m[0][1] = a.m[1][0] = lJ * s2;
    a.m[0][2] = a.m[2][0] = lJ * s1;
    a.m[1][2] = a.m[2][1] = lJ * s0;
    float ev[9] = {0.f, 0.f, 0.f, ev3, ev4, ev5, ev6, ev7, ev8};
#pragma unroll
    for (int i = 3; i < 9; ++i) ev[i] = project_ev(ev[i], projMode);
    v3 u0 = col(U, 0), u1 = col(U, 1), u2 = col(U, 2);
    float vg0 = fc::dot(col(V, 0), gradN), vg1 = fc::dot(col(V, 1), gradN), vg2 = fc::dot(col(V, 2), gradN);
    float Hv[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    // Scaling modes: with a = Va diag(aw) Va^T and the columns vg_j u_j of
    // Ut, their part of the vertex block is Ut Va proj(aw) Va^T Ut^T.  Both
    // projections (clamp and absolute, floor 1e-6) leave every eigenvalue
    // >= 1e-6 unchanged, so when a - 1e-6 I is positive definite (Sylvester's
    // criterion; the usual case) that part is Ut a Ut^T and the eigensystem
    // is not needed.  Otherwise it is projected through Jacobi as in the
    // reference.
    const float e6 = 1e-6f;
    const float b00 = a.m[0][0] - e6, b11 = a.m[1][1] - e6, b22 = a.m[2][2] - e6;
    const float m01 = b00 * b11 - a.m[0][1] * a.m[0][1];
    const float dB = b00 * (b11 * b22 - a.m[1][2] * a.m[1][2]) - a.m[0][1] * (a.m[0][1] * b22 - a.m[1][2] * a.m[0][2]) +
                     a.m[0][2] * (a.m[0][1] * a.m[1][2] - b11 * a.m[0][2]);
    if (b00 > 0.f && m01 > 0.f && dB > 0.f) {
        const v3 t0 = smul(vg0, u0), t1 = smul(vg1, u1), t2 = smul(vg2, u2);
        // rows of Ut a: r_j = sum_k a_jk t_k; block = sum_j t_j r_j^T
        const v3 r0 = fc::lin3(a.m[0][0], t0, a.m[0][1], t1, a.m[0][2], t2);
        const v3 r1 = fc::lin3(a.m[1][0], t0, a.m[1][1], t1, a.m[1][2], t2);
        const v3 r2 = fc::lin3(a.m[2][0], t0, a.m[2][1], t1, a.m[2][2], t2);
        Hv[0] = (t0.x * r0.x + t1.x * r1.x) + t2.x * r2.x;
        Hv[1] = (t0.y * r0.x + t1.y * r1.x) + t2.y * r2.x;
        Hv[2] = (t0.y * r0.y + t1.y * r1.y) + t2.y * r2.y;
        Hv[3] = (t0.z * r0.x + t1.z * r1.x) + t2.z * r2.x;
        Hv[4] = (t0.z * r0.y + t1.z * r1.y) + t2.z * r2.y;
        Hv[5] = (t0.z * r0.z + t1.z * r1.z) + t2.z * r2.z;
    } else {
        v3 aw;
        m3 av;
        sym_eig3(a, aw, av);
        const float pw[3] = {project_ev(aw.x, projMode), project_ev(aw.y, projMode), project_ev(aw.z, projMode)};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v3 c = col(av, k);
            v3 w = fc::lin3(c.x * vg0, u0, c.y * vg1, u1, c.z * vg2, u2);
            add_outer(Hv, pw[k], w);
        }
    }
    // twist (ev3-5) and flip (ev6-8) modes: w = (vg_j u_i -+ vg_i u_j) / sqrt 2
    const float is2 = 0.70710678118f;
    const float a0 = is2 * vg0, a1 = is2 * vg1, a2 = is2 * vg2;
    add_outer(Hv, ev[3], fc::lin2(a1, u0, -a0, u1));
    add_outer(Hv, ev[4], fc::lin2(a2, u0, -a0, u2));
    add_outer(Hv, ev[5], fc::lin2(a2, u1, -a1, u2));
    add_outer(Hv, ev[6], fc::lin2(a1, u0, a0, u1));
    add_outer(Hv, ev[7], fc::lin2(a2, u0, a0, u2));
    add_outer(Hv, ev[8], fc::lin2(a2, u1, a1, u2));
#pragma unroll
    for (int i = 0; i < 6; ++i) H[i] = Hv[i] * V0;
}

// StVK membrane triangle: gradient and projected Hessian for the vertex with
// 2-D shape gradient (gn0, gn1).
AVBD_DEV void stvk_vertex_terms(v3 p0, v3 p1, v3 p2, const float Dm[4], float area, float mu, float lam, float gn0,
                                float gn1, int projMode, v3& g, float H[6]) {
#pragma clang fp contract(on)
    v3 e0 = sub(p1, p0), e1 = sub(p2, p0);
    v3 f0 = add(mul(e0, Dm[0]), mul(e1, Dm[2]));
    v3 f1 = add(mul(e0, Dm[1]), mul(e1, Dm[3]));
    if (norm(cross(e0, e1)) < (2.f * area) * 0.01f) {  // stvkfem.cpp:76-86
        float stiff = mu + lam;
        v3 Pg = add(mul(mul(f0, stiff), gn0), mul(mul(f1, stiff), gn1));
        g = smul(area, Pg);
        float d = area * stiff;
        H[0] = d; H[1] = 0.f; H[2] = d; H[3] = 0.f; H[4] = 0.f; H[5] = d;
        return;
    }
    float L00 = 0.5f * (dot(f0, f0) - 1.f), L01 = 0.5f * dot(f0, f1), L11 = 0.5f * (dot(f1, f1) - 1.f);
    float tr = L00 + L11;
    float S00 = 2.f * mu * L00 + lam * tr, S01 = 2.f * mu * L01, S11 = 2.f * mu * L11 + lam * tr;
    v3 P0 = add(mul(f0, S00), mul(f1, S01));
    v3 P1 = add(mul(f0, S01), mul(f1, S11));
    g = smul(area, add(mul(P0, gn0), mul(P1, gn1)));
    // 3x2 SVD via the eigen-decomposition of F^T F
    float w0, w1, E[2][2];
    sym_eig2(dot(f0, f0), dot(f0, f1), dot(f1, f1), w0, w1, E);
    float s0 = sqrtf(fmaxf(w1, 0.f)), s1 = sqrtf(fmaxf(w0, 0.f));
    float V00 = E[0][1], V10 = E[1][1], V01 = E[0][0], V11 = E[1][0];
    v3 a0 = add(mul(f0, V00), mul(f1, V10));
    v3 a1 = add(mul(f0, V01), mul(f1, V11));
    const float tiny = 1.17549435e-38f * 1e6f;
    v3 U0 = s0 > tiny ? vdiv(a0, s0) : v3{1.f, 0.f, 0.f};
    v3 U1;
    if (s1 > tiny) U1 = vdiv(a1, s1);
    else {
        v3 e = fabsf(U0.x) < 0.9f ? v3{1.f, 0.f, 0.f} : v3{0.f, 1.f, 0.f};
        v3 c = cross(U0, e);
        U1 = vdiv(c, norm(c));
    }
    float detV = V00 * V11 - V01 * V10;
    if (detV < 0.f) { V01 = -V01; V11 = -V11; s1 = -s1; }
    const float trS = s0 * s0 + s1 * s1;
    const float d0 = s0 * (mu * (s0 * s0 - 1.f) + 0.5f * lam * (trS - 2.f));
    const float d1 = s1 * (mu * (s1 * s1 - 1.f) + 0.5f * lam * (trS - 2.f));
    float ev2 = (s0 + s1 > 1e-8f) ? (d0 + d1) / (s0 + s1) : -(mu + lam);
    float ev3 = (fabsf(s0 - s1) > 1e-8f) ? (d0 - d1) / (s0 - s1) : (3.f * mu + 2.f * lam) * s0 * s0 - mu - lam;
    float A00 = 3.f * mu * s0 * s0 + 0.5f * lam * (3.f * s0 * s0 + s1 * s1) - mu - lam;
    float A11 = 3.f * mu * s1 * s1 + 0.5f * lam * (3.f * s1 * s1 + s0 * s0) - mu - lam;
    float A01 = lam * s0 * s1;
    float aw0, aw1, AE[2][2];
    sym_eig2(A00, A01, A11, aw0, aw1, AE);
    float ev0 = project_ev(aw0, projMode), ev1 = project_ev(aw1, projMode);
    ev2 = project_ev(ev2, projMode);
    ev3 = project_ev(ev3, projMode);
    float vg0 = V00 * gn0 + V10 * gn1, vg1 = V01 * gn0 + V11 * gn1;
    float Hv[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    add_outer(Hv, ev0, add(smul(AE[0][0] * vg0, U0), smul(AE[1][0] * vg1, U1)));
    add_outer(Hv, ev1, add(smul(AE[0][1] * vg0, U0), smul(AE[1][1] * vg1, U1)));
    const float is2 = 0.70710678118f;
    add_outer(Hv, ev2, smul(is2, sub(smul(vg1, U0), smul(vg0, U1))));
    add_outer(Hv, ev3, smul(is2, add(smul(vg1, U0), smul(vg0, U1))));
#pragma unroll
    for (int i = 0; i < 6; ++i) H[i] = Hv[i] * area;
}
