// avbd_fem.h -- per-vertex FEM terms on the device: Neo-Hookean tetrahedra
// (neohookeanfem.cpp:40-110, math.cpp:13-230) and StVK membrane triangles
// (stvkfem.cpp:38-86, math.cpp:237-356).
//
// The projected vertex Hessian is the unprojected closed form plus one
// rank-one correction per Hessian mode whose eigenvalue falls below the
// projection floor (analytic eigensystem of Smith et al., "Analytic
// Eigensystems for Isotropic Distortion Energies", 2019).  Those corrections
// need the singular vectors of F, taken from a closed-form eigen-decomposition
// of F^T F (no iteration); the result does not depend on the choice of
// singular vectors, so it is pinned to tolerance, not to the bit.
// Everything stays in registers: loops over matrix indices are unrolled.
//
// This is the compute-heavy part of the step (per tet visit the
// deformation gradient, its cofactor and, off the positive-definite case, a
// 3x3 eigensystem), and nothing here feeds a discrete decision that must
// match the CPU bit for bit, so it is built for throughput: a multiply-add
// written in one expression may be fused and the rotation parameters and
// normalisations use the hardware reciprocal /
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
AVBD_DEV float norm(v3 a) { return fsqrt(fc::dot(a, a)); }
AVBD_DEV v3 lin2(float a, v3 u, float b, v3 v) {
    return v3{fmaf(b, v.x, a * u.x), fmaf(b, v.y, a * u.y), fmaf(b, v.z, a * u.z)};
}
AVBD_DEV v3 lin3(float a, v3 u, float b, v3 v, float c, v3 w) {
    return v3{fmaf(c, w.x, fmaf(b, v.x, a * u.x)), fmaf(c, w.y, fmaf(b, v.y, a * u.y)), fmaf(c, w.z, fmaf(b, v.z, a * u.z))};
}
AVBD_DEV v3 mv(const m3& a, v3 v) { return v3{fc::dot(row(a, 0), v), fc::dot(row(a, 1), v), fc::dot(row(a, 2), v)}; }
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
AVBD_DEV float det3(const m3& a) { return fc::det_cof(a, fc::cofactor(a)); }
AVBD_DEV v3 cross(v3 a, v3 b) { return v3{det2(a.y, a.z, b.y, b.z), det2(a.z, a.x, b.z, b.x), det2(a.x, a.y, b.x, b.y)}; }
}  // namespace fc

// Jacobi rotation parameters that zero g_pq of a 2x2 symmetric block.
AVBD_DEV void jacobi_cs(float app, float aqq, float apq, float& c, float& s) {
#pragma clang fp contract(on)
    float theta = (aqq - app) * frcp(2.f * apq);
    float t = (theta >= 0.f ? 1.f : -1.f) * frcp(fabsf(theta) + fsqrt(theta * theta + 1.f));
    c = frsq(t * t + 1.f);
    s = t * c;
}

// Eigenpairs of the symmetric 3x3 [a00 a01 a02; a01 a11 a12; a02 a12 a22]
// after D. Eberly, "A Robust Eigensolver for 3x3 Symmetric Matrices"
// (Geometric Tools, 2014).  The eigenvalue farther from the other two is a
// root of the normalised characteristic cubic (the largest when
// det(A - qI) >= 0, else the smallest), found by two Newton steps from a
// chord; its vector is the longest cross product of two rows of A - e I,
// and the value is then refined to the vector's Rayleigh quotient.  The
// other pair diagonalises the 2x2 restriction of A to the orthogonal
// complement with one Jacobi rotation.  A fixed number of steps and no
// sorting: the pairs come in no particular order and the vectors are
// orthonormal to rounding, also for (near-)equal eigenvalues.
struct SymEig3 {
    float e[3];
    v3 v[3];
};
AVBD_DEV v3 sym_mv(float a00, float a01, float a02, float a11, float a12, float a22, v3 x) {
    return v3{fmaf(a02, x.z, fmaf(a01, x.y, a00 * x.x)), fmaf(a12, x.z, fmaf(a11, x.y, a01 * x.x)),
              fmaf(a22, x.z, fmaf(a12, x.y, a02 * x.x))};
}
AVBD_DEV void sym_eig3_closed(float a00, float a01, float a02, float a11, float a12, float a22, SymEig3& r) {
#pragma clang fp contract(on)
    const float q = (a00 + a11 + a22) * (1.f / 3.f);
    const float b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
    const float off = a01 * a01 + a02 * a02 + a12 * a12;
    const float p2 = (b00 * b00 + b11 * b11 + b22 * b22 + 2.f * off) * (1.f / 6.f);
    const float p = fsqrt(p2);
    // det(B / p) / 2 with B = A - q I, formed from B / p (entries of order
    // one whatever the spread p); a scalar matrix (p = 0) takes 0
    const float ip = p > 1e-30f ? frcp(p) : 0.f;
    const float n00 = b00 * ip, n11 = b11 * ip, n22 = b22 * ip, n01 = a01 * ip, n02 = a02 * ip, n12 = a12 * ip;
    const float c00 = n11 * n22 - n12 * n12, c01 = n01 * n22 - n12 * n02, c02 = n01 * n12 - n11 * n02;
    float hd = 0.5f * (n00 * c00 - n01 * c01 + n02 * c02);
    hd = fminf(fmaxf(hd, -1.f), 1.f);
    // The eigenvalues of B / p are the roots of beta^3 - 3 beta - 2 hd = 0
    // (beta = 2 cos t turns it into cos 3t = hd).  The one farthest from
    // the others, |beta| in [sqrt 3, 2] with the sign of hd, is at least
    // sqrt 3 from them, so it is well conditioned: Newton's method from the
    // chord through (0, sqrt 3) and (1, 2) in a = |hd| (start error below
    // 0.014) reaches it to within one rounding in two steps, where
    // 2 cos(acos(hd) / 3) would cost an arc cosine and a cosine.
    const float ah = fabsf(hd);
    float beta = fmaf(0.26794919f, ah, 1.73205081f);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const float b2 = beta * beta;
        const float f = fmaf(beta, b2 - 3.f, -2.f * ah);
        const float fp = fmaf(3.f, b2, -3.f);
        beta = fmaf(-f, frcp(fp), beta);
    }
    float es = q + p * (hd >= 0.f ? beta : -beta);
    // rows of A - es I scaled by their largest entry: for a nearly scalar A
    // they are tiny and their cross products' squared norms would fall
    // below the normal range, where the hardware reciprocal square root
    // reads them as zero
    const float d0 = a00 - es, d1 = a11 - es, d2 = a22 - es;
    const float mr = fmaxf(fmaxf(fmaxf(fabsf(d0), fabsf(d1)), fmaxf(fabsf(d2), fabsf(a01))), fmaxf(fabsf(a02), fabsf(a12)));
    const float im = mr > 1e-30f ? frcp(mr) : 0.f;
    const v3 r0{d0 * im, a01 * im, a02 * im}, r1{a01 * im, d1 * im, a12 * im}, r2{a02 * im, a12 * im, d2 * im};
    v3 x = fc::cross(r0, r1);
    float dx = fc::dot(x, x);
    const v3 x02 = fc::cross(r0, r2), x12 = fc::cross(r1, r2);
    const float d02 = fc::dot(x02, x02), d12 = fc::dot(x12, x12);
    if (d02 > dx) { x = x02; dx = d02; }
    if (d12 > dx) { x = x12; dx = d12; }
    const v3 vs = dx > 1e-30f ? smul(frsq(dx), x) : v3{1.f, 0.f, 0.f};
    es = fc::dot(vs, sym_mv(a00, a01, a02, a11, a12, a22, vs));
    // orthonormal basis (U, W) of the plane orthogonal to vs
    v3 U;
    if (fabsf(vs.x) > fabsf(vs.y)) {
        const float il = frsq(fmaf(vs.x, vs.x, vs.z * vs.z));
        U = v3{-vs.z * il, 0.f, vs.x * il};
    } else {
        const float il = frsq(fmaf(vs.y, vs.y, vs.z * vs.z));
        U = v3{0.f, vs.z * il, -vs.y * il};
    }
    const v3 W = fc::cross(vs, U);
    const v3 AU = sym_mv(a00, a01, a02, a11, a12, a22, U), AW = sym_mv(a00, a01, a02, a11, a12, a22, W);
    const float m00 = fc::dot(U, AU), m01 = fc::dot(U, AW), m11 = fc::dot(W, AW);
    // one Jacobi rotation zeroes m01: eigenvalues m00 - t m01, m11 + t m01
    float t = 0.f;
    if (m01 != 0.f) {
        const float theta = (m11 - m00) * frcp(2.f * m01);
        t = (theta >= 0.f ? 1.f : -1.f) * frcp(fabsf(theta) + fsqrt(fmaf(theta, theta, 1.f)));
    }
    const float c = frsq(fmaf(t, t, 1.f)), s = t * c;
    r.e[0] = es;
    r.v[0] = vs;
    r.e[1] = fmaf(-t, m01, m00);
    r.v[1] = fc::lin2(c, U, -s, W);
    r.e[2] = fmaf(t, m01, m11);
    r.v[2] = fc::lin2(s, U, c, W);
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
    const float dg = mu * fc::dot(gradN, gradN);
    const float c1 = diagExtra * invJ * invJ;
    const v3 cc = smul(c1, cg);
    float Hv[6] = {fmaf(cc.x, cg.x, dg), cc.y * cg.x, fmaf(cc.y, cg.y, dg), cc.z * cg.x, cc.z * cg.y, fmaf(cc.z, cg.z, dg)};
    const v3 f0 = col(F, 0), f1 = col(F, 1), f2 = col(F, 2);
    const float C00 = fc::dot(f0, f0), C11 = fc::dot(f1, f1), C22 = fc::dot(f2, f2);
    const float C01 = fc::dot(f0, f1), C02 = fc::dot(f0, f2), C12 = fc::dot(f1, f2);
    const bool soft_ok = mu > 1e-6f && dPhi <= 0.f;  // flip and scaling modes all >= mu
    bool pd = false;
    if (soft_ok) {
        pd = dPhi == 0.f;
        if (!pd) {
            const float T = (mu - 1e-6f) * frcp(-dPhi);
            const float T2 = T * T;
            const float m00 = T2 - C00, m11 = T2 - C11, m22 = T2 - C22;
            const float m01 = -C01, m02 = -C02, m12 = -C12;
            const float d01 = m00 * m11 - m01 * m01;
            const float dM = m00 * (m11 * m22 - m12 * m12) - m01 * (m01 * m22 - m12 * m02) + m02 * (m01 * m12 - m11 * m02);
            pd = m00 > 0.f && d01 > 0.f && dM > 0.f;
        }
    }
    if (!pd) {
        // The projection (math.cpp:137-171 on the 9 modes of math.cpp:86-130)
        // changes only the modes whose eigenvalue ev lies below the floor, so
        // the projected block is the unprojected one above plus
        // (p(ev) - ev) (Q gradN)(Q gradN)^T over those modes, Q = U X V^T the
        // mode's unit eigenmatrix (F = U S V^T).  With the eigenpairs
        // (s_k^2, v_k) of F^T F: w_k = v_k . gradN and
        // u_k = F v_k / s_k = (F + cof F) v_k / (s_k + s_i s_j)
        // (cof F v_k = s_i s_j u_k; the sum stays accurate for a small s_k).
        // Mode (i, j, k a permutation): twist ev = mu + dPhi s_k,
        // Q gradN = (w_j u_i - w_i u_j) / sqrt 2; flip ev = mu - dPhi s_k,
        // (w_j u_i + w_i u_j) / sqrt 2; scaling: eigenpairs (alpha, z) of the
        // 3x3 a below, Q gradN = sum_k z_k w_k u_k.  Any orthonormal set of
        // eigenvectors gives the same block (signs cancel in the outer
        // products; a repeated eigenvalue's modes share one correction).
        SymEig3 E;
        sym_eig3_closed(C00, C01, C02, C11, C12, C22, E);
        float s[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) s[k] = fsqrt(fmaxf(E.e[k], 0.f));
        {
            // the smallest singular value from J / (the other two): its square
            // can be lost to the rounding of F^T F
            const float p12 = s[1] * s[2], p02 = s[0] * s[2], p01 = s[0] * s[1];
            const bool k0 = E.e[0] <= E.e[1] && E.e[0] <= E.e[2];
            const bool k1 = !k0 && E.e[1] <= E.e[2];
            if (k0) { if (p12 > 0.f) s[0] = J * frcp(p12); }
            else if (k1) { if (p02 > 0.f) s[1] = J * frcp(p02); }
            else if (p01 > 0.f) s[2] = J * frcp(p01);
        }
        m3 G;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) G.m[r][c] = F.m[r][c] + cof.m[r][c];
        v3 u[3];
        float w[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int i = (k + 1) % 3, j = (k + 2) % 3;
            u[k] = smul(frcp(fmaxf(fmaf(s[i], s[j], s[k]), 1e-30f)), fc::mv(G, E.v[k]));
            w[k] = fc::dot(E.v[k], gradN);
        }
        // twist modes (the ones that fall below the floor in the usual case);
        // a mode the floor leaves has a correction of exactly 0
        float ct[3], cf[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float et = fmaf(dPhi, s[k], mu), ef = fmaf(-dPhi, s[k], mu);
            ct[k] = project_ev(et, projMode) - et;
            cf[k] = project_ev(ef, projMode) - ef;
        }
        add_outer(Hv, 0.5f * ct[0], fc::lin2(w[2], u[1], -w[1], u[2]));
        // Pair 0 is the eigensolver's best-separated one, the stretch of a
        // uniaxially stretched element and usually the only mode below the
        // floor: the other two (zero outer products otherwise, added in the
        // same order) run only where one of them is needed.
        if (ct[1] != 0.f || ct[2] != 0.f) {
            add_outer(Hv, 0.5f * ct[1], fc::lin2(w[0], u[2], -w[2], u[0]));
            add_outer(Hv, 0.5f * ct[2], fc::lin2(w[1], u[0], -w[0], u[1]));
        }
        if (cf[0] > 0.f || cf[1] > 0.f || cf[2] > 0.f) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int i = (k + 1) % 3, j = (k + 2) % 3;
                add_outer(Hv, 0.5f * cf[k], fc::lin2(w[j], u[i], w[i], u[j]));
            }
        }
        if (!soft_ok) {
            // scaling block a = diag(mu + diagExtra / s_k^2) + lam / J (s_k off the
            // diagonal in row i, column j): projected only when a - 1e-6 I is
            // not positive definite (Sylvester's criterion); with mu > 1e-6
            // and dPhi <= 0 it is >= mu I
            const float lJ = lam * invJ;
            const float a00 = mu + diagExtra * frcp(s[0] * s[0]), a11 = mu + diagExtra * frcp(s[1] * s[1]),
                        a22 = mu + diagExtra * frcp(s[2] * s[2]);
            const float a01 = lJ * s[2], a02 = lJ * s[1], a12 = lJ * s[0];
            const float e6 = 1e-6f;
            const float b00 = a00 - e6, b11 = a11 - e6, b22 = a22 - e6;
            const float m01 = b00 * b11 - a01 * a01;
            const float dB = b00 * (b11 * b22 - a12 * a12) - a01 * (a01 * b22 - a12 * a02) + a02 * (a01 * a12 - b11 * a02);
            if (!(b00 > 0.f && m01 > 0.f && dB > 0.f)) {
                SymEig3 A;
                sym_eig3_closed(a00, a01, a02, a11, a12, a22, A);
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const float ca = project_ev(A.e[m], projMode) - A.e[m];
                    const v3 z = A.v[m];
                    add_outer(Hv, ca, fc::lin3(z.x * w[0], u[0], z.y * w[1], u[1], z.z * w[2], u[2]));
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) H[i] = V0 * Hv[i];
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

// E(x - dx) - E(x) of a Neo-Hookean tet for a move of the vertex with shape
// gradient gradN, the other vertices fixed: the elastic part of the
// trust-region ratio (scene option AVBD_SCENE_TRUST_REGION; the reference
// keeps no energy values).  dF = -dx gradN^T is rank one, so
// J' = J - dx . (cof F gradN) exactly and log J' - log J = log1p(dJ / J); the
// change is formed term by term, V0 [mu (F:dF + |dF|^2 / 2) - mu dlog +
// lam dlog (log J + dlog / 2)], not as a difference of two large energies.
// The energy is the one whose gradient neohookean_vertex_terms uses
// (mu/2 (tr F^T F - 3) - mu log J + lam/2 log^2 J; for J <= 1e-6 the
// penalty (mu + lam)(1e-6 - J)^2 / 2 of neohookeanfem.cpp:86-110); a move
// across that branch takes the two energies' difference.
AVBD_DEV float neohookean_energy_change(v3 p0, v3 p1, v3 p2, v3 p3, const m3& DmInv, float V0, float mu, float lam,
                                        v3 gradN, v3 dx) {
#pragma clang fp contract(on)
    m3 Ds;
    const v3 e1 = sub(p1, p0), e2 = sub(p2, p0), e3 = sub(p3, p0);
    Ds.m[0][0] = e1.x; Ds.m[1][0] = e1.y; Ds.m[2][0] = e1.z;
    Ds.m[0][1] = e2.x; Ds.m[1][1] = e2.y; Ds.m[2][1] = e2.z;
    Ds.m[0][2] = e3.x; Ds.m[1][2] = e3.y; Ds.m[2][2] = e3.z;
    const m3 F = fc::mm(Ds, DmInv);
    const m3 cof = fc::cofactor(F);
    const float J = fc::det_cof(F, cof);
    const float dJ = -fc::dot(dx, fc::mv(cof, gradN));
    const float Jn = J + dJ;
    const float FdF = -fc::dot(dx, fc::mv(F, gradN));
    const float dF2 = fc::dot(dx, dx) * fc::dot(gradN, gradN);
    const float fl = 1e-6f;
    if (J > fl && Jn > fl) {
        const float logJ = logf(J);
        const float dl = log1pf(dJ / J);
        return V0 * ((mu * (FdF + 0.5f * dF2) - mu * dl) + (lam * dl) * (logJ + 0.5f * dl));
    }
    float trC = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) trC = trC + F.m[i][j] * F.m[i][j];
    auto psi = [&](float tc, float Jv) {
        if (Jv <= fl) return 0.5f * (mu + lam) * ((fl - Jv) * (fl - Jv));
        const float l = logf(Jv);
        return (0.5f * mu * (tc - 3.f) - mu * l) + 0.5f * lam * (l * l);
    };
    return V0 * (psi((trC + 2.f * FdF) + dF2, Jn) - psi(trC, J));
}

// The same for a StVK membrane triangle (energy mu E:E + lam/2 tr(E)^2,
// E = (F^T F - I) / 2, the one stvk_vertex_terms differentiates, and
// (mu + lam)|F|^2 / 2 when the area collapses, stvkfem.cpp:76-86): with
// dF = -dx gn^T, dE = (F^T dF + dF^T F + dF^T dF) / 2.  local: the moved
// vertex's slot (0-2).
AVBD_DEV float stvk_energy_change(v3 p0, v3 p1, v3 p2, const float Dm[4], float area, float mu, float lam, int local,
                                  float gn0, float gn1, v3 dx) {
#pragma clang fp contract(on)
    const v3 e0 = sub(p1, p0), e1 = sub(p2, p0);
    const v3 f0 = add(mul(e0, Dm[0]), mul(e1, Dm[2]));
    const v3 f1 = add(mul(e0, Dm[1]), mul(e1, Dm[3]));
    const v3 q0 = local == 0 ? sub(p0, dx) : p0, q1 = local == 1 ? sub(p1, dx) : p1, q2 = local == 2 ? sub(p2, dx) : p2;
    const v3 n0 = sub(q1, q0), n1 = sub(q2, q0);
    const float lim = (2.f * area) * 0.01f;
    const bool degOld = norm(cross(e0, e1)) < lim, degNew = norm(cross(n0, n1)) < lim;
    const float a0 = fc::dot(f0, dx), a1 = fc::dot(f1, dx), d2 = fc::dot(dx, dx);
    const float E00 = 0.5f * (fc::dot(f0, f0) - 1.f), E01 = 0.5f * fc::dot(f0, f1), E11 = 0.5f * (fc::dot(f1, f1) - 1.f);
    const float dE00 = 0.5f * (d2 * (gn0 * gn0)) - a0 * gn0;
    const float dE01 = 0.5f * ((d2 * (gn0 * gn1) - a0 * gn1) - a1 * gn0);
    const float dE11 = 0.5f * (d2 * (gn1 * gn1)) - a1 * gn1;
    if (!degOld && !degNew) {
        const float tr = E00 + E11, dtr = dE00 + dE11;
        const float dEE = 2.f * ((E00 * dE00 + 2.f * (E01 * dE01)) + E11 * dE11) +
                          ((dE00 * dE00 + 2.f * (dE01 * dE01)) + dE11 * dE11);
        return area * (mu * dEE + 0.5f * lam * (2.f * tr * dtr + dtr * dtr));
    }
    auto psi = [&](bool deg, float a, float b, float c) {
        if (deg) return 0.5f * (mu + lam) * ((2.f * a + 1.f) + (2.f * c + 1.f));
        const float t = a + c;
        return mu * ((a * a + 2.f * (b * b)) + c * c) + 0.5f * lam * (t * t);
    };
    return area * (psi(degNew, E00 + dE00, E01 + dE01, E11 + dE11) - psi(degOld, E00, E01, E11));
}
