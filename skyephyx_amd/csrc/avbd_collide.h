// avbd_collide.h -- box-box contact generation on the device
// (CollisionSpace::CollisionBoxBox, collision.cpp:26-447): 15-axis SAT with
// the reference's face bias, Sutherland-Hodgman clipping of the incident face
// against the reference face's side planes, or one edge-edge contact.
// Operation order follows the arithmetic conventions of avbd_math.h so the
// axis choice, clip decisions and contact IDs match the CPU oracle exactly.
#pragma once

#include "avbd_math.h"

struct DevContact {
    v3 position, normal;
    float penetration;
    v3 rA, rB;
    uint32_t id;
};

struct DevCollision {
    DevContact c[8];
    int n;
};

AVBD_DEV uint32_t make_contact_id(int kind, int axisA, int axisB, int vertex) {
    return ((uint32_t)kind & 0x3u) | (((uint32_t)axisA & 0x7u) << 2) | (((uint32_t)axisB & 0x7u) << 5) |
           (((uint32_t)vertex & 0xFu) << 8);
}

AVBD_DEV float hcomp(v3 h, int i) { return comp(h, i); }

// PolygonFaceClip (collision.cpp:81-105)
AVBD_DEV int clip_polygon(const v3* in, int nIn, v3 n, float off, v3* out) {
    if (nIn < 1) return 0;
    int k = 0;
    for (int i = 0; i < nIn; ++i) {
        int j = (i + 1) % nIn;
        float di = dot(n, in[i]) - off;
        float dj = dot(n, in[j]) - off;
        if (di <= 0.f) out[k++] = in[i];
        if (k >= 8) return k;
        if ((di > 0.f) != (dj > 0.f)) {
            float t = di / (di - dj);
            out[k++] = add(in[i], smul(t, sub(in[j], in[i])));
            if (k >= 8) return k;
        }
    }
    return k;
}

// FaceFaceContacts (collision.cpp:108-183); (pA, rotA, hA) is the reference box.
AVBD_DEV void face_face_contacts(v3 pA, const m3& rotA, v3 hA, v3 pB, const m3& rotB, v3 hB, int refAxis,
                                 float refSign, v3 normal, int kind, int axisA, int axisB, DevCollision& out) {
    v3 refN = mul(col(rotA, refAxis), refSign);
    float refOff = dot(refN, pA) + hcomp(hA, refAxis);
    // FindIncidentFace (collision.cpp:65-77)
    float best = 3.402823466e+38f;
    int incAxis = 0;
    float incSign = 0.f;
    for (int a = 0; a < 3; ++a) {
        float d = dot(col(rotB, a), refN);
        if (d < best) { best = d; incAxis = a; incSign = 1.f; }
        if (-d < best) { best = -d; incAxis = a; incSign = -1.f; }
    }
    // GetBoxVerticesWorld (collision.cpp:36-59)
    int u = (incAxis + 1) % 3, v = (incAxis + 2) % 3;
    v3 fc = add(pB, mul(col(rotB, incAxis), incSign * hcomp(hB, incAxis)));
    v3 uv = mul(col(rotB, u), hcomp(hB, u));
    v3 vv = mul(col(rotB, v), hcomp(hB, v));
    v3 b1[8], b2[8];
    if (incSign > 0.f) {
        b1[0] = add(add(fc, uv), vv); b1[1] = sub(add(fc, uv), vv);
        b1[2] = sub(sub(fc, uv), vv); b1[3] = add(sub(fc, uv), vv);
    } else {
        b1[0] = sub(sub(fc, uv), vv); b1[1] = add(sub(fc, uv), vv);
        b1[2] = add(add(fc, uv), vv); b1[3] = sub(add(fc, uv), vv);
    }
    int ru = (refAxis + 1) % 3, rv = (refAxis + 2) % 3;
    v3 cu = col(rotA, ru), cv = col(rotA, rv);
    v3 pn[4] = {cu, neg(cu), cv, neg(cv)};
    float po[4] = {dot(cu, pA) + hcomp(hA, ru), -dot(cu, pA) + hcomp(hA, ru), dot(cv, pA) + hcomp(hA, rv),
                   -dot(cv, pA) + hcomp(hA, rv)};
    int n = 4;
    n = clip_polygon(b1, n, pn[0], po[0], b2);
    if (n < 1) return;
    n = clip_polygon(b2, n, pn[1], po[1], b1);
    if (n < 1) return;
    n = clip_polygon(b1, n, pn[2], po[2], b2);
    if (n < 1) return;
    n = clip_polygon(b2, n, pn[3], po[3], b1);
    if (n < 1) return;
    for (int i = 0; i < n; ++i) {
        float sep = dot(refN, b1[i]) - refOff;
        if (sep <= 0.f) {
            DevContact& c = out.c[out.n];
            v3 proj = sub(b1[i], mul(refN, sep));
            c.position = proj;
            c.normal = normal;
            c.penetration = -sep;
            c.rA = mtv(rotA, sub(proj, pA));
            c.rB = mtv(rotB, sub(b1[i], pB));
            c.id = make_contact_id(kind, axisA, axisB, i);
            out.n++;
            if (out.n >= 8) return;
        }
    }
}

// LineLineContacts + ClosestPointTwoLineSegments (collision.cpp:186-267)
AVBD_DEV void edge_edge_contact(v3 pA, v3 sA, const m3& rotA, v3 pB, v3 sB, const m3& rotB, v3 normal, int axisA,
                                int axisB, DevCollision& out) {
    v3 d = sub(pB, pA);
    int uA = (axisA + 1) % 3, vA = (axisA + 2) % 3;
    float sgnUA = dot(col(rotA, uA), d) > 0.f ? 1.f : -1.f;
    float sgnVA = dot(col(rotA, vA), d) > 0.f ? 1.f : -1.f;
    v3 midA = add(add(pA, mul(col(rotA, uA), (sgnUA * comp(sA, uA)) * 0.5f)), mul(col(rotA, vA), (sgnVA * comp(sA, vA)) * 0.5f));
    v3 dirA = col(rotA, axisA);
    float lenA = comp(sA, axisA);
    v3 startA = sub(midA, mul(dirA, lenA * 0.5f));
    int uB = (axisB + 1) % 3, vB = (axisB + 2) % 3;
    v3 nd = neg(d);
    float sgnUB = dot(col(rotB, uB), nd) > 0.f ? 1.f : -1.f;
    float sgnVB = dot(col(rotB, vB), nd) > 0.f ? 1.f : -1.f;
    v3 midB = add(add(pB, mul(col(rotB, uB), (sgnUB * comp(sB, uB)) * 0.5f)), mul(col(rotB, vB), (sgnVB * comp(sB, vB)) * 0.5f));
    v3 dirB = col(rotB, axisB);
    float lenB = comp(sB, axisB);
    v3 startB = sub(midB, mul(dirB, lenB * 0.5f));
    v3 d1 = mul(dirA, lenA), d2 = mul(dirB, lenB);
    v3 r = sub(startA, startB);
    float a = dot(d1, d1), b = dot(d1, d2), c = dot(d2, d2), dd = dot(d1, r), e = dot(d2, r);
    float denom = a * c - b * b;
    float s, t;
    if (denom < 1e-6f) {
        s = 0.f;
        t = (b > c ? dd / b : e / c);
    } else {
        s = (b * e - c * dd) / denom;
        t = (a * e - b * dd) / denom;
    }
    s = clampf(s, 0.f, 1.f);
    t = clampf(t, 0.f, 1.f);
    v3 onA = add(startA, mul(dirA, s * lenA));
    v3 onB = add(startB, mul(dirB, t * lenB));
    DevContact& ct = out.c[0];
    ct.position = smul(0.5f, add(onA, onB));
    ct.normal = normal;
    ct.penetration = dot(sub(onB, onA), normal);
    if (ct.penetration < 0.f) ct.penetration = -ct.penetration;
    ct.rA = mtv(rotA, sub(onA, pA));
    ct.rB = mtv(rotB, sub(onB, pB));
    ct.id = make_contact_id(2, axisA, axisB, 0);
    out.n = 1;
}

// CollisionBoxBox (collision.cpp:271-447)
AVBD_DEV void collide_boxes(v3 posA, qt qA, v3 scaleA, v3 posB, qt qB, v3 scaleB, DevCollision& res) {
    res.n = 0;
    m3 rotA = rotmat(qA), rotB = rotmat(qB);
    v3 d = sub(posB, posA);
    m3 C, absC;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            C.m[i][j] = dot(col(rotA, i), col(rotB, j));
            absC.m[i][j] = fabsf(C.m[i][j]) + 1e-6f;
        }
    v3 T = mtv(rotA, d);
    v3 hA = mul(scaleA, 0.5f), hB = mul(scaleB, 0.5f);
    float bestPen = 3.402823466e+38f;
    int bestAxis = -1;
    float bestSign = 1.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float ra = comp(scaleA, i) * 0.5f;
        float rb = dot(row(absC, i), hB);
        float Ti = comp(T, i);
        float pen = (ra + rb) - fabsf(Ti);
        float s = Ti > 0.f ? 1.f : -1.f;
        if (pen < 0.f) return;
        if (pen < bestPen) { bestPen = pen; bestAxis = i; bestSign = s; }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float ra = dot(col(absC, i), hA);
        float rb = comp(scaleB, i) * 0.5f;
        float proj = dot(col(C, i), T);
        float pen = (ra + rb) - fabsf(proj);
        float s = proj > 0.f ? 1.f : -1.f;
        if (pen < 0.f) return;
        if (pen < bestPen) { bestPen = pen; bestAxis = 3 + i; bestSign = s; }
    }
    const float bestFacePen = bestPen;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float cl = 1.f - C.m[i][j] * C.m[i][j];
            if (cl < 1e-6f) continue;
            float invLen = 1.f / sqrtf(cl);
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            float ra = (comp(scaleA, i1) * 0.5f) * absC.m[i2][j] + (comp(scaleA, i2) * 0.5f) * absC.m[i1][j];
            float rb = (comp(scaleB, j1) * 0.5f) * absC.m[i][j2] + (comp(scaleB, j2) * 0.5f) * absC.m[i][j1];
            float raw = comp(T, i2) * C.m[i1][j] - comp(T, i1) * C.m[i2][j];
            float pen = ((ra + rb) - fabsf(raw)) * invLen;
            float s = raw > 0.f ? 1.f : -1.f;
            if (pen < 0.95f * bestFacePen + 0.005f) {
                if (pen < 0.f) return;
                if (pen < bestPen) { bestPen = pen; bestAxis = 6 + i * 3 + j; bestSign = s; }
            }
        }
    }
    if (bestAxis < 0) return;
    v3 normal;
    if (bestAxis < 3) {
        normal = mul(col(rotA, bestAxis), bestSign);
    } else if (bestAxis < 6) {
        normal = mul(col(rotB, bestAxis - 3), bestSign);
    } else {
        int i = (bestAxis - 6) / 3, j = (bestAxis - 6) % 3;
        normal = cross(col(rotA, i), col(rotB, j));
        float len = norm(normal);
        if (len < 1e-6f) return;
        normal = vdiv(normal, len);
        if (dot(normal, d) < 0.f) normal = neg(normal);
    }
    if (bestAxis < 3) {
        face_face_contacts(posA, rotA, hA, posB, rotB, hB, bestAxis, bestSign, normal, 0, bestAxis, 0, res);
    } else if (bestAxis < 6) {
        int ax = bestAxis - 3;
        face_face_contacts(posB, rotB, hB, posA, rotA, hA, ax, -bestSign, normal, 1, 0, ax, res);
        for (int i = 0; i < res.n; ++i) {
            v3 t = res.c[i].rA;
            res.c[i].rA = res.c[i].rB;
            res.c[i].rB = t;
        }
    } else {
        edge_edge_contact(posA, scaleA, rotA, posB, scaleB, rotB, normal, (bestAxis - 6) / 3, (bestAxis - 6) % 3, res);
    }
}

// Sphere contacts (scene option AVBD_SCENE_SPHERE_COLLISION; the reference
// builds a sphere's mass and inertia, solver.cpp:84-87, but collides every
// pair as boxes, manifold.cpp:57).  A sphere's radius is scale.x / 2, its
// inertia's radius.  One contact per pair in collide_boxes' conventions:
// normal from A to B, penetration > 0 when overlapping, (A's point) - (B's
// point) = penetration * normal, rA / rB unscaled body-local points,
// position their midpoint; feature id kind 3 with the box region in bits 2-7
// (0-26: the 3x3x3 region of a centre outside the box, clamped onto it;
// 27 + 2 axis + (side > 0): the face a centre inside the box leaves by).
// Same operation order as the oracle, so contacts and ids match it exactly.
#define CONTACT_SPHERE 3u

AVBD_DEV void collide_spheres(v3 pA, qt qA, float radA, v3 pB, qt qB, float radB, DevCollision& res) {
    res.n = 0;
    const v3 d = sub(pB, pA);
    const float dist2 = dot(d, d);
    const float rs = radA + radB;
    if (!(dist2 < rs * rs)) return;
    const float dist = sqrtf(dist2);
    const v3 n = dist > 0.f ? vdiv(d, dist) : v3{0.f, 1.f, 0.f};
    const v3 offA = mul(n, radA), offB = mul(n, -radB);
    DevContact& c = res.c[0];
    c.position = smul(0.5f, add(add(pA, offA), add(pB, offB)));
    c.normal = n;
    c.penetration = rs - dist;
    c.rA = mtv(rotmat(qA), offA);
    c.rB = mtv(rotmat(qB), offB);
    c.id = CONTACT_SPHERE;
    res.n = 1;
}

AVBD_DEV float clamp_box(float x, float hi) { return x < -hi ? -hi : (hi < x ? hi : x); }
AVBD_DEV uint32_t box_region(float qi, float hi) { return qi == -hi ? 0u : (qi == hi ? 2u : 1u); }

// sphere (ps, qs, radius) against the box (pb, qb, scale); sphereIsA: the
// sphere is the pair's body A
AVBD_DEV void collide_sphere_box(v3 ps, qt qs, float rad, v3 pb, qt qb, v3 scale, bool sphereIsA, DevCollision& res) {
    res.n = 0;
    const m3 rb = rotmat(qb);
    const v3 c = mtv(rb, sub(ps, pb));
    const v3 h = mul(scale, 0.5f);
    const v3 q{clamp_box(c.x, h.x), clamp_box(c.y, h.y), clamp_box(c.z, h.z)};
    v3 nl, onBox;
    float pen;
    uint32_t region;
    if (q.x != c.x || q.y != c.y || q.z != c.z) {
        const v3 d = sub(c, q);
        const float dist2 = dot(d, d);
        if (!(dist2 < rad * rad)) return;
        const float dist = sqrtf(dist2);
        nl = vdiv(d, dist);
        pen = rad - dist;
        onBox = q;
        region = box_region(q.x, h.x) + 3u * box_region(q.y, h.y) + 9u * box_region(q.z, h.z);
    } else {
        // centre inside the box: out through the face of least depth (first axis on ties)
        int ax = 0;
        float best = h.x - fabsf(c.x);
        const float dy = h.y - fabsf(c.y), dz = h.z - fabsf(c.z);
        if (dy < best) { best = dy; ax = 1; }
        if (dz < best) { best = dz; ax = 2; }
        const float ca = ax == 0 ? c.x : (ax == 1 ? c.y : c.z);
        const float ha = ax == 0 ? h.x : (ax == 1 ? h.y : h.z);
        const float sg = ca >= 0.f ? 1.f : -1.f;
        nl = v3{ax == 0 ? sg : 0.f, ax == 1 ? sg : 0.f, ax == 2 ? sg : 0.f};
        pen = rad + best;
        const float face = sg * ha;
        onBox = v3{ax == 0 ? face : c.x, ax == 1 ? face : c.y, ax == 2 ? face : c.z};
        region = 27u + 2u * (uint32_t)ax + (sg > 0.f ? 1u : 0u);
    }
    const v3 nW = mv(rb, nl);  // box -> sphere
    const v3 boxW = add(mv(rb, onBox), pb);
    const v3 offS = mul(nW, -rad);
    DevContact& k = res.c[0];
    k.position = smul(0.5f, add(boxW, add(ps, offS)));
    k.penetration = pen;
    const v3 rS = mtv(rotmat(qs), offS);
    if (sphereIsA) {
        k.normal = neg(nW);
        k.rA = rS;
        k.rB = onBox;
    } else {
        k.normal = nW;
        k.rA = onBox;
        k.rB = rS;
    }
    k.id = CONTACT_SPHERE | (region << 2);
    res.n = 1;
}

// The pair's contacts: spheres as spheres when flagged, boxes otherwise.
AVBD_DEV void collide_bodies(v3 pA, qt qA, v3 sA, bool sphA, v3 pB, qt qB, v3 sB, bool sphB, DevCollision& res) {
    if (sphA && sphB) collide_spheres(pA, qA, sA.x * 0.5f, pB, qB, sB.x * 0.5f, res);
    else if (sphA) collide_sphere_box(pA, qA, sA.x * 0.5f, pB, qB, sB, true, res);
    else if (sphB) collide_sphere_box(pB, qB, sB.x * 0.5f, pA, qA, sA, false, res);
    else collide_boxes(pA, qA, sA, pB, qB, sB, res);
}
