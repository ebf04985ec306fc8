// avbd_oracle.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of SkyePhyX's AVBD step (Solver::Step and everything it
// calls), used as the checker for the HIP implementation and as the timed
// CPU baseline.  Only tests, the smoke check and the benchmark's CPU-baseline
// leg may load it.  It follows the reference files line by line:
//   src/physics/solver.{hpp,cpp}, force.{hpp,cpp}, manifold.cpp, joint.cpp,
//   spring.cpp, collision.{hpp,cpp}, energy.{hpp,cpp}, neohookeanfem.cpp,
//   stvkfem.cpp, src/helpers/math.{hpp,cpp}, src/helpers/geometry.{hpp,cpp}.
// Eigen (not available here) is replaced by oracle_math.hpp; where Eigen's
// exact rounding cannot be reproduced the result is pinned to tolerance only.
//
// One deliberate extension: the primal sweep can visit bodies in
// "colour-major" order (greedy colouring, ascending body index, then colour
// by colour).  Bodies of one colour share no force or energy, so this order
// is what a per-colour parallel sweep computes.
#pragma once

#include <chrono>
#include <cstring>
#include <memory>
#include <vector>

#include "../include/avbd_capi.h"
#include "oracle_math.hpp"

namespace orc {

enum class VisitOrder { Natural = 0, ColourMajor = 1 };

template <class R> struct Force;
template <class R> struct Energy;

// Mesh (geometry.hpp:148-256): rigid body or particle state.
template <class R>
struct Body {
    int index = -1;  // solverIndex
    V3<R> pos;
    Q<R> rot;
    V3<R> scale{R(1), R(1), R(1)};
    V3<R> vel, angVel, prevVel, prevAngVel;
    R mass = R(0);
    V3<R> inertia;  // body-frame diagonal
    R friction = R(0);
    R radius = R(0);
    bool isStatic = false;
    bool isParticle = false;
    bool isSphere = false;        // ModelType_Sphere (geometry.hpp:17-22)
    bool collidesAsSphere = false;  // isSphere under the scene's sphere collision option
    bool isDragged = false;       // Mesh::isDragged (geometry.hpp:193)
    V3<R> addedDragVelocity;      // Mesh::addedDragVelocity (geometry.hpp:194)
    V3<R> lastPos, inertialPos;
    Q<R> lastRot, inertialRot;
    M3<R> Iw;  // cached world inertia R I R^T (solver.cpp:214-231)
    std::vector<Force<R>*> forces;
    std::vector<Energy<R>*> energies;

    // GetDisplacementFromInertial (geometry.hpp:244-255)
    V6<R> displacement_from_inertial() const {
        V6<R> d;
        d.set(pos - inertialPos, rotation_difference(rot, inertialRot));
        return d;
    }
    V3<R> transform_point(const V3<R>& p) const { return qrot(rot, cmul(scale, p)) + pos; }
};

// ConstraintPointProperties (force.hpp:9-24)
template <class R>
struct Row {
    R fmin = -std::numeric_limits<R>::infinity();
    R fmax = std::numeric_limits<R>::infinity();
    R stiffness = std::numeric_limits<R>::infinity();
    R fracture = std::numeric_limits<R>::infinity();
    V6<R> J;
    M6<R> H;
    R penalty = R(0);
    R lambda = R(0);
    R C = R(0);
};

enum class ForceKind { Manifold, Joint, Spring };

// Force (force.hpp:27-78)
template <class R>
struct Force {
    ForceKind kind;
    int index = -1;  // solverIndex
    int64_t serial = 0;
    std::vector<Body<R>*> linked;
    Row<R> rows[24];
    bool isManifold = false;
    bool includeHessian = true;
    bool disabled = false;

    virtual ~Force() {
        for (Body<R>* b : linked) {
            auto& f = b->forces;
            f.erase(std::remove(f.begin(), f.end(), this), f.end());
        }
    }
    virtual bool initialize() = 0;
    virtual void compute_constraints(R alpha) = 0;
    virtual void compute_derivatives(Body<R>* b) = 0;
    virtual int num_rows() const = 0;
    void disable() {
        disabled = true;
        for (int i = 0; i < num_rows(); ++i) { rows[i].lambda = R(0); rows[i].penalty = R(0); rows[i].C = R(0); }
    }
    void reset_values() {
        for (int i = 0; i < num_rows(); ++i) { rows[i].lambda = R(0); rows[i].penalty = R(0); rows[i].C = R(0); }
    }
    void attach() {
        for (Body<R>* b : linked) b->forces.push_back(this);
    }
};

// isConstrainedTo (force.hpp:81-89)
template <class R>
bool is_constrained_to(Body<R>* a, Body<R>* b) {
    for (Force<R>* f : a->forces)
        if (std::find(f->linked.begin(), f->linked.end(), b) != f->linked.end()) return true;
    return false;
}

// ---------------------------------------------------------------- collision
// collision.hpp:8-25
template <class R>
struct Contact {
    V3<R> position, normal;
    R penetration = R(0);
    V3<R> rA, rB;
    uint32_t id = 0;
};
template <class R>
struct CollisionResult {
    Contact<R> c[8];
    int n = 0;
};

enum AxisKind { FACE_A = 0, FACE_B = 1, EDGE = 2 };

// createID (collision.cpp:26-32)
inline uint32_t make_contact_id(int kind, int axisA, int axisB, int vertex) {
    return (uint32_t(kind) & 0x3u) | ((uint32_t(axisA) & 0x7u) << 2) | ((uint32_t(axisB) & 0x7u) << 5) |
           ((uint32_t(vertex) & 0xFu) << 8);
}

// GetBoxVerticesWorld (collision.cpp:36-59): the 4 corners of one face.
template <class R>
void face_vertices(const V3<R>& c, const V3<R>& h, const M3<R>& rot, int axis, R sign, V3<R> out[4]) {
    int u = (axis + 1) % 3, v = (axis + 2) % 3;
    V3<R> fc = c + rot.col(axis) * (sign * h[axis]);
    V3<R> uv = rot.col(u) * h[u];
    V3<R> vv = rot.col(v) * h[v];
    if (sign > R(0)) {
        out[0] = (fc + uv) + vv; out[1] = (fc + uv) - vv; out[2] = (fc - uv) - vv; out[3] = (fc - uv) + vv;
    } else {
        out[0] = (fc - uv) - vv; out[1] = (fc - uv) + vv; out[2] = (fc + uv) + vv; out[3] = (fc + uv) - vv;
    }
}

// FindIncidentFace (collision.cpp:65-77)
template <class R>
void incident_face(const M3<R>& rot, const V3<R>& n, int& axis, R& sign) {
    R best = std::numeric_limits<R>::max();
    axis = 0;
    sign = R(0);
    for (int a = 0; a < 3; ++a) {
        R d = dot(rot.col(a), n);
        if (d < best) { best = d; axis = a; sign = R(1); }
        if (-d < best) { best = -d; axis = a; sign = R(-1); }
    }
}

// PolygonFaceClip (collision.cpp:81-105): Sutherland-Hodgman against one plane.
template <class R>
int clip_polygon(const V3<R>* in, int nIn, const V3<R>& n, R off, V3<R>* out) {
    if (nIn < 1) return 0;
    int k = 0;
    for (int i = 0; i < nIn; ++i) {
        int j = (i + 1) % nIn;
        R di = dot(n, in[i]) - off;
        R dj = dot(n, in[j]) - off;
        if (di <= R(0)) out[k++] = in[i];
        if (k >= 8) return k;
        if ((di > R(0)) != (dj > R(0))) {
            R t = di / (di - dj);
            out[k++] = in[i] + t * (in[j] - in[i]);
            if (k >= 8) return k;
        }
    }
    return k;
}

// FaceFaceContacts (collision.cpp:108-183).  (pA, rotA) is the reference box,
// (pB, rotB) the incident one, as passed by the caller.
template <class R>
void face_face_contacts(const V3<R>& pA, const M3<R>& rotA, const V3<R>& hA, const V3<R>& pB, const M3<R>& rotB,
                        const V3<R>& hB, int refAxis, R refSign, const V3<R>& normal, int kind, int axisA,
                        int axisB, CollisionResult<R>& out) {
    V3<R> refN = rotA.col(refAxis) * refSign;
    R refOff = dot(refN, pA) + hA[refAxis];
    int incAxis;
    R incSign;
    incident_face(rotB, refN, incAxis, incSign);
    V3<R> inc[4];
    face_vertices(pB, hB, rotB, incAxis, incSign, inc);
    int u = (refAxis + 1) % 3, v = (refAxis + 2) % 3;
    V3<R> pn[4];
    R po[4];
    pn[0] = rotA.col(u);  po[0] = dot(rotA.col(u), pA) + hA[u];
    pn[1] = -rotA.col(u); po[1] = -dot(rotA.col(u), pA) + hA[u];
    pn[2] = rotA.col(v);  po[2] = dot(rotA.col(v), pA) + hA[v];
    pn[3] = -rotA.col(v); po[3] = -dot(rotA.col(v), pA) + hA[v];
    V3<R> b1[8], b2[8];
    for (int i = 0; i < 4; ++i) b1[i] = inc[i];
    int n = 4;
    for (int p = 0; p < 4; ++p) {
        V3<R>* src = (p % 2 == 0) ? b1 : b2;
        V3<R>* dst = (p % 2 == 0) ? b2 : b1;
        n = clip_polygon(src, n, pn[p], po[p], dst);
        if (n < 1) return;
    }
    for (int i = 0; i < n; ++i) {
        R sep = dot(refN, b1[i]) - refOff;
        if (sep <= R(0)) {
            Contact<R>& c = out.c[out.n];
            V3<R> proj = b1[i] - refN * sep;
            c.position = proj;
            c.normal = normal;
            c.penetration = -sep;
            c.rA = mtv(rotA, proj - pA);
            c.rB = mtv(rotB, b1[i] - pB);
            c.id = make_contact_id(kind, axisA, axisB, i);
            out.n++;
            if (out.n >= 8) return;
        }
    }
}

// ClosestPointTwoLineSegments (collision.cpp:186-213)
template <class R>
void closest_segment_params(const V3<R>& p1, const V3<R>& d1, const V3<R>& p2, const V3<R>& d2, R& s, R& t) {
    V3<R> r = p1 - p2;
    R a = dot(d1, d1), b = dot(d1, d2), c = dot(d2, d2), d = dot(d1, r), e = dot(d2, r);
    R denom = a * c - b * b;
    if (denom < R(1e-6f)) {
        s = R(0);
        t = (b > c ? d / b : e / c);
    } else {
        s = (b * e - c * d) / denom;
        t = (a * e - b * d) / denom;
    }
    s = std::clamp(s, R(0), R(1));
    t = std::clamp(t, R(0), R(1));
}

// LineLineContacts (collision.cpp:216-267)
template <class R>
void edge_edge_contact(const V3<R>& pA, const V3<R>& sA, const M3<R>& rotA, const V3<R>& pB, const V3<R>& sB,
                       const M3<R>& rotB, const V3<R>& normal, int axisA, int axisB, CollisionResult<R>& out) {
    V3<R> d = pB - pA;
    int uA = (axisA + 1) % 3, vA = (axisA + 2) % 3;
    R sgnUA = dot(rotA.col(uA), d) > R(0) ? R(1) : R(-1);
    R sgnVA = dot(rotA.col(vA), d) > R(0) ? R(1) : R(-1);
    V3<R> midA = (pA + rotA.col(uA) * ((sgnUA * sA[uA]) * R(0.5))) + rotA.col(vA) * ((sgnVA * sA[vA]) * R(0.5));
    V3<R> dirA = rotA.col(axisA);
    R lenA = sA[axisA];
    V3<R> startA = midA - dirA * (lenA * R(0.5));
    int uB = (axisB + 1) % 3, vB = (axisB + 2) % 3;
    V3<R> nd = -d;
    R sgnUB = dot(rotB.col(uB), nd) > R(0) ? R(1) : R(-1);
    R sgnVB = dot(rotB.col(vB), nd) > R(0) ? R(1) : R(-1);
    V3<R> midB = (pB + rotB.col(uB) * ((sgnUB * sB[uB]) * R(0.5))) + rotB.col(vB) * ((sgnVB * sB[vB]) * R(0.5));
    V3<R> dirB = rotB.col(axisB);
    R lenB = sB[axisB];
    V3<R> startB = midB - dirB * (lenB * R(0.5));
    R s, t;
    closest_segment_params(startA, dirA * lenA, startB, dirB * lenB, s, t);
    V3<R> onA = startA + dirA * (s * lenA);
    V3<R> onB = startB + dirB * (t * lenB);
    Contact<R>& c = out.c[0];
    c.position = R(0.5) * (onA + onB);
    c.normal = normal;
    c.penetration = dot(onB - onA, normal);
    if (c.penetration < R(0)) c.penetration = -c.penetration;
    c.rA = mtv(rotA, onA - pA);
    c.rB = mtv(rotB, onB - pB);
    c.id = make_contact_id(EDGE, axisA, axisB, 0);
    out.n = 1;
}

// CollisionBoxBox (collision.cpp:271-447): 15-axis SAT, then face clipping or
// an edge-edge contact.
template <class R>
CollisionResult<R> collide_boxes(const V3<R>& posA, const Q<R>& qA, const V3<R>& scaleA, const V3<R>& posB,
                                 const Q<R>& qB, const V3<R>& scaleB) {
    CollisionResult<R> res;
    M3<R> rotA = rotmat(qA), rotB = rotmat(qB);
    V3<R> d = posB - posA;
    M3<R> C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C.m[i][j] = dot(rotA.col(i), rotB.col(j));
    V3<R> T = mtv(rotA, d);
    M3<R> absC;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) absC.m[i][j] = std::abs(C.m[i][j]) + R(1e-6f);
    V3<R> hA = scaleA * R(0.5), hB = scaleB * R(0.5);

    R bestPen = std::numeric_limits<R>::max();
    int bestAxis = -1;
    R bestSign = R(1);
    auto test_axis = [&](int axis, R pen, R sign) -> bool {
        if (pen < R(0)) return false;
        if (pen < bestPen) { bestPen = pen; bestAxis = axis; bestSign = sign; }
        return true;
    };
    for (int i = 0; i < 3; ++i) {
        R ra = scaleA[i] * R(0.5);
        R rb = dot(absC.row(i), hB);
        R dist = std::abs(T[i]);
        R pen = (ra + rb) - dist;
        R s = T[i] > R(0) ? R(1) : R(-1);
        if (!test_axis(i, pen, s)) return res;
    }
    for (int i = 0; i < 3; ++i) {
        R ra = dot(absC.col(i), hA);
        R rb = scaleB[i] * R(0.5);
        R proj = dot(C.col(i), T);
        R dist = std::abs(proj);
        R pen = (ra + rb) - dist;
        R s = proj > R(0) ? R(1) : R(-1);
        if (!test_axis(3 + i, pen, s)) return res;
    }
    const R bestFacePen = bestPen;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            R cl = R(1) - C.m[i][j] * C.m[i][j];
            if (cl < R(1e-6f)) continue;
            R invLen = R(1) / std::sqrt(cl);
            int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            R ra = (scaleA[i1] * R(0.5)) * absC.m[i2][j] + (scaleA[i2] * R(0.5)) * absC.m[i1][j];
            R rb = (scaleB[j1] * R(0.5)) * absC.m[i][j2] + (scaleB[j2] * R(0.5)) * absC.m[i][j1];
            R raw = T[i2] * C.m[i1][j] - T[i1] * C.m[i2][j];
            R dist = std::abs(raw);
            R pen = ((ra + rb) - dist) * invLen;
            R s = raw > R(0) ? R(1) : R(-1);
            if (pen < R(0.95f) * bestFacePen + R(0.005f)) {
                if (!test_axis(6 + i * 3 + j, pen, s)) return res;
            }
        }
    }
    if (bestAxis < 0) return res;
    V3<R> normal;
    if (bestAxis < 3) {
        normal = rotA.col(bestAxis) * bestSign;
    } else if (bestAxis < 6) {
        normal = rotB.col(bestAxis - 3) * bestSign;
    } else {
        int i = (bestAxis - 6) / 3, j = (bestAxis - 6) % 3;
        normal = cross(rotA.col(i), rotB.col(j));
        R len = norm(normal);
        if (len < R(1e-6f)) return res;
        normal = normal / len;
        if (dot(normal, d) < R(0)) normal = -normal;
    }
    if (bestAxis < 3) {
        face_face_contacts(posA, rotA, hA, posB, rotB, hB, bestAxis, bestSign, normal, FACE_A, bestAxis, 0, res);
    } else if (bestAxis < 6) {
        int ax = bestAxis - 3;
        face_face_contacts(posB, rotB, hB, posA, rotA, hA, ax, -bestSign, normal, FACE_B, 0, ax, res);
        for (int i = 0; i < res.n; ++i) std::swap(res.c[i].rA, res.c[i].rB);
    } else {
        edge_edge_contact(posA, scaleA, rotA, posB, scaleB, rotB, normal, (bestAxis - 6) / 3, (bestAxis - 6) % 3,
                          res);
    }
    return res;
}

// ---------------------------------------------------------------- forces
// Sphere contacts (scene option AVBD_SCENE_SPHERE_COLLISION).  The reference
// builds a sphere's mass and inertia (solver.cpp:84-87) but collides every
// pair as boxes (manifold.cpp:57).  A sphere's radius is scale.x / 2, the
// radius of its inertia.  One contact per pair in CollisionBoxBox's
// conventions: normal from A to B, penetration > 0 when overlapping,
// (A's point) - (B's point) = penetration * normal, rA / rB the unscaled
// body-local points, position their midpoint; feature id kind 3 with the
// box region of the contact in bits 2-7 (0-26: the 3x3x3 region of the
// clamped centre outside the box; 27-32: 27 + 2 axis + (side > 0), the face
// a centre inside the box leaves through).
static constexpr uint32_t kSphereContact = 3u;

template <class R>
CollisionResult<R> collide_spheres(const V3<R>& pA, const Q<R>& qA, R radA, const V3<R>& pB, const Q<R>& qB, R radB) {
    CollisionResult<R> res;
    const V3<R> d = pB - pA;
    const R dist2 = dot(d, d);
    const R rs = radA + radB;
    if (!(dist2 < rs * rs)) return res;
    const R dist = std::sqrt(dist2);
    const V3<R> n = dist > R(0) ? d / dist : V3<R>(R(0), R(1), R(0));
    const V3<R> offA = n * radA, offB = n * (-radB);
    Contact<R>& c = res.c[0];
    c.position = R(0.5) * ((pA + offA) + (pB + offB));
    c.normal = n;
    c.penetration = rs - dist;
    c.rA = mtv(rotmat(qA), offA);
    c.rB = mtv(rotmat(qB), offB);
    c.id = kSphereContact;
    res.n = 1;
    return res;
}

// sphere (ps, qs, radius) against the box (pb, qb, scale); sphereIsA: the
// sphere is the pair's body A
template <class R>
CollisionResult<R> collide_sphere_box(const V3<R>& ps, const Q<R>& qs, R rad, const V3<R>& pb, const Q<R>& qb,
                                      const V3<R>& scale, bool sphereIsA) {
    CollisionResult<R> res;
    const M3<R> rb = rotmat(qb);
    const V3<R> c = mtv(rb, ps - pb);  // sphere centre in the box frame
    const V3<R> h = scale * R(0.5);
    auto clampc = [](R x, R hi) { return x < -hi ? -hi : (hi < x ? hi : x); };
    const V3<R> q(clampc(c.x, h.x), clampc(c.y, h.y), clampc(c.z, h.z));
    V3<R> nl, onBox;
    R pen;
    uint32_t region;
    if (q.x != c.x || q.y != c.y || q.z != c.z) {
        const V3<R> d = c - q;
        const R dist2 = dot(d, d);
        if (!(dist2 < rad * rad)) return res;
        const R dist = std::sqrt(dist2);
        nl = d / dist;
        pen = rad - dist;
        onBox = q;
        auto reg = [](R qi, R hi) { return qi == -hi ? 0u : (qi == hi ? 2u : 1u); };
        region = reg(q.x, h.x) + 3u * reg(q.y, h.y) + 9u * reg(q.z, h.z);
    } else {
        // centre inside the box: out through the face of least depth (first axis on ties)
        int ax = 0;
        R best = h.x - std::abs(c.x);
        const R dy = h.y - std::abs(c.y), dz = h.z - std::abs(c.z);
        if (dy < best) { best = dy; ax = 1; }
        if (dz < best) { best = dz; ax = 2; }
        const R sg = c[ax] >= R(0) ? R(1) : R(-1);
        nl = V3<R>(R(0), R(0), R(0));
        nl[ax] = sg;
        pen = rad + best;
        onBox = c;
        onBox[ax] = sg * h[ax];
        region = 27u + 2u * uint32_t(ax) + (sg > R(0) ? 1u : 0u);
    }
    const V3<R> nW = mv(rb, nl);  // box -> sphere
    const V3<R> boxW = mv(rb, onBox) + pb;
    const V3<R> offS = nW * (-rad);  // sphere centre -> its contact point
    Contact<R>& k = res.c[0];
    k.position = R(0.5) * (boxW + (ps + offS));
    k.penetration = pen;
    const V3<R> rS = mtv(rotmat(qs), offS);
    if (sphereIsA) {
        k.normal = -nW;
        k.rA = rS;
        k.rB = onBox;
    } else {
        k.normal = nW;
        k.rA = onBox;
        k.rB = rS;
    }
    k.id = kSphereContact | (region << 2);
    res.n = 1;
    return res;
}

template <class R>
CollisionResult<R> collide_bodies(const Body<R>* A, const Body<R>* B) {
    const R half = R(0.5);
    if (A->collidesAsSphere && B->collidesAsSphere)
        return collide_spheres(A->pos, A->rot, A->scale.x * half, B->pos, B->rot, B->scale.x * half);
    if (A->collidesAsSphere) return collide_sphere_box(A->pos, A->rot, A->scale.x * half, B->pos, B->rot, B->scale, true);
    if (B->collidesAsSphere) return collide_sphere_box(B->pos, B->rot, B->scale.x * half, A->pos, A->rot, A->scale, false);
    return collide_boxes(A->pos, A->rot, A->scale, B->pos, B->rot, B->scale);
}

// A sweep updates the trust-region ratio only when its predicted decrease
// exceeds this fraction of the magnitude its sums cancel (FP32 ratios then
// stay within ~3e-4 of FP64 ones, well inside the 0.01 threshold).
static constexpr double kTrustNoise = 3e-5;
static constexpr float kCollisionMargin = 0.0005f;  // force.hpp:149
static constexpr float kStickThreshold = 0.01f;     // force.hpp:150

// ManifoldContactInfo (force.hpp:153-164)
template <class R>
struct ContactInfo {
    V6<R> JnA, JnB, Jt1A, Jt1B, Jt2A, Jt2B;
    V3<R> C0;
    bool stick = false;
};

// Manifold (force.hpp:167-198, manifold.cpp)
template <class R>
struct Manifold : Force<R> {
    Body<R>* A;
    Body<R>* B;
    Contact<R> contacts[8];
    ContactInfo<R> infos[8];
    int n = 0;
    R friction = R(0);
    R onPenetrationPenalty = R(0);

    Manifold(Body<R>* a, Body<R>* b, R onPen) : A(a), B(b), onPenetrationPenalty(onPen) {
        this->kind = ForceKind::Manifold;
        this->linked = {a, b};
        for (int i = 0; i < 24; i += 3) {
            this->rows[i].fmax = R(0);
            this->rows[i].fmin = -std::numeric_limits<R>::infinity();
        }
        this->isManifold = true;
        this->includeHessian = false;
        this->attach();
    }
    int num_rows() const override { return n * 3; }

    bool initialize() override {  // manifold.cpp:32-194
        friction = std::sqrt(A->friction * B->friction);
        const int oldN = n;
        Contact<R> oldC[8];
        ContactInfo<R> oldI[8];
        R oldK[24], oldL[24];
        for (int i = 0; i < oldN; ++i) { oldC[i] = contacts[i]; oldI[i] = infos[i]; }
        for (int i = 0; i < oldN * 3; ++i) { oldK[i] = this->rows[i].penalty; oldL[i] = this->rows[i].lambda; }
        CollisionResult<R> col = collide_bodies(A, B);
        n = col.n;
        if (n == 0) return false;
        for (int i = 0; i < n; ++i) contacts[i] = col.c[i];
        for (int i = 0; i < n * 3; ++i) { this->rows[i].penalty = R(0); this->rows[i].lambda = R(0); }
        for (int i = 0; i < n; ++i) infos[i] = ContactInfo<R>();
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < oldN; ++j) {
                if (oldC[j].id == contacts[i].id) {
                    for (int k = 0; k < 3; ++k) {
                        this->rows[i * 3 + k].penalty = oldK[j * 3 + k];
                        this->rows[i * 3 + k].lambda = oldL[j * 3 + k];
                    }
                    infos[i].stick = oldI[j].stick;
                    if (infos[i].stick) { contacts[i].rA = oldC[j].rA; contacts[i].rB = oldC[j].rB; }
                    break;
                }
            }
        }
        M3<R> rotA = rotmat(A->rot), rotB = rotmat(B->rot);
        for (int i = 0; i < n; ++i) {
            V3<R> nrm = -contacts[i].normal;
            V3<R> t1 = std::abs(nrm.x) > R(0.9f) ? cross(nrm, V3<R>(R(0), R(1), R(0))) : cross(nrm, V3<R>(R(1), R(0), R(0)));
            t1 = t1 / norm(t1);
            V3<R> t2 = cross(nrm, t1);
            V3<R> rAw = mv(rotA, contacts[i].rA);
            V3<R> rBw = mv(rotB, contacts[i].rB);
            ContactInfo<R>& ci = infos[i];
            ci.JnA.set(nrm, cross(rAw, nrm));
            ci.JnB.set(-nrm, -cross(rBw, nrm));
            ci.Jt1A.set(t1, cross(rAw, t1));
            ci.Jt1B.set(-t1, -cross(rBw, t1));
            ci.Jt2A.set(t2, cross(rAw, t2));
            ci.Jt2B.set(-t2, -cross(rBw, t2));
            V3<R> diff = (A->pos + rAw) - (B->pos + rBw);
            ci.C0 = V3<R>(dot(nrm, diff) + R(kCollisionMargin), dot(t1, diff), dot(t2, diff));
            bool isNew = true;
            for (int j = 0; j < oldN; ++j)
                if (oldC[j].id == contacts[i].id) { isNew = false; break; }
            if (isNew) {
                R depth = contacts[i].penetration;
                R seed = std::clamp(depth * onPenetrationPenalty, R(1), R(100000.0f));
                this->rows[i * 3 + 0].penalty = seed;
                this->rows[i * 3 + 1].penalty = R(1);
                this->rows[i * 3 + 2].penalty = R(1);
            }
        }
        return n > 0;
    }

    void compute_constraints(R alpha) override {  // manifold.cpp:197-236
        V6<R> dA, dB;
        dA.set(A->pos - A->lastPos, rotation_difference(A->rot, A->lastRot));
        dB.set(B->pos - B->lastPos, rotation_difference(B->rot, B->lastRot));
        for (int i = 0; i < n; ++i) {
            ContactInfo<R>& ci = infos[i];
            V3<R> aC0 = (R(1) - alpha) * ci.C0;
            Row<R>* r = &this->rows[i * 3];
            r[0].C = (aC0[0] + dot6(ci.JnA, dA)) + dot6(ci.JnB, dB);
            r[1].C = (aC0[1] + dot6(ci.Jt1A, dA)) + dot6(ci.Jt1B, dB);
            r[2].C = (aC0[2] + dot6(ci.Jt2A, dA)) + dot6(ci.Jt2B, dB);
            R bounds = std::abs(r[0].lambda) * friction;
            r[1].fmax = bounds; r[1].fmin = -bounds;
            r[2].fmax = bounds; r[2].fmin = -bounds;
            R tl = std::sqrt(r[1].lambda * r[1].lambda + r[2].lambda * r[2].lambda);
            R tc = std::sqrt(ci.C0[1] * ci.C0[1] + ci.C0[2] * ci.C0[2]);
            ci.stick = (tl < bounds) && (tc < R(kStickThreshold));
        }
    }
    void compute_derivatives(Body<R>* b) override {  // manifold.cpp:239-256
        for (int i = 0; i < n; ++i) {
            if (b == A) {
                this->rows[i * 3 + 0].J = infos[i].JnA;
                this->rows[i * 3 + 1].J = infos[i].Jt1A;
                this->rows[i * 3 + 2].J = infos[i].Jt2A;
            } else if (b == B) {
                this->rows[i * 3 + 0].J = infos[i].JnB;
                this->rows[i * 3 + 1].J = infos[i].Jt1B;
                this->rows[i * 3 + 2].J = infos[i].Jt2B;
            }
        }
    }
};

// Joint (force.hpp:116-143, joint.cpp)
template <class R>
struct Joint : Force<R> {
    Body<R>* A;  // may be null (world anchor)
    Body<R>* B;
    V3<R> rA, rB;
    Q<R> restRot;
    R torqueArm = R(1);
    V3<R> C0p, C0a;

    Joint(Body<R>* a, const V3<R>& ra, Body<R>* b, const V3<R>& rb) : A(a), B(b), rA(ra), rB(rb) {
        this->kind = ForceKind::Joint;
        if (a) this->linked.push_back(a);
        this->linked.push_back(b);
        this->includeHessian = true;
        this->attach();
    }
    int num_rows() const override { return 6; }

    // AngularError (joint.cpp:13-21)
    V3<R> angular_error(const Q<R>& qa, const Q<R>& qb) const {
        return quaternion_to_rotvec(qmul(qmul(qconj(qa), qb), qconj(restRot))) * torqueArm;
    }
    bool initialize() override {  // joint.cpp:69-94 (forceDisable is never set by the reference)
        if (this->disabled) return false;
        V3<R> wA = A ? A->transform_point(rA) : rA;
        V3<R> wB = B->transform_point(rB);
        C0p = wA - wB;
        C0a = angular_error(A ? A->rot : Q<R>(), B->rot);
        return true;
    }
    void compute_constraints(R alpha) override {  // joint.cpp:97-124
        V3<R> wA = A ? A->transform_point(rA) : rA;
        V3<R> wB = B->transform_point(rB);
        V3<R> cp = wA - wB;
        V3<R> ca = angular_error(A ? A->rot : Q<R>(), B->rot);
        for (int i = 0; i < 3; ++i)
            this->rows[i].C = std::isinf(this->rows[i].stiffness) ? cp[i] - C0p[i] * alpha : cp[i];
        for (int i = 0; i < 3; ++i)
            this->rows[3 + i].C = std::isinf(this->rows[3 + i].stiffness) ? ca[i] - C0a[i] * alpha : ca[i];
    }
    void compute_derivatives(Body<R>* b) override {  // joint.cpp:127-173
        bool isA = (b == A);
        Body<R>* s = isA ? A : B;
        V3<R> r = qrot(s->rot, cmul(s->scale, isA ? rA : rB));
        for (int i = 0; i < 3; ++i) {
            V3<R> e(R(0), R(0), R(0));
            e[i] = R(1);
            Row<R>& lin = this->rows[i];
            Row<R>& ang = this->rows[3 + i];
            lin.H = M6<R>();
            ang.H = M6<R>();
            if (isA) {
                lin.J.set(e, cross(r, e));
                lin.H.m[5][5] = -r[i];
                ang.J.set(V3<R>(R(0), R(0), R(0)), (-torqueArm) * e);
            } else {
                lin.J.set(-e, -cross(r, e));
                lin.H.m[5][5] = r[i];
                ang.J.set(V3<R>(R(0), R(0), R(0)), torqueArm * e);
            }
        }
    }
};

// Spring (force.hpp:92-113, spring.cpp)
template <class R>
struct Spring : Force<R> {
    Body<R>* A;  // may be null
    Body<R>* B;
    V3<R> rA, rB;
    R restLength = R(0);

    Spring(Body<R>* a, const V3<R>& ra, Body<R>* b, const V3<R>& rb) : A(a), B(b), rA(ra), rB(rb) {
        this->kind = ForceKind::Spring;
        if (a) this->linked.push_back(a);
        this->linked.push_back(b);
        this->includeHessian = true;
        this->attach();
    }
    int num_rows() const override { return 1; }
    bool initialize() override { return !this->disabled; }
    void compute_constraints(R) override {  // spring.cpp:41-48
        V3<R> pA = A ? A->transform_point(rA) : rA;
        V3<R> pB = B->transform_point(rB);
        this->rows[0].C = norm(pB - pA) - restLength;
    }
    void compute_derivatives(Body<R>* b) override {  // spring.cpp:51-109
        V3<R> pA = A ? A->transform_point(rA) : rA;
        V3<R> pB = B->transform_point(rB);
        V3<R> d = pB - pA;
        R L = norm(d);
        R L2 = sqnorm(d);
        Row<R>& row = this->rows[0];
        if (L2 < R(1e-6f)) { row.J = V6<R>(); row.H = M6<R>(); return; }
        V3<R> n = d / L;
        M3<R> dxx = msub(M3<R>::identity(), outer(n, n));
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) dxx.m[i][j] = dxx.m[i][j] / L;
        M3<R> dxr, drr;
        if (b == A && A != nullptr) {
            V3<R> r = qrot(A->rot, rA);  // unscaled, as in spring.cpp:76
            M3<R> Sr = mneg(skew(r));
            dxr = mm(dxx, Sr.transpose());
            drr = msub(mm(mm(Sr, dxx), Sr.transpose()), mm(skew(n), skew(r)));
            row.J.set(-n, cross(n, r));
        } else {
            V3<R> r = qrot(B->rot, rB);  // unscaled, as in spring.cpp:94
            M3<R> mSr = skew(r);
            dxr = mm(dxx, mSr.transpose());
            drr = madd(mm(mm(mSr, dxx), mSr.transpose()), mm(skew(n), skew(r)));
            row.J.set(n, cross(r, n));
        }
        M3<R> dxrT = dxr.transpose();
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                row.H.m[i][j] = dxx.m[i][j];
                row.H.m[i][3 + j] = dxr.m[i][j];
                row.H.m[3 + i][j] = dxrT.m[i][j];
                row.H.m[3 + i][3 + j] = drr.m[i][j];
            }
    }
};

// ---------------------------------------------------------------- energies
template <class R>
struct Energy {
    int index = -1;
    int32_t serial = 0;
    std::vector<Body<R>*> linked;
    V6<R> grad;
    M6<R> hess;
    virtual ~Energy() {
        for (Body<R>* b : linked) {
            auto& e = b->energies;
            e.erase(std::remove(e.begin(), e.end(), this), e.end());
        }
    }
    virtual void compute_terms(Body<R>* b, int projectionMode, R rho) = 0;
    // E(vertex b at x - dx) - E(x), the other vertices fixed (the elastic
    // part of the trust-region ratio; not in the reference)
    virtual R energy_change(Body<R>* b, const V3<R>& dx) const = 0;
    void attach() {
        for (Body<R>* b : linked) b->energies.push_back(this);
    }
};

// ProjectEigenvalues (math.cpp:137-171); ADAPTIVE falls to ABSOLUTE unless
// |rho - 1| <= threshold, and rho stays 0 in the reference (solver.cpp:451-455).
template <class R>
void project_eigenvalues(R* ev, int count, int mode, R rho = R(0), R threshold = R(0.01f), R eps = R(1e-6f)) {
    bool absolute = false;
    if (mode == AVBD_PROJ_CLAMP) {
        for (int i = 0; i < count; ++i) ev[i] = std::max(eps, ev[i]);
        return;
    } else if (mode == AVBD_PROJ_ABSOLUTE) {
        absolute = true;
    } else {
        absolute = std::abs(rho - R(1)) > threshold;
        if (!absolute) {
            for (int i = 0; i < count; ++i) ev[i] = std::max(eps, ev[i]);
            return;
        }
    }
    if (absolute)
        for (int i = 0; i < count; ++i) {
            ev[i] = std::abs(ev[i]);
            if (ev[i] < eps) ev[i] = eps;
        }
}

// Per-vertex Neo-Hookean gradient (3) and projected Hessian (3x3) of one tet,
// neohookeanfem.cpp:40-110 with NeoHookeanMath (math.cpp:13-230).
template <class R>
void neohookean_vertex_terms(const V3<R> p[4], const M3<R>& DmInv, R V0, R mu, R lam, int vtx, int projMode, R rho,
                             V3<R>& g, M3<R>& H) {
    M3<R> Ds;
    Ds.set_col(0, p[1] - p[0]);
    Ds.set_col(1, p[2] - p[0]);
    Ds.set_col(2, p[3] - p[0]);
    M3<R> F = mm(Ds, DmInv);
    R J = det3(F);
    V3<R> gN[4];
    gN[1] = DmInv.row(0);
    gN[2] = DmInv.row(1);
    gN[3] = DmInv.row(2);
    gN[0] = (-gN[1] - gN[2]) - gN[3];
    const V3<R> gradN = gN[vtx];
    if (J <= R(1e-6f)) {  // HandleInvertedElement (neohookeanfem.cpp:86-110)
        R stiff = mu + lam;
        V3<R> dJ = mv(cofactor(F), gradN);
        R pm = stiff * (R(1e-6f) - J);
        g = (-V0 * pm) * dJ;
        H = mscale(M3<R>::identity(), V0 * stiff);
        return;
    }
    R logJ = std::log(J);
    R scal = (lam * logJ - mu) / J;
    M3<R> cof = cofactor(F);
    M3<R> P;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P.m[i][j] = mu * F.m[i][j] + scal * cof.m[i][j];
    g = mv(P, gradN) * V0;
    M3<R> U, V;
    V3<R> S;
    svd3(F, U, S, V);
    // ComputeEnergyHessian (math.cpp:86-134)
    R ev[9];
    const R s0 = S[0], s1 = S[1], s2 = S[2];
    R dPhi = (lam * logJ - mu) / J;
    ev[3] = mu + dPhi * s2; ev[4] = mu + dPhi * s1; ev[5] = mu + dPhi * s0;
    ev[6] = mu - dPhi * s2; ev[7] = mu - dPhi * s1; ev[8] = mu - dPhi * s0;
    R diagExtra = mu + lam * (R(1) - logJ);
    M3<R> a;
    a.m[0][0] = mu + diagExtra / (s0 * s0);
    a.m[1][1] = mu + diagExtra / (s1 * s1);
    a.m[2][2] = mu + diagExtra / (s2 * s2);
    a.m[0][1] = a.m[1][0] = lam * s2 / J;
    a.m[0][2] = a.m[2][0] = lam * s1 / J;
    a.m[1][2] = a.m[2][1] = lam * s0 / J;
    V3<R> aw;
    M3<R> av;
    sym_eig3(a, aw, av);
    ev[0] = aw[0]; ev[1] = aw[1]; ev[2] = aw[2];
    project_eigenvalues(ev, 9, projMode, rho);
    // ReconstructVertexHessian (math.cpp:174-230)
    V3<R> u[3] = {U.col(0), U.col(1), U.col(2)};
    R vg[3] = {dot(V.col(0), gradN), dot(V.col(1), gradN), dot(V.col(2), gradN)};
    M3<R> Hv;
    auto add_outer = [&](R l, const V3<R>& w) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Hv.m[i][j] = Hv.m[i][j] + l * (w[i] * w[j]);
    };
    for (int k = 0; k < 3; ++k) {
        V3<R> c = av.col(k);
        V3<R> w = ((c[0] * vg[0]) * u[0] + (c[1] * vg[1]) * u[1]) + (c[2] * vg[2]) * u[2];
        add_outer(ev[k], w);
    }
    static const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    const R invSqrt2 = R(0.70710678118f);
    for (int p = 0; p < 3; ++p) {
        int i = pairs[p][0], j = pairs[p][1];
        add_outer(ev[3 + p], invSqrt2 * (vg[j] * u[i] - vg[i] * u[j]));
    }
    for (int p = 0; p < 3; ++p) {
        int i = pairs[p][0], j = pairs[p][1];
        add_outer(ev[6 + p], invSqrt2 * (vg[j] * u[i] + vg[i] * u[j]));
    }
    H = mscale(Hv, V0);
}

// 3x2 SVD for the membrane: eigen-decomposition of F^T F (2x2), U columns
// F v_i / s_i, third column u0 x u1; V's sign fix as in math.cpp:246-250.
template <class R>
void svd32(const V3<R>& f0, const V3<R>& f1, V3<R> U[3], R S[2], R V[2][2]) {
    R g00 = dot(f0, f0), g01 = dot(f0, f1), g11 = dot(f1, f1);
    R w[2], E[2][2];
    sym_eig2(g00, g01, g11, w, E);
    // descending
    R s0 = std::sqrt(std::max(w[1], R(0))), s1 = std::sqrt(std::max(w[0], R(0)));
    V[0][0] = E[0][1]; V[1][0] = E[1][1];
    V[0][1] = E[0][0]; V[1][1] = E[1][0];
    S[0] = s0; S[1] = s1;
    V3<R> a0 = f0 * V[0][0] + f1 * V[1][0];
    V3<R> a1 = f0 * V[0][1] + f1 * V[1][1];
    const R tiny = std::numeric_limits<R>::min() * R(1e6);
    U[0] = s0 > tiny ? a0 / s0 : V3<R>(R(1), R(0), R(0));
    if (s1 > tiny) U[1] = a1 / s1;
    else {
        V3<R> e = std::abs(U[0].x) < R(0.9) ? V3<R>(R(1), R(0), R(0)) : V3<R>(R(0), R(1), R(0));
        V3<R> c = cross(U[0], e);
        U[1] = c / norm(c);
    }
    U[2] = cross(U[0], U[1]);
    R detV = V[0][0] * V[1][1] - V[0][1] * V[1][0];
    if (detV < R(0)) { V[0][1] = -V[0][1]; V[1][1] = -V[1][1]; S[1] = -S[1]; }
}

// Per-vertex StVK membrane terms of one triangle (stvkfem.cpp:38-86, STVKMath
// math.cpp:237-356).
template <class R>
void stvk_vertex_terms(const V3<R> p[3], const R DmInv[2][2], R area, R mu, R lam, int vtx, int projMode, R rho, V3<R>& g,
                       M3<R>& H) {
    V3<R> e0 = p[1] - p[0], e1 = p[2] - p[0];
    // F = Ds * DmInv (3x2); columns f0, f1
    V3<R> f0 = e0 * DmInv[0][0] + e1 * DmInv[1][0];
    V3<R> f1 = e0 * DmInv[0][1] + e1 * DmInv[1][1];
    R gN[3][2];
    gN[1][0] = DmInv[0][0]; gN[1][1] = DmInv[0][1];
    gN[2][0] = DmInv[1][0]; gN[2][1] = DmInv[1][1];
    gN[0][0] = -gN[1][0] - gN[2][0];
    gN[0][1] = -gN[1][1] - gN[2][1];
    R gn0 = gN[vtx][0], gn1 = gN[vtx][1];
    if (norm(cross(e0, e1)) < (R(2) * area) * R(0.01f)) {  // HandleInvertedElement (stvkfem.cpp:76-86)
        R stiff = mu + lam;
        V3<R> Pg = (f0 * stiff) * gn0 + (f1 * stiff) * gn1;
        g = area * Pg;
        H = mscale(M3<R>::identity(), area * stiff);
        return;
    }
    // P = F (2 mu L + lam tr(L) I), L = (F^T F - I)/2
    R L00 = R(0.5) * (dot(f0, f0) - R(1)), L01 = R(0.5) * dot(f0, f1), L11 = R(0.5) * (dot(f1, f1) - R(1));
    R tr = L00 + L11;
    R S00 = R(2) * mu * L00 + lam * tr, S01 = R(2) * mu * L01, S11 = R(2) * mu * L11 + lam * tr;
    V3<R> P0 = f0 * S00 + f1 * S01;
    V3<R> P1 = f0 * S01 + f1 * S11;
    g = area * (P0 * gn0 + P1 * gn1);
    V3<R> U[3];
    R S[2], V[2][2];
    svd32(f0, f1, U, S, V);
    const R s0 = S[0], s1 = S[1];
    const R trS = s0 * s0 + s1 * s1;
    const R d0 = s0 * (mu * (s0 * s0 - R(1)) + R(0.5) * lam * (trS - R(2)));
    const R d1 = s1 * (mu * (s1 * s1 - R(1)) + R(0.5) * lam * (trS - R(2)));
    R ev[4];
    ev[2] = (s0 + s1 > R(1e-8f)) ? (d0 + d1) / (s0 + s1) : -(mu + lam);
    ev[3] = (std::abs(s0 - s1) > R(1e-8f)) ? (d0 - d1) / (s0 - s1) : (R(3) * mu + R(2) * lam) * s0 * s0 - mu - lam;
    R A00 = R(3) * mu * s0 * s0 + R(0.5) * lam * (R(3) * s0 * s0 + s1 * s1) - mu - lam;
    R A11 = R(3) * mu * s1 * s1 + R(0.5) * lam * (R(3) * s1 * s1 + s0 * s0) - mu - lam;
    R A01 = lam * s0 * s1;
    R w[2], E[2][2];
    sym_eig2(A00, A01, A11, w, E);
    ev[0] = w[0]; ev[1] = w[1];
    project_eigenvalues(ev, 4, projMode, rho);
    R vg[2] = {V[0][0] * gn0 + V[1][0] * gn1, V[0][1] * gn0 + V[1][1] * gn1};
    M3<R> Hv;
    auto add_outer = [&](R l, const V3<R>& w3) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Hv.m[i][j] = Hv.m[i][j] + l * (w3[i] * w3[j]);
    };
    for (int k = 0; k < 2; ++k) {
        V3<R> w3 = (E[0][k] * vg[0]) * U[0] + (E[1][k] * vg[1]) * U[1];
        add_outer(ev[k], w3);
    }
    const R invSqrt2 = R(0.70710678118f);
    add_outer(ev[2], invSqrt2 * (vg[1] * U[0] - vg[0] * U[1]));
    add_outer(ev[3], invSqrt2 * (vg[1] * U[0] + vg[0] * U[1]));
    H = mscale(Hv, area);
}

// Energy change of a Neo-Hookean tet when its vertex vtx moves from p[vtx]
// to p[vtx] - dx, the others fixed (the elastic part of the trust-region
// ratio of AVBD_SCENE_TRUST_REGION; the reference has no energy values).
// dF = -dx gN^T is rank one, so J' = J - dx . (cof F gN) exactly and
// log J' - log J = log1p(dJ / J): the change is formed term by term,
// V0 [mu (F:dF + |dF|^2 / 2) - mu dlog + lam dlog (log J + dlog / 2)], not as
// the difference of two large energies.  The energy is the one whose
// gradient neohookean_vertex_terms uses: mu/2 (tr F^T F - 3) - mu log J +
// lam/2 log^2 J, and for J <= 1e-6 the penalty (mu + lam)(1e-6 - J)^2 / 2
// (neohookeanfem.cpp:86-110); a move across that branch takes the two
// energies' difference.
template <class R>
R neohookean_energy_change(const V3<R> p[4], const M3<R>& DmInv, R V0, R mu, R lam, int vtx, const V3<R>& dx) {
    M3<R> Ds;
    Ds.set_col(0, p[1] - p[0]);
    Ds.set_col(1, p[2] - p[0]);
    Ds.set_col(2, p[3] - p[0]);
    const M3<R> F = mm(Ds, DmInv);
    V3<R> gN[4];
    gN[1] = DmInv.row(0);
    gN[2] = DmInv.row(1);
    gN[3] = DmInv.row(2);
    gN[0] = (-gN[1] - gN[2]) - gN[3];
    const V3<R> g = gN[vtx];
    const R J = det3(F);
    const V3<R> cg = mv(cofactor(F), g);
    const R dJ = -dot(dx, cg);
    const R Jn = J + dJ;
    const R FdF = -dot(dx, mv(F, g));
    const R dF2 = dot(dx, dx) * dot(g, g);
    const R floorJ = R(1e-6f);
    if (J > floorJ && Jn > floorJ) {
        const R logJ = std::log(J);
        const R dl = std::log1p(dJ / J);
        return V0 * ((mu * (FdF + R(0.5) * dF2) - mu * dl) + (lam * dl) * (logJ + R(0.5) * dl));
    }
    R trC = R(0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) trC = trC + F.m[i][j] * F.m[i][j];
    auto psi = [&](R tc, R Jv) {
        if (Jv <= floorJ) return R(0.5) * (mu + lam) * ((floorJ - Jv) * (floorJ - Jv));
        const R l = std::log(Jv);
        return (R(0.5) * mu * (tc - R(3)) - mu * l) + R(0.5) * lam * (l * l);
    };
    return V0 * (psi((trC + R(2) * FdF) + dF2, Jn) - psi(trC, J));
}

// The same for a StVK membrane triangle (energy mu E:E + lam/2 tr(E)^2,
// E = (F^T F - I) / 2, the energy of the gradient stvk_vertex_terms uses, and
// (mu + lam) |F|^2 / 2 when the area collapses, stvkfem.cpp:76-86):
// dE = (F^T dF + dF^T F + dF^T dF) / 2 with dF = -dx gn^T.
template <class R>
R stvk_energy_change(const V3<R> p[3], const R DmInv[2][2], R area, R mu, R lam, int vtx, const V3<R>& dx) {
    const V3<R> e0 = p[1] - p[0], e1 = p[2] - p[0];
    const V3<R> f0 = e0 * DmInv[0][0] + e1 * DmInv[1][0];
    const V3<R> f1 = e0 * DmInv[0][1] + e1 * DmInv[1][1];
    R gN[3][2];
    gN[1][0] = DmInv[0][0]; gN[1][1] = DmInv[0][1];
    gN[2][0] = DmInv[1][0]; gN[2][1] = DmInv[1][1];
    gN[0][0] = -gN[1][0] - gN[2][0];
    gN[0][1] = -gN[1][1] - gN[2][1];
    const R gn0 = gN[vtx][0], gn1 = gN[vtx][1];
    // the moved vertex's edges
    V3<R> q[3] = {p[0], p[1], p[2]};
    q[vtx] = q[vtx] - dx;
    const V3<R> n0 = q[1] - q[0], n1 = q[2] - q[0];
    const R lim = (R(2) * area) * R(0.01f);
    const bool degOld = norm(cross(e0, e1)) < lim, degNew = norm(cross(n0, n1)) < lim;
    const R a0 = dot(f0, dx), a1 = dot(f1, dx), d2 = dot(dx, dx);
    const R E00 = R(0.5) * (dot(f0, f0) - R(1)), E01 = R(0.5) * dot(f0, f1), E11 = R(0.5) * (dot(f1, f1) - R(1));
    const R dE00 = R(0.5) * (d2 * (gn0 * gn0)) - a0 * gn0;
    const R dE01 = R(0.5) * ((d2 * (gn0 * gn1) - a0 * gn1) - a1 * gn0);
    const R dE11 = R(0.5) * (d2 * (gn1 * gn1)) - a1 * gn1;
    if (!degOld && !degNew) {
        const R tr = E00 + E11, dtr = dE00 + dE11;
        const R dEE = R(2) * ((E00 * dE00 + R(2) * (E01 * dE01)) + E11 * dE11) + ((dE00 * dE00 + R(2) * (dE01 * dE01)) + dE11 * dE11);
        return area * (mu * dEE + R(0.5) * lam * (R(2) * tr * dtr + dtr * dtr));
    }
    auto psi = [&](bool deg, R a, R b, R c) {  // E00, E01, E11
        if (deg) return R(0.5) * (mu + lam) * ((R(2) * a + R(1)) + (R(2) * c + R(1)));  // |F|^2 = 2 (E00 + E11) + 2
        const R t = a + c;
        return mu * ((a * a + R(2) * (b * b)) + c * c) + R(0.5) * lam * (t * t);
    };
    return area * (psi(degNew, E00 + dE00, E01 + dE01, E11 + dE11) - psi(degOld, E00, E01, E11));
}

template <class R>
struct NeoHookeanTet : Energy<R> {
    Body<R>* v[4];
    M3<R> DmInv;
    R V0, mu, lam;
    NeoHookeanTet(Body<R>* a, Body<R>* b, Body<R>* c, Body<R>* d) {
        v[0] = a; v[1] = b; v[2] = c; v[3] = d;
        this->linked = {a, b, c, d};
        this->attach();
    }
    void compute_terms(Body<R>* b, int projMode, R rho) override {
        V3<R> p[4] = {v[0]->pos, v[1]->pos, v[2]->pos, v[3]->pos};
        int k = -1;
        for (int i = 0; i < 4; ++i)
            if (v[i] == b) { k = i; break; }
        this->grad = V6<R>();
        this->hess = M6<R>();
        if (k < 0) return;
        V3<R> g;
        M3<R> H;
        neohookean_vertex_terms(p, DmInv, V0, mu, lam, k, projMode, rho, g, H);
        this->grad.set(g, V3<R>());
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) this->hess.m[i][j] = H.m[i][j];
    }
    R energy_change(Body<R>* b, const V3<R>& dx) const override {
        V3<R> p[4] = {v[0]->pos, v[1]->pos, v[2]->pos, v[3]->pos};
        for (int i = 0; i < 4; ++i)
            if (v[i] == b) return neohookean_energy_change(p, DmInv, V0, mu, lam, i, dx);
        return R(0);
    }
};

template <class R>
struct StvkTri : Energy<R> {
    Body<R>* v[3];
    R DmInv[2][2];
    R area, mu, lam;
    StvkTri(Body<R>* a, Body<R>* b, Body<R>* c) {
        v[0] = a; v[1] = b; v[2] = c;
        this->linked = {a, b, c};
        this->attach();
    }
    R energy_change(Body<R>* b, const V3<R>& dx) const override {
        V3<R> p[3] = {v[0]->pos, v[1]->pos, v[2]->pos};
        int k = (b == v[0]) ? 0 : ((b == v[1]) ? 1 : 2);
        return stvk_energy_change(p, DmInv, area, mu, lam, k, dx);
    }
    void compute_terms(Body<R>* b, int projMode, R rho) override {
        V3<R> p[3] = {v[0]->pos, v[1]->pos, v[2]->pos};
        int k = (b == v[0]) ? 0 : ((b == v[1]) ? 1 : 2);  // stvkfem.cpp:50-53
        V3<R> g;
        M3<R> H;
        stvk_vertex_terms(p, DmInv, area, mu, lam, k, projMode, rho, g, H);
        this->grad = V6<R>();
        this->hess = M6<R>();
        this->grad.set(g, V3<R>());
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) this->hess.m[i][j] = H.m[i][j];
    }
};

// ---------------------------------------------------------------- solver
struct Timings {
    double broad = 0, warm = 0, predict = 0, primalDual = 0, velocity = 0, postStab = 0, total = 0;
};

template <class R>
struct World {
    std::vector<std::unique_ptr<Body<R>>> bodies;
    std::vector<std::unique_ptr<Force<R>>> forces;
    std::vector<std::unique_ptr<Energy<R>>> energies;
    // tunables (solver.hpp:69-84) + per-scene extensions
    bool postStabilization = true;
    int iterations = 10;
    R alpha = R(0.95f), beta = R(100000.0f), gamma = R(0.99f), onPenetrationPenalty = R(0);
    R dt = R(1.0f / 60.0f);
    int projectionMode = AVBD_PROJ_ABSOLUTE;
    V3<R> gravity{R(0), R(-9.81f), R(0)};
    V3<R> boundsMin{R(-50), R(-50), R(-50)}, boundsMax{R(50), R(50), R(50)};
    bool emergencyStop = false;
    int nRemoved = 0;
    R trustRegionRho = R(0);  // solver.hpp:80 (updated only under AVBD_SCENE_TRUST_REGION)
    uint32_t options = 0;     // AVBD_SCENE_*
    int64_t nextSerial = 0;
    VisitOrder order = VisitOrder::Natural;
    std::vector<int> colour;  // per body, -1 static (last computed)
    int nColours = 0;
    Timings timings;

    // The renderer's line and point data of the last step (Solver::lineData /
    // debugPointData, solver.hpp:36-37), GPULineData / GPUDebugPointData
    // layouts (geometry.hpp:48-60): 10 and 8 floats, FP32 as the reference.
    std::vector<float> lineData, pointData;
    void add_line(const V3<R>& a, const V3<R>& b, float r, float g, float bl, float al) {
        const float v[10] = {float(a[0]), float(a[1]), float(a[2]), float(b[0]), float(b[1]), float(b[2]), r, g, bl, al};
        lineData.insert(lineData.end(), v, v + 10);
    }
    void add_point(const V3<R>& p, float r, float g, float bl) {
        const float v[8] = {float(p[0]), float(p[1]), float(p[2]), 0.f, r, g, bl, 1.f};
        pointData.insert(pointData.end(), v, v + 8);
    }
    // Force::AddLineData / AddDebugPointData after a successful Initialize
    // (solver.cpp:276-278): manifold.cpp:259-302, joint.cpp:176-190,
    // spring.cpp:112-125.
    void force_debug(Force<R>* f) {
        if (f->kind == ForceKind::Manifold) {
            auto* m = static_cast<Manifold<R>*>(f);
            for (int i = 0; i < m->n; ++i) {
                const Contact<R>& c = m->contacts[i];
                if (c.penetration <= R(0)) continue;
                add_line(c.position, c.position + c.normal * c.penetration, 1.f, 0.f, 0.f, 1.f);
            }
            const M3<R> RA = rotmat(m->A->rot), RB = rotmat(m->B->rot);
            for (int i = 0; i < m->n; ++i) {
                add_point(m->A->pos + mv(RA, m->contacts[i].rA), 1.f, 1.f, 0.f);
                add_point(m->B->pos + mv(RB, m->contacts[i].rB), 0.f, 1.f, 1.f);
            }
        } else if (f->kind == ForceKind::Joint) {
            auto* j = static_cast<Joint<R>*>(f);
            add_line(j->A ? j->A->transform_point(j->rA) : j->rA, j->B->transform_point(j->rB), 0.75f, 0.f, 0.f, 1.f);
        } else {
            auto* sp = static_cast<Spring<R>*>(f);
            add_line(sp->A ? sp->A->transform_point(sp->rA) : sp->rA, sp->B->transform_point(sp->rB), 0.f, 1.f, 0.f,
                     1.f);
        }
    }
    // Energy::AddLineData (solver.cpp:312): the element's edges, coloured
    // with the mean of the bodies' colours (white here; alpha 1, which the
    // reference leaves unset); neohookeanfem.cpp:113-133, stvkfem.cpp:89-105.
    void energy_debug(Energy<R>* e) {
        if (auto* t = dynamic_cast<NeoHookeanTet<R>*>(e)) {
            const int ed[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
            for (auto& p : ed) add_line(t->v[p[0]]->pos, t->v[p[1]]->pos, 1.f, 1.f, 1.f, 1.f);
        } else if (auto* r = dynamic_cast<StvkTri<R>*>(e)) {
            const int ed[3][2] = {{0, 1}, {1, 2}, {2, 0}};
            for (auto& p : ed) add_line(r->v[p[0]]->pos, r->v[p[1]]->pos, 1.f, 1.f, 1.f, 1.f);
        }
    }

    void remove_force(Force<R>* f) {  // Solver::RemoveForce (solver.cpp:143-155)
        int idx = f->index;
        if (idx < 0) return;
        int last = int(forces.size()) - 1;
        if (idx != last) {
            std::swap(forces[idx], forces[last]);
            forces[idx]->index = idx;
        }
        forces.pop_back();
    }
    void remove_energy(Energy<R>* e) {
        int idx = e->index;
        if (idx < 0) return;
        int last = int(energies.size()) - 1;
        if (idx != last) {
            std::swap(energies[idx], energies[last]);
            energies[idx]->index = idx;
        }
        energies.pop_back();
    }
    void remove_body(Body<R>* b) {  // Solver::RemoveBody + ~Mesh (solver.cpp:104-116, geometry.cpp:11-24)
        auto fc = b->forces;
        for (Force<R>* f : fc) remove_force(f);
        auto ec = b->energies;
        for (Energy<R>* e : ec) remove_energy(e);
        int idx = b->index;
        int last = int(bodies.size()) - 1;
        if (idx != last) {
            std::swap(bodies[idx], bodies[last]);
            bodies[idx]->index = idx;
        }
        bodies.pop_back();
    }
    Force<R>* add_force(std::unique_ptr<Force<R>> f) {
        f->index = int(forces.size());
        forces.push_back(std::move(f));
        return forces.back().get();
    }

    // Greedy colouring: dynamic bodies in ascending index each take the
    // smallest colour unused by lower-index dynamic neighbours; neighbours are
    // dynamic bodies sharing any force or energy.
    void compute_colouring() {
        const int N = int(bodies.size());
        colour.assign(N, -1);
        nColours = 0;
        std::vector<uint8_t> used;
        for (int i = 0; i < N; ++i) {
            Body<R>* b = bodies[i].get();
            if (b->isStatic) continue;
            used.assign(used.size(), 0);
            auto mark = [&](Body<R>* o) {
                if (o == b || o->isStatic || o->index >= i) return;
                int c = colour[o->index];
                if (c >= int(used.size())) used.resize(c + 1, 0);
                used[c] = 1;
            };
            for (Force<R>* f : b->forces)
                for (Body<R>* o : f->linked) mark(o);
            for (Energy<R>* e : b->energies)
                for (Body<R>* o : e->linked) mark(o);
            int c = 0;
            while (c < int(used.size()) && used[c]) ++c;
            colour[i] = c;
            nColours = std::max(nColours, c + 1);
        }
    }
    std::vector<Body<R>*> visit_list() {
        std::vector<Body<R>*> out;
        if (order == VisitOrder::Natural) {
            for (auto& b : bodies) out.push_back(b.get());
            return out;
        }
        compute_colouring();
        for (int c = 0; c < nColours; ++c)
            for (auto& b : bodies)
                if (colour[b->index] == c) out.push_back(b.get());
        return out;
    }

    // Solver::CheckExplosion (solver.cpp:182-196)
    bool check_explosion() const {
        const R mx = R(500.0f);
        for (auto& b : bodies) {
            const V3<R>& v = b->vel;
            if (v.x > mx || v.y > mx || v.z > mx || v.x < -mx || v.y < -mx || v.z < -mx) return true;
        }
        return false;
    }

    // One primal block update of one body (solver.cpp:392-448 / 540-595);
    // rhsOut / dxOut receive the block's right-hand side and solution.
    void primal_body(Body<R>* b, R alphaC, V6<R>* rhsOut = nullptr, V6<R>* dxOut = nullptr) {
        const R dt2 = dt * dt;
        M6<R> lhs;
        for (int i = 0; i < 3; ++i) lhs.m[i][i] = b->mass / dt2;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) lhs.m[3 + i][3 + j] = b->Iw.m[i][j] / dt2;
        V6<R> disp = b->displacement_from_inertial();
        V6<R> rhs;
        for (int i = 0; i < 6; ++i) {
            R s = lhs.m[i][0] * disp[0];
            for (int j = 1; j < 6; ++j) s = s + lhs.m[i][j] * disp[j];
            rhs[i] = s;
        }
        for (Force<R>* f : b->forces) {
            f->compute_constraints(alphaC);
            f->compute_derivatives(b);
            const int nr = f->num_rows();
            for (int i = 0; i < nr; ++i) {
                Row<R>& r = f->rows[i];
                R lam = std::isinf(r.stiffness) ? r.lambda : R(0);
                R fv = std::clamp(r.penalty * r.C + lam, r.fmin, r.fmax);
                for (int k = 0; k < 6; ++k) rhs[k] = rhs[k] + r.J[k] * fv;
                for (int k = 0; k < 6; ++k)
                    for (int l = 0; l < 6; ++l) lhs.m[k][l] = lhs.m[k][l] + (r.J[k] * r.J[l]) * r.penalty;
                if (f->includeHessian) {
                    R af = std::abs(fv);
                    for (int j = 0; j < 6; ++j) {
                        R cn = r.H.m[0][j] * r.H.m[0][j];
                        for (int k = 1; k < 6; ++k) cn = cn + r.H.m[k][j] * r.H.m[k][j];
                        lhs.m[j][j] = lhs.m[j][j] + std::sqrt(cn) * af;
                    }
                }
            }
        }
        for (Energy<R>* e : b->energies) {
            e->compute_terms(b, projectionMode, trustRegionRho);
            for (int k = 0; k < 6; ++k) rhs[k] = rhs[k] + e->grad[k];
            for (int k = 0; k < 6; ++k)
                for (int l = 0; l < 6; ++l) lhs.m[k][l] = lhs.m[k][l] + e->hess.m[k][l];
        }
        LDLT6<R> ldlt;
        ldlt.factor(lhs);
        V6<R> dx = ldlt.solve(rhs);
        if (rhsOut) *rhsOut = rhs;
        if (dxOut) *dxOut = dx;
        b->pos = b->pos - V3<R>(dx[0], dx[1], dx[2]);
        Q<R> dq = quaternion_from_difference(V3<R>(dx[3], dx[4], dx[5]), R(-1));
        b->rot = qnormalized(qmul(dq, b->rot));
    }

    // Solver::Step (solver.cpp:199-627)
    void step() {
        if (emergencyStop) return;
        using Clock = std::chrono::steady_clock;
        auto ms = [](Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); };
        auto t_step = Clock::now();
        // 0. world inertia at the start of the step
        for (auto& bp : bodies) {
            Body<R>* b = bp.get();
            if (b->isStatic) continue;
            M3<R> Rm = rotmat(b->rot);
            if (b->mass > R(0)) {
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        b->Iw.m[i][j] = ((Rm.m[i][0] * b->inertia[0]) * Rm.m[j][0] + (Rm.m[i][1] * b->inertia[1]) * Rm.m[j][1]) +
                                        (Rm.m[i][2] * b->inertia[2]) * Rm.m[j][2];
            }
        }
        // 1. broad phase: all pairs i < j
        auto t0 = Clock::now();
        const int N = int(bodies.size());
        for (int i = 0; i < N; ++i) {
            for (int j = i + 1; j < N; ++j) {
                Body<R>* a = bodies[i].get();
                Body<R>* b = bodies[j].get();
                R dist = norm(a->pos - b->pos);
                if (dist < a->radius + b->radius) {
                    if (!is_constrained_to(a, b)) {
                        auto m = std::make_unique<Manifold<R>>(a, b, onPenetrationPenalty);
                        m->serial = nextSerial++;
                        add_force(std::move(m));
                    }
                }
            }
        }
        timings.broad = ms(t0);
        // 2. warm-start forces
        t0 = Clock::now();
        {
            std::vector<Force<R>*> fl;
            for (auto& f : forces) fl.push_back(f.get());
            std::vector<Force<R>*> dead;
            lineData.clear();
            pointData.clear();
            for (Force<R>* f : fl) {
                if (!f->initialize()) { dead.push_back(f); continue; }
                force_debug(f);
                const int nr = f->num_rows();
                for (int i = 0; i < nr; ++i) {
                    Row<R>& r = f->rows[i];
                    if (!postStabilization) r.lambda = r.lambda * (alpha * gamma);
                    R k = r.penalty * gamma;
                    if (k < R(1)) k = R(1);
                    if (k > R(1000000000.0f)) k = R(1000000000.0f);
                    r.penalty = std::min(k, r.stiffness);
                }
            }
            for (Force<R>* f : dead) remove_force(f);
        }
        // 3. energies: Initialize() is always true and the warm-start
        //    ComputeEnergyTerms(nullptr) result is overwritten before use;
        //    each adds its edges to the line data.
        for (auto& e : energies) energy_debug(e.get());
        timings.warm = ms(t0);
        // 4. prediction
        t0 = Clock::now();
        const R maxRot = R(50.0f);
        for (auto& bp : bodies) {
            Body<R>* b = bp.get();
            for (int k = 0; k < 3; ++k) b->angVel[k] = std::max(std::min(b->angVel[k], maxRot), -maxRot);
            V3<R> p = b->pos;
            b->inertialPos = p + b->vel * dt;
            if (!b->isStatic) b->inertialPos = b->inertialPos + gravity * dt * dt;
            Q<R> rq = b->rot;
            b->inertialRot = rq;
            auto integrate = [&](const Q<R>& q0) {
                Q<R> w(R(0), b->angVel.x, b->angVel.y, b->angVel.z);
                Q<R> wq;
                wq.w = w.w * q0.w - dot(w.vec(), q0.vec());
                V3<R> wv = (w.w * q0.vec() + q0.w * w.vec()) + cross(w.vec(), q0.vec());
                Q<R> q = q0;
                q.w = q.w + (R(0.5) * wq.w) * dt;
                q.x = q.x + (R(0.5) * wv.x) * dt;
                q.y = q.y + (R(0.5) * wv.y) * dt;
                q.z = q.z + (R(0.5) * wv.z) * dt;
                return qnormalized(q);
            };
            if (!b->isStatic && !b->isParticle) b->inertialRot = integrate(rq);
            V3<R> acc = (b->vel - b->prevVel) / dt;
            R gmag = norm(gravity);
            R w = R(0);
            if (gmag > R(1e-6f)) {
                V3<R> gdir = gravity / gmag;
                w = dot(acc, gdir) / gmag;
                w = std::clamp(w, R(0), R(1));
            }
            b->lastPos = p;
            b->lastRot = rq;
            b->pos = (p + b->vel * dt) + (gravity * dt * dt) * w;
            if (!b->isStatic && !b->isParticle) b->rot = integrate(rq);
        }
        timings.predict = ms(t0);
        // 5. iterations
        t0 = Clock::now();
        std::vector<Body<R>*> visit = visit_list();
        const bool trustUpdate = projectionMode == AVBD_PROJ_ADAPTIVE && (options & AVBD_SCENE_TRUST_REGION);
        for (int it = 0; it < iterations; ++it) {
            // trust-region ratio of the sweep (step 5.2, solver.cpp:451-455,
            // empty in the reference): over the FEM-only block updates (a
            // body with energies and no forces), the actual decrease of its
            // inertia term plus its elements' energy against the decrease
            // r . dx / 2 its block model predicts, summed in double.  Both
            // sums cancel terms of size `scale` (the summed magnitudes of the
            // per-term changes), so a sweep whose predicted decrease is within
            // kTrustNoise of that scale (a nearly converged solve, where the
            // ratio is rounding noise) leaves rho as it was.
            double actual = 0.0, predicted = 0.0, scale = 0.0;
            for (Body<R>* b : visit) {
                if (b->isStatic) continue;
                if (trustUpdate && b->forces.empty() && !b->energies.empty()) {
                    V6<R> rhs, dx;
                    const V3<R> xOld = b->pos;
                    const V3<R> y = xOld - b->inertialPos;
                    primal_body(b, alpha, &rhs, &dx);
                    const V3<R> xNew = b->pos;
                    const V3<R> d(dx[0], dx[1], dx[2]);
                    b->pos = xOld;  // the energies' change from the old position
                    R dE = (R(0.5) * (b->mass / (dt * dt))) * (dot(d, d) - R(2) * dot(d, y));
                    R dS = std::abs(dE);
                    for (Energy<R>* e : b->energies) {
                        const R de = e->energy_change(b, d);
                        dE = dE + de;
                        dS = dS + std::abs(de);
                    }
                    b->pos = xNew;
                    actual -= double(dE);
                    scale += double(dS);
                    predicted += 0.5 * ((double(rhs[0]) * double(dx[0]) + double(rhs[1]) * double(dx[1])) + double(rhs[2]) * double(dx[2]));
                } else {
                    primal_body(b, alpha);
                }
            }
            if (trustUpdate && predicted > kTrustNoise * scale) trustRegionRho = R(actual / predicted);
            // dual (solver.cpp:457-489)
            std::vector<Force<R>*> fl;
            for (auto& f : forces) fl.push_back(f.get());
            for (Force<R>* f : fl) {
                f->compute_constraints(alpha);
                const int nr = f->num_rows();
                for (int i = 0; i < nr; ++i) {
                    Row<R>& r = f->rows[i];
                    R lam = std::isinf(r.stiffness) ? r.lambda : R(0);
                    r.lambda = lam + r.penalty * r.C;
                    if (r.lambda < r.fmin) r.lambda = r.fmin;
                    else if (r.lambda > r.fmax) r.lambda = r.fmax;
                    if (std::abs(r.lambda) >= r.fracture) {
                        for (Body<R>* lb : f->linked)
                            for (Force<R>* o : lb->forces)
                                if (o->isManifold) o->reset_values();
                        f->disable();
                        break;
                    }
                    if (r.lambda > r.fmin && r.lambda < r.fmax)
                        r.penalty = std::min(r.penalty + beta * std::abs(r.C), std::min(r.stiffness, R(1000000000.0f)));
                }
            }
        }
        timings.primalDual = ms(t0);
        // 6. velocities
        t0 = Clock::now();
        std::vector<Body<R>*> oob;
        for (auto& bp : bodies) {
            Body<R>* b = bp.get();
            if (b->isStatic) continue;
            b->prevVel = b->vel;
            b->prevAngVel = b->angVel;
            V3<R> nv = (b->pos - b->lastPos) / dt;
            if (b->isDragged) {  // solver.cpp:506-510: one-shot added velocity of a dragged body
                nv = nv + b->addedDragVelocity;
                b->addedDragVelocity = V3<R>(R(0), R(0), R(0));
            }
            b->vel = nv;
            if (!b->isParticle) b->angVel = rotation_difference(b->rot, b->lastRot) / dt;
            const V3<R>& p = b->pos;
            if (p.x < boundsMin.x || p.x > boundsMax.x || p.y < boundsMin.y || p.y > boundsMax.y ||
                p.z < boundsMin.z || p.z > boundsMax.z)
                oob.push_back(b);
        }
        emergencyStop = check_explosion();
        if (emergencyStop) return;
        timings.velocity = ms(t0);
        for (Body<R>* b : oob) { remove_body(b); ++nRemoved; }
        // 7. post-stabilisation
        t0 = Clock::now();
        if (postStabilization) {
            std::vector<Body<R>*> vl = oob.empty() ? visit : visit_list();
            for (Body<R>* b : vl) {
                if (b->isStatic) continue;
                primal_body(b, R(0));
            }
        }
        timings.postStab = ms(t0);
        timings.total = ms(t_step);
    }
};

}  // namespace orc
