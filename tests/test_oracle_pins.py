"""Pins of the CPU oracle against facts that do not come from the oracle.

The reference ships no tests or golden data, so each oracle function is
checked against closed forms, independent FP64 computations (numpy/scipy)
written from the mathematics rather than from the oracle's formulas, brute
force, and physical invariants of the step.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_py as O
from parity import colouring_ok
from skyephyx_amd import scenes
from skyephyx_amd.world import MODEL_CUBE, WorldBuilder

RNG = np.random.default_rng(1234)


def box(pos, rot=(1, 0, 0, 0), scale=(1, 1, 1)):
    return np.concatenate([np.asarray(pos, np.float32), np.asarray(rot, np.float32), np.asarray(scale, np.float32)])


def quat_wxyz(r: Rotation):
    x, y, z, w = r.as_quat()
    return np.array([w, x, y, z])


# ----------------------------------------------------------------- rotations
def test_rotation_matrix_matches_scipy():
    for _ in range(50):
        r = Rotation.random(random_state=RNG.integers(1 << 30))
        q = quat_wxyz(r).astype(np.float32)
        R = O.rot_helper(3, q).reshape(3, 3)
        np.testing.assert_allclose(R, r.as_matrix(), atol=2e-6)


def test_quaternion_from_difference_is_exp_map():
    # QFD(v, s) = rotation by s*v (math.hpp:31-48); compare with scipy from_rotvec
    for _ in range(50):
        v = RNG.normal(size=3) * RNG.choice([1e-8, 1e-3, 0.3, 2.0])
        for s in (1.0, -1.0):
            q = O.rot_helper(1, np.concatenate([v, [s]]))
            ref = quat_wxyz(Rotation.from_rotvec(s * v))
            # same rotation up to the quaternion's sign (no hemisphere fix in math.hpp:31-48)
            if np.dot(q, ref) < 0:
                ref = -ref
            np.testing.assert_allclose(q, ref, atol=2e-6)
            if np.linalg.norm(v) < np.pi:
                assert q[0] > 0


def test_rotation_difference_inverts_exp_map():
    # RotDiff(q, p) = rotation vector of q p^* (w >= 0 hemisphere)
    for _ in range(50):
        p = Rotation.random(random_state=RNG.integers(1 << 30))
        v = RNG.normal(size=3)
        v *= RNG.uniform(0.0, 3.0) / np.linalg.norm(v)  # angle below pi: RotDiff returns the w >= 0 branch
        q = Rotation.from_rotvec(v) * p
        d = O.rot_helper(0, quat_wxyz(q).astype(np.float32), quat_wxyz(p).astype(np.float32))
        np.testing.assert_allclose(d, v, atol=5e-6 * max(1.0, np.linalg.norm(v)))
    # small-angle branch: 2 * vec part exactly
    q = np.array([1.0, 1e-8, -2e-8, 3e-8], np.float32)
    d = O.rot_helper(0, q, np.array([1, 0, 0, 0], np.float32))
    np.testing.assert_array_equal(d, 2 * q[1:])


def test_quaternion_to_rotvec():
    for _ in range(30):
        r = Rotation.random(random_state=RNG.integers(1 << 30))
        d = O.rot_helper(2, quat_wxyz(r).astype(np.float32))
        np.testing.assert_allclose(d, r.as_rotvec(), atol=5e-6)
    assert np.all(O.rot_helper(2, np.array([1, 1e-7, 0, 0], np.float32)) == 0)


# ----------------------------------------------------------------- LDLT
def test_ldlt_spd_matches_numpy():
    for _ in range(30):
        A = RNG.normal(size=(6, 6))
        A = A @ A.T + 0.1 * np.eye(6)
        b = RNG.normal(size=6)
        np.testing.assert_allclose(O.ldlt_solve(A, b), np.linalg.solve(A, b), rtol=1e-9, atol=1e-9)


def test_ldlt_pseudo_inverse_on_singular_block():
    # a particle: blockdiag(A3, 0) -> rotational part of the solution is zero
    A3 = RNG.normal(size=(3, 3))
    A3 = A3 @ A3.T + np.eye(3)
    A = np.zeros((6, 6))
    A[:3, :3] = A3
    b = np.concatenate([RNG.normal(size=3), [0.0, 0.0, 0.0]])
    x = O.ldlt_solve(A, b)
    np.testing.assert_allclose(x[:3], np.linalg.solve(A3, b[:3]), rtol=1e-12)
    assert np.all(x[3:] == 0)
    x32 = O.ldlt_solve(A, b, fp64=False)
    np.testing.assert_allclose(x32[:3], np.linalg.solve(A3, b[:3]), rtol=1e-5)
    assert np.all(x32[3:] == 0)


# ----------------------------------------------------------------- SVD
def test_svd3_reconstructs_and_orders():
    for _ in range(100):
        F = RNG.normal(size=(3, 3))
        U, S, V = O.svd3(F)
        np.testing.assert_allclose(U @ np.diag(S) @ V.T, F, atol=1e-12)
        np.testing.assert_allclose(U.T @ U, np.eye(3), atol=1e-12)
        np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-12)
        assert np.linalg.det(U) > 0 and np.linalg.det(V) > 0
        np.testing.assert_allclose(np.abs(S), np.linalg.svd(F, compute_uv=False), rtol=1e-12)
        assert S[0] >= S[1] >= abs(S[2])
        assert np.sign(S[2]) == np.sign(np.linalg.det(F))


# ----------------------------------------------------------------- FEM
def nh_energy(F, mu, lam):
    """Neo-Hookean energy density (math.cpp:52-58 states it as
    mu/2 (I1 - 3) - mu ln J + lam/2 (ln J)^2)."""
    J = np.linalg.det(F)
    lj = math.log(J)
    return 0.5 * mu * (np.sum(F * F) - 3.0) - mu * lj + 0.5 * lam * lj * lj


def stvk_energy(F, mu, lam):
    """StVK membrane: L = (F^T F - I)/2, psi = mu |L|^2 + lam/2 tr(L)^2."""
    L = 0.5 * (F.T @ F - np.eye(2))
    return mu * np.sum(L * L) + 0.5 * lam * np.trace(L) ** 2


def numeric_grad_hess(fun, x, h=1e-5):
    n = len(x)
    g = np.zeros(n)
    H = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        g[i] = (fun(x + e) - fun(x - e)) / (2 * h)
        for j in range(n):
            f = np.zeros(n)
            f[j] = h
            H[i, j] = (fun(x + e + f) - fun(x + e - f) - fun(x - e + f) + fun(x - e - f)) / (4 * h * h)
    return g, 0.5 * (H + H.T)


def project_abs(H, eps=1e-6):
    w, Q = np.linalg.eigh(H)
    w = np.maximum(np.abs(w), eps)
    return (Q * w) @ Q.T


def random_tet():
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64) * 0.5
    Dm = np.stack([base[1] - base[0], base[2] - base[0], base[3] - base[0]], axis=1)
    p = base + RNG.normal(scale=0.05, size=base.shape)
    return base, p, np.linalg.inv(Dm), abs(np.linalg.det(Dm)) / 6.0


@pytest.mark.parametrize("vtx", [0, 1, 2, 3])
def test_neohookean_vertex_terms_match_fd_hessian(vtx):
    mu, lam = 38461.5, 57692.3
    for _ in range(6):
        base, p, Dminv, V0 = random_tet()

        def energy(x):
            q = p.copy()
            q[vtx] = x
            Ds = np.stack([q[1] - q[0], q[2] - q[0], q[3] - q[0]], axis=1)
            return V0 * nh_energy(Ds @ Dminv, mu, lam)

        g_ref, _ = numeric_grad_hess(energy, p[vtx].copy())
        # projected Hessian: V0 G^T proj(d2psi/dF2) G with G = dvec(F)/dx_vtx
        gN = np.concatenate([[-Dminv.sum(axis=0)], Dminv])[vtx]
        F0 = np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]], axis=1) @ Dminv

        def psiF(f):
            return nh_energy(f.reshape(3, 3), mu, lam)

        _, HF = numeric_grad_hess(psiF, F0.reshape(-1), h=1e-4)
        G = np.zeros((9, 3))  # vec(F) row-major: F[a, c] depends on x[a] through gN[c]
        for a in range(3):
            for c in range(3):
                G[a * 3 + c, a] = gN[c]
        H_ref = V0 * G.T @ project_abs(HF) @ G
        g, H = O.neohookean_terms(p, Dminv, V0, mu, lam, vtx)
        np.testing.assert_allclose(g, g_ref, rtol=1e-5, atol=1e-6 * np.abs(g_ref).max())
        np.testing.assert_allclose(H, H_ref, rtol=2e-3, atol=2e-3 * np.abs(H_ref).max())


@pytest.mark.parametrize("vtx", [0, 1, 2, 3])
def test_neohookean_positive_definite_vertex_block_closed_form(vtx):
    """When every eigenvalue of the Neo-Hookean Hessian is >= 1e-6 the
    projection is the identity and the vertex block is
    V0 (mu |gN|^2 I + (mu + lam (1 - log J)) / J^2 (cof F gN)(cof F gN)^T)
    (the d2J/dF2 term vanishes on one vertex).  In the common case
    lam log J <= mu that holds iff s_max(F) < (mu - 1e-6) / -dPhi.  The
    oracle's eigen-based projected block must equal the closed form there."""
    mu, lam = 38461.5, 57692.3
    rng = np.random.default_rng(5 + vtx)
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64) * 0.5
    Dm = np.stack([base[1] - base[0], base[2] - base[0], base[3] - base[0]], axis=1)
    Dminv = np.linalg.inv(Dm)
    V0 = abs(np.linalg.det(Dm)) / 6.0
    gN = np.concatenate([[-Dminv.sum(axis=0)], Dminv])[vtx]
    n_pd = n_not = 0
    for k in range(40):
        R, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        # half near-uniform expansions (positive definite), half general
        s = rng.uniform(1.02, 1.25) * rng.uniform(0.97, 1.03, 3) if k % 2 else rng.uniform(0.75, 1.1, 3)
        F = R @ np.diag(s) @ Q.T
        if np.linalg.det(F) < 0:
            F[:, 0] = -F[:, 0]
        p = base @ F.T
        J = np.linalg.det(F)
        dphi = (lam * np.log(J) - mu) / J
        if dphi > 0:  # the closed form is only used for lam log J <= mu
            continue
        T = (mu - 1e-6) / -dphi
        pd = np.linalg.eigvalsh(F.T @ F).max() < T * T
        g, H = O.neohookean_terms(p, Dminv, V0, mu, lam, vtx)
        cof = J * np.linalg.inv(F).T
        cg = cof @ gN
        c1 = (mu + lam * (1.0 - np.log(J))) / (J * J)
        H_closed = V0 * (mu * (gN @ gN) * np.eye(3) + c1 * np.outer(cg, cg))
        # the oracle's Jacobi tolerances are FLT_EPS-based even in FP64
        if pd:
            n_pd += 1
            np.testing.assert_allclose(H, H_closed, rtol=2e-6, atol=2e-6 * np.abs(H_closed).max())
        else:
            n_not += 1
            assert np.abs(H - H_closed).max() > 1e-4 * np.abs(H_closed).max()
    assert n_pd >= 5 and n_not >= 5, (n_pd, n_not)


def test_neohookean_rest_state():
    base, _, Dminv, V0 = random_tet()
    g, H = O.neohookean_terms(base, Dminv, V0, 1.0, 2.0, 1)
    np.testing.assert_allclose(g, 0, atol=1e-12)
    # rest-state modes (Smith et al.): scaling {2mu+3lam, 2mu, 2mu}, twists mu+(-mu)=0->1e-6, flips 2mu
    w = np.linalg.eigvalsh(H / V0)
    assert np.all(w > 0)


def test_neohookean_inverted_branch():
    base, p, Dminv, V0 = random_tet()
    p = base.copy()
    p[3] = p[0] - (p[3] - p[0])  # invert
    g, H = O.neohookean_terms(p, Dminv, V0, 3.0, 5.0, 2)
    np.testing.assert_allclose(H, V0 * 8.0 * np.eye(3), rtol=1e-12)


@pytest.mark.parametrize("vtx", [0, 1, 2])
def test_stvk_vertex_terms_match_in_plane_projection(vtx):
    mu, lam = 3846.15, 5769.23
    for _ in range(6):
        base = np.array([[0, 0, 0], [0.5, 0, 0], [0.1, 0.4, 0]], np.float64)
        e0, e1 = base[1] - base[0], base[2] - base[0]
        x1 = np.linalg.norm(e0)
        x2 = e1 @ e0 / x1
        y2 = math.sqrt(e1 @ e1 - x2 * x2)
        Dminv = np.linalg.inv(np.array([[x1, x2], [0, y2]]))
        area = 0.5 * x1 * y2
        p = base + RNG.normal(scale=0.03, size=base.shape)

        def energy(x):
            q = p.copy()
            q[vtx] = x
            Ds = np.stack([q[1] - q[0], q[2] - q[0]], axis=1)
            return area * stvk_energy(Ds @ Dminv, mu, lam)

        g_ref, _ = numeric_grad_hess(energy, p[vtx].copy())
        F0 = np.stack([p[1] - p[0], p[2] - p[0]], axis=1) @ Dminv
        _, HF = numeric_grad_hess(lambda f: stvk_energy(f.reshape(3, 2), mu, lam), F0.reshape(-1), h=1e-4)
        # the reference keeps the 4 in-plane modes u_i v_j^T (i, j < 2) and drops the normal ones
        U, S, Vt = np.linalg.svd(F0)
        B = np.stack([np.outer(U[:, i], Vt[j]).reshape(-1) for i in range(2) for j in range(2)], axis=1)
        Hin = B.T @ HF @ B
        Hproj = B @ project_abs(Hin) @ B.T
        gN = np.concatenate([[-Dminv.sum(axis=0)], Dminv])[vtx]
        G = np.zeros((6, 3))
        for a in range(3):
            for c in range(2):
                G[a * 2 + c, a] = gN[c]
        H_ref = area * G.T @ Hproj @ G
        g, H = O.stvk_terms(p, Dminv, area, mu, lam, vtx)
        np.testing.assert_allclose(g, g_ref, rtol=1e-5, atol=1e-6 * np.abs(g_ref).max())
        np.testing.assert_allclose(H, H_ref, rtol=2e-3, atol=2e-3 * np.abs(H_ref).max())


# ----------------------------------------------------------------- collision
def corners(b):
    p, q, s = b[:3].astype(np.float64), b[3:7].astype(np.float64), b[7:].astype(np.float64)
    R = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()
    signs = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64)
    return p + (signs * (s / 2)) @ R.T, R


def sat_separated(a, b):
    """Brute-force SAT in FP64: project all 8 corners of both boxes on the 15
    candidate axes; separated iff some axis shows a positive gap."""
    ca, RA = corners(a)
    cb, RB = corners(b)
    axes = [RA[:, i] for i in range(3)] + [RB[:, i] for i in range(3)]
    for i in range(3):
        for j in range(3):
            c = np.cross(RA[:, i], RB[:, j])
            n = np.linalg.norm(c)
            if n > 1e-6:
                axes.append(c / n)
    gap = -np.inf
    for ax in axes:
        pa, pb = ca @ ax, cb @ ax
        gap = max(gap, max(pa.min() - pb.max(), pb.min() - pa.max()))
    return gap


def test_resting_box_known_answer():
    ground = box((0, -10, 0), scale=(20, 1, 20))
    cube = box((0.3, -9.0 - 0.01, 0.2))
    n, c, ids = O.collide(ground, cube)
    assert n == 4
    assert sorted(ids.tolist()) == [4 + 256 * i for i in range(4)]  # FACE_A, axis 1, vertex i
    np.testing.assert_allclose(c[:, 3:6], np.tile([0, 1, 0], (4, 1)))
    np.testing.assert_allclose(c[:, 6], 0.01, atol=1e-6)


def test_yawed_box_on_ground():
    ground = box((0, -10, 0), scale=(20, 1, 20))
    q = quat_wxyz(Rotation.from_euler("y", 45, degrees=True))
    n, c, ids = O.collide(ground, box((0, -9.0 - 0.02, 0), q))
    assert n == 4
    np.testing.assert_allclose(c[:, 6], 0.02, atol=1e-5)


def test_crossed_edges_give_one_edge_contact():
    # two boxes rotated 45 degrees about orthogonal axes, edge to edge
    qa = quat_wxyz(Rotation.from_euler("z", 45, degrees=True))
    qb = quat_wxyz(Rotation.from_euler("x", 45, degrees=True))
    h = math.sqrt(2) / 2
    n, c, ids = O.collide(box((0, 0, 0), qa), box((0, 2 * h - 0.01, 0), qb))
    assert n == 1
    assert ids[0] & 3 == 2  # EDGE
    np.testing.assert_allclose(c[0, 6], 0.01, atol=1e-4)


def test_sat_against_brute_force():
    agree = 0
    for _ in range(400):
        a = box(RNG.normal(scale=0.6, size=3), quat_wxyz(Rotation.random(random_state=RNG.integers(1 << 30))),
                RNG.uniform(0.4, 1.6, 3))
        b = box(RNG.normal(scale=0.6, size=3), quat_wxyz(Rotation.random(random_state=RNG.integers(1 << 30))),
                RNG.uniform(0.4, 1.6, 3))
        gap = sat_separated(a, b)
        n, c, ids = O.collide(a, b, fp64=True)
        if abs(gap) < 1e-3:
            continue  # too close to call in FP64 vs the reference's 1e-6 tolerance on |C|
        if gap > 0:
            assert n == 0
        else:
            assert n >= 1
            d = b[:3] - a[:3]
            for k in range(n):
                assert c[k, 6] >= -1e-6  # penetration
                assert abs(np.linalg.norm(c[k, 3:6]) - 1) < 1e-5
            # face normals point from A to B
            if ids[0] & 3 != 2:
                assert c[0, 3:6] @ d >= -1e-6
        agree += 1
    assert agree > 300


# ----------------------------------------------------------------- colouring
@pytest.mark.parametrize("make", [lambda: scenes.pyramid_2d(), lambda: scenes.fem_block(n=5),
                                  lambda: scenes.box_lattice(shape=(3, 3, 2), spacing=3.0, n_chain=6, n_springs=4,
                                                             bounds=50.0)])
def test_greedy_colouring_spec(make):
    o = O.Oracle(make())
    o.step(30)
    w = o.save()
    col, nc = o.colours(len(w.bodies))
    assert colouring_ok(w, col)


def test_colour_major_equals_natural_when_no_coupling():
    # mass_springs: each body only linked to the world -> any visit order is the same
    w = scenes.mass_springs(4)
    a = O.Oracle(w, order=O.NATURAL)
    b = O.Oracle(w, order=O.COLOUR_MAJOR)
    a.step(5)
    b.step(5)
    np.testing.assert_array_equal(a.save().bodies["pos"], b.save().bodies["pos"])


# ----------------------------------------------------------------- step invariants
def single_box(pos, vel=(0, 0, 0), ground=True, **scene):
    b = WorldBuilder(**scene)
    if ground:
        scenes.add_ground(b)
    b.add_body(MODEL_CUBE, 1.0, 0.5, pos, (1, 1, 1), velocity=vel)
    return b.build()


def test_free_fall_first_step():
    w = single_box((0, 5, 0), ground=False)
    o = O.Oracle(w)
    o.step(1)
    s = o.save()
    dt = np.float32(1 / 60)
    y = np.float32(5) + np.float32(-9.81) * dt * dt
    assert abs(s.bodies["pos"][0][1] - y) <= 2e-6
    # v = (y - x0) / dt: the subtraction at |x0| = 5 leaves ~ulp(5)/dt of absolute error
    np.testing.assert_allclose(s.bodies["vel"][0], [0, -9.81 * (1 / 60), 0], atol=1e-4)


def test_resting_box_normal_force_balances_weight():
    o = O.Oracle(single_box((0.3, -8.99, 0.2), alpha=0.99), fp64=True)
    o.step(200)
    s = o.save()
    m = s.manifolds[0]
    lam = m["contacts"]["lambda"][: m["n_contacts"], 0]
    assert np.all(lam <= 0)
    np.testing.assert_allclose(lam.sum(), -1.0 * 9.81, rtol=1e-4)
    np.testing.assert_allclose(s.bodies["vel"][1], 0, atol=1e-6)


def test_soft_spring_equilibrium_stretch():
    b = WorldBuilder()
    k = 200.0
    i = b.add_body(MODEL_CUBE, 1.0, 0.5, (0, 0, 0), (1, 1, 1))
    b.add_spring(None, (0, 0.5, 0), i, (0, 0.5, 0), k)  # rest length 0
    o = O.Oracle(b.build(), fp64=True)
    o.step(600)
    s = o.save()
    stretch = -float(s.bodies["pos"][0][1])
    np.testing.assert_allclose(stretch, 1.0 * 9.81 / k, rtol=2e-3)


def test_pendulum_joint_residual_small():
    o = O.Oracle(scenes.joint_chain(3), fp64=True)
    o.step(60)
    s = o.save()
    j = s.joints[0]
    anchor = j["r_a"]
    b = s.bodies[j["body_b"]]
    R = Rotation.from_quat([b["rot"][1], b["rot"][2], b["rot"][3], b["rot"][0]]).as_matrix()
    pin = R @ (b["scale"] * j["r_b"]) + b["pos"]
    assert np.linalg.norm(pin - anchor) < 1e-3


def test_fracture_disables_then_removes_spring():
    b = WorldBuilder()
    i = b.add_body(MODEL_CUBE, 50.0, 0.5, (0, 0, 0), (1, 1, 1))
    s = b.add_spring(None, (0, 0.5, 0), i, (0, 0.5, 0), np.float32(np.inf))
    b.spring(s)["fracture"] = 100.0  # weight 490 N > 100 N
    o = O.Oracle(b.build())
    for _ in range(60):  # the penalty ramps up over a few steps until |lambda| >= fracture
        o.step(1)
        w1 = o.save()
        if len(w1.springs) == 0 or w1.springs["flags"][0] & 1:
            break
    assert len(w1.springs) == 1 and w1.springs["flags"][0] & 1
    assert w1.springs["lambda"][0] == 0 and w1.springs["penalty"][0] >= 0
    o.step(1)
    assert len(o.save().springs) == 0


def test_out_of_bounds_body_is_removed():
    w = single_box((0, 0, 0), ground=False)
    w.bodies["pos"][0] = (60, 0, 0)
    o = O.Oracle(w)
    o.step(1)
    s = o.save()
    assert len(s.bodies) == 0 and s.scenes["n_removed"][0] == 1


def test_explosion_stops_the_scene():
    w = single_box((0, 0, 0), vel=(600, 0, 0), ground=False)
    o = O.Oracle(w)
    o.step(1)
    s = o.save()
    assert s.scenes["emergency_stop"][0] == 1
    before = s.bodies["pos"].copy()
    o.step(3)
    np.testing.assert_array_equal(o.save().bodies["pos"], before)


# ----------------------------------------------------------- RemoveBody
def test_oracle_remove_bodies_is_swap_with_last():
    """RemoveBody (solver.cpp:104-116): each listed body (indices before the
    call) is swapped with the current last body and dropped, with every
    force and energy that touches it (~Mesh, geometry.cpp:11-24)."""
    w = scenes.box_lattice(shape=(3, 2, 2), spacing=3.0, n_chain=4, n_springs=3, bounds=50.0)
    o = O.Oracle(w)
    o.step(8)
    before = o.save()
    ids = [3, 0, len(before.bodies) - 2]
    o.remove_bodies(ids)
    after = o.save()
    # emulate the order: a list of original indices
    order = list(range(len(before.bodies)))
    for i in ids:
        k = order.index(i)
        order[k] = order[-1]
        order.pop()
    assert len(after.bodies) == len(order)
    np.testing.assert_array_equal(after.bodies["pos"], before.bodies["pos"][order])
    gone = set(ids)
    keep = lambda a, b: a not in gone and b not in gone  # noqa: E731
    assert len(after.joints) == sum(keep(int(j["body_a"]), int(j["body_b"])) for j in before.joints)
    assert len(after.springs) == sum(keep(int(s["body_a"]), int(s["body_b"])) for s in before.springs)
    assert len(after.manifolds) == sum(keep(int(m["body_a"]), int(m["body_b"])) for m in before.manifolds)
    # surviving forces point at the same bodies (renumbered)
    new_of = {old: new for new, old in enumerate(order)}
    pairs_before = sorted((new_of[int(m["body_a"])], new_of[int(m["body_b"])]) for m in before.manifolds
                          if keep(int(m["body_a"]), int(m["body_b"])))
    pairs_after = sorted((int(m["body_a"]), int(m["body_b"])) for m in after.manifolds)
    assert pairs_before == pairs_after
