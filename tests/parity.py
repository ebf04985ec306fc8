"""Comparison helpers: GPU world vs oracle world after the same step(s)."""
from __future__ import annotations

import numpy as np

POS_TOL = 1e-4   # relative, positions and orientations (single step)
LAMBDA_TOL = 1e-3  # relative, lambda (single step)


def pos_rel_err(a, b):
    """max_i |x_a - x_b| / max(|x_b|, 1 m)."""
    d = np.linalg.norm(a["pos"].astype(np.float64) - b["pos"].astype(np.float64), axis=1)
    s = np.maximum(np.linalg.norm(b["pos"].astype(np.float64), axis=1), 1.0)
    return float((d / s).max()) if len(d) else 0.0


def rot_err(a, b):
    """max_i min(|q_a - q_b|, |q_a + q_b|) (unit quaternions)."""
    qa = a["rot"].astype(np.float64)
    qb = b["rot"].astype(np.float64)
    d = np.minimum(np.linalg.norm(qa - qb, axis=1), np.linalg.norm(qa + qb, axis=1))
    return float(d.max()) if len(d) else 0.0


def manifold_table(w):
    out = {}
    for m in w.manifolds:
        out[(int(m["body_a"]), int(m["body_b"]))] = m
    return out


def compare_manifolds(g, o):
    """Returns dict with: same pair set, contact counts/ids equal, lambda rel err."""
    tg, to = manifold_table(g), manifold_table(o)
    res = {"pairs_equal": set(tg) == set(to), "ids_equal": True, "n_equal": True, "lambda_rel": 0.0,
           "n_manifolds": len(to), "missing": sorted(set(to) - set(tg))[:5], "extra": sorted(set(tg) - set(to))[:5]}
    lam_all = [np.abs(m["contacts"]["lambda"][: m["n_contacts"]]).max() for m in o.manifolds if m["n_contacts"]]
    lam_scale = max(lam_all) if lam_all and max(lam_all) > 0 else 1.0  # all zero: static pairs only
    worst = 0.0
    d2 = r2 = 0.0
    for key, mo in to.items():
        mg = tg.get(key)
        if mg is None:
            continue
        n = int(mo["n_contacts"])
        if int(mg["n_contacts"]) != n:
            res["n_equal"] = False
            continue
        if not np.array_equal(mg["contacts"]["id"][:n], mo["contacts"]["id"][:n]):
            res["ids_equal"] = False
        lg = mg["contacts"]["lambda"][:n].astype(np.float64)
        lo = mo["contacts"]["lambda"][:n].astype(np.float64)
        rel = np.abs(lg - lo) / (np.abs(lo) + 1e-3 * lam_scale)
        worst = max(worst, float(rel.max()) if rel.size else 0.0)
        d2 += float(((lg - lo) ** 2).sum())
        r2 += float((lo ** 2).sum())
    # per entry: |dl| / (|l| + 1e-3 max|l|) -- a diagnostic, dominated by
    # small friction rows whose FP32 rounding sensitivity can exceed 1e-3
    res["lambda_rel"] = worst
    # norm-wise over every contact row of the step: ||dl|| / ||l||
    res["lambda_norm_rel"] = float(np.sqrt(d2 / r2)) if r2 > 0 else float(np.sqrt(d2))
    return res


def compare_rows(g_rows, o_rows):
    """(per entry max |dl| / (|l| + 1e-3 max|l|), norm-wise ||dl|| / ||l||)."""
    g = np.asarray(g_rows, np.float64)
    o = np.asarray(o_rows, np.float64)
    if g.size == 0:
        return 0.0, 0.0
    scale = np.abs(o).max() if o.size else 1.0
    per = float((np.abs(g - o) / (np.abs(o) + 1e-3 * max(scale, 1e-12))).max())
    r = float(np.sqrt((o ** 2).sum()))
    d = float(np.sqrt(((g - o) ** 2).sum()))
    return per, (d / r if r > 0 else d)


def summary(g, o):
    s = {"pos_rel": pos_rel_err(g.bodies, o.bodies), "rot": rot_err(g.bodies, o.bodies)}
    s.update(compare_manifolds(g, o))
    if len(o.joints):
        s["joint_lambda_rel"], s["joint_lambda_norm_rel"] = compare_rows(g.joints["lambda"], o.joints["lambda"])
    if len(o.springs):
        s["spring_lambda_rel"], s["spring_lambda_norm_rel"] = compare_rows(g.springs["lambda"], o.springs["lambda"])
    return s


def colouring_ok(world, colours):
    """Independent check of the greedy spec: proper, and every dynamic body
    has the smallest colour unused by its lower-index dynamic neighbours."""
    dyn = (world.bodies["flags"] & 1) == 0
    nb = [set() for _ in range(len(world.bodies))]

    def link(ids):
        ids = [int(i) for i in ids if i >= 0 and dyn[i]]
        for x in ids:
            for y in ids:
                if x != y:
                    nb[x].add(y)

    for m in world.manifolds:
        link([m["body_a"], m["body_b"]])
    for j in world.joints:
        link([j["body_a"], j["body_b"]])
    for s in world.springs:
        link([s["body_a"], s["body_b"]])
    for t in world.tets:
        link(t["v"])
    for t in world.tris:
        link(t["v"])
    for v in range(len(world.bodies)):
        if not dyn[v]:
            continue
        lower = {colours[u] for u in nb[v] if u < v}
        mex = 0
        while mex in lower:
            mex += 1
        if colours[v] != mex:
            return False
        if any(colours[u] == colours[v] for u in nb[v]):
            return False
    return True


def one_step_report(state, g, gcol, natural=True):
    """The GPU's result g (and its colours gcol) of one step from `state`
    against the FP32 oracle in colour-major order (what the GPU computes),
    with the FP64 oracle's distance beside it and, as a report only, the
    distance from the FP32 oracle in natural order (the reference's own
    bodyPtrs visit order, solver.cpp:392-449)."""
    import oracle_py as O

    o = O.Oracle(state, fp64=False, order=O.COLOUR_MAJOR)
    o.step(1)
    ref = o.save()
    ocol, _ = o.colours(len(ref.bodies))
    del o
    d = O.Oracle(state, fp64=True, order=O.COLOUR_MAJOR)
    d.step(1)
    ref64 = d.save()
    del d
    res = summary(g, ref)
    dyn = (g.bodies["flags"] & 1) == 0
    res["colours_equal"] = bool(len(gcol) == len(ocol) and np.array_equal(gcol[dyn], ocol[dyn]))
    g64 = summary(g, ref64)
    res["fp64_pos_rel_gpu"] = g64["pos_rel"]
    res["fp64_rot_gpu"] = g64["rot"]
    spread = summary(ref, ref64)  # the FP32 oracle against the FP64 oracle
    res["fp64_pos_rel_oracle32"] = spread["pos_rel"]
    res["fp64_rot_oracle32"] = spread["rot"]
    # the FP32 oracle's own distance from FP64, per lambda figure
    for k in ("lambda_rel", "lambda_norm_rel", "joint_lambda_rel", "joint_lambda_norm_rel", "spring_lambda_rel",
              "spring_lambda_norm_rel"):
        if k in spread:
            res["spread_" + k] = spread[k]
    res["oracle32_vs_64_lambda_rel"] = spread["lambda_rel"]
    if natural:
        n = O.Oracle(state, fp64=False, order=O.NATURAL)
        n.step(1)
        nat = n.save()
        del n
        if len(nat.bodies) == len(g.bodies):
            sn = summary(g, nat)
            res["natural_pos_rel"] = sn["pos_rel"]
            res["natural_rot"] = sn["rot"]
            res["natural_lambda_norm_rel"] = sn["lambda_norm_rel"]
    return res


def assert_parity(res, strict_lambda=False):
    """The single-step bars.  Positions and orientations within POS_TOL;
    pairs, contact counts, feature ids and colours bit-exact.  Each lambda
    figure (contacts, joints, springs; per entry |dl| / (|l| + 1e-3 max|l|)
    and norm-wise ||dl|| / ||l||) against the FP32 oracle within
    max(LAMBDA_TOL, k x the FP32 oracle's own distance from FP64 in the same
    figure): k = 1 with strict_lambda (the GPU accumulating in the oracle's
    sequential order, so only the rounding of transcendentals differs), else
    k = 2.  Where FP32 itself is well conditioned (spread below LAMBDA_TOL)
    that is LAMBDA_TOL outright; where it is not (friction rows of settling
    piles, the statically indeterminate contacts of a resting box, the
    near-zero forces of a slack chain), the GPU may not move further from
    the FP32 oracle than FP32 moves from FP64."""
    assert res["pairs_equal"], res
    assert res["n_equal"] and res["ids_equal"], res
    assert res["colours_equal"], res
    assert res["pos_rel"] <= POS_TOL, res
    assert res["rot"] <= POS_TOL, res
    k = 1.0 if strict_lambda else 2.0
    for key in ("lambda_rel", "lambda_norm_rel", "joint_lambda_rel", "joint_lambda_norm_rel", "spring_lambda_rel",
                "spring_lambda_norm_rel"):
        if key in res:
            assert res[key] <= max(LAMBDA_TOL, k * res.get("spread_" + key, 0.0)), (key, res)
