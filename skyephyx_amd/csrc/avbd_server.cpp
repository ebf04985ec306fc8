// avbd_server.cpp -- served form of the AVBD step: requests for k steps of a
// registered scene, scenes resident on the device under an HBM byte budget,
// continuous batching (every tick steps all in-flight scenes in one batched
// device step; requests join and leave between ticks).  Host-only C++ over
// the solver's C ABI.
//
// Scene state round trip: a resident scene's state lives in the device world.
// When the resident set changes, the device world is downloaded, split back
// into per-scene host copies and a world of the new set is uploaded.  Within
// a scene the relative order of forces (joints, springs, contact manifolds:
// the order Solver::Step visits them, solver.cpp:269-292) and of energies is
// kept by renumbering their serials by rank, so a scene's steps do not depend
// on which other scenes share the device world.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "avbd_internal.h"

namespace {

struct SceneHost {
    avbd_scene sc{};
    std::vector<avbd_body> bodies;       // scene field 0, local body indices
    std::vector<avbd_joint> joints;      // serials = rank in the scene's force order
    std::vector<avbd_spring> springs;
    std::vector<avbd_manifold> manifolds;
    std::vector<avbd_tet> tets;          // static; serials = rank in energy order
    std::vector<avbd_tri> tris;
    long long nonpart = 0;
    long long last_used = -1;  // tick of the last step (eviction order)
    int slot = -1;             // scene index in the device world, -1 = host only
    bool busy = false;         // a request of this scene is running
    bool want = false;         // scratch: in the planned resident set

    WorldCounts counts() const {
        WorldCounts k;
        k.S = 1;
        k.N = (long long)bodies.size();
        k.nonpart = nonpart;
        k.J = (long long)joints.size();
        k.SP = (long long)springs.size();
        k.M = (long long)manifolds.size();
        k.T = (long long)tets.size();
        k.TR = (long long)tris.size();
        k.tiles = broad_tiles(k.N);
        k.items = broad_items(k.N);
        return k;
    }
};

void add_counts(WorldCounts& a, const WorldCounts& b) {
    a.S += b.S; a.N += b.N; a.nonpart += b.nonpart; a.J += b.J; a.SP += b.SP;
    a.M += b.M; a.T += b.T; a.TR += b.TR; a.tiles += b.tiles; a.items += b.items;
}

// Renumber the force serials of one scene by rank (stable on ties), keeping
// their relative order.
void renumber_forces(SceneHost& h) {
    struct Ref { long long serial; int kind, idx; };
    std::vector<Ref> refs;
    refs.reserve(h.joints.size() + h.springs.size() + h.manifolds.size());
    for (size_t i = 0; i < h.joints.size(); ++i) refs.push_back({h.joints[i].serial, 0, (int)i});
    for (size_t i = 0; i < h.springs.size(); ++i) refs.push_back({h.springs[i].serial, 1, (int)i});
    for (size_t i = 0; i < h.manifolds.size(); ++i) refs.push_back({h.manifolds[i].serial, 2, (int)i});
    std::stable_sort(refs.begin(), refs.end(), [](const Ref& a, const Ref& b) { return a.serial < b.serial; });
    for (size_t r = 0; r < refs.size(); ++r) {
        const Ref& f = refs[r];
        if (f.kind == 0) h.joints[f.idx].serial = (long long)r;
        else if (f.kind == 1) h.springs[f.idx].serial = (long long)r;
        else h.manifolds[f.idx].serial = (long long)r;
    }
}

void renumber_energies(SceneHost& h) {
    struct Ref { int serial, kind, idx; };
    std::vector<Ref> refs;
    refs.reserve(h.tets.size() + h.tris.size());
    for (size_t i = 0; i < h.tets.size(); ++i) refs.push_back({h.tets[i].serial, 0, (int)i});
    for (size_t i = 0; i < h.tris.size(); ++i) refs.push_back({h.tris[i].serial, 1, (int)i});
    std::stable_sort(refs.begin(), refs.end(), [](const Ref& a, const Ref& b) { return a.serial < b.serial; });
    for (size_t r = 0; r < refs.size(); ++r) {
        if (refs[r].kind == 0) h.tets[refs[r].idx].serial = (int)r;
        else h.tris[refs[r].idx].serial = (int)r;
    }
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct avbd_server {
    int device = 0;
    long long budget = 0;
    int max_slots = 0;
    avbd_ctx* ctx = nullptr;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::vector<SceneHost> scenes;
    std::vector<int> resident;  // device scene index -> scene id
    long long resident_bytes = 0;

    struct Req {
        avbd_request_info info{};
        int remaining = 0;
    };
    std::unordered_map<long long, Req> reqs;  // queued, running and unpolled finished requests
    std::deque<long long> queue;              // FIFO of queued tickets
    std::vector<long long> running;
    std::vector<long long> finished;          // in completion order, until polled
    long long next_ticket = 1;
    avbd_server_stats st{};

    long long set_bytes(const std::vector<int>& ids) const {
        WorldCounts k;
        for (int id : ids) add_counts(k, scenes[id].counts());
        return world_device_bytes(k, 0);
    }

    // Pull the state of every resident scene back into its host copy.
    int sync_to_host() {
        if (resident.empty()) return 0;
        avbd_world cnt{};
        int rc = avbd_get_counts(ctx, &cnt);
        if (rc) return rc;
        std::vector<avbd_scene> sc(cnt.n_scenes);
        std::vector<avbd_body> bodies(cnt.n_bodies);
        std::vector<avbd_joint> joints(cnt.n_joints);
        std::vector<avbd_spring> springs(cnt.n_springs);
        std::vector<avbd_manifold> mans(cnt.n_manifolds);
        std::vector<avbd_tet> tets(cnt.n_tets);
        std::vector<avbd_tri> tris(cnt.n_tris);
        avbd_world w = cnt;
        w.scenes = sc.data(); w.bodies = bodies.data(); w.joints = joints.data(); w.springs = springs.data();
        w.manifolds = mans.data(); w.tets = tets.data(); w.tris = tris.data();
        if ((rc = avbd_download(ctx, &w))) return rc;
        if (w.n_scenes != (int)resident.size()) return avbd_set_error(AVBD_E_STATE, "device scene count changed");
        const int S = w.n_scenes;
        std::vector<int> off(S + 1, 0);
        for (int i = 0; i < w.n_bodies; ++i) off[bodies[i].scene + 1]++;
        for (int s = 0; s < S; ++s) off[s + 1] += off[s];
        for (int s = 0; s < S; ++s) {
            // (bodies that left the world bounds were removed on the device)
            SceneHost& h = scenes[resident[s]];
            h.sc = sc[s];
            h.bodies.assign(bodies.begin() + off[s], bodies.begin() + off[s + 1]);
            h.nonpart = 0;
            for (avbd_body& b : h.bodies) {
                b.scene = 0;
                if (!(b.flags & AVBD_BODY_PARTICLE)) ++h.nonpart;
            }
            h.joints.clear();
            h.springs.clear();
            h.manifolds.clear();
            h.tets.clear();
            h.tris.clear();
        }
        auto local = [&](int b, int s) { return b < 0 ? b : b - off[s]; };
        for (int i = 0; i < w.n_joints; ++i) {
            avbd_joint j = joints[i];
            int s = bodies[j.body_b].scene;
            j.body_a = local(j.body_a, s);
            j.body_b = local(j.body_b, s);
            scenes[resident[s]].joints.push_back(j);
        }
        for (int i = 0; i < w.n_springs; ++i) {
            avbd_spring p = springs[i];
            int s = bodies[p.body_b].scene;
            p.body_a = local(p.body_a, s);
            p.body_b = local(p.body_b, s);
            scenes[resident[s]].springs.push_back(p);
        }
        for (int i = 0; i < w.n_manifolds; ++i) {
            avbd_manifold m = mans[i];
            int s = bodies[m.body_a].scene;
            m.body_a = local(m.body_a, s);
            m.body_b = local(m.body_b, s);
            scenes[resident[s]].manifolds.push_back(m);
        }
        for (int i = 0; i < w.n_tets; ++i) {
            avbd_tet t = tets[i];
            const int s = bodies[t.v[0]].scene;
            for (int k = 0; k < 4; ++k) t.v[k] = local(t.v[k], s);
            scenes[resident[s]].tets.push_back(t);
        }
        for (int i = 0; i < w.n_tris; ++i) {
            avbd_tri t = tris[i];
            const int s = bodies[t.v[0]].scene;
            for (int k = 0; k < 3; ++k) t.v[k] = local(t.v[k], s);
            scenes[resident[s]].tris.push_back(t);
        }
        for (int id : resident) {
            renumber_forces(scenes[id]);
            renumber_energies(scenes[id]);
        }
        return 0;
    }

    // Make `ids` (ascending scene ids) the device world.
    int rebuild(const std::vector<int>& ids) {
        int rc = sync_to_host();
        if (rc) return rc;
        for (int id : resident) scenes[id].slot = -1;
        resident.clear();
        resident_bytes = 0;
        if (ids.empty()) return 0;
        avbd_world w{};
        std::vector<avbd_scene> sc;
        std::vector<avbd_body> bodies;
        std::vector<avbd_joint> joints;
        std::vector<avbd_spring> springs;
        std::vector<avbd_manifold> mans;
        std::vector<avbd_tet> tets;
        std::vector<avbd_tri> tris;
        long long fser = 0;
        int eser = 0;
        for (size_t s = 0; s < ids.size(); ++s) {
            const SceneHost& h = scenes[ids[s]];
            const int boff = (int)bodies.size();
            auto g = [&](int b) { return b < 0 ? b : b + boff; };
            sc.push_back(h.sc);
            for (avbd_body b : h.bodies) {
                b.scene = (int)s;
                bodies.push_back(b);
            }
            long long nf = 0;
            for (avbd_joint j : h.joints) {
                j.body_a = g(j.body_a); j.body_b = g(j.body_b); nf = std::max<long long>(nf, j.serial + 1); j.serial += fser;
                joints.push_back(j);
            }
            for (avbd_spring p : h.springs) {
                p.body_a = g(p.body_a); p.body_b = g(p.body_b); nf = std::max<long long>(nf, p.serial + 1); p.serial += fser;
                springs.push_back(p);
            }
            for (avbd_manifold m : h.manifolds) {
                m.body_a = g(m.body_a); m.body_b = g(m.body_b); nf = std::max<long long>(nf, m.serial + 1); m.serial += fser;
                mans.push_back(m);
            }
            fser += nf;
            int ne = 0;
            for (avbd_tet t : h.tets) {
                for (int k = 0; k < 4; ++k) t.v[k] = g(t.v[k]);
                ne = std::max(ne, t.serial + 1);
                t.serial += eser;
                tets.push_back(t);
            }
            for (avbd_tri t : h.tris) {
                for (int k = 0; k < 3; ++k) t.v[k] = g(t.v[k]);
                ne = std::max(ne, t.serial + 1);
                t.serial += eser;
                tris.push_back(t);
            }
            eser += ne;
        }
        w.n_scenes = (int)sc.size(); w.n_bodies = (int)bodies.size(); w.n_joints = (int)joints.size();
        w.n_springs = (int)springs.size(); w.n_manifolds = (int)mans.size(); w.n_tets = (int)tets.size();
        w.n_tris = (int)tris.size();
        w.scenes = sc.data(); w.bodies = bodies.data(); w.joints = joints.data(); w.springs = springs.data();
        w.manifolds = mans.data(); w.tets = tets.data(); w.tris = tris.data();
        if ((rc = avbd_upload(ctx, &w, 0))) return rc;
        resident = ids;
        for (size_t s = 0; s < ids.size(); ++s) scenes[ids[s]].slot = (int)s;
        int64_t bytes = 0;
        if ((rc = avbd_get_memory(ctx, &bytes))) return rc;
        resident_bytes = bytes;
        st.layout_rebuilds += 1;
        return 0;
    }

    // Admit queued requests into free slots; returns the resident set the
    // admitted requests need (or the current one).
    int admit(std::vector<int>* plan, bool* changed) {
        for (SceneHost& h : scenes) h.want = h.slot >= 0;
        std::vector<int> cur = resident;
        *changed = false;
        std::deque<long long> keep;
        while (!queue.empty()) {
            long long t = queue.front();
            queue.pop_front();
            Req& r = reqs[t];
            SceneHost& h = scenes[r.info.scene];
            if ((int)running.size() >= max_slots || h.busy) {
                keep.push_back(t);
                continue;
            }
            if (!h.want) {
                // bring the scene in, evicting idle scenes least recently used first
                std::vector<int> trial = cur;
                trial.push_back(r.info.scene);
                std::vector<int> idle;
                for (int id : cur)
                    if (!scenes[id].busy) idle.push_back(id);
                std::stable_sort(idle.begin(), idle.end(),
                                 [&](int a, int b) { return scenes[a].last_used < scenes[b].last_used; });
                size_t ev = 0;
                long long bytes = set_bytes(trial);
                while ((bytes < 0 || bytes > budget) && ev < idle.size()) {
                    trial.erase(std::find(trial.begin(), trial.end(), idle[ev++]));
                    bytes = set_bytes(trial);
                }
                if (bytes < 0 || bytes > budget) {
                    keep.push_back(t);  // waits until running scenes finish
                    continue;
                }
                for (size_t k = 0; k < ev; ++k) scenes[idle[k]].want = false;
                cur = trial;
                h.want = true;
                *changed = true;
            }
            h.busy = true;
            r.info.state = 1;
            running.push_back(t);
        }
        queue.swap(keep);
        std::sort(cur.begin(), cur.end());
        *plan = cur;
        return 0;
    }
};

extern "C" {

int avbd_server_create(int32_t device, int64_t hbm_budget_bytes, int32_t max_slots, avbd_server** out) {
    if (!out) return avbd_set_error(AVBD_E_ARG, "null output pointer");
    if (hbm_budget_bytes <= 0 || max_slots <= 0) return avbd_set_error(AVBD_E_ARG, "budget and slots must be positive");
    avbd_server* s = new avbd_server();
    s->device = device;
    s->budget = hbm_budget_bytes;
    s->max_slots = max_slots;
    int rc = avbd_create(device, &s->ctx);
    if (rc) {
        delete s;
        return rc;
    }
    // A body's accumulation order must not depend on the size of the launch
    // it lands in, so that a scene's result is the same whichever scenes
    // share its ticks: bodies with forces at a fixed 4 lanes, FEM-only
    // particles with canonical lane sums (the same tree at every width), so
    // their width is tuned to each tick's launch sizes.
    if ((rc = avbd_set_option(s->ctx, "lanes", 4)) || (rc = avbd_set_option(s->ctx, "fem_canonical_sums", 1)) ||
        (rc = avbd_set_option(s->ctx, "fem_lanes", 0))) {
        avbd_server_destroy(s);
        return rc;
    }
    *out = s;
    return 0;
}

int avbd_server_destroy(avbd_server* s) {
    if (!s) return 0;
    if (s->ctx) avbd_destroy(s->ctx);
    delete s;
    return 0;
}

avbd_ctx* avbd_server_context(avbd_server* s) { return s ? s->ctx : nullptr; }

double avbd_server_now(avbd_server* s) { return s ? seconds_since(s->t0) : 0.0; }

int avbd_server_add_scenes(avbd_server* s, const avbd_world* w, int32_t* first_id) {
    if (!s || !w || !first_id) return avbd_set_error(AVBD_E_ARG, "null argument");
    const int S = w->n_scenes, N = w->n_bodies;
    if (S < 1) return avbd_set_error(AVBD_E_ARG, "world needs at least one scene");
    if (N < 0 || w->n_joints < 0 || w->n_springs < 0 || w->n_manifolds < 0 || w->n_tets < 0 || w->n_tris < 0)
        return avbd_set_error(AVBD_E_ARG, "negative count");
    std::vector<int> off(S + 1, 0);
    for (int i = 0; i < N; ++i) {
        int sc = w->bodies[i].scene;
        if (sc < 0 || sc >= S) return avbd_set_error(AVBD_E_ARG, "body scene out of range");
        if (i > 0 && sc < w->bodies[i - 1].scene) return avbd_set_error(AVBD_E_ARG, "bodies must be grouped by scene in order");
        off[sc + 1]++;
    }
    for (int k = 0; k < S; ++k) off[k + 1] += off[k];
    auto ok = [&](int b) { return b >= 0 && b < N; };
    auto scene_of = [&](int b) { return w->bodies[b].scene; };
    std::vector<SceneHost> hs(S);
    for (int k = 0; k < S; ++k) {
        SceneHost& h = hs[k];
        h.sc = w->scenes[k];
        h.bodies.assign(w->bodies + off[k], w->bodies + off[k + 1]);
        for (avbd_body& b : h.bodies) {
            b.scene = 0;
            if (!(b.flags & AVBD_BODY_PARTICLE)) ++h.nonpart;
        }
    }
    auto local = [&](int b, int k) { return b < 0 ? b : b - off[k]; };
    for (int i = 0; i < w->n_joints; ++i) {
        avbd_joint j = w->joints[i];
        if (!ok(j.body_b) || (j.body_a != -1 && !ok(j.body_a))) return avbd_set_error(AVBD_E_ARG, "joint body index");
        int k = scene_of(j.body_b);
        if (j.body_a >= 0 && scene_of(j.body_a) != k) return avbd_set_error(AVBD_E_ARG, "joint links two scenes");
        j.body_a = local(j.body_a, k);
        j.body_b = local(j.body_b, k);
        hs[k].joints.push_back(j);
    }
    for (int i = 0; i < w->n_springs; ++i) {
        avbd_spring p = w->springs[i];
        if (!ok(p.body_b) || (p.body_a != -1 && !ok(p.body_a))) return avbd_set_error(AVBD_E_ARG, "spring body index");
        int k = scene_of(p.body_b);
        if (p.body_a >= 0 && scene_of(p.body_a) != k) return avbd_set_error(AVBD_E_ARG, "spring links two scenes");
        p.body_a = local(p.body_a, k);
        p.body_b = local(p.body_b, k);
        hs[k].springs.push_back(p);
    }
    for (int i = 0; i < w->n_manifolds; ++i) {
        avbd_manifold m = w->manifolds[i];
        if (!ok(m.body_a) || !ok(m.body_b) || m.n_contacts < 0 || m.n_contacts > 8)
            return avbd_set_error(AVBD_E_ARG, "manifold record");
        int k = scene_of(m.body_a);
        if (scene_of(m.body_b) != k) return avbd_set_error(AVBD_E_ARG, "manifold links two scenes");
        m.body_a = local(m.body_a, k);
        m.body_b = local(m.body_b, k);
        hs[k].manifolds.push_back(m);
    }
    for (int i = 0; i < w->n_tets; ++i) {
        avbd_tet t = w->tets[i];
        for (int v = 0; v < 4; ++v)
            if (!ok(t.v[v]) || scene_of(t.v[v]) != scene_of(t.v[0])) return avbd_set_error(AVBD_E_ARG, "tet vertex index");
        int k = scene_of(t.v[0]);
        for (int v = 0; v < 4; ++v) t.v[v] = local(t.v[v], k);
        hs[k].tets.push_back(t);
    }
    for (int i = 0; i < w->n_tris; ++i) {
        avbd_tri t = w->tris[i];
        for (int v = 0; v < 3; ++v)
            if (!ok(t.v[v]) || scene_of(t.v[v]) != scene_of(t.v[0])) return avbd_set_error(AVBD_E_ARG, "triangle vertex index");
        int k = scene_of(t.v[0]);
        for (int v = 0; v < 3; ++v) t.v[v] = local(t.v[v], k);
        hs[k].tris.push_back(t);
    }
    for (SceneHost& h : hs) {
        renumber_forces(h);
        renumber_energies(h);
        long long bytes = world_device_bytes(h.counts(), 0);
        if (bytes < 0) return (int)bytes;
        if (bytes > s->budget) return avbd_set_error(AVBD_E_CAPACITY, "a scene alone exceeds the HBM budget");
    }
    *first_id = (int32_t)s->scenes.size();
    for (SceneHost& h : hs) s->scenes.push_back(std::move(h));
    return 0;
}

int avbd_server_preload(avbd_server* s) {
    if (!s) return avbd_set_error(AVBD_E_ARG, "null server");
    if (!s->running.empty()) return avbd_set_error(AVBD_E_STATE, "preload while requests are running");
    std::vector<int> ids;
    for (int id = 0; id < (int)s->scenes.size(); ++id) {
        ids.push_back(id);
        long long b = s->set_bytes(ids);
        if (b < 0 || b > s->budget) {
            ids.pop_back();
            break;
        }
    }
    if (ids == s->resident) return 0;
    return s->rebuild(ids);
}

int avbd_server_submit(avbd_server* s, int32_t scene_id, int32_t n_steps, double t_arrival, int64_t* ticket) {
    if (!s || !ticket) return avbd_set_error(AVBD_E_ARG, "null argument");
    if (scene_id < 0 || scene_id >= (int)s->scenes.size()) return avbd_set_error(AVBD_E_ARG, "unknown scene id");
    if (n_steps < 1) return avbd_set_error(AVBD_E_ARG, "a request asks for at least one step");
    avbd_server::Req r;
    r.info.ticket = s->next_ticket++;
    r.info.scene = scene_id;
    r.info.steps = n_steps;
    r.info.state = 0;
    r.info.t_arrival = t_arrival >= 0 ? t_arrival : seconds_since(s->t0);
    r.info.t_start = -1.0;
    r.info.t_done = -1.0;
    r.remaining = n_steps;
    s->reqs[r.info.ticket] = r;
    s->queue.push_back(r.info.ticket);
    *ticket = r.info.ticket;
    return 0;
}

int avbd_server_tick(avbd_server* s, int32_t* n_running) {
    if (!s) return avbd_set_error(AVBD_E_ARG, "null server");
    std::vector<int> plan;
    bool changed = false;
    int rc = s->admit(&plan, &changed);
    if (rc) return rc;
    if (changed && (rc = s->rebuild(plan))) return rc;
    if (n_running) *n_running = (int32_t)s->running.size();
    if (s->running.empty()) return 0;
    std::vector<int32_t> slots;
    slots.reserve(s->running.size());
    const double t_begin = seconds_since(s->t0);
    for (long long t : s->running) {
        avbd_server::Req& r = s->reqs[t];
        slots.push_back(s->scenes[r.info.scene].slot);
        if (r.info.t_start < 0) r.info.t_start = t_begin;
    }
    std::sort(slots.begin(), slots.end());
    if ((rc = avbd_step_scenes(s->ctx, slots.data(), (int32_t)slots.size(), 1))) return rc;
    const double t_end = seconds_since(s->t0);
    s->st.ticks += 1;
    s->st.scene_steps += (long long)slots.size();
    std::vector<long long> still;
    for (long long t : s->running) {
        avbd_server::Req& r = s->reqs[t];
        SceneHost& h = s->scenes[r.info.scene];
        h.last_used = s->st.ticks;
        if (--r.remaining == 0) {
            r.info.state = 2;
            r.info.t_done = t_end;
            h.busy = false;
            s->finished.push_back(t);
            s->st.completed += 1;
        } else {
            still.push_back(t);
        }
    }
    s->running.swap(still);
    return 0;
}

int avbd_server_request(avbd_server* s, int64_t ticket, avbd_request_info* out) {
    if (!s || !out) return avbd_set_error(AVBD_E_ARG, "null argument");
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return avbd_set_error(AVBD_E_ARG, "unknown or already polled ticket");
    *out = it->second.info;
    return 0;
}

int avbd_server_poll(avbd_server* s, avbd_request_info* out, int32_t cap, int32_t* n_out) {
    if (!s || !n_out || (!out && cap > 0)) return avbd_set_error(AVBD_E_ARG, "null argument");
    int n = 0;
    size_t k = 0;
    for (; k < s->finished.size() && n < cap; ++k) {
        auto it = s->reqs.find(s->finished[k]);
        out[n++] = it->second.info;
        s->reqs.erase(it);
    }
    s->finished.erase(s->finished.begin(), s->finished.begin() + (long)k);
    *n_out = n;
    return 0;
}

int avbd_server_get_stats(avbd_server* s, avbd_server_stats* out) {
    if (!s || !out) return avbd_set_error(AVBD_E_ARG, "null argument");
    *out = s->st;
    out->resident_bytes = s->resident_bytes;
    out->resident_scenes = (int32_t)s->resident.size();
    out->queued = (int32_t)s->queue.size();
    out->running = (int32_t)s->running.size();
    return 0;
}

int avbd_server_download_scene(avbd_server* s, int32_t scene_id, avbd_world* out) {
    if (!s || !out) return avbd_set_error(AVBD_E_ARG, "null argument");
    if (scene_id < 0 || scene_id >= (int)s->scenes.size()) return avbd_set_error(AVBD_E_ARG, "unknown scene id");
    if (s->scenes[scene_id].slot >= 0) {
        int rc = s->sync_to_host();
        if (rc) return rc;
    }
    const SceneHost& h = s->scenes[scene_id];
    const bool sizes_only = out->n_bodies == 0 && out->n_joints == 0 && out->n_springs == 0 &&
                            out->n_manifolds == 0 && out->n_tets == 0 && out->n_tris == 0 && out->n_scenes == 0;
    if (!sizes_only) {
        if (out->n_scenes < 1 || out->n_bodies < (int)h.bodies.size() || out->n_joints < (int)h.joints.size() ||
            out->n_springs < (int)h.springs.size() || out->n_manifolds < (int)h.manifolds.size() ||
            out->n_tets < (int)h.tets.size() || out->n_tris < (int)h.tris.size())
            return avbd_set_error(AVBD_E_ARG, "output arrays smaller than the scene");
        out->scenes[0] = h.sc;
        std::copy(h.bodies.begin(), h.bodies.end(), out->bodies);
        std::copy(h.joints.begin(), h.joints.end(), out->joints);
        std::copy(h.springs.begin(), h.springs.end(), out->springs);
        std::copy(h.manifolds.begin(), h.manifolds.end(), out->manifolds);
        std::copy(h.tets.begin(), h.tets.end(), out->tets);
        std::copy(h.tris.begin(), h.tris.end(), out->tris);
    }
    out->n_scenes = 1;
    out->n_bodies = (int)h.bodies.size();
    out->n_joints = (int)h.joints.size();
    out->n_springs = (int)h.springs.size();
    out->n_manifolds = (int)h.manifolds.size();
    out->n_tets = (int)h.tets.size();
    out->n_tris = (int)h.tris.size();
    return 0;
}

}  // extern "C"
