// avbd_host.hip -- context, world upload/download and step orchestration of
// the device-resident AVBD solver; implements the C ABI of avbd_capi.h.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "avbd_internal.h"
#include "avbd_kernels.h"
#include "avbd_tune.h"
#include "avbd_world.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

// Temporary storage given to the device-wide scans (n_scan items) and the
// new-pair sort (n_sort keys, data in double buffers of its own): an upper
// bound of their needs (a 16-byte lookback state per 256 scan items; digit
// histograms and lookback states of the radix passes, well under a byte per
// key).  The upload takes the larger of this and the library's own queries,
// so the byte count of a world is known without a device.
size_t cub_tmp_bound(long long n_scan, long long n_sort) {
    return (size_t)std::max<long long>(16 * (n_scan / 256 + 64) + 1024, n_sort + (1 << 20));
}

// Capacity of the broad-phase candidate and new-pair buffers.
// Broad-phase candidate buffer: every overlapping pair is a candidate,
// those already in contact (at most the manifold capacity) included.
long long pair_capacity(long long N, long long mcap) { return std::max<long long>(65536, 4 * N) + mcap; }

// Significant bits of the new-pair sort keys (i << 32 | j, i < j < N).
int pair_key_bits(int N) {
    int b = 0;
    while (b < 31 && (1ll << b) < N) ++b;
    return 32 + b;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(AVBD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

int grid_for(long long n, int block = 256, int max_blocks = 4096) {
    long long b = (n + block - 1) / block;
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return int(b);
}

}  // namespace

// Device scratch of avbd_get_instances_noalloc: a file-scope array, one
// copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the
// symbol): its device address is looked up once per device, with that device
// current, by noalloc_scratch_init (called from avbd_create), and kept in
// g_noalloc_ptr.  Layout for n body slots: n instances (128 bytes each), then
// n live flags, n live offsets and one total (ints).
constexpr int kNoallocMaxBodies = AVBD_NOALLOC_MAX_BODIES;
constexpr size_t kNoallocBytes =
    (size_t)kNoallocMaxBodies * (sizeof(avbd_instance) + 2 * sizeof(int)) + 256;
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

// Called with `device` current.
int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(AVBD_E_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(AVBD_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc_ptr[device] = p;
    return 0;
}

void* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    return device >= 0 && device < kMaxDevices ? g_noalloc_ptr[device] : nullptr;
}
}  // namespace

struct avbd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // second stream for the FEM bucket of a colour
    hipEvent_t fork = nullptr, join = nullptr;
    hipStream_t side2 = nullptr;  // third stream: the per-scene kernel of the small scenes
    hipEvent_t fork2 = nullptr, join2 = nullptr;
    // Small scenes run phases 4-7 in one workgroup each (k_scene_step):
    // scenes of at most fused_max bodies with no breakable joint or spring,
    // when the two lane widths agree (the kernel has one).
    bool use_fused = false;
    int fused_max = 512;
    // a world with at most SMALL_PAIR_SORT possible candidate pairs sorts its
    // new pairs in one workgroup, without reading their count back
    bool small_pairs = false;
    // a scene whose constraint graph changed a little updates its last
    // colouring from the changed bodies instead of colouring from scratch
    bool colour_updates = true;
    // FEM-only sweeps read each vertex's neighbour positions from LDS (its
    // neighbour list) instead of gathering four positions per element
    bool use_nb = true;
    int spare_bodies = 0;  // dead slots per scene added at upload (avbd_add_bodies fills them)
    // Debug/render output (option debug_output): counters, offsets and the
    // contact points come from alloc (an upload frees them; set up again on
    // the next use), the output arrays are grown here and owned here.
    bool dbg_on = false;
    struct Debug {
        int* lc = nullptr;
        int* loff = nullptr;
        int* pc = nullptr;
        int* poff = nullptr;
        void* cub = nullptr;
        size_t cub_bytes = 0;
        float4* rgb = nullptr;  // per body colour (avbd_set_body_colours), null = white
        float* lines = nullptr;
        size_t lcap = 0;
        float* points = nullptr;
        size_t pcap = 0;
        long long nl = 0, np = 0;
        bool ready = false;
        void* inst = nullptr;  // avbd_get_instances scratch (owned here)
        size_t inst_bytes = 0;
    } dbg;
    std::vector<int> scene_bodies;      // per scene, at upload
    int max_scene_bodies = 0;
    std::vector<char> scene_breakable;  // per scene: a joint or spring with a finite fracture
    std::vector<char> fused_on;         // per scene: ST_FUSED as last written
    bool fuse_single = false;           // option fused_single_scene
    int scene_lds = 60 * 1024;          // option scene_lds_bytes
    std::vector<char> scene_trust;      // per scene: ADAPTIVE with AVBD_SCENE_TRUST_REGION
    bool any_trust = false;
    std::vector<int> fusable_ids;
    int* fused_list = nullptr;          // device copy of fusable_ids
    std::vector<void*> allocs;
    DevWorld w{};
    ManifoldBuf mb[2]{};
    int cur = 0;
    int* host_word = nullptr;  // pinned: one status word read back mid-step
    int4* items = nullptr;  // broad-phase work items (scene, i-tile, j-tile)
    int n_items = 0;
    int n_tiles = 0;
    // The same on the host, grouped by scene (item_off, tile_off: per-scene
    // ranges), and the buffers a step over a subset of the scenes fills with
    // the running scenes' items and tiles (pinned staging, device copy).
    std::vector<int4> h_items;
    std::vector<int> h_item_off, h_tile_off;
    int4* sub_items = nullptr;
    int* sub_tiles = nullptr;
    int* run_dev = nullptr;    // the running scene ids of a subset step (device, S)
    int* run_pin = nullptr;    // their pinned staging (S)
    hipEvent_t run_ev = nullptr;  // recorded after the copy out of run_pin
    int4* pin_items = nullptr;
    int* pin_tiles = nullptr;
    std::vector<int> run_ids;  // the running scenes of a subset step, ascending
    void* cub_tmp = nullptr;
    size_t cub_bytes = 0;
    int lanes = 0;      // 0 = chosen by the tuner
    int fem_lanes = 0;  // 0 = chosen by the tuner
    // option fem_canonical_sums: FEM-only buckets sum their lanes in one
    // fixed tree (k_primal_canon), so the FEM width (4 to 32, tuner class 2)
    // can change with the launch size without changing results
    bool fem_canon = false;
    LaneTuner tuner;
    bool have_world = false;
    std::vector<int> bucket_off = std::vector<int>(AVBD_BUCKETS + 1, 0);
    // host mirrors of the static parts
    std::vector<avbd_scene> scenes;
    std::vector<int> sc_off;
    std::vector<int32_t> tet_serial, tri_serial;  // creation serials of the device's elements
    // Removed bodies stay on the device as dead slots at the end of their
    // scene's index range (static, no broad-phase sphere, no energies);
    // removed elements keep their index.  Downloads skip both.
    std::vector<char> body_dead, tet_dead, tri_dead;
    // bodies removed by the current avbd_step call: scene, index in the scene
    // at the moment of removal (RemoveBody order)
    std::vector<int32_t> removed_scene, removed_index;
    std::vector<int32_t> host_colours;
    hipEvent_t ev[8]{};
    avbd_timings timings{};
    int last_ncol = 0;
    // set by avbd_step_scenes when some scenes sit the step out: the dual
    // and stick passes then walk a list of the running scenes' manifolds
    bool subset = false;
    // Per-launch timing (option kernel_timing): HIP events recorded on the
    // stream the work runs on, around each dual launch (K_DUAL) and each
    // primal bucket launch (K_FORCES, K_FEM).  Two events per launch and no
    // more: on small launches every extra event shows in the step time.
    enum { K_DUAL = 0, K_FORCES = 1, K_FEM = 2, K_KINDS = 3 };
    struct Mark {
        int a, b, kind;
    };
    bool ktiming = false;
    std::vector<hipEvent_t> kev;
    std::vector<Mark> kmarks;
    int kev_used = 0;
    long long kstat_n[K_KINDS] = {};
    double kstat_ms[K_KINDS] = {};

    int next_event() {
        if (kev_used >= (int)kev.size()) {
            hipEvent_t e;
            hipEventCreate(&e);
            kev.push_back(e);
        }
        return kev_used++;
    }
    // returns a mark id for mark_end (-1 when timing is off)
    int mark_begin(int kind) {
        if (!ktiming) return -1;
        int e = next_event();
        hipEventRecord(kev[e], stream);
        kmarks.push_back({e, -1, kind});
        return (int)kmarks.size() - 1;
    }
    void mark_end(int id) {
        if (id < 0) return;
        int e = next_event();
        hipEventRecord(kev[e], stream);
        kmarks[id].b = e;
    }
    // called after both streams have been synchronised
    void collect_kernel_times() {
        for (const Mark& m : kmarks) {
            if (m.b < 0) continue;
            float ms = 0.f;
            hipEventElapsedTime(&ms, kev[m.a], kev[m.b]);
            kstat_n[m.kind] += 1;
            kstat_ms[m.kind] += ms;
        }
        kev_used = 0;
        kmarks.clear();
    }

    // Step graphs: the launches of phase 5 (the iterations) and phase 7
    // (post-stabilisation) captured once per launch shape and replayed;
    // prediction and velocities, one kernel each, launch between them with
    // the phase events recorded from the host (events recorded inside a
    // replayed graph do not give ordered timestamps).  A shape is captured
    // the second time it occurs, so shapes that never repeat (a served
    // tick's changing scene set) stay eager.
    struct StepGraph {
        std::vector<int> key;
        hipGraphExec_t iters = nullptr;  // phase 5: the iterations
        hipGraphExec_t post = nullptr;   // phase 7: post-stabilisation (if any scene has it)
        bool captured = false;
        long long used = 0;
    };
    static constexpr size_t kMaxGraphs = 16;
    bool use_graphs = true;
    std::vector<StepGraph> graphs;
    long long graph_clock = 0;
    long long graph_launches = 0, graph_captures = 0;
    static void destroy(StepGraph& g) {
        if (g.iters) hipGraphExecDestroy(g.iters);
        if (g.post) hipGraphExecDestroy(g.post);
        g.iters = g.post = nullptr;
        g.captured = false;
    }
    void drop_graphs() {
        for (StepGraph& g : graphs) destroy(g);
        graphs.clear();
    }

    // dry: allocations are only counted (world_device_bytes)
    bool dry = false;
    size_t alloc_bytes = 0;

    template <class T>
    int alloc(T** p, size_t count) {
        void* q = nullptr;
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        alloc_bytes += bytes;
        if (dry) {
            *p = nullptr;
            return 0;
        }
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) return fail(AVBD_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        e = hipMemsetAsync(q, 0, bytes, stream);
        if (e != hipSuccess) return fail(AVBD_E_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
        allocs.push_back(q);
        *p = static_cast<T*>(q);
        return 0;
    }
    void release() {
        if (stream) hipStreamSynchronize(stream);
        drop_graphs();
        for (void* p : allocs) hipFree(p);
        allocs.clear();
        alloc_bytes = 0;
        kev_used = 0;
        kmarks.clear();
        w = DevWorld{};
        mb[0] = mb[1] = ManifoldBuf{};
        items = nullptr;
        n_items = n_tiles = 0;
        sub_items = nullptr;
        sub_tiles = nullptr;
        if (pin_items) hipHostFree(pin_items);
        if (pin_tiles) hipHostFree(pin_tiles);
        if (run_pin) hipHostFree(run_pin);
        pin_items = nullptr;
        pin_tiles = nullptr;
        run_pin = nullptr;
        run_dev = nullptr;
        cub_tmp = nullptr;
        have_world = false;
        if (dbg.lines) hipFree(dbg.lines);
        if (dbg.points) hipFree(dbg.points);
        if (dbg.inst) hipFree(dbg.inst);
        dbg = Debug{};
    }
};

namespace {

int alloc_manifolds(avbd_ctx* c, ManifoldBuf& m, int cap) {
    int rc;
    if ((rc = c->alloc(&m.a, cap))) return rc;
    if ((rc = c->alloc(&m.b, cap))) return rc;
    if ((rc = c->alloc(&m.serial, cap))) return rc;
    if ((rc = c->alloc(&m.n, cap))) return rc;
    if ((rc = c->alloc(&m.fric, cap))) return rc;
    if ((rc = c->alloc(&m.hot, (size_t)cap * 8 * CONTACT_HOT_F4))) return rc;
    if ((rc = c->alloc(&m.cold, (size_t)cap * 8 * CONTACT_COLD_F4))) return rc;
    return 0;
}

template <class T>
int upload_vec(avbd_ctx* c, T* dst, const std::vector<T>& v) {
    if (v.empty()) return 0;
    HIP_TRY(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return 0;
}

template <class T>
int download_vec(avbd_ctx* c, std::vector<T>& v, const T* src, size_t n) {
    v.resize(n);
    if (n == 0) return 0;
    HIP_TRY(hipMemcpyAsync(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    return 0;
}

float4 f4(const float* p, float w) { return make_float4(p[0], p[1], p[2], w); }

// Exclusive prefix sum of n ints (in != out).  One tile: a single launch;
// more: tile sums (held in the sort/scan temporary buffer), their scan, the
// tiles.  A few launches of no more than n/16 threads each, where a
// decoupled-lookback scan over a world-sized array spends most of its time
// in the chained look-back between thousands of partitions.
int scan_ints(avbd_ctx* c, const int* in, int* out, int n) {
    if (n <= 0) return 0;
    if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) % 16)
        return fail(AVBD_E_ARG, "scan buffers must be 16-byte aligned");
    const int tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles == 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(256), 0, c->stream, in, out, n, (const int*)nullptr);
        return 0;
    }
    if ((size_t)(tiles + 1) * sizeof(int) > c->cub_bytes) return fail(AVBD_E_CAPACITY, "scan temporary too small");
    int* sums = reinterpret_cast<int*>(c->cub_tmp);
    hipLaunchKernelGGL(k_scan_reduce, dim3(tiles), dim3(256), 0, c->stream, in, n, sums);
    hipLaunchKernelGGL(k_scan_flags_1wg, dim3(1), dim3(1024), 0, c->stream, (const int*)sums, sums, tiles,
                       sums + tiles);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(256), 0, c->stream, in, out, n, (const int*)sums);
    return 0;
}

// Build the force incidence lists of the current manifold buffer, joints and
// springs: count, scan, fill, order by creation.
int build_incidence(avbd_ctx* c) {
    DevWorld& w = c->w;
    const int N = w.N;
    const int Fcap = c->w.Mcap + w.J + w.SP;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(N + 1)), dim3(256), 0, c->stream, w.cnt_i, N + 1, 0);
    hipLaunchKernelGGL(k_inc_count, dim3(grid_for(Fcap)), dim3(256), 0, c->stream, w);
    int rc = scan_ints(c, w.cnt_i, w.f_off, N + 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(N + 1)), dim3(256), 0, c->stream, w.cur_i, N + 1, 0);
    hipLaunchKernelGGL(k_inc_fill, dim3(grid_for(Fcap)), dim3(256), 0, c->stream, w);
    hipLaunchKernelGGL(k_inc_sort, dim3(grid_for(N)), dim3(256), 0, c->stream, w);
    return 0;
}

// Blocks of a bucket launch: enough for n bodies at G lanes, rounded up to a
// power of two (the kernel strides over the bucket, whose size it reads from
// device memory), so that a step graph keyed on it survives small changes.
int bucket_grid(const avbd_ctx* c, int n, int G, int block = LaneTuner::kDefaultBlock) {
    const int g = grid_for((long long)n * G, block, 16384);
    // In a world of several scenes at least 64 blocks (the surplus ones exit
    // at once): the launch shape of small buckets then does not change with
    // their exact size, so a served tick's step graph repeats although its
    // scene set changes.  A single small scene keeps the smallest grid.
    int p = c->w.S > 1 ? 64 : 1;
    while (p < g) p <<= 1;
    return std::min(p, 16384);
}

// Tuner class of a bucket launch: 0 bodies with forces, 1 FEM-only, 2
// FEM-only with canonical lane sums (widths 4 to 32).
int tuner_class(const avbd_ctx* c, bool fem) { return fem ? (c->fem_canon ? 2 : 1) : 0; }

template <bool FEM>
void launch_primal_g(avbd_ctx* c, int G, int grid, int block, int bucket, int iter, int poststab) {
    const DevWorld& w = c->w;
    if (FEM && c->fem_canon) {
        switch (G) {
            case 4: hipLaunchKernelGGL(k_primal_canon<4>, dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
            case 8: hipLaunchKernelGGL(k_primal_canon<8>, dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
            case 16: hipLaunchKernelGGL(k_primal_canon<16>, dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
            default: hipLaunchKernelGGL(k_primal_canon<32>, dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
        }
        return;
    }
    if (FEM) {
        switch (G) {
            case 1: hipLaunchKernelGGL(k_primal_fem<1>, dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
            case 2: hipLaunchKernelGGL(k_primal_fem<2>, dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
            case 4: hipLaunchKernelGGL(k_primal_fem<4>, dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
            case 8: hipLaunchKernelGGL(k_primal_fem<8>, dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
            case 16: hipLaunchKernelGGL(k_primal_fem<16>, dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
            default: hipLaunchKernelGGL(k_primal_fem<32>, dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
        }
        return;
    }
    switch (G) {
        case 1: hipLaunchKernelGGL((k_primal<1, FEM>), dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
        case 2: hipLaunchKernelGGL((k_primal<2, FEM>), dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
        case 4: hipLaunchKernelGGL((k_primal<4, FEM>), dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
        case 8: hipLaunchKernelGGL((k_primal<8, FEM>), dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
        case 16: hipLaunchKernelGGL((k_primal<16, FEM>), dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
        default: hipLaunchKernelGGL((k_primal<32, FEM>), dim3(grid), dim3(block), 0, c->stream, w, bucket, iter, poststab); break;
    }
}

// Lane width of a bucket launch of n bodies (class 0 = bodies with forces,
// 1 = FEM-only) without starting a tuning sample: the fixed width, the tuned
// one, or 0 while the tuner still samples this size.
int bucket_lanes(const avbd_ctx* c, bool fem, int n, int* block) {
    const int fixed = fem ? c->fem_lanes : c->lanes;
    *block = LaneTuner::kDefaultBlock;
    return fixed > 0 ? fixed : c->tuner.peek(tuner_class(c, fem), n, block);
}

// One bucket launch on c->stream with the lane width fixed by the user or
// picked by the tuner.
template <bool FEM>
void launch_bucket(avbd_ctx* c, int bucket, int n, int iter, int poststab) {
    int fixed = FEM ? c->fem_lanes : c->lanes;
    int measure = -1, block = LaneTuner::kDefaultBlock;
    int G = fixed > 0 ? fixed : c->tuner.choose(tuner_class(c, FEM), n, &measure, &block);
    const int mark = c->mark_begin(FEM ? avbd_ctx::K_FEM : avbd_ctx::K_FORCES);
    if (measure >= 0) (void)hipEventRecord(c->tuner.pending[measure].a, c->stream);
    launch_primal_g<FEM>(c, G, bucket_grid(c, n, G, block), block, bucket, iter, poststab);
    if (measure >= 0) (void)hipEventRecord(c->tuner.pending[measure].b, c->stream);
    c->mark_end(mark);
}

// Both buckets of a colour in one k_primal_pair launch, when both are small
// (together at most kPairBlocks blocks, well within one pass of the chip),
// their shapes are settled (fixed or tuned widths, 256-thread blocks) and
// an instantiation exists: forces at 4 lanes, FEM at 4 to 32.  Returns
// false when the colour takes the two-launch path.
constexpr int kPairBlocks = 512;
bool launch_pair(avbd_ctx* c, int colour, int n_gen, int n_fem, int iter, int poststab) {
    if (c->ktiming) return false;  // per-launch timing keeps the two classes apart
    int br = 0, bf = 0;
    const int gr = bucket_lanes(c, false, n_gen, &br), gf = bucket_lanes(c, true, n_fem, &bf);
    if (gr != 4 || gf < 4 || br != LaneTuner::kDefaultBlock || bf != LaneTuner::kDefaultBlock) return false;
    const int rb = bucket_grid(c, n_gen, gr), fb = bucket_grid(c, n_fem, gf);
    if (rb + fb > kPairBlocks) return false;
    const DevWorld& w = c->w;
    const dim3 grid(rb + fb), block(256);
    hipStream_t st = c->stream;
    if (c->fem_canon) {
        switch (gf) {
            case 4: hipLaunchKernelGGL((k_primal_pair<4, 4, true>), grid, block, 0, st, w, colour, iter, poststab, rb); break;
            case 8: hipLaunchKernelGGL((k_primal_pair<4, 8, true>), grid, block, 0, st, w, colour, iter, poststab, rb); break;
            case 16: hipLaunchKernelGGL((k_primal_pair<4, 16, true>), grid, block, 0, st, w, colour, iter, poststab, rb); break;
            default: hipLaunchKernelGGL((k_primal_pair<4, 32, true>), grid, block, 0, st, w, colour, iter, poststab, rb); break;
        }
        return true;
    }
    switch (gf) {
        case 4: hipLaunchKernelGGL((k_primal_pair<4, 4, false>), grid, block, 0, st, w, colour, iter, poststab, rb); break;
        case 8: hipLaunchKernelGGL((k_primal_pair<4, 8, false>), grid, block, 0, st, w, colour, iter, poststab, rb); break;
        case 16: hipLaunchKernelGGL((k_primal_pair<4, 16, false>), grid, block, 0, st, w, colour, iter, poststab, rb); break;
        default: hipLaunchKernelGGL((k_primal_pair<4, 32, false>), grid, block, 0, st, w, colour, iter, poststab, rb); break;
    }
    return true;
}

// One colour of a sweep: the bodies with forces and the FEM-only bodies, two
// instantiations that run concurrently (the FEM bucket on the side stream,
// forked from and joined back into the main stream with events, so the next
// colour still sees both), or one paired launch for small buckets.  Empty
// buckets are not launched.
int launch_colour(avbd_ctx* c, const int* off, int colour, int iter, int poststab) {
    const int n_gen = off[2 * colour + 1] - off[2 * colour], n_fem = off[2 * colour + 2] - off[2 * colour + 1];
    if (n_gen == 0 && n_fem == 0) return 0;
    if (n_gen > 0 && n_fem > 0 && launch_pair(c, colour, n_gen, n_fem, iter, poststab)) return 0;
    if (n_gen > 0 && n_fem > 0) {
        HIP_TRY(hipEventRecord(c->fork, c->stream));
        HIP_TRY(hipStreamWaitEvent(c->side, c->fork, 0));
        std::swap(c->stream, c->side);
        launch_bucket<true>(c, 2 * colour + 1, n_fem, iter, poststab);
        std::swap(c->stream, c->side);
        launch_bucket<false>(c, 2 * colour, n_gen, iter, poststab);
        HIP_TRY(hipEventRecord(c->join, c->side));
        HIP_TRY(hipStreamWaitEvent(c->stream, c->join, 0));
    } else if (n_gen > 0) {
        launch_bucket<false>(c, 2 * colour, n_gen, iter, poststab);
    } else {
        launch_bucket<true>(c, 2 * colour + 1, n_fem, iter, poststab);
    }
    return 0;
}

// Lane width of the per-scene kernel: the common width of the two bucket
// classes (4 when both are tuned), 0 when they differ.
int fused_lanes(const avbd_ctx* c) {
    if (c->lanes != c->fem_lanes) return 0;
    return c->lanes == 0 ? 4 : c->lanes;
}

// Mark on the device which scenes run in the per-scene kernel (ST_FUSED) and
// keep the list of those scenes; touches device memory only on a change.
int sync_fused(avbd_ctx* c) {
    DevWorld& w = c->w;
    const int S = w.S;
    // option fused_single_scene: a world of one small scene (the plain
    // per-scene step) in the per-scene kernel as well
    const bool on = (c->use_fused || (S == 1 && c->fuse_single)) && fused_lanes(c) > 0;
    std::vector<char> want(S, 0);
    // scenes updating a trust-region ratio need its kernel between sweeps
    for (int s = 0; s < S; ++s)
        want[s] = on && !c->scene_breakable[s] && !(s < (int)c->scene_trust.size() && c->scene_trust[s]) &&
                  c->scene_bodies[s] <= c->fused_max;
    if (want == c->fused_on) return 0;
    std::vector<int> st;
    int rc;
    if ((rc = download_vec(c, st, w.st, (size_t)S * ST_WORDS))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->fusable_ids.clear();
    for (int s = 0; s < S; ++s) {
        st[s * ST_WORDS + ST_FUSED] = want[s];
        if (want[s]) c->fusable_ids.push_back(s);
    }
    if ((rc = upload_vec(c, w.st, st)) || (rc = upload_vec(c, c->fused_list, c->fusable_ids))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->fused_on = want;
    c->drop_graphs();
    return 0;
}

// The per-scene kernel over the marked scenes (mode 0: phases 4-7, 1:
// post-stabilisation only); its blocks leave at once for scenes not running.
void launch_scene_step(avbd_ctx* c, hipStream_t st, int mode) {
    const int n = (int)c->fusable_ids.size();
    if (n == 0) return;
    const DevWorld& w = c->w;
    const int* list = c->fused_list;
    // dynamic LDS for a scene's body arrays and incidence lists (scenes
    // that need more run on global memory; 0: every scene on global memory)
    const int lds = c->scene_lds;
    switch (fused_lanes(c)) {
        case 1: hipLaunchKernelGGL((k_scene_step<1>), dim3(n), dim3(256), lds, st, w, list, mode, lds); break;
        case 2: hipLaunchKernelGGL((k_scene_step<2>), dim3(n), dim3(256), lds, st, w, list, mode, lds); break;
        case 4: hipLaunchKernelGGL((k_scene_step<4>), dim3(n), dim3(256), lds, st, w, list, mode, lds); break;
        case 8: hipLaunchKernelGGL((k_scene_step<8>), dim3(n), dim3(256), lds, st, w, list, mode, lds); break;
        case 16: hipLaunchKernelGGL((k_scene_step<16>), dim3(n), dim3(256), lds, st, w, list, mode, lds); break;
        case 32: hipLaunchKernelGGL((k_scene_step<32>), dim3(n), dim3(256), lds, st, w, list, mode, lds); break;
        default: break;
    }
}

// Sizes of every device array of a world.
struct WorldDims {
    int S = 0, N = 0, J = 0, SP = 0, T = 0, TR = 0, Mcap = 0;
    int n_tiles = 0, n_items = 0;
    long long pair_cap = 0;
    size_t scan_bytes = 0;
};

int make_dims(const WorldCounts& k, long long mcap_req, WorldDims* d) {
    if (k.S < 1) return fail(AVBD_E_ARG, "world needs at least one scene");
    const long long lim = (1ll << 29) - 1;
    if (k.N > lim || k.J > lim || k.SP > lim || k.T > lim || k.TR > lim || k.items > (1ll << 31) - 1)
        return fail(AVBD_E_CAPACITY, "world too large for 29-bit element indices");
    long long mcap = mcap_req > 0 ? mcap_req : 8 * k.nonpart + (k.N - k.nonpart) / 2 + 1024;
    mcap = std::max<long long>(mcap, k.M + 1024);
    if (mcap * 8 >= (1ll << 29)) return fail(AVBD_E_CAPACITY, "manifold capacity too large");
    d->S = int(k.S); d->N = int(k.N); d->J = int(k.J); d->SP = int(k.SP); d->T = int(k.T); d->TR = int(k.TR);
    d->Mcap = int(mcap);
    d->n_tiles = int(k.tiles);
    d->n_items = int(k.items);
    d->pair_cap = pair_capacity(k.N, mcap);
    if (d->pair_cap > (1ll << 31) - 1) return fail(AVBD_E_CAPACITY, "broad-phase pair buffer too large");
    d->scan_bytes = cub_tmp_bound(std::max<long long>(k.N + 1, mcap + 1), d->pair_cap);
    return 0;
}

// Allocate every device array of a world of the given dimensions (or, on a
// dry context, only count their bytes in c->alloc_bytes).
int allocate_world(avbd_ctx* c, const WorldDims& d, avbd_scene** d_sc, int** d_off, int** e_off, uint32_t** e_ent) {
    const int S = d.S, N = d.N, J = d.J, SP = d.SP, T = d.T, TR = d.TR, Mcap = d.Mcap;
    DevWorld& w = c->w;
    w.S = S; w.N = N; w.J = J; w.SP = SP; w.T = T; w.TR = TR; w.Mcap = Mcap;
    w.use_nb = c->use_nb ? 1 : 0;
    int rc = 0;
#define A(p, n) if ((rc = c->alloc(p, n))) return rc
    A(d_sc, S);
    A(&w.st, (size_t)S * ST_WORDS);
    A(&w.rho, S); A(&w.rho_acc, (size_t)3 * S);
    A(&c->fused_list, S);
    A(d_off, S + 1);
    w.sc = *d_sc;
    w.sc_off = *d_off;
    A(&w.pos, N); A(&w.rot, N); A(&w.scale, N); A(&w.vel, N); A(&w.angvel, N); A(&w.pvel, N); A(&w.pangvel, N);
    A(&w.drag, N);
    A(&w.inertia, N); A(&w.flags, N); A(&w.scene, N); A(&w.last_pos, N); A(&w.last_rot, N); A(&w.inert_pos, N);
    A(&w.inert_rot, N); A(&w.diff0, N); A(&w.diff1, N);
    if ((rc = alloc_manifolds(c, c->mb[0], Mcap))) return rc;
    if ((rc = alloc_manifolds(c, c->mb[1], Mcap))) return rc;
    c->cur = 0;
    w.m = c->mb[0];
    A(&w.j_a, J); A(&w.j_b, J); A(&w.j_serial, J); A(&w.j_flags, J); A(&w.j_ra, J); A(&w.j_rb, J); A(&w.j_rest, J);
    A(&w.j_c0, (size_t)J * 6); A(&w.j_row, (size_t)J * 36);
    A(&w.s_a, SP); A(&w.s_b, SP); A(&w.s_serial, SP); A(&w.s_flags, SP); A(&w.s_ra, SP); A(&w.s_rb, SP);
    A(&w.s_row, (size_t)SP * 6);
    A(&w.t_v, T); A(&w.t_d, (size_t)T * 3);
    A(&w.r_v, TR); A(&w.r_d, (size_t)TR * 2);
    w.f_cap = 16 * Mcap + 2 * (J + SP) + 1;
    A(&w.f_off, N + 1);
    A(&w.f_ent, w.f_cap);
    A(e_off, N + 1);
    A(e_ent, (size_t)4 * T + 3 * TR + 1);
    w.e_off = *e_off;
    w.e_ent = *e_ent;
    A(&w.ev_a, (size_t)4 * T + 3 * TR + 1); A(&w.ev_b, (size_t)4 * T + 3 * TR + 1);
    A(&w.ev_c, (size_t)4 * T + 3 * TR + 1); A(&w.ev_v, (size_t)4 * T + 3 * TR + 1);
    A(&w.nb_id, 2 * ((size_t)4 * T + 3 * TR) + 1); A(&w.nb_n, N); A(&w.ev_l, (size_t)4 * T + 3 * TR + 1);
    A(&w.colour, N); A(&w.col_cnt, AVBD_BUCKETS + 1); A(&w.col_off, AVBD_BUCKETS + 1); A(&w.col_list, N); A(&w.col_fem, N);
    A(&w.ccount, N); A(&w.cseed, N); A(&w.queue, (size_t)2 * N + 1); A(&w.cnt_i, N + 1); A(&w.cur_i, N + 1);
    w.pair_cap = (int)d.pair_cap;
    A(&w.pairs, w.pair_cap); A(&w.pairs_tmp, w.pair_cap);
    int* toff;
    int* tscene;
    A(&toff, S + 1); A(&tscene, std::max(d.n_tiles, 1));
    w.tile_off = toff;
    w.tile_scene = tscene;
    A(&w.tile_lo, std::max(d.n_tiles, 1)); A(&w.tile_hi, std::max(d.n_tiles, 1));
    A(&w.chunk_lo, (size_t)8 * std::max(d.n_tiles, 1)); A(&w.chunk_hi, (size_t)8 * std::max(d.n_tiles, 1));
    c->n_tiles = d.n_tiles;
    c->n_items = d.n_items;
    A(&c->items, std::max(d.n_items, 1));
    A(&c->sub_items, std::max(d.n_items, 1));
    A(&c->sub_tiles, std::max(d.n_tiles, 1));
    A(&c->run_dev, std::max(d.S, 1));
    A(&w.alive, Mcap + 1); A(&w.alive_off, Mcap + 1);
    A(&w.fracture_bodies, (size_t)2 * (J + SP) + 2);
    A(&w.g, G_WORDS);
    A(&w.next_serial, 1);
    char* tmp;
    A(&tmp, d.scan_bytes);
    c->cub_tmp = tmp;
    c->cub_bytes = d.scan_bytes;
#undef A
    return 0;
}

// Energy records -> device: the element arrays, the per-vertex incidence
// lists in creation (serial) order and the incidence-ordered element stream.
// Vertex indices must be < N.
int upload_energies(avbd_ctx* c, int N, const avbd_tet* tets, int T, const avbd_tri* tris, int TR) {
    DevWorld& w = c->w;
    struct ERef { int serial, kind, idx; };
    std::vector<ERef> refs;
    refs.reserve(T + TR);
    for (int i = 0; i < T; ++i) refs.push_back({tets[i].serial, 0, i});
    for (int i = 0; i < TR; ++i) refs.push_back({tris[i].serial, 1, i});
    const auto by_serial = [](const ERef& x, const ERef& y) { return x.serial < y.serial; };
    if (!std::is_sorted(refs.begin(), refs.end(), by_serial))  // builders write them in creation order
        std::stable_sort(refs.begin(), refs.end(), by_serial);
    std::vector<int> cnt(N + 1, 0);
    for (const ERef& r : refs) {
        if (r.kind == 0) for (int k = 0; k < 4; ++k) cnt[tets[r.idx].v[k]]++;
        else for (int k = 0; k < 3; ++k) cnt[tris[r.idx].v[k]]++;
    }
    // a FEM-only bucket entry packs a vertex's element count in 23 bits
    if (N > 0 && *std::max_element(cnt.begin(), cnt.end()) >= (1 << 23))
        return fail(AVBD_E_CAPACITY, "a vertex with 2^23 or more elements");
    std::vector<int> eoff(N + 1, 0);
    for (int i = 0; i < N; ++i) eoff[i + 1] = eoff[i] + cnt[i];
    std::vector<uint32_t> ent(std::max(1, eoff[N]));
    std::vector<int> cur(eoff.begin(), eoff.end() - 1);
    for (const ERef& r : refs) {
        if (r.kind == 0) {
            for (int k = 0; k < 4; ++k) ent[cur[tets[r.idx].v[k]]++] = MAKE_EENT(0, k, r.idx);
        } else {
            for (int k = 0; k < 3; ++k) ent[cur[tris[r.idx].v[k]]++] = MAKE_EENT(1, k, r.idx);
        }
    }
    std::vector<int4> tv(T), rv(TR);
    std::vector<float4> td((size_t)T * 3), rd((size_t)TR * 2);
    for (int i = 0; i < T; ++i) {
        const avbd_tet& t = tets[i];
        tv[i] = make_int4(t.v[0], t.v[1], t.v[2], t.v[3]);
        const float* d = t.dm_inv;
        td[i * 3 + 0] = make_float4(d[0], d[1], d[2], d[3]);
        td[i * 3 + 1] = make_float4(d[4], d[5], d[6], d[7]);
        td[i * 3 + 2] = make_float4(d[8], t.rest_volume, t.mu, t.lambda);
    }
    for (int i = 0; i < TR; ++i) {
        const avbd_tri& t = tris[i];
        rv[i] = make_int4(t.v[0], t.v[1], t.v[2], 0);
        rd[i * 2 + 0] = make_float4(t.dm_inv[0], t.dm_inv[1], t.dm_inv[2], t.dm_inv[3]);
        rd[i * 2 + 1] = make_float4(t.rest_area, t.mu, t.lambda, 0.f);
    }
    int rc;
    if ((rc = upload_vec(c, const_cast<int*>(w.e_off), eoff)) || (rc = upload_vec(c, const_cast<uint32_t*>(w.e_ent), ent)) ||
        (rc = upload_vec(c, w.t_v, tv)) || (rc = upload_vec(c, w.t_d, td)) || (rc = upload_vec(c, w.r_v, rv)) ||
        (rc = upload_vec(c, w.r_d, rd)))
        return rc;
    w.T = T;
    w.TR = TR;
    // the element stream in incidence order, gathered on the device (a
    // quarter of the host work and of the bytes over PCIe: each element's
    // records cross once, not once per vertex)
    if (eoff[N] > 0) {
        hipLaunchKernelGGL(k_element_stream, dim3(grid_for(eoff[N])), dim3(256), 0, c->stream, w, eoff[N]);
        hipLaunchKernelGGL(k_nb_build, dim3(grid_for(N)), dim3(256), 0, c->stream, w, 0, N);
        HIP_TRY(hipGetLastError());
    } else {
        // no elements: the one padding entry of each stream array is zero
        const std::vector<float4> z4(1, make_float4(0.f, 0.f, 0.f, 0.f));
        const std::vector<int4> zi(1, make_int4(0, 0, 0, 0));
        if ((rc = upload_vec(c, w.ev_a, z4)) || (rc = upload_vec(c, w.ev_b, z4)) || (rc = upload_vec(c, w.ev_c, z4)) ||
            (rc = upload_vec(c, w.ev_v, zi)))
            return rc;
    }
    c->tet_serial.resize(T);
    c->tri_serial.resize(TR);
    c->tet_dead.assign(T, 0);
    c->tri_dead.assign(TR, 0);
    for (int i = 0; i < T; ++i) c->tet_serial[i] = tets[i].serial;
    for (int i = 0; i < TR; ++i) c->tri_serial[i] = tris[i].serial;
    return 0;
}

// Broad-phase tiles and work items (scene, i-tile, j-tile >= i-tile) of the
// scenes with body offsets `off`.
int upload_tiles(avbd_ctx* c, const std::vector<int>& off) {
    DevWorld& w = c->w;
    const int S = (int)off.size() - 1;
    std::vector<int> toff(S + 1, 0), tscene, ioff(S + 1, 0);
    std::vector<int4> items;
    for (int s = 0; s < S; ++s) {
        const int tiles = (int)broad_tiles(off[s + 1] - off[s]);
        toff[s + 1] = toff[s] + tiles;
        for (int ti = 0; ti < tiles; ++ti) {
            tscene.push_back(s);
            for (int tj = ti; tj < tiles; ++tj) items.push_back(make_int4(s, ti, tj, 0));
        }
        ioff[s + 1] = (int)items.size();
    }
    int rc;
    if ((rc = upload_vec(c, const_cast<int*>(w.tile_off), toff)) ||
        (rc = upload_vec(c, const_cast<int*>(w.tile_scene), tscene)) || (rc = upload_vec(c, c->items, items)))
        return rc;
    c->n_tiles = (int)tscene.size();
    c->n_items = (int)items.size();
    if (c->pin_items) hipHostFree(c->pin_items);
    if (c->pin_tiles) hipHostFree(c->pin_tiles);
    if (c->run_pin) hipHostFree(c->run_pin);
    c->pin_items = nullptr;
    c->pin_tiles = nullptr;
    c->run_pin = nullptr;
    HIP_TRY(hipHostMalloc((void**)&c->pin_items, std::max<size_t>(items.size(), 1) * sizeof(int4), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&c->pin_tiles, std::max<size_t>(tscene.size(), 1) * sizeof(int), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&c->run_pin, std::max<size_t>(S, 1) * sizeof(int), hipHostMallocDefault));
    c->h_items = std::move(items);
    c->h_item_off = std::move(ioff);
    c->h_tile_off = std::move(toff);
    return 0;
}

int do_upload(avbd_ctx* c, const avbd_world* in, int64_t mcap_req) {
    const int S = in->n_scenes, N = in->n_bodies, J = in->n_joints, SP = in->n_springs;
    const int M = in->n_manifolds, T = in->n_tets, TR = in->n_tris;
    if (S < 1) return fail(AVBD_E_ARG, "world needs at least one scene");
    if (N < 0 || J < 0 || SP < 0 || M < 0 || T < 0 || TR < 0) return fail(AVBD_E_ARG, "negative count");
    if (!in->scenes || (N > 0 && !in->bodies) || (J > 0 && !in->joints) || (SP > 0 && !in->springs) ||
        (M > 0 && !in->manifolds) || (T > 0 && !in->tets) || (TR > 0 && !in->tris))
        return fail(AVBD_E_ARG, "null record array");
    // bodies of a scene are contiguous, scenes in order
    std::vector<int> off(S + 1, 0);
    for (int i = 0; i < N; ++i) {
        int s = in->bodies[i].scene;
        if (s < 0 || s >= S) return fail(AVBD_E_ARG, "body scene out of range");
        if (i > 0 && s < in->bodies[i - 1].scene) return fail(AVBD_E_ARG, "bodies must be grouped by scene in order");
        off[s + 1]++;
    }
    for (int s = 0; s < S; ++s) off[s + 1] += off[s];
    auto body_ok = [&](int b) { return b >= 0 && b < N; };
    for (int i = 0; i < J; ++i) {
        const avbd_joint& j = in->joints[i];
        if (!body_ok(j.body_b) || (j.body_a != -1 && !body_ok(j.body_a))) return fail(AVBD_E_ARG, "joint body index");
        if (j.body_a >= 0 && in->bodies[j.body_a].scene != in->bodies[j.body_b].scene)
            return fail(AVBD_E_ARG, "joint links two scenes");
    }
    for (int i = 0; i < SP; ++i) {
        const avbd_spring& s = in->springs[i];
        if (!body_ok(s.body_b) || (s.body_a != -1 && !body_ok(s.body_a))) return fail(AVBD_E_ARG, "spring body index");
        if (s.body_a >= 0 && in->bodies[s.body_a].scene != in->bodies[s.body_b].scene)
            return fail(AVBD_E_ARG, "spring links two scenes");
    }
    for (int i = 0; i < M; ++i) {
        const avbd_manifold& m = in->manifolds[i];
        if (!body_ok(m.body_a) || !body_ok(m.body_b)) return fail(AVBD_E_ARG, "manifold body index");
        if (in->bodies[m.body_a].scene != in->bodies[m.body_b].scene) return fail(AVBD_E_ARG, "manifold links two scenes");
        if (m.n_contacts < 0 || m.n_contacts > 8) return fail(AVBD_E_ARG, "manifold contact count");
    }
    // an element's vertices lie in one scene: colouring, removal and the
    // energy lists index a scene's own body range
    for (int i = 0; i < T; ++i)
        for (int k = 0; k < 4; ++k) {
            if (!body_ok(in->tets[i].v[k])) return fail(AVBD_E_ARG, "tet vertex index");
            if (in->bodies[in->tets[i].v[k]].scene != in->bodies[in->tets[i].v[0]].scene)
                return fail(AVBD_E_ARG, "tet links two scenes");
        }
    for (int i = 0; i < TR; ++i)
        for (int k = 0; k < 3; ++k) {
            if (!body_ok(in->tris[i].v[k])) return fail(AVBD_E_ARG, "triangle vertex index");
            if (in->bodies[in->tris[i].v[k]].scene != in->bodies[in->tris[i].v[0]].scene)
                return fail(AVBD_E_ARG, "triangle links two scenes");
        }

    c->release();
    WorldDims dims;
    int rc = make_dims(world_counts(in), mcap_req, &dims);
    if (rc) return rc;
    {
        size_t b1 = 0, b2 = 0, b3 = 0;
        int* dummy = nullptr;
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b1, dummy, dummy, N + 1, c->stream));
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, dummy, dummy, dims.Mcap + 1, c->stream));
        hipcub::DoubleBuffer<unsigned long long> kb(nullptr, nullptr);
        HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, b3, kb, (int)dims.pair_cap, 0, pair_key_bits(N), c->stream));
        dims.scan_bytes = std::max(dims.scan_bytes, std::max(b1, std::max(b2, b3)));
        // the tile sums of scan_ints
        const size_t tile_sums = ((size_t)std::max(N + 1, dims.Mcap + 1) / SCAN_TILE + 2) * sizeof(int);
        dims.scan_bytes = std::max(dims.scan_bytes, tile_sums);
    }
    DevWorld& w = c->w;
    avbd_scene* d_sc;
    int* d_off;
    int* e_off;
    uint32_t* e_ent;
    if ((rc = allocate_world(c, dims, &d_sc, &d_off, &e_off, &e_ent))) return rc;

    // ---- host staging (records -> device SoA) ----
    c->scenes.assign(in->scenes, in->scenes + S);
    c->sc_off = off;
    c->body_dead.assign(N, 0);
    c->scene_bodies.assign(S, 0);
    c->scene_breakable.assign(S, 0);
    for (int s = 0; s < S; ++s) c->scene_bodies[s] = off[s + 1] - off[s];
    {
        // every candidate pair lies inside one scene
        long long most = 0;
        for (int s = 0; s < S; ++s) most += (long long)c->scene_bodies[s] * (c->scene_bodies[s] - 1) / 2;
        c->small_pairs = most <= SMALL_PAIR_SORT;
    }
    c->max_scene_bodies = S > 0 ? *std::max_element(c->scene_bodies.begin(), c->scene_bodies.end()) : 0;
    for (int i = 0; i < J; ++i)
        for (int r = 0; r < 6; ++r)
            if (std::isfinite(in->joints[i].fracture[r])) c->scene_breakable[in->bodies[in->joints[i].body_b].scene] = 1;
    for (int i = 0; i < SP; ++i)
        if (std::isfinite(in->springs[i].fracture)) c->scene_breakable[in->bodies[in->springs[i].body_b].scene] = 1;
    c->fused_on.assign(S, 0);
    c->fusable_ids.clear();
    std::vector<int> st((size_t)S * ST_WORDS, 0);
    for (int s = 0; s < S; ++s) {
        st[s * ST_WORDS + ST_EMERGENCY] = in->scenes[s].emergency_stop;
        st[s * ST_WORDS + ST_REMOVED] = in->scenes[s].n_removed;
        st[s * ST_WORDS + ST_ACTIVE] = 1;
        st[s * ST_WORDS + ST_DIRTY] = DIRTY_FULL;
    }
    if ((rc = upload_vec(c, d_sc, c->scenes))) return rc;
    if ((rc = upload_vec(c, w.st, st))) return rc;
    if ((rc = upload_vec(c, d_off, off))) return rc;
    {
        // trust-region state: rho per scene, the sweep's two sums cleared
        std::vector<float> rho(S);
        c->scene_trust.assign(S, 0);
        c->any_trust = false;
        for (int s = 0; s < S; ++s) {
            rho[s] = in->scenes[s].trust_region_rho;
            c->scene_trust[s] = in->scenes[s].projection_mode == AVBD_PROJ_ADAPTIVE &&
                                (in->scenes[s].options & AVBD_SCENE_TRUST_REGION) != 0;
            c->any_trust = c->any_trust || c->scene_trust[s];
        }
        if ((rc = upload_vec(c, w.rho, rho)) || (rc = upload_vec(c, w.rho_acc, std::vector<double>((size_t)3 * S, 0.0))))
            return rc;
    }
    {
        std::vector<float4> pos(N), rot(N), scale(N), vel(N), av(N), pv(N), pav(N), inert(N);
        std::vector<uint32_t> flags(N);
        std::vector<int> scene(N);
        for (int i = 0; i < N; ++i) {
            const avbd_body& b = in->bodies[i];
            pos[i] = f4(b.pos, 0.f);
            rot[i] = make_float4(b.rot[0], b.rot[1], b.rot[2], b.rot[3]);
            scale[i] = f4(b.scale, b.radius);
            vel[i] = f4(b.vel, b.mass);
            av[i] = f4(b.ang_vel, b.friction);
            pv[i] = f4(b.prev_vel, 0.f);
            pav[i] = f4(b.prev_ang_vel, 0.f);
            inert[i] = f4(b.inertia, 0.f);
            flags[i] = b.flags;
            scene[i] = b.scene;
        }
        if ((rc = upload_vec(c, w.pos, pos)) || (rc = upload_vec(c, w.rot, rot)) || (rc = upload_vec(c, w.scale, scale)) ||
            (rc = upload_vec(c, w.vel, vel)) || (rc = upload_vec(c, w.angvel, av)) || (rc = upload_vec(c, w.pvel, pv)) ||
            (rc = upload_vec(c, w.pangvel, pav)) || (rc = upload_vec(c, w.inertia, inert)) ||
            (rc = upload_vec(c, w.flags, flags)) || (rc = upload_vec(c, w.scene, scene)))
            return rc;
    }
    long long max_serial = -1;
    {
        std::vector<int> a(J), b(J);
        std::vector<long long> ser(J);
        std::vector<uint32_t> fl(J);
        std::vector<float4> ra(J), rb(J), rest(J);
        std::vector<float> rows((size_t)J * 36);
        for (int i = 0; i < J; ++i) {
            const avbd_joint& j = in->joints[i];
            a[i] = j.body_a; b[i] = j.body_b; ser[i] = j.serial; fl[i] = j.flags & AVBD_FORCE_DISABLED;
            max_serial = std::max<long long>(max_serial, j.serial);
            ra[i] = f4(j.r_a, j.torque_arm);
            rb[i] = f4(j.r_b, 0.f);
            rest[i] = make_float4(j.rest_rot[0], j.rest_rot[1], j.rest_rot[2], j.rest_rot[3]);
            for (int r = 0; r < 6; ++r) {
                float* row = &rows[(size_t)(i * 6 + r) * 6];
                row[0] = j.stiffness[r]; row[1] = j.fracture[r]; row[2] = j.fmin[r]; row[3] = j.fmax[r];
                row[4] = j.lambda[r]; row[5] = j.penalty[r];
            }
        }
        if ((rc = upload_vec(c, w.j_a, a)) || (rc = upload_vec(c, w.j_b, b)) || (rc = upload_vec(c, w.j_serial, ser)) ||
            (rc = upload_vec(c, w.j_flags, fl)) || (rc = upload_vec(c, w.j_ra, ra)) || (rc = upload_vec(c, w.j_rb, rb)) ||
            (rc = upload_vec(c, w.j_rest, rest)) || (rc = upload_vec(c, w.j_row, rows)))
            return rc;
    }
    {
        std::vector<int> a(SP), b(SP);
        std::vector<long long> ser(SP);
        std::vector<uint32_t> fl(SP);
        std::vector<float4> ra(SP), rb(SP);
        std::vector<float> rows((size_t)SP * 6);
        for (int i = 0; i < SP; ++i) {
            const avbd_spring& s = in->springs[i];
            a[i] = s.body_a; b[i] = s.body_b; ser[i] = s.serial; fl[i] = s.flags & AVBD_FORCE_DISABLED;
            max_serial = std::max<long long>(max_serial, s.serial);
            ra[i] = f4(s.r_a, s.rest_length);
            rb[i] = f4(s.r_b, 0.f);
            float* row = &rows[(size_t)i * 6];
            row[0] = s.stiffness; row[1] = s.fracture; row[2] = s.fmin; row[3] = s.fmax; row[4] = s.lambda; row[5] = s.penalty;
        }
        if ((rc = upload_vec(c, w.s_a, a)) || (rc = upload_vec(c, w.s_b, b)) || (rc = upload_vec(c, w.s_serial, ser)) ||
            (rc = upload_vec(c, w.s_flags, fl)) || (rc = upload_vec(c, w.s_ra, ra)) || (rc = upload_vec(c, w.s_rb, rb)) ||
            (rc = upload_vec(c, w.s_row, rows)))
            return rc;
    }
    {
        // manifolds must be in creation order
        std::vector<int> order(M);
        for (int i = 0; i < M; ++i) order[i] = i;
        const auto by_serial = [&](int x, int y) { return in->manifolds[x].serial < in->manifolds[y].serial; };
        if (!std::is_sorted(order.begin(), order.end(), by_serial)) std::stable_sort(order.begin(), order.end(), by_serial);
        std::vector<int> a(M), b(M), n(M);
        std::vector<long long> ser(M);
        std::vector<float> fr(M);
        std::vector<float4> hot((size_t)M * 8 * CONTACT_HOT_F4, make_float4(0, 0, 0, 0));
        std::vector<float4> cold((size_t)M * 8 * CONTACT_COLD_F4, make_float4(0, 0, 0, 0));
        for (int k = 0; k < M; ++k) {
            const avbd_manifold& m = in->manifolds[order[k]];
            a[k] = m.body_a; b[k] = m.body_b; n[k] = m.n_contacts; ser[k] = m.serial; fr[k] = m.friction;
            max_serial = std::max<long long>(max_serial, m.serial);
            for (int ci = 0; ci < m.n_contacts; ++ci) {
                const avbd_contact& ct = m.contacts[ci];
                float4* h = &hot[((size_t)k * 8 + ci) * CONTACT_HOT_F4];
                h[3].w = ct.lambda[0];
                h[4].w = ct.lambda[1];
                h[5] = make_float4(ct.lambda[2], ct.penalty[0], ct.penalty[1], ct.penalty[2]);
                float4* cd = &cold[((size_t)k * 8 + ci) * CONTACT_COLD_F4];
                uint32_t id = ct.id;
                float idf;
                std::memcpy(&idf, &id, 4);
                cd[0] = make_float4(idf, (float)(ct.stick != 0), ct.r_a[0], ct.r_a[1]);
                cd[1] = make_float4(ct.r_a[2], ct.r_b[0], ct.r_b[1], ct.r_b[2]);
                cd[2] = f4(ct.normal, ct.penetration);
            }
        }
        ManifoldBuf& m0 = c->mb[0];
        if ((rc = upload_vec(c, m0.a, a)) || (rc = upload_vec(c, m0.b, b)) || (rc = upload_vec(c, m0.n, n)) ||
            (rc = upload_vec(c, m0.serial, ser)) || (rc = upload_vec(c, m0.fric, fr)) || (rc = upload_vec(c, m0.hot, hot)) ||
            (rc = upload_vec(c, m0.cold, cold)))
            return rc;
    }
    if ((rc = upload_energies(c, N, in->tets, T, in->tris, TR))) return rc;
    {
        std::vector<int> g(G_WORDS, 0);
        g[G_MANIFOLDS] = M;
        if ((rc = upload_vec(c, w.g, g))) return rc;
        std::vector<long long> ns(1, max_serial + 1);
        if ((rc = upload_vec(c, w.next_serial, ns))) return rc;
    }
    if ((rc = upload_tiles(c, off))) return rc;
    if (c->n_tiles != dims.n_tiles || c->n_items != dims.n_items) return fail(AVBD_E_ARG, "broad-phase tile count mismatch");
    c->have_world = true;
    if ((rc = build_incidence(c))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipGetLastError());
    return 0;
}

int check_errors(avbd_ctx* c, const int* g) {
    (void)c;
    if (g[G_ERR_MANIFOLD_CAP]) return fail(AVBD_E_CAPACITY, "manifold capacity exceeded (raise manifold_capacity)");
    if (g[G_ERR_PAIR_CAP]) return fail(AVBD_E_CAPACITY, "new-pair buffer exceeded in the broad phase");
    if (g[G_ERR_COLOURS]) return fail(AVBD_E_CAPACITY, "more than 256 colours needed");
    return 0;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    return ms;
}

// Greedy colouring of the dirty running scenes and the sweep bucket lists;
// reads back the status words (g) and the bucket offsets (c->bucket_off).
int colour_and_buckets(avbd_ctx* c, int* g) {
    DevWorld& w = c->w;
    hipStream_t st = c->stream;
    const int N = w.N;
    // scenes up to COLOUR_LDS_MAX bodies with LDS words, larger ones without
    const size_t lds = sizeof(unsigned int) * (size_t)std::min(c->max_scene_bodies, COLOUR_LDS_MAX);
    const int full = c->colour_updates ? 0 : 1;
    hipLaunchKernelGGL(k_colour<true>, dim3(w.S), dim3(1024), lds, st, w, full);
    if (c->max_scene_bodies > COLOUR_LDS_MAX)
        hipLaunchKernelGGL(k_colour<false>, dim3(w.S), dim3(1024), 0, st, w, full);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, st, w.col_cnt, AVBD_BUCKETS + 1, 0);
    hipLaunchKernelGGL(k_colour_hist, dim3(grid_for(N)), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_colour_scan, dim3(1), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_colour_fill, dim3(grid_for(N)), dim3(256), 0, st, w);
    HIP_TRY(hipMemcpyAsync(g, w.g, G_WORDS * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(c->bucket_off.data(), w.col_off, (AVBD_BUCKETS + 1) * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    return check_errors(c, g);
}

// Post-stabilisation sweep (solver.cpp:536-597) of the scenes marked in
// `only` (the others are made inactive meanwhile), with their colouring
// recomputed first.
int poststab_scenes(avbd_ctx* c, const std::vector<char>& only) {
    DevWorld& w = c->w;
    const int S = w.S;
    std::vector<int> st;
    int rc;
    if ((rc = download_vec(c, st, w.st, (size_t)S * ST_WORDS))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<int> active(S);
    bool any_post = false;
    for (int s = 0; s < S; ++s) {
        active[s] = st[s * ST_WORDS + ST_ACTIVE];
        st[s * ST_WORDS + ST_ACTIVE] = only[s] ? active[s] : 0;
        any_post = any_post || (only[s] && c->scenes[s].post_stabilization);
    }
    if ((rc = upload_vec(c, w.st, st))) return rc;
    int g[G_WORDS];
    if ((rc = colour_and_buckets(c, g))) return rc;
    if (any_post) {
        for (int col = 0; col < g[G_NCOLOURS]; ++col)
            if ((rc = launch_colour(c, c->bucket_off.data(), col, -1, 1))) return rc;
        hipLaunchKernelGGL(k_stick, dim3(grid_for((long long)w.Mcap * 8)), dim3(256), 0, c->stream, w);
        launch_scene_step(c, c->stream, 1);
    }
    // restore the active set
    if ((rc = download_vec(c, st, w.st, (size_t)S * ST_WORDS))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int s = 0; s < S; ++s) st[s * ST_WORDS + ST_ACTIVE] = active[s];
    if ((rc = upload_vec(c, w.st, st))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Solver::RemoveBody for lists of bodies: explicit_lists[s] holds, for each
// scene marked in aff, its bodies to remove (device indices) in removal
// order; without explicit lists, the bodies of the scenes marked in aff that
// are outside their scene's bounds, in index order.  stw: the scenes' status
// words (as downloaded; uploaded again here with the removal counts).  Each
// removal swaps the body with the scene's last live body
// (solver.cpp:104-116); its forces and energies go with it (~Mesh,
// geometry.cpp:11-24).  The device state is renumbered in place, so a
// step's own state (inertial targets, poses at step start, contact rows)
// carries into a post-stabilisation that follows.
int remove_bodies(avbd_ctx* c, const std::vector<char>& aff, std::vector<int>& stw,
                  const std::vector<std::vector<int>>* explicit_lists) {
    DevWorld& w = c->w;
    hipStream_t st = c->stream;
    const int S = w.S;
    int rc;
    c->drop_graphs();
    std::vector<int> aff_ids;
    for (int s = 0; s < S; ++s)
        if (aff[s]) aff_ids.push_back(s);
    if (aff_ids.empty()) return 0;
    // Only the affected scenes' body ranges are read back and renumbered;
    // every other index stays where it is.
    struct Seg {
        int s, a, e, n_live;
        std::vector<float4> pos;
        std::vector<uint32_t> flags;
        std::vector<int> perm;  // new local -> old global
    };
    std::vector<Seg> segs(aff_ids.size());
    for (size_t q = 0; q < aff_ids.size(); ++q) {
        Seg& g = segs[q];
        g.s = aff_ids[q];
        g.a = c->sc_off[g.s];
        g.e = c->sc_off[g.s + 1];
        if ((rc = download_vec(c, g.pos, w.pos + g.a, (size_t)(g.e - g.a))) ||
            (rc = download_vec(c, g.flags, w.flags + g.a, (size_t)(g.e - g.a))))
            return rc;
    }
    HIP_TRY(hipStreamSynchronize(st));
    // New order inside each affected scene: its live bodies (earlier dead
    // slots are at the end of its index range), the removals applied one by
    // one as swap with last, then the removed bodies and the earlier dead
    // slots as the scene's trailing dead slots.  aoff[s]: where scene s's
    // part of lmap starts (-1: not affected); lmap: old local -> new global
    // index, -1 for a removed body.  mv_dst[k] <- mv_src[k] moves the bodies.
    std::vector<int> aoff(S, -1), lmap, mv_src, mv_dst, dead_new;
    for (Seg& g : segs) {
        const int a = g.a, e = g.e;
        aoff[g.s] = (int)lmap.size();
        const size_t L0 = lmap.size();
        lmap.resize(L0 + (e - a), -1);
        std::vector<int> order, where(e - a, -1), gone;
        for (int i = a; i < e; ++i)
            if (!c->body_dead[i]) {
                where[i - a] = (int)order.size();
                order.push_back(i);
            }
        const avbd_scene& sc = c->scenes[g.s];
        auto remove = [&](int i) {
            const int k = where[i - a], last = (int)order.size() - 1;
            c->removed_scene.push_back(g.s);
            c->removed_index.push_back(k);
            if (k != last) {
                order[k] = order[last];
                where[order[k] - a] = k;
            }
            order.pop_back();
            where[i - a] = -1;
            gone.push_back(i);
        };
        if (explicit_lists) {
            for (int i : (*explicit_lists)[g.s]) remove(i);
        } else {
            for (int i = a; i < e; ++i) {
                if (c->body_dead[i] || (g.flags[i - a] & AVBD_BODY_STATIC)) continue;
                const float4 p = g.pos[i - a];
                if (p.x < sc.bounds_min[0] || p.x > sc.bounds_max[0] || p.y < sc.bounds_min[1] ||
                    p.y > sc.bounds_max[1] || p.z < sc.bounds_min[2] || p.z > sc.bounds_max[2])
                    remove(i);
            }
        }
        if (!explicit_lists) stw[g.s * ST_WORDS + ST_REMOVED] += (int)gone.size();  // bounds removals only
        stw[g.s * ST_WORDS + ST_DIRTY] = DIRTY_FULL;
        g.n_live = (int)order.size();
        int k = a;
        auto place = [&](int o, bool live) {
            g.perm.push_back(o);
            mv_dst.push_back(k);
            mv_src.push_back(o);
            lmap[L0 + (o - a)] = live ? k : -1;
            ++k;
        };
        for (int o : order) place(o, true);
        for (int o : gone) {
            dead_new.push_back(k);
            place(o, false);
        }
        for (int i = a; i < e; ++i)
            if (c->body_dead[i]) place(i, false);
    }
    // old global index -> new global index (-1 removed) inside a segment
    auto new_index = [&](const Seg& g, int o) -> int {
        return o >= g.a && o < g.e ? lmap[aoff[g.s] + (o - g.a)] : -1;
    };
    // energy lists of the affected scenes: each is the contiguous range
    // [e_off[a], e_off[e]) of the incidence arrays (an element's vertices
    // share its scene).  Rebuilt on the host in the new vertex order without
    // the elements that lost a vertex; the trailing dead slots get empty
    // lists (the last one keeps the rest of the range, never read).
    struct SegData {
        std::vector<int> off;
        std::vector<uint32_t> ent;
        std::vector<float4> ea, eb, ec;
        std::vector<int4> ev;
    };
    std::vector<SegData> sd(segs.size());
    for (size_t q = 0; q < segs.size(); ++q) {
        const Seg& g = segs[q];
        if ((rc = download_vec(c, sd[q].off, w.e_off + g.a, (size_t)(g.e - g.a + 1)))) return rc;
    }
    HIP_TRY(hipStreamSynchronize(st));
    for (size_t q = 0; q < segs.size(); ++q) {
        SegData& d = sd[q];
        const int E0 = d.off[0], E1 = d.off[segs[q].e - segs[q].a];
        if ((rc = download_vec(c, d.ent, w.e_ent + E0, (size_t)(E1 - E0))) ||
            (rc = download_vec(c, d.ea, w.ev_a + E0, (size_t)(E1 - E0))) ||
            (rc = download_vec(c, d.eb, w.ev_b + E0, (size_t)(E1 - E0))) ||
            (rc = download_vec(c, d.ec, w.ev_c + E0, (size_t)(E1 - E0))) ||
            (rc = download_vec(c, d.ev, w.ev_v + E0, (size_t)(E1 - E0))))
            return rc;
    }
    HIP_TRY(hipStreamSynchronize(st));
    for (size_t q = 0; q < segs.size(); ++q) {
        const Seg& g = segs[q];
        SegData& d = sd[q];
        const int E0 = d.off[0], E1 = d.off[g.e - g.a];
        std::vector<int> noff(g.e - g.a, 0);
        std::vector<uint32_t> nent;
        std::vector<float4> na, nb, nc;
        std::vector<int4> nv;
        for (int k = 0; k < g.e - g.a; ++k) {
            noff[k] = E0 + (int)nent.size();
            if (k >= g.n_live) continue;
            const int o = g.perm[k] - g.a;  // old local index
            for (int x = d.off[o] - E0; x < d.off[o + 1] - E0; ++x) {
                const uint32_t ent = d.ent[x];
                const int4 v = d.ev[x];
                const bool tri = (ent >> 31) != 0;
                const int m0 = new_index(g, v.x), m1 = new_index(g, v.y), m2 = new_index(g, v.z);
                const int m3 = tri ? m0 : new_index(g, v.w);  // slot 3 of a triangle repeats slot 0
                if (m0 < 0 || m1 < 0 || m2 < 0 || m3 < 0) {
                    const uint32_t idx = ent & 0x1FFFFFFFu;
                    if (tri) c->tri_dead[idx] = 1;
                    else c->tet_dead[idx] = 1;
                    continue;
                }
                nent.push_back(ent);
                na.push_back(d.ea[x]);
                nb.push_back(d.eb[x]);
                nc.push_back(d.ec[x]);
                nv.push_back(make_int4(m0, m1, m2, m3));
            }
        }
        if (E0 + (int)nent.size() > E1) return fail(AVBD_E_STATE, "energy list grew during removal");
        if ((rc = upload_vec(c, const_cast<int*>(w.e_off) + g.a, noff)) ||
            (rc = upload_vec(c, const_cast<uint32_t*>(w.e_ent) + E0, nent)) || (rc = upload_vec(c, w.ev_a + E0, na)) ||
            (rc = upload_vec(c, w.ev_b + E0, nb)) || (rc = upload_vec(c, w.ev_c + E0, nc)) ||
            (rc = upload_vec(c, w.ev_v + E0, nv)))
            return rc;
        // the segment's neighbour lists from its rebuilt streams (new indices)
        hipLaunchKernelGGL(k_nb_build, dim3(grid_for(g.e - g.a)), dim3(256), 0, st, w, g.a, g.e);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(st));
    }
    // device renumbering of the moved bodies, then of the forces' and
    // elements' body ids and the manifolds' endpoints
    const int K = (int)mv_src.size();
    struct Scratch {  // freed on every return path
        void* p = nullptr;
        ~Scratch() { if (p) hipFree(p); }
    } scratch;
    const size_t ints = 3 * (size_t)K + (size_t)S + (size_t)dead_new.size() + 4;
    HIP_TRY(hipMalloc(&scratch.p, sizeof(float4) * (size_t)K + sizeof(int) * ints));
    float4* d_tmp = static_cast<float4*>(scratch.p);
    int* d_src = reinterpret_cast<int*>(d_tmp + K);
    int* d_dst = d_src + K;
    int* d_lmap = d_dst + K;
    int* d_aoff = d_lmap + K;
    int* d_dead = d_aoff + S;
    if ((rc = upload_vec(c, d_src, mv_src)) || (rc = upload_vec(c, d_dst, mv_dst)) || (rc = upload_vec(c, d_lmap, lmap)) ||
        (rc = upload_vec(c, d_aoff, aoff)) || (rc = upload_vec(c, d_dead, dead_new)))
        return rc;
    {
        float4* f4[] = {w.pos, w.rot, w.scale, w.vel, w.angvel, w.pvel, w.pangvel, w.inertia,
                        w.last_pos, w.last_rot, w.inert_pos, w.inert_rot, w.diff0, w.diff1, w.drag};
        for (float4* arr : f4) {
            hipLaunchKernelGGL(k_gather_f4, dim3(grid_for(K)), dim3(256), 0, st, d_tmp, arr, d_src, K);
            hipLaunchKernelGGL(k_scatter_f4, dim3(grid_for(K)), dim3(256), 0, st, arr, d_tmp, d_dst, K);
        }
        if (c->dbg.rgb) {
            hipLaunchKernelGGL(k_gather_f4, dim3(grid_for(K)), dim3(256), 0, st, d_tmp, c->dbg.rgb, d_src, K);
            hipLaunchKernelGGL(k_scatter_f4, dim3(grid_for(K)), dim3(256), 0, st, c->dbg.rgb, d_tmp, d_dst, K);
        }
        uint32_t* u32[] = {w.flags, reinterpret_cast<uint32_t*>(w.scene)};
        uint32_t* t32 = reinterpret_cast<uint32_t*>(d_tmp);
        for (uint32_t* arr : u32) {
            hipLaunchKernelGGL(k_gather_u32, dim3(grid_for(K)), dim3(256), 0, st, t32, arr, d_src, K);
            hipLaunchKernelGGL(k_scatter_u32, dim3(grid_for(K)), dim3(256), 0, st, arr, t32, d_dst, K);
        }
    }
    if (!dead_new.empty())
        // the new dead slots: static, no broad-phase sphere, at rest
        hipLaunchKernelGGL(k_kill_bodies, dim3(grid_for((int)dead_new.size())), dim3(256), 0, st, w, (const int*)d_dead,
                           (int)dead_new.size());
    const BodyRemap rm{d_aoff, d_lmap};
    if (w.J + w.SP > 0) hipLaunchKernelGGL(k_remap_forces, dim3(grid_for(w.J + w.SP)), dim3(256), 0, st, w, rm);
    if (w.T + w.TR > 0) hipLaunchKernelGGL(k_remap_elements, dim3(grid_for(w.T + w.TR)), dim3(256), 0, st, w, rm);
    hipLaunchKernelGGL(k_remap_manifolds, dim3(grid_for(w.Mcap)), dim3(256), 0, st, w, rm);
    hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, st, w.alive + w.Mcap, 0);
    if ((rc = scan_ints(c, w.alive, w.alive_off, w.Mcap + 1))) return rc;
    ManifoldBuf& dst = c->mb[c->cur ^ 1];
    hipLaunchKernelGGL(k_compact, dim3(grid_for(w.Mcap)), dim3(256), 0, st, w, dst);
    hipLaunchKernelGGL(k_compact_finish, dim3(1), dim3(1), 0, st, w);
    c->cur ^= 1;
    w.m = c->mb[c->cur];
    for (const Seg& g : segs)
        for (int i = g.a; i < g.e; ++i) c->body_dead[i] = (i - g.a) >= g.n_live;
    if ((rc = upload_vec(c, w.st, stw))) return rc;
    if ((rc = build_incidence(c))) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

// Out-of-bounds removal (solver.cpp:520-534) for the scenes k_velocity
// flagged, then their post-stabilisation, which the step skipped for them.
// A scene whose explosion check fired removes nothing (the reference returns
// before, solver.cpp:528-529).
int remove_out_of_bounds(avbd_ctx* c) {
    DevWorld& w = c->w;
    hipStream_t st = c->stream;
    const int S = w.S;
    std::vector<int> stw;
    int rc;
    if ((rc = download_vec(c, stw, w.st, (size_t)S * ST_WORDS))) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<char> aff(S, 0);
    bool any = false;
    for (int s = 0; s < S; ++s) {
        int* sw = &stw[s * ST_WORDS];
        if (sw[ST_OOB] > 0 && !sw[ST_EMERGENCY] && sw[ST_ACTIVE]) {
            aff[s] = 1;
            any = true;
        }
        sw[ST_OOB] = 0;
    }
    hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, st, w.g + G_ANY_OOB, 0);
    if (!any) {
        if ((rc = upload_vec(c, w.st, stw))) return rc;
        HIP_TRY(hipStreamSynchronize(st));
        return 0;
    }
    if ((rc = remove_bodies(c, aff, stw, nullptr))) return rc;
    return poststab_scenes(c, aff);
}

// Phase 5 of a step for the bucket offsets off: the iterations, per colour
// the primal buckets, then the dual and the fracture reset.
int launch_iterations(avbd_ctx* c, const int* off, int ncol, int max_iters, int dgrid) {
    DevWorld& w = c->w;
    hipStream_t st = c->stream;
    for (int it = 0; it < max_iters; ++it) {
        for (int col = 0; col < ncol; ++col) launch_colour(c, off, col, it, 0);
        // step 5.2 (solver.cpp:451-455): the sweep's trust-region ratio
        if (c->any_trust) hipLaunchKernelGGL(k_rho_update, dim3(grid_for(w.S)), dim3(256), 0, st, w);
        const int mark = c->mark_begin(avbd_ctx::K_DUAL);
        hipLaunchKernelGGL(k_dual, dim3(dgrid), dim3(256), 0, st, w, it);
        c->mark_end(mark);
        if (w.J + w.SP > 0) {
            hipLaunchKernelGGL(k_fracture_reset, dim3(grid_for(2 * (w.J + w.SP) + 2)), dim3(256), 0, st, w);
            hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, st, w.g + G_FRACTURES, 0);
        }
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// Phase 7: the post-stabilisation sweep and the stick flags it leaves.
int launch_poststab(avbd_ctx* c, const int* off, int ncol) {
    DevWorld& w = c->w;
    for (int col = 0; col < ncol; ++col) launch_colour(c, off, col, -1, 1);
    hipLaunchKernelGGL(k_stick, dim3(grid_for((long long)w.Mcap * 8)), dim3(256), 0, c->stream, w);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Captures the launches `body` makes on the context's stream into an
// executable graph.
template <class F>
int capture(avbd_ctx* c, hipGraphExec_t* exec, F body) {
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(AVBD_E_HIP, std::string("step graph capture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        *exec = nullptr;
        return fail(AVBD_E_HIP, std::string("step graph instantiate: ") + hipGetErrorString(ei));
    }
    return 0;
}

// Phases 4-7 of a step: prediction, the iterations, velocities and
// post-stabilisation, with the phase events ev[3] (after the prediction),
// ev[4] (after the iterations), ev[5] (after the velocities) recorded on
// the host side of the launches.  sg: the step graph of this launch shape,
// whose captured iterations and post-stabilisation replace the launches of
// phases 5 and 7 (prediction and velocities are one kernel each); null =
// everything eagerly.
int sweep_phases(avbd_ctx* c, const int* off, int ncol, int max_iters, bool any_post, int dgrid,
                 avbd_ctx::StepGraph* sg) {
    DevWorld& w = c->w;
    hipStream_t st = c->stream;
    const int N = w.N;
    // the small scenes' whole phases 4-7 on their own stream, beside the buckets
    const bool fused = !c->fusable_ids.empty();
    if (fused && (int)c->fusable_ids.size() == w.S) {
        // every scene in the per-scene kernel: the bucket, prediction and
        // velocity launches would find nothing to do (the phases are not
        // timed apart: all of it counts as the iterations)
        HIP_TRY(hipEventRecord(c->ev[3], st));
        launch_scene_step(c, st, 0);
        HIP_TRY(hipEventRecord(c->ev[4], st));
        HIP_TRY(hipEventRecord(c->ev[5], st));
        HIP_TRY(hipGetLastError());
        return 0;
    }
    if (fused) {
        HIP_TRY(hipEventRecord(c->fork2, st));
        HIP_TRY(hipStreamWaitEvent(c->side2, c->fork2, 0));
        launch_scene_step(c, c->side2, 0);
    }
    hipLaunchKernelGGL(k_predict, dim3(grid_for(N)), dim3(256), 0, st, w);
    HIP_TRY(hipEventRecord(c->ev[3], st));
    int rc;
    if (sg) {
        HIP_TRY(hipGraphLaunch(sg->iters, st));
    } else if ((rc = launch_iterations(c, off, ncol, max_iters, dgrid))) {
        return rc;
    }
    HIP_TRY(hipEventRecord(c->ev[4], st));
    hipLaunchKernelGGL(k_velocity, dim3(grid_for(N)), dim3(256), 0, st, w);
    HIP_TRY(hipEventRecord(c->ev[5], st));
    if (any_post) {
        if (sg) {
            HIP_TRY(hipGraphLaunch(sg->post, st));
        } else if ((rc = launch_poststab(c, off, ncol))) {
            return rc;
        }
    }
    if (fused) {
        HIP_TRY(hipEventRecord(c->join2, c->side2));
        HIP_TRY(hipStreamWaitEvent(st, c->join2, 0));
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// Launch shape of phases 4-7: everything a captured graph fixes (manifold
// buffer, colours, iterations, per bucket the lane width and grid, the dual
// grid).  False while a bucket's lane width is still being tuned.
bool step_graph_key(const avbd_ctx* c, const int* off, int ncol, int max_iters, bool any_post, int dgrid,
                    std::vector<int>* key) {
    key->assign({c->cur, ncol, max_iters, any_post ? 1 : 0, dgrid, fused_lanes(c), (int)c->fusable_ids.size(),
                 c->fem_canon ? 1 : 0});
    for (int k = 0; k < 2 * ncol; ++k) {
        const int n = off[k + 1] - off[k];
        if (n == 0) {
            key->push_back(0);
            continue;
        }
        int block = 0;
        const int G = bucket_lanes(c, (k & 1) != 0, n, &block);
        if (G == 0) return false;
        key->push_back(G * 65536 + bucket_grid(c, n, G, block));
        key->push_back(block);
    }
    return true;
}

int do_step(avbd_ctx* c) {
    DevWorld& w = c->w;
    hipStream_t st = c->stream;
    const int N = w.N;
    int max_iters = 0;
    bool any_post = false;
    for (const avbd_scene& s : c->scenes) {
        max_iters = std::max(max_iters, s.iterations);
        any_post = any_post || s.post_stabilization != 0;
    }
    int rc = sync_fused(c);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(c->ev[0], st));
    // ---- 1. broad phase
    hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, st, w.g + G_NEW_PAIRS, 0);
    hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, st, w.g + G_CANDIDATES, 0);
    // With a subset of the scenes running, only their tiles and work items
    // are launched (the others' would exit at once, but a world of
    // thousands of scenes has hundreds of thousands of them).
    int n_tiles = c->n_tiles, n_items = c->n_items;
    const int* tiles = nullptr;
    const int4* items = c->items;
    if (c->subset && n_items > 0) {
        n_tiles = n_items = 0;
        for (int s : c->run_ids) {
            for (int t = c->h_tile_off[s]; t < c->h_tile_off[s + 1]; ++t) c->pin_tiles[n_tiles++] = t;
            const int i0 = c->h_item_off[s], i1 = c->h_item_off[s + 1];
            std::copy(c->h_items.begin() + i0, c->h_items.begin() + i1, c->pin_items + n_items);
            n_items += i1 - i0;
        }
        if (n_items > 0) {
            HIP_TRY(hipMemcpyAsync(c->sub_tiles, c->pin_tiles, n_tiles * sizeof(int), hipMemcpyHostToDevice, st));
            HIP_TRY(hipMemcpyAsync(c->sub_items, c->pin_items, n_items * sizeof(int4), hipMemcpyHostToDevice, st));
        }
        tiles = c->sub_tiles;
        items = c->sub_items;
    }
    if (n_items > 0) {
        hipLaunchKernelGGL(k_tile_bounds, dim3(n_tiles), dim3(256), 0, st, w, tiles);
        hipLaunchKernelGGL(k_broad, dim3(n_items), dim3(256), 0, st, w, items);
        hipLaunchKernelGGL(k_pair_filter, dim3(grid_for(N)), dim3(256), 0, st, w);
        if (c->small_pairs) {
            hipLaunchKernelGGL(k_pair_sort_small, dim3(1), dim3(1024), 0, st, w);
            hipLaunchKernelGGL(k_pair_append, dim3(grid_for(SMALL_PAIR_SORT)), dim3(256), 0, st, w,
                               (const unsigned long long*)w.pairs);
        } else {
            // the new-pair count sizes the sort
            HIP_TRY(hipMemcpyAsync(c->host_word, w.g + G_NEW_PAIRS, sizeof(int), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            const int K = std::min(*c->host_word, w.pair_cap);
            if (K > 0) {
                hipcub::DoubleBuffer<unsigned long long> kb(reinterpret_cast<unsigned long long*>(w.pairs),
                                                             reinterpret_cast<unsigned long long*>(w.pairs_tmp));
                size_t bytes = c->cub_bytes;
                HIP_TRY(hipcub::DeviceRadixSort::SortKeys(c->cub_tmp, bytes, kb, K, 0, pair_key_bits(N), st));
                hipLaunchKernelGGL(k_pair_append, dim3(grid_for(K)), dim3(256), 0, st, w,
                                   (const unsigned long long*)kb.Current());
            }
        }
    }
    hipLaunchKernelGGL(k_pair_finish, dim3(1), dim3(1), 0, st, w);
    HIP_TRY(hipEventRecord(c->ev[1], st));
    // ---- 2. force warm start, compaction, incidence, colouring
    hipLaunchKernelGGL(k_manifold_init, dim3(grid_for(w.Mcap, 128, 8192)), dim3(128), 0, st, w);
    if (w.J + w.SP > 0) hipLaunchKernelGGL(k_force_init, dim3(grid_for(w.J + w.SP)), dim3(256), 0, st, w);
    hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, st, w.alive + w.Mcap, 0);
    if ((rc = scan_ints(c, w.alive, w.alive_off, w.Mcap + 1))) return rc;
    ManifoldBuf& dst = c->mb[c->cur ^ 1];
    hipLaunchKernelGGL(k_compact, dim3(grid_for(w.Mcap)), dim3(256), 0, st, w, dst);
    hipLaunchKernelGGL(k_compact_finish, dim3(1), dim3(1), 0, st, w);
    c->cur ^= 1;
    w.m = c->mb[c->cur];
    if ((rc = build_incidence(c))) return rc;
    int g[G_WORDS];
    if ((rc = colour_and_buckets(c, g))) return rc;
    int* off = c->bucket_off.data();
    const int ncol = g[G_NCOLOURS];
    c->last_ncol = ncol;
    HIP_TRY(hipEventRecord(c->ev[2], st));
    int dgrid = grid_for((long long)g[G_MANIFOLDS] * 8 + w.J + w.SP, 256, 8192);
    {
        int p = 1;
        while (p < dgrid) p <<= 1;
        dgrid = std::min(p, 8192);
    }
    // ---- 4.-7. with the iterations and post-stabilisation as step graphs
    // when this launch shape has been seen before
    std::vector<int> key;
    avbd_ctx::StepGraph* use = nullptr;
    if (c->use_graphs && !c->ktiming && step_graph_key(c, off, ncol, max_iters, any_post, dgrid, &key)) {
        avbd_ctx::StepGraph* hit = nullptr;
        for (avbd_ctx::StepGraph& sg : c->graphs)
            if (sg.key == key) hit = &sg;
        if (hit && !hit->captured) {
            rc = capture(c, &hit->iters, [&] { return launch_iterations(c, off, ncol, max_iters, dgrid); });
            if (!rc && any_post) rc = capture(c, &hit->post, [&] { return launch_poststab(c, off, ncol); });
            if (rc) {
                avbd_ctx::destroy(*hit);
                return rc;
            }
            hit->captured = true;
            ++c->graph_captures;
        }
        if (hit) {
            hit->used = ++c->graph_clock;
            use = hit;
            ++c->graph_launches;
        } else {
            // first sighting: remember the shape, run eagerly
            if (c->graphs.size() >= avbd_ctx::kMaxGraphs) {
                auto lru = std::min_element(c->graphs.begin(), c->graphs.end(),
                                            [](const avbd_ctx::StepGraph& a, const avbd_ctx::StepGraph& b) {
                                                return a.used < b.used;
                                            });
                avbd_ctx::destroy(*lru);
                c->graphs.erase(lru);
            }
            avbd_ctx::StepGraph sg;
            sg.key = key;
            sg.used = ++c->graph_clock;
            c->graphs.push_back(std::move(sg));
        }
    }
    if ((rc = sweep_phases(c, off, ncol, max_iters, any_post, dgrid, use))) return rc;
    HIP_TRY(hipEventRecord(c->ev[6], st));
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

// Device index -> index among the live bodies (dead slots skipped), -1 for
// a dead slot; negative ids (the world anchor) pass through.
static std::vector<int> live_index(const avbd_ctx* c) {
    std::vector<int> idx(c->w.N, -1);
    int k = 0;
    for (int i = 0; i < c->w.N; ++i)
        if (!c->body_dead[i]) idx[i] = k++;
    return idx;
}

int avbd_set_error(int code, const std::string& msg) { return fail(code, msg); }

WorldCounts world_counts(const avbd_world* w) {
    WorldCounts k;
    k.S = w->n_scenes; k.N = w->n_bodies; k.J = w->n_joints; k.SP = w->n_springs;
    k.M = w->n_manifolds; k.T = w->n_tets; k.TR = w->n_tris;
    long long run = 0;
    for (long long i = 0; i < k.N; ++i) {
        if (!(w->bodies[i].flags & AVBD_BODY_PARTICLE)) ++k.nonpart;
        ++run;
        if (i + 1 == k.N || w->bodies[i + 1].scene != w->bodies[i].scene) {
            k.tiles += broad_tiles(run);
            k.items += broad_items(run);
            run = 0;
        }
    }
    return k;
}

long long world_device_bytes(const WorldCounts& k, long long manifold_capacity) {
    WorldDims d;
    int rc = make_dims(k, manifold_capacity, &d);
    if (rc) return rc;
    avbd_ctx dry;
    dry.dry = true;
    avbd_scene* sc;
    int *off, *eoff;
    uint32_t* eent;
    rc = allocate_world(&dry, d, &sc, &off, &eoff, &eent);
    if (rc) return rc;
    return (long long)dry.alloc_bytes;
}

extern "C" {

int avbd_abi_version(void) { return AVBD_ABI_VERSION; }

int avbd_footprint(const avbd_world* w, int64_t manifold_capacity, int64_t* bytes) {
    if (!w || !bytes) return fail(AVBD_E_ARG, "null argument");
    if (w->n_scenes < 1 || w->n_bodies < 0 || w->n_joints < 0 || w->n_springs < 0 || w->n_manifolds < 0 ||
        w->n_tets < 0 || w->n_tris < 0)
        return fail(AVBD_E_ARG, "bad counts");
    if (w->n_bodies > 0 && !w->bodies) return fail(AVBD_E_ARG, "null body array");
    long long b = world_device_bytes(world_counts(w), manifold_capacity);
    if (b < 0) return (int)b;
    *bytes = b;
    return 0;
}

int avbd_get_memory(avbd_ctx* c, int64_t* bytes) {
    if (!c || !bytes) return fail(AVBD_E_ARG, "null argument");
    *bytes = (int64_t)c->alloc_bytes;
    return 0;
}

int avbd_record_sizes(int32_t* out, int32_t n) {
    const int32_t sizes[] = {(int32_t)sizeof(avbd_scene),   (int32_t)sizeof(avbd_body),     (int32_t)sizeof(avbd_joint),
                             (int32_t)sizeof(avbd_spring),  (int32_t)sizeof(avbd_contact),  (int32_t)sizeof(avbd_manifold),
                             (int32_t)sizeof(avbd_tet),     (int32_t)sizeof(avbd_tri),      (int32_t)sizeof(avbd_world),
                             (int32_t)sizeof(avbd_timings), (int32_t)sizeof(avbd_request_info),
                             (int32_t)sizeof(avbd_server_stats)};
    int k = (int)(sizeof(sizes) / sizeof(sizes[0]));
    for (int i = 0; i < n && i < k; ++i) out[i] = sizes[i];
    return k;
}

const char* avbd_last_error(void) { return g_last_error.c_str(); }

int avbd_create(int32_t device, avbd_ctx** out) {
    if (!out) return fail(AVBD_E_ARG, "null output pointer");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(AVBD_E_NOGPU, "no HIP device");
    if (device < 0 || device >= count) return fail(AVBD_E_ARG, "device index out of range");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
        return fail(AVBD_E_NOGPU, std::string("library built for gfx950, device is ") + prop.gcnArchName);
    HIP_TRY(hipSetDevice(device));
    if (int rc = noalloc_scratch_init(device)) return rc;
    avbd_ctx* c = new avbd_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(AVBD_E_HIP, "hipStreamCreate failed");
    }
    for (auto& e : c->ev) hipEventCreate(&e);
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->run_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork2, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join2, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&c->host_word, sizeof(int), hipHostMallocDefault) != hipSuccess) {
        avbd_destroy(c);
        return fail(AVBD_E_HIP, "stream/event/pinned buffer creation failed");
    }
    // the LDS colouring takes up to 128 KiB of dynamic LDS per workgroup
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(k_colour<true>), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(unsigned int) * COLOUR_LDS_MAX)) != hipSuccess) {
        avbd_destroy(c);
        return fail(AVBD_E_HIP, "cannot reserve LDS for the colouring kernel");
    }
    *out = c;
    return 0;
}

int avbd_destroy(avbd_ctx* c) {
    if (!c) return 0;
    hipSetDevice(c->device);
    c->release();
    for (auto& e : c->ev)
        if (e) hipEventDestroy(e);
    for (auto& e : c->kev) hipEventDestroy(e);
    if (c->side) hipStreamSynchronize(c->side);
    if (c->side2) hipStreamSynchronize(c->side2);
    if (c->fork) hipEventDestroy(c->fork);
    if (c->join) hipEventDestroy(c->join);
    if (c->fork2) hipEventDestroy(c->fork2);
    if (c->join2) hipEventDestroy(c->join2);
    if (c->run_ev) hipEventDestroy(c->run_ev);
    if (c->side) hipStreamDestroy(c->side);
    if (c->side2) hipStreamDestroy(c->side2);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->host_word) hipHostFree(c->host_word);
    delete c;
    return 0;
}

int avbd_set_option(avbd_ctx* c, const char* name, int64_t value) {
    if (!c || !name) return fail(AVBD_E_ARG, "null argument");
    std::string n(name);
    if (n == "lanes" || n == "fem_lanes") {
        if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8 && value != 16 && value != 32)
            return fail(AVBD_E_ARG, n + " must be 0 (tuned), 1, 2, 4, 8, 16 or 32");
        if (n == "fem_lanes" && c->fem_canon && value > 0 && value < 4)
            return fail(AVBD_E_ARG, "fem_lanes must be 0 or 4 to 32 with fem_canonical_sums");
        (n == "lanes" ? c->lanes : c->fem_lanes) = int(value);
        return 0;
    }
    if (n == "fem_canonical_sums") {
        if (value != 0 && c->fem_lanes > 0 && c->fem_lanes < 4)
            return fail(AVBD_E_ARG, "fem_canonical_sums needs fem_lanes 0 or 4 to 32");
        c->fem_canon = value != 0;
        return 0;
    }
    if (n == "kernel_timing") {
        c->ktiming = value != 0;
        return 0;
    }
    if (n == "fused_scenes") {
        c->use_fused = value != 0;
        return 0;
    }
    if (n == "fused_single_scene") {
        c->fuse_single = value != 0;
        return 0;
    }
    if (n == "scene_lds_bytes") {
        if (value < 0 || value > 65536) return fail(AVBD_E_ARG, "scene_lds_bytes out of range (0 to 65536)");
        c->scene_lds = int(value);
        c->drop_graphs();  // captured per-scene launches carry the size
        return 0;
    }
    if (n == "fused_max_bodies") {
        if (value < 0 || value > (1 << 20)) return fail(AVBD_E_ARG, "fused_max_bodies out of range");
        c->fused_max = int(value);
        return 0;
    }
    if (n == "spare_bodies") {
        if (value < 0 || value > (1 << 16)) return fail(AVBD_E_ARG, "spare_bodies out of range");
        c->spare_bodies = int(value);
        return 0;
    }
    if (n == "debug_output") {
        c->dbg_on = value != 0;
        return 0;
    }
    if (n == "colour_updates") {
        c->colour_updates = value != 0;
        return 0;
    }
    if (n == "fem_neighbour_lists") {
        c->use_nb = value != 0;
        c->w.use_nb = c->use_nb ? 1 : 0;
        c->drop_graphs();  // captured launches carry the world by value
        return 0;
    }
    if (n == "step_graphs") {
        c->use_graphs = value != 0;
        if (!c->use_graphs) c->drop_graphs();
        return 0;
    }
    return fail(AVBD_E_ARG, "unknown option " + n);
}

int avbd_tune_table(avbd_ctx* c, char* out, int32_t len) {
    if (!c) return fail(AVBD_E_ARG, "null context");
    std::string s = c->tuner.to_json();
    if (out && len > 0) {
        std::strncpy(out, s.c_str(), (size_t)len - 1);
        out[len - 1] = 0;
    }
    return (int)s.size();
}

int avbd_tune_load(avbd_ctx* c, const char* json) {
    if (!c || !json) return fail(AVBD_E_ARG, "null argument");
    c->tuner.from_json(json);
    return 0;
}

int avbd_kernel_stats(avbd_ctx* c, const char* kernel, int64_t* launches, double* total_ms, int32_t reset) {
    if (!c || !kernel) return fail(AVBD_E_ARG, "null argument");
    std::string k(kernel);
    if (k == "step_graph") {  // replays of captured steps (launches), captures (total_ms field: count)
        if (launches) *launches = c->graph_launches;
        if (total_ms) *total_ms = double(c->graph_captures);
        if (reset) c->graph_launches = c->graph_captures = 0;
        return 0;
    }
    // "primal" = both bucket classes together
    int lo = k == "primal" ? avbd_ctx::K_FORCES
           : k == "dual" ? avbd_ctx::K_DUAL
           : k == "primal_forces" ? avbd_ctx::K_FORCES
           : k == "primal_fem" ? avbd_ctx::K_FEM : -1;
    if (lo < 0) return fail(AVBD_E_ARG, "unknown kernel class " + k);
    const int hi = k == "primal" ? avbd_ctx::K_FEM : lo;
    long long n = 0;
    double ms = 0.0;
    for (int kind = lo; kind <= hi; ++kind) {
        n += c->kstat_n[kind];
        ms += c->kstat_ms[kind];
        if (reset) {
            c->kstat_n[kind] = 0;
            c->kstat_ms[kind] = 0.0;
        }
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    return 0;
}

int avbd_upload(avbd_ctx* c, const avbd_world* w, int64_t manifold_capacity) {
    if (!c || !w) return fail(AVBD_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    const int k = c->spare_bodies;
    if (k == 0) return do_upload(c, w, manifold_capacity);
    // Option spare_bodies: every scene gets k dead slots after its bodies
    // (static, no broad-phase sphere), which avbd_add_bodies can fill.  The
    // records are expanded on the host and every body index shifted by k per
    // earlier scene.
    if (w->n_scenes < 1 || w->n_bodies < 0 || (w->n_bodies > 0 && !w->bodies)) return fail(AVBD_E_ARG, "bad world");
    if (w->n_joints < 0 || w->n_springs < 0 || w->n_manifolds < 0 || w->n_tets < 0 || w->n_tris < 0)
        return fail(AVBD_E_ARG, "negative count");
    if ((w->n_joints > 0 && !w->joints) || (w->n_springs > 0 && !w->springs) || (w->n_manifolds > 0 && !w->manifolds) ||
        (w->n_tets > 0 && !w->tets) || (w->n_tris > 0 && !w->tris))
        return fail(AVBD_E_ARG, "null record array");
    const int S = w->n_scenes, N = w->n_bodies;
    if ((long long)N + (long long)k * S >= INT32_MAX) return fail(AVBD_E_ARG, "too many bodies with spare slots");
    for (int i = 0; i < N; ++i)
        if (w->bodies[i].scene < 0 || w->bodies[i].scene >= S || (i > 0 && w->bodies[i].scene < w->bodies[i - 1].scene))
            return fail(AVBD_E_ARG, "bodies must be grouped by scene in ascending order");
    // Every index is checked against the caller's N before the expansion: an
    // index in [N, N + k S) would otherwise name a spare or shifted body and
    // pass the range checks of the expanded world.  -1 (the world anchor of
    // a joint or spring) is the only negative index kept.
    auto in_range = [&](int b, bool anchor_ok) { return (b >= 0 && b < N) || (anchor_ok && b == -1); };
    for (int q = 0; q < w->n_joints; ++q)
        if (!in_range(w->joints[q].body_a, true) || !in_range(w->joints[q].body_b, false))
            return fail(AVBD_E_ARG, "joint body index");
    for (int q = 0; q < w->n_springs; ++q)
        if (!in_range(w->springs[q].body_a, true) || !in_range(w->springs[q].body_b, false))
            return fail(AVBD_E_ARG, "spring body index");
    for (int q = 0; q < w->n_manifolds; ++q)
        if (!in_range(w->manifolds[q].body_a, false) || !in_range(w->manifolds[q].body_b, false))
            return fail(AVBD_E_ARG, "manifold body index");
    for (int q = 0; q < w->n_tets; ++q)
        for (int v = 0; v < 4; ++v)
            if (!in_range(w->tets[q].v[v], false)) return fail(AVBD_E_ARG, "tet vertex index");
    for (int q = 0; q < w->n_tris; ++q)
        for (int v = 0; v < 3; ++v)
            if (!in_range(w->tris[q].v[v], false)) return fail(AVBD_E_ARG, "triangle vertex index");
    auto shift = [&](int i) { return i < 0 ? i : i + k * w->bodies[i].scene; };
    std::vector<avbd_body> bodies;
    bodies.reserve((size_t)N + (size_t)k * S);
    std::vector<int> spare;
    int i = 0;
    for (int s = 0; s < S; ++s) {
        for (; i < N && w->bodies[i].scene == s; ++i) bodies.push_back(w->bodies[i]);
        for (int q = 0; q < k; ++q) {
            avbd_body d{};
            d.rot[0] = 1.f;
            d.scale[0] = d.scale[1] = d.scale[2] = 1.f;
            d.radius = -1e30f;
            d.flags = AVBD_BODY_STATIC;
            d.scene = s;
            spare.push_back((int)bodies.size());
            bodies.push_back(d);
        }
    }
    std::vector<avbd_joint> joints(w->joints, w->joints + w->n_joints);
    for (avbd_joint& j : joints) { j.body_a = shift(j.body_a); j.body_b = shift(j.body_b); }
    std::vector<avbd_spring> springs(w->springs, w->springs + w->n_springs);
    for (avbd_spring& sp : springs) { sp.body_a = shift(sp.body_a); sp.body_b = shift(sp.body_b); }
    std::vector<avbd_manifold> mans(w->manifolds, w->manifolds + w->n_manifolds);
    for (avbd_manifold& m : mans) { m.body_a = shift(m.body_a); m.body_b = shift(m.body_b); }
    std::vector<avbd_tet> tets(w->tets, w->tets + w->n_tets);
    for (avbd_tet& t : tets)
        for (int q = 0; q < 4; ++q) t.v[q] = shift(t.v[q]);
    std::vector<avbd_tri> tris(w->tris, w->tris + w->n_tris);
    for (avbd_tri& t : tris)
        for (int q = 0; q < 3; ++q) t.v[q] = shift(t.v[q]);
    avbd_world x = *w;
    x.n_bodies = (int32_t)bodies.size();
    x.bodies = bodies.data();
    x.joints = joints.data();
    x.springs = springs.data();
    x.manifolds = mans.data();
    x.tets = tets.data();
    x.tris = tris.data();
    const int rc = do_upload(c, &x, manifold_capacity);
    if (rc) return rc;
    for (int d : spare) c->body_dead[d] = 1;
    return 0;
}

// Debug output buffers for the current world (see DebugOut).
static int ensure_debug(avbd_ctx* c) {
    auto& d = c->dbg;
    if (d.ready) return 0;
    DevWorld& w = c->w;
    const size_t slots = (size_t)w.J + w.SP + w.Mcap + w.T + w.TR;
    // the line and point offsets are 32-bit scans over per-slot counts: a
    // joint or spring draws 1 line, a manifold up to 8 lines and 16 points,
    // a tet 6 lines, a triangle 3
    const size_t most_lines = (size_t)w.J + w.SP + 8 * (size_t)w.Mcap + 6 * (size_t)w.T + 3 * (size_t)w.TR;
    const size_t most_points = 16 * (size_t)w.Mcap;
    if (slots >= (size_t)INT32_MAX || most_lines >= (size_t)INT32_MAX || most_points >= (size_t)INT32_MAX)
        return fail(AVBD_E_CAPACITY, "world too large for debug output");
    int rc;
    if ((rc = c->alloc(&d.lc, slots + 1)) || (rc = c->alloc(&d.loff, slots + 1)) ||
        (rc = c->alloc(&d.pc, (size_t)w.Mcap + 1)) || (rc = c->alloc(&d.poff, (size_t)w.Mcap + 1)))
        return rc;
    for (ManifoldBuf& m : c->mb)
        if (!m.cpos && (rc = c->alloc(&m.cpos, (size_t)w.Mcap * 8))) return rc;
    w.m = c->mb[c->cur];
    size_t b1 = 0, b2 = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b1, d.lc, d.loff, (int)slots, c->stream));
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, d.pc, d.poff, w.Mcap, c->stream));
    d.cub_bytes = std::max<size_t>(std::max(b1, b2), 1);
    unsigned char* cub;
    if ((rc = c->alloc(&cub, d.cub_bytes))) return rc;
    d.cub = cub;
    d.ready = true;
    return 0;
}

// The step's line and point data (k_dbg_count / k_dbg_write), after the step
// and before its bounds removal.
static int extract_debug(avbd_ctx* c) {
    int rc;
    if ((rc = ensure_debug(c))) return rc;
    auto& d = c->dbg;
    DevWorld& w = c->w;
    hipStream_t st = c->stream;
    const int slots = w.J + w.SP + w.Mcap + w.T + w.TR;
    hipLaunchKernelGGL(k_dbg_count, dim3(grid_for(slots)), dim3(256), 0, st, w, d.lc, d.pc);
    size_t bytes = d.cub_bytes;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(d.cub, bytes, d.lc, d.loff, slots, st));
    bytes = d.cub_bytes;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(d.cub, bytes, d.pc, d.poff, w.Mcap, st));
    int tail[4];
    HIP_TRY(hipMemcpyAsync(&tail[0], d.loff + slots - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&tail[1], d.lc + slots - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&tail[2], d.poff + w.Mcap - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&tail[3], d.pc + w.Mcap - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    d.nl = (long long)tail[0] + tail[1];
    d.np = (long long)tail[2] + tail[3];
    // grown by a quarter beyond the need: the counts drift up step by step
    // as contacts appear, and a reallocation per step would dominate
    auto grow = [&](float*& buf, size_t& cap, size_t need) -> int {
        if (need <= cap) return 0;
        if (buf) HIP_TRY(hipFree(buf));
        buf = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(need + need / 4, 1024);
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&buf), sizeof(float) * want));
        cap = want;
        return 0;
    };
    if ((rc = grow(d.lines, d.lcap, (size_t)d.nl * 10)) || (rc = grow(d.points, d.pcap, (size_t)d.np * 8))) return rc;
    const DebugOut out{d.lines, d.points, d.loff, d.poff, d.rgb};
    const int blocks = std::max(1, std::min((slots + 127) / 128, 65536));
    hipLaunchKernelGGL(k_dbg_write, dim3(blocks), dim3(128), 0, st, w, out, (const int*)d.lc, (const int*)d.pc);
    HIP_TRY(hipGetLastError());
    return 0;
}

static int step_impl(avbd_ctx* c, int32_t n_steps) {
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    if (n_steps < 0) return fail(AVBD_E_ARG, "negative step count");
    HIP_TRY(hipSetDevice(c->device));
    avbd_timings acc{};
    c->removed_scene.clear();
    c->removed_index.clear();
    for (int k = 0; k < n_steps; ++k) {
        int rc;
        // the contact points are recorded by the step's warm start
        if (c->dbg_on && (rc = ensure_debug(c))) return rc;
        rc = do_step(c);
        if (rc) return rc;
        int g[G_WORDS];
        HIP_TRY(hipMemcpyAsync(g, c->w.g, sizeof(g), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->collect_kernel_times();
        c->tuner.collect();
        if (c->dbg_on && (rc = extract_debug(c))) return rc;
        if (g[G_ANY_OOB] && (rc = remove_out_of_bounds(c))) return rc;
        acc.broad_phase_ms = elapsed(c->ev[0], c->ev[1]);
        acc.warmstart_ms = elapsed(c->ev[1], c->ev[2]);
        acc.prediction_ms = elapsed(c->ev[2], c->ev[3]);
        acc.primal_dual_ms = elapsed(c->ev[3], c->ev[4]);
        acc.velocity_update_ms = elapsed(c->ev[4], c->ev[5]);
        acc.post_stab_ms = elapsed(c->ev[5], c->ev[6]);
        acc.total_ms = elapsed(c->ev[0], c->ev[6]);
        acc.n_colours = c->last_ncol;
    }
    if (n_steps > 0) c->timings = acc;
    return 0;
}

int avbd_step(avbd_ctx* c, int32_t n_steps) {
    if (!c) return fail(AVBD_E_ARG, "null context");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    // every scene runs: the flags are set on the device, in stream order
    // (no read-back of the status words)
    hipLaunchKernelGGL(k_set_active, dim3(grid_for(c->w.S)), dim3(256), 0, c->stream, c->w, (const int*)nullptr, 0, 1);
    c->subset = false;
    return step_impl(c, n_steps);
}

int avbd_step_scenes(avbd_ctx* c, const int32_t* ids, int32_t n_ids, int32_t n_steps) {
    if (!c || (!ids && n_ids > 0)) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    const int S = c->w.S;
    std::vector<char> on(S, 0);
    for (int i = 0; i < n_ids; ++i) {
        if (ids[i] < 0 || ids[i] >= S) return fail(AVBD_E_ARG, "scene id out of range");
        on[ids[i]] = 1;
    }
    c->run_ids.clear();
    for (int s = 0; s < S; ++s)
        if (on[s]) c->run_ids.push_back(s);
    // the run flags are set on the device, in stream order (no read-back of
    // the status words): clear every scene's, then mark the listed ones from
    // the pinned staging (which the previous call's copy has left by now)
    HIP_TRY(hipEventSynchronize(c->run_ev));
    const int n = (int)c->run_ids.size();
    std::copy(c->run_ids.begin(), c->run_ids.end(), c->run_pin);
    hipLaunchKernelGGL(k_set_active, dim3(grid_for(S)), dim3(256), 0, c->stream, c->w, (const int*)nullptr, 0, 0);
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(c->run_dev, c->run_pin, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipEventRecord(c->run_ev, c->stream));
        hipLaunchKernelGGL(k_set_active, dim3(grid_for(n)), dim3(256), 0, c->stream, c->w, (const int*)c->run_dev, n, 0);
    }
    c->subset = n < S;
    const int rc = step_impl(c, n_steps);
    c->subset = false;
    return rc;
}

int avbd_synchronize(avbd_ctx* c) {
    if (!c) return fail(AVBD_E_ARG, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int avbd_get_counts(avbd_ctx* c, avbd_world* out) {
    if (!c || !out) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    const DevWorld& w = c->w;
    int g[G_WORDS];
    HIP_TRY(hipMemcpyAsync(g, w.g, sizeof(g), hipMemcpyDeviceToHost, c->stream));
    std::vector<uint32_t> jf, sf;
    int rc;
    if ((rc = download_vec(c, jf, w.j_flags, w.J)) || (rc = download_vec(c, sf, w.s_flags, w.SP))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    out->n_scenes = w.S;
    out->n_bodies = w.N - (int)std::count(c->body_dead.begin(), c->body_dead.end(), 1);
    out->n_joints = (int)std::count_if(jf.begin(), jf.end(), [](uint32_t f) { return !(f & FORCE_DEAD); });
    out->n_springs = (int)std::count_if(sf.begin(), sf.end(), [](uint32_t f) { return !(f & FORCE_DEAD); });
    out->n_manifolds = g[G_MANIFOLDS];
    out->n_tets = w.T - (int)std::count(c->tet_dead.begin(), c->tet_dead.end(), 1);
    out->n_tris = w.TR - (int)std::count(c->tri_dead.begin(), c->tri_dead.end(), 1);
    return 0;
}

int avbd_download(avbd_ctx* c, avbd_world* out) {
    if (!c || !out) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    avbd_world cnt{};
    int rc = avbd_get_counts(c, &cnt);
    if (rc) return rc;
    if (out->n_scenes < cnt.n_scenes || out->n_bodies < cnt.n_bodies || out->n_joints < cnt.n_joints ||
        out->n_springs < cnt.n_springs || out->n_manifolds < cnt.n_manifolds || out->n_tets < cnt.n_tets ||
        out->n_tris < cnt.n_tris)
        return fail(AVBD_E_ARG, "output arrays smaller than avbd_get_counts");
    const DevWorld& w = c->w;
    const int N = w.N, J = w.J, SP = w.SP, M = cnt.n_manifolds;
    std::vector<float4> pos, rot, scale, vel, av, pv, pav, inert;
    std::vector<uint32_t> flags, jf, sf;
    std::vector<int> scene, st, ja, jb, sa, sb, ma, mbv, mn;
    std::vector<long long> jser, sser, mser;
    std::vector<float4> jra, jrb, jrest, sra, srb, hot, cold;
    std::vector<float> jrow, srow, mfr;
    if ((rc = download_vec(c, pos, w.pos, N)) || (rc = download_vec(c, rot, w.rot, N)) ||
        (rc = download_vec(c, scale, w.scale, N)) || (rc = download_vec(c, vel, w.vel, N)) ||
        (rc = download_vec(c, av, w.angvel, N)) || (rc = download_vec(c, pv, w.pvel, N)) ||
        (rc = download_vec(c, pav, w.pangvel, N)) || (rc = download_vec(c, inert, w.inertia, N)) ||
        (rc = download_vec(c, flags, w.flags, N)) || (rc = download_vec(c, scene, w.scene, N)) ||
        (rc = download_vec(c, st, w.st, (size_t)w.S * ST_WORDS)) || (rc = download_vec(c, ja, w.j_a, J)) ||
        (rc = download_vec(c, jb, w.j_b, J)) || (rc = download_vec(c, jser, w.j_serial, J)) ||
        (rc = download_vec(c, jf, w.j_flags, J)) || (rc = download_vec(c, jra, w.j_ra, J)) ||
        (rc = download_vec(c, jrb, w.j_rb, J)) || (rc = download_vec(c, jrest, w.j_rest, J)) ||
        (rc = download_vec(c, jrow, w.j_row, (size_t)J * 36)) || (rc = download_vec(c, sa, w.s_a, SP)) ||
        (rc = download_vec(c, sb, w.s_b, SP)) || (rc = download_vec(c, sser, w.s_serial, SP)) ||
        (rc = download_vec(c, sf, w.s_flags, SP)) || (rc = download_vec(c, sra, w.s_ra, SP)) ||
        (rc = download_vec(c, srb, w.s_rb, SP)) || (rc = download_vec(c, srow, w.s_row, (size_t)SP * 6)) ||
        (rc = download_vec(c, ma, w.m.a, M)) || (rc = download_vec(c, mbv, w.m.b, M)) ||
        (rc = download_vec(c, mser, w.m.serial, M)) || (rc = download_vec(c, mn, w.m.n, M)) ||
        (rc = download_vec(c, mfr, w.m.fric, M)) ||
        (rc = download_vec(c, hot, w.m.hot, (size_t)M * 8 * CONTACT_HOT_F4)) ||
        (rc = download_vec(c, cold, w.m.cold, (size_t)M * 8 * CONTACT_COLD_F4)))
        return rc;
    std::vector<int4> tv, rv;
    std::vector<float4> td, rd;
    if ((rc = download_vec(c, tv, w.t_v, w.T)) || (rc = download_vec(c, td, w.t_d, (size_t)w.T * 3)) ||
        (rc = download_vec(c, rv, w.r_v, w.TR)) || (rc = download_vec(c, rd, w.r_d, (size_t)w.TR * 2)))
        return rc;
    std::vector<float> rho;
    if ((rc = download_vec(c, rho, w.rho, w.S))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int s = 0; s < w.S; ++s) {
        out->scenes[s] = c->scenes[s];
        out->scenes[s].emergency_stop = st[s * ST_WORDS + ST_EMERGENCY];
        out->scenes[s].n_removed = st[s * ST_WORDS + ST_REMOVED];
        out->scenes[s].trust_region_rho = rho[s];
    }
    auto put3 = [](float* d, float4 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; };
    const std::vector<int> live = live_index(c);
    auto L = [&](int b) { return b < 0 ? b : live[b]; };
    for (int i = 0; i < N; ++i) {
        if (live[i] < 0) continue;
        avbd_body& b = out->bodies[live[i]];
        std::memset(&b, 0, sizeof(b));
        put3(b.pos, pos[i]);
        b.rot[0] = rot[i].x; b.rot[1] = rot[i].y; b.rot[2] = rot[i].z; b.rot[3] = rot[i].w;
        put3(b.scale, scale[i]);
        b.radius = scale[i].w;
        put3(b.vel, vel[i]);
        b.mass = vel[i].w;
        put3(b.ang_vel, av[i]);
        b.friction = av[i].w;
        put3(b.prev_vel, pv[i]);
        put3(b.prev_ang_vel, pav[i]);
        put3(b.inertia, inert[i]);
        b.flags = flags[i];
        b.scene = scene[i];
    }
    int k = 0;
    for (int i = 0; i < J; ++i) {
        if (jf[i] & FORCE_DEAD) continue;
        avbd_joint& j = out->joints[k++];
        std::memset(&j, 0, sizeof(j));
        j.body_a = L(ja[i]); j.body_b = L(jb[i]); j.serial = jser[i]; j.flags = jf[i] & AVBD_FORCE_DISABLED;
        put3(j.r_a, jra[i]); j.torque_arm = jra[i].w;
        put3(j.r_b, jrb[i]);
        j.rest_rot[0] = jrest[i].x; j.rest_rot[1] = jrest[i].y; j.rest_rot[2] = jrest[i].z; j.rest_rot[3] = jrest[i].w;
        for (int r = 0; r < 6; ++r) {
            const float* row = &jrow[(size_t)(i * 6 + r) * 6];
            j.stiffness[r] = row[0]; j.fracture[r] = row[1]; j.fmin[r] = row[2]; j.fmax[r] = row[3];
            j.lambda[r] = row[4]; j.penalty[r] = row[5];
        }
    }
    k = 0;
    for (int i = 0; i < SP; ++i) {
        if (sf[i] & FORCE_DEAD) continue;
        avbd_spring& s = out->springs[k++];
        std::memset(&s, 0, sizeof(s));
        s.body_a = L(sa[i]); s.body_b = L(sb[i]); s.serial = sser[i]; s.flags = sf[i] & AVBD_FORCE_DISABLED;
        put3(s.r_a, sra[i]); s.rest_length = sra[i].w;
        put3(s.r_b, srb[i]);
        const float* row = &srow[(size_t)i * 6];
        s.stiffness = row[0]; s.fracture = row[1]; s.fmin = row[2]; s.fmax = row[3]; s.lambda = row[4]; s.penalty = row[5];
    }
    for (int i = 0; i < M; ++i) {
        avbd_manifold& m = out->manifolds[i];
        std::memset(&m, 0, sizeof(m));
        m.body_a = L(ma[i]); m.body_b = L(mbv[i]); m.serial = mser[i]; m.n_contacts = mn[i]; m.friction = mfr[i];
        for (int ci = 0; ci < mn[i]; ++ci) {
            const float4* h = &hot[((size_t)i * 8 + ci) * CONTACT_HOT_F4];
            const float4* cd = &cold[((size_t)i * 8 + ci) * CONTACT_COLD_F4];
            avbd_contact& ct = m.contacts[ci];
            std::memcpy(&ct.id, &cd[0].x, 4);
            ct.stick = cd[0].y != 0.f;
            ct.r_a[0] = cd[0].z; ct.r_a[1] = cd[0].w; ct.r_a[2] = cd[1].x;
            ct.r_b[0] = cd[1].y; ct.r_b[1] = cd[1].z; ct.r_b[2] = cd[1].w;
            put3(ct.normal, cd[2]);
            ct.penetration = cd[2].w;
            ct.lambda[0] = h[3].w; ct.lambda[1] = h[4].w; ct.lambda[2] = h[5].x;
            ct.penalty[0] = h[5].y; ct.penalty[1] = h[5].z; ct.penalty[2] = h[5].w;
        }
    }
    k = 0;
    for (int i = 0; i < w.T; ++i) {
        if (c->tet_dead[i]) continue;
        avbd_tet& t = out->tets[k++];
        t.v[0] = L(tv[i].x); t.v[1] = L(tv[i].y); t.v[2] = L(tv[i].z); t.v[3] = L(tv[i].w);
        const float d[12] = {td[i * 3].x, td[i * 3].y, td[i * 3].z, td[i * 3].w, td[i * 3 + 1].x, td[i * 3 + 1].y,
                             td[i * 3 + 1].z, td[i * 3 + 1].w, td[i * 3 + 2].x, td[i * 3 + 2].y, td[i * 3 + 2].z,
                             td[i * 3 + 2].w};
        for (int q = 0; q < 9; ++q) t.dm_inv[q] = d[q];
        t.rest_volume = d[9]; t.mu = d[10]; t.lambda = d[11];
        t.serial = c->tet_serial[i];
    }
    k = 0;
    for (int i = 0; i < w.TR; ++i) {
        if (c->tri_dead[i]) continue;
        avbd_tri& t = out->tris[k++];
        t.v[0] = L(rv[i].x); t.v[1] = L(rv[i].y); t.v[2] = L(rv[i].z);
        t.dm_inv[0] = rd[i * 2].x; t.dm_inv[1] = rd[i * 2].y; t.dm_inv[2] = rd[i * 2].z; t.dm_inv[3] = rd[i * 2].w;
        t.rest_area = rd[i * 2 + 1].x; t.mu = rd[i * 2 + 1].y; t.lambda = rd[i * 2 + 1].z;
        t.serial = c->tri_serial[i];
    }
    out->n_scenes = cnt.n_scenes;
    out->n_bodies = cnt.n_bodies;
    out->n_joints = cnt.n_joints;
    out->n_springs = cnt.n_springs;
    out->n_manifolds = cnt.n_manifolds;
    out->n_tets = cnt.n_tets;
    out->n_tris = cnt.n_tris;
    return 0;
}

int avbd_get_status(avbd_ctx* c, int32_t* emergency, int32_t* removed, int32_t n) {
    if (!c) return fail(AVBD_E_ARG, "null context");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    std::vector<int> st((size_t)c->w.S * ST_WORDS);
    HIP_TRY(hipMemcpy(st.data(), c->w.st, st.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int s = 0; s < n && s < c->w.S; ++s) {
        if (emergency) emergency[s] = st[s * ST_WORDS + ST_EMERGENCY];
        if (removed) removed[s] = st[s * ST_WORDS + ST_REMOVED];
    }
    return 0;
}

int avbd_get_trust_region_rho(avbd_ctx* c, float* rho, int32_t n) {
    if (!c || (!rho && n > 0)) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    const int k = std::max(0, std::min(n, c->w.S));
    HIP_TRY(hipMemcpy(rho, c->w.rho, (size_t)k * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int avbd_remove_bodies(avbd_ctx* c, const int32_t* bodies, int32_t n) {
    if (!c || (!bodies && n > 0)) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    if (n < 0) return fail(AVBD_E_ARG, "negative body count");
    HIP_TRY(hipSetDevice(c->device));
    DevWorld& w = c->w;
    const int S = w.S;
    // live index -> device index
    std::vector<int> dev;
    dev.reserve(w.N);
    for (int i = 0; i < w.N; ++i)
        if (!c->body_dead[i]) dev.push_back(i);
    std::vector<char> aff(S, 0), taken(w.N, 0);
    std::vector<std::vector<int>> lists(S);
    for (int k = 0; k < n; ++k) {
        const int b = bodies[k];
        if (b < 0 || b >= (int)dev.size()) return fail(AVBD_E_ARG, "body index out of range");
        const int d = dev[b];
        if (taken[d]) return fail(AVBD_E_ARG, "body listed twice");
        taken[d] = 1;
        int s = (int)(std::upper_bound(c->sc_off.begin(), c->sc_off.end(), d) - c->sc_off.begin()) - 1;
        aff[s] = 1;
        lists[s].push_back(d);
    }
    c->removed_scene.clear();
    c->removed_index.clear();
    if (n == 0) return 0;
    std::vector<int> stw;
    int rc;
    if ((rc = download_vec(c, stw, w.st, (size_t)S * ST_WORDS))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return remove_bodies(c, aff, stw, &lists);
}

int avbd_add_bodies(avbd_ctx* c, const avbd_body* bodies, int32_t n) {
    if (!c || (!bodies && n > 0)) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    if (n < 0) return fail(AVBD_E_ARG, "negative body count");
    if (n == 0) return 0;
    HIP_TRY(hipSetDevice(c->device));
    DevWorld& w = c->w;
    const int S = w.S;
    // the slots: per scene its first dead slots (they trail the live bodies)
    std::vector<int> next(S, -1), slots(n);
    for (int k = 0; k < n; ++k) {
        const int s = bodies[k].scene;
        if (s < 0 || s >= S) return fail(AVBD_E_ARG, "body scene out of range");
        const int a = c->sc_off[s], e = c->sc_off[s + 1];
        if (next[s] < 0) {
            next[s] = a;
            while (next[s] < e && !c->body_dead[next[s]]) ++next[s];
        }
        if (next[s] >= e)
            return fail(AVBD_E_CAPACITY, "scene has no free body slot (slots come from removals and from the "
                                         "spare_bodies option at upload; otherwise upload the world again)");
        slots[k] = next[s]++;
    }
    std::vector<int> stw;
    int rc;
    if ((rc = download_vec(c, stw, w.st, (size_t)S * ST_WORDS))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->drop_graphs();
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int k = 0; k < n; ++k) {
        const avbd_body& b = bodies[k];
        const int i = slots[k];
        // the record as avbd_upload converts it; the step's caches cleared
        const float4 f4s[] = {f4(b.pos, 0.f), make_float4(b.rot[0], b.rot[1], b.rot[2], b.rot[3]), f4(b.scale, b.radius),
                              f4(b.vel, b.mass), f4(b.ang_vel, b.friction), f4(b.prev_vel, 0.f),
                              f4(b.prev_ang_vel, 0.f), f4(b.inertia, 0.f), z, z, z, z, z, z, z};
        float4* dst[] = {w.pos, w.rot, w.scale, w.vel, w.angvel, w.pvel, w.pangvel, w.inertia,
                         w.last_pos, w.last_rot, w.inert_pos, w.inert_rot, w.diff0, w.diff1, w.drag};
        for (int a = 0; a < 15; ++a)
            HIP_TRY(hipMemcpyAsync(dst[a] + i, &f4s[a], sizeof(float4), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(w.flags + i, &b.flags, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
        const float4 white = make_float4(1.f, 1.f, 1.f, 0.f);
        if (c->dbg.rgb) HIP_TRY(hipMemcpyAsync(c->dbg.rgb + i, &white, sizeof(float4), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));  // the sources are on this frame
        c->body_dead[i] = 0;
        stw[b.scene * ST_WORDS + ST_DIRTY] = DIRTY_FULL;
    }
    if ((rc = upload_vec(c, w.st, stw))) return rc;
    if ((rc = build_incidence(c))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int avbd_set_body_colours(avbd_ctx* c, const float* rgb, int32_t n) {
    if (!c || (!rgb && n > 0)) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    DevWorld& w = c->w;
    std::vector<float4> col(w.N, make_float4(1.f, 1.f, 1.f, 0.f));
    int k = 0;
    for (int i = 0; i < w.N; ++i)
        if (!c->body_dead[i]) {
            if (k >= n) return fail(AVBD_E_ARG, "colour count differs from the live body count");
            col[i] = make_float4(rgb[3 * k], rgb[3 * k + 1], rgb[3 * k + 2], 0.f);
            ++k;
        }
    if (k != n) return fail(AVBD_E_ARG, "colour count differs from the live body count");
    int rc;
    if (!c->dbg.rgb && (rc = c->alloc(&c->dbg.rgb, (size_t)w.N))) return rc;
    if ((rc = upload_vec(c, c->dbg.rgb, col))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int avbd_get_debug_lines(avbd_ctx* c, avbd_line* out, int64_t cap, int64_t* n) {
    if (!c || !n || (!out && cap > 0)) return fail(AVBD_E_ARG, "null argument");
    HIP_TRY(hipStreamSynchronize(c->stream));
    *n = c->dbg.nl;
    const long long k = std::min<long long>(cap, c->dbg.nl);
    if (k > 0) HIP_TRY(hipMemcpy(out, c->dbg.lines, sizeof(avbd_line) * (size_t)k, hipMemcpyDeviceToHost));
    return 0;
}

int avbd_get_instances(avbd_ctx* c, avbd_instance* out, int32_t cap, int32_t* n) {
    if (!c || !n || (!out && cap > 0)) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    HIP_TRY(hipSetDevice(c->device));
    DevWorld& w = c->w;
    const int n_live = w.N - (int)std::count(c->body_dead.begin(), c->body_dead.end(), 1);
    *n = n_live;
    if (!out || cap <= 0 || n_live == 0) return 0;
    hipStream_t st = c->stream;
    size_t cub_bytes = 0;
    int* nul = nullptr;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, cub_bytes, nul, nul, w.N, st));
    const size_t ints = 2 * (size_t)w.N + 4;
    // one scratch buffer for the instances, the live flags, their scan and
    // the scan's temporary storage; kept for the next call (a renderer asks
    // every frame)
    const size_t need = sizeof(avbd_instance) * (size_t)n_live + sizeof(int) * ints + cub_bytes + 256;
    if (need > c->dbg.inst_bytes) {
        if (c->dbg.inst) HIP_TRY(hipFree(c->dbg.inst));
        c->dbg.inst = nullptr;
        c->dbg.inst_bytes = 0;
        HIP_TRY(hipMalloc(&c->dbg.inst, need + need / 4));
        c->dbg.inst_bytes = need + need / 4;
    }
    float4* d_out = static_cast<float4*>(c->dbg.inst);
    int* flag = reinterpret_cast<int*>(d_out + (size_t)n_live * 8);
    int* off = flag + w.N + 2;
    void* cub = reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(off + w.N + 2) + 255) & ~(uintptr_t)255);
    hipLaunchKernelGGL(k_live_flag, dim3(grid_for(w.N)), dim3(256), 0, st, w, flag);
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(cub, cub_bytes, flag, off, w.N, st));
    hipLaunchKernelGGL(k_instances, dim3(grid_for(w.N)), dim3(256), 0, st, w, (const int*)flag, (const int*)off,
                       (const float4*)c->dbg.rgb, d_out);
    HIP_TRY(hipMemcpyAsync(out, d_out, sizeof(avbd_instance) * (size_t)std::min(cap, n_live), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int avbd_get_instances_noalloc(avbd_ctx* c, avbd_instance* out, int32_t cap, int32_t* n) {
    if (!c || !n || (!out && cap > 0)) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    DevWorld& w = c->w;
    if (w.N > kNoallocMaxBodies)
        return fail(AVBD_E_CAPACITY, "avbd_get_instances_noalloc: the world has " + std::to_string(w.N) +
                                         " body slots, at most " + std::to_string(kNoallocMaxBodies) +
                                         " fit its scratch (use avbd_get_instances)");
    HIP_TRY(hipSetDevice(c->device));
    unsigned char* base = static_cast<unsigned char*>(noalloc_scratch(c->device));
    if (!base) return fail(AVBD_E_STATE, "scratch of this device not set up");
    const int n_live = w.N - (int)std::count(c->body_dead.begin(), c->body_dead.end(), 1);
    *n = n_live;
    if (!out || cap <= 0 || n_live == 0) return 0;
    hipStream_t st = c->stream;
    // every word read below is written first by this call: nothing to clear
    float4* d_out = reinterpret_cast<float4*>(base);
    int* flag = reinterpret_cast<int*>(base + sizeof(avbd_instance) * (size_t)kNoallocMaxBodies);
    int* off = flag + kNoallocMaxBodies;
    int* total = off + kNoallocMaxBodies;
    hipLaunchKernelGGL(k_live_flag, dim3(grid_for(w.N)), dim3(256), 0, st, w, flag);
    hipLaunchKernelGGL(k_scan_flags_1wg, dim3(1), dim3(1024), 0, st, (const int*)flag, off, w.N, total);
    hipLaunchKernelGGL(k_instances, dim3(grid_for(w.N)), dim3(256), 0, st, w, (const int*)flag, (const int*)off,
                       (const float4*)c->dbg.rgb, d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, d_out, sizeof(avbd_instance) * (size_t)std::min(cap, n_live), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

int avbd_get_debug_points(avbd_ctx* c, avbd_debug_point* out, int64_t cap, int64_t* n) {
    if (!c || !n || (!out && cap > 0)) return fail(AVBD_E_ARG, "null argument");
    HIP_TRY(hipStreamSynchronize(c->stream));
    *n = c->dbg.np;
    const long long k = std::min<long long>(cap, c->dbg.np);
    if (k > 0) HIP_TRY(hipMemcpy(out, c->dbg.points, sizeof(avbd_debug_point) * (size_t)k, hipMemcpyDeviceToHost));
    return 0;
}

int avbd_last_removed(avbd_ctx* c, int32_t* scene, int32_t* index, int32_t cap, int32_t* n_out) {
    if (!c || !n_out) return fail(AVBD_E_ARG, "null argument");
    const int n = (int)c->removed_index.size();
    *n_out = n;
    for (int i = 0; i < n && i < cap; ++i) {
        if (scene) scene[i] = c->removed_scene[i];
        if (index) index[i] = c->removed_index[i];
    }
    return 0;
}

int avbd_download_bodies(avbd_ctx* c, avbd_body* out, int32_t n) {
    if (!c || (!out && n > 0)) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    const DevWorld& w = c->w;
    const std::vector<int> live = live_index(c);
    const int n_live = w.N - (int)std::count(c->body_dead.begin(), c->body_dead.end(), 1);
    if (n != n_live) return fail(AVBD_E_ARG, "body count differs from the device's live bodies");
    std::vector<float4> pos, rot, scale, vel, av, pv, pav, inert;
    std::vector<uint32_t> flags;
    std::vector<int> scene;
    int rc;
    if ((rc = download_vec(c, pos, w.pos, w.N)) || (rc = download_vec(c, rot, w.rot, w.N)) ||
        (rc = download_vec(c, scale, w.scale, w.N)) || (rc = download_vec(c, vel, w.vel, w.N)) ||
        (rc = download_vec(c, av, w.angvel, w.N)) || (rc = download_vec(c, pv, w.pvel, w.N)) ||
        (rc = download_vec(c, pav, w.pangvel, w.N)) || (rc = download_vec(c, inert, w.inertia, w.N)) ||
        (rc = download_vec(c, flags, w.flags, w.N)) || (rc = download_vec(c, scene, w.scene, w.N)))
        return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    auto put3 = [](float* d, float4 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; };
    for (int i = 0; i < w.N; ++i) {
        if (live[i] < 0) continue;
        avbd_body& b = out[live[i]];
        std::memset(&b, 0, sizeof(b));
        put3(b.pos, pos[i]);
        b.rot[0] = rot[i].x; b.rot[1] = rot[i].y; b.rot[2] = rot[i].z; b.rot[3] = rot[i].w;
        put3(b.scale, scale[i]);
        b.radius = scale[i].w;
        put3(b.vel, vel[i]);
        b.mass = vel[i].w;
        put3(b.ang_vel, av[i]);
        b.friction = av[i].w;
        put3(b.prev_vel, pv[i]);
        put3(b.prev_ang_vel, pav[i]);
        put3(b.inertia, inert[i]);
        b.flags = flags[i];
        b.scene = scene[i];
    }
    return 0;
}

int avbd_upload_bodies(avbd_ctx* c, const avbd_body* in, int32_t n) {
    if (!c || (!in && n > 0)) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    DevWorld& w = c->w;
    const std::vector<int> live = live_index(c);
    const int n_live = w.N - (int)std::count(c->body_dead.begin(), c->body_dead.end(), 1);
    if (n != n_live) return fail(AVBD_E_ARG, "body count differs from the device's live bodies");
    std::vector<float4> pos, rot, vel, av, pv, pav;
    int rc;
    if ((rc = download_vec(c, pos, w.pos, w.N)) || (rc = download_vec(c, rot, w.rot, w.N)) ||
        (rc = download_vec(c, vel, w.vel, w.N)) || (rc = download_vec(c, av, w.angvel, w.N)) ||
        (rc = download_vec(c, pv, w.pvel, w.N)) || (rc = download_vec(c, pav, w.pangvel, w.N)))
        return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int i = 0; i < w.N; ++i) {
        if (live[i] < 0) continue;
        const avbd_body& b = in[live[i]];
        pos[i] = make_float4(b.pos[0], b.pos[1], b.pos[2], pos[i].w);
        rot[i] = make_float4(b.rot[0], b.rot[1], b.rot[2], b.rot[3]);
        vel[i] = make_float4(b.vel[0], b.vel[1], b.vel[2], vel[i].w);
        av[i] = make_float4(b.ang_vel[0], b.ang_vel[1], b.ang_vel[2], av[i].w);
        pv[i] = make_float4(b.prev_vel[0], b.prev_vel[1], b.prev_vel[2], 0.f);
        pav[i] = make_float4(b.prev_ang_vel[0], b.prev_ang_vel[1], b.prev_ang_vel[2], 0.f);
    }
    if ((rc = upload_vec(c, w.pos, pos)) || (rc = upload_vec(c, w.rot, rot)) || (rc = upload_vec(c, w.vel, vel)) ||
        (rc = upload_vec(c, w.angvel, av)) || (rc = upload_vec(c, w.pvel, pv)) || (rc = upload_vec(c, w.pangvel, pav)))
        return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int avbd_set_drag(avbd_ctx* c, const int32_t* bodies, const float* added_velocity, const int32_t* dragged,
                  int32_t n) {
    if (!c || (n > 0 && (!bodies || !added_velocity || !dragged)) || n < 0) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    if (n == 0) return 0;
    HIP_TRY(hipSetDevice(c->device));
    DevWorld& w = c->w;
    // live index -> device slot
    std::vector<int> slot;
    slot.reserve(w.N);
    for (int i = 0; i < w.N; ++i)
        if (!c->body_dead[i]) slot.push_back(i);
    std::vector<float4> drag;
    int rc;
    if ((rc = download_vec(c, drag, w.drag, w.N))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int k = 0; k < n; ++k) {
        if (bodies[k] < 0 || bodies[k] >= (int)slot.size()) return fail(AVBD_E_ARG, "body index out of range");
        const float* v = added_velocity + 3 * k;
        drag[slot[bodies[k]]] = make_float4(v[0], v[1], v[2], dragged[k] ? 1.f : 0.f);
    }
    if ((rc = upload_vec(c, w.drag, drag))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int avbd_get_timings(avbd_ctx* c, avbd_timings* out) {
    if (!c || !out) return fail(AVBD_E_ARG, "null argument");
    *out = c->timings;
    return 0;
}

int avbd_get_colours(avbd_ctx* c, int32_t* out, int32_t n) {
    if (!c || !out) return fail(AVBD_E_ARG, "null argument");
    if (!c->have_world) return fail(AVBD_E_STATE, "no world uploaded");
    std::vector<int> col;
    int rc = download_vec(c, col, c->w.colour, c->w.N);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    const std::vector<int> live = live_index(c);
    for (int i = 0; i < c->w.N; ++i)
        if (live[i] >= 0 && live[i] < n) out[live[i]] = col[i];
    return 0;
}

}  // extern "C"
