// snapshot_fuzz.cpp -- the snapshot reader (skyephyx_amd/csrc/avbd_snapshot.cpp)
// on damaged files, built with -fsanitize=address,undefined by
// tests/test_sanitizers.py.
//
//   snapshot_fuzz <valid snapshot> <scratch file>
//
// Reads the valid file (must succeed), then every truncation of its header
// and a spread of truncations of its body, seeded byte flips, and headers
// whose counts or record sizes are set to huge, negative or mismatched
// values: each read must either succeed (a flip in alignment padding) or
// return an AVBD_E_* code, with no sanitizer report.  Prints a summary.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "avbd_capi.h"

// The library's error sink (avbd_host.hip); this harness links the snapshot
// code alone.
int avbd_set_error(int code, const std::string&) { return code; }

namespace {

std::vector<unsigned char> slurp(const char* path) {
    std::vector<unsigned char> d;
    FILE* f = std::fopen(path, "rb");
    if (!f) return d;
    unsigned char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) d.insert(d.end(), buf, buf + n);
    std::fclose(f);
    return d;
}

void spill(const char* path, const std::vector<unsigned char>& d, size_t n) {
    FILE* f = std::fopen(path, "wb");
    if (n) std::fwrite(d.data(), 1, n, f);
    std::fclose(f);
}

// counts, then a full read into arrays sized from them; returns the first
// nonzero status
int read_all(const char* path) {
    avbd_world c{};
    int rc = avbd_snapshot_counts(path, &c);
    if (rc) return rc;
    const int32_t n[7] = {c.n_scenes, c.n_bodies, c.n_joints, c.n_springs, c.n_manifolds, c.n_tets, c.n_tris};
    for (int32_t v : n)
        if (v < 0) {
            std::printf("negative record count accepted\n");
            std::exit(6);
        }
    std::vector<avbd_scene> s((size_t)c.n_scenes);
    std::vector<avbd_body> b((size_t)c.n_bodies);
    std::vector<avbd_joint> j((size_t)c.n_joints);
    std::vector<avbd_spring> sp((size_t)c.n_springs);
    std::vector<avbd_manifold> m((size_t)c.n_manifolds);
    std::vector<avbd_tet> t((size_t)c.n_tets);
    std::vector<avbd_tri> tr((size_t)c.n_tris);
    avbd_world w = c;
    w.scenes = s.data();
    w.bodies = b.data();
    w.joints = j.data();
    w.springs = sp.data();
    w.manifolds = m.data();
    w.tets = t.data();
    w.tris = tr.data();
    return avbd_snapshot_read(path, &w);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    const std::vector<unsigned char> good = slurp(argv[1]);
    if (good.size() < 128) return 3;
    if (read_all(argv[1]) != 0) {
        std::printf("valid snapshot refused\n");
        return 4;
    }
    const char* tmp = argv[2];
    int refused = 0, accepted = 0, cases = 0;
    auto run = [&](const std::vector<unsigned char>& d, size_t n, bool must_fail) {
        spill(tmp, d, n);
        const int rc = read_all(tmp);
        ++cases;
        if (rc) ++refused;
        else ++accepted;
        if (must_fail && rc == 0) {
            std::printf("damaged file accepted (%zu bytes)\n", n);
            std::exit(5);
        }
    };
    // truncations: every length inside the header, a spread over the body
    for (size_t n = 0; n < 128; ++n) run(good, n, true);
    for (size_t k = 1; k < 257; ++k) run(good, 128 + (good.size() - 129) * k / 257, true);
    // byte flips (seeded): a flip in a checked byte must be refused
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (int k = 0; k < 600; ++k) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        std::vector<unsigned char> d = good;
        const size_t at = (size_t)(x % d.size());
        d[at] ^= (unsigned char)(1u << ((x >> 32) % 8));
        run(d, d.size(), false);
    }
    // header fields: counts and record sizes set to extreme values
    const size_t fields[] = {8, 12, 16, 20};  // version, header bytes, ABI version, kinds
    for (size_t off : fields)
        for (uint32_t v : {0u, 1u, 0x7fffffffu, 0xffffffffu}) {
            std::vector<unsigned char> d = good;
            std::memcpy(&d[off], &v, 4);
            run(d, d.size(), false);
        }
    for (size_t off = 24; off + 8 <= 128; off += 4)
        for (int64_t v : {(int64_t)-1, (int64_t)1 << 40, (int64_t)INT32_MAX + 1, (int64_t)3}) {
            std::vector<unsigned char> d = good;
            std::memcpy(&d[off], &v, off + 8 <= 128 ? 8 : 4);
            run(d, d.size(), false);
        }
    std::printf("%d cases: %d refused, %d read\n", cases, refused, accepted);
    return 0;
}
