// avbd_snapshot.cpp -- world snapshots: the ABI records of a world in one
// binary file, for fixtures, reproducible configurations and server preload.
//
// Layout (little-endian, the byte layout of the records of this ABI):
//   [0, 128)   header: magic "AVBDSNAP", format version, header size, ABI
//              version, record kinds, size of each record kind, count of each
//              kind, and a 64-bit checksum of the sections' bytes
//   sections   scenes, bodies, joints, springs, manifolds (contacts inline),
//              tets, tris: each the records back to back, starting at a
//              64-byte aligned offset (zero padding between); the file ends
//              with the last record
// A reader checks magic, versions, record sizes (a snapshot is only read by
// a build with the same record layout), counts, file size and checksum.
// Host code only.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "avbd_internal.h"

namespace {

constexpr char kMagic[8] = {'A', 'V', 'B', 'D', 'S', 'N', 'A', 'P'};
constexpr uint32_t kHeaderBytes = 128;
constexpr int kKinds = 7;

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    uint32_t abi_version;
    uint32_t kinds;
    uint32_t record_bytes[kKinds];
    uint32_t pad0;
    int64_t counts[kKinds];
    uint64_t checksum;
    uint64_t reserved;  // 0
};
static_assert(sizeof(Header) == kHeaderBytes, "snapshot header is 128 bytes");

const uint32_t kRecordBytes[kKinds] = {sizeof(avbd_scene),    sizeof(avbd_body), sizeof(avbd_joint),
                                       sizeof(avbd_spring),   sizeof(avbd_manifold), sizeof(avbd_tet),
                                       sizeof(avbd_tri)};

// Word-wise multiply-xorshift hash of the section bytes, fed section by
// section in file order (not a cryptographic digest: it catches truncation
// and corruption).
struct Hash {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    void feed(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        size_t i = 0;
        for (; i + 8 <= n; i += 8) {
            uint64_t v;
            std::memcpy(&v, b + i, 8);
            mix(v);
        }
        if (i < n) {
            uint64_t v = 0;
            std::memcpy(&v, b + i, n - i);
            mix(v ^ (uint64_t)(n - i) << 56);
        }
    }
    void mix(uint64_t v) {
        h ^= v;
        h *= 0xFF51AFD7ED558CCDull;
        h ^= h >> 29;
    }
};

size_t align64(size_t x) { return (x + 63) & ~(size_t)63; }

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

void sections(const avbd_world* w, const void* ptr[kKinds], int64_t n[kKinds]) {
    ptr[0] = w->scenes; n[0] = w->n_scenes;
    ptr[1] = w->bodies; n[1] = w->n_bodies;
    ptr[2] = w->joints; n[2] = w->n_joints;
    ptr[3] = w->springs; n[3] = w->n_springs;
    ptr[4] = w->manifolds; n[4] = w->n_manifolds;
    ptr[5] = w->tets; n[5] = w->n_tets;
    ptr[6] = w->tris; n[6] = w->n_tris;
}

int io_error(const std::string& what, const char* path) {
    return avbd_set_error(AVBD_E_IO, what + ": " + (path ? path : "(null)"));
}

// Reads and checks the header; on success the file position is at the end
// of the header and *bytes holds the file size.
int read_header(FILE* f, const char* path, Header* h, long long* bytes) {
    if (std::fseek(f, 0, SEEK_END) != 0) return io_error("cannot seek", path);
    *bytes = std::ftell(f);
    if (*bytes < (long long)kHeaderBytes) return io_error("not a world snapshot (too short)", path);
    std::rewind(f);
    if (std::fread(h, sizeof(Header), 1, f) != 1) return io_error("cannot read the snapshot header", path);
    if (std::memcmp(h->magic, kMagic, 8) != 0) return io_error("not a world snapshot (bad magic)", path);
    if (h->version != AVBD_SNAPSHOT_VERSION || h->header_bytes != kHeaderBytes)
        return io_error("unsupported snapshot version", path);
    if (h->abi_version != AVBD_ABI_VERSION || h->kinds != (uint32_t)kKinds)
        return io_error("snapshot written by another ABI version", path);
    for (int k = 0; k < kKinds; ++k)
        if (h->record_bytes[k] != kRecordBytes[k])
            return io_error("snapshot record layout differs from this build's", path);
    size_t at = kHeaderBytes, end = kHeaderBytes;
    for (int k = 0; k < kKinds; ++k) {
        if (h->counts[k] < 0 || h->counts[k] > 0x7FFFFFFFll) return io_error("bad record count in snapshot", path);
        const size_t len = (size_t)h->counts[k] * kRecordBytes[k];
        at = align64(at) + len;
        if (len) end = at;
    }
    if ((long long)end > *bytes) return io_error("truncated snapshot", path);
    if ((long long)end < *bytes) return io_error("trailing bytes after the snapshot", path);
    return 0;
}

}  // namespace

extern "C" {

int avbd_snapshot_write(const char* path, const avbd_world* w) {
    if (!path || !w) return avbd_set_error(AVBD_E_ARG, "null argument");
    const void* ptr[kKinds];
    int64_t n[kKinds];
    sections(w, ptr, n);
    Header h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = AVBD_SNAPSHOT_VERSION;
    h.header_bytes = kHeaderBytes;
    h.abi_version = AVBD_ABI_VERSION;
    h.kinds = kKinds;
    Hash hash;
    for (int k = 0; k < kKinds; ++k) {
        if (n[k] < 0 || (n[k] > 0 && !ptr[k])) return avbd_set_error(AVBD_E_ARG, "world record array missing");
        h.record_bytes[k] = kRecordBytes[k];
        h.counts[k] = n[k];
        hash.feed(ptr[k], (size_t)n[k] * kRecordBytes[k]);
    }
    h.checksum = hash.h;
    std::vector<char> buf(1 << 22);  // outlives the stream that uses it
    File f(std::fopen(path, "wb"));
    if (!f) return io_error("cannot create", path);
    std::setvbuf(f.get(), buf.data(), _IOFBF, buf.size());
    if (std::fwrite(&h, sizeof(Header), 1, f.get()) != 1) return io_error("write failed", path);
    // at: the layout cursor (as the reader computes it); pos: bytes written.
    // Padding is written only in front of a non-empty section, so the file
    // ends with the last record.
    size_t at = kHeaderBytes, pos = kHeaderBytes;
    static const char zeros[64] = {0};
    for (int k = 0; k < kKinds; ++k) {
        const size_t start = align64(at);
        const size_t bytes = (size_t)n[k] * kRecordBytes[k];
        if (bytes) {
            const size_t pad = start - pos;
            if (pad && std::fwrite(zeros, 1, pad, f.get()) != pad) return io_error("write failed", path);
            if (std::fwrite(ptr[k], 1, bytes, f.get()) != bytes) return io_error("write failed", path);
            pos = start + bytes;
        }
        at = start + bytes;
    }
    if (std::fclose(f.release()) != 0) return io_error("write failed", path);
    return 0;
}

int avbd_snapshot_counts(const char* path, avbd_world* counts) {
    if (!path || !counts) return avbd_set_error(AVBD_E_ARG, "null argument");
    File f(std::fopen(path, "rb"));
    if (!f) return io_error("cannot open", path);
    Header h;
    long long bytes = 0;
    int rc = read_header(f.get(), path, &h, &bytes);
    if (rc) return rc;
    counts->n_scenes = (int32_t)h.counts[0];
    counts->n_bodies = (int32_t)h.counts[1];
    counts->n_joints = (int32_t)h.counts[2];
    counts->n_springs = (int32_t)h.counts[3];
    counts->n_manifolds = (int32_t)h.counts[4];
    counts->n_tets = (int32_t)h.counts[5];
    counts->n_tris = (int32_t)h.counts[6];
    return 0;
}

int avbd_snapshot_read(const char* path, avbd_world* out) {
    if (!path || !out) return avbd_set_error(AVBD_E_ARG, "null argument");
    File f(std::fopen(path, "rb"));
    if (!f) return io_error("cannot open", path);
    Header h;
    long long bytes = 0;
    int rc = read_header(f.get(), path, &h, &bytes);
    if (rc) return rc;
    const void* cptr[kKinds];
    int64_t n[kKinds];
    sections(out, cptr, n);
    for (int k = 0; k < kKinds; ++k) {
        if (n[k] != h.counts[k]) return avbd_set_error(AVBD_E_ARG, "world arrays not sized from avbd_snapshot_counts");
        if (n[k] > 0 && !cptr[k]) return avbd_set_error(AVBD_E_ARG, "world record array missing");
    }
    Hash hash;
    size_t at = kHeaderBytes;
    for (int k = 0; k < kKinds; ++k) {
        const size_t start = align64(at);
        const size_t len = (size_t)n[k] * kRecordBytes[k];
        if (len) {
            if (std::fseek(f.get(), (long)start, SEEK_SET) != 0) return io_error("cannot seek", path);
            void* dst = const_cast<void*>(cptr[k]);
            if (std::fread(dst, 1, len, f.get()) != len) return io_error("truncated snapshot", path);
            hash.feed(dst, len);
        }
        at = start + len;
    }
    if (hash.h != h.checksum) return io_error("snapshot checksum mismatch (corrupt file)", path);
    return 0;
}

int avbd_snapshot_save(avbd_ctx* ctx, const char* path) {
    if (!ctx || !path) return avbd_set_error(AVBD_E_ARG, "null argument");
    avbd_world w{};
    int rc = avbd_get_counts(ctx, &w);
    if (rc) return rc;
    std::vector<avbd_scene> sc(w.n_scenes);
    std::vector<avbd_body> bo(w.n_bodies);
    std::vector<avbd_joint> jo(w.n_joints);
    std::vector<avbd_spring> sp(w.n_springs);
    std::vector<avbd_manifold> ma(w.n_manifolds);
    std::vector<avbd_tet> te(w.n_tets);
    std::vector<avbd_tri> tr(w.n_tris);
    w.scenes = sc.data();
    w.bodies = bo.data();
    w.joints = jo.data();
    w.springs = sp.data();
    w.manifolds = ma.data();
    w.tets = te.data();
    w.tris = tr.data();
    if ((rc = avbd_download(ctx, &w))) return rc;
    return avbd_snapshot_write(path, &w);
}

int avbd_snapshot_load(avbd_ctx* ctx, const char* path, int64_t manifold_capacity) {
    if (!ctx || !path) return avbd_set_error(AVBD_E_ARG, "null argument");
    avbd_world w{};
    int rc = avbd_snapshot_counts(path, &w);
    if (rc) return rc;
    std::vector<avbd_scene> sc(w.n_scenes);
    std::vector<avbd_body> bo(w.n_bodies);
    std::vector<avbd_joint> jo(w.n_joints);
    std::vector<avbd_spring> sp(w.n_springs);
    std::vector<avbd_manifold> ma(w.n_manifolds);
    std::vector<avbd_tet> te(w.n_tets);
    std::vector<avbd_tri> tr(w.n_tris);
    w.scenes = sc.data();
    w.bodies = bo.data();
    w.joints = jo.data();
    w.springs = sp.data();
    w.manifolds = ma.data();
    w.tets = te.data();
    w.tris = tr.data();
    if ((rc = avbd_snapshot_read(path, &w))) return rc;
    return avbd_upload(ctx, &w, manifold_capacity);
}

}  // extern "C"
