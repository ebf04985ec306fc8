// avbd_tune.h -- autotuner for the launch shape of the primal sweep.
//
// Two parameters of a bucket launch are tuned.  The lanes per body (G) trade
// per-body latency (more lanes share a body's incidence list) against
// redundant per-body work and reduction cost (every lane of the group solves
// the block system); which wins depends on how much parallel work a launch
// has: a 32^3 FEM block alone wants 16 lanes, the 4096-scene mix wants 4.
// The workgroup size (64, 128 or 256 threads) then sets how many bodies a
// workgroup holds and how the launch spreads over the CUs.  Neither changes
// a body's arithmetic (a lane group is always inside one wave), so the
// choice never changes results.
//
// The tuner keys on (kernel class, log2 of the bodies in the launch).  Class
// 2 (FEM-only buckets with canonical lane sums) samples only the widths that
// kernel has, 4 to 32; there the width is free to change with the launch
// size, because results do not depend on it.  A key
// first samples every candidate width at 256 threads, a few times each, on
// live launches (timed with HIP events on the stream that runs the kernel),
// keeps the fastest, then samples the three workgroup sizes at that width
// and keeps the fastest.  The table can be written to and read from a small
// JSON object so later runs start tuned.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

struct LaneTuner {
    static constexpr int kNumCands = 6;
    static constexpr int kCands[kNumCands] = {1, 2, 4, 8, 16, 32};
    static constexpr int kNumBlocks = 3;
    static constexpr int kBlocks[kNumBlocks] = {64, 128, 256};
    static constexpr int kDefaultBlock = 256;
    static constexpr int kSamples = 2;

    struct Entry {
        int best = 0;       // chosen width, 0 while tuning
        int next = 0;       // next candidate index to sample
        unsigned mask = 0;  // candidates sampled (bit c: kCands[c]); set at first use
        int count[kNumCands] = {0};
        double us_per_body[kNumCands] = {0.0};
        int block = 0;      // chosen workgroup size (after best), 0 while tuning
        int bnext = 0;
        int bcount[kNumBlocks] = {0};
        double bus[kNumBlocks] = {0.0};
    };
    struct Pending {
        int key;
        int cand;   // width candidate (stage 1) or workgroup candidate (stage 2)
        bool block; // stage 2
        int n;
        hipEvent_t a, b;
    };

    bool enabled = true;
    std::map<int, Entry> table;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;

    static int band(int n) {
        int b = 0;
        while ((1 << (b + 1)) <= n && b < 30) ++b;
        return b;
    }
    static int key(int cls, int n) { return cls * 64 + band(n); }

    ~LaneTuner() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }

    hipEvent_t event() {
        if (pool_used == pool.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            pool.push_back(e);
        }
        return pool[pool_used++];
    }

    // Width and workgroup size for a launch of class cls over n bodies.
    // While the key is being tuned, *measure receives the index of a pending
    // record whose events the caller records around the launch; else -1.
    int choose(int cls, int n, int* measure, int* block) {
        *measure = -1;
        *block = kDefaultBlock;
        const int k = key(cls, n);
        Entry& e = table[k];
        if (!enabled) {
            if (e.best && e.block) *block = e.block;
            return e.best ? e.best : 8;
        }
        if (e.best && e.block) {
            *block = e.block;
            return e.best;
        }
        if (e.best) {  // stage 2: workgroup sizes at the chosen width
            const int bi = e.bnext % kNumBlocks;
            e.bnext = bi + 1;
            pending.push_back(Pending{k, bi, true, n, event(), event()});
            *measure = int(pending.size()) - 1;
            *block = kBlocks[bi];
            return e.best;
        }
        if (!e.mask) e.mask = seed_mask(k);
        int ci = e.next % kNumCands;
        while (!((e.mask >> ci) & 1u)) ci = (ci + 1) % kNumCands;
        e.next = ci + 1;
        pending.push_back(Pending{k, ci, false, n, event(), event()});
        *measure = int(pending.size()) - 1;
        return kCands[ci];
    }

    // Widths a class can take: every candidate, 4 to 32 for class 2.
    static unsigned class_mask(int cls) { return cls == 2 ? 0x3Cu : (1u << kNumCands) - 1; }

    // Candidates for a key first seen: next to a tuned key of the same class
    // (launches within a factor of two in size) only that key's width and
    // its half and double, so a bucket that crosses a power of two while the
    // world runs costs a few sampled launches, not a sweep of every width.
    unsigned seed_mask(int k) const {
        const unsigned all = class_mask(k / 64);
        for (int d : {-1, 1}) {
            auto it = table.find(k + d);
            if (it == table.end() || !it->second.best || (k + d) / 64 != k / 64) continue;
            const int b = it->second.best;
            unsigned m = 0;
            for (int c = 0; c < kNumCands; ++c)
                if (kCands[c] == b || 2 * kCands[c] == b || kCands[c] == 2 * b) m |= 1u << c;
            m &= all;
            return m ? m : all;
        }
        return all;
    }

    // The width choose() would return without sampling (and its workgroup
    // size in *block): 0 while the key is still being tuned.
    int peek(int cls, int n, int* block) const {
        auto it = table.find(key(cls, n));
        const int best = it == table.end() ? 0 : it->second.best;
        const int blk = it == table.end() ? 0 : it->second.block;
        *block = blk ? blk : kDefaultBlock;
        if (!enabled) return best ? best : 8;
        return blk ? best : 0;
    }

    // After the stream(s) have been synchronised.
    void collect() {
        for (const Pending& p : pending) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, p.a, p.b) != hipSuccess) continue;
            Entry& e = table[p.key];
            if (p.block) {
                if (e.block) continue;
                e.bcount[p.cand]++;
                e.bus[p.cand] += 1e3 * ms / (p.n > 0 ? p.n : 1);
                bool done = true;
                for (int c = 0; c < kNumBlocks; ++c) done = done && e.bcount[c] >= kSamples;
                if (done) {
                    int bi = 0;
                    for (int c = 1; c < kNumBlocks; ++c)
                        if (e.bus[c] / e.bcount[c] < e.bus[bi] / e.bcount[bi]) bi = c;
                    e.block = kBlocks[bi];
                }
                continue;
            }
            if (e.best) continue;
            e.count[p.cand]++;
            e.us_per_body[p.cand] += 1e3 * ms / (p.n > 0 ? p.n : 1);
            bool done = true;
            for (int c = 0; c < kNumCands; ++c) done = done && (!((e.mask >> c) & 1u) || e.count[c] >= kSamples);
            if (done) {
                int bi = -1;
                for (int c = 0; c < kNumCands; ++c)
                    if (((e.mask >> c) & 1u) &&
                        (bi < 0 || e.us_per_body[c] / e.count[c] < e.us_per_body[bi] / e.count[bi]))
                        bi = c;
                e.best = kCands[bi];
            }
        }
        pending.clear();
        pool_used = 0;
    }

    // {"key": [width, workgroup], ...} for every fully tuned key.
    std::string to_json() const {
        std::string s = "{";
        bool first = true;
        for (const auto& kv : table) {
            if (!kv.second.best || !kv.second.block) continue;
            char buf[96];
            std::snprintf(buf, sizeof(buf), "%s\"%d\": [%d, %d]", first ? "" : ", ", kv.first, kv.second.best,
                          kv.second.block);
            s += buf;
            first = false;
        }
        return s + "}";
    }

    // Parses what to_json writes; a bare width ("key": width, the earlier
    // format) takes the default workgroup size.  Values outside the
    // candidate sets are ignored.
    void from_json(const std::string& s) {
        size_t i = 0;
        while ((i = s.find('"', i)) != std::string::npos) {
            size_t j = s.find('"', i + 1);
            if (j == std::string::npos) break;
            int k = std::atoi(s.substr(i + 1, j - i - 1).c_str());
            size_t c = s.find(':', j);
            if (c == std::string::npos) break;
            size_t v0 = s.find_first_not_of(" \t\n", c + 1);
            if (v0 == std::string::npos) break;
            int v = 0, blk = kDefaultBlock;
            if (s[v0] == '[') {
                v = std::atoi(s.c_str() + v0 + 1);
                size_t comma = s.find(',', v0);
                size_t close = s.find(']', v0);
                if (comma != std::string::npos && close != std::string::npos && comma < close)
                    blk = std::atoi(s.c_str() + comma + 1);
                i = close == std::string::npos ? v0 + 1 : close + 1;
            } else {
                v = std::atoi(s.c_str() + v0);
                i = v0 + 1;
            }
            bool ok = false, okb = false;
            for (int q = 0; q < kNumCands; ++q) ok = ok || (kCands[q] == v && ((class_mask(k / 64) >> q) & 1u));
            for (int q = 0; q < kNumBlocks; ++q) okb = okb || kBlocks[q] == blk;
            if (ok && okb) {
                table[k].best = v;
                table[k].block = blk;
            }
        }
    }
};
