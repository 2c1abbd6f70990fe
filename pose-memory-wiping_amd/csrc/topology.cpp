// topology.cpp — scaffold emission, step scheduling and location assignment.
#include "topology.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "b3_host.h"

namespace pose {

int graph_parameter_n(uint64_t m) {
    if (m <= 1) return 0;
    int bits = 64 - __builtin_clzll(m - 1);
    return bits > 1 ? bits - 1 : 0;
}

static uint64_t butterfly_nodes(int d) { return (uint64_t)(d + 1) << d; }
static uint64_t connected_nodes(int l) {
    return l == 0 ? 1 : 2 * connected_nodes(l - 1) + 2 * butterfly_nodes(l - 1);
}
static uint64_t standalone_nodes(int l) {
    return l == 0 ? 1 : standalone_nodes(l - 1) + butterfly_nodes(l - 1) + connected_nodes(l - 1);
}

uint64_t scaffold_node_count(uint64_t m) {
    return 2 * standalone_nodes(graph_parameter_n(m) + 1);
}

namespace {

// Set of integers in [0, n) with insert, erase and max in O(levels): a bitset
// with one summary bit per 64-bit word, twice.
class HeightSet {
public:
    explicit HeightSet(uint32_t n)
        : l0_((n + 63) / 64, 0), l1_((l0_.size() + 63) / 64, 0), l2_((l1_.size() + 63) / 64, 0) {}
    bool empty() const { return count_ == 0; }
    void insert(uint32_t h) {
        if (l0_[h >> 6] & bit(h)) return;
        ++count_;
        l0_[h >> 6] |= bit(h);
        l1_[h >> 12] |= bit(h >> 6);
        l2_[h >> 18] |= bit(h >> 12);
    }
    void erase(uint32_t h) {
        if (!(l0_[h >> 6] & bit(h))) return;
        --count_;
        l0_[h >> 6] &= ~bit(h);
        if (l0_[h >> 6]) return;
        l1_[h >> 12] &= ~bit(h >> 6);
        if (l1_[h >> 12]) return;
        l2_[h >> 18] &= ~bit(h >> 12);
    }
    uint32_t max() const {   // requires !empty()
        size_t i = l2_.size();
        while (!l2_[--i]) {}
        const uint32_t w1 = (uint32_t)(i * 64 + top(l2_[i]));
        const uint32_t w0 = w1 * 64 + top(l1_[w1]);
        return w0 * 64 + top(l0_[w0]);
    }

private:
    static uint64_t bit(uint32_t h) { return 1ull << (h & 63); }
    static uint32_t top(uint64_t w) { return 63 - (uint32_t)__builtin_clzll(w); }
    std::vector<uint64_t> l0_, l1_, l2_;
    uint64_t count_ = 0;
};

// Node ids are handed out in emission order; the emission order is the graph
// definition (common/src/graph.c:99-520, with the layer-aligned ingress order of
// :414-426).  A "butterfly" here is one first layer -- one predecessor each, or
// (primary[i], in[i]) pairs -- followed by `dim` exchange layers in which node i
// reads the previous layer at i with bit (dim-k) cleared and set.
class Emitter {
public:
    Emitter(uint8_t* np, uint32_t* p0, uint32_t* p1) : np_(np), p0_(p0), p1_(p1) {}
    uint64_t next = 0;

    uint32_t node(int k, uint32_t a = 0, uint32_t b = 0) {
        uint32_t id = (uint32_t)next++;
        if (k == 2 && a > b) std::swap(a, b);
        np_[id] = (uint8_t)k;
        p0_[id] = k >= 1 ? a : 0;
        p1_[id] = k >= 2 ? b : 0;
        return id;
    }

    void butterfly(int dim, const uint32_t* in, const uint32_t* primary, uint32_t* out) {
        const int w = 1 << dim;
        std::vector<uint32_t> a(w), b(w);
        for (int i = 0; i < w; ++i) a[i] = primary ? node(2, primary[i], in[i]) : node(1, in[i]);
        for (int k = 1; k <= dim; ++k) {
            const int bit = 1 << (dim - k);
            for (int i = 0; i < w; ++i) b[i] = node(2, a[i & ~bit], a[i | bit]);
            a.swap(b);
        }
        std::copy(a.begin(), a.end(), out);
    }

    // Ingress butterfly over `ingress`, then a centre butterfly whose first
    // layer pairs `primary` with the ingress outputs.
    void merged(int dim, const uint32_t* primary, const uint32_t* ingress, uint32_t* out) {
        std::vector<uint32_t> g((size_t)1 << dim);
        butterfly(dim, ingress, nullptr, g.data());
        butterfly(dim, g.data(), primary, out);
    }

    void connected(int level, const uint32_t* in, uint32_t* base) {
        if (level == 0) { base[0] = node(1, in[0]); return; }
        const int h = 1 << (level - 1);
        std::vector<uint32_t> centre(h);
        connected(level - 1, in, base);
        merged(level - 1, base, in + h, centre.data());
        connected(level - 1, centre.data(), base + h);
    }

    void standalone(int level, uint32_t* base) {
        if (level == 0) { base[0] = node(0); return; }
        const int h = 1 << (level - 1);
        std::vector<uint32_t> centre(h);
        standalone(level - 1, base);
        butterfly(level - 1, base, nullptr, centre.data());
        connected(level - 1, centre.data(), base + h);
    }

private:
    uint8_t* np_;
    uint32_t* p0_;
    uint32_t* p1_;
};

}  // namespace

int emit_topology(uint32_t cb, U8Buf& np, U32Buf& p0, U32Buf& p1, I32Buf& rank) {
    if (cb == 0 || cb > (1u << 20)) return -1;
    const int n = graph_parameter_n(cb);
    const uint64_t half_w = 1ull << n, retained = cb - half_w;
    const uint64_t count = scaffold_node_count(cb);
    // The emitter writes np, p0 and p1 of every id (absent predecessors as 0),
    // so they are not cleared first; rank is -1 except at outputs.
    np.resize(count);
    p0.resize(count);
    p1.resize(count);
    rank.resize(count);
    // The reference emits the standalone construction twice, the right copy's
    // ids after the left's (common/src/graph.c:1089-1093).  The two emissions
    // write disjoint id ranges, so they run concurrently.
    Emitter e_left(np.data(), p0.data(), p1.data()), e_right(np.data(), p0.data(), p1.data());
    e_right.next = count / 2;
    std::vector<uint32_t> left(2 * half_w), right(2 * half_w);
    std::thread right_copy([&]() {
        std::fill(rank.begin() + (long)(count / 2), rank.end(), -1);
        e_right.standalone(n + 1, right.data());
    });
    std::fill(rank.begin(), rank.begin() + (long)(count / 2), -1);
    e_left.standalone(n + 1, left.data());
    right_copy.join();
    if (e_left.next != count / 2 || e_right.next != count) return -1;
    // Output set: second half of each copy's base, left copy first
    // (common/src/graph.c:1089-1093, :993-998).
    for (uint64_t i = 0; i < half_w; ++i) rank[left[half_w + i]] = (int32_t)i;
    for (uint64_t i = 0; i < retained; ++i) rank[right[half_w + i]] = (int32_t)(half_w + i);
    return 0;
}

namespace {

// The emitted scaffold with its per-node ASAP level and height (longest path
// to a sink), shared by the schedules built from it.
struct Graph {
    U8Buf np;
    U32Buf p0, p1, level, height;
    I32Buf rank;
};

// Host threads for the schedule passes: POSE_GPU_BUILD_THREADS (1..64) if
// set, else the machine's, at most 16.  The schedule does not depend on it.
unsigned build_threads() {
    if (const char* e = std::getenv("POSE_GPU_BUILD_THREADS")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v >= 1 && v <= 64) return (unsigned)v;
    }
    const unsigned h = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, h ? h : 4u));
}

// Runs f(chunk, begin, end) over [0, n) in up to `threads` contiguous chunks
// (fewer for small n); returns the number of chunks.
template <typename F>
unsigned parallel_chunks(uint64_t n, unsigned threads, F f) {
    threads = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(threads, n / 65536 + 1));
    std::vector<std::thread> th;
    for (unsigned t = 1; t < threads; ++t) th.emplace_back([=]() { f(t, n * t / threads, n * (t + 1) / threads); });
    f(0u, 0, n / threads);
    for (std::thread& x : th) x.join();
    return threads;
}

// last = max(last, v) for a word other threads update the same way.
inline void atomic_max_u32(uint32_t* p, uint32_t v) {
    uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
}

// Successor lists of a node range [lo, hi) in range-local ids (CSR), every
// list in ascending order.  Built in parallel: counts and slots by atomic
// increments, then each (short) list sorted, so the result is the one a
// serial pass over ascending node ids gives.
struct Successors {
    U32Buf off, succ;
};

void build_successors(const Graph& g, uint32_t lo, uint32_t hi, unsigned threads, Successors& S) {
    const uint32_t n = hi - lo;
    const uint8_t* np = g.np.data() + lo;
    auto q0 = [&](uint32_t i) { return g.p0[lo + i] - lo; };
    auto q1 = [&](uint32_t i) { return g.p1[lo + i] - lo; };
    U32Buf& off = S.off;
    off.resize((uint64_t)n + 1);
    parallel_chunks((uint64_t)n + 1, threads, [&](unsigned, uint64_t i0, uint64_t i1) {
        for (uint64_t i = i0; i < i1; ++i) off[i] = 0;
    });
    parallel_chunks(n, threads, [&](unsigned, uint64_t i0, uint64_t i1) {
        for (uint32_t i = (uint32_t)i0; i < i1; ++i) {
            if (np[i] >= 1) __atomic_fetch_add(&off[q0(i) + 1], 1u, __ATOMIC_RELAXED);
            if (np[i] == 2) __atomic_fetch_add(&off[q1(i) + 1], 1u, __ATOMIC_RELAXED);
        }
    });
    for (uint32_t i = 0; i < n; ++i) off[i + 1] += off[i];
    U32Buf cur(n);
    parallel_chunks(n, threads, [&](unsigned, uint64_t i0, uint64_t i1) {
        for (uint64_t i = i0; i < i1; ++i) cur[i] = off[i];
    });
    S.succ.resize(off[n]);
    uint32_t* succ = S.succ.data();
    parallel_chunks(n, threads, [&](unsigned, uint64_t i0, uint64_t i1) {
        for (uint32_t i = (uint32_t)i0; i < i1; ++i) {
            if (np[i] >= 1) succ[__atomic_fetch_add(&cur[q0(i)], 1u, __ATOMIC_RELAXED)] = i;
            if (np[i] == 2) succ[__atomic_fetch_add(&cur[q1(i)], 1u, __ATOMIC_RELAXED)] = i;
        }
    });
    parallel_chunks(n, threads, [&](unsigned, uint64_t v0, uint64_t v1) {
        for (uint64_t v = v0; v < v1; ++v) {
            uint32_t* a = succ + off[v];
            const uint32_t len = off[v + 1] - off[v];
            for (uint32_t x = 1; x < len; ++x)   // insertion sort: lists are short
                for (uint32_t y = x; y > 0 && a[y - 1] > a[y]; --y) std::swap(a[y - 1], a[y]);
        }
    });
}

// A schedule order of a node range: the range-local node at each position and
// the step start offsets (n_steps + 1 entries, the last one the range size).
struct Order {
    std::vector<uint32_t> order, steps;
};

// Orders the nodes [lo, hi) of g, a range closed under predecessors, into
// steps of at most `lanes` nodes: dependency levels split into steps, or a
// list schedule that fills each step with the ready nodes of greatest height.
// List mode uses `sc`, the range's successor lists, when given.
int order_range(const Graph& g, uint32_t lo, uint32_t hi, uint32_t lanes, int mode, Order& O,
                const char** err, const Successors* sc = nullptr) {
    auto fail = [&](const char* m) { *err = m; return -1; };
    const uint32_t n = hi - lo;
    const uint8_t* np = g.np.data() + lo;
    for (uint32_t i = 0; i < n; ++i)
        if ((np[i] >= 1 && g.p0[lo + i] < lo) || (np[i] == 2 && g.p1[lo + i] < lo))
            return fail("node range not closed under predecessors");
    std::vector<uint32_t>& order = O.order;
    std::vector<uint32_t>& steps = O.steps;
    order.clear();
    steps.clear();
    order.reserve(n);
    if (mode == kSchedLevels) {
        uint32_t maxlevel = 0;
        for (uint32_t i = 0; i < n; ++i) maxlevel = std::max(maxlevel, g.level[lo + i]);
        std::vector<uint32_t> cnt(maxlevel + 2, 0);
        for (uint32_t i = 0; i < n; ++i) cnt[g.level[lo + i] + 1]++;
        for (uint32_t l = 0; l <= maxlevel; ++l) cnt[l + 1] += cnt[l];
        order.resize(n);
        std::vector<uint32_t> cur(cnt.begin(), cnt.end() - 1);
        for (uint32_t i = 0; i < n; ++i) order[cur[g.level[lo + i]]++] = i;
        for (uint32_t l = 0; l <= maxlevel; ++l)
            for (uint32_t s = cnt[l]; s < cnt[l + 1]; s += lanes) steps.push_back(s);
    } else if (mode == kSchedList) {
        Successors own;
        if (!sc) {
            build_successors(g, lo, hi, 1, own);
            sc = &own;
        }
        const uint32_t* soff = sc->off.data();
        const uint32_t* succ = sc->succ.data();
        // Ready nodes in per-height FIFO queues; each step takes up to `lanes`
        // ready nodes, highest first.  Nodes made ready by a step join the
        // queues only after it closes.  The non-empty heights are kept in a
        // three-level bitset so the highest one is found in O(1).
        const uint32_t* height = g.height.data() + lo;
        const uint32_t H = n ? *std::max_element(height, height + n) : 0;
        constexpr uint32_t NIL = 0xFFFFFFFFu;
        std::vector<uint32_t> qhead(H + 1, NIL), qtail(H + 1, NIL), link(n, NIL);
        std::vector<uint8_t> remaining(np, np + n);
        HeightSet hs(H + 1);
        auto push = [&](uint32_t v) {
            const uint32_t h = height[v];
            link[v] = NIL;
            if (qtail[h] == NIL) { qhead[h] = v; hs.insert(h); }
            else link[qtail[h]] = v;
            qtail[h] = v;
        };
        for (uint32_t i = 0; i < n; ++i) if (np[i] == 0) push(i);
        std::vector<uint32_t> fresh;
        uint64_t placed = 0;
        while (placed < n) {
            if (hs.empty()) return fail("schedule stalled");
            steps.push_back((uint32_t)order.size());
            fresh.clear();
            for (uint32_t k = 0; k < lanes && !hs.empty(); ++k) {
                const uint32_t h = hs.max();
                const uint32_t v = qhead[h];
                qhead[h] = link[v];
                if (qhead[h] == NIL) { qtail[h] = NIL; hs.erase(h); }
                order.push_back(v);
                ++placed;
                for (uint32_t e = soff[v]; e < soff[v + 1]; ++e)
                    if (--remaining[succ[e]] == 0) fresh.push_back(succ[e]);
            }
            for (uint32_t v : fresh) push(v);
        }
    } else {
        return fail("unknown schedule mode");
    }
    steps.push_back(n);
    return 0;
}

// Places an ordered range [lo, hi) of g: assigns every node a location and
// writes the padded step records and the near table.  Node ids in the records
// are the scaffold ids; outputs keep their absolute rank; scratch slots are
// numbered from 0 within the range's own pool.  Every pass but the recycling
// of long-lived slots runs on `threads` host threads; the result does not
// depend on their number.
int place_range(const Graph& g, uint32_t lo, uint32_t hi, uint32_t lanes, const Order& O, Schedule& S,
                const char** err, unsigned threads) {
    auto fail = [&](const char* m) { *err = m; return -1; };
    const uint32_t n = hi - lo;
    const uint8_t* np = g.np.data() + lo;
    const int32_t* rank = g.rank.data() + lo;
    auto q0 = [&](uint32_t i) { return g.p0[lo + i] - lo; };
    auto q1 = [&](uint32_t i) { return g.p1[lo + i] - lo; };
    const std::vector<uint32_t>& order = O.order;
    const std::vector<uint32_t>& steps = O.steps;
    if (order.size() != n || steps.empty() || steps.back() != n) return fail("schedule order does not cover its range");
    const uint32_t NS = (uint32_t)steps.size() - 1;

    // Step and lane of each node (the order is a permutation of the range),
    // then the last step that reads it.
    U32Buf step_of(n), lane_of(n), last_use(n);
    parallel_chunks(NS, threads, [&](unsigned, uint64_t s0, uint64_t s1) {
        for (uint32_t s = (uint32_t)s0; s < s1; ++s)
            for (uint32_t k = steps[s]; k < steps[s + 1]; ++k) {
                step_of[order[k]] = s;
                lane_of[order[k]] = k - steps[s];
            }
    });
    parallel_chunks(n, threads, [&](unsigned, uint64_t i0, uint64_t i1) {
        for (uint32_t i = (uint32_t)i0; i < i1; ++i) last_use[i] = step_of[i];
    });
    parallel_chunks(n, threads, [&](unsigned, uint64_t i0, uint64_t i1) {
        for (uint32_t i = (uint32_t)i0; i < i1; ++i) {
            if (np[i] >= 1) atomic_max_u32(&last_use[q0(i)], step_of[i]);
            if (np[i] == 2) atomic_max_u32(&last_use[q1(i)], step_of[i]);
        }
    });

    // Scratch slots for non-output nodes.  Short-lived nodes (last read at
    // most L steps after they are written) take slot (step mod (L+1))*lanes +
    // lane in a ring of step pages, so a step's labels sit side by side and the
    // next layer's predecessor reads coalesce into whole cache lines.  The
    // rest take individually recycled slots above the ring: a slot freed after
    // step s goes to a node born in step s+1 or later (allocation of a step
    // precedes its frees, so a step never overwrites a label it still reads).
    // L is the power of two (256..8192) that covers 90% of scratch reads.
    uint32_t L = 256;
    {
        std::vector<std::vector<uint64_t>> hist(threads, std::vector<uint64_t>(64, 0));
        const unsigned used = parallel_chunks(n, threads, [&](unsigned t, uint64_t i0, uint64_t i1) {
            uint64_t* h = hist[t].data();
            auto add_read = [&](uint32_t pred) {
                if (rank[pred] >= 0) return;
                const uint32_t life = last_use[pred] - step_of[pred];
                h[life ? 64 - __builtin_clzll(life) : 0]++;
            };
            for (uint32_t i = (uint32_t)i0; i < i1; ++i) {
                if (np[i] >= 1) add_read(q0(i));
                if (np[i] == 2) add_read(q1(i));
            }
        });
        std::vector<uint64_t> sum(64, 0);
        uint64_t total = 0;
        for (unsigned t = 0; t < used; ++t)
            for (int bkt = 0; bkt < 64; ++bkt) { sum[bkt] += hist[t][bkt]; total += hist[t][bkt]; }
        uint64_t acc = 0;
        for (int bkt = 0; bkt < 64; ++bkt) {
            acc += sum[bkt];
            if (acc * 10 >= total * 9) { L = bkt >= 1 ? (1u << bkt) : 1u; break; }
        }
        L = std::min<uint32_t>(8192, std::max<uint32_t>(256, L));
    }
    const uint32_t ring = L + 1;
    const uint64_t long_base = (uint64_t)kFirstSlot + (uint64_t)ring * lanes;
    if (long_base >= kPersistBit) return fail("scratch ring too large");
    U32Buf loc(n);
    {
        auto is_long = [&](uint32_t v) { return rank[v] < 0 && last_use[v] - step_of[v] > L; };
        // Outputs and ring-page nodes in parallel; positions of long-lived
        // nodes marked for the serial pass below.
        std::vector<uint8_t> long_at(n);
        parallel_chunks(NS, threads, [&](unsigned, uint64_t s0, uint64_t s1) {
            for (uint32_t s = (uint32_t)s0; s < s1; ++s)
                for (uint32_t k = steps[s]; k < steps[s + 1]; ++k) {
                    const uint32_t v = order[k];
                    long_at[k] = 0;
                    if (rank[v] >= 0) loc[v] = kPersistBit | (uint32_t)rank[v];
                    else if (!is_long(v)) loc[v] = kFirstSlot + (s % ring) * lanes + (k - steps[s]);
                    else long_at[k] = 1;
                }
        });
        std::vector<uint32_t> fcnt(NS + 1, 0);
        for (uint32_t i = 0; i < n; ++i) if (is_long(i)) fcnt[last_use[i] + 1]++;
        for (uint32_t s = 0; s < NS; ++s) fcnt[s + 1] += fcnt[s];
        std::vector<uint32_t> fnodes(fcnt[NS]);
        {
            std::vector<uint32_t> cur(fcnt.begin(), fcnt.end() - 1);
            for (uint32_t i = 0; i < n; ++i) if (is_long(i)) fnodes[cur[last_use[i]]++] = i;
        }
        std::vector<uint32_t> freelist;
        uint64_t next_slot = long_base;
        for (uint32_t s = 0; s < NS; ++s) {
            for (uint32_t k = steps[s]; k < steps[s + 1]; ++k) {
                if (!long_at[k]) continue;
                const uint32_t v = order[k];
                if (!freelist.empty()) { loc[v] = freelist.back(); freelist.pop_back(); }
                else loc[v] = (uint32_t)next_slot++;
            }
            // Free in reverse so the next step pops slots in ascending order.
            for (uint32_t k = fcnt[s + 1]; k-- > fcnt[s];) freelist.push_back(loc[fnodes[k]]);
        }
        if (next_slot >= kPersistBit) return fail("scratch pool too large");
        S.pool = (uint32_t)next_slot;
        S.ring_steps = L;
    }

    // Records and near table, rows built in parallel (every row is independent;
    // every word is written here).  lanes <= 1024, so a lane index never
    // collides with kFarPred.
    S.rec.resize(4 * (uint64_t)(NS + 1) * lanes);
    S.near.resize((uint64_t)(NS + 1) * lanes);
    parallel_chunks(NS + 1, threads, [&](unsigned, uint64_t s0, uint64_t s1) {
        for (uint32_t s = (uint32_t)s0; s < s1; ++s) {
            for (uint32_t t = 0; t < lanes; ++t) {
                uint32_t* r = &S.rec[4 * ((uint64_t)s * lanes + t)];
                uint32_t& nr = S.near[(uint64_t)s * lanes + t];
                nr = kFarPred | (kFarPred << 16);
                const uint32_t k = s < NS ? steps[s] + t : 0;
                if (s == NS || k >= steps[s + 1]) {
                    r[0] = kDiscardSlot; r[1] = kZeroSlot; r[2] = kZeroSlot; r[3] = kEmptyRecord;
                    continue;
                }
                const uint32_t v = order[k];
                r[0] = loc[v];
                r[1] = np[v] >= 1 ? loc[q0(v)] : kZeroSlot;
                r[2] = np[v] == 2 ? loc[q1(v)] : kZeroSlot;
                r[3] = ((lo + v) << 2) | np[v];
                if (s == 0) continue;
                uint32_t h0 = kFarPred, h1 = kFarPred;
                if (np[v] >= 1 && step_of[q0(v)] + 1 == s) h0 = lane_of[q0(v)];
                if (np[v] == 2 && step_of[q1(v)] + 1 == s) h1 = lane_of[q1(v)];
                nr = h0 | (h1 << 16);
            }
        }
    });
    // Output ranks of the range: one contiguous run.
    uint32_t rlo = 0xFFFFFFFFu, rhi = 0, nout = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (rank[i] < 0) continue;
        rlo = std::min(rlo, (uint32_t)rank[i]);
        rhi = std::max(rhi, (uint32_t)rank[i] + 1);
        ++nout;
    }
    if (!nout) rlo = rhi = 0;
    if (rhi - rlo != nout) return fail("output ranks of a node range are not contiguous");
    S.id_lo = lo;
    S.id_hi = hi;
    S.rank_lo = rlo;
    S.rank_hi = rhi;
    S.n_steps = NS;
    S.step_start.assign(steps.begin(), steps.end());
    return 0;
}

// Copy 1's schedule from copy 0's when copy 1 is copy 0 shifted by `id_shift`
// node ids and `rank_shift` output ranks (same order, same steps, same slots):
// records keep their slots and shift their node ids and output ranks.
void shift_schedule(const Schedule& a, uint32_t id_shift, uint32_t rank_shift, Schedule& b, unsigned threads) {
    b.id_lo = a.id_lo + id_shift;
    b.id_hi = a.id_hi + id_shift;
    b.rank_lo = a.rank_lo + rank_shift;
    b.rank_hi = a.rank_hi + rank_shift;
    b.n_steps = a.n_steps;
    b.pool = a.pool;
    b.ring_steps = a.ring_steps;
    b.step_start = a.step_start;
    b.rec.resize(a.rec.size());
    b.near.resize(a.near.size());
    parallel_chunks(a.near.size(), threads, [&](unsigned, uint64_t p0, uint64_t p1) {
        for (uint64_t p = p0; p < p1; ++p) {
            const uint32_t* r = &a.rec[4 * p];
            uint32_t* w = &b.rec[4 * p];
            if (r[3] == kEmptyRecord) {
                w[0] = r[0]; w[1] = r[1]; w[2] = r[2]; w[3] = r[3];
            } else {
                for (int c = 0; c < 3; ++c) w[c] = (r[c] & kPersistBit) ? r[c] + rank_shift : r[c];
                w[3] = r[3] + (id_shift << 2);
            }
            b.near[p] = a.near[p];
        }
    });
}

// True when the copy of ids [half, 2 half) is the copy of ids [0, half)
// shifted by half: the same predecessor counts, every predecessor shifted.
// Both copies come from the same standalone construction, so this holds; it
// is checked because the schedules below rely on it.
bool copies_mirror(const Graph& g, uint32_t half) {
    if (g.np.size() != 2 * (uint64_t)half) return false;
    for (uint32_t i = 0; i < half; ++i) {
        const uint8_t k = g.np[i];
        if (g.np[half + i] != k) return false;
        if (k >= 1 && g.p0[half + i] != g.p0[i] + half) return false;
        if (k == 2 && g.p1[half + i] != g.p1[i] + half) return false;
    }
    return true;
}

// The whole scaffold's order from copy 0's order at half the width: step s
// holds copy 0's step s followed by the same nodes of copy 1.  By symmetry
// this is what a list schedule of the whole scaffold does (it takes the same
// number of steps: 568,762 at 2^17 and 64 lanes either way).
void interleave_copies(const Order& o, uint32_t half, Order& w, unsigned threads) {
    const uint32_t NS = (uint32_t)o.steps.size() - 1;
    w.order.resize(2 * o.order.size());
    w.steps.resize(NS + 1);
    for (uint32_t s = 0; s <= NS; ++s) w.steps[s] = 2 * o.steps[s];
    parallel_chunks(NS, threads, [&](unsigned, uint64_t s0, uint64_t s1) {
        for (uint32_t s = (uint32_t)s0; s < s1; ++s) {
            const uint32_t a = o.steps[s], b = o.steps[s + 1];
            for (uint32_t k = a; k < b; ++k) {
                w.order[2 * a + (k - a)] = o.order[k];
                w.order[a + b + (k - a)] = o.order[k] + half;
            }
        }
    });
}

}  // namespace

int build_host_topology(uint32_t cb, uint32_t lanes, int mode, HostTopology& T,
                        std::string* err, bool with_parts) {
    auto t0 = std::chrono::steady_clock::now();
    auto fail = [&](const char* m) { if (err) *err = m; return -1; };
    if (lanes == 0 || lanes > 1024) return fail("lanes must be in 1..1024");
    if (mode != kSchedLevels && mode != kSchedList) return fail("unknown schedule mode");
    Graph g;
    if (emit_topology(cb, g.np, g.p0, g.p1, g.rank) != 0) return fail("chunk_blocks out of range");
    const uint64_t N = g.np.size();
    if (N >= (1ull << 30)) return fail("scaffold too large for 30-bit node ids");
    const uint32_t half = (uint32_t)(N / 2);
    const unsigned threads = build_threads();
    // A failed host allocation ends that schedule with an error, not the process.
    auto guarded = [](auto&& f, const char** e) {
        try {
            return f();
        } catch (const std::exception&) {
            *e = "out of host memory while scheduling";
            return -1;
        }
    };

    // The whole scaffold, and (optionally) each of its two copies on its own:
    // copy 0 is ids [0, N/2), copy 1 [N/2, N), and no edge joins them
    // (emit_topology emits two standalone graphs).  Copy 1 mirrors copy 0, so
    // one list schedule of copy 0 serves both copies (range-local orders are
    // equal), and copy 0's schedule at half the width, interleaved with its
    // mirror, is the whole scaffold's.
    const bool mirror = mode == kSchedList && lanes % 2 == 0 && half > 0 && copies_mirror(g, half);
    // With the same order, copy 1 is placed exactly as copy 0 when its output
    // ranks are copy 0's shifted by one constant (every power-of-two
    // chunk_blocks); its schedule is then copy 0's, shifted.
    uint32_t rank_shift = 0;
    bool shift_part1 = with_parts && mirror;
    if (shift_part1) {
        int32_t top = -1;
        for (uint32_t i = 0; i < half; ++i) top = std::max(top, g.rank[i]);
        rank_shift = (uint32_t)(top + 1);
        for (uint32_t i = 0; i < half && shift_part1; ++i)
            shift_part1 = g.rank[half + i] == (g.rank[i] < 0 ? -1 : g.rank[i] + (int32_t)rank_shift);
    }

    // ASAP level and longest path to a sink ("height"); ids are topological.
    // The two copies are independent, so each pass runs per copy, concurrently.
    // With mirrored copies only copy 0's are read (its orders serve both), and
    // copy 0's successor lists are built meanwhile.
    g.level.resize(N);
    g.height.resize(N);
    uint32_t maxlevel_part[2] = {0, 0};
    Successors sc0;   // copy 0's successor lists, shared by its two list schedules
    {
        auto passes = [&](uint32_t lo, uint32_t hi, uint32_t* maxlevel) {
            std::fill(g.height.begin() + lo, g.height.begin() + hi, 0u);
            for (uint32_t i = lo; i < hi; ++i) {
                uint32_t lv = 0;
                if (g.np[i] >= 1) lv = g.level[g.p0[i]] + 1;
                if (g.np[i] == 2) lv = std::max(lv, g.level[g.p1[i]] + 1);
                g.level[i] = lv;
                *maxlevel = std::max(*maxlevel, lv);
            }
            for (uint32_t i = hi; i-- > lo;) {
                if (g.np[i] >= 1) g.height[g.p0[i]] = std::max(g.height[g.p0[i]], g.height[i] + 1);
                if (g.np[i] == 2) g.height[g.p1[i]] = std::max(g.height[g.p1[i]], g.height[i] + 1);
            }
        };
        if (mirror) {
            const char* e_sc = nullptr;
            int rc_sc = 0;
            std::thread succ([&]() {
                rc_sc = guarded([&]() { build_successors(g, 0, half, threads, sc0); return 0; }, &e_sc);
            });
            passes(0, half, &maxlevel_part[0]);
            // Copy 1's levels and heights are copy 0's (kept for completeness).
            parallel_chunks(half, threads, [&](unsigned, uint64_t i0, uint64_t i1) {
                for (uint64_t i = i0; i < i1; ++i) {
                    g.level[half + i] = g.level[i];
                    g.height[half + i] = g.height[i];
                }
            });
            maxlevel_part[1] = maxlevel_part[0];
            succ.join();
            if (rc_sc != 0) return fail(e_sc);
        } else {
            // Copy 0's predecessors precede it (ids are topological); the passes
            // run per copy only when no predecessor of copy 1 lies in copy 0.
            bool closed = true;
            for (uint32_t i = half; i < N && closed; ++i)
                closed = !((g.np[i] >= 1 && g.p0[i] < half) || (g.np[i] == 2 && g.p1[i] < half));
            if (closed) {
                std::thread right([&]() { passes(half, (uint32_t)N, &maxlevel_part[1]); });
                passes(0, half, &maxlevel_part[0]);
                right.join();
            } else {
                passes(0, (uint32_t)N, &maxlevel_part[0]);
            }
        }
    }
    const uint32_t maxlevel = std::max(maxlevel_part[0], maxlevel_part[1]);

    Schedule whole;
    const char* e_step[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int rc_step[6] = {0, 0, 0, 0, 0, 0};   // orders 0-2, placements 3-5 (whole, copy 0, copy 1)
    if (mirror) {
        // Two chains run concurrently: copy 0's order at full width, its
        // placement and copy 1's (or the shifted copy); and copy 0's order at
        // half width, interleaved into the whole scaffold's order and placed.
        Order o_copy, o_half, o_whole;
        std::thread copies;
        if (with_parts)
            copies = std::thread([&]() {
                rc_step[1] = guarded([&]() { return order_range(g, 0, half, lanes, mode, o_copy, &e_step[1], &sc0); },
                                     &e_step[1]);
                if (rc_step[1] != 0) return;
                rc_step[4] = guarded([&]() { return place_range(g, 0, half, lanes, o_copy, T.part[0], &e_step[4], threads); },
                                     &e_step[4]);
                if (rc_step[4] != 0) return;
                if (shift_part1)
                    rc_step[5] = guarded([&]() { shift_schedule(T.part[0], half, rank_shift, T.part[1], threads); return 0; },
                                         &e_step[5]);
                else
                    rc_step[5] = guarded([&]() { return place_range(g, half, (uint32_t)N, lanes, o_copy, T.part[1], &e_step[5],
                                                                    threads); },
                                         &e_step[5]);
            });
        rc_step[0] = guarded([&]() { return order_range(g, 0, half, lanes / 2, mode, o_half, &e_step[0], &sc0); },
                             &e_step[0]);
        if (rc_step[0] == 0)
            rc_step[0] = guarded([&]() { interleave_copies(o_half, half, o_whole, threads); return 0; }, &e_step[0]);
        if (rc_step[0] == 0)
            rc_step[3] = guarded([&]() { return place_range(g, 0, (uint32_t)N, lanes, o_whole, whole, &e_step[3], threads); },
                                 &e_step[3]);
        if (copies.joinable()) copies.join();
    } else {
        // The orders concurrently, then the placements concurrently.
        Order o_whole, o_copy[2];
        {
            std::vector<std::thread> th;
            th.emplace_back([&]() {
                rc_step[0] = guarded([&]() { return order_range(g, 0, (uint32_t)N, lanes, mode, o_whole, &e_step[0]); },
                                     &e_step[0]);
            });
            if (with_parts) {
                th.emplace_back([&]() {
                    rc_step[2] = guarded([&]() { return order_range(g, half, (uint32_t)N, lanes, mode, o_copy[1], &e_step[2]); },
                                         &e_step[2]);
                });
                rc_step[1] = guarded([&]() { return order_range(g, 0, half, lanes, mode, o_copy[0], &e_step[1]); },
                                     &e_step[1]);
            }
            for (std::thread& x : th) x.join();
        }
        for (int i = 0; i < 3; ++i) if (rc_step[i] != 0) return fail(e_step[i]);
        std::vector<std::thread> th;
        if (with_parts) {
            for (uint32_t c = 0; c < 2; ++c)
                th.emplace_back([&, c]() {
                    rc_step[4 + c] = guarded([&]() { return place_range(g, c ? half : 0, c ? (uint32_t)N : half, lanes,
                                                                        o_copy[c], T.part[c], &e_step[4 + c], threads); },
                                             &e_step[4 + c]);
                });
        }
        rc_step[3] = guarded([&]() { return place_range(g, 0, (uint32_t)N, lanes, o_whole, whole, &e_step[3], threads); },
                             &e_step[3]);
        for (std::thread& x : th) x.join();
    }
    for (int i = 0; i < 6; ++i) if (rc_step[i] != 0) return fail(e_step[i]);
    T.has_parts = with_parts;
    if (!with_parts) T.part[0] = T.part[1] = Schedule();
    T.part1_shifted = with_parts && shift_part1;
    T.part1_id_shift = T.part1_shifted ? half : 0;
    T.part1_rank_shift = T.part1_shifted ? rank_shift : 0;

    T.cb = cb;
    T.n = graph_parameter_n(cb);
    T.count = N;
    T.lanes = lanes;
    T.mode = mode;
    T.levels = maxlevel + 1;
    T.n_steps = whole.n_steps;
    T.pool = whole.pool;
    T.ring_steps = whole.ring_steps;
    T.step_start = std::move(whole.step_start);
    T.rec = std::move(whole.rec);
    T.near = std::move(whole.near);
    graph_descriptor(T.desc, cb, (uint64_t)T.n);
    node_keys(T.key_src, T.key_int);
    T.build_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

}  // namespace pose
