// topo_check.cpp — builds host topologies (whole scaffold and both copy
// schedules) for the shapes on the command line and prints a digest of each
// schedule.  tests/test_sanitizers.py compiles it with the product's
// topology.cpp under ThreadSanitizer and AddressSanitizer/UBSan and compares
// its digests across thread counts; it is test infrastructure, not product.
//
//   topo_check CB LANES MODE [CB LANES MODE ...]
#include <cstdio>
#include <cstdlib>
#include <string>

#include "topology.h"

namespace {

uint64_t fnv(const void* p, size_t n, uint64_t h) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) { h ^= c[i]; h *= 1099511628211ull; }
    return h;
}

template <class V>
uint64_t fnv_vec(const V& v, uint64_t h) { return fnv(v.data(), v.size() * sizeof(v[0]), h); }

uint64_t digest(const pose::Schedule& s) {
    const uint32_t m[7] = {s.id_lo, s.id_hi, s.rank_lo, s.rank_hi, s.n_steps, s.pool, s.ring_steps};
    uint64_t h = fnv(m, sizeof m, 1469598103934665603ull);
    return fnv_vec(s.near, fnv_vec(s.rec, fnv_vec(s.step_start, h)));
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4 || (argc - 1) % 3 != 0) {
        std::fprintf(stderr, "usage: %s CB LANES MODE [CB LANES MODE ...]\n", argv[0]);
        return 2;
    }
    for (int i = 1; i + 2 < argc; i += 3) {
        const uint32_t cb = (uint32_t)std::strtoul(argv[i], nullptr, 10);
        const uint32_t lanes = (uint32_t)std::strtoul(argv[i + 1], nullptr, 10);
        const int mode = std::atoi(argv[i + 2]);
        pose::HostTopology T;
        std::string err;
        if (pose::build_host_topology(cb, lanes, mode, T, &err, true) != 0) {
            std::printf("%u %u %d error %s\n", cb, lanes, mode, err.c_str());
            return 1;
        }
        const uint32_t m[6] = {T.cb, (uint32_t)T.count, T.lanes, T.levels, T.pool, T.ring_steps};
        uint64_t h = fnv(m, sizeof m, 1469598103934665603ull);
        h = fnv_vec(T.near, fnv_vec(T.rec, fnv_vec(T.step_start, h)));
        std::printf("%u %u %d %016llx %016llx %016llx\n", cb, lanes, mode, (unsigned long long)h,
                    (unsigned long long)digest(T.part[0]), (unsigned long long)digest(T.part[1]));
    }
    return 0;
}
