// wipe_bench.cpp — single-GPU producer: fill a target (file or block device)
// with PoSE-DB labels from the MI355X labeler and report throughput.
//
// Producer interface of the reference's tools/disk-wipe-bench (flags of
// disk_wipe_bench.cu:711-744; worker :563-708; verify :300-366; verify-only
// :899-920).  Several --disk targets are wiped in turn, largest first, as the
// reference's single team does with one GPU; --gpus takes one ordinal (one
// process drives one GPU), and --mode / --gpus-per-disk are accepted with the
// reference's validation, both meaning one GPU per disk here.  Labeling, device->host copies and writes are
// overlapped by the label server (pose_gpu_server_*): the target range is
// submitted as window-sized requests and each finished super-chunk is written
// at disk_off + j*cb*32 through pinned memory, O_DIRECT when the target
// allows it.  --verify / --verify-only recompute sampled super-chunks with the
// reference CPU library's pose_graph_label_region_pooled_inplace
// (common/include/graph.h:227-234), loaded from --cpu-lib, else from
// $POSE_COMMON_LIB, else as libpose_common.so from the loader's search path;
// it is loaded before anything is written, so a missing library ends the run
// before the target is touched.  --challenge N
// spot-checks N random 32-byte blocks the way a PoSE verifier does: each block
// read back from the target must equal the GPU's challenge answer
// (pose_gpu_challenge_batch) for that absolute block.
#include <dlfcn.h>
#include <fcntl.h>
#include <sys/ioctl.h>
#include <sys/stat.h>
#include <unistd.h>
#include <linux/fs.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gpu_label.h"

namespace {

constexpr uint64_t MiB = 1ull << 20;

// Fixed benchmark seed, as in the reference producer ("DiskWipe" ...); timing is
// seed-independent.  --seed-hex overrides it.
uint8_t g_seed[32] = {0x44, 0x69, 0x73, 0x6b, 0x57, 0x69, 0x70, 0x65, 0x01, 0x23, 0x45,
                      0x67, 0x89, 0xab, 0xcd, 0xef, 0xfe, 0xed, 0xfa, 0xce, 0x0b, 0xad,
                      0xf0, 0x0d, 0xaa, 0xbb, 0xcc, 0xdd, 0xee, 0xff, 0x00, 0x11};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Config {
    std::vector<std::string> disks;   // --disk, repeatable
    int gpu = -1;                     // --gpus ordinal (one GPU per process)
    uint32_t chunk_blocks = 4096;
    uint64_t per_disk_bytes = 0;
    uint64_t hbm_window_mib = 0;
    uint64_t stage_mib = 0;        // 0: sized from the launch size (stage_bytes)
    bool dry_run = false, confirm = false, verify_only = false, verbose = false;
    std::string json;                 // --json PATH ("-" = stdout): machine-readable results
    uint64_t verify_n = 0, challenge_n = 0;
    std::string cpu_lib;
    int cpu_threads = 16;
};

void usage(const char* a0) {
    std::fprintf(stderr,
        "Usage: %s --disk PATH [--disk PATH ...] [options]\n"
        "Fills each PATH (block device or file) with MI355X-generated PoSE-DB labels,\n"
        "one target after another (largest first) on one GPU.\n"
        "  --gpus N             GPU ordinal (one per process; default: the current device)\n"
        "  --mode MODE          disk-per-gpu | gpus-per-disk (the same with one GPU)\n"
        "  --gpus-per-disk K    accepted; one GPU per disk here\n"
        "  --chunk-blocks N     power of two 1..2^20 (default 4096)\n"
        "  --per-disk-mib N     cap bytes written (default: whole target)\n"
        "  --hbm-window-mib N   device arena for the label server (default: a third of free HBM)\n"
        "  --stage-mib N        pinned host staging (default: room for two of the label\n"
        "                       server's largest launches, 1..8 GiB, at most a quarter of RAM)\n"
        "  --dry-run            label + device->host copy only, no writes\n"
        "  --confirm            required for real (destructive) writes\n"
        "  --verify N           read back N sampled super-chunks and CPU-recompute them\n"
        "  --verify-only        only verify an existing wipe (read-only)\n"
        "  --challenge N        read back N random 32-byte blocks and check them against\n"
        "                       GPU challenge answers (no CPU library needed)\n"
        "  --cpu-lib PATH       CPU library exporting the reference graph.h API (for --verify;\n"
        "                       default $POSE_COMMON_LIB, else libpose_common.so on the loader path)\n"
        "  --cpu-threads N      its POSE_REGION_THREADS (default 16)\n"
        "  --seed-hex HEX       64 hex chars = 32-byte session seed\n"
        "  --verbose            per-window progress lines\n"
        "  --json PATH          write the run's results as one JSON object to PATH (- = stdout)\n", a0);
}

int parse_hex(const char* hex, uint8_t out[32]) {
    if (std::strlen(hex) != 64) return -1;
    for (int i = 0; i < 32; ++i) {
        unsigned v;
        if (std::sscanf(hex + 2 * i, "%2x", &v) != 1) return -1;
        out[i] = (uint8_t)v;
    }
    return 0;
}

int pread_all(int fd, uint8_t* buf, size_t len, off_t off) {
    size_t done = 0;
    while (done < len) {
        ssize_t n = pread(fd, buf + done, len - done, off + (off_t)done);
        if (n < 0) { if (errno == EINTR) continue; return -1; }
        if (n == 0) { errno = EIO; return -1; }
        done += (size_t)n;
    }
    return 0;
}

// The reference CPU labeler, resolved at run time.
struct CpuLib {
    void* h = nullptr;
    size_t (*scratch_bytes)(uint32_t) = nullptr;
    void* (*pool_create)(void) = nullptr;
    void (*pool_destroy)(void*) = nullptr;
    int (*label)(const uint8_t*, size_t, void*, size_t, uint64_t, uint8_t**, void*, int, uint32_t) = nullptr;
    std::string path;
    bool load(const std::string& p) {
        path = p;
        h = dlopen(p.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) return false;
        scratch_bytes = (size_t (*)(uint32_t))dlsym(h, "pose_graph_scratch_bytes_inplace");
        pool_create = (void* (*)(void))dlsym(h, "pose_graph_pool_create");
        pool_destroy = (void (*)(void*))dlsym(h, "pose_graph_pool_destroy");
        label = (int (*)(const uint8_t*, size_t, void*, size_t, uint64_t, uint8_t**, void*, int, uint32_t))
            dlsym(h, "pose_graph_label_region_pooled_inplace");
        return scratch_bytes && pool_create && pool_destroy && label;
    }
    // --cpu-lib, else $POSE_COMMON_LIB, else libpose_common.so (loader path).
    bool load_default(const std::string& flag) {
        if (!flag.empty()) return load(flag);
        const char* env = std::getenv("POSE_COMMON_LIB");
        return load(env && *env ? std::string(env) : std::string("libpose_common.so"));
    }
    ~CpuLib() {
        if (h) dlclose(h);
    }
};

// Sampled read-back + CPU recompute (the reference's verify_disk).  Samples are
// evenly spaced over the wiped range, first and last always included.
int verify(int fd, uint64_t len, uint32_t cb, uint64_t samples, const Config& cfg, const CpuLib& lib,
           uint64_t* checked, uint64_t* mism) {
    *checked = *mism = 0;
    const uint64_t sb = (uint64_t)cb * 32, n_super = len / sb;
    if (!n_super) return -1;
    if (samples > n_super) samples = n_super;
    if (!samples) samples = 1;
    uint8_t* disk = nullptr;
    if (posix_memalign((void**)&disk, 4096, sb) != 0) return -1;
    std::vector<uint8_t> ref(sb);
    std::vector<std::vector<uint8_t>> scr((size_t)cfg.cpu_threads, std::vector<uint8_t>(lib.scratch_bytes(cb)));
    std::vector<uint8_t*> sp;
    for (auto& v : scr) sp.push_back(v.data());
    void* pool = lib.pool_create();
    int rc = 0;
    for (uint64_t k = 0; k < samples; ++k) {
        const uint64_t J = samples == 1 ? 0 : k * (n_super - 1) / (samples - 1);
        if (pread_all(fd, disk, sb, (off_t)(J * sb)) != 0) { rc = -1; break; }
        if (lib.label(g_seed, 32, ref.data(), sb, J * cb, sp.data(), pool, 0, cb) != 0) { rc = -1; break; }
        ++*checked;
        if (std::memcmp(ref.data(), disk, sb) != 0) {
            ++*mism;
            std::fprintf(stderr, "  verify: MISMATCH super-chunk %llu\n", (unsigned long long)J);
        }
    }
    lib.pool_destroy(pool);
    std::free(disk);
    return rc;
}

// Random block challenges: block b (absolute 32-byte block of the target) is
// rank b % cb of super-chunk b / cb.  Reads go through a 4 KiB-aligned buffer
// so an O_DIRECT descriptor works.
int challenge_check(int fd, uint64_t len, uint32_t cb, uint64_t n, uint64_t* checked, uint64_t* mism,
                    double* gpu_s) {
    *checked = *mism = 0;
    const uint64_t blocks = len / ((uint64_t)cb * 32) * cb;
    if (!blocks || !n) return -1;
    uint64_t x = 0x9E3779B97F4A7C15ull ^ ((uint64_t)g_seed[0] << 8 | g_seed[1]);
    std::vector<uint64_t> pick(n);
    std::vector<pose_gpu_challenge_t> ch(n);
    for (uint64_t k = 0; k < n; ++k) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;          // xorshift64
        pick[k] = x % blocks;
        if (pose_gpu_chunk_seed(g_seed, 32, pick[k] / cb, ch[k].chunk_seed) != 0) return -1;
        ch[k].chunk_blocks = cb;
        ch[k].rank = (uint32_t)(pick[k] % cb);
    }
    std::vector<uint8_t> ans(n * 32);
    const double t0 = now_s();
    if (pose_gpu_challenge_batch(ch.data(), n, ans.data(), nullptr) != 0) return -1;
    *gpu_s = now_s() - t0;
    uint8_t* sec = nullptr;
    if (posix_memalign((void**)&sec, 4096, 4096) != 0) return -1;
    int rc = 0;
    for (uint64_t k = 0; k < n; ++k) {
        const uint64_t off = pick[k] * 32, base = off & ~4095ull;
        if (pread_all(fd, sec, 4096, (off_t)base) != 0) { rc = -1; break; }
        ++*checked;
        if (std::memcmp(sec + (off - base), ans.data() + k * 32, 32) != 0) {
            ++*mism;
            std::fprintf(stderr, "  challenge: MISMATCH block %llu\n", (unsigned long long)pick[k]);
        }
    }
    std::free(sec);
    return rc;
}

int open_target(const std::string& p, bool ro, uint64_t* size, bool* direct) {
    int flags = ro ? O_RDONLY : O_RDWR;
    int fd = open(p.c_str(), flags | O_DIRECT);
    *direct = fd >= 0;
    if (fd < 0) fd = open(p.c_str(), flags);     // e.g. tmpfs has no O_DIRECT
    if (fd < 0) return -1;
    struct stat st;
    if (fstat(fd, &st) != 0) { close(fd); return -1; }
    if (S_ISBLK(st.st_mode)) {
        uint64_t b = 0;
        if (ioctl(fd, BLKGETSIZE64, &b) != 0) { close(fd); return -1; }
        *size = b;
    } else {
        *size = (uint64_t)st.st_size;
    }
    return fd;
}

// A target opened and sized for the run.
struct Target {
    std::string path;
    int fd = -1;
    uint64_t size = 0, wipe_len = 0;
    bool direct = false;
    // results
    bool wiped = false;
    double seconds = 0;
    pose_gpu_server_stats_t st{};
    int64_t verify_checked = -1, verify_mism = 0, chal_checked = -1, chal_mism = 0;
};

std::string json_str(const std::string& v) {
    std::string o = "\"";
    for (unsigned char c : v) {
        if (c == '"' || c == '\\') { o += '\\'; o += (char)c; }
        else if (c < 0x20) { char b[8]; std::snprintf(b, sizeof b, "\\u%04x", c); o += b; }
        else o += (char)c;
    }
    return o + "\"";
}

// One JSON object with the run's configuration, per-target results and totals.
bool write_json(const Config& cfg, uint32_t cb, int dev, const std::vector<Target>& ts, uint64_t total,
                double wall, bool ok) {
    FILE* f = cfg.json == "-" ? stdout : std::fopen(cfg.json.c_str(), "w");
    if (!f) { std::fprintf(stderr, "error: --json %s: %s\n", cfg.json.c_str(), std::strerror(errno)); return false; }
    std::fprintf(f, "{\"tool\": \"pose_wipe_bench\", \"chunk_blocks\": %u, \"gpu\": %d, \"dry_run\": %s, \"targets\": [",
                 cb, dev, cfg.dry_run ? "true" : "false");
    for (size_t i = 0; i < ts.size(); ++i) {
        const Target& t = ts[i];
        const double mib = (double)t.st.bytes_done / MiB;
        std::fprintf(f, "%s{\"path\": %s, \"size_bytes\": %llu, \"wipe_bytes\": %llu, \"bytes_done\": %llu, "
                     "\"seconds\": %.3f, \"mib_per_s\": %.2f, \"super_chunks\": %llu, \"launches\": %llu, "
                     "\"kernel_s\": %.3f, \"copy_s\": %.3f, \"sink_s\": %.3f, \"p50_ms\": %.1f, \"p99_ms\": %.1f, "
                     "\"o_direct\": %s",
                     i ? ", " : "", json_str(t.path).c_str(), (unsigned long long)t.size, (unsigned long long)t.wipe_len,
                     (unsigned long long)t.st.bytes_done, t.seconds, t.seconds > 0 ? mib / t.seconds : 0.0,
                     (unsigned long long)t.st.super_chunks_done, (unsigned long long)t.st.launches,
                     t.st.kernel_ms_total / 1e3, t.st.copy_ms_total / 1e3, t.st.sink_ms_total / 1e3, t.st.p50_ms,
                     t.st.p99_ms, t.direct ? "true" : "false");
        if (t.verify_checked >= 0)
            std::fprintf(f, ", \"verify\": {\"checked\": %lld, \"mismatches\": %lld}", (long long)t.verify_checked,
                         (long long)t.verify_mism);
        if (t.chal_checked >= 0)
            std::fprintf(f, ", \"challenge\": {\"checked\": %lld, \"mismatches\": %lld}", (long long)t.chal_checked,
                         (long long)t.chal_mism);
        std::fprintf(f, "}");
    }
    std::fprintf(f, "], \"total_bytes\": %llu, \"wall_s\": %.3f, \"aggregate_mib_per_s\": %.2f, \"ok\": %s}\n",
                 (unsigned long long)total, wall, wall > 0 ? (double)total / MiB / wall : 0.0, ok ? "true" : "false");
    const bool good = !std::ferror(f);
    if (f != stdout) std::fclose(f); else std::fflush(f);
    return good;
}

void close_all(std::vector<Target>& ts) {
    for (Target& t : ts)
        if (t.fd >= 0) { close(t.fd); t.fd = -1; }
}

// --gpus LIST, as the reference parses it (disk_wipe_bench.cu:786-806):
// comma-separated ordinals, no duplicates.  Returns false on a bad list.
bool parse_gpus(const std::string& l, std::vector<int>* out) {
    size_t p = 0;
    while (p <= l.size()) {
        const size_t c = l.find(',', p);
        const std::string tok = l.substr(p, c == std::string::npos ? std::string::npos : c - p);
        char* end = nullptr;
        const long v = std::strtol(tok.c_str(), &end, 10);
        if (tok.empty() || !end || *end != '\0' || v < 0 || v > 1 << 20) {
            std::fprintf(stderr, "--gpus: bad ordinal '%s' (expected e.g. 0,1,2,3)\n", tok.c_str());
            return false;
        }
        for (int g : *out)
            if (g == (int)v) {
                std::fprintf(stderr, "--gpus: ordinal %ld listed twice\n", v);
                return false;
            }
        out->push_back((int)v);
        if (c == std::string::npos) break;
        p = c + 1;
    }
    return true;
}

// Pinned staging for a target: --stage-mib when given, else room for two of
// the label server's largest launches (four super-chunks per CU each), so one
// launch's output drains to the sink while the next is labeled.  Measured at
// chunk_blocks = 2^17 on 16 GiB (dry run): 1 GiB of staging cuts launches to
// about 140 super-chunks, which run in the latency form, 1,602 MiB/s; 4 GiB
// 2,736 MiB/s; 8 GiB 2,953 MiB/s.  Between 1 and 8 GiB, no more than the
// target needs, and at most a quarter of the host's memory.
uint64_t stage_bytes(const Config& cfg, uint64_t sb, uint64_t wipe_len, int dev) {
    if (cfg.stage_mib) return std::max<uint64_t>(cfg.stage_mib * MiB, sb);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    uint64_t want = 2ull * 4 * (uint64_t)cus * sb;
    want = std::min<uint64_t>(std::max<uint64_t>(want, 1ull << 30), 8ull << 30);
    want = std::min<uint64_t>(want, std::max<uint64_t>(wipe_len, sb));
    const long pages = sysconf(_SC_PHYS_PAGES), page = sysconf(_SC_PAGESIZE);
    if (pages > 0 && page > 0) want = std::min<uint64_t>(want, (uint64_t)pages * (uint64_t)page / 4);
    want = want / sb * sb;
    return std::max<uint64_t>(want, sb);
}

// The GPU of this process: --gpus' ordinal if given, else the current device.
// Called before the first GPU call, so argument errors never touch the GPU.
bool select_gpu(const Config& cfg, int* dev) {
    if (cfg.gpu >= 0 && hipSetDevice(cfg.gpu) != hipSuccess) {
        std::fprintf(stderr, "error: --gpus %d: no such GPU\n", cfg.gpu);
        return false;
    }
    return hipGetDevice(dev) == hipSuccess;
}

}  // namespace

int main(int argc, char** argv) {
    Config cfg;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto need = [&](const char* n) -> const char* {
            if (i + 1 >= argc) { std::fprintf(stderr, "%s requires a value\n", n); std::exit(2); }
            return argv[++i];
        };
        if (a == "--disk") cfg.disks.push_back(need("--disk"));
        else if (a == "--chunk-blocks") cfg.chunk_blocks = (uint32_t)std::strtoul(need("--chunk-blocks"), nullptr, 10);
        else if (a == "--per-disk-mib") cfg.per_disk_bytes = std::strtoull(need("--per-disk-mib"), nullptr, 10) * MiB;
        else if (a == "--hbm-window-mib") cfg.hbm_window_mib = std::strtoull(need("--hbm-window-mib"), nullptr, 10);
        else if (a == "--stage-mib") cfg.stage_mib = std::strtoull(need("--stage-mib"), nullptr, 10);
        else if (a == "--dry-run") cfg.dry_run = true;
        else if (a == "--confirm") cfg.confirm = true;
        else if (a == "--verify") cfg.verify_n = std::strtoull(need("--verify"), nullptr, 10);
        else if (a == "--verify-only") cfg.verify_only = true;
        else if (a == "--challenge") cfg.challenge_n = std::strtoull(need("--challenge"), nullptr, 10);
        else if (a == "--cpu-lib") cfg.cpu_lib = need("--cpu-lib");
        else if (a == "--cpu-threads") cfg.cpu_threads = std::atoi(need("--cpu-threads"));
        else if (a == "--verbose") cfg.verbose = true;
        else if (a == "--json") cfg.json = need("--json");
        else if (a == "--gpus") {
            std::vector<int> gpus;
            const std::string l = need("--gpus");
            if (!parse_gpus(l, &gpus)) return 2;
            if (gpus.size() != 1) {
                std::fprintf(stderr, "error: --gpus %s: this producer drives one GPU per process; "
                             "run one process per GPU (--gpus 0, --gpus 1, ...) over disjoint targets\n", l.c_str());
                return 2;
            }
            cfg.gpu = gpus[0];
        } else if (a == "--mode") {
            const std::string m = need("--mode");
            if (m != "gpus-per-disk" && m != "disk-per-gpu") {
                std::fprintf(stderr, "unknown --mode %s (use disk-per-gpu | gpus-per-disk)\n", m.c_str());
                return 2;
            }
        } else if (a == "--gpus-per-disk") {
            const int k = std::atoi(need("--gpus-per-disk"));
            if (k > 1)
                std::fprintf(stderr, "warning: --gpus-per-disk %d exceeds 1 selected GPU; using 1\n", k);
        } else if (a == "--seed-hex") {
            if (parse_hex(need("--seed-hex"), g_seed) != 0) { std::fprintf(stderr, "--seed-hex must be 64 hex characters\n"); return 2; }
        } else if (a == "-h" || a == "--help") { usage(argv[0]); return 0; }
        else { std::fprintf(stderr, "unknown argument: %s\n", a.c_str()); usage(argv[0]); return 2; }
    }
    if (cfg.disks.empty()) {
        std::fprintf(stderr, "error: at least one --disk is required (no auto-selection)\n");
        usage(argv[0]);
        return 2;
    }
    const uint32_t cb = cfg.chunk_blocks ? cfg.chunk_blocks : 4096;
    if ((cb & (cb - 1)) != 0 || cb > (1u << 20)) {
        std::fprintf(stderr, "error: --chunk-blocks must be a power of two in [1, %u]\n", 1u << 20);
        return 2;
    }
    if (cfg.cpu_threads < 1 || cfg.cpu_threads > 256) { std::fprintf(stderr, "error: --cpu-threads\n"); return 2; }
    if (cfg.verify_only && cfg.verify_n == 0 && cfg.challenge_n == 0) cfg.verify_n = 128;
    const uint64_t sb = (uint64_t)cb * 32;
    // The CPU verifier is resolved before any target is opened, so a wipe
    // never starts that its --verify could not check.
    CpuLib cpu;
    if (cfg.verify_n && (cfg.verify_only || !cfg.dry_run) && !cpu.load_default(cfg.cpu_lib)) {
        const char* why = dlerror();
        std::fprintf(stderr, "error: --verify needs the reference CPU library (graph.h API): cannot load '%s' (%s); "
                     "pass --cpu-lib PATH or set POSE_COMMON_LIB\n", cpu.path.c_str(), why ? why : "missing symbols");
        return 2;
    }

    // Open and size every target first (the reference does the same before
    // its banner): a bad target ends the run before anything is written.
    std::vector<Target> targets;
    for (const std::string& path : cfg.disks) {
        Target t;
        t.path = path;
        t.fd = open_target(path, cfg.verify_only, &t.size, &t.direct);
        if (t.fd < 0) {
            std::fprintf(stderr, "error: %s: %s\n", path.c_str(), std::strerror(errno));
            close_all(targets);
            return 1;
        }
        targets.push_back(t);
        if (t.direct && sb % 4096 != 0) {
            std::fprintf(stderr, "error: super-chunk %llu B is not a multiple of 4096 for O_DIRECT; "
                         "use --chunk-blocks >= 128\n", (unsigned long long)sb);
            close_all(targets);
            return 1;
        }
        const uint64_t want = cfg.per_disk_bytes ? std::min(cfg.per_disk_bytes, t.size) : t.size;
        targets.back().wipe_len = want / sb * sb;
        if (targets.back().wipe_len == 0) {
            std::fprintf(stderr, "error: %s: smaller than one super-chunk (%llu B)\n", path.c_str(), (unsigned long long)sb);
            close_all(targets);
            return 1;
        }
    }

    if (cfg.verify_only) {
        int rc = 0;
        uint64_t mism_all = 0;
        std::printf("=== pose-wipe-bench: verify-only ===\nchunk_blocks  : %u  samples/disk: %llu  challenges/disk: %llu\n",
                    cb, (unsigned long long)cfg.verify_n, (unsigned long long)cfg.challenge_n);
        int dev = -1;
        if (cfg.challenge_n && !select_gpu(cfg, &dev)) { close_all(targets); return 1; }
        for (const Target& t : targets) {
            uint64_t checked = 0, mism = 0;
            int trc = 0;
            if (cfg.verify_n) {
                trc = verify(t.fd, t.wipe_len, cb, cfg.verify_n, cfg, cpu, &checked, &mism);
                if (trc) std::printf("  %s: verify ERROR\n", t.path.c_str());
                else std::printf("  %s: %llu/%llu sampled super-chunks match%s\n", t.path.c_str(),
                                 (unsigned long long)(checked - mism), (unsigned long long)checked,
                                 mism ? "   <-- MISMATCH" : "");
            }
            if (trc == 0 && cfg.challenge_n) {
                uint64_t cc = 0, cm = 0;
                double gs = 0;
                trc = challenge_check(t.fd, t.wipe_len, cb, cfg.challenge_n, &cc, &cm, &gs);
                std::printf("  %s: %llu/%llu challenged blocks match (GPU answers in %.2f s)%s\n", t.path.c_str(),
                            (unsigned long long)(cc - cm), (unsigned long long)cc, gs, cm ? "   <-- MISMATCH" : "");
                mism += cm;
            }
            if (trc) rc = 1;
            mism_all += mism;
        }
        close_all(targets);
        std::printf("%s\n", (rc || mism_all) ? "VERIFY FAILED" : "VERIFY OK");
        return (rc || mism_all) ? 1 : 0;
    }

    // One GPU wipes the targets in turn, largest first (the reference's
    // assignment for one team, disk_wipe_bench.cu:989-996).
    std::vector<size_t> order(targets.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return targets[x].wipe_len > targets[y].wipe_len; });
    uint64_t total_wipe = 0;
    for (const Target& t : targets) total_wipe += t.wipe_len;
    std::printf("=== pose-wipe-bench ===\n");
    std::printf("schedule      : 1 GPU(s)/disk, 1 disk(s) in parallel  [disk-per-gpu]%s\n",
                targets.size() > 1 ? ", targets in turn" : "");
    std::printf("chunk_blocks  : %u  (super-chunk %llu B)\n", cb, (unsigned long long)sb);
    std::printf("write mode    : %s\n", cfg.dry_run ? "DRY-RUN (no writes)"
                                       : (cfg.confirm ? "REAL WRITES (destructive)" : "(blocked: need --confirm or --dry-run)"));
    for (const Target& t : targets)
        std::printf("  target      : %s  size %.2f MiB  wipe %.2f MiB%s%s\n", t.path.c_str(), (double)t.size / MiB,
                    (double)t.wipe_len / MiB, cfg.per_disk_bytes ? "  [capped]" : "", t.direct ? "  [O_DIRECT]" : "");
    std::printf("total to write: %.2f MiB\n", (double)total_wipe / MiB);
    std::fflush(stdout);
    if (!cfg.dry_run && !cfg.confirm) {
        std::fprintf(stderr, "\nREFUSING TO WRITE: this will DESTROY all data on the target(s). "
                     "Re-run with --dry-run, or --confirm to wipe.\n");
        close_all(targets);
        return 1;
    }
    int dev = -1;
    if (!select_gpu(cfg, &dev)) { close_all(targets); return 1; }

    int rc = 0;
    uint64_t total_bytes = 0;
    double busy_label = 0, busy_copy = 0, busy_sink = 0, busy = 0;
    const double t_all = now_s();
    for (size_t oi = 0; oi < order.size() && rc == 0; ++oi) {
        Target& t = targets[order[oi]];
        pose_gpu_server_config_t sc{};
        sc.hbm_budget_bytes = cfg.hbm_window_mib * MiB;   // 0: the server's default
        sc.pinned_bytes = stage_bytes(cfg, sb, t.wipe_len, dev);
        sc.sink_fd = cfg.dry_run ? -1 : t.fd;
        sc.preload_mask = 1u << (31 - __builtin_clz(cb));
        const double t_build = now_s();
        pose_gpu_server_t* srv = pose_gpu_server_create(&sc);
        if (!srv) { rc = 1; break; }
        const double t0 = now_s();
        std::printf("%s: topology built in %.2f s on GPU %d\n", t.path.c_str(), t0 - t_build, dev);
        // Window-sized requests keep the pipeline full while bounding host bookkeeping.
        const uint64_t window = std::max<uint64_t>(sb, std::min<uint64_t>(t.wipe_len, 4096 * sb));
        for (uint64_t pos = 0; pos < t.wipe_len; pos += window) {
            const uint64_t len = std::min(window, t.wipe_len - pos);
            pose_gpu_request_t r{};
            r.session_seed = g_seed;
            r.seed_len = 32;
            r.chunk_blocks = cb;
            r.region_block_offset = pos / 32;
            r.n_super = len / sb;
            r.sink_offset = pos;
            uint64_t id;
            if (pose_gpu_server_submit(srv, &r, &id) != 0) { rc = 1; break; }
            if (cfg.verbose) std::printf("  submitted window at %.0f MiB (%llu super-chunks)\n", (double)pos / MiB,
                                         (unsigned long long)r.n_super);
        }
        if (pose_gpu_server_drain(srv) != 0) rc = 1;
        const double wall = now_s() - t0;
        pose_gpu_server_stats_t st{};
        pose_gpu_server_stats(srv, &st);
        pose_gpu_server_destroy(srv);
        if (!cfg.dry_run && rc == 0 && fsync(t.fd) != 0 && errno != EINVAL) rc = 1;
        const double mib = (double)st.bytes_done / MiB;
        std::printf("  GPU wipe: %.0f MiB in %.3f s = %.1f MiB/s\n", mib, wall, wall > 0 ? mib / wall : 0.0);
        std::printf("  %s: %llu super-chunks, %llu launches, kernel %.1f s, p50 %.0f ms, p99 %.0f ms per super-chunk%s\n",
                    t.path.c_str(), (unsigned long long)st.super_chunks_done, (unsigned long long)st.launches,
                    st.kernel_ms_total / 1e3, st.p50_ms, st.p99_ms, rc ? "  ERROR" : "");
        std::fflush(stdout);
        t.wiped = true;
        t.seconds = wall;
        t.st = st;
        total_bytes += st.bytes_done;
        busy += wall;
        busy_label += st.kernel_ms_total / 1e3;
        busy_copy += st.copy_ms_total / 1e3;
        busy_sink += st.sink_ms_total / 1e3;
    }
    const double wall_all = now_s() - t_all;
    std::printf("\n=== results ===\n");
    std::printf("  [gpu %d] %.2f MiB in %.1fs = %.1f MiB/s  (label %.1fs  d2h %.1fs  write %.1fs)%s\n", dev,
                (double)total_bytes / MiB, busy, busy > 0 ? (double)total_bytes / MiB / busy : 0.0, busy_label, busy_copy,
                busy_sink, rc ? "  ERROR" : "");
    std::printf("total written : %.2f MiB\n", (double)total_bytes / MiB);
    std::printf("wall clock    : %.1f s\n", wall_all);
    std::printf("aggregate     : %.1f MiB/s%s\n", wall_all > 0 ? (double)total_bytes / MiB / wall_all : 0.0,
                cfg.dry_run ? "  [dry-run: label+D2H only, no writes]" : "");
    if (rc == 0 && cfg.verify_n && !cfg.dry_run) {
        for (Target& t : targets) {
            uint64_t checked = 0, mism = 0;
            const int vrc = verify(t.fd, t.wipe_len, cb, cfg.verify_n, cfg, cpu, &checked, &mism);
            t.verify_checked = (int64_t)checked;
            t.verify_mism = (int64_t)mism;
            std::printf("verify        : %s: %llu/%llu sampled super-chunks match%s\n", t.path.c_str(),
                        (unsigned long long)(checked - mism), (unsigned long long)checked, mism ? "   <-- MISMATCH" : "");
            if (vrc || mism) rc = 1;
        }
    } else if (cfg.verify_n && cfg.dry_run) {
        std::fprintf(stderr, "warning: --verify is ignored with --dry-run (nothing was written)\n");
    }
    if (rc == 0 && cfg.challenge_n && !cfg.dry_run) {
        for (Target& t : targets) {
            uint64_t cc = 0, cm = 0;
            double gs = 0;
            const int crc = challenge_check(t.fd, t.wipe_len, cb, cfg.challenge_n, &cc, &cm, &gs);
            t.chal_checked = (int64_t)cc;
            t.chal_mism = (int64_t)cm;
            std::printf("challenge     : %s: %llu/%llu random blocks match GPU answers (%.2f s)%s\n", t.path.c_str(),
                        (unsigned long long)(cc - cm), (unsigned long long)cc, gs, cm ? "   <-- MISMATCH" : "");
            if (crc || cm) rc = 1;
        }
    }
    close_all(targets);
    if (!cfg.json.empty() && !write_json(cfg, cb, dev, targets, total_bytes, wall_all, rc == 0)) rc = 1;
    std::printf("%s\n", rc ? "FAILED" : "OK");
    return rc;
}
