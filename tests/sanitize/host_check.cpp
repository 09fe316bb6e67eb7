// Host-side sanitizer check (test infrastructure).
//
// The library's host code (descriptor validation, instance resolution, tuning-cache
// parsing, the batched queue builder, the sparse and utility entry points' argument
// checks) and the CPU oracle, built with AddressSanitizer and UndefinedBehaviorSanitizer
// on the host side only, and driven through every path that runs without a device:
// argument checks at and beyond their limits, the batched work queue of random problem
// sets (checked to tile every C exactly once), malformed and oversized tuning files, and
// the oracle's routines on small inputs against brute-force products.  Any sanitizer
// report aborts the program (-fno-sanitize-recover); a failed CHECK makes it exit 1.
//
// The batched translation unit is included rather than linked so that its internal
// queue builder (check_problem, bucket_of, count_items, build_items, queue_shift,
// prior_variant) can be called directly.
#include "batched.hip"

#include <unistd.h>

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

extern "C" {
#include "../../oracle/kami_oracle.h"
}

namespace {

int g_failed = 0;
int g_checks = 0;

#define CHECK(cond, ...)                                                       \
    do {                                                                       \
        ++g_checks;                                                            \
        if (!(cond)) {                                                         \
            ++g_failed;                                                        \
            std::fprintf(stderr, "CHECK failed %s:%d: %s: ", __FILE__, __LINE__, #cond); \
            std::fprintf(stderr, __VA_ARGS__);                                 \
            std::fprintf(stderr, "\n");                                        \
        }                                                                      \
    } while (0)

bool g_have_device = false;

const char* tmpdir() {
    const char* t = std::getenv("TMPDIR");
    return t && *t ? t : "/tmp";
}

std::string write_temp(const std::string& body) {
    std::string path = std::string(tmpdir()) + "/kami_tune_XXXXXX";
    std::vector<char> buf(path.begin(), path.end());
    buf.push_back('\0');
    const int fd = mkstemp(buf.data());
    if (fd < 0) return "";
    size_t off = 0;
    while (off < body.size()) {
        const ssize_t w = write(fd, body.data() + off, body.size() - off);
        if (w <= 0) break;
        off += (size_t)w;
    }
    close(fd);
    return std::string(buf.data());
}

// ------------------------------------------------------------------ status, descriptors

void check_status_and_version() {
    for (int s = -3; s < 12; ++s) {
        const char* p = kami_status_string((kami_status)s);
        CHECK(p && std::strlen(p) > 0, "status %d", s);
    }
    CHECK(kami_version() && std::strlen(kami_version()) > 0, "version");
    const kami_status dc = kami_device_check();
    CHECK(dc == KAMI_OK || dc == KAMI_ENODEV || dc == KAMI_EHIP, "device check %d", (int)dc);
    g_have_device = dc == KAMI_OK;
    (void)kami_last_hip_error();
}

bool config_sane(const kami_gemm_config& c) {
    return c.threads > 0 && c.threads <= 1024 && c.threads % 64 == 0 && c.lds_bytes >= 0 &&
           c.lds_bytes <= 160 * 1024 && c.waves_m > 0 && c.waves_n > 0 && c.waves_k > 0 &&
           c.waves_m * c.waves_n * c.waves_k * 64 <= c.threads && c.pipeline >= 1 &&
           c.pipeline <= 2 && c.partition >= KAMI_PART_1D && c.partition <= KAMI_PART_3D;
}

void check_descriptors() {
    const int32_t dims[] = {INT32_MIN, -16, 0, 8, 16, 24, 32, 48, 64, 128, 192, 208, 224,
                            256, 512, 4096, 1 << 20, INT32_MAX};
    const int64_t lds[] = {-1, 0, 1, 64, 65, 1 << 20, (int64_t)1 << 40, INT64_MAX};
    int queried = 0, resolved = 0;
    for (int prec = -1; prec <= 6; ++prec)
        for (int32_t m : dims)
            for (int32_t n : dims)
                for (int32_t k : dims) {
                    kami_gemm_desc d;
                    const kami_status st = kami_gemm_desc_init(&d, m, n, k, prec);
                    CHECK(st == KAMI_OK || st == KAMI_EINVAL || st == KAMI_ENOTSUP,
                          "desc_init %d %d %d %d -> %d", m, n, k, prec, (int)st);
                    if (st != KAMI_OK) continue;
                    kami_gemm_config c;
                    kami_status q = kami_block_gemm_query(&d, &c);
                    ++queried;
                    if (q == KAMI_OK) {
                        ++resolved;
                        CHECK(config_sane(c), "config of %d^3 prec %d", m, prec);
                    }
                    // one field at a time through its range, beyond its limits included
                    for (int part = -1; part <= 4; ++part) {
                        kami_gemm_desc e = d;
                        e.partition = part;
                        for (int ranks : {-1, 0, 1, 2, 3, 4, 8, 13, 14, 16, 64, INT32_MAX}) {
                            e.ranks = ranks;
                            for (int pipe = -1; pipe <= 3; ++pipe) {
                                e.pipeline = pipe;
                                q = kami_block_gemm_query(&e, &c);
                                ++queried;
                                if (q == KAMI_OK) CHECK(config_sane(c), "config");
                            }
                        }
                    }
                    for (int32_t v : {INT32_MIN, -1, 0, 1, 2, 7, 8, 9, 1 << 20, INT32_MAX}) {
                        kami_gemm_desc e = d;
                        e.iters = v;
                        (void)kami_block_gemm_query(&e, &c);
                        e = d;
                        e.slices = v;
                        (void)kami_block_gemm_query(&e, &c);
                        e = d;
                        e.b_layout = v;
                        (void)kami_block_gemm_query(&e, &c);
                        queried += 3;
                    }
                    for (int64_t v : lds) {
                        kami_gemm_desc e = d;
                        e.lda = v;
                        (void)kami_block_gemm_query(&e, &c);
                        e = d;
                        e.ldb = v;
                        (void)kami_block_gemm_query(&e, &c);
                        e = d;
                        e.ldc = v;
                        (void)kami_block_gemm_query(&e, &c);
                        queried += 3;
                        if (!g_have_device) {
                            // argument checks of the compute entry points; without a
                            // device a valid request stops at the device check
                            e = d;
                            e.lda = v;
                            void* p = reinterpret_cast<void*>(uintptr_t(1) << 20);
                            for (int64_t count : {(int64_t)-1, (int64_t)0, (int64_t)1,
                                                  (int64_t)1 << 40, INT64_MAX}) {
                                const kami_status r =
                                    kami_block_gemm(&e, p, p, p, count, 0, 0, 0, nullptr);
                                CHECK(r != KAMI_OK || count == 0, "block_gemm without device -> OK");
                                const kami_status r2 =
                                    kami_split_gemm(&e, p, p, p, count, 0, 0, 0, nullptr);
                                CHECK(r2 != KAMI_OK || count == 0, "split_gemm without device -> OK");
                            }
                        }
                    }
                }
    // shapes with compiled instances resolve from a plain descriptor
    const int known[][4] = {{KAMI_PREC_F64, 64, 64, 64},        {KAMI_PREC_F64, 128, 128, 128},
                            {KAMI_PREC_F16_F32, 256, 256, 256}, {KAMI_PREC_F16_F32, 256, 256, 16},
                            {KAMI_PREC_BF16_F32, 224, 224, 224}, {KAMI_PREC_F64_SPLIT_F16, 128, 128, 128}};
    for (const auto& s : known) {
        kami_gemm_desc d;
        kami_gemm_config c;
        CHECK(kami_gemm_desc_init(&d, s[1], s[2], s[3], s[0]) == KAMI_OK &&
                  kami_block_gemm_query(&d, &c) == KAMI_OK,
              "%d x %d x %d precision %d does not resolve", s[1], s[2], s[3], s[0]);
    }
    CHECK(kami_gemm_desc_init(nullptr, 64, 64, 64, 0) == KAMI_EINVAL, "null desc");
    kami_gemm_config c;
    CHECK(kami_block_gemm_query(nullptr, &c) == KAMI_EINVAL, "null query desc");
    std::printf("descriptors: %d queries, %d plain descriptors resolved\n", queried, resolved);
}

// ------------------------------------------------------------------ tuning cache

std::string slurp(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return "";
    std::string s;
    char buf[4096];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, r);
    std::fclose(f);
    return s;
}

void check_tune_cache() {
    CHECK(kami_tune_clear() == KAMI_OK, "clear");
    CHECK(kami_tune_load(nullptr) == KAMI_EINVAL, "null load");
    CHECK(kami_tune_save(nullptr) == KAMI_EINVAL, "null save");
    CHECK(kami_tune_load("/nonexistent/dir/kami.tune") == KAMI_EINVAL, "missing file");
    CHECK(kami_tune_save(tmpdir()) == KAMI_EINVAL, "save to a directory");

    // a compiled instance, named the way the file names it
    kami_gemm_desc d;
    CHECK(kami_gemm_desc_init(&d, 64, 64, 64, KAMI_PREC_F64) == KAMI_OK, "init");
    d.partition = KAMI_PART_2D;
    kami_gemm_config c;
    CHECK(kami_block_gemm_query(&d, &c) == KAMI_OK, "2D 64^3 query");
    char rec[128];
    std::snprintf(rec, sizeof rec, "block %d 64 64 64 %d %d %d\n", KAMI_PREC_F64, c.partition,
                  c.ranks, c.pipeline);
    const std::string good = std::string("kami-tune 1\n# comment\n\n") + rec + "batched 0 3 1\n";
    std::string path = write_temp(good);
    CHECK(!path.empty(), "temp file");
    CHECK(kami_tune_load(path.c_str()) == KAMI_OK, "good file");
    unlink(path.c_str());
    kami_gemm_desc a = d;
    a.partition = KAMI_PART_AUTO;
    kami_gemm_config ca;
    CHECK(kami_block_gemm_query(&a, &ca) == KAMI_OK && ca.partition == c.partition &&
              ca.ranks == c.ranks,
          "loaded choice drives AUTO");
    CHECK(batched_tuned_variant(KAMI_PREC_F64, 3) == 1, "loaded batched variant");
    path = std::string(tmpdir()) + "/kami_tune_roundtrip.txt";
    CHECK(kami_tune_save(path.c_str()) == KAMI_OK, "save");
    const std::string saved = slurp(path);
    CHECK(saved.find(rec) != std::string::npos && saved.find("batched 0 3 1\n") != std::string::npos,
          "saved file holds the records");
    CHECK(kami_tune_clear() == KAMI_OK && batched_tuned_variant(KAMI_PREC_F64, 3) == -1, "clear");
    CHECK(kami_tune_load(path.c_str()) == KAMI_OK && batched_tuned_variant(KAMI_PREC_F64, 3) == 1,
          "round trip");
    unlink(path.c_str());
    kami_tune_clear();

    std::string garbage;
    for (int i = 0; i < 4096; ++i) garbage.push_back((char)(i * 131 % 256));
    const std::string longline = "kami-tune 1\nbatched 0 1 " + std::string(10000, '7') + "\n";
    const std::vector<std::string> bad = {
        "",
        "kami-tune 2\n",
        "kami-tune\n",
        "kami-tune 1\nblock 0 64 64\n",
        "kami-tune 1\nblock 0 64 64 64 9 2 1\n",
        "kami-tune 1\nblock 0 64 64 64 2 0 1\n",
        "kami-tune 1\nblock 0 64 64 64 2 2 3\n",
        "kami-tune 1\nblock 3 64 64 64 2 2 1\n",
        "kami-tune 1\nblock 0 -64 64 64 2 2 1\n",
        "kami-tune 1\nblock 0 99999999999999999999 64 64 2 2 1\n",
        "kami-tune 1\nbatched 0 125 0\n",
        "kami-tune 1\nbatched 0 -1 0\n",
        "kami-tune 1\nbatched 9 0 0\n",
        "kami-tune 1\nbatched 0 0 99\n",
        "kami-tune 1\nbatched 0 0 -2147483648\n",
        "kami-tune 1\nsomething else\n",
        "kami-tune 1\n" + garbage,
        garbage,
        longline,
        std::string("kami-tune 1\nbatched 0 1 1\0 junk\n", 30),
        "kami-tune 1\nbatched 0 1 1 junk\n",
        "kami-tune 1\nblock 0 64 64 64 2 2 1 7\n",
        "kami-tune 1 trailing\n",
    };
    for (size_t i = 0; i < bad.size(); ++i) {
        path = write_temp(bad[i]);
        CHECK(kami_tune_load(path.c_str()) == KAMI_EINVAL, "malformed file %zu accepted", i);
        unlink(path.c_str());
        CHECK(batched_tuned_variant(KAMI_PREC_F64, 0) == -1, "malformed file %zu applied", i);
    }
    // many records, each bucket written many times: the last one wins
    std::string many = "kami-tune 1\n";
    for (int r = 0; r < 40; ++r)
        for (int b = 0; b < kBatchedBuckets; ++b)
            many += "batched 1 " + std::to_string(b) + " " + std::to_string((b + r) % 3) + "\n";
    path = write_temp(many);
    CHECK(kami_tune_load(path.c_str()) == KAMI_OK, "many records");
    unlink(path.c_str());
    for (int b = 0; b < kBatchedBuckets; ++b)
        CHECK(batched_tuned_variant(KAMI_PREC_F16_F32, b) == (b + 39) % 3, "bucket %d", b);
    kami_tune_clear();
    std::printf("tuning cache: %zu malformed files refused\n", bad.size());
}

// ------------------------------------------------------------------ batched queue

uintptr_t fake(uint64_t region) { return (uintptr_t)(region + 1) << 32; }

void check_batched_queue() {
    std::mt19937_64 rng(12345);
    const int precs[] = {KAMI_PREC_F64, KAMI_PREC_F16_F32, KAMI_PREC_BF16_F32};
    int64_t total_items = 0;
    for (int trial = 0; trial < 60; ++trial) {
        const int prec = precs[trial % 3];
        const int layout = (trial / 3) % 2 ? KAMI_B_KN : KAMI_B_NK;
        const int ei = elt_in(prec), eo = elt_out(prec);
        const int64_t count = 1 + (int64_t)(rng() % 300);
        std::vector<kami_problem> probs((size_t)count);
        for (int64_t p = 0; p < count; ++p) {
            kami_problem& P = probs[(size_t)p];
            std::memset(&P, 0, sizeof P);
            P.m = 16 * (1 + (int)(rng() % 24));
            P.n = 16 * (1 + (int)(rng() % 24));
            P.k = 16 * (1 + (int)(rng() % 40));
            if (rng() % 4 == 0) P.lda = P.k + 16 * (int64_t)(rng() % 4);
            if (rng() % 4 == 0) P.ldc = P.n + 16 * (int64_t)(rng() % 4);
            P.A = reinterpret_cast<const void*>(fake(3 * (uint64_t)p));
            P.B = reinterpret_cast<const void*>(fake(3 * (uint64_t)p + 1));
            P.C = reinterpret_cast<void*>(fake(3 * (uint64_t)p + 2));
            CHECK(check_problem(P, layout, ei, eo) == KAMI_OK, "problem %d x %d x %d", P.m, P.n, P.k);
        }
        std::vector<int> bucket((size_t)count);
        for (int64_t p = 0; p < count; ++p) {
            bucket[(size_t)p] = bucket_of(probs[(size_t)p]);
            CHECK(bucket[(size_t)p] >= 0 && bucket[(size_t)p] < kBatchedBuckets, "bucket");
        }
        int vob[125];
        for (int b = 0; b < 125; ++b) {
            vob[b] = trial % 5 == 4 ? (int)(rng() % variants_of(prec)) : prior_variant(prec, b);
            CHECK(vob[b] >= 0 && vob[b] < variants_of(prec), "variant of bucket %d", b);
        }
        std::vector<WorkItem> items;
        build_items(probs, bucket, vob, ei, eo, layout == KAMI_B_KN, items);
        CHECK((int64_t)items.size() == count_items(probs.data(), count, bucket.data(), vob),
              "item count");
        total_items += (int64_t)items.size();
        // every C element of every problem in exactly one item, with the item's A and B
        // origins at that tile's rows and columns
        std::vector<std::vector<unsigned char>> cover((size_t)count);
        for (int64_t p = 0; p < count; ++p)
            cover[(size_t)p].assign((size_t)probs[(size_t)p].m * probs[(size_t)p].n, 0);
        for (const WorkItem& w : items) {
            const uint64_t region = ((uintptr_t)w.C >> 32) - 1;
            CHECK(region % 3 == 2 && region / 3 < (uint64_t)count, "C region");
            if (region % 3 != 2 || region / 3 >= (uint64_t)count) continue;
            const int64_t p = (int64_t)(region / 3);
            const kami_problem& P = probs[(size_t)p];
            const int64_t coff = (int64_t)((uintptr_t)w.C - (uintptr_t)P.C) / eo;
            const int64_t r0 = coff / P.ldc, c0 = coff % P.ldc;
            const int v = w.variant;
            CHECK(v >= 0 && v < variants_of(prec), "item variant");
            const int tm = kVariantTM[v], tn = kVariantTN[v];
            CHECK(r0 % tm == 0 && c0 % tn == 0 && r0 < P.m && c0 < P.n, "tile origin");
            CHECK(w.mlim == P.m - r0 && w.nlim == P.n - c0 && w.k == P.k, "tile limits");
            CHECK((uintptr_t)w.A == (uintptr_t)P.A + (uintptr_t)(r0 * P.lda * ei), "A origin");
            const int64_t boff = layout == KAMI_B_KN ? c0 : c0 * P.ldb;
            CHECK((uintptr_t)w.B == (uintptr_t)P.B + (uintptr_t)(boff * ei), "B origin");
            for (int64_t r = r0; r < std::min<int64_t>(P.m, r0 + tm); ++r)
                for (int64_t c = c0; c < std::min<int64_t>(P.n, c0 + tn); ++c)
                    ++cover[(size_t)p][(size_t)(r * P.n + c)];
        }
        for (int64_t p = 0; p < count; ++p)
            for (unsigned char x : cover[(size_t)p]) CHECK(x == 1, "problem %lld covered %d times", (long long)p, x);
    }
    for (int64_t n : {(int64_t)0, (int64_t)1, (int64_t)4096, (int64_t)1 << 20, (int64_t)1 << 40,
                      INT64_MAX / 4})
        for (int grid : {0, 1, 256, 768, 4096}) {
            const int32_t sh = queue_shift(n, grid);
            CHECK(sh >= 1 && sh <= kQueueShift, "queue_shift(%lld, %d) = %d", (long long)n, grid, sh);
        }
    std::printf("batched queue: %lld items over 60 random problem sets tile every C once\n",
                (long long)total_items);
}

void check_batched_arguments() {
    kami_problem base;
    std::memset(&base, 0, sizeof base);
    base.m = base.n = base.k = 32;
    base.A = base.B = reinterpret_cast<const void*>(uintptr_t(1) << 20);
    base.C = reinterpret_cast<void*>(uintptr_t(1) << 21);
    const int64_t extremes[] = {-1, 0, 31, 32, 48, (int64_t)1 << 24, (int64_t)1 << 40,
                                INT64_MAX / 8, INT64_MAX};
    for (int prec : {KAMI_PREC_F64, KAMI_PREC_F16_F32, KAMI_PREC_BF16_F32})
        for (int layout : {KAMI_B_NK, KAMI_B_KN})
            for (int field = 0; field < 3; ++field)
                for (int64_t v : extremes) {
                    kami_problem p = base;
                    (field == 0 ? p.lda : field == 1 ? p.ldb : p.ldc) = v;
                    const kami_status st = check_problem(p, layout, elt_in(prec), elt_out(prec));
                    CHECK(st == KAMI_OK || st == KAMI_EINVAL || st == KAMI_ENOTSUP, "status");
                    if (v < 0) CHECK(st == KAMI_EINVAL, "negative leading dimension accepted");
                }
    for (int32_t v : {INT32_MIN, -16, 0, 8, 16, 1 << 20, INT32_MAX - 15}) {
        kami_problem p = base;
        p.m = v;
        (void)check_problem(p, KAMI_B_NK, 8, 8);
        p = base;
        p.k = v;
        (void)check_problem(p, KAMI_B_KN, 2, 4);
    }

    std::vector<kami_problem> arr(2, base);
    kami_batched_opts o;
    std::memset(&o, 0, sizeof o);
    o.alpha = 1.0;
    kami_batched_plan* plan = nullptr;
    CHECK(kami_batched_plan_create(nullptr, 2, 0, &o, nullptr, &plan) == KAMI_EINVAL, "null");
    CHECK(kami_batched_plan_create(arr.data(), 0, 0, &o, nullptr, &plan) == KAMI_EINVAL, "zero");
    CHECK(kami_batched_plan_create(arr.data(), -5, 0, &o, nullptr, &plan) == KAMI_EINVAL, "neg");
    CHECK(kami_batched_plan_create(arr.data(), 2, 3, &o, nullptr, &plan) == KAMI_ENOTSUP, "split");
    CHECK(kami_batched_plan_create(arr.data(), 2, 99, &o, nullptr, &plan) == KAMI_ENOTSUP, "prec");
    o.b_layout = 7;
    CHECK(kami_batched_plan_create(arr.data(), 2, 0, &o, nullptr, &plan) == KAMI_EINVAL, "layout");
    o.b_layout = 0;
    if (!g_have_device) {
        CHECK(kami_batched_plan_create(arr.data(), 2, 0, &o, nullptr, &plan) == KAMI_ENODEV,
              "valid plan without device");
        CHECK(kami_batched_plan_create(arr.data(), 2, 0, nullptr, nullptr, &plan) == KAMI_ENODEV,
              "default options");
    }
    CHECK(kami_batched_execute(nullptr, nullptr) == KAMI_EINVAL, "null execute");
    CHECK(kami_batched_plan_info(nullptr, nullptr) == KAMI_EINVAL, "null info");
    CHECK(kami_batched_plan_destroy(nullptr) == KAMI_OK, "null destroy");

    // the allocation-free entry point's host staging, filled to its limits
    CHECK(kami_batched_gemm_static(arr.data(), 0, 0, &o, nullptr) == KAMI_EINVAL, "static zero");
    o.autotune = 1;
    CHECK(kami_batched_gemm_static(arr.data(), 2, 0, &o, nullptr) == KAMI_EINVAL, "static tune");
    o.autotune = 0;
    std::vector<kami_problem> big((size_t)kStaticItems + 1, base);
    for (kami_problem& p : big) p.m = p.n = p.k = 16;
    CHECK(kami_batched_gemm_static(big.data(), kStaticItems + 1, 0, &o, nullptr) == KAMI_ENOSPC,
          "static too many problems");
    for (kami_problem& p : big) p.m = p.n = 256;  // four tiles or more each
    CHECK(kami_batched_gemm_static(big.data(), kStaticItems, 0, &o, nullptr) == KAMI_ENOSPC,
          "static too many tiles");
    if (!g_have_device) {
        for (kami_problem& p : big) p.m = p.n = 16;
        CHECK(kami_batched_gemm_static(big.data(), kStaticItems, 0, &o, nullptr) == KAMI_ENODEV,
              "static at its limit reaches the device check");
    }
    std::printf("batched arguments: checked\n");
}

// ------------------------------------------------------------------ sparse, utility

void check_sparse_and_utility() {
    size_t bytes = 0;
    for (int32_t mb : {INT32_MIN, -1, 0, 1, 33, 1 << 20, INT32_MAX})
        for (int32_t nb : {INT32_MIN, -1, 0, 1, 31, 32, 33, 1 << 20, INT32_MAX}) {
            const kami_status st = kami_bsr_spgemm_workspace_size(mb, nb, &bytes);
            if (mb > 0 && nb > 0)
                CHECK(st == KAMI_OK && bytes == (size_t)mb * (((size_t)nb + 31) / 32) * 4,
                      "workspace %d %d", mb, nb);
            else
                CHECK(st == KAMI_EINVAL, "workspace %d %d", mb, nb);
        }
    CHECK(kami_bsr_spgemm_workspace_size(4, 4, nullptr) == KAMI_EINVAL, "null bytes");

    const int32_t fake_ptr_i = 0;
    (void)fake_ptr_i;
    const int32_t* rp = reinterpret_cast<const int32_t*>(uintptr_t(1) << 20);
    const void* vals = reinterpret_cast<const void*>(uintptr_t(1) << 21);
    int32_t* out = reinterpret_cast<int32_t*>(uintptr_t(1) << 22);
    int64_t nnz = 0;
    for (int32_t mb : {INT32_MIN, -1, 0, 1, 64, INT32_MAX})
        for (int32_t nb : {-1, 0, 1, 64, INT32_MAX})
            for (int32_t bs : {-1, 0, 8, 16, 32, 64, INT32_MAX})
                for (int32_t nnzb : {INT32_MIN, -1, 0, 1, 1000, INT32_MAX}) {
                    kami_bsr A{mb, nb, bs, nnzb, rp, rp, vals};
                    kami_bsr B{nb, mb, bs, nnzb, rp, rp, vals};
                    kami_status st;
                    if (!g_have_device) {
                        st = kami_bsr_spgemm_symbolic(&A, &B, out, &nnz, nullptr, 0, nullptr);
                        CHECK(st != KAMI_OK, "symbolic without device");
                        st = kami_bsr_spgemm_fill(&A, &B, rp, out, nullptr, 0, nullptr);
                        CHECK(st != KAMI_OK, "fill without device");
                        for (int prec : {-1, 0, 1, 2, 3})
                            for (int64_t nc : {(int64_t)-1, (int64_t)0, (int64_t)5, INT64_MAX}) {
                                st = kami_bsr_spgemm_numeric(&A, &B, rp, rp, nc, out, prec, 1.0,
                                                             nullptr);
                                CHECK(st != KAMI_OK || nc == 0, "numeric without device");
                                st = kami_bsr_spgemm_numeric_repeat(&A, &B, rp, rp, nc, out, prec,
                                                                    1.0, 3, nullptr);
                                CHECK(st != KAMI_OK || nc == 0, "repeat without device");
                            }
                        for (int32_t n : {-16, 0, 15, 16, 256, INT32_MAX - 15})
                            for (int64_t ld : {(int64_t)-1, (int64_t)0, (int64_t)1 << 40, INT64_MAX})
                                for (int prec : {0, 1, 2, 5}) {
                                    st = kami_bsr_spmm(&A, vals, n, ld, KAMI_B_NK, out, ld, prec,
                                                       1.0, 0.0, nullptr);
                                    CHECK(st != KAMI_OK, "spmm without device");
                                    st = kami_bsr_spmm(&A, vals, n, ld, KAMI_B_KN, out, 0, prec,
                                                       1.0, 1.0, nullptr);
                                    CHECK(st != KAMI_OK, "spmm KN without device");
                                }
                    }
                }
    CHECK(kami_bsr_spgemm_symbolic(nullptr, nullptr, out, &nnz, nullptr, 0, nullptr) == KAMI_EINVAL,
          "null symbolic");
    CHECK(kami_bsr_spmm(nullptr, vals, 16, 0, 0, out, 0, 0, 1.0, 0.0, nullptr) == KAMI_EINVAL,
          "null spmm");

    const double* X = reinterpret_cast<const double*>(uintptr_t(1) << 20);
    uint16_t* dig = reinterpret_cast<uint16_t*>(uintptr_t(1) << 21);
    int32_t* ex = reinterpret_cast<int32_t*>(uintptr_t(1) << 22);
    if (!g_have_device)
        for (int64_t rows : {(int64_t)-1, (int64_t)0, (int64_t)1, (int64_t)1 << 40, INT64_MAX})
            for (int64_t k : {(int64_t)-1, (int64_t)0, (int64_t)33, (int64_t)1 << 33, INT64_MAX})
                for (int64_t ldx : {(int64_t)-1, (int64_t)0, (int64_t)64, INT64_MAX})
                    for (int prec : {0, 3, 4, 5})
                        for (int slices : {-1, 0, 1, 8, 9}) {
                            const kami_status st =
                                kami_split_slices(X, rows, k, ldx, prec, slices, dig, ex, nullptr);
                            CHECK(st != KAMI_OK || rows == 0 || k == 0, "slices without device");
                        }
    void* dst = reinterpret_cast<void*>(uintptr_t(1) << 20);
    if (!g_have_device)
        for (int kind : {-1, 0, 1, 2})
            for (int prec : {-1, 0, 1, 2, 3})
                for (int64_t count : {(int64_t)-1, (int64_t)0, (int64_t)1, (int64_t)1 << 40, INT64_MAX})
                    for (int64_t off : {(int64_t)-1, (int64_t)0, INT64_MAX}) {
                        const kami_status st = kami_generate(kind, prec, dst, count, off, nullptr);
                        CHECK(st != KAMI_OK || count == 0, "generate without device");
                    }
    std::printf("sparse and utility arguments: checked\n");
}

// ------------------------------------------------------------------ oracle

void check_oracle() {
    // binary16 / bfloat16: every finite pattern round-trips
    int h_ok = 0, b_ok = 0;
    for (int h = 0; h < 65536; ++h) {
        const double x = oracle_half_to_double((uint16_t)h);
        if (x == x && x - x == 0) h_ok += oracle_double_to_half(x) == (uint16_t)h || x == 0;
        else h_ok += 1;
        const double y = oracle_bf16_to_double((uint16_t)h);
        if (y == y && y - y == 0) b_ok += oracle_double_to_bf16(y) == (uint16_t)h || y == 0;
        else b_ok += 1;
    }
    CHECK(h_ok == 65536 && b_ok == 65536, "half %d bf16 %d round trips", h_ok, b_ok);

    std::mt19937 rng(7);
    auto small_int = [&]() { return (double)((int)(rng() % 7) - 3); };
    const int m = 48, n = 40, k = 56;
    std::vector<double> A((size_t)m * k), B((size_t)n * k), Bt((size_t)k * n);
    for (double& a : A) a = small_int();
    for (int j = 0; j < n; ++j)
        for (int q = 0; q < k; ++q) B[(size_t)j * k + q] = Bt[(size_t)q * n + j] = small_int();
    std::vector<double> ref((size_t)m * n, 0.0);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j)
            for (int q = 0; q < k; ++q) ref[(size_t)i * n + j] += A[(size_t)i * k + q] * B[(size_t)j * k + q];
    std::vector<double> C((size_t)m * n), absC((size_t)m * n);
    oracle_gemm_f64(m, n, k, A.data(), k, B.data(), k, 0, C.data(), n);
    CHECK(C == ref, "gemm_f64 NK");
    oracle_gemm_f64(m, n, k, A.data(), k, Bt.data(), n, 1, C.data(), n);
    CHECK(C == ref, "gemm_f64 KN");
    oracle_gemm_dot2_f64(m, n, k, A.data(), k, B.data(), k, 0, C.data(), n, absC.data());
    CHECK(C == ref, "gemm_dot2");
    std::vector<uint16_t> Ah(A.size()), Bh(B.size());
    for (size_t i = 0; i < A.size(); ++i) Ah[i] = oracle_double_to_half(A[i]);
    for (size_t i = 0; i < B.size(); ++i) Bh[i] = oracle_double_to_half(B[i]);
    oracle_gemm_f16(m, n, k, Ah.data(), k, Bh.data(), k, 0, C.data(), n);
    CHECK(C == ref, "gemm_f16");
    for (size_t i = 0; i < A.size(); ++i) Ah[i] = oracle_double_to_bf16(A[i]);
    for (size_t i = 0; i < B.size(); ++i) Bh[i] = oracle_double_to_bf16(B[i]);
    oracle_gemm_bf16(m, n, k, Ah.data(), k, Bh.data(), k, 0, C.data(), n);
    CHECK(C == ref, "gemm_bf16");

    // batched: three copies of the same problem, and an uneven pair
    std::vector<double> A3, B3, C3((size_t)3 * m * n);
    for (int r = 0; r < 3; ++r) {
        A3.insert(A3.end(), A.begin(), A.end());
        B3.insert(B3.end(), B.begin(), B.end());
    }
    oracle_batched_gemm_f64(3, m, n, k, A3.data(), (int64_t)m * k, B3.data(), (int64_t)n * k,
                            C3.data(), (int64_t)m * n, 0, 2);
    for (int r = 0; r < 3; ++r)
        CHECK(std::equal(ref.begin(), ref.end(), C3.begin() + (size_t)r * m * n), "batched %d", r);
    const int32_t mnk[6] = {m, n, k, 16, 16, 16};
    std::vector<double> Au(A), Bu(B), Cu((size_t)m * n + 256);
    Au.resize(A.size() + 256, 1.0);
    Bu.resize(B.size() + 256, 1.0);
    oracle_batched_uneven_f64(2, mnk, Au.data(), Bu.data(), Cu.data(), 2);
    CHECK(std::equal(ref.begin(), ref.end(), Cu.begin()) && Cu[(size_t)m * n] == 16.0, "uneven");

    std::vector<double> slab((size_t)3 * 4 * 5, 1.0);
    oracle_slab_sum_f64(4, 5, 3, slab.data());
    CHECK(slab[0] == 3.0 && slab[19] == 3.0, "slab sum");
    std::vector<double> fill(1000);
    oracle_fill_ramp(fill.data(), 1000);
    CHECK(fill[999] == 999.0, "ramp");
    oracle_fill_randr_mod3(fill.data(), 1000);
    for (double v : fill) CHECK(v == 0.0 || v == 1.0 || v == 2.0, "rand_r mod 3");

    // block SpGEMM on a generated matrix against a dense product of the block patterns
    const int mb = 24, bs = 16;
    std::vector<int32_t> rpA(mb + 1), ciA((size_t)mb * mb);
    std::vector<uint16_t> vA((size_t)mb * mb * bs * bs);
    const int nnzb = oracle_bsr_generate(mb, mb, 0.8f, bs, rpA.data(), ciA.data(), vA.data());
    CHECK(nnzb > 0 && nnzb < mb * mb && rpA[mb] == nnzb, "generate %d", nnzb);
    std::vector<int32_t> rpC(mb + 1);
    const int64_t nnzbC = oracle_spgemm_symbolic(mb, mb, rpA.data(), ciA.data(), rpA.data(),
                                                 ciA.data(), rpC.data());
    std::vector<unsigned char> pat((size_t)mb * mb, 0);
    for (int i = 0; i < mb; ++i)
        for (int a = rpA[i]; a < rpA[i + 1]; ++a)
            for (int b = rpA[ciA[a]]; b < rpA[ciA[a] + 1]; ++b) pat[(size_t)i * mb + ciA[b]] = 1;
    int64_t dense_nnz = 0;
    for (unsigned char x : pat) dense_nnz += x;
    CHECK(nnzbC == dense_nnz && rpC[mb] == nnzbC, "symbolic %lld vs %lld", (long long)nnzbC,
          (long long)dense_nnz);
    std::vector<int32_t> ciC((size_t)std::max<int64_t>(nnzbC, 1));
    oracle_spgemm_fill(mb, mb, rpA.data(), ciA.data(), rpA.data(), ciA.data(), rpC.data(), ciC.data());
    for (int i = 0; i < mb; ++i)
        for (int q = rpC[i]; q < rpC[i + 1]; ++q) {
            CHECK(pat[(size_t)i * mb + ciC[q]] == 1, "fill column");
            if (q > rpC[i]) CHECK(ciC[q] > ciC[q - 1], "fill order");
        }
    std::vector<double> vC((size_t)std::max<int64_t>(nnzbC, 1) * bs * bs);
    oracle_spgemm_numeric_u16(mb, mb, bs, rpA.data(), ciA.data(), vA.data(), rpA.data(), ciA.data(),
                              vA.data(), rpC.data(), vC.data(), 0);
    // dense check of block row 0
    const int N = mb * bs;
    std::vector<double> dA((size_t)N * N, 0.0);
    for (int i = 0; i < mb; ++i)
        for (int a = rpA[i]; a < rpA[i + 1]; ++a)
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c)
                    dA[(size_t)(i * bs + r) * N + ciA[a] * bs + c] =
                        oracle_half_to_double(vA[(size_t)a * bs * bs + r * bs + c]);
    bool num_ok = true;
    for (int q = rpC[0]; q < rpC[1]; ++q)
        for (int r = 0; r < bs; ++r)
            for (int c = 0; c < bs; ++c) {
                double s = 0;
                for (int t = 0; t < N; ++t) s += dA[(size_t)r * N + t] * dA[(size_t)t * N + ciC[q] * bs + c];
                num_ok &= s == vC[(size_t)q * bs * bs + r * bs + c];
            }
    CHECK(num_ok, "numeric block row 0");
    int64_t cub = 0;
    for (int a = 0; a < nnzb; ++a) cub += rpA[ciA[a] + 1] - rpA[ciA[a]];
    CHECK(oracle_spgemm_nnzcub(nnzb, ciA.data(), rpA.data()) == cub, "nnzcub");
    std::vector<int32_t> sc = {3, 1, 4, 1, 5};
    oracle_exclusive_scan_i32(sc.data(), 5);
    CHECK(sc == std::vector<int32_t>({0, 3, 4, 8, 9}), "scan");

    // block SpMM against the dense product
    const int sn = 32;
    std::vector<int32_t> srp(mb + 1), sci((size_t)mb * mb);
    std::vector<double> sv((size_t)mb * mb * bs * bs), sB((size_t)sn * N);
    const int snnz = oracle_spmm_generate(mb, mb, 0.8f, bs, srp.data(), sci.data(), sv.data(), sn,
                                          sB.data());
    CHECK(snnz == srp[mb], "spmm generate");
    std::vector<double> sC((size_t)N * sn, 0.0);
    oracle_bsr_spmm_f64(mb, mb, bs, srp.data(), sci.data(), sv.data(), sB.data(), sn, N, 0,
                        sC.data(), sn);
    bool spmm_ok = true;
    for (int i = 0; i < mb && spmm_ok; ++i)
        for (int r = 0; r < bs; ++r)
            for (int j = 0; j < sn; ++j) {
                double s = 0;
                for (int a = srp[i]; a < srp[i + 1]; ++a)
                    for (int c = 0; c < bs; ++c)
                        s += sv[(size_t)a * bs * bs + r * bs + c] * sB[(size_t)j * N + sci[a] * bs + c];
                spmm_ok &= s == sC[(size_t)(i * bs + r) * sn + j];
            }
    CHECK(spmm_ok, "spmm");
    CHECK(oracle_max_threads() >= 1, "threads");
    std::printf("oracle: conversions, products, SpGEMM and SpMM checked\n");
}

}  // namespace

int main() {
    check_status_and_version();
    check_descriptors();
    check_tune_cache();
    check_batched_queue();
    check_batched_arguments();
    check_sparse_and_utility();
    check_oracle();
    std::printf("%d checks, %d failed%s\n", g_checks, g_failed,
                g_have_device ? " (device present: compute calls skipped)" : "");
    return g_failed ? 1 : 0;
}
