// Tuning choices kept across processes.
//
// The reference fixes its per-size choices in the sweep scripts' tables
// (scripts/fp64_block_square_GH200.sh:11-25, scripts/fp64_batched_GH200.sh:96-123).  Here
// kami_block_gemm_autotune and batched plans created with autotune record their choices in
// process-wide tables; kami_tune_save writes them to a text file and kami_tune_load merges
// a file back, after which KAMI_PART_AUTO descriptors and batched plans created without
// autotune start from the recorded choices.
//
// File format, one record per line:
//   kami-tune 1
//   block <precision> <m> <n> <k> <partition> <ranks> <pipeline>
//   batched <precision> <bucket> <variant>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kami_internal.hpp"

namespace kami {

static std::mutex g_batched_mu;
static int g_batched_variant[5][kBatchedBuckets];
static bool g_batched_init = false;

static void batched_init_locked() {
    if (g_batched_init) return;
    for (auto& row : g_batched_variant)
        for (int& v : row) v = -1;
    g_batched_init = true;
}

static bool batched_prec_ok(int prec) {
    return prec == KAMI_PREC_F64 || prec == KAMI_PREC_F16_F32 || prec == KAMI_PREC_BF16_F32;
}

int batched_tuned_variant(int prec, int bucket) {
    if (!batched_prec_ok(prec) || bucket < 0 || bucket >= kBatchedBuckets) return -1;
    std::lock_guard<std::mutex> lk(g_batched_mu);
    batched_init_locked();
    return g_batched_variant[prec][bucket];
}

void set_batched_tuned_variant(int prec, int bucket, int variant) {
    if (!batched_prec_ok(prec) || bucket < 0 || bucket >= kBatchedBuckets) return;
    if (variant < 0 || variant >= kBatchedVariants) return;
    std::lock_guard<std::mutex> lk(g_batched_mu);
    batched_init_locked();
    g_batched_variant[prec][bucket] = variant;
}

void clear_batched_tuned_variants() {
    std::lock_guard<std::mutex> lk(g_batched_mu);
    g_batched_init = false;
    batched_init_locked();
}

}  // namespace kami

using namespace kami;

extern "C" {

kami_status kami_tune_save(const char* path) {
    if (!path) return KAMI_EINVAL;
    FILE* f = std::fopen(path, "w");
    if (!f) return KAMI_EINVAL;
    std::fprintf(f, "kami-tune 1\n");
    for (const TunedBlock& t : tuned_block_choices())
        std::fprintf(f, "block %d %d %d %d %d %d %d\n", t.prec, t.m, t.n, t.k, t.part, t.ranks,
                     t.pipe);
    for (int prec = 0; prec < 5; ++prec)
        for (int b = 0; b < kBatchedBuckets; ++b) {
            const int v = batched_tuned_variant(prec, b);
            if (v >= 0) std::fprintf(f, "batched %d %d %d\n", prec, b, v);
        }
    const bool ok = std::fflush(f) == 0;
    return (std::fclose(f) == 0 && ok) ? KAMI_OK : KAMI_EINVAL;
}

kami_status kami_tune_load(const char* path) {
    if (!path) return KAMI_EINVAL;
    FILE* f = std::fopen(path, "r");
    if (!f) return KAMI_EINVAL;
    std::vector<TunedBlock> blocks;
    std::vector<int> batched;  // prec, bucket, variant triples
    char* line = nullptr;
    size_t cap = 0;
    ssize_t len;
    bool header = false, ok = true;
    // a record is the whole line: its fields, then only white space
    auto rest_blank = [](const char* p) {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
        return *p == '\0';
    };
    while (ok && (len = getline(&line, &cap, f)) != -1) {
        if ((size_t)len != std::strlen(line)) {  // a NUL byte inside the line
            ok = false;
            break;
        }
        if (line[0] == '\n' || line[0] == '#') continue;
        int used = -1;
        if (!header) {
            int ver = 0;
            ok = std::sscanf(line, "kami-tune %d%n", &ver, &used) == 1 && used > 0 &&
                 rest_blank(line + used) && ver == 1;
            header = true;
            continue;
        }
        TunedBlock t;
        int prec, bucket, variant;
        if (std::sscanf(line, "block %d %d %d %d %d %d %d%n", &t.prec, &t.m, &t.n, &t.k, &t.part,
                        &t.ranks, &t.pipe, &used) == 7 &&
            used > 0 && rest_blank(line + used)) {
            // the choice must name a compiled instance (checked like a query)
            kami_gemm_desc d;
            ok = kami_gemm_desc_init(&d, t.m, t.n, t.k, t.prec) == KAMI_OK && !is_split(t.prec) &&
                 t.part >= KAMI_PART_1D && t.part <= KAMI_PART_3D && t.ranks > 0 &&
                 (t.pipe == 1 || t.pipe == 2);
            if (ok) {
                d.partition = t.part;
                d.ranks = t.ranks;
                d.pipeline = t.pipe;
                kami_status st;
                ok = resolve_block(&d, &st) != nullptr;
            }
            if (ok) blocks.push_back(t);
        } else if (std::sscanf(line, "batched %d %d %d%n", &prec, &bucket, &variant, &used) == 3 &&
                   used > 0 && rest_blank(line + used)) {
            ok = batched_prec_ok(prec) && bucket >= 0 && bucket < kBatchedBuckets &&
                 variant >= 0 && variant < kBatchedVariants;
            if (ok) batched.insert(batched.end(), {prec, bucket, variant});
        } else {
            ok = false;
        }
    }
    std::free(line);
    std::fclose(f);
    if (!ok || !header) return KAMI_EINVAL;  // nothing is applied from a bad file
    for (const TunedBlock& t : blocks) set_tuned_choice(t.prec, t.m, t.n, t.k, t.part, t.ranks, t.pipe);
    for (size_t i = 0; i < batched.size(); i += 3)
        set_batched_tuned_variant(batched[i], batched[i + 1], batched[i + 2]);
    return KAMI_OK;
}

kami_status kami_tune_clear(void) {
    clear_tuned_block_choices();
    clear_batched_tuned_variants();
    return KAMI_OK;
}

}  // extern "C"
