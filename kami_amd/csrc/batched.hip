// Batched GEMM over problems of uneven size: one persistent launch, a device work queue,
// per-size-bucket tile variants chosen by an autotuner, optional hipGraph execution.
//
// Reference: batched_gemm_{1d,2d,3d}_double_mma (src/batched_gemm/KAMI/1d/
// batched_gemm_1d_double_mma.cu:59-218) runs one uniform-size problem per CTA at fixed
// strides (:71-74) with the partition picked per size by the sweep script
// (scripts/fp64_batched_GH200.sh:96-123).  Here:
//   * problems carry their own {m, n, k, lda, ldb, ldc, A, B, C} (the kami_problem array);
//   * every problem is cut into output tiles; a tile is one queue entry
//     {problem, row0, col0, variant};
//   * a persistent grid (CUs x resident workgroups) pops entries with one atomic add
//     per tile, in descending-cost order (longest first), until the queue is empty;
//   * the variant of a tile is the tile shape and wave partition of its size bucket,
//     picked by timing candidates on the device at plan creation (or from priors);
//   * one execution is [reset the queue counter] -> [persistent kernel], captured once
//     into a hipGraph and replayed.  The counter is cleared by a one-thread kernel,
//     never by a memset node.
// M, N and K must be multiples of 16.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "kami_internal.hpp"

namespace kami {

struct WorkItem {
    int32_t prob, row0, col0, variant;
};

// Staging shared by every tile variant: a stage holds up to TMAX rows of A and TNMAX rows
// of B (or KC rows of TNMAX columns when B is K x N), KC elements of K each, in LDS rows
// padded by 16 bytes.  Pieces are 16 bytes; a thread owns pieces tid + i * NT.
template <class AT>
struct StreamGeom {
    static constexpr int NT = 256;
    static constexpr int ELT = (int)sizeof(typename AT::in_t), VE = 16 / ELT;
    static constexpr int KC = ELT == 8 ? 16 : 64;
    static constexpr int TMAX = 128, TNMAX = 64;
    static constexpr int ROWB = KC * ELT + 16;
    static constexpr int A_REGION = TMAX * ROWB;
    static constexpr int STAGE = (TMAX + TNMAX) * ROWB;
    static constexpr int LDS = 2 * STAGE;
    static constexpr int CPR = KC / VE;  // pieces per staged row
    static constexpr int NP = ((TMAX + TNMAX) * CPR + NT - 1) / NT;
    static_assert((KC * TNMAX) / VE <= TNMAX * CPR, "K x N staging fits the B region");
};

// Tile geometry of a variant: TM x TN output tile, WM x WN waves (4 waves).
template <class AT, int TM_, int TN_, int WM_, int WN_>
struct TileGeom {
    using SG = StreamGeom<AT>;
    static constexpr int TM = TM_, TN = TN_, WM = WM_, WN = WN_;
    static_assert(WM * WN * 64 == SG::NT, "four waves");
    static_assert(TM <= SG::TMAX && TN <= SG::TNMAX, "tile fits the stage");
    static constexpr int SUBM = TM / WM, SUBN = TN / WN, IM = SUBM / 16, JN = SUBN / 16;
    static_assert(SUBM % 16 == 0 && SUBN % 16 == 0 && SG::KC % AT::KSTEP == 0, "tile");
};

// Tile variants.  Index = variant id stored in the work items.
//   0: 64 x 64 tile, 2 x 2 waves (2D)       1: 32 x 32 tile, 2 x 2 waves (2D)
//   2: 64 x 64 tile, 4 x 1 waves (1D)       3: 128 x 64 tile, 2 x 2 waves (2D)
constexpr int kVariants = 4;
static const int kVariantTM[kVariants] = {64, 32, 64, 128};
static const int kVariantTN[kVariants] = {64, 32, 64, 64};

template <class AT>
struct Variants {
    using V0 = TileGeom<AT, 64, 64, 2, 2>;
    using V1 = TileGeom<AT, 32, 32, 2, 2>;
    using V2 = TileGeom<AT, 64, 64, 4, 1>;
    using V3 = TileGeom<AT, 128, 64, 2, 2>;
};

__device__ inline int variant_tm(int v) { return v == 1 ? 32 : (v == 3 ? 128 : 64); }
__device__ inline int variant_tn(int v) { return v == 1 ? 32 : 64; }

struct BatchArgs {
    const kami_problem* probs;
    const WorkItem* items;
    unsigned int* counter;
    int64_t n_items;
    double alpha, beta;
    int32_t b_kn;
};

// A queue entry as the stream sees it (identical in every thread of a workgroup).
struct TileRef {
    int prob, row0, col0, variant, nch;
    bool valid;
};

// Loads the pieces of stage (tile t, K chunk c) into registers; rows outside C and K
// beyond the problem are zero.
template <class AT>
__device__ __forceinline__ void stream_load(const BatchArgs& a, const TileRef& t, int c, bool b_kn,
                            u4_t (&r)[StreamGeom<AT>::NP]) {
    using SG = StreamGeom<AT>;
    using in_t = typename AT::in_t;
    const kami_problem P = a.probs[t.prob];
    const int tm = variant_tm(t.variant), tn = variant_tn(t.variant);
    const int mlim = P.m - t.row0, nlim = P.n - t.col0, K = P.k, k0 = c * SG::KC;
    const in_t* A = reinterpret_cast<const in_t*>(P.A);
    const in_t* B = reinterpret_cast<const in_t*>(P.B);
    const int ca = tm * SG::CPR;
#pragma unroll
    for (int i = 0; i < SG::NP; ++i) {
        const int idx = threadIdx.x + i * SG::NT;
        u4_t v = {0, 0, 0, 0};
        if (idx < ca) {
            const int row = idx / SG::CPR, kc = (idx % SG::CPR) * SG::VE;
            if (row < mlim && k0 + kc < K) v = load_raw<16>(A + (int64_t)(t.row0 + row) * P.lda + k0 + kc);
        } else if (!b_kn) {
            const int j = idx - ca;
            const int row = j / SG::CPR, kc = (j % SG::CPR) * SG::VE;
            if (row < tn && row < nlim && k0 + kc < K)
                v = load_raw<16>(B + (int64_t)(t.col0 + row) * P.ldb + k0 + kc);
        } else {
            const int j = idx - ca, ppr = tn / SG::VE;  // pieces per K row of the tile
            const int kr = j / ppr, jc = (j % ppr) * SG::VE;
            if (kr < SG::KC && jc < nlim && k0 + kr < K)
                v = load_raw<16>(B + (int64_t)(k0 + kr) * P.ldb + t.col0 + jc);
        }
        r[i] = v;
    }
}

template <class AT>
__device__ __forceinline__ void stream_write(int variant, bool b_kn, const u4_t (&r)[StreamGeom<AT>::NP],
                             unsigned char* buf) {
    using SG = StreamGeom<AT>;
    using in_t = typename AT::in_t;
    const int tm = variant_tm(variant), tn = variant_tn(variant);
    const int ca = tm * SG::CPR;
#pragma unroll
    for (int i = 0; i < SG::NP; ++i) {
        const int idx = threadIdx.x + i * SG::NT;
        if (idx < ca) {
            store_raw<16>(buf + (idx / SG::CPR) * SG::ROWB + (idx % SG::CPR) * 16, r[i]);
        } else if (!b_kn) {
            const int j = idx - ca;
            if (j < tn * SG::CPR)
                store_raw<16>(buf + SG::A_REGION + (j / SG::CPR) * SG::ROWB + (j % SG::CPR) * 16, r[i]);
        } else {
            const int j = idx - ca, ppr = tn / SG::VE;
            const int kr = j / ppr, jc = (j % ppr) * SG::VE;
            if (kr < SG::KC) {
                const in_t* e = reinterpret_cast<const in_t*>(&r[i]);
#pragma unroll
                for (int q = 0; q < SG::VE; ++q)
                    *reinterpret_cast<in_t*>(buf + SG::A_REGION + (jc + q) * SG::ROWB + kr * SG::ELT) = e[q];
            }
        }
    }
}

// Stream state of a workgroup: the current tile and two queued behind it, the position
// of the next stage to load, and a two-deep register ring of staged pieces (slot 0 =
// the stage after the one in LDS, slot 1 = the stage after that).
template <class AT>
struct Stream {
    TileRef T[3];
    int lt, lc;  // load cursor: tile index into T, K chunk
    u4_t ring0[StreamGeom<AT>::NP], ring1[StreamGeom<AT>::NP];
    int v0, v1;  // variant of the stage in each slot, -1 = empty
    int buf;     // LDS buffer holding the stage being computed
};

__device__ __forceinline__ TileRef fetch_tile(const BatchArgs& a, unsigned int* s_item) {
    if (threadIdx.x == 0) *s_item = atomicAdd(a.counter, 1u);
    __syncthreads();
    const int64_t it = *s_item;
    __syncthreads();
    TileRef t;
    t.valid = it < a.n_items;  // identical in every thread: the exit is uniform
    if (t.valid) {
        const WorkItem w = a.items[it];
        t.prob = w.prob;
        t.row0 = w.row0;
        t.col0 = w.col0;
        t.variant = w.variant;
        t.nch = 0;
    } else {
        t.prob = t.row0 = t.col0 = t.variant = t.nch = 0;
    }
    return t;
}

template <class AT>
__device__ __forceinline__ void set_chunks(const BatchArgs& a, TileRef& t) {
    if (t.valid) t.nch = (a.probs[t.prob].k + StreamGeom<AT>::KC - 1) / StreamGeom<AT>::KC;
}

// Field-wise selection: a runtime index into the private array, or a select between
// their addresses, would keep the whole stream state in scratch memory.
__device__ __forceinline__ TileRef pick_tile(int i, const TileRef& x, const TileRef& y,
                                             const TileRef& z) {
    TileRef t;
    t.prob = i == 0 ? x.prob : (i == 1 ? y.prob : z.prob);
    t.row0 = i == 0 ? x.row0 : (i == 1 ? y.row0 : z.row0);
    t.col0 = i == 0 ? x.col0 : (i == 1 ? y.col0 : z.col0);
    t.variant = i == 0 ? x.variant : (i == 1 ? y.variant : z.variant);
    t.nch = i == 0 ? x.nch : (i == 1 ? y.nch : z.nch);
    t.valid = i == 0 ? x.valid : (i == 1 ? y.valid : z.valid);
    return t;
}

// Load the stage at the cursor into `dst` (if any) and advance the cursor.
template <class AT>
__device__ __forceinline__ int load_at_cursor(const BatchArgs& a, Stream<AT>& st, bool b_kn,
                              u4_t (&dst)[StreamGeom<AT>::NP]) {
    if (st.lt > 2) return -1;
    const TileRef t = pick_tile(st.lt, st.T[0], st.T[1], st.T[2]);
    if (!t.valid) return -1;
    stream_load<AT>(a, t, st.lc, b_kn, dst);
    const int v = t.variant;
    if (++st.lc >= t.nch) {
        st.lc = 0;
        ++st.lt;
    }
    return v;
}

template <class AT, class V>
__device__ __forceinline__ void run_tile(const BatchArgs& a, Stream<AT>& st, unsigned char* smem, bool b_kn) {
    using SG = StreamGeom<AT>;
    using acc_t = typename AT::acc_t;
    using out_t = typename AT::out_t;
    using frag_t = typename AT::frag_t;
    using accv_t = typename AT::accv_t;
    constexpr int FB = AT::FRAG_BYTES;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / V::WN, wn = wave % V::WN;
    const TileRef cur = st.T[0];
    const kami_problem P = a.probs[cur.prob];
    const int mlim = P.m - cur.row0, nlim = P.n - cur.col0;

    accv_t acc[V::IM][V::JN];
#pragma unroll
    for (int i = 0; i < V::IM; ++i)
#pragma unroll
        for (int j = 0; j < V::JN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[i][j][v] = 0;

    for (int c = 0; c < cur.nch; ++c) {
        const unsigned char* base = smem + st.buf * SG::STAGE;
        const unsigned char* bb = base + SG::A_REGION;
#pragma unroll
        for (int ks = 0; ks < SG::KC / AT::KSTEP; ++ks) {
            const int kk = (ks * AT::KSTEP + AT::KV * (lane >> 4)) * SG::ELT;
            frag_t af[V::IM];
#pragma unroll
            for (int im = 0; im < V::IM; ++im)
                af[im] = bitcast<frag_t>(
                    load_raw<FB>(base + (wm * V::SUBM + im * 16 + (lane & 15)) * SG::ROWB + kk));
#pragma unroll
            for (int jn = 0; jn < V::JN; ++jn) {
                if (wn * V::SUBN + jn * 16 >= nlim) continue;  // atom outside C: skip
                const frag_t bf = bitcast<frag_t>(
                    load_raw<FB>(bb + (wn * V::SUBN + jn * 16 + (lane & 15)) * SG::ROWB + kk));
#pragma unroll
                for (int im = 0; im < V::IM; ++im)
                    if (wm * V::SUBM + im * 16 < mlim) AT::mma(acc[im][jn], af[im], bf);
            }
        }
        // the stage after this one goes to the other buffer; the ring shifts and the
        // stage two ahead (possibly of a later tile) is loaded
        if (st.v0 >= 0) stream_write<AT>(st.v0, b_kn, st.ring0, smem + (st.buf ^ 1) * SG::STAGE);
#pragma unroll
        for (int i = 0; i < SG::NP; ++i) st.ring0[i] = st.ring1[i];
        st.v0 = st.v1;
        st.v1 = load_at_cursor<AT>(a, st, b_kn, st.ring1);
        __syncthreads();
        st.buf ^= 1;
    }

    const acc_t alpha = (acc_t)a.alpha, beta = (acc_t)a.beta;
    out_t* C = reinterpret_cast<out_t*>(P.C);
#pragma unroll
    for (int im = 0; im < V::IM; ++im)
#pragma unroll
        for (int jn = 0; jn < V::JN; ++jn) {
            const int rb = wm * V::SUBM + im * 16, cb = wn * V::SUBN + jn * 16;
            if (rb >= mlim || cb >= nlim) continue;
            const int col = cur.col0 + cb + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = cur.row0 + rb + AT::out_row(lane, v);
                out_t* dst = C + (int64_t)row * P.ldc + col;
                acc_t val = alpha * acc[im][jn][v];
                if (beta != (acc_t)0) val += beta * (acc_t)(*dst);
                *dst = (out_t)val;
            }
        }
}

template <class AT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) batched_persistent_kernel(BatchArgs args) {
    using SG = StreamGeom<AT>;
    using VS = Variants<AT>;
    __shared__ __attribute__((aligned(16))) unsigned char smem[SG::LDS];
    __shared__ unsigned int s_item;
    const bool b_kn = args.b_kn != 0;
    Stream<AT> st;
    st.T[0] = fetch_tile(args, &s_item);
    if (!st.T[0].valid) return;  // uniform
    st.T[1] = fetch_tile(args, &s_item);
    st.T[2] = fetch_tile(args, &s_item);
    set_chunks<AT>(args, st.T[0]);
    set_chunks<AT>(args, st.T[1]);
    set_chunks<AT>(args, st.T[2]);
    st.lt = 0;
    st.lc = 0;
    st.buf = 0;
    // prologue: first stage into LDS buffer 0, the next two into the ring
    const int vfirst = load_at_cursor<AT>(args, st, b_kn, st.ring0);
    stream_write<AT>(vfirst, b_kn, st.ring0, smem);
    st.v0 = load_at_cursor<AT>(args, st, b_kn, st.ring0);
    st.v1 = load_at_cursor<AT>(args, st, b_kn, st.ring1);
    __syncthreads();
    while (st.T[0].valid) {
        switch (st.T[0].variant) {
            case 0: run_tile<AT, typename VS::V0>(args, st, smem, b_kn); break;
            case 1: run_tile<AT, typename VS::V1>(args, st, smem, b_kn); break;
            case 2: run_tile<AT, typename VS::V2>(args, st, smem, b_kn); break;
            default: run_tile<AT, typename VS::V3>(args, st, smem, b_kn); break;
        }
        // the next queued tile becomes current; one more is popped behind it
        st.T[0] = st.T[1];
        st.T[1] = st.T[2];
        st.T[2] = fetch_tile(args, &s_item);
        set_chunks<AT>(args, st.T[2]);
        st.lt -= 1;
        // the cursor may have stopped at the end of the queue; if a tile appeared, the
        // ring is refilled from it
        if (st.v1 < 0 && st.lt >= 0) {
            if (st.v0 < 0) st.v0 = load_at_cursor<AT>(args, st, b_kn, st.ring0);
            if (st.v0 >= 0) st.v1 = load_at_cursor<AT>(args, st, b_kn, st.ring1);
        }
    }
}

__global__ void reset_counter_kernel(unsigned int* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

static int size_class(int x) { return x <= 32 ? 0 : x <= 64 ? 1 : x <= 128 ? 2 : x <= 256 ? 3 : 4; }
static int bucket_of(const kami_problem& p) {
    return (size_class(p.m) * 5 + size_class(p.n)) * 5 + size_class(p.k);
}

// Prior variant of a bucket: the largest tile that the problem fills, following the
// reference's choice of fewer, larger warps for larger sizes
// (scripts/fp64_batched_GH200.sh:7-14).
static int prior_variant(int bucket) {
    const int cm = bucket / 25, cn = (bucket / 5) % 5;
    if (cm == 0 && cn == 0) return 1;
    if (cm >= 3 && cn >= 2) return 3;
    return 0;
}

}  // namespace kami

using namespace kami;

struct kami_batched_plan {
    int prec = 0;
    int64_t problems = 0;
    int64_t n_items = 0;
    int grid = 0;
    double flops = 0;
    int b_kn = 0;
    double alpha = 1.0, beta = 0.0;
    kami_problem* d_probs = nullptr;
    WorkItem* d_items = nullptr;
    unsigned int* d_counter = nullptr;
    bool use_graph = false;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> bucket_ids;      // non-empty buckets, ascending
    std::vector<int> bucket_variant;  // chosen variant per non-empty bucket
};

namespace kami {

static hipError_t launch_batched(int prec, const BatchArgs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, s, a.counter);
    if (prec == KAMI_PREC_F64)
        hipLaunchKernelGGL(batched_persistent_kernel<AtomF64>, dim3(grid), dim3(256), 0, s, a);
    else if (prec == KAMI_PREC_F16_F32)
        hipLaunchKernelGGL(batched_persistent_kernel<AtomF16x32>, dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(batched_persistent_kernel<AtomBF16x32>, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

static hipError_t persistent_grid(int prec, int* grid) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if (prec == KAMI_PREC_F64)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_persistent_kernel<AtomF64>, 256, 0);
    else if (prec == KAMI_PREC_F16_F32)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_persistent_kernel<AtomF16x32>, 256, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_persistent_kernel<AtomBF16x32>, 256, 0);
    if (e != hipSuccess) return e;
    *grid = std::max(1, cus * std::max(1, per_cu));
    return hipSuccess;
}

// Cut every problem into tiles of its bucket's variant; longest tiles first.
static void build_items(const std::vector<kami_problem>& probs, const std::vector<int>& bucket,
                        const int* variant_of_bucket, std::vector<WorkItem>& items) {
    struct Keyed {
        int64_t cost;
        WorkItem w;
    };
    std::vector<Keyed> tmp;
    for (size_t p = 0; p < probs.size(); ++p) {
        const int v = variant_of_bucket[bucket[p]];
        const int tm = kVariantTM[v], tn = kVariantTN[v];
        for (int r = 0; r < probs[p].m; r += tm)
            for (int c = 0; c < probs[p].n; c += tn) {
                const int64_t cost = (int64_t)std::min(tm, probs[p].m - r) *
                                     std::min(tn, probs[p].n - c) * probs[p].k;
                tmp.push_back({cost, WorkItem{(int32_t)p, r, c, v}});
            }
    }
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const Keyed& a, const Keyed& b) { return a.cost > b.cost; });
    items.clear();
    items.reserve(tmp.size());
    for (auto& k : tmp) items.push_back(k.w);
}

static kami_status upload(const void* host, size_t bytes, void** dev) {
    hipError_t e = hipMalloc(dev, bytes ? bytes : 16);
    if (e != hipSuccess) {
        g_last_hip = (int32_t)e;
        return KAMI_ENOMEM;
    }
    if (bytes) {
        e = hipMemcpy(*dev, host, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e);
    }
    return KAMI_OK;
}

static void free_plan(kami_batched_plan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    if (p->d_probs) (void)hipFree(p->d_probs);
    if (p->d_items) (void)hipFree(p->d_items);
    if (p->d_counter) (void)hipFree(p->d_counter);
    delete p;
}

}  // namespace kami

extern "C" {

kami_status kami_batched_plan_create(const kami_problem* probs, int64_t count, int32_t precision,
                                     const kami_batched_opts* opts, void* stream,
                                     kami_batched_plan** out) {
    if (!probs || !out || count <= 0 || count > 0x7fffffffLL) return KAMI_EINVAL;
    if (precision != KAMI_PREC_F64 && precision != KAMI_PREC_F16_F32 &&
        precision != KAMI_PREC_BF16_F32)
        return KAMI_ENOTSUP;
    kami_batched_opts o;
    std::memset(&o, 0, sizeof(o));
    o.alpha = 1.0;
    if (opts) o = *opts;
    if (o.b_layout != KAMI_B_NK && o.b_layout != KAMI_B_KN) return KAMI_EINVAL;
    if (o.autotune && o.beta != 0.0) return KAMI_EINVAL;  // tuning runs overwrite C
    const int ei = elt_in(precision), eo = elt_out(precision);
    std::vector<kami_problem> hp(probs, probs + count);
    double flops = 0;
    for (auto& p : hp) {
        if (p.m <= 0 || p.n <= 0 || p.k <= 0 || p.m % 16 || p.n % 16 || p.k % 16) return KAMI_EINVAL;
        if (!p.A || !p.B || !p.C) return KAMI_EINVAL;
        if (!p.lda) p.lda = p.k;
        if (!p.ldb) p.ldb = o.b_layout == KAMI_B_NK ? p.k : p.n;
        if (!p.ldc) p.ldc = p.n;
        if (p.lda < p.k || p.ldb < (o.b_layout == KAMI_B_NK ? p.k : p.n) || p.ldc < p.n)
            return KAMI_EINVAL;
        if ((p.lda * ei) % 16 || (p.ldb * ei) % 16 || ((uintptr_t)p.A % 16) ||
            ((uintptr_t)p.B % 16) || ((uintptr_t)p.C % eo))
            return KAMI_EINVAL;
        flops += 2.0 * p.m * p.n * (double)p.k;
    }
    kami_status st = device_check();
    if (st != KAMI_OK) return st;
    hipStream_t s = (hipStream_t)stream;

    kami_batched_plan* plan = new (std::nothrow) kami_batched_plan();
    if (!plan) return KAMI_ENOMEM;
    plan->prec = precision;
    plan->problems = count;
    plan->flops = flops;
    plan->b_kn = o.b_layout == KAMI_B_KN;
    plan->alpha = o.alpha;
    plan->beta = o.beta;
    plan->use_graph = o.use_graph != 0;

    std::vector<int> bucket(hp.size());
    int variant_of_bucket[125];
    bool used[125] = {false};
    for (size_t i = 0; i < hp.size(); ++i) {
        bucket[i] = bucket_of(hp[i]);
        used[bucket[i]] = true;
    }
    for (int b = 0; b < 125; ++b) {
        const int tuned = batched_tuned_variant(precision, b);
        variant_of_bucket[b] = tuned >= 0 ? tuned : prior_variant(b);
    }

    hipError_t e = persistent_grid(precision, &plan->grid);
    if (e != hipSuccess) {
        free_plan(plan);
        return hip_fail(e);
    }
    st = upload(hp.data(), hp.size() * sizeof(kami_problem), (void**)&plan->d_probs);
    if (st == KAMI_OK) st = upload(nullptr, 0, (void**)&plan->d_counter);
    if (st != KAMI_OK) {
        free_plan(plan);
        return st;
    }

    if (o.autotune) {
        // time every variant on up to 256 problems of each bucket, keep the fastest
        hipEvent_t e0, e1;
        e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e != hipSuccess) {
            free_plan(plan);
            return hip_fail(e);
        }
        for (int b = 0; b < 125 && st == KAMI_OK; ++b) {
            if (!used[b]) continue;
            std::vector<kami_problem> sample;
            std::vector<int> sb;
            for (size_t i = 0; i < hp.size() && sample.size() < 256; ++i)
                if (bucket[i] == b) {
                    sample.push_back(hp[i]);
                    sb.push_back(0);
                }
            kami_problem* d_sample = nullptr;
            st = upload(sample.data(), sample.size() * sizeof(kami_problem), (void**)&d_sample);
            if (st != KAMI_OK) break;
            float best = 1e30f;
            for (int v = 0; v < kVariants; ++v) {
                int vb[1] = {v};
                std::vector<WorkItem> items;
                build_items(sample, sb, vb, items);
                WorkItem* d_it = nullptr;
                st = upload(items.data(), items.size() * sizeof(WorkItem), (void**)&d_it);
                if (st != KAMI_OK) break;
                BatchArgs a{d_sample, d_it, plan->d_counter, (int64_t)items.size(), o.alpha, 0.0,
                            plan->b_kn};
                e = launch_batched(precision, a, plan->grid, s);  // warm-up
                if (e == hipSuccess) e = hipEventRecord(e0, s);
                for (int rep = 0; rep < 3 && e == hipSuccess; ++rep)
                    e = launch_batched(precision, a, plan->grid, s);
                if (e == hipSuccess) e = hipEventRecord(e1, s);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                (void)hipFree(d_it);
                if (e != hipSuccess) {
                    st = hip_fail(e);
                    break;
                }
                if (ms < best) {
                    best = ms;
                    variant_of_bucket[b] = v;
                }
            }
            (void)hipFree(d_sample);
            if (st == KAMI_OK) set_batched_tuned_variant(precision, b, variant_of_bucket[b]);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (st != KAMI_OK) {
            free_plan(plan);
            return st;
        }
    }

    std::vector<WorkItem> items;
    build_items(hp, bucket, variant_of_bucket, items);
    plan->n_items = (int64_t)items.size();
    st = upload(items.data(), items.size() * sizeof(WorkItem), (void**)&plan->d_items);
    if (st != KAMI_OK) {
        free_plan(plan);
        return st;
    }
    for (int b = 0; b < 125; ++b)
        if (used[b]) {
            plan->bucket_ids.push_back(b);
            plan->bucket_variant.push_back(variant_of_bucket[b]);
        }

    if (plan->use_graph) {
        e = hipStreamCreateWithFlags(&plan->cap_stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamBeginCapture(plan->cap_stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            BatchArgs a{plan->d_probs, plan->d_items, plan->d_counter, plan->n_items, plan->alpha,
                        plan->beta, plan->b_kn};
            hipError_t le = launch_batched(precision, a, plan->grid, plan->cap_stream);
            e = hipStreamEndCapture(plan->cap_stream, &plan->graph);
            if (le != hipSuccess) e = le;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            free_plan(plan);
            return hip_fail(e);
        }
    }
    *out = plan;
    return KAMI_OK;
}

kami_status kami_batched_execute(kami_batched_plan* plan, void* stream) {
    if (!plan) return KAMI_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (plan->use_graph) {
        e = hipGraphLaunch(plan->exec, s);
    } else {
        BatchArgs a{plan->d_probs, plan->d_items, plan->d_counter, plan->n_items, plan->alpha,
                    plan->beta, plan->b_kn};
        e = launch_batched(plan->prec, a, plan->grid, s);
    }
    if (e != hipSuccess) return hip_fail(e);
    return KAMI_OK;
}

kami_status kami_batched_plan_info(const kami_batched_plan* plan, kami_plan_info* out) {
    if (!plan || !out) return KAMI_EINVAL;
    std::memset(out, 0, sizeof(*out));
    out->problems = plan->problems;
    out->work_items = plan->n_items;
    out->buckets = (int32_t)plan->bucket_ids.size();
    out->workgroups = plan->grid;
    out->flops = plan->flops;
    for (size_t i = 0; i < plan->bucket_variant.size() && i < 16; ++i)
        out->variant_of_bucket[i] = plan->bucket_variant[i];
    return KAMI_OK;
}

kami_status kami_batched_plan_destroy(kami_batched_plan* plan) {
    free_plan(plan);
    return KAMI_OK;
}

}  // extern "C"
